// double instantiation of the single-workgroup LU panel kernel (2 rows of 32 columns per
// thread, panels of up to 1024 rows).
#include "lu_panel.hpp"

namespace pvl
{

int panel_wg(double *P, int lda, int m, int jb, int row_base, int *piv, int *status)
{
    if (m > 1024 || jb > kPanelWgCols || jb < 1)
    {
        set_error("panel_wg: %d x %d panel exceeds the single-workgroup kernel", m, jb);
        return PVL_ERR_ARG;
    }
    lu_panel::panel_wg_kernel<double, kPanelWgCols, 2><<<1, 512, 0, stream()>>>(P, lda, m, jb, row_base, piv, status);
    PVL_HIP(hipGetLastError());
    return PVL_OK;
}

} // namespace pvl
