// float instantiation of the single-workgroup LU panel kernel (4 rows of 32 columns per
// thread, panels of up to 2048 rows).
#include "lu_panel.hpp"

namespace pvl
{

int panel_wg(float *P, int lda, int m, int jb, int row_base, int *piv, int *status)
{
    if (m > 2048 || jb > kPanelWgCols || jb < 1)
    {
        set_error("panel_wg: %d x %d panel exceeds the single-workgroup kernel", m, jb);
        return PVL_ERR_ARG;
    }
    lu_panel::panel_wg_kernel<float, kPanelWgCols, 4><<<1, 512, 0, stream()>>>(P, lda, m, jb, row_base, piv, status);
    PVL_HIP(hipGetLastError());
    return PVL_OK;
}

} // namespace pvl
