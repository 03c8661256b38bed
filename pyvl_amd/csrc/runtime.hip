// Runtime services of libpvl_hip.so: error reporting, device selection, the library
// stream, device memory helpers, scratch buffers and event-based kernel timing.
#include "pvl_internal.hpp"

#include <cstdarg>
#include <cstdlib>
#include <map>
#include <mutex>
#include <string>
#include <vector>

namespace pvl
{

static thread_local char g_err[1024] = "";
static hipStream_t g_stream = nullptr;
static int g_device = -1;
static std::mutex g_mutex;

std::recursive_mutex &api_mutex()
{
    static std::recursive_mutex m;
    return m;
}

void set_error(const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
}

int check_hip(hipError_t e, const char *what)
{
    if (e == hipSuccess)
        return PVL_OK;
    set_error("%s failed: %s", what, hipGetErrorString(e));
    if (e == hipErrorOutOfMemory)
        return PVL_ERR_NOMEM;
    return PVL_ERR_HIP;
}

static int init_locked(int device)
{
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0)
    {
        set_error("no HIP device is visible");
        return PVL_ERR_NODEV;
    }
    if (device < 0 || device >= n)
    {
        set_error("device %d out of range (%d visible)", device, n);
        return PVL_ERR_ARG;
    }
    hipDeviceProp_t prop;
    PVL_HIP(hipGetDeviceProperties(&prop, device));
    if (std::string(prop.gcnArchName).rfind("gfx950", 0) != 0)
    {
        set_error("device %d is %s; this library is built for gfx950 only", device, prop.gcnArchName);
        return PVL_ERR_NODEV;
    }
    PVL_HIP(hipSetDevice(device));
    if (g_stream == nullptr)
        PVL_HIP(hipStreamCreateWithFlags(&g_stream, hipStreamNonBlocking));
    // Failure is not fatal to the rest of the library: the no-allocation entry point repeats
    // the set-up on its first call and reports the error itself.
    (void)noalloc_setup(device);
    g_device = device;
    return PVL_OK;
}

int initialised_device()
{
    std::lock_guard<std::mutex> lk(g_mutex);
    return g_device;
}

int ensure_init()
{
    std::lock_guard<std::mutex> lk(g_mutex);
    if (g_device >= 0)
    {
        if (hipSetDevice(g_device) != hipSuccess)
        {
            set_error("hipSetDevice(%d) failed", g_device);
            return PVL_ERR_HIP;
        }
        return PVL_OK;
    }
    return init_locked(0);
}

hipStream_t stream() { return g_stream; }

// Workgroups per CU given to work on the side stream (environment PVL_SIDE_GROUPS_PER_CU,
// default per_cu): the rest of each CU's slots stay free for the library stream's kernels.
int side_groups(int per_cu)
{
    static int cus = -1;
    if (cus < 0)
    {
        int n = 0;
        if (hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, g_device < 0 ? 0 : g_device) !=
                hipSuccess ||
            n <= 0)
            n = 256;
        cus = n;
    }
    if (const char *e = getenv("PVL_SIDE_GROUPS_PER_CU"))
        per_cu = atoi(e) > 0 ? atoi(e) : per_cu;
    return per_cu * cus;
}

// Ordinary launches by default: every grid-synced kernel runs on the library stream (never
// two at once), no other kernel waits on one of them from the device, and each grid fits on
// the otherwise idle device (checked below), so every group becomes resident once the CUs
// that other streams' kernels hold are free; every wait on another group is bounded.  A
// cooperative launch (PVL_COOPERATIVE=1) adds ~25 us per launch, 6 ms over the 256 panels of
// a 16k factorisation.
static bool cooperative_launches()
{
    static int on = -1;
    if (on < 0)
    {
        const char *e = getenv("PVL_COOPERATIVE");
        on = e && e[0] == '1';
    }
    return on != 0;
}

int launch_grid_synced(const void *func, unsigned grid, unsigned block, void **args, size_t lds, hipStream_t s)
{
    int per_cu = 0, cus = 0;
    PVL_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, func, (int)block, lds));
    PVL_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, g_device < 0 ? 0 : g_device));
    if ((long)grid > (long)per_cu * cus)
    {
        set_error("grid of %u workgroups exceeds the %d x %d that can be resident at once", grid, per_cu, cus);
        return PVL_ERR_ARG;
    }
    if (cooperative_launches())
        PVL_HIP(hipLaunchCooperativeKernel(func, dim3(grid), dim3(block), args, (unsigned)lds, s));
    else
        PVL_HIP(hipLaunchKernel(func, dim3(grid), dim3(block), args, lds, s));
    return PVL_OK;
}

// Second library stream (default priority) for work that runs beside the library stream;
// ordered against it with events only.
static hipStream_t g_side = nullptr;
hipStream_t side_stream()
{
    if (g_side == nullptr && hipStreamCreateWithFlags(&g_side, hipStreamNonBlocking) != hipSuccess)
    {
        set_error("hipStreamCreateWithFlags(side stream) failed");
        g_side = nullptr;
    }
    return g_side;
}

// ---------------------------------------------------------------------------------------
// Device-to-device copies and fills as the library's own small kernels (at the library
// stream's priority, a few VGPRs): the runtime's blit kernels can wait for CU room while the
// side stream's kernel fills the device.
__global__ void __launch_bounds__(256) copy_words_kernel(uint32_t *__restrict__ dst, const uint32_t *__restrict__ src,
                                                         size_t n)
{
    main_stream_priority();
    for (size_t i = blockIdx.x * (size_t)256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256)
        dst[i] = src[i];
}
__global__ void __launch_bounds__(256) copy_bytes_kernel(uint8_t *__restrict__ dst, const uint8_t *__restrict__ src,
                                                         size_t n)
{
    main_stream_priority();
    for (size_t i = blockIdx.x * (size_t)256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256)
        dst[i] = src[i];
}
__global__ void __launch_bounds__(256) fill_words_kernel(uint32_t *__restrict__ dst, uint32_t v, size_t n)
{
    main_stream_priority();
    for (size_t i = blockIdx.x * (size_t)256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256)
        dst[i] = v;
}

static unsigned blocks_for(size_t n)
{
    const size_t b = (n + 255) / 256;
    return (unsigned)(b < 1024 ? (b ? b : 1) : 1024);
}

int device_copy(void *dst, const void *src, size_t bytes, hipStream_t s)
{
    if (bytes == 0)
        return PVL_OK;
    if (!s)
        s = g_stream;
    if (((uintptr_t)dst | (uintptr_t)src | bytes) % 4 == 0)
        copy_words_kernel<<<blocks_for(bytes / 4), 256, 0, s>>>((uint32_t *)dst, (const uint32_t *)src, bytes / 4);
    else
        copy_bytes_kernel<<<blocks_for(bytes), 256, 0, s>>>((uint8_t *)dst, (const uint8_t *)src, bytes);
    PVL_HIP(hipGetLastError());
    return PVL_OK;
}

int device_fill32(void *dst, uint32_t value, size_t bytes, hipStream_t s)
{
    if (bytes == 0)
        return PVL_OK;
    if (((uintptr_t)dst | bytes) % 4 != 0)
    {
        set_error("device_fill32: %zu bytes at %p are not whole aligned 32-bit words", bytes, dst);
        return PVL_ERR_ARG;
    }
    fill_words_kernel<<<blocks_for(bytes / 4), 256, 0, s ? s : g_stream>>>((uint32_t *)dst, value, bytes / 4);
    PVL_HIP(hipGetLastError());
    return PVL_OK;
}

// ---------------------------------------------------------------------------------------
// Scratch buffers: grown on demand, never shrunk; freed at process end by the runtime.
static std::vector<std::pair<void *, size_t>> g_scratch(16, {nullptr, 0});

void *scratch(int slot, size_t bytes)
{
    auto &s = g_scratch.at(slot);
    if (s.second >= bytes && s.first)
        return s.first;
    if (s.first)
    {
        // Make sure no queued kernel (on either library stream) still reads the old buffer.
        (void)hipStreamSynchronize(g_stream);
        if (g_side)
            (void)hipStreamSynchronize(g_side);
        (void)hipFree(s.first);
        s = {nullptr, 0};
    }
    void *p = nullptr;
    const size_t rounded = (bytes + 0xFFFFF) & ~size_t(0xFFFFF);
    if (hipMalloc(&p, rounded) != hipSuccess)
    {
        set_error("scratch allocation of %zu bytes failed", rounded);
        return nullptr;
    }
    s = {p, rounded};
    return p;
}

// ---------------------------------------------------------------------------------------
// Profiling: event pairs are recorded on the library stream without waiting; elapsed times
// are resolved when queried, so enabling it does not serialise the host with the GPU.
struct ProfEntry
{
    double ms = 0;
    long count = 0;
    double work = 0;
};
struct Pending
{
    std::string name;
    hipEvent_t ev0, ev1;
    double work;
};
static bool g_prof = false;
static std::string g_prof_only; // empty: every scope
static std::map<std::string, ProfEntry> g_prof_tab;
static std::vector<Pending> g_pending;
static std::vector<hipEvent_t> g_event_pool;

static hipEvent_t take_event()
{
    if (!g_event_pool.empty())
    {
        hipEvent_t e = g_event_pool.back();
        g_event_pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess)
        return nullptr;
    return e;
}

static void resolve_pending()
{
    for (auto &p : g_pending)
    {
        (void)hipEventSynchronize(p.ev1);
        float ms = 0;
        if (hipEventElapsedTime(&ms, p.ev0, p.ev1) != hipSuccess)
            ms = 0;
        auto &e = g_prof_tab[p.name];
        e.ms += ms;
        e.count += 1;
        e.work += p.work;
        g_event_pool.push_back(p.ev0);
        g_event_pool.push_back(p.ev1);
    }
    g_pending.clear();
}

ProfScope::ProfScope(const char *n, double work_units, hipStream_t s)
    : name(n), work(work_units), strm(s ? s : g_stream)
{
    if (!g_prof || (!g_prof_only.empty() && g_prof_only != n))
        return;
    ev0 = take_event();
    ev1 = take_event();
    if (!ev0 || !ev1 || hipEventRecord(ev0, strm) != hipSuccess)
    {
        if (ev0)
            g_event_pool.push_back(ev0);
        if (ev1)
            g_event_pool.push_back(ev1);
        ev0 = ev1 = nullptr;
    }
}

ProfScope::~ProfScope()
{
    if (!ev0)
        return;
    if (hipEventRecord(ev1, strm) != hipSuccess)
    {
        g_event_pool.push_back(ev0);
        g_event_pool.push_back(ev1);
        return;
    }
    g_pending.push_back({name, ev0, ev1, work});
    if (g_pending.size() > 4096)
        resolve_pending();
}

} // namespace pvl

using namespace pvl;

static hipEvent_t g_t0 = nullptr, g_t1 = nullptr;

extern "C" {

const char *pvl_last_error(void) { return g_err; }

int pvl_init(int device)
{
    PVL_API_LOCK;
    std::lock_guard<std::mutex> lk(g_mutex);
    if (g_device >= 0)
    {
        if (device == g_device)
            return PVL_OK;
        set_error("library already bound to device %d", g_device);
        return PVL_ERR_ARG;
    }
    return init_locked(device);
}

int pvl_device_count(int *count)
{
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess)
        n = 0;
    *count = n;
    return PVL_OK;
}

int pvl_synchronize(void)
{
    PVL_API_LOCK;
    if (int st = ensure_init())
        return st;
    PVL_HIP(hipStreamSynchronize(g_stream));
    if (g_side)
        PVL_HIP(hipStreamSynchronize(g_side));
    return take_async_error();
}

int pvl_shutdown(void)
{
    PVL_API_LOCK;
    std::lock_guard<std::mutex> lk(g_mutex);
    if (g_device < 0)
        return PVL_OK;
    (void)hipSetDevice(g_device);
    if (g_stream)
        (void)hipStreamSynchronize(g_stream);
    if (g_side)
        (void)hipStreamSynchronize(g_side);
    resolve_pending();
    for (hipEvent_t e : g_event_pool)
        (void)hipEventDestroy(e);
    g_event_pool.clear();
    g_prof_tab.clear();
    if (g_t0)
        (void)hipEventDestroy(g_t0);
    if (g_t1)
        (void)hipEventDestroy(g_t1);
    g_t0 = g_t1 = nullptr;
    for (auto &sl : g_scratch)
    {
        if (sl.first)
            (void)hipFree(sl.first);
        sl = {nullptr, 0};
    }
    lu_release_workspace();
    gmres_release_workspace();
    if (g_side)
        (void)hipStreamDestroy(g_side);
    if (g_stream)
        (void)hipStreamDestroy(g_stream);
    g_side = g_stream = nullptr;
    g_device = -1;
    return PVL_OK;
}

int pvl_malloc(void **dev_ptr, size_t bytes)
{
    PVL_API_LOCK;
    if (int st = ensure_init())
        return st;
    *dev_ptr = nullptr;
    if (bytes == 0)
        bytes = 256;
    PVL_HIP(hipMalloc(dev_ptr, bytes));
    return PVL_OK;
}

int pvl_free(void *dev_ptr)
{
    PVL_API_LOCK;
    if (!dev_ptr)
        return PVL_OK;
    if (int st = ensure_init())
        return st;
    PVL_HIP(hipStreamSynchronize(g_stream));
    PVL_HIP(hipFree(dev_ptr));
    return PVL_OK;
}

int pvl_memcpy_h2d(void *dst, const void *src, size_t bytes)
{
    PVL_API_LOCK;
    if (int st = ensure_init())
        return st;
    if (bytes == 0)
        return PVL_OK;
    PVL_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, g_stream));
    PVL_HIP(hipStreamSynchronize(g_stream));
    return PVL_OK;
}

int pvl_memcpy_d2h(void *dst, const void *src, size_t bytes)
{
    PVL_API_LOCK;
    if (int st = ensure_init())
        return st;
    if (bytes == 0)
        return PVL_OK;
    PVL_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, g_stream));
    PVL_HIP(hipStreamSynchronize(g_stream));
    return take_async_error();
}

int pvl_memcpy_d2d(void *dst, const void *src, size_t bytes)
{
    PVL_API_LOCK;
    if (int st = ensure_init())
        return st;
    return device_copy(dst, src, bytes, g_stream);
}

int pvl_memset_zero(void *dst, size_t bytes)
{
    PVL_API_LOCK;
    if (int st = ensure_init())
        return st;
    if (bytes == 0)
        return PVL_OK;
    if (((uintptr_t)dst | bytes) % 4 == 0)
        return device_fill32(dst, 0u, bytes, g_stream);
    PVL_HIP(hipMemsetAsync(dst, 0, bytes, g_stream));
    return PVL_OK;
}

int pvl_prof_filter(const char *name)
{
    PVL_API_LOCK;
    g_prof_only = name ? name : "";
    return PVL_OK;
}

int pvl_prof_enable(int on)
{
    PVL_API_LOCK;
    if (int st = ensure_init())
        return st;
    g_prof = on != 0;
    return PVL_OK;
}

int pvl_prof_reset(void)
{
    PVL_API_LOCK;
    resolve_pending();
    g_prof_tab.clear();
    return PVL_OK;
}

int pvl_prof_query(const char *name, double *total_ms, long *count)
{
    PVL_API_LOCK;
    resolve_pending();
    auto it = g_prof_tab.find(name);
    if (it == g_prof_tab.end())
    {
        *total_ms = 0;
        *count = 0;
        return PVL_OK;
    }
    *total_ms = it->second.ms;
    *count = it->second.count;
    return PVL_OK;
}

int pvl_prof_work(const char *name, double *work)
{
    PVL_API_LOCK;
    resolve_pending();
    auto it = g_prof_tab.find(name);
    *work = it == g_prof_tab.end() ? 0.0 : it->second.work;
    return PVL_OK;
}

int pvl_timer_start(void)
{
    PVL_API_LOCK;
    if (int st = ensure_init())
        return st;
    if (!g_t0)
    {
        PVL_HIP(hipEventCreate(&g_t0));
        PVL_HIP(hipEventCreate(&g_t1));
    }
    PVL_HIP(hipEventRecord(g_t0, g_stream));
    return PVL_OK;
}

int pvl_timer_stop(double *ms)
{
    PVL_API_LOCK;
    if (!g_t0)
    {
        set_error("pvl_timer_stop without pvl_timer_start");
        return PVL_ERR_ARG;
    }
    PVL_HIP(hipEventRecord(g_t1, g_stream));
    PVL_HIP(hipEventSynchronize(g_t1));
    float f = 0;
    PVL_HIP(hipEventElapsedTime(&f, g_t0, g_t1));
    *ms = f;
    return PVL_OK;
}

} // extern "C"
