// Error reporting, allocation and event timing shared by the 3-D and
// P-SV contexts.
#include "swpc_common.h"

#include <algorithm>

namespace swpc {

static thread_local std::string g_err;

int fail(int code, const char *fmt, ...)
{
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

void clear_error() { g_err.clear(); }

// Buffers of 16 MiB and more start 2304 * (n mod 32) bytes into their
// allocation (n counts such allocations), so that equal indices of the
// many arrays a kernel streams together do not all fall on the same DRAM
// channel phase: 0.5-1 % on the 512 x 512 x 256 step.
constexpr size_t STAGGER = 2304;
static unsigned g_stagger = 0;

int dalloc(DBuf &b, size_t bytes, long long &total, size_t lead)
{
    dfree(b);
    if (bytes == 0) return SWPC_OK;
    const size_t off = (bytes >= ((size_t)1 << 24) ? (size_t)(g_stagger++ % 32) * STAGGER : 0) + lead;
    void *base = nullptr;
    hipError_t e = hipMalloc(&base, bytes + off);
    if (e != hipSuccess) {
        b.p = nullptr;
        return fail(SWPC_ENOMEM, "hipMalloc(%zu) failed: %s", bytes, hipGetErrorString(e));
    }
    e = hipMemset(base, 0, bytes + off);
    b.base = base;
    b.p = (char *)base + off;
    if (e != hipSuccess) return fail(SWPC_EHIP, "hipMemset failed: %s", hipGetErrorString(e));
    b.bytes = bytes;
    total += (long long)(bytes + off);
    return SWPC_OK;
}

void dfree(DBuf &b)
{
    if (b.base) (void)hipFree(b.base);
    b.base = nullptr;
    b.p = nullptr;
    b.bytes = 0;
}

void Prof::reset()
{
    for (auto &v : ev) {
        for (auto e : v) (void)hipEventDestroy(e);
        v.clear();
    }
}

int prof_begin(Prof &p, int cls, hipStream_t s, hipEvent_t *out)
{
    *out = nullptr;
    if (!p.on) return SWPC_OK;
    hipEvent_t e;
    SWPC_HIP(hipEventCreate(&e));
    SWPC_HIP(hipEventRecord(e, s));
    p.ev[cls].push_back(e);
    *out = e;
    return SWPC_OK;
}

int prof_end(Prof &p, int cls, hipStream_t s, hipEvent_t start)
{
    if (!p.on || !start) return SWPC_OK;
    hipEvent_t e;
    SWPC_HIP(hipEventCreate(&e));
    SWPC_HIP(hipEventRecord(e, s));
    p.ev[cls].push_back(e);
    return SWPC_OK;
}

int prof_read(Prof &p, double *ms5, long long *count5)
{
    for (int c = 0; c < 5; c++) {
        double sum = 0;
        auto &v = p.ev[c];
        for (size_t n = 0; n + 1 < v.size(); n += 2) {
            SWPC_HIP(hipEventSynchronize(v[n + 1]));
            float ms = 0;
            SWPC_HIP(hipEventElapsedTime(&ms, v[n], v[n + 1]));
            sum += ms;
        }
        if (ms5) ms5[c] = sum;
        if (count5) count5[c] = (long long)(v.size() / 2);
    }
    return SWPC_OK;
}

}  // namespace swpc

extern "C" {

const char *swpc_last_error(void) { return swpc::g_err.c_str(); }

int swpc_version(void) { return 10000; }

int swpc_device_count(void)
{
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Stream-copy probe: the measured HBM bandwidth that the stencil kernels are
// compared against (and the scale for rocprof FETCH/WRITE counters).
// ---------------------------------------------------------------------------
// One 16-byte non-temporal load and store per thread and pass, the access
// form of the stencil kernels' streamed arrays.
typedef float f4c __attribute__((ext_vector_type(4)));

__global__ void __launch_bounds__(256) k_stream_copy(const f4c *__restrict__ a, f4c *__restrict__ b, size_t n)
{
    size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    const size_t stride = (size_t)gridDim.x * 256;
    for (; i < n; i += stride) __builtin_nontemporal_store(__builtin_nontemporal_load(a + i), b + i);
}

extern "C" int swpc_membw(long long bytes, int iters, double *gbs)
{
    swpc::clear_error();
    if (bytes < 1024 || iters < 1 || !gbs) return swpc::fail(SWPC_EINVAL, "bad argument");
    size_t n = (size_t)bytes / 16;
    void *a = nullptr, *b = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = SWPC_OK;
    if (hipMalloc(&a, n * 16) != hipSuccess || hipMalloc(&b, n * 16) != hipSuccess) {
        rc = swpc::fail(SWPC_ENOMEM, "hipMalloc failed");
    } else {
        (void)hipMemset(a, 0, n * 16);
        int blocks = (int)std::min<size_t>((n + 255) / 256, (size_t)1 << 20);
        hipLaunchKernelGGL(k_stream_copy, dim3(blocks), dim3(256), 0, 0, (const f4c *)a, (f4c *)b, n);
        (void)hipEventCreate(&e0);
        (void)hipEventCreate(&e1);
        (void)hipEventRecord(e0, 0);
        for (int q = 0; q < iters; q++)
            hipLaunchKernelGGL(k_stream_copy, dim3(blocks), dim3(256), 0, 0, (const f4c *)a, (f4c *)b, n);
        (void)hipEventRecord(e1, 0);
        if (hipEventSynchronize(e1) != hipSuccess) {
            rc = swpc::fail(SWPC_EHIP, "copy probe failed");
        } else {
            float ms = 0;
            (void)hipEventElapsedTime(&ms, e0, e1);
            *gbs = 2.0 * (double)n * 16 * iters / (ms * 1e-3) / 1e9;
        }
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (a) (void)hipFree(a);
    if (b) (void)hipFree(b);
    return rc;
}
