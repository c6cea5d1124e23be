// Random-media synthesis on the GPU, the method of rmedia__3dgen_ani /
// rmedia__2dgen_ani (src/shared/m_rmedia.f90:40-173) as
// openswpc_amd/rmedia.py states it: noise -> FFT -> spectral whitening ->
// multiply by sqrt(PSDF) -> zero DC -> inverse FFT -> demean -> scale to
// rms epsilon.  The noise comes from the caller (the host generator's
// stream), so a realisation equals the host one up to FP32 FFT rounding.
// The FFTs are hipFFT's; the filter, the reductions (FP64 vector atomics)
// and the scaling are kernels here.
#include "swpc_common.h"

#include <hipfft/hipfft.h>

#include <cmath>
#include <vector>

namespace swpc {

struct RmSpec {
    int nd;
    int n[3];                 // shape, last axis fastest
    double d[3], a[3];        // grid widths, correlation lengths (km)
    double epsil, kappa;
    int ptype;                // 0 von Karman, 1 Gauss, 2 exponential
    double amp;               // N * prod(d): the spectrum's scale before sqrt(PSDF)
    double c;                 // the PSDF's constant factor
    long long nc;             // complex coefficients of the R2C transform
    int nh;                   // n[last] / 2 + 1
};

// numpy fftfreq / rfftfreq index -> angular wavenumber
__device__ inline double rm_k(int i, int n, double d, bool half)
{
    const int f = half ? i : (i <= (n - 1) / 2 ? i : i - n);
    return 2.0 * M_PI * (double)f / ((double)n * d);
}

__global__ void __launch_bounds__(256) k_rm_filter(hipfftComplex *__restrict__ B, const RmSpec s)
{
    const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= s.nc) return;
    const int last = s.nd - 1;
    long long r = q;
    int idx[3] = {0, 0, 0};
    idx[last] = (int)(r % s.nh);
    r /= s.nh;
    for (int ax = last - 1; ax >= 0; ax--) {
        idx[ax] = (int)(r % s.n[ax]);
        r /= s.n[ax];
    }
    double m2a2 = 0.0;
    for (int ax = 0; ax <= last; ax++) {
        const double k = rm_k(idx[ax], s.n[ax], s.d[ax], ax == last) * s.a[ax];
        m2a2 += k * k;
    }
    double p;   // rmedia.py psdf_vonkarman / psdf_gauss / psdf_exp
    if (s.ptype == 0) p = s.c / pow(1.0 + m2a2, s.kappa + (s.nd == 3 ? 1.5 : 1.0));
    else if (s.ptype == 1) p = s.c * exp(-m2a2 / 4.0);
    else p = s.c / pow(1.0 + m2a2, s.nd == 3 ? 2.0 : 1.5);
    hipfftComplex v = B[q];
    double mag = sqrt((double)v.x * v.x + (double)v.y * v.y);
    if (mag == 0.0) mag = 1.0;
    const double g = q == 0 ? 0.0 : sqrt(p * s.amp) / mag;
    v.x = (float)((double)v.x * g);
    v.y = (float)((double)v.y * g);
    B[q] = v;
}

__global__ void __launch_bounds__(256) k_rm_sum(const float *__restrict__ f, long long n, double mean,
                                                 int square, double *__restrict__ out)
{
    double s = 0.0;
    for (long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (long long)gridDim.x * blockDim.x) {
        const double v = (double)f[q] - mean;
        s += square ? v * v : v;
    }
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
    if ((threadIdx.x & 63) == 0) atomicAdd(out, s);
}

__global__ void __launch_bounds__(256) k_rm_scale(float *__restrict__ f, long long n, double mean, double g)
{
    const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q < n) f[q] = (float)(((double)f[q] - mean) * g);
}

__global__ void k_rm_zero(double *p, int n)
{
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q < n) p[q] = 0.0;
}

}  // namespace swpc

using namespace swpc;

extern "C" int swpc_rmedia_generate(int nd, const int *shape, const double *d, const double *a, double epsil,
                                    double kappa, int ptype, const float *noise, float *out)
{
    clear_error();
    if ((nd != 2 && nd != 3) || !shape || !d || !a || !noise || !out) return fail(SWPC_EINVAL, "bad argument");
    if (ptype < 0 || ptype > 2) return fail(SWPC_EINVAL, "unknown PSDF type %d", ptype);
    if (!(epsil > 0) || !(kappa > 0 || ptype != 0)) return fail(SWPC_EINVAL, "epsilon and kappa must be positive");
    RmSpec s;
    s.nd = nd;
    long long N = 1;
    double a3 = 1.0, dv = 1.0;
    for (int q = 0; q < nd; q++) {
        if (shape[q] < 2 || !(d[q] > 0) || !(a[q] > 0)) return fail(SWPC_EINVAL, "axis %d: shape, d and a must be positive", q);
        s.n[q] = shape[q];
        s.d[q] = d[q];
        s.a[q] = a[q];
        N *= shape[q];
        a3 *= a[q];
        dv *= d[q];
    }
    if (N >= (1LL << 31)) return fail(SWPC_EINVAL, "random field too large for one transform");
    s.epsil = epsil;
    s.kappa = kappa;
    s.ptype = ptype;
    s.amp = (double)N * dv;
    const double e2 = epsil * epsil;
    if (ptype == 0)
        s.c = nd == 3 ? 8 * std::pow(M_PI, 1.5) * e2 * a3 * std::tgamma(kappa + 1.5) / std::tgamma(kappa)
                      : 4 * M_PI * kappa * e2 * a3;
    else if (ptype == 1) s.c = e2 * a3 * std::pow(std::sqrt(M_PI), nd);
    else s.c = nd == 3 ? 8 * M_PI * e2 * a3 : 2 * M_PI * e2 * a3;
    s.nh = shape[nd - 1] / 2 + 1;
    s.nc = N / shape[nd - 1] * s.nh;

    hipStream_t st = nullptr;
    SWPC_HIP(hipStreamCreate(&st));
    float *f = nullptr;
    hipfftComplex *B = nullptr;
    double *red = nullptr;
    hipfftHandle p1 = 0, p2 = 0;
    bool h1 = false, h2 = false;
    int rc = SWPC_OK;
    auto bail = [&](int code, const char *what) {
        if (rc == SWPC_OK) rc = fail(code, "random media: %s", what);
    };
    if (hipMalloc(&f, (size_t)N * 4) != hipSuccess || hipMalloc(&B, (size_t)s.nc * sizeof(hipfftComplex)) != hipSuccess ||
        hipMalloc(&red, 16) != hipSuccess)
        bail(SWPC_ENOMEM, "device allocation failed");
    if (rc == SWPC_OK && hipMemcpy(f, noise, (size_t)N * 4, hipMemcpyHostToDevice) != hipSuccess) bail(SWPC_EHIP, "upload failed");
    if (rc == SWPC_OK) {
        const hipfftResult r1 = nd == 3 ? hipfftPlan3d(&p1, shape[0], shape[1], shape[2], HIPFFT_R2C)
                                        : hipfftPlan2d(&p1, shape[0], shape[1], HIPFFT_R2C);
        const hipfftResult r2 = nd == 3 ? hipfftPlan3d(&p2, shape[0], shape[1], shape[2], HIPFFT_C2R)
                                        : hipfftPlan2d(&p2, shape[0], shape[1], HIPFFT_C2R);
        h1 = r1 == HIPFFT_SUCCESS;
        h2 = r2 == HIPFFT_SUCCESS;
        if (!h1 || !h2 || hipfftSetStream(p1, st) != HIPFFT_SUCCESS || hipfftSetStream(p2, st) != HIPFFT_SUCCESS)
            bail(SWPC_EHIP, "hipFFT plan creation failed");
    }
    if (rc == SWPC_OK && hipfftExecR2C(p1, f, B) != HIPFFT_SUCCESS) bail(SWPC_EHIP, "forward FFT failed");
    if (rc == SWPC_OK) {
        hipLaunchKernelGGL(k_rm_filter, dim3((unsigned)((s.nc + 255) / 256)), dim3(256), 0, st, B, s);
        if (hipGetLastError() != hipSuccess) bail(SWPC_EHIP, "filter launch failed");
    }
    if (rc == SWPC_OK && hipfftExecC2R(p2, B, f) != HIPFFT_SUCCESS) bail(SWPC_EHIP, "inverse FFT failed");
    double h[2] = {0.0, 0.0};
    const unsigned nb = (unsigned)std::min<long long>((N + 255) / 256, 4096);
    // the mean, then the rms about it, in FP64 (rmedia.py: f -= mean; f *= eps / rms)
    for (int pass = 0; pass < 2 && rc == SWPC_OK; pass++) {
        const double mean = pass == 0 ? 0.0 : h[0] / (double)N;
        hipLaunchKernelGGL(k_rm_zero, dim3(1), dim3(64), 0, st, red, 2);
        hipLaunchKernelGGL(k_rm_sum, dim3(nb), dim3(256), 0, st, (const float *)f, N, mean, pass, red + pass);
        if (hipGetLastError() != hipSuccess || hipMemcpyAsync(h + pass, red + pass, 8, hipMemcpyDeviceToHost, st) != hipSuccess ||
            hipStreamSynchronize(st) != hipSuccess)
            bail(SWPC_EHIP, "reduction failed");
    }
    if (rc == SWPC_OK) {
        const double mean = h[0] / (double)N, rms = std::sqrt(h[1] / (double)N);
        if (!(rms > 0)) bail(SWPC_EINVAL, "degenerate field (zero rms)");
        if (rc == SWPC_OK) {
            hipLaunchKernelGGL(k_rm_scale, dim3((unsigned)((N + 255) / 256)), dim3(256), 0, st, f, N, mean, epsil / rms);
            if (hipGetLastError() != hipSuccess || hipMemcpyAsync(out, f, (size_t)N * 4, hipMemcpyDeviceToHost, st) != hipSuccess ||
                hipStreamSynchronize(st) != hipSuccess)
                bail(SWPC_EHIP, "download failed");
        }
    }
    if (h1) (void)hipfftDestroy(p1);
    if (h2) (void)hipfftDestroy(p2);
    if (f) (void)hipFree(f);
    if (B) (void)hipFree(B);
    if (red) (void)hipFree(red);
    (void)hipStreamSynchronize(st);
    (void)hipStreamDestroy(st);
    return rc;
}
