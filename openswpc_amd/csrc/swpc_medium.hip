// medium__setup on the GPU (src/swpc_3d/m_medium.f90:36-467).
//
// The host form (openswpc_amd/medium.py medium_setup3) runs O(N) numpy
// passes over the memory box; at 1024 x 1024 x 512 they take about a
// minute, as long as the run itself.  Here each pass is one kernel over
// the device-layout arrays:
//   k_med_ingest   transpose of the model's (k,i,j) box + homogenisation
//   k_med_relax    taup, taus from Q; unrelaxed lam, mu at fq_ref
//   k_med_detect   kfs, kob per column; k_med_window their 6 x 6 extremes
//   k_med_average  buoyancy and harmonic shear means
//   k_med_minmax   vmin, vmax below the free surface
//   k_med_planes   planes on which an array is constant (K3::cpl)
// Products and sums follow the host form's order with no contraction,
// so everything but the relaxation factor is bit-identical to it; the
// factor 1/Re(1/sqrt(w)) = |w| / Re sqrt(w) is formed in FP64 and rounded
// once (the host form rounds through complex FP32 steps).
#include "swpc_medium.h"

#include <cfloat>
#include <climits>
#include <cstring>

#pragma clang fp contract(off)

namespace swpc {

__device__ __forceinline__ int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

// one 32 (k) x 32 (i) tile of plane j per block, through LDS
__global__ void __launch_bounds__(256) k_med_ingest(const MedSetup m, const float *__restrict__ src,
                                                    float *__restrict__ dst)
{
    __shared__ float t[32][33];
    const int nib = m.nx + 6, nkb = m.nz + 6;
    const int k0 = (int)blockIdx.x * 32 - 2, i0 = (int)blockIdx.y * 32 - 2, j = (int)blockIdx.z - 2;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
    const int sj = clampi(j, m.na + 1, m.ny - m.na);
#pragma unroll
    for (int r = 0; r < 4; r++) {
        const int i = i0 + ty + 8 * r, k = k0 + tx;
        if (i <= m.nx + 3 && k <= m.nz + 3) {
            const int si = clampi(i, m.na + 1, m.nx - m.na), sk = k < m.nz - m.na ? k : m.nz - m.na;
            t[ty + 8 * r][tx] = src[((size_t)(sj + 2) * nib + (size_t)(si + 2)) * nkb + (size_t)(sk + 2)];
        }
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < 4; r++) {
        const int k = k0 + ty + 8 * r, i = i0 + tx;
        if (i <= m.nx + 3 && k <= m.nz + 3) dst[(size_t)at3(m.px, m.sxy, i, j, k)] = t[tx][ty + 8 * r];
    }
}

// 1 / Re(1 / sqrt(w)) for w = 1 - cc tau, with w's two parts formed in
// FP32 as the complex FP32 product and difference give them
__device__ __forceinline__ float med_chi(float cr, float ci, float tau)
{
    const float wr = 1.0f - cr * tau;
    const float wi = -(ci * tau);
    const double aw = sqrt((double)wr * (double)wr + (double)wi * (double)wi);
    const double x = sqrt(0.5 * (aw + (double)wr));
    return (float)(aw / x);
}

__global__ void __launch_bounds__(256) k_med_relax(const MedSetup m)
{
    const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long vol = (long long)m.sxy * (m.nz + 6);
    if (q >= vol) return;
    const int i = (int)(q % m.px) - IO;
    if (i < -2 || i > m.nx + 3) return;   // row padding stays zero
    const float taup = m.zeta3 / m.taup[q];
    const float taus = m.zeta3 / m.taus[q];
    m.taup[q] = taup;
    m.taus[q] = taus;
    const float mu = m.mu[q], lam = m.lam[q];
    const float chi_mu = med_chi(m.cc_re, m.cc_im, taus);
    const float chi_lam = med_chi(m.cc_re, m.cc_im, taup);
    const float rho_alpha2 = lam + 2.0f * mu;
    const float mu_n = mu / (chi_mu * chi_mu);
    m.mu[q] = mu_n;
    m.lam[q] = rho_alpha2 / (chi_lam * chi_lam) - 2.0f * mu_n;
}

// last k in [1, nz-1] with |x(k)| < eps < |x(k+1)|, 0 if none
__device__ __forceinline__ int med_last(const float *__restrict__ x, int px, int sxy, int i, int j, int nz)
{
    for (int k = nz - 1; k >= 1; k--)
        if (fabsf(x[at3(px, sxy, i, j, k)]) < FLT_EPSILON && fabsf(x[at3(px, sxy, i, j, k + 1)]) > FLT_EPSILON)
            return k;
    return 0;
}

__global__ void __launch_bounds__(256) k_med_detect(const MedSetup m, int *__restrict__ kfs, int *__restrict__ kob)
{
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    const int w = m.nx + 3;   // columns i in [0, nx+2]
    if (q >= w * (m.ny + 3)) return;
    const int i = q % w, j = q / w;
    const size_t h = (size_t)(j + 2) * (m.nx + 6) + (i + 2);
    kfs[h] = med_last(m.lam, m.px, m.sxy, i, j, m.nz);
    kob[h] = med_last(m.mu, m.px, m.sxy, i, j, m.nz);
}

__global__ void __launch_bounds__(256) k_med_window(const MedSetup m, const int *__restrict__ kfs,
                                                    const int *__restrict__ kob, int4 *__restrict__ surf,
                                                    int *__restrict__ kobc)
{
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= m.nx * m.ny) return;
    const int i = q % m.nx + 1, j = q / m.nx + 1, nib = m.nx + 6;
    int fmn = INT_MAX, fmx = INT_MIN, omn = INT_MAX, omx = INT_MIN;
    for (int b = j - 2; b <= j + 3; b++)
        for (int a = i - 2; a <= i + 3; a++) {
            const size_t h = (size_t)(b + 2) * nib + (a + 2);
            fmn = min(fmn, kfs[h]); fmx = max(fmx, kfs[h]);
            omn = min(omn, kob[h]); omx = max(omx, kob[h]);
        }
    surf[q] = make_int4(max(fmn - 2, 1), min(fmx + 2, m.nz), max(omn - 2, 1), min(omx + 2, m.nz));
    kobc[q] = kob[(size_t)(j + 2) * nib + (i + 2)];
}

__global__ void __launch_bounds__(256) k_med_average(const MedSetup m)
{
    const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long vol = (long long)m.sxy * (m.nz + 6);
    if (q >= vol) return;
    const int i = (int)(q % m.px) - IO, j = (int)((q / m.px) % (m.sxy / m.px)) - JO, k = (int)(q / m.sxy) - KO;
    if (i < -2 || i > m.nx + 3) return;
    if (i < 1 || i > m.nx || j < 1 || j > m.ny || k < 1 || k > m.nz) {
        m.bx[q] = m.by[q] = m.bz[q] = 0.0f;
        m.muyz[q] = m.muxz[q] = m.muxy[q] = 0.0f;
        return;
    }
    const int px = m.px, sxy = m.sxy;
    const float *__restrict__ r = m.rho;
    const float r0 = r[q];
    m.bx[q] = 2.0f / (r0 + r[q + 1]);
    m.by[q] = 2.0f / (r0 + r[q + px]);
    m.bz[q] = 2.0f / (r0 + r[q + sxy]);
    const float *__restrict__ u = m.mu;
    const float nnn = u[q], pnn = u[q + sxy], npn = u[q + 1], ppn = u[q + 1 + sxy];
    const float nnp = u[q + px], npp = u[q + 1 + px], pnp = u[q + px + sxy];
    const float e = FLT_EPSILON;
    m.muxz[q] = 4.0f * nnn * pnn * npn * ppn / (nnn * pnn * npn + nnn * pnn * ppn + nnn * npn * ppn + pnn * npn * ppn + e);
    m.muxy[q] = 4.0f * nnn * npn * nnp * npp / (nnn * npn * nnp + nnn * npn * npp + nnn * nnp * npp + npn * nnp * npp + e);
    m.muyz[q] = 4.0f * nnn * pnn * nnp * pnp / (nnn * pnn * nnp + nnn * pnn * pnp + nnn * nnp * pnp + pnn * nnp * pnp + e);
}

// out[0]: bits of min vs (vs >= eps), out[1]: bits of max vp, out[2]: any
// cell below the surface (vector atomics on non-negative float bits)
__global__ void __launch_bounds__(256) k_med_minmax(const MedSetup m, const int *__restrict__ kfs,
                                                    unsigned *__restrict__ out)
{
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= m.nx * m.ny) return;
    const int i = q % m.nx + 1, j = q / m.nx + 1;
    const int k0 = kfs[(size_t)(j + 2) * (m.nx + 6) + (i + 2)] + 1;
    float vmx = -1.0f, vmn = INFINITY;
    for (int k = k0 < 1 ? 1 : k0; k <= m.nz; k++) {
        const size_t c = at3(m.px, m.sxy, i, j, k);
        const float lam = m.lam[c], mu = m.mu[c], rho = m.rho[c];
        const float vp = sqrtf((lam + 2.0f * mu) / rho);
        const float vs = sqrtf(mu / rho);
        vmx = fmaxf(vmx, vp);
        if (vs >= FLT_EPSILON) vmn = fminf(vmn, vs);
    }
    if (vmx >= 0.0f) {
        atomicMax(out + 1, __float_as_uint(vmx));
        atomicOr(out + 2, 1u);
    }
    if (vmn < INFINITY) atomicMin(out, __float_as_uint(vmn));
}

// one block row of plane k per block: clear the plane's flag on any
// cell whose bits differ from cell (1, 1, k)
__global__ void __launch_bounds__(256) k_med_planes(const MedSetup m, const float *__restrict__ a,
                                                    int *__restrict__ flags)
{
    const int k = blockIdx.y + 1;
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= m.nx * m.ny) return;
    const int i = q % m.nx + 1, j = q / m.nx + 1;
    const float v0 = a[at3(m.px, m.sxy, 1, 1, k)];
    const float v = a[at3(m.px, m.sxy, i, j, k)];
    if (__float_as_uint(v) != __float_as_uint(v0) || v0 != v0) atomicAnd(flags + k + KO, 0);
}

__global__ void k_med_tab(const MedSetup m, const float *__restrict__ a, int *__restrict__ flags,
                          float *__restrict__ tab, int pz)
{
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= pz) return;
    const int k = q - KO;
    const bool in = k >= 1 && k <= m.nz;
    if (!in) flags[q] = 0;
    tab[q] = in && flags[q] ? a[at3(m.px, m.sxy, 1, 1, k)] : __int_as_float(0x7fc00000);
}

__global__ void k_med_fill_i32(int *p, int n, int v)
{
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q < n) p[q] = v;
}

static unsigned blocks(long long n, int b) { return (unsigned)((n + b - 1) / b); }

int med_ingest(const MedSetup &m, const float *src, float *dst, hipStream_t s)
{
    const dim3 grid(blocks(m.nz + 6, 32), blocks(m.nx + 6, 32), m.ny + 6);
    hipLaunchKernelGGL(k_med_ingest, grid, dim3(256), 0, s, m, src, dst);
    SWPC_CHECK_LAUNCH();
    return SWPC_OK;
}

int med_relax(const MedSetup &m, hipStream_t s)
{
    const long long vol = (long long)m.sxy * (m.nz + 6);
    hipLaunchKernelGGL(k_med_relax, dim3(blocks(vol, 256)), dim3(256), 0, s, m);
    SWPC_CHECK_LAUNCH();
    return SWPC_OK;
}

int med_surface(const MedSetup &m, int *kfs, int *kob, int4 *surf, int *kobc, hipStream_t s)
{
    const int nmap = (m.nx + 6) * (m.ny + 6);
    hipLaunchKernelGGL(k_med_fill_i32, dim3(blocks(nmap, 256)), dim3(256), 0, s, kfs, nmap, 0);
    hipLaunchKernelGGL(k_med_fill_i32, dim3(blocks(nmap, 256)), dim3(256), 0, s, kob, nmap, 0);
    hipLaunchKernelGGL(k_med_detect, dim3(blocks((long long)(m.nx + 3) * (m.ny + 3), 256)), dim3(256), 0, s, m, kfs,
                       kob);
    hipLaunchKernelGGL(k_med_window, dim3(blocks((long long)m.nx * m.ny, 256)), dim3(256), 0, s, m, kfs, kob, surf,
                       kobc);
    SWPC_CHECK_LAUNCH();
    return SWPC_OK;
}

int med_average(const MedSetup &m, hipStream_t s)
{
    const long long vol = (long long)m.sxy * (m.nz + 6);
    hipLaunchKernelGGL(k_med_average, dim3(blocks(vol, 256)), dim3(256), 0, s, m);
    SWPC_CHECK_LAUNCH();
    return SWPC_OK;
}

int med_minmax(const MedSetup &m, const int *kfs, float *out, hipStream_t s)
{
    unsigned *d = nullptr;
    SWPC_HIP(hipMalloc(&d, 16));
    const unsigned init[4] = {0x7f800000u, 0u, 0u, 0u};
    hipError_t e = hipMemcpyAsync(d, init, 16, hipMemcpyHostToDevice, s);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_med_minmax, dim3(blocks((long long)m.nx * m.ny, 256)), dim3(256), 0, s, m, kfs, d);
        e = hipGetLastError();
    }
    unsigned h[4] = {0, 0, 0, 0};
    if (e == hipSuccess) e = hipMemcpyAsync(h, d, 16, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    (void)hipFree(d);
    if (e != hipSuccess) return fail(SWPC_EHIP, "velocity min/max: %s", hipGetErrorString(e));
    float vmin, vmax;
    memcpy(&vmin, &h[0], 4);
    memcpy(&vmax, &h[1], 4);
    out[0] = h[0] == 0x7f800000u ? 1e30f : vmin;
    out[1] = h[2] ? vmax : -1.0f;
    return SWPC_OK;
}

int med_planes(const MedSetup &m, const float *a, int *flags, float *tab, int pz, hipStream_t s)
{
    hipLaunchKernelGGL(k_med_fill_i32, dim3(blocks(pz, 256)), dim3(256), 0, s, flags, pz, 1);
    hipLaunchKernelGGL(k_med_planes, dim3(blocks((long long)m.nx * m.ny, 256), m.nz), dim3(256), 0, s, m, a, flags);
    hipLaunchKernelGGL(k_med_tab, dim3(blocks(pz, 256)), dim3(256), 0, s, m, a, flags, tab, pz);
    SWPC_CHECK_LAUNCH();
    return SWPC_OK;
}

}  // namespace swpc
