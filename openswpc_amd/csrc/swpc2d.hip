// 2-D P-SV viscoelastic velocity-stress core (swpc_psv) for MI355X.
//
//   k2_stress_int  interior stress + memory variables, normal and shear
//                  parts fused (src/swpc_psv/m_kernel.f90:139-306); the
//                  harmonic mu_xz is formed in the kernel as in the reference
//   k2_vel_int     interior velocity with in-kernel buoyancy
//                  (src/swpc_psv/m_kernel.f90:76-137)
//   k2_stress_pml / k2_vel_pml   ADE-CFS PML (src/swpc_psv/m_absorb_p.f90)
//   k2_stressglut / k2_bodyforce / k2_receivers / k2_vmax
//
// One thread per cell in an x-fastest (k, i) layout: a P-SV grid has too
// few columns to stream z per thread and still fill 256 CUs.
#include "swpc_common.h"
#include "swpc_batch.h"

#include <algorithm>
#include <memory>

namespace swpc {

// Every P-SV kernel rounds operation by operation in the reference's
// expression order (no contraction into FMA), so that a run matches the
// reference loops evaluated in IEEE single precision bit for bit.  The
// P-SV step is bound by launch gaps and its state sits in the Infinity
// Cache: the extra multiplies and adds cost nothing measurable.
#pragma clang fp contract(off)

constexpr int NM2 = 3;

struct Box2 {
    int i0, k0, ni, nk;
    unsigned start;
};

struct Boxes2 {
    Box2 b[3];
    int n;
    unsigned total;
};

template <typename T>
struct K2 {
    int px, nx, nz, na;
    T *Vx, *Vz, *Sxx, *Szz, *Sxz;
    T *R[9];    // Rxx(m) Rzz(m) Rxz(m)
    T *A[8];    // axVx azVx axVz azVz axSxx azSxz axSxz azSzz
    const float *rho, *lam, *mu, *taup, *taus;
    const int4 *surf;   // per column i in [1, nx]
    const T *gx, *gz;   // [8][n]
    T e4[2][2], e2[2][2];
    T r2[2];
    T c1[NM2], c2[NM2], d1[NM2], d2;
    T dt;
    int ti0;
    Boxes2 pml;
};

// Evaluated in IEEE single precision operation by operation, as the
// reference's real(SP) expression: no contraction into FMA.
__device__ inline float harm4(float nnn, float pnn, float npn, float ppn)
{
#pragma clang fp contract(off)
    const float epsl = 1.1920928955078125e-07f;  // epsilon(1.0)
    return 4 * nnn * pnn * npn * ppn / (nnn * pnn * npn + nnn * pnn * ppn + nnn * npn * ppn + pnn * npn * ppn + epsl);
}

template <typename T>
__device__ __forceinline__ void stress_int_cell(const K2<T> &a, const int i, const int k)
{
    const int px = a.px;
    const int c = at2(px, i, k);
    const int4 fs = a.surf[i - 1];
    const int o1 = (k - fs.x) * (fs.y - k), o2 = (k - fs.z) * (fs.w - k);
    const bool lo = max(o1, o2) >= 0;
    const T e0x = lo ? a.e2[0][0] : a.e4[0][0], e1x = lo ? a.e2[0][1] : a.e4[0][1];
    const T e0z = lo ? a.e2[1][0] : a.e4[1][0], e1z = lo ? a.e2[1][1] : a.e4[1][1];
    const T *__restrict__ Vx = a.Vx;
    const T *__restrict__ Vz = a.Vz;
    const T dt = a.dt;
    const T vx0 = Vx[c], vz0 = Vz[c];

    // normal (m_kernel.f90:196-222)
    const T dxVx = (vx0 - Vx[c - 1]) * e0x - (Vx[c + 1] - Vx[c - 2]) * e1x;
    const T dzVz = (vz0 - Vz[c - px]) * e0z - (Vz[c + px] - Vz[c - 2 * px]) * e1z;
    const float mu00 = a.mu[c];
    const T mu2 = 2 * (T)mu00;
    const T lam2mu = (T)a.lam[c] + mu2;
    const T taup1 = (T)a.taup[c], taus1 = (T)a.taus[c];
    const T d2v2 = dxVx + dzVz;
    const T f_Rxx = lam2mu * taup1 * d2v2 - mu2 * taus1 * dzVz;
    const T f_Rzz = lam2mu * taup1 * d2v2 - mu2 * taus1 * dxVx;
    T Rxx_n = 0, Rzz_n = 0;
#pragma unroll
    for (int m = 0; m < NM2; m++) {
        T rxx = a.c1[m] * a.R[m][c] - a.c2[m] * f_Rxx * dt;
        T rzz = a.c1[m] * a.R[3 + m][c] - a.c2[m] * f_Rzz * dt;
        a.R[m][c] = rxx;
        a.R[3 + m][c] = rzz;
        Rxx_n += a.d1[m] * rxx;
        Rzz_n += a.d1[m] * rzz;
    }
    const T taup_plus1 = 1 + taup1 * (1 + a.d2);
    const T taus_plus1 = 1 + taus1 * (1 + a.d2);
    a.Sxx[c] = a.Sxx[c] + (lam2mu * taup_plus1 * d2v2 - mu2 * taus_plus1 * dzVz + Rxx_n) * dt;
    a.Szz[c] = a.Szz[c] + (lam2mu * taup_plus1 * d2v2 - mu2 * taus_plus1 * dxVx + Rzz_n) * dt;

    // shear (m_kernel.f90:263-289)
    const T dxVz = (Vz[c + 1] - vz0) * e0x - (Vz[c + 2] - Vz[c - 1]) * e1x;
    const T dzVx = (Vx[c + px] - vx0) * e0z - (Vx[c + 2 * px] - Vx[c - px]) * e1z;
    const T mu_xz = (T)harm4(mu00, a.mu[c + px], a.mu[c + 1], a.mu[c + px + 1]);
    const T dxVz_dzVx = dxVz + dzVx;
    const T f_Rxz = mu_xz * taus1 * dxVz_dzVx;
    T Rxz_n = 0;
#pragma unroll
    for (int m = 0; m < NM2; m++) {
        T rxz = a.c1[m] * a.R[6 + m][c] - a.c2[m] * f_Rxz * dt;
        a.R[6 + m][c] = rxz;
        Rxz_n += a.d1[m] * rxz;
    }
    a.Sxz[c] = a.Sxz[c] + (mu_xz * taus_plus1 * dxVz_dzVx + Rxz_n) * dt;
}

// interior cell of tile thread (bx, bz, tx, tz), or false outside the interior
template <typename T>
__device__ __forceinline__ bool int_cell2(const K2<T> &a, int bx, int bz, int tx, int tz, int &i, int &k)
{
    i = a.ti0 + bx * 64 + tx;
    k = 1 + bz * 4 + tz;
    return !(i <= a.na || i > a.nx - a.na || k > a.nz - a.na);
}

template <typename T>
__global__ void __launch_bounds__(256) k2_stress_int(const K2<T> a)
{
    int i, k;
    if (int_cell2(a, (int)blockIdx.x, (int)blockIdx.y, (int)threadIdx.x, (int)threadIdx.y, i, k))
        stress_int_cell(a, i, k);
}

template <typename T>
__device__ __forceinline__ void vel_int_cell(const K2<T> &a, const int i, const int k)
{
    const int px = a.px;
    const int c = at2(px, i, k);
    const int4 fs = a.surf[i - 1];
    const int o1 = (k - fs.x) * (fs.y - k), o2 = (k - fs.z) * (fs.w - k);
    const bool lo = max(o1, o2) >= 0;
    const T e0x = lo ? a.e2[0][0] : a.e4[0][0], e1x = lo ? a.e2[0][1] : a.e4[0][1];
    const T e0z = lo ? a.e2[1][0] : a.e4[1][0], e1z = lo ? a.e2[1][1] : a.e4[1][1];
    const T *__restrict__ Sxx = a.Sxx;
    const T *__restrict__ Szz = a.Szz;
    const T *__restrict__ Sxz = a.Sxz;
    // m_kernel.f90:110-123
    const T sxz0 = Sxz[c];
    const T dxSxx = (Sxx[c + 1] - Sxx[c]) * e0x - (Sxx[c + 2] - Sxx[c - 1]) * e1x;
    const T dzSzz = (Szz[c + px] - Szz[c]) * e0z - (Szz[c + 2 * px] - Szz[c - px]) * e1z;
    const T dxSxz = (sxz0 - Sxz[c - 1]) * e0x - (Sxz[c + 1] - Sxz[c - 2]) * e1x;
    const T dzSxz = (sxz0 - Sxz[c - px]) * e0z - (Sxz[c + px] - Sxz[c - 2 * px]) * e1z;
    const float r0 = a.rho[c];
    const float bx = 2.0f / (r0 + a.rho[c + 1]);
    const float bz = 2.0f / (r0 + a.rho[c + px]);
    a.Vx[c] = a.Vx[c] + (T)bx * (dxSxx + dzSxz) * a.dt;
    a.Vz[c] = a.Vz[c] + (T)bz * (dxSxz + dzSzz) * a.dt;
}

template <typename T>
__global__ void __launch_bounds__(256) k2_vel_int(const K2<T> a)
{
    int i, k;
    if (int_cell2(a, (int)blockIdx.x, (int)blockIdx.y, (int)threadIdx.x, (int)threadIdx.y, i, k))
        vel_int_cell(a, i, k);
}

__device__ inline bool pml_cell2(const Boxes2 &B, unsigned id, int &i, int &k)
{
    if (id >= B.total) return false;
    int b = 0;
#pragma unroll
    for (int q = 1; q < 3; q++)
        if (q < B.n && id >= B.b[q].start) b = q;
    const Box2 &x = B.b[b];
    unsigned l = id - x.start;
    i = x.i0 + (int)(l % (unsigned)x.ni);
    k = x.k0 + (int)(l / (unsigned)x.ni);
    return true;
}

template <typename T>
__device__ __forceinline__ void stress_pml_cell2(const K2<T> &a, const int i, const int k)
{
    const int px = a.px, nx = a.nx, nz = a.nz;
    const int c = at2(px, i, k);
    const T *__restrict__ Vx = a.Vx;
    const T *__restrict__ Vz = a.Vz;
    const T gxc1 = a.gx[0 * nx + i - 1], gxc2 = a.gx[1 * nx + i - 1], gxc3 = a.gx[2 * nx + i - 1], gxc4 = a.gx[3 * nx + i - 1];
    const T gxe1 = a.gx[4 * nx + i - 1], gxe2 = a.gx[5 * nx + i - 1], gxe3 = a.gx[6 * nx + i - 1], gxe4 = a.gx[7 * nx + i - 1];
    const T gzc1 = a.gz[0 * nz + k - 1], gzc2 = a.gz[1 * nz + k - 1], gzc3 = a.gz[2 * nz + k - 1], gzc4 = a.gz[3 * nz + k - 1];
    const T gze1 = a.gz[4 * nz + k - 1], gze2 = a.gz[5 * nz + k - 1], gze3 = a.gz[6 * nz + k - 1], gze4 = a.gz[7 * nz + k - 1];
    const T r20x = a.r2[0], r20z = a.r2[1], dt = a.dt;
    const T vx0 = Vx[c], vz0 = Vz[c];
    // src/swpc_psv/m_absorb_p.f90:347-384
    const T dxVx = (vx0 - Vx[c - 1]) * r20x;
    const T dzVz = (vz0 - Vz[c - px]) * r20z;
    const float mu00 = a.mu[c];
    const T lam2mu_R = (T)(a.lam[c] + 2 * mu00);
    const T lam_R = lam2mu_R - 2 * (T)mu00;
    const T axvx = a.A[0][c], azvz = a.A[3][c];
    const T dxVx_ade = gxc1 * dxVx + gxc2 * axvx;
    const T dzVz_ade = gzc1 * dzVz + gzc2 * azvz;
    a.Sxx[c] = a.Sxx[c] + (lam2mu_R * dxVx_ade + lam_R * dzVz_ade) * dt;
    a.Szz[c] = a.Szz[c] + (lam2mu_R * dzVz_ade + lam_R * dxVx_ade) * dt;
    a.A[0][c] = gxc3 * axvx + gxc4 * dxVx * dt;
    a.A[3][c] = gzc3 * azvz + gzc4 * dzVz * dt;
    // shear (m_absorb_p.f90:453-476)
    const T dzVx = (Vx[c + px] - vx0) * r20z;
    const T dxVz = (Vz[c + 1] - vz0) * r20x;
    const T muxz = (T)harm4(mu00, a.mu[c + px], a.mu[c + 1], a.mu[c + px + 1]);
    const T azvx = a.A[1][c], axvz = a.A[2][c];
    a.Sxz[c] = a.Sxz[c] + muxz * (gxe1 * dxVz + gze1 * dzVx + gxe2 * axvz + gze2 * azvx) * dt;
    a.A[1][c] = gze3 * azvx + gze4 * dzVx * dt;
    a.A[2][c] = gxe3 * axvz + gxe4 * dxVz * dt;
}

template <typename T>
__global__ void __launch_bounds__(256) k2_stress_pml(const K2<T> a)
{
    int i, k;
    if (pml_cell2(a.pml, blockIdx.x * 256u + threadIdx.x, i, k)) stress_pml_cell2(a, i, k);
}

template <typename T>
__device__ __forceinline__ void vel_pml_cell2(const K2<T> &a, const int i, const int k)
{
    const int px = a.px, nx = a.nx, nz = a.nz;
    const int c = at2(px, i, k);
    const T gxc1 = a.gx[0 * nx + i - 1], gxc2 = a.gx[1 * nx + i - 1], gxc3 = a.gx[2 * nx + i - 1], gxc4 = a.gx[3 * nx + i - 1];
    const T gxe1 = a.gx[4 * nx + i - 1], gxe2 = a.gx[5 * nx + i - 1], gxe3 = a.gx[6 * nx + i - 1], gxe4 = a.gx[7 * nx + i - 1];
    const T gzc1 = a.gz[0 * nz + k - 1], gzc2 = a.gz[1 * nz + k - 1], gzc3 = a.gz[2 * nz + k - 1], gzc4 = a.gz[3 * nz + k - 1];
    const T gze1 = a.gz[4 * nz + k - 1], gze2 = a.gz[5 * nz + k - 1], gze3 = a.gz[6 * nz + k - 1], gze4 = a.gz[7 * nz + k - 1];
    const T r20x = a.r2[0], r20z = a.r2[1], dt = a.dt;
    // src/swpc_psv/m_absorb_p.f90:177-201
    const T sxx0 = a.Sxx[c], szz0 = a.Szz[c], sxz0 = a.Sxz[c];
    const T dxSxx = (a.Sxx[c + 1] - sxx0) * r20x;
    const T dzSzz = (a.Szz[c + px] - szz0) * r20z;
    const T dxSxz = (sxz0 - a.Sxz[c - 1]) * r20x;
    const T dzSxz = (sxz0 - a.Sxz[c - px]) * r20z;
    const float r0 = a.rho[c];
    const float bx = 2.0f / (r0 + a.rho[c + 1]);
    const float bz = 2.0f / (r0 + a.rho[c + px]);
    const T axsxx = a.A[4][c], azsxz = a.A[5][c], axsxz = a.A[6][c], azszz = a.A[7][c];
    a.Vx[c] = a.Vx[c] + (T)bx * (gxe1 * dxSxx + gzc1 * dzSxz + gxe2 * axsxx + gzc2 * azsxz) * dt;
    a.Vz[c] = a.Vz[c] + (T)bz * (gxc1 * dxSxz + gze1 * dzSzz + gxc2 * axsxz + gze2 * azszz) * dt;
    a.A[4][c] = gxe3 * axsxx + gxe4 * dxSxx * dt;
    a.A[5][c] = gzc3 * azsxz + gzc4 * dzSxz * dt;
    a.A[6][c] = gxc3 * axsxz + gxc4 * dxSxz * dt;
    a.A[7][c] = gze3 * azszz + gze4 * dzSzz * dt;
}

template <typename T>
__global__ void __launch_bounds__(256) k2_vel_pml(const K2<T> a)
{
    int i, k;
    if (pml_cell2(a.pml, blockIdx.x * 256u + threadIdx.x, i, k)) vel_pml_cell2(a, i, k);
}

template <typename T>
__global__ void k2_stressglut(const K2<T> a, int nsrc, const int *__restrict__ src, const double *__restrict__ m3,
                              const double *__restrict__ sdrop_all, int nt, const int *__restrict__ dit)
{
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    const int it = *dit;
    if (it < 1 || it > nt) return;
    const double *__restrict__ sdrop = sdrop_all + (size_t)(it - 1) * nsrc;
    const int px = a.px;
    for (int n = 0; n < nsrc; n++) {
        const int c = at2(px, src[2 * n], src[2 * n + 1]);
        const T sd = (T)sdrop[n];
        const T mxx = (T)m3[3 * n], mzz = (T)m3[3 * n + 1], mxz = (T)m3[3 * n + 2];
        a.Sxx[c] -= mxx * sd;
        a.Szz[c] -= mzz * sd;
        a.Sxz[c] -= mxz * sd / 4;
        a.Sxz[c - px] -= mxz * sd / 4;
        a.Sxz[c - 1] -= mxz * sd / 4;
        a.Sxz[c - 1 - px] -= mxz * sd / 4;
    }
}

template <typename T>
__global__ void k2_bodyforce(const K2<T> a, int nsrc, const int *__restrict__ src, const double *__restrict__ f2,
                             const double *__restrict__ fdt_all, int nt, const int *__restrict__ dit)
{
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    const int it = *dit;
    if (it < 1 || it > nt) return;
    const double *__restrict__ fdt = fdt_all + (size_t)(it - 1) * nsrc;
    const int px = a.px;
    for (int n = 0; n < nsrc; n++) {
        const int c = at2(px, src[2 * n], src[2 * n + 1]);
        const T f = (T)fdt[n] / (T)a.rho[c];
        const T fx = (T)f2[2 * n], fz = (T)f2[2 * n + 1];
        a.Vx[c] += fx * f;
        a.Vx[c - 1] += fx * f;
        a.Vz[c] += fz * f;
        a.Vz[c - px] += fz * f;
    }
}

template <typename T>
__device__ __forceinline__ void receiver2(const K2<T> &a, int n, int nst, const int *__restrict__ st,
                                          double *__restrict__ uacc, float *__restrict__ wvel,
                                          double *__restrict__ wdisp, int ntdec_w, int ntw,
                                          const int *__restrict__ dit)
{
    const int it = *dit;
    const int itw = (it - 1) / ntdec_w;
    const int store = ((it - 1) % ntdec_w) == 0 && itw < ntw;
    float *__restrict__ vel_out = wvel + (size_t)(store ? itw : 0) * nst * 2;
    double *__restrict__ disp_out = wdisp + (size_t)(store ? itw : 0) * nst * 2;
    const int px = a.px;
    const int c = at2(px, st[2 * n], st[2 * n + 1]);
    const T vx = a.Vx[c] + a.Vx[c - 1];
    const T vz = a.Vz[c] + a.Vz[c - px];
    const double hdt = 0.5 * (double)a.dt;
    const double ux = uacc[2 * n] + (double)vx * hdt;
    const double uz = uacc[2 * n + 1] - (double)vz * hdt;
    uacc[2 * n] = ux;
    uacc[2 * n + 1] = uz;
    if (store) {
        vel_out[2 * n] = (float)(vx / 2);
        vel_out[2 * n + 1] = (float)(-vz / 2);
        disp_out[2 * n] = ux;
        disp_out[2 * n + 1] = uz;
    }
}

template <typename T>
__global__ void k2_receivers(const K2<T> a, int nst, const int *__restrict__ st, double *__restrict__ uacc,
                             float *__restrict__ wvel, double *__restrict__ wdisp, int ntdec_w, int ntw,
                             const int *__restrict__ dit)
{
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n < nst) receiver2(a, n, nst, st, uacc, wvel, wdisp, ntdec_w, ntw, dit);
}

// ---------------------------------------------------------------------------
// Fused step kernels.  A P-SV grid of a few hundred thousand cells is
// updated in microseconds, so the step's seven launches cost more in gaps
// and tails than in work.  Two launches do the same work: one workgroup
// range per part, each part's workgroups uniform in what they run.
//   k2_stress_step  receivers (they read V only), interior and absorber
//                   stress, and the stress glut of each source cell by
//                   the thread that updated that cell, after its update
//   k2_vel_step     interior and absorber velocity, and the body force of
//                   each force cell by the thread that updated that cell
// The per-cell operations and their order are those of the separate
// kernels (k2_stressglut applies its sources in order, and each cell's
// share of a source after the cell's update), so both forms agree bit for
// bit.  The host fuses only sources whose footprint lies in [1, nx] x
// [1, nz] and at most MAXFUSE of them; otherwise it launches the separate
// kernels.
constexpr int MAXFUSE = 8;

struct Src2 {
    int n, nt;                      // sources (0: none fused), table steps
    const int *__restrict__ ik;     // (i, k) per source
    const double *__restrict__ m;   // m3 (stress glut) or f2 (body force) per source
    const double *__restrict__ tab; // [it-1][n]
};

struct Recv2 {
    int nst, ntdec_w, ntw;
    const int *__restrict__ st;
    double *__restrict__ uacc;
    float *__restrict__ wvel;
    double *__restrict__ wdisp;
};

// k2_stressglut's share for cell (i, k): Sxx, Szz and a quarter of Sxz at
// the source cell, a quarter of Sxz at its -x, -z and -x-z neighbours
template <typename T>
__device__ __forceinline__ void glut_at2(const K2<T> &a, const Src2 &s, const int *__restrict__ dit, int i, int k)
{
    for (int n = 0; n < s.n; n++) {
        const int di = s.ik[2 * n] - i, dk = s.ik[2 * n + 1] - k;
        if ((unsigned)di > 1u || (unsigned)dk > 1u) continue;
        const int it = *dit;
        if (it < 1 || it > s.nt) return;
        const int c = at2(a.px, i, k);
        const T sd = (T)s.tab[(size_t)(it - 1) * s.n + n];
        const T mxx = (T)s.m[3 * n], mzz = (T)s.m[3 * n + 1], mxz = (T)s.m[3 * n + 2];
        if (di == 0 && dk == 0) {
            a.Sxx[c] -= mxx * sd;
            a.Szz[c] -= mzz * sd;
        }
        a.Sxz[c] -= mxz * sd / 4;
    }
}

// k2_bodyforce's share for cell (i, k): Vx at the force cell and its -x
// neighbour, Vz at the force cell and its -z neighbour, f = fdt / rho of
// the force cell
template <typename T>
__device__ __forceinline__ void bf_at2(const K2<T> &a, const Src2 &s, const int *__restrict__ dit, int i, int k)
{
    for (int n = 0; n < s.n; n++) {
        const int si = s.ik[2 * n], sk = s.ik[2 * n + 1];
        const int di = si - i, dk = sk - k;
        if ((unsigned)di > 1u || (unsigned)dk > 1u || (di == 1 && dk == 1)) continue;
        const int it = *dit;
        if (it < 1 || it > s.nt) return;
        const int c = at2(a.px, i, k);
        const T f = (T)s.tab[(size_t)(it - 1) * s.n + n] / (T)a.rho[at2(a.px, si, sk)];
        const T fx = (T)s.m[2 * n], fz = (T)s.m[2 * n + 1];
        if (dk == 0) a.Vx[c] += fx * f;
        if (di == 0) a.Vz[c] += fz * f;
    }
}

// workgroups: [0, nint) interior tiles (ntx per tile row), [nint, nint +
// npml) absorber cells, then receivers
template <typename T>
__device__ __forceinline__ void stress_step_block(const K2<T> &a, int ntx, unsigned nint, unsigned npml,
                                                  const Src2 &s, const Recv2 &r, const int *__restrict__ dit,
                                                  const unsigned b)
{
    int i, k;
    if (b < nint) {
        if (!int_cell2(a, (int)(b % (unsigned)ntx), (int)(b / (unsigned)ntx), (int)(threadIdx.x & 63),
                       (int)(threadIdx.x >> 6), i, k))
            return;
        stress_int_cell(a, i, k);
    } else if (b < nint + npml) {
        if (!pml_cell2(a.pml, (b - nint) * 256u + threadIdx.x, i, k)) return;
        stress_pml_cell2(a, i, k);
    } else {
        const int n = (int)((b - nint - npml) * 256u + threadIdx.x);
        if (n < r.nst) receiver2(a, n, r.nst, r.st, r.uacc, r.wvel, r.wdisp, r.ntdec_w, r.ntw, dit);
        return;
    }
    if (s.n) glut_at2(a, s, dit, i, k);
}

template <typename T>
__device__ __forceinline__ void vel_step_block(const K2<T> &a, int ntx, unsigned nint, const Src2 &s,
                                               const int *__restrict__ dit, const unsigned b)
{
    int i, k;
    if (b < nint) {
        if (!int_cell2(a, (int)(b % (unsigned)ntx), (int)(b / (unsigned)ntx), (int)(threadIdx.x & 63),
                       (int)(threadIdx.x >> 6), i, k))
            return;
        vel_int_cell(a, i, k);
    } else {
        if (!pml_cell2(a.pml, (b - nint) * 256u + threadIdx.x, i, k)) return;
        vel_pml_cell2(a, i, k);
    }
    if (s.n) bf_at2(a, s, dit, i, k);
}

template <typename T>
__global__ void __launch_bounds__(256) k2_stress_step(const K2<T> a, int ntx, unsigned nint, unsigned npml, Src2 s,
                                                       Recv2 r, const int *__restrict__ dit)
{
    stress_step_block(a, ntx, nint, npml, s, r, dit, blockIdx.x);
}

template <typename T>
__global__ void __launch_bounds__(256) k2_vel_step(const K2<T> a, int ntx, unsigned nint, Src2 s,
                                                    const int *__restrict__ dit)
{
    vel_step_block(a, ntx, nint, s, dit, blockIdx.x);
}

// ---------------------------------------------------------------------------
// Groups of P-SV runs in one launch per phase (batches): the workgroups of
// every member's fused step kernel, end to end.  A workgroup finds its
// member by bisection of the members' first workgroups and runs that
// member's block of k2_stress_step / k2_vel_step, with the member's own
// arguments, sources, receivers and step counter, so every member's cells
// see exactly the operations of its own run.
template <typename T>
struct Member2 {
    K2<T> a;
    Src2 sg, sb;
    Recv2 rv;
    int *dit;
    int ntx;
    unsigned nint, npml;
};

__device__ __forceinline__ int member_of(const unsigned *__restrict__ first, int nm, unsigned b)
{
    int lo = 0, hi = nm - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (first[mid] <= b) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

template <typename T>
__global__ void __launch_bounds__(256) k2_stress_group(const Member2<T> *__restrict__ mem,
                                                        const unsigned *__restrict__ first, int nm)
{
    const int m = member_of(first, nm, blockIdx.x);
    const Member2<T> &g = mem[m];
    stress_step_block(g.a, g.ntx, g.nint, g.npml, g.sg, g.rv, g.dit, blockIdx.x - first[m]);
}

template <typename T>
__global__ void __launch_bounds__(256) k2_vel_group(const Member2<T> *__restrict__ mem,
                                                     const unsigned *__restrict__ first, int nm)
{
    const int m = member_of(first, nm, blockIdx.x);
    const Member2<T> &g = mem[m];
    vel_step_block(g.a, g.ntx, g.nint, g.sb, g.dit, blockIdx.x - first[m]);
}

__global__ void k2_group_advance(int *const *__restrict__ dits, int n)
{
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q < n) *dits[q] = *dits[q] + 1;
}

__global__ void k2_fill_u32(unsigned *p, int n, unsigned v)
{
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) p[t] = v;
}

// device-side step number (see the 3-D solver): set per run, advanced per step
__global__ void k2_step_set(int *dit, int it)
{
    if (blockIdx.x == 0 && threadIdx.x == 0) *dit = it;
}

__global__ void k2_step_advance(int *dit)
{
    if (blockIdx.x == 0 && threadIdx.x == 0) *dit = *dit + 1;
}

template <typename T>
__global__ void k2_vmax(const K2<T> a, const int *__restrict__ kob, unsigned *__restrict__ out)
{
    const int i = a.na + 1 + (int)(blockIdx.x * blockDim.x + threadIdx.x);
    float m0 = 0, m1 = 0;
    if (i <= a.nx - a.na) {
        const int c = at2(a.px, i, kob[i - 1] + 1);
        m0 = fabsf((float)a.Vx[c]);
        m1 = fabsf((float)a.Vz[c]);
    }
    for (int off = 32; off > 0; off >>= 1) {
        m0 = fmaxf(m0, __shfl_xor(m0, off));
        m1 = fmaxf(m1, __shfl_xor(m1, off));
    }
    if ((threadIdx.x & 63) == 0) {
        atomicMax(out + 0, __float_as_uint(m0));
        atomicMax(out + 1, __float_as_uint(m1));
    }
}

// P-SV snapshot products at a list of (i,k) cells (src/swpc_psv/m_snap.f90):
// kind 0 'ps' follows wbuf_xz_ps (461-494): div and rot masked by lam and
// by the harmonic mu_xz, then x UC*M0*1e-3; kind 1 'v' wbuf_xz_v (552-553);
// kind 2 'u' wbuf_xz_u (612-613), advanced at every call.  Variable-major
// output out[n*npts + p]; the reference's operation order, no contraction.
template <typename T>
__global__ void k2_snap(const K2<T> a, int kind, int npts, const int2 *__restrict__ pts, float uc, float m0, float dt,
                        float *__restrict__ acc, float *__restrict__ frames, int ntdec, int nbuf,
                        const int *__restrict__ dit)
{
#pragma clang fp contract(off)
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= npts) return;
    const int it = *dit;
    const int store = ((it - 1) % ntdec) == 0;
    if (!store && kind != 2) return;
    float *__restrict__ out = frames + (size_t)(((it - 1) / ntdec) % nbuf) * 2 * npts;
    const int2 q = pts[p];
    const int px = a.px;
    const int c = at2(px, q.x, q.y);
    const T *vx = a.Vx, *vz = a.Vz;
    if (kind == 0) {
        const T rx = a.r2[0], rz = a.r2[1];
        float dv = (float)((vx[c] - vx[c - 1]) * rx + (vz[c] - vz[c - px]) * rz);
        float rt = (float)((vx[c + px] - vx[c]) * rz - (vz[c + 1] - vz[c]) * rx);
        const float eps = 1.1920929e-07f;  // epsilon(1.0)
        const float muxz = harm4(a.mu[c], a.mu[c + px], a.mu[c + 1], a.mu[c + px + 1]);
        const float lam = a.lam[c];
        dv = dv * lam / (fabsf(lam) + eps);
        rt = rt * muxz / fabsf(muxz + eps);
        out[p] = dv * uc * m0 * 1e-3f;
        out[npts + p] = rt * uc * m0 * 1e-3f;
    } else if (kind == 1) {
        out[p] = (float)(vx[c] * (T)uc * (T)m0);
        out[npts + p] = (float)(vz[c] * (T)uc * (T)m0);
    } else {
        const float ux = (float)((T)acc[p] + vx[c] * (T)uc * (T)m0 * (T)dt);
        const float uz = (float)((T)acc[npts + p] + vz[c] * (T)uc * (T)m0 * (T)dt);
        acc[p] = ux;
        acc[npts + p] = uz;
        if (store) {
            out[p] = ux;
            out[npts + p] = uz;
        }
    }
}

}  // namespace swpc

struct swpc2d_ctx {
    swpc2d_desc d;
    swpc::Lay2 L;
    size_t esz;
    hipStream_t stream = nullptr;
    long long bytes = 0;
    swpc::DBuf V[2], S[3], R[9], A[8];
    swpc::DBuf med[5];  // rho lam mu taup taus
    swpc::DBuf surf, kob, gx, gz;
    swpc::DBuf src, m3, sdrop, bsrc, f2, fdt;
    swpc::DBuf st, uacc, wvel, wdisp, vmax;
    int nsrc = 0, nt_src = 0, nbf = 0, nt_bf = 0;
    int nst = 0, ntdec_w = 1, ntw = 0;
    bool have_medium = false, have_surface = false, have_pml = false;
    swpc::Boxes2 boxes;
    int ti0 = 1, ntx = 0, ntz = 0;
    int graph_k = 16;       // steps per captured step graph (0 = launch step by step)
    bool fused = true;      // steps in the fused kernels (swpc2d_set_fusion)
    bool fuse_src = true, fuse_bf = true;   // the stress glut / body force inside them
    hipGraph_t graph = nullptr;
    unsigned gen = 0;       // bumped by every graph drop (batches recapture on change)
    hipGraphExec_t gexec = nullptr;
    swpc::DBuf dit;         // device step number
    struct Snap {
        int kind, npts, ntdec, nbuf;
        float uc, m0, dt;
        swpc::DBuf pts, acc, frames;
        long long last = -1;
    };
    std::vector<Snap> snaps;
};

namespace swpc {

template <typename T>
static K2<T> make_args2(swpc2d_ctx *c)
{
    K2<T> a;
    memset(&a, 0, sizeof a);
    const swpc2d_desc &d = c->d;
    a.px = c->L.px; a.nx = d.nx; a.nz = d.nz; a.na = d.na;
    a.Vx = (T *)c->V[0].p; a.Vz = (T *)c->V[1].p;
    a.Sxx = (T *)c->S[0].p; a.Szz = (T *)c->S[1].p; a.Sxz = (T *)c->S[2].p;
    for (int q = 0; q < 9; q++) a.R[q] = (T *)c->R[q].p;
    for (int q = 0; q < 8; q++) a.A[q] = (T *)c->A[q].p;
    a.rho = (const float *)c->med[0].p; a.lam = (const float *)c->med[1].p; a.mu = (const float *)c->med[2].p;
    a.taup = (const float *)c->med[3].p; a.taus = (const float *)c->med[4].p;
    a.surf = (const int4 *)c->surf.p;
    a.gx = (const T *)c->gx.p; a.gz = (const T *)c->gz.p;
    const double h[2] = {d.dx, d.dz};
    for (int q = 0; q < 2; q++) {
        T hh = (T)h[q];
        T rc40 = (T)17.0 / (T)16.0 / hh, rc41 = (T)1.0 / (T)48.0 / hh;
        T rd40 = (T)(-1.0) / (T)16.0 / hh, rd41 = (T)(-1.0) / (T)48.0 / hh;
        a.e4[q][0] = rc40 - rd40; a.e4[q][1] = rc41 - rd41;
        a.e2[q][0] = rc40 + rd40; a.e2[q][1] = rc41 + rd41;
    }
    // r20 is real(SP) in the P-SV absorber (src/swpc_psv/m_absorb_p.f90:69-71)
    a.r2[0] = (T)(1.0 / d.dx);
    a.r2[1] = (T)(1.0 / d.dz);
    T t = (T)(float)d.dt, s = 0;
    for (int m = 0; m < NM2; m++) {
        T tm = (T)d.ts[m];
        a.c1[m] = (2 * tm - t) / (2 * tm + t);
        a.c2[m] = (T)2 / (2 * tm + t) / NM2;
        s += t / (2 * tm - t);
        a.d1[m] = 2 * tm / (2 * tm - t);
    }
    a.d2 = s / NM2;
    a.dt = t;
    a.ti0 = c->ti0;
    a.pml = c->boxes;
    return a;
}

static bool fp64(const swpc2d_ctx *c) { return c->d.precision == SWPC_FP64; }

static int ready2(const swpc2d_ctx *c)
{
    if (!c) return fail(SWPC_EINVAL, "null context");
    if (!c->have_medium || !c->have_surface || !c->have_pml)
        return fail(SWPC_ESTATE, "set_medium, set_surface and set_pml must precede time stepping");
    return SWPC_OK;
}

template <typename T>
static int l2_stress(swpc2d_ctx *c)
{
    K2<T> a = make_args2<T>(c);
    if (c->ntx > 0 && c->ntz > 0) {
        hipLaunchKernelGGL((k2_stress_int<T>), dim3(c->ntx, c->ntz), dim3(64, 4), 0, c->stream, a);
        SWPC_CHECK_LAUNCH();
    }
    if (c->boxes.total > 0) {
        hipLaunchKernelGGL((k2_stress_pml<T>), dim3((c->boxes.total + 255) / 256), dim3(256), 0, c->stream, a);
        SWPC_CHECK_LAUNCH();
    }
    return SWPC_OK;
}

template <typename T>
static int l2_vel(swpc2d_ctx *c)
{
    K2<T> a = make_args2<T>(c);
    if (c->ntx > 0 && c->ntz > 0) {
        hipLaunchKernelGGL((k2_vel_int<T>), dim3(c->ntx, c->ntz), dim3(64, 4), 0, c->stream, a);
        SWPC_CHECK_LAUNCH();
    }
    if (c->boxes.total > 0) {
        hipLaunchKernelGGL((k2_vel_pml<T>), dim3((c->boxes.total + 255) / 256), dim3(256), 0, c->stream, a);
        SWPC_CHECK_LAUNCH();
    }
    return SWPC_OK;
}

template <typename T>
static int l2_glut(swpc2d_ctx *c)
{
    if (c->nsrc == 0) return SWPC_OK;
    K2<T> a = make_args2<T>(c);
    hipLaunchKernelGGL((k2_stressglut<T>), dim3(1), dim3(64), 0, c->stream, a, c->nsrc, (const int *)c->src.p,
                       (const double *)c->m3.p, (const double *)c->sdrop.p, c->nt_src, (const int *)c->dit.p);
    SWPC_CHECK_LAUNCH();
    return SWPC_OK;
}

template <typename T>
static int l2_bf(swpc2d_ctx *c)
{
    if (c->nbf == 0) return SWPC_OK;
    K2<T> a = make_args2<T>(c);
    hipLaunchKernelGGL((k2_bodyforce<T>), dim3(1), dim3(64), 0, c->stream, a, c->nbf, (const int *)c->bsrc.p,
                       (const double *)c->f2.p, (const double *)c->fdt.p, c->nt_bf, (const int *)c->dit.p);
    SWPC_CHECK_LAUNCH();
    return SWPC_OK;
}

template <typename T>
static int l2_recv(swpc2d_ctx *c)
{
    if (c->nst == 0) return SWPC_OK;
    K2<T> a = make_args2<T>(c);
    hipLaunchKernelGGL((k2_receivers<T>), dim3((c->nst + 63) / 64), dim3(64), 0, c->stream, a, c->nst,
                       (const int *)c->st.p, (double *)c->uacc.p, (float *)c->wvel.p, (double *)c->wdisp.p,
                       c->ntdec_w, c->ntw, (const int *)c->dit.p);
    SWPC_CHECK_LAUNCH();
    return SWPC_OK;
}

template <typename T>
static int l2_snap(swpc2d_ctx *c)
{
    if (c->snaps.empty()) return SWPC_OK;
    K2<T> a = make_args2<T>(c);
    for (auto &sn : c->snaps) {
        hipLaunchKernelGGL((k2_snap<T>), dim3((sn.npts + 255) / 256), dim3(256), 0, c->stream, a, sn.kind, sn.npts,
                           (const int2 *)sn.pts.p, sn.uc, sn.m0, sn.dt, (float *)sn.acc.p, (float *)sn.frames.p,
                           sn.ntdec, sn.nbuf, (const int *)c->dit.p);
        SWPC_CHECK_LAUNCH();
    }
    return SWPC_OK;
}

static void note_snapshots2(swpc2d_ctx *c, int it0, int it1)
{
    for (auto &sn : c->snaps) {
        const long long last_due = (long long)(it1 - 1) / sn.ntdec * sn.ntdec + 1;
        if (last_due >= it0) sn.last = std::max(sn.last, (last_due - 1) / sn.ntdec);
    }
}

static int set_step2(swpc2d_ctx *c, int it)
{
    hipLaunchKernelGGL(k2_step_set, dim3(1), dim3(64), 0, c->stream, (int *)c->dit.p, it);
    SWPC_CHECK_LAUNCH();
    return SWPC_OK;
}

// P-SV step order (src/swpc_psv/main.f90:95-114): snapshots, receivers, stress,
// absorber, stress glut, velocity, body force, absorber.  The additive
// body force commutes with the absorber update.  The step number is the
// device counter, advanced at the end.
template <typename T>
static int step2_split(swpc2d_ctx *c)
{
    int r;
    if ((r = l2_snap<T>(c))) return r;
    if ((r = l2_recv<T>(c))) return r;
    if ((r = l2_stress<T>(c))) return r;
    if ((r = l2_glut<T>(c))) return r;
    if ((r = l2_vel<T>(c))) return r;
    if ((r = l2_bf<T>(c))) return r;
    hipLaunchKernelGGL(k2_step_advance, dim3(1), dim3(64), 0, c->stream, (int *)c->dit.p);
    SWPC_CHECK_LAUNCH();
    return SWPC_OK;
}

// Arguments of the fused step kernels for one context.
template <typename T>
static void member2_args(swpc2d_ctx *c, Member2<T> &g)
{
    memset(&g, 0, sizeof g);
    g.a = make_args2<T>(c);
    g.nint = (unsigned)c->ntx * (unsigned)c->ntz;
    g.npml = (c->boxes.total + 255) / 256;
    g.ntx = std::max(c->ntx, 1);
    g.dit = (int *)c->dit.p;
    if (c->nst > 0) {
        g.rv.nst = c->nst; g.rv.ntdec_w = c->ntdec_w; g.rv.ntw = c->ntw;
        g.rv.st = (const int *)c->st.p; g.rv.uacc = (double *)c->uacc.p;
        g.rv.wvel = (float *)c->wvel.p; g.rv.wdisp = (double *)c->wdisp.p;
    }
    if (c->nsrc > 0 && c->fuse_src) {
        g.sg.n = c->nsrc; g.sg.nt = c->nt_src;
        g.sg.ik = (const int *)c->src.p; g.sg.m = (const double *)c->m3.p; g.sg.tab = (const double *)c->sdrop.p;
    }
    if (c->nbf > 0 && c->fuse_bf) {
        g.sb.n = c->nbf; g.sb.nt = c->nt_bf;
        g.sb.ik = (const int *)c->bsrc.p; g.sb.m = (const double *)c->f2.p; g.sb.tab = (const double *)c->fdt.p;
    }
}

template <typename T>
static unsigned stress_blocks2(const Member2<T> &g) { return g.nint + g.npml + ((unsigned)g.rv.nst + 255u) / 256u; }

// The same step in the fused kernels (k2_stress_step, k2_vel_step): the
// snapshots, a stress glut or body force that is not fused, and the step
// counter keep launches of their own.
template <typename T>
static int step2_fused(swpc2d_ctx *c)
{
    int r;
    if ((r = l2_snap<T>(c))) return r;
    Member2<T> g;
    member2_args(c, g);
    const unsigned ns = stress_blocks2(g), nv = g.nint + g.npml;
    if (ns > 0) {
        hipLaunchKernelGGL((k2_stress_step<T>), dim3(ns), dim3(256), 0, c->stream, g.a, g.ntx, g.nint, g.npml, g.sg,
                           g.rv, (const int *)g.dit);
        SWPC_CHECK_LAUNCH();
    }
    if (!c->fuse_src && (r = l2_glut<T>(c))) return r;
    if (nv > 0) {
        hipLaunchKernelGGL((k2_vel_step<T>), dim3(nv), dim3(256), 0, c->stream, g.a, g.ntx, g.nint, g.sb,
                           (const int *)g.dit);
        SWPC_CHECK_LAUNCH();
    }
    if (!c->fuse_bf && (r = l2_bf<T>(c))) return r;
    hipLaunchKernelGGL(k2_step_advance, dim3(1), dim3(64), 0, c->stream, (int *)c->dit.p);
    SWPC_CHECK_LAUNCH();
    return SWPC_OK;
}

template <typename T>
static int step2_dev(swpc2d_ctx *c)
{
    return c->fused ? step2_fused<T>(c) : step2_split<T>(c);
}

static int step2_any(swpc2d_ctx *c) { return fp64(c) ? step2_dev<double>(c) : step2_dev<float>(c); }

static void graph_drop2(swpc2d_ctx *c)
{
    c->gen++;
    if (c->gexec) (void)hipGraphExecDestroy(c->gexec);
    if (c->graph) (void)hipGraphDestroy(c->graph);
    c->gexec = nullptr;
    c->graph = nullptr;
}

static int graph_ready2(swpc2d_ctx *c)
{
    if (c->gexec) return SWPC_OK;
    SWPC_HIP(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    int r = SWPC_OK;
    for (int q = 0; q < c->graph_k && r == SWPC_OK; q++) r = step2_any(c);
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(c->stream, &g);
    if (r != SWPC_OK) {
        if (g) (void)hipGraphDestroy(g);
        return r;
    }
    if (e != hipSuccess || !g) return fail(SWPC_EHIP, "step graph capture failed: %s", hipGetErrorString(e));
    hipGraphExec_t ge = nullptr;
    const hipError_t ei = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
    if (ei != hipSuccess) {
        (void)hipGraphDestroy(g);
        return fail(SWPC_EHIP, "step graph instantiation failed: %s", hipGetErrorString(ei));
    }
    c->graph = g;
    c->gexec = ge;
    return SWPC_OK;
}

static int run_steps2(swpc2d_ctx *c, int it0, int n)
{
    if (n <= 0) return SWPC_OK;
    if (c->nsrc > 0 && it0 + n - 1 > c->nt_src)
        return fail(SWPC_EINVAL, "stressglut: steps %d..%d outside 1..%d", it0, it0 + n - 1, c->nt_src);
    if (c->nbf > 0 && it0 + n - 1 > c->nt_bf)
        return fail(SWPC_EINVAL, "bodyforce: steps %d..%d outside 1..%d", it0, it0 + n - 1, c->nt_bf);
    int r;
    if ((r = set_step2(c, it0))) return r;
    int done = 0;
    if (c->graph_k > 0 && n >= c->graph_k) {
        if ((r = graph_ready2(c))) return r;
        for (; done + c->graph_k <= n; done += c->graph_k) SWPC_HIP(hipGraphLaunch(c->gexec, c->stream));
    }
    for (; done < n; done++)
        if ((r = step2_any(c))) return r;
    note_snapshots2(c, it0, it0 + n - 1);
    return SWPC_OK;
}

// Hooks of swpc_batch.hip.
int batch_member2(swpc2d_ctx *c, BatchMember &m)
{
    if (int r = ready2(c)) return r;
    m.stream = c->stream;
    m.dit = (int *)c->dit.p;
    m.gen = c->gen;
    return SWPC_OK;
}

int batch_check2(swpc2d_ctx *c, int it0, int n)
{
    if (c->nsrc > 0 && it0 + n - 1 > c->nt_src)
        return fail(SWPC_EINVAL, "stressglut: steps %d..%d outside 1..%d", it0, it0 + n - 1, c->nt_src);
    if (c->nbf > 0 && it0 + n - 1 > c->nt_bf)
        return fail(SWPC_EINVAL, "bodyforce: steps %d..%d outside 1..%d", it0, it0 + n - 1, c->nt_bf);
    return SWPC_OK;
}

int batch_step2(swpc2d_ctx *c) { return step2_any(c); }

void batch_done2(swpc2d_ctx *c, int it0, int it1) { note_snapshots2(c, it0, it1); }

struct Group2 {
    DBuf mem, first_s, first_v, dits;
    unsigned ns = 0, nv = 0;
    int nm = 0;
};

bool group2_eligible(const swpc2d_ctx *c)
{
    return c && !fp64(c) && c->fused && c->snaps.empty() && (c->nsrc == 0 || c->fuse_src) &&
           (c->nbf == 0 || c->fuse_bf) && ready2(c) == SWPC_OK;
}

void group2_free(Group2 *g)
{
    if (!g) return;
    dfree(g->mem);
    dfree(g->first_s);
    dfree(g->first_v);
    dfree(g->dits);
    delete g;
}

int group2_build(const std::vector<swpc2d_ctx *> &cs, Group2 **out, long long &bytes)
{
    *out = nullptr;
    if (cs.empty()) return SWPC_OK;
    std::vector<Member2<float>> mem(cs.size());
    std::vector<unsigned> fs(cs.size()), fv(cs.size());
    std::vector<int *> dits(cs.size());
    unsigned long long ns = 0, nv = 0;
    for (size_t q = 0; q < cs.size(); q++) {
        if (!group2_eligible(cs[q])) return fail(SWPC_EINVAL, "P-SV member %zu cannot join a group", q);
        member2_args(cs[q], mem[q]);
        fs[q] = (unsigned)ns;
        fv[q] = (unsigned)nv;
        ns += stress_blocks2(mem[q]);
        nv += mem[q].nint + mem[q].npml;
        dits[q] = mem[q].dit;
    }
    if (ns >= (1ULL << 31)) return fail(SWPC_EINVAL, "P-SV group too large for one launch");
    auto *g = new Group2();
    g->nm = (int)cs.size();
    g->ns = (unsigned)ns;
    g->nv = (unsigned)nv;
    int r = dalloc(g->mem, mem.size() * sizeof(Member2<float>), bytes);
    if (!r) r = dalloc(g->first_s, fs.size() * 4, bytes);
    if (!r) r = dalloc(g->first_v, fv.size() * 4, bytes);
    if (!r) r = dalloc(g->dits, dits.size() * sizeof(int *), bytes);
    hipError_t e = hipSuccess;
    if (!r) e = hipMemcpy(g->mem.p, mem.data(), mem.size() * sizeof(Member2<float>), hipMemcpyHostToDevice);
    if (!r && e == hipSuccess) e = hipMemcpy(g->first_s.p, fs.data(), fs.size() * 4, hipMemcpyHostToDevice);
    if (!r && e == hipSuccess) e = hipMemcpy(g->first_v.p, fv.data(), fv.size() * 4, hipMemcpyHostToDevice);
    if (!r && e == hipSuccess) e = hipMemcpy(g->dits.p, dits.data(), dits.size() * sizeof(int *), hipMemcpyHostToDevice);
    if (r || e != hipSuccess) {
        group2_free(g);
        return r ? r : fail(SWPC_EHIP, "P-SV group tables: %s", hipGetErrorString(e));
    }
    *out = g;
    return SWPC_OK;
}

int group2_step(Group2 *g, hipStream_t s)
{
    if (g->ns > 0) {
        hipLaunchKernelGGL((k2_stress_group<float>), dim3(g->ns), dim3(256), 0, s, (const Member2<float> *)g->mem.p,
                           (const unsigned *)g->first_s.p, g->nm);
        SWPC_CHECK_LAUNCH();
    }
    if (g->nv > 0) {
        hipLaunchKernelGGL((k2_vel_group<float>), dim3(g->nv), dim3(256), 0, s, (const Member2<float> *)g->mem.p,
                           (const unsigned *)g->first_v.p, g->nm);
        SWPC_CHECK_LAUNCH();
    }
    hipLaunchKernelGGL(k2_group_advance, dim3((g->nm + 63) / 64), dim3(64), 0, s, (int *const *)g->dits.p, g->nm);
    SWPC_CHECK_LAUNCH();
    return SWPC_OK;
}

template <typename S, typename D>
static void box2_to_dev(const Lay2 &L, const S *src, D *dst)
{
    const int nkb = L.nz + 6;
    for (int k = -2; k <= L.nz + 3; k++)
        for (int i = -2; i <= L.nx + 3; i++) dst[at2(L.px, i, k)] = (D)src[(size_t)(k + 2) + (size_t)nkb * (size_t)(i + 2)];
}

template <typename S, typename D>
static void dev_to_box2(const Lay2 &L, const S *src, D *dst)
{
    const int nkb = L.nz + 6;
    for (int k = -2; k <= L.nz + 3; k++)
        for (int i = -2; i <= L.nx + 3; i++) dst[(size_t)(k + 2) + (size_t)nkb * (size_t)(i + 2)] = (D)src[at2(L.px, i, k)];
}

static DBuf *field_buf2(swpc2d_ctx *c, int f)
{
    if (f >= SWPC2D_VX && f <= SWPC2D_VZ) return &c->V[f];
    if (f >= SWPC2D_SXX && f <= SWPC2D_SXZ) return &c->S[f - SWPC2D_SXX];
    if (f >= SWPC2D_R0 && f < SWPC2D_ADE0) return &c->R[f - SWPC2D_R0];
    if (f >= SWPC2D_ADE0 && f < SWPC2D_NFIELD) return &c->A[f - SWPC2D_ADE0];
    return nullptr;
}

static int upload_profile2(swpc2d_ctx *c, DBuf &dst, const float *gc, const float *ge, int n)
{
    std::vector<double> h((size_t)8 * n);
    for (int q = 0; q < 4; q++)
        for (int i = 0; i < n; i++) {
            h[(size_t)q * n + i] = gc[q + 4 * i];
            h[(size_t)(4 + q) * n + i] = ge[q + 4 * i];
        }
    if (int r = dalloc(dst, (size_t)8 * n * c->esz, c->bytes)) return r;
    if (fp64(c)) {
        SWPC_HIP(hipMemcpy(dst.p, h.data(), h.size() * 8, hipMemcpyHostToDevice));
    } else {
        std::vector<float> f(h.begin(), h.end());
        SWPC_HIP(hipMemcpy(dst.p, f.data(), f.size() * 4, hipMemcpyHostToDevice));
    }
    return SWPC_OK;
}

}  // namespace swpc

using namespace swpc;

extern "C" {

int swpc2d_create(const swpc2d_desc *desc, swpc2d_ctx **out)
{
    clear_error();
    if (!desc || !out) return fail(SWPC_EINVAL, "null argument");
    *out = nullptr;
    const swpc2d_desc &d = *desc;
    if (d.nx < 8 || d.nz < 8) return fail(SWPC_EINVAL, "grid %dx%d too small", d.nx, d.nz);
    if (d.na < 0 || 2 * d.na >= d.nx || d.na >= d.nz) return fail(SWPC_EINVAL, "absorber thickness na=%d does not fit", d.na);
    if (d.nm != NM2) return fail(SWPC_EINVAL, "nm=%d: the kernels are built for NM = 3", d.nm);
    if (d.precision != SWPC_FP32 && d.precision != SWPC_FP64)
        return fail(SWPC_EINVAL, "precision mode %d not supported by this build", d.precision);
    if (!(d.dx > 0 && d.dz > 0 && d.dt > 0)) return fail(SWPC_EINVAL, "non-positive dx/dz/dt");
    for (int m = 0; m < d.nm; m++)
        if (!(d.ts[m] > 0)) return fail(SWPC_EINVAL, "ts(%d) must be positive", m + 1);
    Lay2 L = make_lay2(d.nx, d.nz);
    if (L.vol >= (1LL << 31)) return fail(SWPC_EINVAL, "grid too large");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return fail(SWPC_EHIP, "no HIP device available");
    std::unique_ptr<swpc2d_ctx> c(new swpc2d_ctx());
    c->d = d;
    c->L = L;
    c->esz = d.precision == SWPC_FP64 ? 8 : 4;
    SWPC_HIP(hipStreamCreate(&c->stream));
    const size_t vb = (size_t)L.vol * c->esz;
    int r;
    for (auto &b : c->V) if ((r = dalloc(b, vb, c->bytes))) return r;
    for (auto &b : c->S) if ((r = dalloc(b, vb, c->bytes))) return r;
    for (auto &b : c->R) if ((r = dalloc(b, vb, c->bytes))) return r;
    for (auto &b : c->A) if ((r = dalloc(b, vb, c->bytes))) return r;
    for (auto &b : c->med) if ((r = dalloc(b, (size_t)L.vol * 4, c->bytes))) return r;
    if ((r = dalloc(c->surf, (size_t)d.nx * 16, c->bytes))) return r;
    if ((r = dalloc(c->kob, (size_t)d.nx * 4, c->bytes))) return r;
    if ((r = dalloc(c->vmax, 64, c->bytes))) return r;
    if ((r = dalloc(c->dit, 64, c->bytes))) return r;
    const int ilo = d.na + 1, ihi = d.nx - d.na;
    c->ti0 = 1 + ((ilo - 1) / 64) * 64;
    c->ntx = ihi >= ilo ? (ihi - c->ti0) / 64 + 1 : 0;
    const int nk = d.nz - d.na;
    c->ntz = nk > 0 ? (nk + 3) / 4 : 0;
    Boxes2 &B = c->boxes;
    memset(&B, 0, sizeof B);
    auto add = [&](int i0, int i1, int k0, int k1) {
        int ni = i1 - i0 + 1, nk = k1 - k0 + 1;
        if (ni <= 0 || nk <= 0) return;
        Box2 &x = B.b[B.n++];
        x.i0 = i0; x.k0 = k0; x.ni = ni; x.nk = nk;
        x.start = B.total;
        B.total += (unsigned)ni * (unsigned)nk;
    };
    if (d.na > 0) {
        add(1, d.na, 1, d.nz);
        add(d.nx - d.na + 1, d.nx, 1, d.nz);
        add(d.na + 1, d.nx - d.na, d.nz - d.na + 1, d.nz);
    }
    *out = c.release();
    return SWPC_OK;
}

int swpc2d_destroy(swpc2d_ctx *c)
{
    if (!c) return SWPC_OK;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    for (auto &b : c->V) dfree(b);
    for (auto &b : c->S) dfree(b);
    for (auto &b : c->R) dfree(b);
    for (auto &b : c->A) dfree(b);
    for (auto &b : c->med) dfree(b);
    graph_drop2(c);
    DBuf *rest[] = {&c->surf, &c->kob, &c->gx, &c->gz, &c->src, &c->m3, &c->sdrop, &c->bsrc, &c->f2,
                    &c->fdt, &c->st, &c->uacc, &c->wvel, &c->wdisp, &c->vmax, &c->dit};
    for (auto b : rest) dfree(*b);
    for (auto &sn : c->snaps) {
        dfree(sn.pts);
        dfree(sn.acc);
        dfree(sn.frames);
    }
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return SWPC_OK;
}

int swpc2d_set_medium(swpc2d_ctx *c, const float *rho, const float *lam, const float *mu, const float *taup,
                      const float *taus)
{
    clear_error();
    if (c) graph_drop2(c);  // the captured steps would launch stale arguments
    if (!c || !rho || !lam || !mu || !taup || !taus) return fail(SWPC_EINVAL, "null argument");
    const float *in[5] = {rho, lam, mu, taup, taus};
    SWPC_HIP(hipStreamSynchronize(c->stream));
    std::vector<float> h((size_t)c->L.vol, 0.0f);
    for (int q = 0; q < 5; q++) {
        box2_to_dev(c->L, in[q], h.data());
        SWPC_HIP(hipMemcpy(c->med[q].p, h.data(), h.size() * 4, hipMemcpyHostToDevice));
    }
    c->have_medium = true;
    return SWPC_OK;
}

int swpc2d_set_surface(swpc2d_ctx *c, const int *kfs_top, const int *kfs_bot, const int *kob_top, const int *kob_bot,
                       const int *kob)
{
    clear_error();
    if (c) graph_drop2(c);  // the captured steps would launch stale arguments
    if (!c || !kfs_top || !kfs_bot || !kob_top || !kob_bot || !kob) return fail(SWPC_EINVAL, "null argument");
    const int nx = c->d.nx;
    std::vector<int> s4((size_t)nx * 4), kb(nx);
    for (int i = 1; i <= nx; i++) {
        size_t h = (size_t)(i + 2);
        s4[4 * (i - 1) + 0] = kfs_top[h];
        s4[4 * (i - 1) + 1] = kfs_bot[h];
        s4[4 * (i - 1) + 2] = kob_top[h];
        s4[4 * (i - 1) + 3] = kob_bot[h];
        kb[i - 1] = kob[h];
        if (kob[h] + 1 < -2 || kob[h] + 1 > c->d.nz + 3) return fail(SWPC_EINVAL, "kob(%d)=%d outside the memory box", i, kob[h]);
    }
    SWPC_HIP(hipStreamSynchronize(c->stream));
    SWPC_HIP(hipMemcpy(c->surf.p, s4.data(), s4.size() * 4, hipMemcpyHostToDevice));
    SWPC_HIP(hipMemcpy(c->kob.p, kb.data(), kb.size() * 4, hipMemcpyHostToDevice));
    c->have_surface = true;
    return SWPC_OK;
}

int swpc2d_set_pml(swpc2d_ctx *c, const float *gxc, const float *gxe, const float *gzc, const float *gze)
{
    clear_error();
    if (c) graph_drop2(c);  // the captured steps would launch stale arguments
    if (!c || !gxc || !gxe || !gzc || !gze) return fail(SWPC_EINVAL, "null argument");
    SWPC_HIP(hipStreamSynchronize(c->stream));
    int r;
    if ((r = upload_profile2(c, c->gx, gxc, gxe, c->d.nx))) return r;
    if ((r = upload_profile2(c, c->gz, gzc, gze, c->d.nz))) return r;
    c->have_pml = true;
    return SWPC_OK;
}

static int pack_src2(swpc2d_ctx *c, int nsrc, const int *is, const int *ks, std::vector<int> &pack)
{
    pack.resize((size_t)2 * nsrc);
    for (int n = 0; n < nsrc; n++) {
        if (is[n] < -1 || is[n] > c->d.nx + 3 || ks[n] < -1 || ks[n] > c->d.nz + 3)
            return fail(SWPC_EINVAL, "source %d at (%d,%d) outside the memory box", n + 1, is[n], ks[n]);
        pack[2 * n] = is[n];
        pack[2 * n + 1] = ks[n];
    }
    return SWPC_OK;
}

// The fused step kernels apply a source's share in the threads of the
// updated cells [1, nx] x [1, nz]: a few sources whose footprint (the
// cell and its -x, -z neighbours) lies there.
static bool fusable2(const swpc2d_ctx *c, int nsrc, const std::vector<int> &pack)
{
    if (nsrc > MAXFUSE) return false;
    for (int n = 0; n < nsrc; n++) {
        const int i = pack[2 * n], k = pack[2 * n + 1];
        if (i < 2 || i > c->d.nx || k < 2 || k > c->d.nz) return false;
    }
    return true;
}

int swpc2d_set_sources(swpc2d_ctx *c, int nsrc, const int *isrc, const int *ksrc, const double *m3, int nt,
                       const double *sdrop)
{
    clear_error();
    if (c) graph_drop2(c);  // the captured steps would launch stale arguments
    if (!c || nsrc < 0 || nt < 0) return fail(SWPC_EINVAL, "bad argument");
    if (nsrc > 0 && (!isrc || !ksrc || !m3 || !sdrop || nt < 1)) return fail(SWPC_EINVAL, "null source array");
    std::vector<int> pack;
    if (int r = pack_src2(c, nsrc, isrc, ksrc, pack)) return r;
    SWPC_HIP(hipStreamSynchronize(c->stream));
    int r;
    if ((r = dalloc(c->src, pack.size() * 4, c->bytes))) return r;
    if ((r = dalloc(c->m3, (size_t)3 * nsrc * 8, c->bytes))) return r;
    if ((r = dalloc(c->sdrop, (size_t)nt * nsrc * 8, c->bytes))) return r;
    if (nsrc > 0) {
        SWPC_HIP(hipMemcpy(c->src.p, pack.data(), pack.size() * 4, hipMemcpyHostToDevice));
        SWPC_HIP(hipMemcpy(c->m3.p, m3, (size_t)3 * nsrc * 8, hipMemcpyHostToDevice));
        SWPC_HIP(hipMemcpy(c->sdrop.p, sdrop, (size_t)nt * nsrc * 8, hipMemcpyHostToDevice));
    }
    c->nsrc = nsrc;
    c->nt_src = nt;
    c->fuse_src = fusable2(c, nsrc, pack);
    return SWPC_OK;
}

int swpc2d_set_bodyforce(swpc2d_ctx *c, int nsrc, const int *isrc, const int *ksrc, const double *f2, int nt,
                         const double *fdt)
{
    clear_error();
    if (c) graph_drop2(c);  // the captured steps would launch stale arguments
    if (!c || nsrc < 0 || nt < 0) return fail(SWPC_EINVAL, "bad argument");
    if (nsrc > 0 && (!isrc || !ksrc || !f2 || !fdt || nt < 1)) return fail(SWPC_EINVAL, "null source array");
    std::vector<int> pack;
    if (int r = pack_src2(c, nsrc, isrc, ksrc, pack)) return r;
    SWPC_HIP(hipStreamSynchronize(c->stream));
    int r;
    if ((r = dalloc(c->bsrc, pack.size() * 4, c->bytes))) return r;
    if ((r = dalloc(c->f2, (size_t)2 * nsrc * 8, c->bytes))) return r;
    if ((r = dalloc(c->fdt, (size_t)nt * nsrc * 8, c->bytes))) return r;
    if (nsrc > 0) {
        SWPC_HIP(hipMemcpy(c->bsrc.p, pack.data(), pack.size() * 4, hipMemcpyHostToDevice));
        SWPC_HIP(hipMemcpy(c->f2.p, f2, (size_t)2 * nsrc * 8, hipMemcpyHostToDevice));
        SWPC_HIP(hipMemcpy(c->fdt.p, fdt, (size_t)nt * nsrc * 8, hipMemcpyHostToDevice));
    }
    c->nbf = nsrc;
    c->nt_bf = nt;
    c->fuse_bf = fusable2(c, nsrc, pack);
    return SWPC_OK;
}

int swpc2d_set_receivers(swpc2d_ctx *c, int nst, const int *ist, const int *kst, int ntdec_w, int ntw)
{
    clear_error();
    if (c) graph_drop2(c);  // the captured steps would launch stale arguments
    if (!c || nst < 0 || ntdec_w < 1 || ntw < 0) return fail(SWPC_EINVAL, "bad argument");
    if (nst > 0 && (!ist || !kst)) return fail(SWPC_EINVAL, "null station array");
    std::vector<int> pack((size_t)2 * nst);
    for (int n = 0; n < nst; n++) {
        if (ist[n] < 1 || ist[n] > c->d.nx || kst[n] < 1 || kst[n] > c->d.nz)
            return fail(SWPC_EINVAL, "station %d at (%d,%d) outside the grid", n + 1, ist[n], kst[n]);
        pack[2 * n] = ist[n];
        pack[2 * n + 1] = kst[n];
    }
    SWPC_HIP(hipStreamSynchronize(c->stream));
    int r;
    if ((r = dalloc(c->st, pack.size() * 4, c->bytes))) return r;
    if ((r = dalloc(c->uacc, (size_t)2 * nst * 8, c->bytes))) return r;
    if ((r = dalloc(c->wvel, (size_t)2 * nst * std::max(ntw, 1) * 4, c->bytes))) return r;
    if ((r = dalloc(c->wdisp, (size_t)2 * nst * std::max(ntw, 1) * 8, c->bytes))) return r;
    if (nst > 0) SWPC_HIP(hipMemcpy(c->st.p, pack.data(), pack.size() * 4, hipMemcpyHostToDevice));
    c->nst = nst;
    c->ntdec_w = ntdec_w;
    c->ntw = ntw;
    return SWPC_OK;
}

int swpc2d_update_stress(swpc2d_ctx *c)
{
    clear_error();
    if (int r = ready2(c)) return r;
    return fp64(c) ? l2_stress<double>(c) : l2_stress<float>(c);
}

int swpc2d_update_vel(swpc2d_ctx *c)
{
    clear_error();
    if (int r = ready2(c)) return r;
    return fp64(c) ? l2_vel<double>(c) : l2_vel<float>(c);
}

int swpc2d_stressglut(swpc2d_ctx *c, int it)
{
    clear_error();
    if (int r = ready2(c)) return r;
    if (c->nsrc > 0 && (it < 1 || it > c->nt_src)) return fail(SWPC_EINVAL, "stressglut: it=%d outside 1..%d", it, c->nt_src);
    if (int r = set_step2(c, it)) return r;
    return fp64(c) ? l2_glut<double>(c) : l2_glut<float>(c);
}

int swpc2d_bodyforce(swpc2d_ctx *c, int it)
{
    clear_error();
    if (int r = ready2(c)) return r;
    if (c->nbf > 0 && (it < 1 || it > c->nt_bf)) return fail(SWPC_EINVAL, "bodyforce: it=%d outside 1..%d", it, c->nt_bf);
    if (int r = set_step2(c, it)) return r;
    return fp64(c) ? l2_bf<double>(c) : l2_bf<float>(c);
}

int swpc2d_store_receivers(swpc2d_ctx *c, int it)
{
    clear_error();
    if (int r = ready2(c)) return r;
    if (it < 1) return fail(SWPC_EINVAL, "it must be >= 1");
    if (int r = set_step2(c, it)) return r;
    return fp64(c) ? l2_recv<double>(c) : l2_recv<float>(c);
}

int swpc2d_add_snapshot(swpc2d_ctx *c, int kind, int npts, const int *ik, int ntdec_s, int nbuf, float uc, float m0,
                        float dt, int *id)
{
    clear_error();
    if (!c || !ik || !id) return fail(SWPC_EINVAL, "null argument");
    if (kind < SWPC_SNAP_PS || kind > SWPC_SNAP_U) return fail(SWPC_EINVAL, "unknown snapshot kind %d", kind);
    if (npts < 1 || ntdec_s < 1 || nbuf < 1) return fail(SWPC_EINVAL, "npts, ntdec_s and nbuf must be >= 1");
    if ((long long)2 * npts * nbuf >= (1LL << 31)) return fail(SWPC_EINVAL, "snapshot store too large");
    std::vector<int> p2((size_t)2 * npts);
    for (int n = 0; n < npts; n++) {
        const int i = ik[2 * n], k = ik[2 * n + 1];
        if (i < -1 || i > c->d.nx + 2 || k < -1 || k > c->d.nz + 2)
            return fail(SWPC_EINVAL, "snapshot point %d at (%d,%d) outside the grid", n, i, k);
        p2[2 * n] = i;
        p2[2 * n + 1] = k;
    }
    SWPC_HIP(hipStreamSynchronize(c->stream));
    swpc2d_ctx::Snap sn;
    sn.kind = kind; sn.npts = npts; sn.ntdec = ntdec_s; sn.nbuf = nbuf; sn.uc = uc; sn.m0 = m0; sn.dt = dt;
    int r;
    if ((r = dalloc(sn.pts, p2.size() * 4, c->bytes))) return r;
    if ((r = dalloc(sn.frames, (size_t)2 * npts * nbuf * 4, c->bytes))) { dfree(sn.pts); return r; }
    if (kind == SWPC_SNAP_U && (r = dalloc(sn.acc, (size_t)2 * npts * 4, c->bytes))) {
        dfree(sn.pts);
        dfree(sn.frames);
        return r;
    }
    SWPC_HIP(hipMemcpy(sn.pts.p, p2.data(), p2.size() * 4, hipMemcpyHostToDevice));
    c->snaps.push_back(sn);
    *id = (int)c->snaps.size() - 1;
    graph_drop2(c);
    return SWPC_OK;
}

int swpc2d_snapshots(swpc2d_ctx *c, int it)
{
    clear_error();
    if (int r = ready2(c)) return r;
    if (it < 1) return fail(SWPC_EINVAL, "it must be >= 1");
    if (int r = set_step2(c, it)) return r;
    if (int r = fp64(c) ? l2_snap<double>(c) : l2_snap<float>(c)) return r;
    note_snapshots2(c, it, it);
    return SWPC_OK;
}

int swpc2d_get_snapshot(swpc2d_ctx *c, int id, int frame, float *out)
{
    clear_error();
    if (!c || !out) return fail(SWPC_EINVAL, "null argument");
    if (id < 0 || id >= (int)c->snaps.size()) return fail(SWPC_EINVAL, "unknown snapshot id %d", id);
    const auto &sn = c->snaps[id];
    if (frame < 0 || frame > sn.last || frame <= sn.last - sn.nbuf)
        return fail(SWPC_EINVAL, "snapshot %d frame %d not held (latest %lld, %d buffered)", id, frame, sn.last,
                    sn.nbuf);
    const size_t n = (size_t)2 * sn.npts;
    SWPC_HIP(hipStreamSynchronize(c->stream));
    SWPC_HIP(hipMemcpy(out, (const float *)sn.frames.p + (size_t)(frame % sn.nbuf) * n, n * 4, hipMemcpyDeviceToHost));
    return SWPC_OK;
}

int swpc2d_step(swpc2d_ctx *c, int it)
{
    clear_error();
    if (int r = ready2(c)) return r;
    if (it < 1) return fail(SWPC_EINVAL, "it must be >= 1");
    return run_steps2(c, it, 1);
}

int swpc2d_run(swpc2d_ctx *c, int it0, int nsteps)
{
    clear_error();
    if (int r = ready2(c)) return r;
    if (it0 < 1 || nsteps < 0) return fail(SWPC_EINVAL, "bad step range");
    return run_steps2(c, it0, nsteps);
}

int swpc2d_set_fusion(swpc2d_ctx *c, int on)
{
    clear_error();
    if (!c) return fail(SWPC_EINVAL, "null context");
    graph_drop2(c);
    c->fused = on != 0;
    return SWPC_OK;
}

int swpc2d_set_graph(swpc2d_ctx *c, int steps)
{
    clear_error();
    if (!c) return fail(SWPC_EINVAL, "null context");
    if (steps < 0 || steps > 1024) return fail(SWPC_EINVAL, "steps per graph must be in 0..1024");
    graph_drop2(c);
    c->graph_k = steps;
    return SWPC_OK;
}

int swpc2d_get_receivers(swpc2d_ctx *c, float *vel, double *disp)
{
    clear_error();
    if (!c) return fail(SWPC_EINVAL, "null context");
    SWPC_HIP(hipStreamSynchronize(c->stream));
    size_t n = (size_t)2 * c->nst * c->ntw;
    if (n == 0) return SWPC_OK;
    if (vel) SWPC_HIP(hipMemcpy(vel, c->wvel.p, n * 4, hipMemcpyDeviceToHost));
    if (disp) SWPC_HIP(hipMemcpy(disp, c->wdisp.p, n * 8, hipMemcpyDeviceToHost));
    return SWPC_OK;
}

int swpc2d_get_field(swpc2d_ctx *c, int f, double *out)
{
    clear_error();
    if (!c || !out) return fail(SWPC_EINVAL, "null argument");
    DBuf *b = field_buf2(c, f);
    if (!b) return fail(SWPC_EINVAL, "unknown field %d", f);
    SWPC_HIP(hipStreamSynchronize(c->stream));
    if (fp64(c)) {
        std::vector<double> h((size_t)c->L.vol);
        SWPC_HIP(hipMemcpy(h.data(), b->p, h.size() * 8, hipMemcpyDeviceToHost));
        dev_to_box2(c->L, h.data(), out);
    } else {
        std::vector<float> h((size_t)c->L.vol);
        SWPC_HIP(hipMemcpy(h.data(), b->p, h.size() * 4, hipMemcpyDeviceToHost));
        dev_to_box2(c->L, h.data(), out);
    }
    return SWPC_OK;
}

int swpc2d_set_field(swpc2d_ctx *c, int f, const double *in)
{
    clear_error();
    if (c) graph_drop2(c);  // the captured steps would launch stale arguments
    if (!c || !in) return fail(SWPC_EINVAL, "null argument");
    DBuf *b = field_buf2(c, f);
    if (!b) return fail(SWPC_EINVAL, "unknown field %d", f);
    SWPC_HIP(hipStreamSynchronize(c->stream));
    if (fp64(c)) {
        std::vector<double> h((size_t)c->L.vol, 0.0);
        box2_to_dev(c->L, in, h.data());
        SWPC_HIP(hipMemcpy(b->p, h.data(), h.size() * 8, hipMemcpyHostToDevice));
    } else {
        std::vector<float> h((size_t)c->L.vol, 0.0f);
        box2_to_dev(c->L, in, h.data());
        SWPC_HIP(hipMemcpy(b->p, h.data(), h.size() * 4, hipMemcpyHostToDevice));
    }
    return SWPC_OK;
}

int swpc2d_vmax(swpc2d_ctx *c, float *out2)
{
    clear_error();
    if (int r = ready2(c)) return r;
    if (!out2) return fail(SWPC_EINVAL, "null argument");
    hipLaunchKernelGGL(k2_fill_u32, dim3(1), dim3(64), 0, c->stream, (unsigned *)c->vmax.p, 2, 0u);
    SWPC_CHECK_LAUNCH();
    int ni = c->d.nx - 2 * c->d.na;
    if (ni > 0) {
        if (fp64(c)) {
            hipLaunchKernelGGL((k2_vmax<double>), dim3((ni + 255) / 256), dim3(256), 0, c->stream, make_args2<double>(c),
                               (const int *)c->kob.p, (unsigned *)c->vmax.p);
        } else {
            hipLaunchKernelGGL((k2_vmax<float>), dim3((ni + 255) / 256), dim3(256), 0, c->stream, make_args2<float>(c),
                               (const int *)c->kob.p, (unsigned *)c->vmax.p);
        }
        SWPC_CHECK_LAUNCH();
    }
    unsigned u[2];
    SWPC_HIP(hipMemcpyAsync(u, c->vmax.p, 8, hipMemcpyDeviceToHost, c->stream));
    SWPC_HIP(hipStreamSynchronize(c->stream));
    memcpy(&out2[0], &u[0], 4);
    memcpy(&out2[1], &u[1], 4);
    return SWPC_OK;
}

int swpc2d_sync(swpc2d_ctx *c)
{
    clear_error();
    if (!c) return fail(SWPC_EINVAL, "null context");
    SWPC_HIP(hipStreamSynchronize(c->stream));
    return SWPC_OK;
}

}  // extern "C"
