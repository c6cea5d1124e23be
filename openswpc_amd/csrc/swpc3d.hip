// 3-D viscoelastic velocity-stress core (swpc_3d) for MI355X.
//
// Kernels
//   k3_stress_int  interior stress: 4th/2nd-order staggered derivatives of
//                  V and the generalized-Zener memory-variable update,
//                  normal and shear parts fused in one pass
//                  (src/swpc_3d/m_kernel.f90:142-326)
//   k3_vel_int     interior velocity (m_kernel.f90:75-140)
//   k3_stress_pml  ADE-CFS PML stress (m_absorb_p.f90:423-503)
//   k3_vel_pml     ADE-CFS PML velocity (m_absorb_p.f90:247-309)
//   k3_stressglut / k3_bodyforce / k3_receivers / k3_vmax
//
// The interior kernels give each thread one (i,j) column of a 64 x BY
// tile and stream it through z: the z-neighbours of the stencil sit in a
// register window that shifts by one plane per step, the x-neighbours
// come from the same wavefront's coalesced row and the y-neighbours from
// the rows of the same workgroup (L1/L2).  Every array is read and written
// once per cell, in 128-byte aligned rows.  The PML kernels enumerate the
// five boxes of the absorber region (two x faces, two y faces, the
// bottom slab) with one thread per cell.  There is no MFMA: the stencil
// is a gather, not a contraction, and is bound by HBM bandwidth.
#include "swpc_common.h"
#include "swpc_lab.h"
#include "swpc_batch.h"
#include "swpc_medium.h"

#include <hip/hip_bf16.h>
#include <hip/hip_fp16.h>

#include <algorithm>
#include <cmath>
#include <memory>

namespace swpc {

constexpr int NM3 = 3;  // Zener bodies; the reference fixes NM = 3 (m_global.f90:31)

struct Box3 {
    int i0, j0, k0, ni, nj, nk;
    unsigned start;  // first flat cell id of this box
};

struct Boxes3 {
    Box3 b[5];
    int n;
    unsigned total;
};

// The 18 memory variables, and apart from them the 6 stresses, are
// interleaved by tiles of 256 cells (tix): a wave's four cells per lane
// then read and write one region of 18 KiB (R) or 6 KiB (S) instead of
// arrays far apart in HBM.  The stress update takes 2247 instead of
// 2369 us at 512 x 512 x 256 with R tiled, and 2194 with S tiled too.
// K3::R[q] and STile::p point at component q's slot of tile 0.
__host__ __device__ __forceinline__ long long rix(int c) { return tix<NR3>(c); }

// FP16 memory variables (reduced storage) in their tiles: the pairs (0,1) ..
// (6,7) of the normal components and (9,10) .. (15,16) of the shear ones
// interleaved by groups of 4 cells, so that one 16-byte access moves a
// lane's four cells of both components of a pair; 8 and 17 follow the
// eight pairs, each contiguous.  The stress update takes 1563 instead of
// 1652 us at 512 x 512 x 256 with the pairs (18 8-byte accesses become
// 8 16-byte and 2 8-byte ones).  Element of component q at cell c,
// counted from the region's start:
__host__ __device__ __forceinline__ long long r16ix(int q, int c)
{
    const long long tb = (long long)((unsigned)(c + RB) >> 8) * (NR3 * RTILE);
    const int s = (c + RB) & (RTILE - 1);
    if (q == 8 || q == 17) return tb + (q == 8 ? 16 : 17) * RTILE + s;
    const int k = q < 8 ? q : q - 9;   // 0..7 within the normal or shear half
    const int p = (q < 8 ? 0 : 4) + (k >> 1);
    return tb + (long long)p * 2 * RTILE + (s >> 2) * 8 + (k & 1) * 4 + (s & 3);
}

// Global-memory address space: every state array lives there, and
// accesses that say so compile to global_* instructions also where the
// compiler cannot trace a pointer back to a kernel argument.
#define GLB __attribute__((address_space(1)))

// p as a global-memory pointer
template <typename T>
__device__ __forceinline__ GLB T *gp(T *p) { return (GLB T *)p; }

// Cell c of one stress component in the tiled layout: a cursor that moves
// by cells (p + c - 2 is cell c - 2, wherever its tile lies) and becomes
// an address only where it is loaded or stored (ld4, ld2, ld4nt, st4m,
// at()).  It has no conversion to a pointer, so an address cannot be
// offset by cells after the tile mapping by mistake.
template <typename T>
struct SCell {
    T *p;
    int c;
    __device__ __forceinline__ SCell operator+(int d) const { return {p, c + d}; }
    __device__ __forceinline__ SCell operator-(int d) const { return {p, c - d}; }
    __device__ __forceinline__ T *at() const { return p + tix<NS3>(c); }
};

// One stress component: s[c] is its value at cell c, s + c its cursor.
template <typename T>
struct STile {
    T *p;
    __device__ __forceinline__ GLB T &operator[](int c) const { return ((GLB T *)p)[tix<NS3>(c)]; }
    __device__ __forceinline__ SCell<T> operator+(int c) const { return {p, c}; }
};

template <typename T>
struct K3 {
    int px, sxy, nx, ny, nz, na;
    T *V[3];    // Vx Vy Vz
    STile<T> S[6];  // Sxx Syy Szz Syz Sxz Sxy, tiled
    T *R[18];   // Rxx(m) Ryy(m) Rzz(m) Ryz(m) Rxz(m) Rxy(m), m = 0..2
    T *A[18];   // ADE arrays (order of enum swpc3d_field)
    const float *med[10];   // bx by bz lam mu taup taus muyz muxz muxy (the reference's real(SP) arrays)
    const float *rho;       // density (FP32 box) when set: the 4-wide velocity kernel forms bx, by, bz itself
    int m16;                // medium stored as 16-bit residuals (SWPC_FP32_S16); else FP32 arrays
    // reduced storage (SWPC_FP32_S16): FP32 arithmetic, V and S in FP32
    const __half *hmed[10];     // medium: FP16 residual about a per-plane FP32 base
    const float2 *mbs[10];      // per-plane (base, scale), indexed k + KO
    // FP32 medium: per-plane value of an array that is constant over the
    // plane's cells i, j in [1, n] (NaN where the plane varies), indexed
    // k + KO; null when no plane of the array is constant
    const float *cpl[10];
    __half *hR;                 // memory variables stored as R * r_inv, in the tiles' pair layout (r16ix)
    float r_scale, r_inv;
    int sr;                     // 16-bit stores round stochastically (else to nearest)
    __hip_bfloat16 *hA[18];     // ADE arrays of the PML
    const int4 *surf;   // (kfs_top, kfs_bot, kob_top, kob_bot) per column (i,j) in [1,nx]x[1,ny]
    const T *gx, *gy, *gz;  // profiles [8][n]: c1..c4 at centres then e1..e4 at edges
    const float *gxv;       // x profiles for the 4-wide PML: [8][gxp], 16-byte aligned rows
    int gxp;
    T e4[3][2], e2[3][2];   // (re40, re41) for isign = -1 (4th) and +1 (2nd), per axis
    T r2[3];                // 1/dx, 1/dy, 1/dz
    T c1[NM3], c2[NM3], d1[NM3], d2;
    T dt;
    int ti0;  // i of the first interior tile
    int kc;   // z-chunk length of the interior kernels
    int ntx, nty, nkc;          // interior tiles along x, y and z chunks
    unsigned nblk, nblk_pad;    // ntx*nty*nkc and its round-up to a multiple of 8
    int xcd;                    // workgroup order: 0 as launched, 1 XCD-contiguous, >= 2 y-groups (velz_blk)
    int ade_sparse;             // ADE arrays hold zeros wherever their axis is undamped
    Boxes3 pml;
};

enum { MBX = 0, MBY, MBZ, MLAM, MMU, MTAUP, MTAUS, MMUYZ, MMUXZ, MMUXY };

// Medium value at cell c of plane k.  Reduced storage reconstructs
// base(k) + scale(k) * q; a plane whose residual is identically zero
// (a layered medium) is not read at all.
template <bool S16, typename T>
__device__ __forceinline__ T medv(const K3<T> &a, int q, int c, int k)
{
    if constexpr (S16) {
        if (!a.m16) return (T)gp(a.med[q])[c];
        const float2 bs = a.mbs[q][k + KO];
        float v = bs.x;
        if (bs.y != 0.0f) v += bs.y * __half2float(a.hmed[q][c]);
        return (T)v;
    } else {
        return (T)gp(a.med[q])[c];
    }
}

// Stochastic rounding of the 16-bit stores: the dropped low bits of the
// FP32 value are compared with a hashed threshold, so a store rounds up
// with probability equal to the dropped fraction.  Small increments of a
// slowly varying accumulator (memory variables with c1 close to 1, ADE
// arrays) then survive on average instead of being rounded away every
// step.  The hash depends on the cell, the array and the value, so runs
// stay deterministic and graph replays identical.
__device__ __forceinline__ unsigned sr_hash(unsigned c, unsigned q, unsigned vb)
{
    unsigned x = c * 0x9E3779B1u ^ (q << 24) ^ vb;
    x ^= x >> 16;
    x *= 0x7FEB352Du;
    x ^= x >> 15;
    x *= 0x846CA68Bu;
    x ^= x >> 16;
    return x;
}

// FP32 -> FP16 (finite, within +-65504 after scaling), stochastic
__device__ __forceinline__ __half f2h_sr(float v, unsigned c, unsigned q)
{
    const unsigned vb = __float_as_uint(v);
    return __float2half_rz(__uint_as_float(vb + (sr_hash(c, q, vb) & 0x1FFFu)));
}

// FP32 -> BF16 bits, stochastic
__device__ __forceinline__ unsigned short f2bf_sr(float v, unsigned c, unsigned q)
{
    const unsigned vb = __float_as_uint(v);
    return (unsigned short)((vb + (sr_hash(c, q, vb) & 0xFFFFu)) >> 16);
}


template <bool S16, typename T>
__device__ __forceinline__ T ldR(const K3<T> &a, int q, int c)
{
    if constexpr (S16) return (T)(__half2float(a.hR[r16ix(q, c)]) * a.r_scale);
    else return gp(a.R[q])[rix(c)];
}

template <bool S16, typename T>
__device__ __forceinline__ void stR(const K3<T> &a, int q, int c, T v)
{
    if constexpr (S16) {
        const float w = fminf(fmaxf((float)v * a.r_inv, -65504.0f), 65504.0f);
        a.hR[r16ix(q, c)] = a.sr ? f2h_sr(w, (unsigned)c, (unsigned)q) : __float2half_rn(w);
    } else {
        gp(a.R[q])[rix(c)] = v;
    }
}

template <bool S16, typename T>
__device__ __forceinline__ T ldA(const K3<T> &a, int q, int c)
{
    if constexpr (S16) return (T)__bfloat162float(a.hA[q][c]);
    else return gp(a.A[q])[c];
}

template <bool S16, typename T>
__device__ __forceinline__ void stA(const K3<T> &a, int q, int c, T v)
{
    if constexpr (S16) {
        if (a.sr) {
            const unsigned short b = f2bf_sr((float)v, (unsigned)c, 32u + (unsigned)q);
            a.hA[q][c] = *(const __hip_bfloat16 *)&b;
        } else {
            a.hA[q][c] = __float2bfloat16((float)v);
        }
    } else {
        gp(a.A[q])[c] = v;
    }
}

// ---------------------------------------------------------------------------
// interior tiles
// ---------------------------------------------------------------------------
// The interior box is cut into 64 x BY column tiles and nkc z-chunks,
// enumerated by one linear workgroup id.  Workgroups are dealt to the 8
// XCDs round-robin; with xcd != 0 the id is remapped so that each XCD owns
// one contiguous run of tiles and its L2 holds the y-neighbour rows that
// adjacent tiles share.
template <typename T>
__device__ inline bool tile_of(const K3<T> &a, int &bx, int &by, int &bz)
{
    const unsigned b = blockIdx.x;
    unsigned L = b;
    if (a.xcd) {
        const unsigned per = a.nblk_pad >> 3;
        L = (b & 7u) * per + (b >> 3);
    }
    if (L >= a.nblk) return false;
    bx = (int)(L % (unsigned)a.ntx);
    const unsigned t = L / (unsigned)a.ntx;
    by = (int)(t % (unsigned)a.nty);
    bz = (int)(t / (unsigned)a.nty);
    return true;
}

// ---------------------------------------------------------------------------
// interior stress
// ---------------------------------------------------------------------------
// PART 0: normal and shear fused (V and taus read once per cell);
// PART 1: normal only; PART 2: shear only (the reference's two loops).
template <typename T, int BY, int PART, bool S16>
__device__ __forceinline__ void stress_column(const K3<T> &a, int bx, int by, int bz)
{
    const int i = a.ti0 + bx * 64 + (int)threadIdx.x;
    const int j = a.na + 1 + by * BY + (int)threadIdx.y;
    if (i <= a.na || i > a.nx - a.na || j > a.ny - a.na) return;
    const int kb = 1 + bz * a.kc;
    const int ke = min(kb + a.kc - 1, a.nz - a.na);
    const int4 fs = a.surf[(j - 1) * a.nx + (i - 1)];
    const int px = a.px, sxy = a.sxy;

    const T *__restrict__ Vx = a.V[0];
    const T *__restrict__ Vy = a.V[1];
    const T *__restrict__ Vz = a.V[2];
    const STile<T> Sxx = a.S[0];
    const STile<T> Syy = a.S[1];
    const STile<T> Szz = a.S[2];
    const STile<T> Syz = a.S[3];
    const STile<T> Sxz = a.S[4];
    const STile<T> Sxy = a.S[5];
    constexpr bool NORMAL = PART != 2, SHEAR = PART != 1;

    int c = at3(px, sxy, i, j, kb);
    T vxm1 = 0, vx0 = Vx[c], vxp1 = 0;
    T vym1 = 0, vy0 = Vy[c], vyp1 = 0;
    T vzm2 = 0, vzm1 = Vz[c - sxy], vz0 = Vz[c];
    if (SHEAR) {
        vxm1 = Vx[c - sxy]; vxp1 = Vx[c + sxy];
        vym1 = Vy[c - sxy]; vyp1 = Vy[c + sxy];
    }
    if (NORMAL) vzm2 = Vz[c - 2 * sxy];
    const T dt = a.dt;

    for (int k = kb; k <= ke; ++k, c += sxy) {
        T vxp2 = 0, vyp2 = 0;
        if (SHEAR) {
            vxp2 = Vx[c + 2 * sxy];
            vyp2 = Vy[c + 2 * sxy];
        }
        const T vzp1 = Vz[c + sxy];

        // order switch of m_kernel.f90:185-186
        const int o1 = (k - fs.x) * (fs.y - k), o2 = (k - fs.z) * (fs.w - k);
        const bool lo = max(o1, o2) >= 0;
        const T e0x = lo ? a.e2[0][0] : a.e4[0][0], e1x = lo ? a.e2[0][1] : a.e4[0][1];
        const T e0y = lo ? a.e2[1][0] : a.e4[1][0], e1y = lo ? a.e2[1][1] : a.e4[1][1];
        const T e0z = lo ? a.e2[2][0] : a.e4[2][0], e1z = lo ? a.e2[2][1] : a.e4[2][1];
        const T taus1 = medv<S16>(a, MTAUS, c, k);
        const T taus_plus1 = 1 + taus1 * (1 + a.d2);

        if (NORMAL) {
            // m_kernel.f90:195-235
            const T dxVx = (vx0 - Vx[c - 1]) * e0x - (Vx[c + 1] - Vx[c - 2]) * e1x;
            const T dyVy = (vy0 - Vy[c - px]) * e0y - (Vy[c + px] - Vy[c - 2 * px]) * e1y;
            const T dzVz = (vz0 - vzm1) * e0z - (vzp1 - vzm2) * e1z;
            const T mu2 = 2 * medv<S16>(a, MMU, c, k);
            const T lam2mu = medv<S16>(a, MLAM, c, k) + mu2;
            const T taup1 = medv<S16>(a, MTAUP, c, k);
            const T d3v3 = dxVx + dyVy + dzVz;
            const T dyVy_dzVz = dyVy + dzVz, dxVx_dzVz = dxVx + dzVz, dxVx_dyVy = dxVx + dyVy;
            const T fxx = lam2mu * taup1 * d3v3 - mu2 * taus1 * dyVy_dzVz;
            const T fyy = lam2mu * taup1 * d3v3 - mu2 * taus1 * dxVx_dzVz;
            const T fzz = lam2mu * taup1 * d3v3 - mu2 * taus1 * dxVx_dyVy;
            T Rxx_n = 0, Ryy_n = 0, Rzz_n = 0;
#pragma unroll
            for (int m = 0; m < NM3; m++) {
                T rxx = a.c1[m] * ldR<S16>(a, 0 + m, c) - a.c2[m] * fxx * dt;
                T ryy = a.c1[m] * ldR<S16>(a, 3 + m, c) - a.c2[m] * fyy * dt;
                T rzz = a.c1[m] * ldR<S16>(a, 6 + m, c) - a.c2[m] * fzz * dt;
                stR<S16>(a, 0 + m, c, rxx);
                stR<S16>(a, 3 + m, c, ryy);
                stR<S16>(a, 6 + m, c, rzz);
                Rxx_n += a.d1[m] * rxx;
                Ryy_n += a.d1[m] * ryy;
                Rzz_n += a.d1[m] * rzz;
            }
            const T taup_plus1 = 1 + taup1 * (1 + a.d2);
            Sxx[c] = Sxx[c] + (lam2mu * taup_plus1 * d3v3 - mu2 * taus_plus1 * dyVy_dzVz + Rxx_n) * dt;
            Syy[c] = Syy[c] + (lam2mu * taup_plus1 * d3v3 - mu2 * taus_plus1 * dxVx_dzVz + Ryy_n) * dt;
            Szz[c] = Szz[c] + (lam2mu * taup_plus1 * d3v3 - mu2 * taus_plus1 * dxVx_dyVy + Rzz_n) * dt;
        }
        if (SHEAR) {
            // m_kernel.f90:282-311
            const T dxVy_dyVx = (Vy[c + 1] - vy0) * e0x - (Vy[c + 2] - Vy[c - 1]) * e1x
                              + (Vx[c + px] - vx0) * e0y - (Vx[c + 2 * px] - Vx[c - px]) * e1y;
            const T dxVz_dzVx = (Vz[c + 1] - vz0) * e0x - (Vz[c + 2] - Vz[c - 1]) * e1x
                              + (vxp1 - vx0) * e0z - (vxp2 - vxm1) * e1z;
            const T dyVz_dzVy = (Vz[c + px] - vz0) * e0y - (Vz[c + 2 * px] - Vz[c - px]) * e1y
                              + (vyp1 - vy0) * e0z - (vyp2 - vym1) * e1z;
            const T myz = medv<S16>(a, MMUYZ, c, k), mxz = medv<S16>(a, MMUXZ, c, k), mxy = medv<S16>(a, MMUXY, c, k);
            const T fyz = myz * taus1 * dyVz_dzVy;
            const T fxz = mxz * taus1 * dxVz_dzVx;
            const T fxy = mxy * taus1 * dxVy_dyVx;
            T Ryz_n = 0, Rxz_n = 0, Rxy_n = 0;
#pragma unroll
            for (int m = 0; m < NM3; m++) {
                T ryz = a.c1[m] * ldR<S16>(a, 9 + m, c) - a.c2[m] * fyz * dt;
                T rxz = a.c1[m] * ldR<S16>(a, 12 + m, c) - a.c2[m] * fxz * dt;
                T rxy = a.c1[m] * ldR<S16>(a, 15 + m, c) - a.c2[m] * fxy * dt;
                stR<S16>(a, 9 + m, c, ryz);
                stR<S16>(a, 12 + m, c, rxz);
                stR<S16>(a, 15 + m, c, rxy);
                Ryz_n += a.d1[m] * ryz;
                Rxz_n += a.d1[m] * rxz;
                Rxy_n += a.d1[m] * rxy;
            }
            Syz[c] = Syz[c] + (myz * taus_plus1 * dyVz_dzVy + Ryz_n) * dt;
            Sxz[c] = Sxz[c] + (mxz * taus_plus1 * dxVz_dzVx + Rxz_n) * dt;
            Sxy[c] = Sxy[c] + (mxy * taus_plus1 * dxVy_dyVx + Rxy_n) * dt;
        }

        vxm1 = vx0; vx0 = vxp1; vxp1 = vxp2;
        vym1 = vy0; vy0 = vyp1; vyp1 = vyp2;
        vzm2 = vzm1; vzm1 = vz0; vz0 = vzp1;
        if (!SHEAR) {
            // the normal part needs Vx and Vy only at the centre plane
            vx0 = Vx[c + sxy];
            vy0 = Vy[c + sxy];
        }
    }
}

// WPE > 0 asks the compiler for at least WPE waves per SIMD (a register
// budget of 512/WPE VGPRs).
template <typename T, int BY, int PART, int WPE, bool S16>
__global__ void __launch_bounds__(64 * BY) __attribute__((amdgpu_waves_per_eu(WPE > 0 ? WPE : 1, 8)))
k3_stress_int(const K3<T> a)
{
    int bx, by, bz;
    if (!tile_of(a, bx, by, bz)) return;
    stress_column<T, BY, PART, S16>(a, bx, by, bz);
}

// ---------------------------------------------------------------------------
// interior velocity
// ---------------------------------------------------------------------------
template <typename T, int BY, int WPE, bool S16>
__global__ void __launch_bounds__(64 * BY) __attribute__((amdgpu_waves_per_eu(WPE > 0 ? WPE : 1, 8)))
k3_vel_int(const K3<T> a)
{
    int bx, by, bz;
    if (!tile_of(a, bx, by, bz)) return;
    const int i = a.ti0 + bx * 64 + (int)threadIdx.x;
    const int j = a.na + 1 + by * BY + (int)threadIdx.y;
    if (i <= a.na || i > a.nx - a.na || j > a.ny - a.na) return;
    const int kb = 1 + bz * a.kc;
    const int ke = min(kb + a.kc - 1, a.nz - a.na);
    const int4 fs = a.surf[(j - 1) * a.nx + (i - 1)];
    const int px = a.px, sxy = a.sxy;

    T *__restrict__ Vx = a.V[0];
    T *__restrict__ Vy = a.V[1];
    T *__restrict__ Vz = a.V[2];
    const STile<T> Sxx = a.S[0];
    const STile<T> Syy = a.S[1];
    const STile<T> Szz = a.S[2];
    const STile<T> Syz = a.S[3];
    const STile<T> Sxz = a.S[4];
    const STile<T> Sxy = a.S[5];

    int c = at3(px, sxy, i, j, kb);
    T xzm2 = Sxz[c - 2 * sxy], xzm1 = Sxz[c - sxy], xz0 = Sxz[c];
    T yzm2 = Syz[c - 2 * sxy], yzm1 = Syz[c - sxy], yz0 = Syz[c];
    T zzm1 = Szz[c - sxy], zz0 = Szz[c], zzp1 = Szz[c + sxy];
    const T dt = a.dt;

    for (int k = kb; k <= ke; ++k, c += sxy) {
        const T xzp1 = Sxz[c + sxy];
        const T yzp1 = Syz[c + sxy];
        const T zzp2 = Szz[c + 2 * sxy];

        const int o1 = (k - fs.x) * (fs.y - k), o2 = (k - fs.z) * (fs.w - k);
        const bool lo = max(o1, o2) >= 0;
        const T e0x = lo ? a.e2[0][0] : a.e4[0][0], e1x = lo ? a.e2[0][1] : a.e4[0][1];
        const T e0y = lo ? a.e2[1][0] : a.e4[1][0], e1y = lo ? a.e2[1][1] : a.e4[1][1];
        const T e0z = lo ? a.e2[2][0] : a.e4[2][0], e1z = lo ? a.e2[2][1] : a.e4[2][1];

        const T sxy0 = Sxy[c];
        // m_kernel.f90:113-121
        const T d3Sx3 = (Sxx[c + 1] - Sxx[c]) * e0x - (Sxx[c + 2] - Sxx[c - 1]) * e1x
                      + (sxy0 - Sxy[c - px]) * e0y - (Sxy[c + px] - Sxy[c - 2 * px]) * e1y
                      + (xz0 - xzm1) * e0z - (xzp1 - xzm2) * e1z;
        const T d3Sy3 = (sxy0 - Sxy[c - 1]) * e0x - (Sxy[c + 1] - Sxy[c - 2]) * e1x
                      + (Syy[c + px] - Syy[c]) * e0y - (Syy[c + 2 * px] - Syy[c - px]) * e1y
                      + (yz0 - yzm1) * e0z - (yzp1 - yzm2) * e1z;
        const T d3Sz3 = (xz0 - Sxz[c - 1]) * e0x - (Sxz[c + 1] - Sxz[c - 2]) * e1x
                      + (yz0 - Syz[c - px]) * e0y - (Syz[c + px] - Syz[c - 2 * px]) * e1y
                      + (zzp1 - zz0) * e0z - (zzp2 - zzm1) * e1z;

        Vx[c] = Vx[c] + medv<S16>(a, MBX, c, k) * d3Sx3 * dt;
        Vy[c] = Vy[c] + medv<S16>(a, MBY, c, k) * d3Sy3 * dt;
        Vz[c] = Vz[c] + medv<S16>(a, MBZ, c, k) * d3Sz3 * dt;

        xzm2 = xzm1; xzm1 = xz0; xz0 = xzp1;
        yzm2 = yzm1; yzm1 = yz0; yz0 = yzp1;
        zzm1 = zz0; zz0 = zzp1; zzp1 = zzp2;
    }
}

// ---------------------------------------------------------------------------
// 4-wide interior kernels (FP32 arithmetic, FP32 or reduced storage)
// ---------------------------------------------------------------------------
// Each thread owns 4 consecutive x-cells of one (j, k) row, so every
// wavefront moves 1 KiB per streamed array with one 16-byte (non-temporal)
// access per lane; a wavefront covers a 256-cell x tile.  The grid is flat
// over k (one plane per workgroup row of tiles); the stencil's x
// neighbours come from 8-byte halo loads of the adjacent cells, y and z
// neighbours from 16-byte loads of the adjacent rows and planes, which the
// L2 and the Infinity Cache serve.  With xcd != 0 every XCD owns one
// contiguous band of y tiles and walks it plane by plane, so the planes
// k-2..k+2 that a band needs stay in that XCD's L2.
typedef float f4v __attribute__((ext_vector_type(4)));

// Vector accesses of state arrays, as global-memory instructions (GLB;
// flat_* would also wait on the LDS counter).
__device__ __forceinline__ f4v ld4(const float *p) { return *(const GLB f4v *)p; }
__device__ __forceinline__ f4v ld4nt(const float *p) { return __builtin_nontemporal_load((const GLB f4v *)p); }
__device__ __forceinline__ void st4nt(float *p, f4v v) { __builtin_nontemporal_store(v, (GLB f4v *)p); }
__device__ __forceinline__ float2 ld2(const float *p)
{
    typedef float f2v __attribute__((ext_vector_type(2)));
    const f2v v = *(const GLB f2v *)p;
    return make_float2(v.x, v.y);
}
__device__ __forceinline__ f4v ld4(SCell<float> s) { return ld4(s.at()); }
__device__ __forceinline__ f4v ld4nt(SCell<float> s) { return ld4nt(s.at()); }
__device__ __forceinline__ float2 ld2(SCell<float> s) { return ld2(s.at()); }
__device__ __forceinline__ float ld1(const float *p) { return *(const GLB float *)p; }
__device__ __forceinline__ float ld1(SCell<float> s) { return ld1(s.at()); }

typedef _Float16 h4v __attribute__((ext_vector_type(4)));

// four FP16 values (8 bytes, non-temporal) widened to FP32
__device__ __forceinline__ f4v h4(const __half *p)
{
    const h4v h = __builtin_nontemporal_load((const GLB h4v *)p);
    return __builtin_convertvector(h, f4v);
}

// four FP32 values rounded to nearest FP16 and stored non-temporally
__device__ __forceinline__ void sth4(__half *p, f4v v)
{
    const h4v h = __builtin_convertvector(v, h4v);
    __builtin_nontemporal_store(h, (GLB h4v *)p);
}

// medium value of 4 cells of plane k
template <bool S16>
__device__ __forceinline__ f4v med4(const K3<float> &a, int q, int c, int k)
{
    if constexpr (S16) {
        if (a.m16) {
            const float2 bs = a.mbs[q][k + KO];
            f4v v = {bs.x, bs.x, bs.x, bs.x};
            if (bs.y != 0.0f) v += bs.y * h4(a.hmed[q] + c);
            return v;
        }
    }
    if (a.cpl[q]) {   // a plane-constant array costs no stream
        const float v = a.cpl[q][k + KO];
        if (v == v) return f4v{v, v, v, v};
    }
    return ld4nt(a.med[q] + c);
}

template <bool S16>
__device__ __forceinline__ f4v ldR4(const K3<float> &a, int q, int c)
{
    if constexpr (S16) return h4(a.hR + r16ix(q, c)) * a.r_scale;
    else return ld4nt(a.R[q] + rix(c));
}

// masked stores: all four cells, or the valid ones one by one
__device__ __forceinline__ void st4m(float *q, f4v v, unsigned m)
{
    if (m == 15u) {
        st4nt(q, v);
    } else {
        GLB float *p = (GLB float *)q;
        if (m & 1u) p[0] = v.x;
        if (m & 2u) p[1] = v.y;
        if (m & 4u) p[2] = v.z;
        if (m & 8u) p[3] = v.w;
    }
}

__device__ __forceinline__ void st4m(SCell<float> s, f4v v, unsigned m) { st4m(s.at(), v, m); }

template <bool S16>
__device__ __forceinline__ void stR4(const K3<float> &a, int q, int c, f4v v, unsigned m)
{
    if constexpr (S16) {
        const float lim = 65504.0f;
        v *= a.r_inv;
        v.x = fminf(fmaxf(v.x, -lim), lim);
        v.y = fminf(fmaxf(v.y, -lim), lim);
        v.z = fminf(fmaxf(v.z, -lim), lim);
        v.w = fminf(fmaxf(v.w, -lim), lim);
        if (a.sr) {
            __half h[4] = {f2h_sr(v.x, (unsigned)c, (unsigned)q), f2h_sr(v.y, (unsigned)c + 1u, (unsigned)q),
                           f2h_sr(v.z, (unsigned)c + 2u, (unsigned)q), f2h_sr(v.w, (unsigned)c + 3u, (unsigned)q)};
            if (m == 15u) {
                h4v hv;
                for (int e = 0; e < 4; e++) hv[e] = *(const _Float16 *)&h[e];
                __builtin_nontemporal_store(hv, (GLB h4v *)(a.hR + r16ix(q, c)));
            } else {
                for (int e = 0; e < 4; e++)
                    if (m & (1u << e)) a.hR[r16ix(q, c) + e] = h[e];
            }
        } else if (m == 15u) {
            sth4(a.hR + r16ix(q, c), v);
        } else {
            __half *p = a.hR + r16ix(q, c);
            if (m & 1u) p[0] = __float2half_rn(v.x);
            if (m & 2u) p[1] = __float2half_rn(v.y);
            if (m & 4u) p[2] = __float2half_rn(v.z);
            if (m & 8u) p[3] = __float2half_rn(v.w);
        }
    } else {
        st4m(a.R[q] + rix(c), v, m);
    }
}

typedef _Float16 h8v __attribute__((ext_vector_type(8)));

// Both components of the FP16 pair (q, q + 1) at a lane's four cells (q the
// pair's first component, r16ix), as stored (ldR4/stR4 give the same values
// and stores one component at a time).
__device__ __forceinline__ void ldR4pair(const K3<float> &a, int q, int c, f4v &lo, f4v &hi)
{
    const h8v h = __builtin_nontemporal_load((const GLB h8v *)(a.hR + r16ix(q, c)));
    lo = f4v{(float)h[0], (float)h[1], (float)h[2], (float)h[3]} * a.r_scale;
    hi = f4v{(float)h[4], (float)h[5], (float)h[6], (float)h[7]} * a.r_scale;
}

__device__ __forceinline__ void stR4pair(const K3<float> &a, int q, int c, f4v lo, f4v hi, unsigned m)
{
    if (m != 15u || a.sr) {
        stR4<true>(a, q, c, lo, m);
        stR4<true>(a, q + 1, c, hi, m);
        return;
    }
    const float lim = 65504.0f;
    lo *= a.r_inv;
    hi *= a.r_inv;
    lo.x = fminf(fmaxf(lo.x, -lim), lim); lo.y = fminf(fmaxf(lo.y, -lim), lim);
    lo.z = fminf(fmaxf(lo.z, -lim), lim); lo.w = fminf(fmaxf(lo.w, -lim), lim);
    hi.x = fminf(fmaxf(hi.x, -lim), lim); hi.y = fminf(fmaxf(hi.y, -lim), lim);
    hi.z = fminf(fmaxf(hi.z, -lim), lim); hi.w = fminf(fmaxf(hi.w, -lim), lim);
    const h4v l = __builtin_convertvector(lo, h4v), u = __builtin_convertvector(hi, h4v);
    const h8v h = {l[0], l[1], l[2], l[3], u[0], u[1], u[2], u[3]};
    __builtin_nontemporal_store(h, (GLB h8v *)(a.hR + r16ix(q, c)));
}

// Tile decode of the 4-wide kernels: (workgroup b, thread) -> (i0, j, k)
// and the mask of its 4 cells inside the interior box.
__device__ __forceinline__ bool x4_cell(const K3<float> &a, int by, unsigned b, int &i0, int &j, int &k, unsigned &m)
{
    unsigned xt, yt, kk;
    if (a.xcd) {
        const unsigned nyc = ((unsigned)a.nty + 7u) >> 3;
        const unsigned loc = b >> 3;
        xt = loc % (unsigned)a.ntx;
        const unsigned t = loc / (unsigned)a.ntx;
        yt = (b & 7u) * nyc + t % nyc;
        kk = t / nyc;
        if (yt >= (unsigned)a.nty) return false;
    } else {
        xt = b % (unsigned)a.ntx;
        const unsigned t = b / (unsigned)a.ntx;
        yt = t % (unsigned)a.nty;
        kk = t / (unsigned)a.nty;
    }
    k = 1 + (int)kk;
    if (k > a.nz - a.na) return false;
    i0 = a.ti0 + (int)xt * 256 + (int)threadIdx.x * 4;
    j = a.na + 1 + (int)yt * by + (int)threadIdx.y;
    if (j > a.ny - a.na) return false;
    m = 0;
#pragma unroll
    for (int q = 0; q < 4; q++)
        if (i0 + q > a.na && i0 + q <= a.nx - a.na) m |= 1u << q;
    return m != 0;
}

struct Ord {
    float e0x, e1x, e0y, e1y, e0z, e1z;
};

__device__ __forceinline__ Ord order4(const K3<float> &a, int i, int j, int k)
{
    i = min(max(i, 1), a.nx);
    const int4 fs = a.surf[(j - 1) * a.nx + (i - 1)];
    const int o1 = (k - fs.x) * (fs.y - k), o2 = (k - fs.z) * (fs.w - k);
    const bool lo = max(o1, o2) >= 0;
    Ord o;
    o.e0x = lo ? a.e2[0][0] : a.e4[0][0];
    o.e1x = lo ? a.e2[0][1] : a.e4[0][1];
    o.e0y = lo ? a.e2[1][0] : a.e4[1][0];
    o.e1y = lo ? a.e2[1][1] : a.e4[1][1];
    o.e0z = lo ? a.e2[2][0] : a.e4[2][0];
    o.e1z = lo ? a.e2[2][1] : a.e4[2][1];
    return o;
}

// x neighbours by wavefront shifts: lane l holds cells i0 .. i0+3 and
// lane l - 1 the four before them, so the cells i0-2, i0-1 (left) and
// i0+4, i0+5 (right) a stencil needs come from the neighbour lanes'
// registers.  A lane whose neighbour lane holds no interior cell (it has
// returned) or lies outside the wave loads them instead (pass the loads
// in ldl / ldr).
struct XNb {
    bool l, r;  // left / right neighbour lane active
};

__device__ __forceinline__ XNb x4_nb(const K3<float> &a, int i0)
{
    const unsigned lane = threadIdx.x;
    XNb n;
    n.l = lane > 0u && i0 - 1 > a.na;
    n.r = lane < 63u && i0 + 4 <= a.nx - a.na;
    return n;
}

// The value of lane l - 1 (from_left) or l + 1 (from_right) by a DPP
// wavefront shift (wave_shr:1 / wave_shl:1): a VALU move, no LDS crossbar.
__device__ __forceinline__ float from_left(float v)
{
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x138, 0xF, 0xF, false));
}

__device__ __forceinline__ float from_right(float v)
{
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x130, 0xF, 0xF, false));
}

// cells i0-2, i0-1 of array p (v holds i0 .. i0+3 of this lane); p is a
// pointer or a tiled stress component
template <typename P>
__device__ __forceinline__ float2 x4_left(const XNb &n, const f4v v, const P &p, int c)
{
    const float z = from_left(v.z), w = from_left(v.w);
    if (n.l) return make_float2(z, w);
    return ld2(p + c - 2);
}

// cells i0+4, i0+5
template <typename P>
__device__ __forceinline__ float2 x4_right(const XNb &n, const f4v v, const P &p, int c)
{
    const float x = from_right(v.x), y = from_right(v.y);
    if (n.r) return make_float2(x, y);
    return ld2(p + c + 4);
}

// cell i0+4
template <typename P>
__device__ __forceinline__ float x4_right1(const XNb &n, const f4v v, const P &p, int c)
{
    const float x = from_right(v.x);
    if (n.r) return x;
    return ld1(p + c + 4);
}

// normal stresses and their memory variables (m_kernel.f90:176-239)
template <int BY, bool S16>
__global__ void __launch_bounds__(64 * BY) k3x_normal(const K3<float> a)
{
    int i0, j, k;
    unsigned msk;
    if (!x4_cell(a, BY, blockIdx.x, i0, j, k, msk)) return;
    const int px = a.px, sxy = a.sxy;
    const int c = at3(px, sxy, i0, j, k);
    const float *__restrict__ Vx = a.V[0];
    const float *__restrict__ Vy = a.V[1];
    const float *__restrict__ Vz = a.V[2];
    const f4v vx = ld4(Vx + c);
    const float2 vxl = ld2(Vx + c - 2);
    const float vxr = Vx[c + 4];
    const f4v vy = ld4(Vy + c), vym1 = ld4(Vy + c - px), vym2 = ld4(Vy + c - 2 * px), vyp1 = ld4(Vy + c + px);
    const f4v vz = ld4(Vz + c), vzm1 = ld4(Vz + c - sxy), vzm2 = ld4(Vz + c - 2 * sxy), vzp1 = ld4(Vz + c + sxy);
    const f4v lam = med4<S16>(a, MLAM, c, k), mu = med4<S16>(a, MMU, c, k);
    const f4v taup = med4<S16>(a, MTAUP, c, k), taus = med4<S16>(a, MTAUS, c, k);
    f4v R[9];
#pragma unroll
    for (int q = 0; q < 9; q++) R[q] = ldR4<S16>(a, q, c);
    f4v sxx = ld4nt(a.S[0] + c), syy = ld4nt(a.S[1] + c), szz = ld4nt(a.S[2] + c);
    const float xs[7] = {vxl.x, vxl.y, vx.x, vx.y, vx.z, vx.w, vxr};
    const float dt = a.dt;
#pragma unroll
    for (int q = 0; q < 4; q++) {
        const Ord o = order4(a, i0 + q, j, k);
        const float dxVx = (xs[q + 2] - xs[q + 1]) * o.e0x - (xs[q + 3] - xs[q]) * o.e1x;
        const float dyVy = (vy[q] - vym1[q]) * o.e0y - (vyp1[q] - vym2[q]) * o.e1y;
        const float dzVz = (vz[q] - vzm1[q]) * o.e0z - (vzp1[q] - vzm2[q]) * o.e1z;
        const float mu2 = 2 * mu[q];
        const float lam2mu = lam[q] + mu2;
        const float taup1 = taup[q], taus1 = taus[q];
        const float d3v3 = dxVx + dyVy + dzVz;
        const float dyVy_dzVz = dyVy + dzVz, dxVx_dzVz = dxVx + dzVz, dxVx_dyVy = dxVx + dyVy;
        const float fxx = lam2mu * taup1 * d3v3 - mu2 * taus1 * dyVy_dzVz;
        const float fyy = lam2mu * taup1 * d3v3 - mu2 * taus1 * dxVx_dzVz;
        const float fzz = lam2mu * taup1 * d3v3 - mu2 * taus1 * dxVx_dyVy;
        float Rxx_n = 0, Ryy_n = 0, Rzz_n = 0;
#pragma unroll
        for (int mm = 0; mm < NM3; mm++) {
            const float rxx = a.c1[mm] * R[0 + mm][q] - a.c2[mm] * fxx * dt;
            const float ryy = a.c1[mm] * R[3 + mm][q] - a.c2[mm] * fyy * dt;
            const float rzz = a.c1[mm] * R[6 + mm][q] - a.c2[mm] * fzz * dt;
            R[0 + mm][q] = rxx;
            R[3 + mm][q] = ryy;
            R[6 + mm][q] = rzz;
            Rxx_n += a.d1[mm] * rxx;
            Ryy_n += a.d1[mm] * ryy;
            Rzz_n += a.d1[mm] * rzz;
        }
        const float taup_plus1 = 1 + taup1 * (1 + a.d2);
        const float taus_plus1 = 1 + taus1 * (1 + a.d2);
        sxx[q] = sxx[q] + (lam2mu * taup_plus1 * d3v3 - mu2 * taus_plus1 * dyVy_dzVz + Rxx_n) * dt;
        syy[q] = syy[q] + (lam2mu * taup_plus1 * d3v3 - mu2 * taus_plus1 * dxVx_dzVz + Ryy_n) * dt;
        szz[q] = szz[q] + (lam2mu * taup_plus1 * d3v3 - mu2 * taus_plus1 * dxVx_dyVy + Rzz_n) * dt;
    }
#pragma unroll
    for (int q = 0; q < 9; q++) stR4<S16>(a, q, c, R[q], msk);
    st4m(a.S[0] + c, sxx, msk);
    st4m(a.S[1] + c, syy, msk);
    st4m(a.S[2] + c, szz, msk);
}

// shear stresses and their memory variables (m_kernel.f90:263-315)
template <int BY, bool S16>
__global__ void __launch_bounds__(64 * BY) k3x_shear(const K3<float> a)
{
    int i0, j, k;
    unsigned msk;
    if (!x4_cell(a, BY, blockIdx.x, i0, j, k, msk)) return;
    const int px = a.px, sxy = a.sxy;
    const int c = at3(px, sxy, i0, j, k);
    const float *__restrict__ Vx = a.V[0];
    const float *__restrict__ Vy = a.V[1];
    const float *__restrict__ Vz = a.V[2];
    const f4v vx = ld4(Vx + c), vxyp1 = ld4(Vx + c + px), vxyp2 = ld4(Vx + c + 2 * px), vxym1 = ld4(Vx + c - px);
    const f4v vxzp1 = ld4(Vx + c + sxy), vxzp2 = ld4(Vx + c + 2 * sxy), vxzm1 = ld4(Vx + c - sxy);
    const f4v vy = ld4(Vy + c), vyzp1 = ld4(Vy + c + sxy), vyzp2 = ld4(Vy + c + 2 * sxy), vyzm1 = ld4(Vy + c - sxy);
    const float2 vyl = ld2(Vy + c - 2), vyr = ld2(Vy + c + 4);
    const f4v vz = ld4(Vz + c), vzyp1 = ld4(Vz + c + px), vzyp2 = ld4(Vz + c + 2 * px), vzym1 = ld4(Vz + c - px);
    const float2 vzl = ld2(Vz + c - 2), vzr = ld2(Vz + c + 4);
    const f4v taus = med4<S16>(a, MTAUS, c, k);
    const f4v myz = med4<S16>(a, MMUYZ, c, k), mxz = med4<S16>(a, MMUXZ, c, k), mxy = med4<S16>(a, MMUXY, c, k);
    f4v R[9];
#pragma unroll
    for (int q = 0; q < 9; q++) R[q] = ldR4<S16>(a, 9 + q, c);
    f4v syz = ld4nt(a.S[3] + c), sxz = ld4nt(a.S[4] + c), sxy_ = ld4nt(a.S[5] + c);
    // x rows i0-1 .. i0+5 of Vy and Vz
    const float ys[7] = {vyl.y, vy.x, vy.y, vy.z, vy.w, vyr.x, vyr.y};
    const float zs[7] = {vzl.y, vz.x, vz.y, vz.z, vz.w, vzr.x, vzr.y};
    const float dt = a.dt;
#pragma unroll
    for (int q = 0; q < 4; q++) {
        const Ord o = order4(a, i0 + q, j, k);
        const float dxVy_dyVx = (ys[q + 2] - ys[q + 1]) * o.e0x - (ys[q + 3] - ys[q]) * o.e1x
                              + (vxyp1[q] - vx[q]) * o.e0y - (vxyp2[q] - vxym1[q]) * o.e1y;
        const float dxVz_dzVx = (zs[q + 2] - zs[q + 1]) * o.e0x - (zs[q + 3] - zs[q]) * o.e1x
                              + (vxzp1[q] - vx[q]) * o.e0z - (vxzp2[q] - vxzm1[q]) * o.e1z;
        const float dyVz_dzVy = (vzyp1[q] - vz[q]) * o.e0y - (vzyp2[q] - vzym1[q]) * o.e1y
                              + (vyzp1[q] - vy[q]) * o.e0z - (vyzp2[q] - vyzm1[q]) * o.e1z;
        const float taus1 = taus[q];
        const float fyz = myz[q] * taus1 * dyVz_dzVy;
        const float fxz = mxz[q] * taus1 * dxVz_dzVx;
        const float fxy = mxy[q] * taus1 * dxVy_dyVx;
        float Ryz_n = 0, Rxz_n = 0, Rxy_n = 0;
#pragma unroll
        for (int mm = 0; mm < NM3; mm++) {
            const float ryz = a.c1[mm] * R[0 + mm][q] - a.c2[mm] * fyz * dt;
            const float rxz = a.c1[mm] * R[3 + mm][q] - a.c2[mm] * fxz * dt;
            const float rxy = a.c1[mm] * R[6 + mm][q] - a.c2[mm] * fxy * dt;
            R[0 + mm][q] = ryz;
            R[3 + mm][q] = rxz;
            R[6 + mm][q] = rxy;
            Ryz_n += a.d1[mm] * ryz;
            Rxz_n += a.d1[mm] * rxz;
            Rxy_n += a.d1[mm] * rxy;
        }
        const float taus_plus1 = 1 + taus1 * (1 + a.d2);
        syz[q] = syz[q] + (myz[q] * taus_plus1 * dyVz_dzVy + Ryz_n) * dt;
        sxz[q] = sxz[q] + (mxz[q] * taus_plus1 * dxVz_dzVx + Rxz_n) * dt;
        sxy_[q] = sxy_[q] + (mxy[q] * taus_plus1 * dxVy_dyVx + Rxy_n) * dt;
    }
#pragma unroll
    for (int q = 0; q < 9; q++) stR4<S16>(a, 9 + q, c, R[q], msk);
    st4m(a.S[3] + c, syz, msk);
    st4m(a.S[4] + c, sxz, msk);
    st4m(a.S[5] + c, sxy_, msk);
}

// 4-point harmonic mean of averaged_medium (m_medium.f90:446-459) in the
// reference's single-precision operation order, without contraction, so
// that it equals the host-computed muxz/muxy/muyz bit for bit.
__device__ __forceinline__ float harm4s(float a, float b, float c, float d)
{
#pragma clang fp contract(off)
    const float epsl = 1.1920928955078125e-07f;  // epsilon(1.0)
    return 4 * a * b * c * d / (a * b * c + a * b * d + a * c * d + b * c * d + epsl);
}

// Normal and shear stresses in one pass (m_kernel.f90:142-326): the
// velocity differences of both parts are formed first from one set of V
// loads, then the normal and the shear memory variables are streamed in
// turn.  BAR = 1 keeps the compiler from hoisting the shear loads above the
// normal stores, which bounds the registers held at once.
template <int BY, bool S16, int BAR, bool XS = false>
__device__ __forceinline__ void stress4_blk(const K3<float> &a, unsigned blk)
{
    int i0, j, k;
    unsigned msk;
    if (!x4_cell(a, BY, blk, i0, j, k, msk)) return;
    const int px = a.px, sxy = a.sxy;
    const int c = at3(px, sxy, i0, j, k);
    const float *__restrict__ Vx = a.V[0];
    const float *__restrict__ Vy = a.V[1];
    const float *__restrict__ Vz = a.V[2];
    const XNb nb = XS ? x4_nb(a, i0) : XNb{false, false};
    f4v dxVx, dyVy, dzVz, dxVy_dyVx, dxVz_dzVx, dyVz_dzVy;
    {
        const f4v vx = ld4(Vx + c);
        const float2 vxl = XS ? x4_left(nb, vx, Vx, c) : ld2(Vx + c - 2);
        const float vxr = XS ? x4_right1(nb, vx, Vx, c) : Vx[c + 4];
        const f4v vxyp1 = ld4(Vx + c + px), vxyp2 = ld4(Vx + c + 2 * px), vxym1 = ld4(Vx + c - px);
        const f4v vxzp1 = ld4(Vx + c + sxy), vxzp2 = ld4(Vx + c + 2 * sxy), vxzm1 = ld4(Vx + c - sxy);
        const f4v vy = ld4(Vy + c), vym1 = ld4(Vy + c - px), vym2 = ld4(Vy + c - 2 * px), vyp1 = ld4(Vy + c + px);
        const f4v vyzp1 = ld4(Vy + c + sxy), vyzp2 = ld4(Vy + c + 2 * sxy), vyzm1 = ld4(Vy + c - sxy);
        const float2 vyl = XS ? x4_left(nb, vy, Vy, c) : ld2(Vy + c - 2);
        const float2 vyr = XS ? x4_right(nb, vy, Vy, c) : ld2(Vy + c + 4);
        const f4v vz = ld4(Vz + c), vzm1 = ld4(Vz + c - sxy), vzm2 = ld4(Vz + c - 2 * sxy), vzp1 = ld4(Vz + c + sxy);
        const f4v vzyp1 = ld4(Vz + c + px), vzyp2 = ld4(Vz + c + 2 * px), vzym1 = ld4(Vz + c - px);
        const float2 vzl = XS ? x4_left(nb, vz, Vz, c) : ld2(Vz + c - 2);
        const float2 vzr = XS ? x4_right(nb, vz, Vz, c) : ld2(Vz + c + 4);
        // x rows: Vx i0-2 .. i0+4, Vy and Vz i0-1 .. i0+5
        const float xs[7] = {vxl.x, vxl.y, vx.x, vx.y, vx.z, vx.w, vxr};
        const float ys[7] = {vyl.y, vy.x, vy.y, vy.z, vy.w, vyr.x, vyr.y};
        const float zs[7] = {vzl.y, vz.x, vz.y, vz.z, vz.w, vzr.x, vzr.y};
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const Ord o = order4(a, i0 + q, j, k);
            dxVx[q] = (xs[q + 2] - xs[q + 1]) * o.e0x - (xs[q + 3] - xs[q]) * o.e1x;
            dyVy[q] = (vy[q] - vym1[q]) * o.e0y - (vyp1[q] - vym2[q]) * o.e1y;
            dzVz[q] = (vz[q] - vzm1[q]) * o.e0z - (vzp1[q] - vzm2[q]) * o.e1z;
            dxVy_dyVx[q] = (ys[q + 2] - ys[q + 1]) * o.e0x - (ys[q + 3] - ys[q]) * o.e1x
                         + (vxyp1[q] - vx[q]) * o.e0y - (vxyp2[q] - vxym1[q]) * o.e1y;
            dxVz_dzVx[q] = (zs[q + 2] - zs[q + 1]) * o.e0x - (zs[q + 3] - zs[q]) * o.e1x
                         + (vxzp1[q] - vx[q]) * o.e0z - (vxzp2[q] - vxzm1[q]) * o.e1z;
            dyVz_dzVy[q] = (vzyp1[q] - vz[q]) * o.e0y - (vzyp2[q] - vzym1[q]) * o.e1y
                         + (vyzp1[q] - vy[q]) * o.e0z - (vyzp2[q] - vyzm1[q]) * o.e1z;
        }
    }
    const float dt = a.dt;
    const f4v taus = med4<S16>(a, MTAUS, c, k);
    {
        // FP32: mu with cached loads, its neighbours feed the shear means below
        const f4v lam = med4<S16>(a, MLAM, c, k), taup = med4<S16>(a, MTAUP, c, k);
        f4v mu;
        if (S16 && a.m16) mu = med4<S16>(a, MMU, c, k);
        else mu = ld4(a.med[MMU] + c);
        f4v R[9];
        if constexpr (S16) {
            // FP16 pairs (0,1) .. (6,7) in one 16-byte access each, 8 alone
#pragma unroll
            for (int q = 0; q < 8; q += 2) ldR4pair(a, q, c, R[q], R[q + 1]);
            R[8] = ldR4<S16>(a, 8, c);
        } else {
#pragma unroll
            for (int q = 0; q < 9; q++) R[q] = ldR4<S16>(a, q, c);
        }
        f4v sxx = ld4nt(a.S[0] + c), syy = ld4nt(a.S[1] + c), szz = ld4nt(a.S[2] + c);
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const float mu2 = 2 * mu[q];
            const float lam2mu = lam[q] + mu2;
            const float taup1 = taup[q], taus1 = taus[q];
            const float d3v3 = dxVx[q] + dyVy[q] + dzVz[q];
            const float dyVy_dzVz = dyVy[q] + dzVz[q], dxVx_dzVz = dxVx[q] + dzVz[q], dxVx_dyVy = dxVx[q] + dyVy[q];
            const float fxx = lam2mu * taup1 * d3v3 - mu2 * taus1 * dyVy_dzVz;
            const float fyy = lam2mu * taup1 * d3v3 - mu2 * taus1 * dxVx_dzVz;
            const float fzz = lam2mu * taup1 * d3v3 - mu2 * taus1 * dxVx_dyVy;
            float Rxx_n = 0, Ryy_n = 0, Rzz_n = 0;
#pragma unroll
            for (int mm = 0; mm < NM3; mm++) {
                const float rxx = a.c1[mm] * R[0 + mm][q] - a.c2[mm] * fxx * dt;
                const float ryy = a.c1[mm] * R[3 + mm][q] - a.c2[mm] * fyy * dt;
                const float rzz = a.c1[mm] * R[6 + mm][q] - a.c2[mm] * fzz * dt;
                R[0 + mm][q] = rxx;
                R[3 + mm][q] = ryy;
                R[6 + mm][q] = rzz;
                Rxx_n += a.d1[mm] * rxx;
                Ryy_n += a.d1[mm] * ryy;
                Rzz_n += a.d1[mm] * rzz;
            }
            const float taup_plus1 = 1 + taup1 * (1 + a.d2);
            const float taus_plus1 = 1 + taus1 * (1 + a.d2);
            sxx[q] = sxx[q] + (lam2mu * taup_plus1 * d3v3 - mu2 * taus_plus1 * dyVy_dzVz + Rxx_n) * dt;
            syy[q] = syy[q] + (lam2mu * taup_plus1 * d3v3 - mu2 * taus_plus1 * dxVx_dzVz + Ryy_n) * dt;
            szz[q] = szz[q] + (lam2mu * taup_plus1 * d3v3 - mu2 * taus_plus1 * dxVx_dyVy + Rzz_n) * dt;
        }
        if constexpr (S16) {
#pragma unroll
            for (int q = 0; q < 8; q += 2) stR4pair(a, q, c, R[q], R[q + 1], msk);
            stR4<S16>(a, 8, c, R[8], msk);
        } else {
#pragma unroll
            for (int q = 0; q < 9; q++) stR4<S16>(a, q, c, R[q], msk);
        }
        st4m(a.S[0] + c, sxx, msk);
        st4m(a.S[1] + c, syy, msk);
        st4m(a.S[2] + c, szz, msk);
    }
    if constexpr (BAR != 0) __builtin_amdgcn_sched_barrier(0);
    {
        f4v myz, mxz, mxy;
        if (S16 && a.m16) {
            myz = med4<S16>(a, MMUYZ, c, k);
            mxz = med4<S16>(a, MMUXZ, c, k);
            mxy = med4<S16>(a, MMUXY, c, k);
        } else {
            // interior cells: muxz, muxy, muyz from mu and its +1 neighbours
            // (cached loads), instead of three more streamed arrays
            const float *__restrict__ M = a.med[MMU];
            const f4v m0 = ld4(M + c), mk = ld4(M + c + sxy), mj = ld4(M + c + px), mkj = ld4(M + c + sxy + px);
            const float mi[5] = {m0.x, m0.y, m0.z, m0.w, XS ? x4_right1(nb, m0, M, c) : M[c + 4]};
            const float mki[5] = {mk.x, mk.y, mk.z, mk.w, XS ? x4_right1(nb, mk, M + sxy, c) : M[c + sxy + 4]};
            const float mji[5] = {mj.x, mj.y, mj.z, mj.w, XS ? x4_right1(nb, mj, M + px, c) : M[c + px + 4]};
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const float nnn = mi[q], pnn = mk[q], npn = mi[q + 1], ppn = mki[q + 1];
                const float nnp = mj[q], npp = mji[q + 1], pnp = mkj[q];
                mxz[q] = harm4s(nnn, pnn, npn, ppn);
                mxy[q] = harm4s(nnn, npn, nnp, npp);
                myz[q] = harm4s(nnn, pnn, nnp, pnp);
            }
        }
        f4v R[9];
        if constexpr (S16) {
            // FP16 pairs (9,10) .. (15,16), 17 alone
#pragma unroll
            for (int q = 0; q < 8; q += 2) ldR4pair(a, 9 + q, c, R[q], R[q + 1]);
            R[8] = ldR4<S16>(a, 17, c);
        } else {
#pragma unroll
            for (int q = 0; q < 9; q++) R[q] = ldR4<S16>(a, 9 + q, c);
        }
        f4v syz = ld4nt(a.S[3] + c), sxz = ld4nt(a.S[4] + c), sxy_ = ld4nt(a.S[5] + c);
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const float taus1 = taus[q];
            const float fyz = myz[q] * taus1 * dyVz_dzVy[q];
            const float fxz = mxz[q] * taus1 * dxVz_dzVx[q];
            const float fxy = mxy[q] * taus1 * dxVy_dyVx[q];
            float Ryz_n = 0, Rxz_n = 0, Rxy_n = 0;
#pragma unroll
            for (int mm = 0; mm < NM3; mm++) {
                const float ryz = a.c1[mm] * R[0 + mm][q] - a.c2[mm] * fyz * dt;
                const float rxz = a.c1[mm] * R[3 + mm][q] - a.c2[mm] * fxz * dt;
                const float rxy = a.c1[mm] * R[6 + mm][q] - a.c2[mm] * fxy * dt;
                R[0 + mm][q] = ryz;
                R[3 + mm][q] = rxz;
                R[6 + mm][q] = rxy;
                Ryz_n += a.d1[mm] * ryz;
                Rxz_n += a.d1[mm] * rxz;
                Rxy_n += a.d1[mm] * rxy;
            }
            const float taus_plus1 = 1 + taus1 * (1 + a.d2);
            syz[q] = syz[q] + (myz[q] * taus_plus1 * dyVz_dzVy[q] + Ryz_n) * dt;
            sxz[q] = sxz[q] + (mxz[q] * taus_plus1 * dxVz_dzVx[q] + Rxz_n) * dt;
            sxy_[q] = sxy_[q] + (mxy[q] * taus_plus1 * dxVy_dyVx[q] + Rxy_n) * dt;
        }
        if constexpr (S16) {
#pragma unroll
            for (int q = 0; q < 8; q += 2) stR4pair(a, 9 + q, c, R[q], R[q + 1], msk);
            stR4<S16>(a, 17, c, R[8], msk);
        } else {
#pragma unroll
            for (int q = 0; q < 9; q++) stR4<S16>(a, 9 + q, c, R[q], msk);
        }
        st4m(a.S[3] + c, syz, msk);
        st4m(a.S[4] + c, sxz, msk);
        st4m(a.S[5] + c, sxy_, msk);
    }
}

template <int BY, bool S16, int BAR, bool XS>
__global__ void __launch_bounds__(64 * BY) k3x_stress(const K3<float> a)
{
    stress4_blk<BY, S16, BAR, XS>(a, blockIdx.x);
}

// One-pass stress update, z-marching: each thread owns 4 x-cells of one
// row and walks a chunk of a.kc planes.  The z-stencils of Vz (planes
// k-2..k+1), Vx and Vy (k-1..k+2) come from register windows, one new
// plane of each per step; the memory variables, stresses and medium stream
// through once.  Workgroups: 256-cell x tiles x BY rows x z-chunks.
template <int BY, bool S16>
__global__ void __launch_bounds__(64 * BY) k3x_stressz(const K3<float> a)
{
    const unsigned b = blockIdx.x;
    const unsigned xt = b % (unsigned)a.ntx;
    const unsigned t = b / (unsigned)a.ntx;
    const unsigned yt = t % (unsigned)a.nty;
    const unsigned zc = t / (unsigned)a.nty;
    const int i0 = a.ti0 + (int)xt * 256 + (int)threadIdx.x * 4;
    const int j = a.na + 1 + (int)yt * BY + (int)threadIdx.y;
    if (j > a.ny - a.na) return;
    unsigned msk = 0;
#pragma unroll
    for (int q = 0; q < 4; q++)
        if (i0 + q > a.na && i0 + q <= a.nx - a.na) msk |= 1u << q;
    if (msk == 0) return;
    const int kb = 1 + (int)zc * a.kc;
    const int ke = min(kb + a.kc - 1, a.nz - a.na);
    if (kb > ke) return;
    const int px = a.px, sxy = a.sxy;
    const float *__restrict__ Vx = a.V[0];
    const float *__restrict__ Vy = a.V[1];
    const float *__restrict__ Vz = a.V[2];
    int4 fs[4];
#pragma unroll
    for (int q = 0; q < 4; q++) fs[q] = a.surf[(j - 1) * a.nx + (min(max(i0 + q, 1), a.nx) - 1)];
    int c = at3(px, sxy, i0, j, kb);
    f4v vz_m2 = ld4(Vz + c - 2 * sxy), vz_m1 = ld4(Vz + c - sxy), vz_0 = ld4(Vz + c);
    f4v vx_m1 = ld4(Vx + c - sxy), vx_0 = ld4(Vx + c), vx_p1 = ld4(Vx + c + sxy);
    f4v vy_m1 = ld4(Vy + c - sxy), vy_0 = ld4(Vy + c), vy_p1 = ld4(Vy + c + sxy);
    const float dt = a.dt;
    for (int k = kb; k <= ke; k++, c += sxy) {
        const f4v vz_p1 = ld4(Vz + c + sxy), vx_p2 = ld4(Vx + c + 2 * sxy), vy_p2 = ld4(Vy + c + 2 * sxy);
        f4v dxVx, dyVy, dzVz, dxVy_dyVx, dxVz_dzVx, dyVz_dzVy;
        {
            const float2 vxl = ld2(Vx + c - 2);
            const float vxr = Vx[c + 4];
            const f4v vxyp1 = ld4(Vx + c + px), vxyp2 = ld4(Vx + c + 2 * px), vxym1 = ld4(Vx + c - px);
            const f4v vym1 = ld4(Vy + c - px), vym2 = ld4(Vy + c - 2 * px), vyp1 = ld4(Vy + c + px);
            const float2 vyl = ld2(Vy + c - 2), vyr = ld2(Vy + c + 4);
            const f4v vzyp1 = ld4(Vz + c + px), vzyp2 = ld4(Vz + c + 2 * px), vzym1 = ld4(Vz + c - px);
            const float2 vzl = ld2(Vz + c - 2), vzr = ld2(Vz + c + 4);
            const float xs[7] = {vxl.x, vxl.y, vx_0.x, vx_0.y, vx_0.z, vx_0.w, vxr};
            const float ys[7] = {vyl.y, vy_0.x, vy_0.y, vy_0.z, vy_0.w, vyr.x, vyr.y};
            const float zs[7] = {vzl.y, vz_0.x, vz_0.y, vz_0.z, vz_0.w, vzr.x, vzr.y};
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const int o1 = (k - fs[q].x) * (fs[q].y - k), o2 = (k - fs[q].z) * (fs[q].w - k);
                const bool lo = max(o1, o2) >= 0;
                const float e0x = lo ? a.e2[0][0] : a.e4[0][0], e1x = lo ? a.e2[0][1] : a.e4[0][1];
                const float e0y = lo ? a.e2[1][0] : a.e4[1][0], e1y = lo ? a.e2[1][1] : a.e4[1][1];
                const float e0z = lo ? a.e2[2][0] : a.e4[2][0], e1z = lo ? a.e2[2][1] : a.e4[2][1];
                dxVx[q] = (xs[q + 2] - xs[q + 1]) * e0x - (xs[q + 3] - xs[q]) * e1x;
                dyVy[q] = (vy_0[q] - vym1[q]) * e0y - (vyp1[q] - vym2[q]) * e1y;
                dzVz[q] = (vz_0[q] - vz_m1[q]) * e0z - (vz_p1[q] - vz_m2[q]) * e1z;
                dxVy_dyVx[q] = (ys[q + 2] - ys[q + 1]) * e0x - (ys[q + 3] - ys[q]) * e1x
                             + (vxyp1[q] - vx_0[q]) * e0y - (vxyp2[q] - vxym1[q]) * e1y;
                dxVz_dzVx[q] = (zs[q + 2] - zs[q + 1]) * e0x - (zs[q + 3] - zs[q]) * e1x
                             + (vx_p1[q] - vx_0[q]) * e0z - (vx_p2[q] - vx_m1[q]) * e1z;
                dyVz_dzVy[q] = (vzyp1[q] - vz_0[q]) * e0y - (vzyp2[q] - vzym1[q]) * e1y
                             + (vy_p1[q] - vy_0[q]) * e0z - (vy_p2[q] - vy_m1[q]) * e1z;
            }
        }
        const f4v taus = med4<S16>(a, MTAUS, c, k);
        {
            const f4v lam = med4<S16>(a, MLAM, c, k), mu = med4<S16>(a, MMU, c, k), taup = med4<S16>(a, MTAUP, c, k);
            f4v R[9];
#pragma unroll
            for (int q = 0; q < 9; q++) R[q] = ldR4<S16>(a, q, c);
            f4v sxx = ld4nt(a.S[0] + c), syy = ld4nt(a.S[1] + c), szz = ld4nt(a.S[2] + c);
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const float mu2 = 2 * mu[q];
                const float lam2mu = lam[q] + mu2;
                const float taup1 = taup[q], taus1 = taus[q];
                const float d3v3 = dxVx[q] + dyVy[q] + dzVz[q];
                const float dyVy_dzVz = dyVy[q] + dzVz[q], dxVx_dzVz = dxVx[q] + dzVz[q], dxVx_dyVy = dxVx[q] + dyVy[q];
                const float fxx = lam2mu * taup1 * d3v3 - mu2 * taus1 * dyVy_dzVz;
                const float fyy = lam2mu * taup1 * d3v3 - mu2 * taus1 * dxVx_dzVz;
                const float fzz = lam2mu * taup1 * d3v3 - mu2 * taus1 * dxVx_dyVy;
                float Rxx_n = 0, Ryy_n = 0, Rzz_n = 0;
#pragma unroll
                for (int mm = 0; mm < NM3; mm++) {
                    const float rxx = a.c1[mm] * R[0 + mm][q] - a.c2[mm] * fxx * dt;
                    const float ryy = a.c1[mm] * R[3 + mm][q] - a.c2[mm] * fyy * dt;
                    const float rzz = a.c1[mm] * R[6 + mm][q] - a.c2[mm] * fzz * dt;
                    R[0 + mm][q] = rxx;
                    R[3 + mm][q] = ryy;
                    R[6 + mm][q] = rzz;
                    Rxx_n += a.d1[mm] * rxx;
                    Ryy_n += a.d1[mm] * ryy;
                    Rzz_n += a.d1[mm] * rzz;
                }
                const float taup_plus1 = 1 + taup1 * (1 + a.d2);
                const float taus_plus1 = 1 + taus1 * (1 + a.d2);
                sxx[q] = sxx[q] + (lam2mu * taup_plus1 * d3v3 - mu2 * taus_plus1 * dyVy_dzVz + Rxx_n) * dt;
                syy[q] = syy[q] + (lam2mu * taup_plus1 * d3v3 - mu2 * taus_plus1 * dxVx_dzVz + Ryy_n) * dt;
                szz[q] = szz[q] + (lam2mu * taup_plus1 * d3v3 - mu2 * taus_plus1 * dxVx_dyVy + Rzz_n) * dt;
            }
#pragma unroll
            for (int q = 0; q < 9; q++) stR4<S16>(a, q, c, R[q], msk);
            st4m(a.S[0] + c, sxx, msk);
            st4m(a.S[1] + c, syy, msk);
            st4m(a.S[2] + c, szz, msk);
        }
        {
            const f4v myz = med4<S16>(a, MMUYZ, c, k), mxz = med4<S16>(a, MMUXZ, c, k), mxy = med4<S16>(a, MMUXY, c, k);
            f4v R[9];
#pragma unroll
            for (int q = 0; q < 9; q++) R[q] = ldR4<S16>(a, 9 + q, c);
            f4v syz = ld4nt(a.S[3] + c), sxz = ld4nt(a.S[4] + c), sxy_ = ld4nt(a.S[5] + c);
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const float taus1 = taus[q];
                const float fyz = myz[q] * taus1 * dyVz_dzVy[q];
                const float fxz = mxz[q] * taus1 * dxVz_dzVx[q];
                const float fxy = mxy[q] * taus1 * dxVy_dyVx[q];
                float Ryz_n = 0, Rxz_n = 0, Rxy_n = 0;
#pragma unroll
                for (int mm = 0; mm < NM3; mm++) {
                    const float ryz = a.c1[mm] * R[0 + mm][q] - a.c2[mm] * fyz * dt;
                    const float rxz = a.c1[mm] * R[3 + mm][q] - a.c2[mm] * fxz * dt;
                    const float rxy = a.c1[mm] * R[6 + mm][q] - a.c2[mm] * fxy * dt;
                    R[0 + mm][q] = ryz;
                    R[3 + mm][q] = rxz;
                    R[6 + mm][q] = rxy;
                    Ryz_n += a.d1[mm] * ryz;
                    Rxz_n += a.d1[mm] * rxz;
                    Rxy_n += a.d1[mm] * rxy;
                }
                const float taus_plus1 = 1 + taus1 * (1 + a.d2);
                syz[q] = syz[q] + (myz[q] * taus_plus1 * dyVz_dzVy[q] + Ryz_n) * dt;
                sxz[q] = sxz[q] + (mxz[q] * taus_plus1 * dxVz_dzVx[q] + Rxz_n) * dt;
                sxy_[q] = sxy_[q] + (mxy[q] * taus_plus1 * dxVy_dyVx[q] + Rxy_n) * dt;
            }
#pragma unroll
            for (int q = 0; q < 9; q++) stR4<S16>(a, 9 + q, c, R[q], msk);
            st4m(a.S[3] + c, syz, msk);
            st4m(a.S[4] + c, sxz, msk);
            st4m(a.S[5] + c, sxy_, msk);
        }
        vz_m2 = vz_m1; vz_m1 = vz_0; vz_0 = vz_p1;
        vx_m1 = vx_0; vx_0 = vx_p1; vx_p1 = vx_p2;
        vy_m1 = vy_0; vy_0 = vy_p1; vy_p1 = vy_p2;
    }
}

// velocity (m_kernel.f90:99-129)
template <int BY, bool S16>
__global__ void __launch_bounds__(64 * BY) k3x_vel(const K3<float> a)
{
    int i0, j, k;
    unsigned msk;
    if (!x4_cell(a, BY, blockIdx.x, i0, j, k, msk)) return;
    const int px = a.px, sxy = a.sxy;
    const int c = at3(px, sxy, i0, j, k);
    const STile<float> Sxx = a.S[0];
    const STile<float> Syy = a.S[1];
    const STile<float> Szz = a.S[2];
    const STile<float> Syz = a.S[3];
    const STile<float> Sxz = a.S[4];
    const STile<float> Sxy = a.S[5];
    const f4v sxx = ld4(Sxx + c);
    const float2 sxxl = ld2(Sxx + c - 2), sxxr = ld2(Sxx + c + 4);
    const f4v syy = ld4(Syy + c), syyp1 = ld4(Syy + c + px), syyp2 = ld4(Syy + c + 2 * px), syym1 = ld4(Syy + c - px);
    const f4v szz = ld4(Szz + c), szzp1 = ld4(Szz + c + sxy), szzp2 = ld4(Szz + c + 2 * sxy), szzm1 = ld4(Szz + c - sxy);
    const f4v sxyv = ld4(Sxy + c), sxyym1 = ld4(Sxy + c - px), sxyym2 = ld4(Sxy + c - 2 * px), sxyyp1 = ld4(Sxy + c + px);
    const float2 sxyl = ld2(Sxy + c - 2), sxyr = ld2(Sxy + c + 4);
    const f4v sxzv = ld4(Sxz + c), sxzzm1 = ld4(Sxz + c - sxy), sxzzm2 = ld4(Sxz + c - 2 * sxy), sxzzp1 = ld4(Sxz + c + sxy);
    const float2 sxzl = ld2(Sxz + c - 2), sxzr = ld2(Sxz + c + 4);
    const f4v syzv = ld4(Syz + c), syzym1 = ld4(Syz + c - px), syzym2 = ld4(Syz + c - 2 * px), syzyp1 = ld4(Syz + c + px);
    const f4v syzzm1 = ld4(Syz + c - sxy), syzzm2 = ld4(Syz + c - 2 * sxy), syzzp1 = ld4(Syz + c + sxy);
    const f4v bx = med4<S16>(a, MBX, c, k), by = med4<S16>(a, MBY, c, k), bz = med4<S16>(a, MBZ, c, k);
    f4v vx = ld4nt(a.V[0] + c), vy = ld4nt(a.V[1] + c), vz = ld4nt(a.V[2] + c);
    // x rows i0-2 .. i0+5
    const float xx[8] = {sxxl.x, sxxl.y, sxx.x, sxx.y, sxx.z, sxx.w, sxxr.x, sxxr.y};
    const float xy[8] = {sxyl.x, sxyl.y, sxyv.x, sxyv.y, sxyv.z, sxyv.w, sxyr.x, sxyr.y};
    const float xz[8] = {sxzl.x, sxzl.y, sxzv.x, sxzv.y, sxzv.z, sxzv.w, sxzr.x, sxzr.y};
    const float dt = a.dt;
#pragma unroll
    for (int q = 0; q < 4; q++) {
        const Ord o = order4(a, i0 + q, j, k);
        const int x = q + 2;  // index of cell i0+q in the x rows
        const float d3Sx3 = (xx[x + 1] - xx[x]) * o.e0x - (xx[x + 2] - xx[x - 1]) * o.e1x
                          + (sxyv[q] - sxyym1[q]) * o.e0y - (sxyyp1[q] - sxyym2[q]) * o.e1y
                          + (sxzv[q] - sxzzm1[q]) * o.e0z - (sxzzp1[q] - sxzzm2[q]) * o.e1z;
        const float d3Sy3 = (xy[x] - xy[x - 1]) * o.e0x - (xy[x + 1] - xy[x - 2]) * o.e1x
                          + (syyp1[q] - syy[q]) * o.e0y - (syyp2[q] - syym1[q]) * o.e1y
                          + (syzv[q] - syzzm1[q]) * o.e0z - (syzzp1[q] - syzzm2[q]) * o.e1z;
        const float d3Sz3 = (xz[x] - xz[x - 1]) * o.e0x - (xz[x + 1] - xz[x - 2]) * o.e1x
                          + (syzv[q] - syzym1[q]) * o.e0y - (syzyp1[q] - syzym2[q]) * o.e1y
                          + (szzp1[q] - szz[q]) * o.e0z - (szzp2[q] - szzm1[q]) * o.e1z;
        vx[q] = vx[q] + bx[q] * d3Sx3 * dt;
        vy[q] = vy[q] + by[q] * d3Sy3 * dt;
        vz[q] = vz[q] + bz[q] * d3Sz3 * dt;
    }
    st4m(a.V[0] + c, vx, msk);
    st4m(a.V[1] + c, vy, msk);
    st4m(a.V[2] + c, vz, msk);
}

// Velocity, z-marching: each thread owns 4 x-cells of one row and walks a
// chunk of a.kc planes.  The z-stencils of Szz, Sxz and Syz come from a
// register window (one new plane per step), so each of their planes is
// read once per chunk instead of four times; y- and x-neighbours are read
// as in k3x_vel.  Workgroups: 256-cell x tiles x BY rows x z-chunks.
template <int BY, bool S16, bool RHO, bool XS = false>
__device__ __forceinline__ void velz_blk(const K3<float> &a, unsigned b)
{
    unsigned xt, yt, zc;
    if (a.xcd >= 2) {
        // groups of G = a.xcd consecutive y tiles (all x tiles, one
        // z-chunk), group g on XCD g mod 8 (workgroups are dealt to the
        // XCDs in turn): the y-halo rows that a group's tiles share stay in
        // one L2, while the XCDs still work on neighbouring rows
        const unsigned G = (unsigned)a.xcd, per = G * (unsigned)a.ntx;
        const unsigned ngy = ((unsigned)a.nty + G - 1u) / G;
        const unsigned s = b >> 3, gi = (b & 7u) + 8u * (s / per), r = s % per;
        if (gi >= ngy * (unsigned)a.nkc) return;
        zc = gi / ngy;
        yt = (gi % ngy) * G + r / (unsigned)a.ntx;
        xt = r % (unsigned)a.ntx;
        if (yt >= (unsigned)a.nty) return;
    } else {
        if (a.xcd) {
            // each XCD walks one contiguous run of tiles, so the y-neighbour
            // rows of a tile were fetched into the same L2
            b = (b & 7u) * (a.nblk_pad >> 3) + (b >> 3);
            if (b >= a.nblk) return;
        }
        xt = b % (unsigned)a.ntx;
        const unsigned t = b / (unsigned)a.ntx;
        yt = t % (unsigned)a.nty;
        zc = t / (unsigned)a.nty;
    }
    const int i0 = a.ti0 + (int)xt * 256 + (int)threadIdx.x * 4;
    const int j = a.na + 1 + (int)yt * BY + (int)threadIdx.y;
    if (j > a.ny - a.na) return;
    unsigned msk = 0;
#pragma unroll
    for (int q = 0; q < 4; q++)
        if (i0 + q > a.na && i0 + q <= a.nx - a.na) msk |= 1u << q;
    if (msk == 0) return;
    const int kb = 1 + (int)zc * a.kc;
    const int ke = min(kb + a.kc - 1, a.nz - a.na);
    if (kb > ke) return;
    const int px = a.px, sxy = a.sxy;
    const STile<float> Sxx = a.S[0];
    const STile<float> Syy = a.S[1];
    const STile<float> Szz = a.S[2];
    const STile<float> Syz = a.S[3];
    const STile<float> Sxz = a.S[4];
    const STile<float> Sxy = a.S[5];
    int4 fs[4];
#pragma unroll
    for (int q = 0; q < 4; q++) fs[q] = a.surf[(j - 1) * a.nx + (min(max(i0 + q, 1), a.nx) - 1)];
    int c = at3(px, sxy, i0, j, kb);
    const XNb nb = XS ? x4_nb(a, i0) : XNb{false, false};
    f4v zz_m1 = ld4(Szz + c - sxy), zz_0 = ld4(Szz + c), zz_p1 = ld4(Szz + c + sxy);
    f4v xz_m2 = ld4(Sxz + c - 2 * sxy), xz_m1 = ld4(Sxz + c - sxy), xz_0 = ld4(Sxz + c);
    f4v yz_m2 = ld4(Syz + c - 2 * sxy), yz_m1 = ld4(Syz + c - sxy), yz_0 = ld4(Syz + c);
    // RHO: buoyancies 2/(rho + rho') of averaged_medium (m_medium.f90:442-444)
    // from a density window (planes k, k+1) instead of three streamed arrays
    const float *__restrict__ Rh = a.rho;
    f4v r_0;
    if constexpr (RHO) r_0 = ld4(Rh + c);
    const float dt = a.dt;
    for (int k = kb; k <= ke; k++, c += sxy) {
        const f4v zz_p2 = ld4(Szz + c + 2 * sxy), xz_p1 = ld4(Sxz + c + sxy), yz_p1 = ld4(Syz + c + sxy);
        const f4v sxx = ld4(Sxx + c);
        const float2 sxxl = XS ? x4_left(nb, sxx, Sxx, c) : ld2(Sxx + c - 2);
        const float2 sxxr = XS ? x4_right(nb, sxx, Sxx, c) : ld2(Sxx + c + 4);
        const f4v syy = ld4(Syy + c), syyp1 = ld4(Syy + c + px), syyp2 = ld4(Syy + c + 2 * px), syym1 = ld4(Syy + c - px);
        const f4v sxyv = ld4(Sxy + c), sxyym1 = ld4(Sxy + c - px), sxyym2 = ld4(Sxy + c - 2 * px), sxyyp1 = ld4(Sxy + c + px);
        const float2 sxyl = XS ? x4_left(nb, sxyv, Sxy, c) : ld2(Sxy + c - 2);
        const float2 sxyr = XS ? x4_right(nb, sxyv, Sxy, c) : ld2(Sxy + c + 4);
        const float2 sxzl = XS ? x4_left(nb, xz_0, Sxz, c) : ld2(Sxz + c - 2);
        const float2 sxzr = XS ? x4_right(nb, xz_0, Sxz, c) : ld2(Sxz + c + 4);
        const f4v syzym1 = ld4(Syz + c - px), syzym2 = ld4(Syz + c - 2 * px), syzyp1 = ld4(Syz + c + px);
        f4v bx, by, bz, r_p1;
        if constexpr (RHO) {
            r_p1 = ld4(Rh + c + sxy);
            const f4v rj = ld4(Rh + c + px);
            const float ri[5] = {r_0.x, r_0.y, r_0.z, r_0.w, XS ? x4_right1(nb, r_0, Rh, c) : Rh[c + 4]};
#pragma unroll
            for (int q = 0; q < 4; q++) {
                bx[q] = 2.0f / (ri[q] + ri[q + 1]);
                by[q] = 2.0f / (r_0[q] + rj[q]);
                bz[q] = 2.0f / (r_0[q] + r_p1[q]);
            }
        } else {
            bx = med4<S16>(a, MBX, c, k);
            by = med4<S16>(a, MBY, c, k);
            bz = med4<S16>(a, MBZ, c, k);
        }
        f4v vx = ld4nt(a.V[0] + c), vy = ld4nt(a.V[1] + c), vz = ld4nt(a.V[2] + c);
        // x rows i0-2 .. i0+5
        const float xx[8] = {sxxl.x, sxxl.y, sxx.x, sxx.y, sxx.z, sxx.w, sxxr.x, sxxr.y};
        const float xy[8] = {sxyl.x, sxyl.y, sxyv.x, sxyv.y, sxyv.z, sxyv.w, sxyr.x, sxyr.y};
        const float xz[8] = {sxzl.x, sxzl.y, xz_0.x, xz_0.y, xz_0.z, xz_0.w, sxzr.x, sxzr.y};
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const int o1 = (k - fs[q].x) * (fs[q].y - k), o2 = (k - fs[q].z) * (fs[q].w - k);
            const bool lo = max(o1, o2) >= 0;
            const float e0x = lo ? a.e2[0][0] : a.e4[0][0], e1x = lo ? a.e2[0][1] : a.e4[0][1];
            const float e0y = lo ? a.e2[1][0] : a.e4[1][0], e1y = lo ? a.e2[1][1] : a.e4[1][1];
            const float e0z = lo ? a.e2[2][0] : a.e4[2][0], e1z = lo ? a.e2[2][1] : a.e4[2][1];
            const int x = q + 2;  // index of cell i0+q in the x rows
            const float d3Sx3 = (xx[x + 1] - xx[x]) * e0x - (xx[x + 2] - xx[x - 1]) * e1x
                              + (sxyv[q] - sxyym1[q]) * e0y - (sxyyp1[q] - sxyym2[q]) * e1y
                              + (xz_0[q] - xz_m1[q]) * e0z - (xz_p1[q] - xz_m2[q]) * e1z;
            const float d3Sy3 = (xy[x] - xy[x - 1]) * e0x - (xy[x + 1] - xy[x - 2]) * e1x
                              + (syyp1[q] - syy[q]) * e0y - (syyp2[q] - syym1[q]) * e1y
                              + (yz_0[q] - yz_m1[q]) * e0z - (yz_p1[q] - yz_m2[q]) * e1z;
            const float d3Sz3 = (xz[x] - xz[x - 1]) * e0x - (xz[x + 1] - xz[x - 2]) * e1x
                              + (yz_0[q] - syzym1[q]) * e0y - (syzyp1[q] - syzym2[q]) * e1y
                              + (zz_p1[q] - zz_0[q]) * e0z - (zz_p2[q] - zz_m1[q]) * e1z;
            vx[q] = vx[q] + bx[q] * d3Sx3 * dt;
            vy[q] = vy[q] + by[q] * d3Sy3 * dt;
            vz[q] = vz[q] + bz[q] * d3Sz3 * dt;
        }
        st4m(a.V[0] + c, vx, msk);
        st4m(a.V[1] + c, vy, msk);
        st4m(a.V[2] + c, vz, msk);
        zz_m1 = zz_0; zz_0 = zz_p1; zz_p1 = zz_p2;
        xz_m2 = xz_m1; xz_m1 = xz_0; xz_0 = xz_p1;
        yz_m2 = yz_m1; yz_m1 = yz_0; yz_0 = yz_p1;
        if constexpr (RHO) r_0 = r_p1;
    }
}

template <int BY, bool S16, bool RHO, bool XS>
__global__ void __launch_bounds__(64 * BY) k3x_velz(const K3<float> a)
{
    velz_blk<BY, S16, RHO, XS>(a, blockIdx.x);
}


// Velocity, z-marching with the y-stencil rows staged in LDS: as k3x_velz,
// but per plane the workgroup first copies rows j0-2 .. j0+BY+1 of Syy, Sxy
// and Syz (256 + 8 cells each, x halo included) into LDS, and every thread
// takes its y neighbours (and Sxy's x neighbours) from there instead of
// from the L2.  NB = 2 double-buffers the planes (one barrier per plane),
// NB = 1 reuses one buffer (two).  The arithmetic and its order are those
// of k3x_velz, so both give the same bits.
template <int BY, int NB, bool S16, bool RHO>
__global__ void __launch_bounds__(64 * BY) k3x_velzl(const K3<float> a)
{
    constexpr int NR = BY + 4;        // staged rows j0-2 .. j0+BY+1
    constexpr int RL = 264;           // cells i0b-4 .. i0b+259 per row
    __shared__ __attribute__((aligned(16))) float L[NB][3][NR][RL];
    const unsigned b = blockIdx.x;
    const unsigned xt = b % (unsigned)a.ntx;
    const unsigned t = b / (unsigned)a.ntx;
    const unsigned yt = t % (unsigned)a.nty;
    const unsigned zc = t / (unsigned)a.nty;
    const int i0b = a.ti0 + (int)xt * 256;
    const int j0 = a.na + 1 + (int)yt * BY;
    const int tx = (int)threadIdx.x, ty = (int)threadIdx.y;
    const int i0 = i0b + tx * 4;
    const int j = j0 + ty;
    unsigned msk = 0;
#pragma unroll
    for (int q = 0; q < 4; q++)
        if (i0 + q > a.na && i0 + q <= a.nx - a.na) msk |= 1u << q;
    // every thread takes part in the staging and the barriers
    const bool act = j <= a.ny - a.na && msk != 0;
    const int kb = 1 + (int)zc * a.kc;
    const int ke = min(kb + a.kc - 1, a.nz - a.na);
    if (kb > ke) return;
    const int px = a.px, sxy = a.sxy;
    const STile<float> Sxx = a.S[0];
    const STile<float> Syy = a.S[1];
    const STile<float> Szz = a.S[2];
    const STile<float> Syz = a.S[3];
    const STile<float> Sxz = a.S[4];
    const STile<float> Sxy = a.S[5];
    const STile<float> Sy[3] = {Syy, Sxy, Syz};
    int4 fs[4];
#pragma unroll
    for (int q = 0; q < 4; q++)
        fs[q] = a.surf[(min(j, a.ny) - 1) * a.nx + (min(max(i0 + q, 1), a.nx) - 1)];
    int c = at3(px, sxy, i0, min(j, a.ny + 3), kb);
    f4v zz_m1 = {0, 0, 0, 0}, zz_0 = zz_m1, zz_p1 = zz_m1, xz_m2 = zz_m1, xz_m1 = zz_m1, xz_0 = zz_m1;
    f4v yz_m2 = zz_m1, yz_m1 = zz_m1, yz_0 = zz_m1, r_0 = zz_m1;
    if (act) {
        zz_m1 = ld4(Szz + c - sxy); zz_0 = ld4(Szz + c); zz_p1 = ld4(Szz + c + sxy);
        xz_m2 = ld4(Sxz + c - 2 * sxy); xz_m1 = ld4(Sxz + c - sxy); xz_0 = ld4(Sxz + c);
        yz_m2 = ld4(Syz + c - 2 * sxy); yz_m1 = ld4(Syz + c - sxy); yz_0 = ld4(Syz + c);
        if constexpr (RHO) r_0 = ld4(a.rho + c);
    }
    const float *__restrict__ Rh = a.rho;
    const float dt = a.dt;
    // staging: NR rows x 66 float4 groups x 3 fields per plane; a group
    // lies inside its row of the memory box or is filled with zeros (no
    // active thread reads those)
    constexpr int NG = 66, NSTAGE = 3 * NR * NG;
    const int tid = ty * 64 + tx;
    for (int k = kb; k <= ke; k++, c += sxy) {
        const int buf = NB == 2 ? (k - kb) & 1 : 0;
        if (NB == 1 && k > kb) __syncthreads();   // the previous plane's reads are done
        for (int e = tid; e < NSTAGE; e += 64 * BY) {
            const int f = e / (NR * NG), rem = e - f * (NR * NG);
            const int r = rem / NG, g = rem - r * NG;
            const int jj = j0 - 2 + r, ii = i0b - 4 + 4 * g;
            f4v v = {0, 0, 0, 0};
            if (jj >= -2 && jj <= a.ny + 3 && ii + 31 >= 0 && ii + 31 + 3 < px)
                v = ld4(Sy[f] + at3(px, sxy, ii, jj, k));
            *(f4v *)&L[buf][f][r][4 * g] = v;
        }
        __syncthreads();
        if (!act) continue;
        const float *Lyy = &L[buf][0][0][0], *Lxy = &L[buf][1][0][0], *Lyz = &L[buf][2][0][0];
        const int o = 4 + 4 * tx;     // own float4 in a staged row
        const f4v zz_p2 = ld4(Szz + c + 2 * sxy), xz_p1 = ld4(Sxz + c + sxy), yz_p1 = ld4(Syz + c + sxy);
        const f4v sxx = ld4(Sxx + c);
        const float2 sxxl = ld2(Sxx + c - 2), sxxr = ld2(Sxx + c + 4);
        const f4v syy = *(const f4v *)&Lyy[(ty + 2) * RL + o], syyp1 = *(const f4v *)&Lyy[(ty + 3) * RL + o];
        const f4v syyp2 = *(const f4v *)&Lyy[(ty + 4) * RL + o], syym1 = *(const f4v *)&Lyy[(ty + 1) * RL + o];
        const f4v sxyv = *(const f4v *)&Lxy[(ty + 2) * RL + o], sxyym1 = *(const f4v *)&Lxy[(ty + 1) * RL + o];
        const f4v sxyym2 = *(const f4v *)&Lxy[ty * RL + o], sxyyp1 = *(const f4v *)&Lxy[(ty + 3) * RL + o];
        const float2 sxyl = *(const float2 *)&Lxy[(ty + 2) * RL + o - 2];
        const float2 sxyr = *(const float2 *)&Lxy[(ty + 2) * RL + o + 4];
        const float2 sxzl = ld2(Sxz + c - 2), sxzr = ld2(Sxz + c + 4);
        const f4v syzym1 = *(const f4v *)&Lyz[(ty + 1) * RL + o], syzym2 = *(const f4v *)&Lyz[ty * RL + o];
        const f4v syzyp1 = *(const f4v *)&Lyz[(ty + 3) * RL + o];
        f4v bx, by, bz, r_p1;
        if constexpr (RHO) {
            r_p1 = ld4(Rh + c + sxy);
            const f4v rj = ld4(Rh + c + px);
            const float ri[5] = {r_0.x, r_0.y, r_0.z, r_0.w, Rh[c + 4]};
#pragma unroll
            for (int q = 0; q < 4; q++) {
                bx[q] = 2.0f / (ri[q] + ri[q + 1]);
                by[q] = 2.0f / (r_0[q] + rj[q]);
                bz[q] = 2.0f / (r_0[q] + r_p1[q]);
            }
        } else {
            bx = med4<S16>(a, MBX, c, k);
            by = med4<S16>(a, MBY, c, k);
            bz = med4<S16>(a, MBZ, c, k);
        }
        f4v vx = ld4nt(a.V[0] + c), vy = ld4nt(a.V[1] + c), vz = ld4nt(a.V[2] + c);
        // x rows i0-2 .. i0+5
        const float xx[8] = {sxxl.x, sxxl.y, sxx.x, sxx.y, sxx.z, sxx.w, sxxr.x, sxxr.y};
        const float xy[8] = {sxyl.x, sxyl.y, sxyv.x, sxyv.y, sxyv.z, sxyv.w, sxyr.x, sxyr.y};
        const float xz[8] = {sxzl.x, sxzl.y, xz_0.x, xz_0.y, xz_0.z, xz_0.w, sxzr.x, sxzr.y};
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const int o1 = (k - fs[q].x) * (fs[q].y - k), o2 = (k - fs[q].z) * (fs[q].w - k);
            const bool lo = max(o1, o2) >= 0;
            const float e0x = lo ? a.e2[0][0] : a.e4[0][0], e1x = lo ? a.e2[0][1] : a.e4[0][1];
            const float e0y = lo ? a.e2[1][0] : a.e4[1][0], e1y = lo ? a.e2[1][1] : a.e4[1][1];
            const float e0z = lo ? a.e2[2][0] : a.e4[2][0], e1z = lo ? a.e2[2][1] : a.e4[2][1];
            const int x = q + 2;  // index of cell i0+q in the x rows
            const float d3Sx3 = (xx[x + 1] - xx[x]) * e0x - (xx[x + 2] - xx[x - 1]) * e1x
                              + (sxyv[q] - sxyym1[q]) * e0y - (sxyyp1[q] - sxyym2[q]) * e1y
                              + (xz_0[q] - xz_m1[q]) * e0z - (xz_p1[q] - xz_m2[q]) * e1z;
            const float d3Sy3 = (xy[x] - xy[x - 1]) * e0x - (xy[x + 1] - xy[x - 2]) * e1x
                              + (syyp1[q] - syy[q]) * e0y - (syyp2[q] - syym1[q]) * e1y
                              + (yz_0[q] - yz_m1[q]) * e0z - (yz_p1[q] - yz_m2[q]) * e1z;
            const float d3Sz3 = (xz[x] - xz[x - 1]) * e0x - (xz[x + 1] - xz[x - 2]) * e1x
                              + (yz_0[q] - syzym1[q]) * e0y - (syzyp1[q] - syzym2[q]) * e1y
                              + (zz_p1[q] - zz_0[q]) * e0z - (zz_p2[q] - zz_m1[q]) * e1z;
            vx[q] = vx[q] + bx[q] * d3Sx3 * dt;
            vy[q] = vy[q] + by[q] * d3Sy3 * dt;
            vz[q] = vz[q] + bz[q] * d3Sz3 * dt;
        }
        st4m(a.V[0] + c, vx, msk);
        st4m(a.V[1] + c, vy, msk);
        st4m(a.V[2] + c, vz, msk);
        zz_m1 = zz_0; zz_0 = zz_p1; zz_p1 = zz_p2;
        xz_m2 = xz_m1; xz_m1 = xz_0; xz_0 = xz_p1;
        yz_m2 = yz_m1; yz_m1 = yz_0; yz_0 = yz_p1;
        if constexpr (RHO) r_0 = r_p1;
    }
}

// ---------------------------------------------------------------------------
// PML (ADE-CFS, 2nd order, elastic)
// ---------------------------------------------------------------------------
__device__ inline bool pml_cell(const Boxes3 &B, unsigned id, int &i, int &j, int &k)
{
    if (id >= B.total) return false;
    int b = 0;
#pragma unroll
    for (int q = 1; q < 5; q++)
        if (q < B.n && id >= B.b[q].start) b = q;
    const Box3 &x = B.b[b];
    unsigned l = id - x.start;
    unsigned li = l % (unsigned)x.ni;
    unsigned t = l / (unsigned)x.ni;
    unsigned lj = t % (unsigned)x.nj;
    unsigned lk = t / (unsigned)x.nj;
    i = x.i0 + (int)li;
    j = x.j0 + (int)lj;
    k = x.k0 + (int)lk;
    return true;
}

// The absorber boxes that span whole rows i in [1, nx] (the y faces and
// the bottom), walked one 256-cell row segment per workgroup: j and k are
// then uniform across the workgroup, and so are the y and z profiles.
struct Rows3 {
    int n, nseg;
    int j0[5], nj[5], k0[5];
    unsigned start[5];  // first row of each box
    unsigned total;     // rows
};

// Workgroups go to the XCDs round robin.  With per != 0 (a grid of
// 8 * per workgroups) XCD x takes the contiguous run of logical workgroups
// [x * per, (x + 1) * per), so that the neighbour rows and planes a cell
// reads were fetched into that XCD's L2 by its own neighbours; the
// logical ids past the launch's count exit at once.
__device__ __forceinline__ unsigned pml_block(unsigned per)
{
    const unsigned b = blockIdx.x;
    return per ? (b & 7u) * per + (b >> 3) : b;
}

__device__ __forceinline__ bool pml_row(const Rows3 &R, unsigned blk, int nx, int &i, int &j, int &k)
{
    const unsigned row = blk / (unsigned)R.nseg;
    if (row >= R.total) return false;
    i = 1 + (int)(blk - row * (unsigned)R.nseg) * 256 + (int)threadIdx.x;
    int b = 0;
#pragma unroll
    for (int q = 1; q < 5; q++)
        if (q < R.n && row >= R.start[q]) b = q;
    const unsigned l = row - R.start[b];
    j = R.j0[b] + (int)(l % (unsigned)R.nj[b]);
    k = R.k0[b] + (int)(l / (unsigned)R.nj[b]);
    return i <= nx;
}

template <typename T, bool S16>
__device__ __forceinline__ void stress_pml_cell(const K3<T> &a, const int i, const int j, const int k)
{
    const int px = a.px, sxy = a.sxy;
    const int c = at3(px, sxy, i, j, k);
    const GLB T *__restrict__ Vx = gp(a.V[0]);
    const GLB T *__restrict__ Vy = gp(a.V[1]);
    const GLB T *__restrict__ Vz = gp(a.V[2]);
    const int nx = a.nx, ny = a.ny, nz = a.nz;
    const T gxc1 = gp(a.gx)[0 * nx + i - 1], gxc2 = gp(a.gx)[1 * nx + i - 1], gxc3 = gp(a.gx)[2 * nx + i - 1], gxc4 = gp(a.gx)[3 * nx + i - 1];
    const T gxe1 = gp(a.gx)[4 * nx + i - 1], gxe2 = gp(a.gx)[5 * nx + i - 1], gxe3 = gp(a.gx)[6 * nx + i - 1], gxe4 = gp(a.gx)[7 * nx + i - 1];
    const T gyc1 = gp(a.gy)[0 * ny + j - 1], gyc2 = gp(a.gy)[1 * ny + j - 1], gyc3 = gp(a.gy)[2 * ny + j - 1], gyc4 = gp(a.gy)[3 * ny + j - 1];
    const T gye1 = gp(a.gy)[4 * ny + j - 1], gye2 = gp(a.gy)[5 * ny + j - 1], gye3 = gp(a.gy)[6 * ny + j - 1], gye4 = gp(a.gy)[7 * ny + j - 1];
    const T gzc1 = gp(a.gz)[0 * nz + k - 1], gzc2 = gp(a.gz)[1 * nz + k - 1], gzc3 = gp(a.gz)[2 * nz + k - 1], gzc4 = gp(a.gz)[3 * nz + k - 1];
    const T gze1 = gp(a.gz)[4 * nz + k - 1], gze2 = gp(a.gz)[5 * nz + k - 1], gze3 = gp(a.gz)[6 * nz + k - 1], gze4 = gp(a.gz)[7 * nz + k - 1];
    const T r20x = a.r2[0], r20y = a.r2[1], r20z = a.r2[2];
    const T dt = a.dt;
    // Along an axis where this cell is undamped (d = 0, so g4 = 0 at the
    // centre and the edge) the ADE update a = g3 a + g4 dv dt keeps a zero
    // array zero, and its term g2 a adds zero.  While the context knows the
    // arrays started from zero (ade_sparse), such ADE values are neither
    // read nor written; the arithmetic is unchanged.
    const bool lx = !a.ade_sparse || gxc4 != (T)0 || gxe4 != (T)0;
    const bool ly = !a.ade_sparse || gyc4 != (T)0 || gye4 != (T)0;
    const bool lz = !a.ade_sparse || gzc4 != (T)0 || gze4 != (T)0;
    const T vx0 = Vx[c], vy0 = Vy[c], vz0 = Vz[c];
    // normal (m_absorb_p.f90:445-462)
    const T dxVx = (vx0 - Vx[c - 1]) * r20x;
    const T dyVy = (vy0 - Vy[c - px]) * r20y;
    const T dzVz = (vz0 - Vz[c - sxy]) * r20z;
    const T mu0 = medv<S16>(a, MMU, c, k);
    T lam2mu_R;
    if constexpr (S16) lam2mu_R = medv<S16>(a, MLAM, c, k) + 2 * mu0;
    else lam2mu_R = (T)(gp(a.med[MLAM])[c] + 2 * gp(a.med[MMU])[c]);
    const T lam_R = lam2mu_R - 2 * mu0;
    const T axvx = lx ? ldA<S16>(a, 0, c) : (T)0, ayvy = ly ? ldA<S16>(a, 4, c) : (T)0;
    const T azvz = lz ? ldA<S16>(a, 8, c) : (T)0;
    const T dxVx_ade = gxc1 * dxVx + gxc2 * axvx;
    const T dyVy_ade = gyc1 * dyVy + gyc2 * ayvy;
    const T dzVz_ade = gzc1 * dzVz + gzc2 * azvz;
    a.S[0][c] = a.S[0][c] + (lam2mu_R * dxVx_ade + lam_R * (dyVy_ade + dzVz_ade)) * dt;
    a.S[1][c] = a.S[1][c] + (lam2mu_R * dyVy_ade + lam_R * (dxVx_ade + dzVz_ade)) * dt;
    a.S[2][c] = a.S[2][c] + (lam2mu_R * dzVz_ade + lam_R * (dxVx_ade + dyVy_ade)) * dt;
    if (lx) stA<S16>(a, 0, c, gxc3 * axvx + gxc4 * dxVx * dt);
    if (ly) stA<S16>(a, 4, c, gyc3 * ayvy + gyc4 * dyVy * dt);
    if (lz) stA<S16>(a, 8, c, gzc3 * azvz + gzc4 * dzVz * dt);

    // shear (m_absorb_p.f90:469-491)
    const T dxVy = (Vy[c + 1] - vy0) * r20x;
    const T dxVz = (Vz[c + 1] - vz0) * r20x;
    const T dyVx = (Vx[c + px] - vx0) * r20y;
    const T dyVz = (Vz[c + px] - vz0) * r20y;
    const T dzVx = (Vx[c + sxy] - vx0) * r20z;
    const T dzVy = (Vy[c + sxy] - vy0) * r20z;
    const T ayvx = ly ? ldA<S16>(a, 1, c) : (T)0, azvx = lz ? ldA<S16>(a, 2, c) : (T)0;
    const T axvy = lx ? ldA<S16>(a, 3, c) : (T)0, azvy = lz ? ldA<S16>(a, 5, c) : (T)0;
    const T axvz = lx ? ldA<S16>(a, 6, c) : (T)0, ayvz = ly ? ldA<S16>(a, 7, c) : (T)0;
    T myz, mxz, mxy;
    if (S16 && a.m16) {
        myz = medv<S16>(a, MMUYZ, c, k);
        mxz = medv<S16>(a, MMUXZ, c, k);
        mxy = medv<S16>(a, MMUXY, c, k);
    } else {
        // the averaged moduli of averaged_medium from mu and its +1 neighbours
        const GLB float *__restrict__ M = gp(a.med[MMU]);
        const float nnn = M[c], pnn = M[c + sxy], npn = M[c + 1], ppn = M[c + sxy + 1];
        const float nnp = M[c + px], npp = M[c + px + 1], pnp = M[c + sxy + px];
        mxz = (T)harm4s(nnn, pnn, npn, ppn);
        mxy = (T)harm4s(nnn, npn, nnp, npp);
        myz = (T)harm4s(nnn, pnn, nnp, pnp);
    }
    a.S[3][c] = a.S[3][c] + myz * (gye1 * dyVz + gze1 * dzVy + gye2 * ayvz + gze2 * azvy) * dt;
    a.S[4][c] = a.S[4][c] + mxz * (gxe1 * dxVz + gze1 * dzVx + gxe2 * axvz + gze2 * azvx) * dt;
    a.S[5][c] = a.S[5][c] + mxy * (gxe1 * dxVy + gye1 * dyVx + gxe2 * axvy + gye2 * ayvx) * dt;
    if (ly) stA<S16>(a, 1, c, gye3 * ayvx + gye4 * dyVx * dt);
    if (lz) stA<S16>(a, 2, c, gze3 * azvx + gze4 * dzVx * dt);
    if (lx) stA<S16>(a, 3, c, gxe3 * axvy + gxe4 * dxVy * dt);
    if (lz) stA<S16>(a, 5, c, gze3 * azvy + gze4 * dzVy * dt);
    if (lx) stA<S16>(a, 6, c, gxe3 * axvz + gxe4 * dxVz * dt);
    if (ly) stA<S16>(a, 7, c, gye3 * ayvz + gye4 * dyVz * dt);
}

template <typename T, bool S16>
__global__ void __launch_bounds__(256) k3_stress_pml(const K3<T> a, unsigned per)
{
    int i, j, k;
    if (pml_cell(a.pml, pml_block(per) * 256u + threadIdx.x, i, j, k)) stress_pml_cell<T, S16>(a, i, j, k);
}

template <typename T, bool S16>
__global__ void __launch_bounds__(256) k3_stress_pml_rows(const K3<T> a, const Rows3 R, unsigned per)
{
    int i, j, k;
    if (pml_row(R, pml_block(per), a.nx, i, j, k)) stress_pml_cell<T, S16>(a, i, j, k);
}

template <typename T, bool S16>
__device__ __forceinline__ void vel_pml_cell(const K3<T> &a, const int i, const int j, const int k)
{
    const int px = a.px, sxy = a.sxy;
    const int c = at3(px, sxy, i, j, k);
    const STile<T> Sxx = a.S[0];
    const STile<T> Syy = a.S[1];
    const STile<T> Szz = a.S[2];
    const STile<T> Syz = a.S[3];
    const STile<T> Sxz = a.S[4];
    const STile<T> Sxy = a.S[5];
    const int nx = a.nx, ny = a.ny, nz = a.nz;
    const T gxc1 = gp(a.gx)[0 * nx + i - 1], gxc2 = gp(a.gx)[1 * nx + i - 1], gxc3 = gp(a.gx)[2 * nx + i - 1], gxc4 = gp(a.gx)[3 * nx + i - 1];
    const T gxe1 = gp(a.gx)[4 * nx + i - 1], gxe2 = gp(a.gx)[5 * nx + i - 1], gxe3 = gp(a.gx)[6 * nx + i - 1], gxe4 = gp(a.gx)[7 * nx + i - 1];
    const T gyc1 = gp(a.gy)[0 * ny + j - 1], gyc2 = gp(a.gy)[1 * ny + j - 1], gyc3 = gp(a.gy)[2 * ny + j - 1], gyc4 = gp(a.gy)[3 * ny + j - 1];
    const T gye1 = gp(a.gy)[4 * ny + j - 1], gye2 = gp(a.gy)[5 * ny + j - 1], gye3 = gp(a.gy)[6 * ny + j - 1], gye4 = gp(a.gy)[7 * ny + j - 1];
    const T gzc1 = gp(a.gz)[0 * nz + k - 1], gzc2 = gp(a.gz)[1 * nz + k - 1], gzc3 = gp(a.gz)[2 * nz + k - 1], gzc4 = gp(a.gz)[3 * nz + k - 1];
    const T gze1 = gp(a.gz)[4 * nz + k - 1], gze2 = gp(a.gz)[5 * nz + k - 1], gze3 = gp(a.gz)[6 * nz + k - 1], gze4 = gp(a.gz)[7 * nz + k - 1];
    const T r20x = a.r2[0], r20y = a.r2[1], r20z = a.r2[2];
    const T dt = a.dt;
    // undamped axes: see k3_stress_pml
    const bool lx = !a.ade_sparse || gxc4 != (T)0 || gxe4 != (T)0;
    const bool ly = !a.ade_sparse || gyc4 != (T)0 || gye4 != (T)0;
    const bool lz = !a.ade_sparse || gzc4 != (T)0 || gze4 != (T)0;
    // m_absorb_p.f90:266-298
    const T sxx0 = Sxx[c], syy0 = Syy[c], szz0 = Szz[c], syz0 = Syz[c], sxz0 = Sxz[c], sxy0 = Sxy[c];
    const T dxSxx = (Sxx[c + 1] - sxx0) * r20x;
    const T dySyy = (Syy[c + px] - syy0) * r20y;
    const T dzSzz = (Szz[c + sxy] - szz0) * r20z;
    const T dySyz = (syz0 - Syz[c - px]) * r20y;
    const T dzSyz = (syz0 - Syz[c - sxy]) * r20z;
    const T dxSxz = (sxz0 - Sxz[c - 1]) * r20x;
    const T dzSxz = (sxz0 - Sxz[c - sxy]) * r20z;
    const T dxSxy = (sxy0 - Sxy[c - 1]) * r20x;
    const T dySxy = (sxy0 - Sxy[c - px]) * r20y;
    const T axsxx = lx ? ldA<S16>(a, 9, c) : (T)0, aysxy = ly ? ldA<S16>(a, 10, c) : (T)0;
    const T azsxz = lz ? ldA<S16>(a, 11, c) : (T)0, axsxy = lx ? ldA<S16>(a, 12, c) : (T)0;
    const T aysyy = ly ? ldA<S16>(a, 13, c) : (T)0, azsyz = lz ? ldA<S16>(a, 14, c) : (T)0;
    const T axsxz = lx ? ldA<S16>(a, 15, c) : (T)0, aysyz = ly ? ldA<S16>(a, 16, c) : (T)0;
    const T azszz = lz ? ldA<S16>(a, 17, c) : (T)0;
    T bx, by, bz;
    if (a.rho) {
        // buoyancies 2/(rho + rho') of averaged_medium from the density
        const GLB float *__restrict__ Rh = gp(a.rho);
        const float r0 = Rh[c];
        bx = (T)(2.0f / (r0 + Rh[c + 1]));
        by = (T)(2.0f / (r0 + Rh[c + px]));
        bz = (T)(2.0f / (r0 + Rh[c + sxy]));
    } else {
        bx = medv<S16>(a, MBX, c, k);
        by = medv<S16>(a, MBY, c, k);
        bz = medv<S16>(a, MBZ, c, k);
    }
    gp(a.V[0])[c] = gp(a.V[0])[c] + bx * (gxe1 * dxSxx + gyc1 * dySxy + gzc1 * dzSxz
                                  + gxe2 * axsxx + gyc2 * aysxy + gzc2 * azsxz) * dt;
    gp(a.V[1])[c] = gp(a.V[1])[c] + by * (gxc1 * dxSxy + gye1 * dySyy + gzc1 * dzSyz
                                  + gxc2 * axsxy + gye2 * aysyy + gzc2 * azsyz) * dt;
    gp(a.V[2])[c] = gp(a.V[2])[c] + bz * (gxc1 * dxSxz + gyc1 * dySyz + gze1 * dzSzz
                                  + gxc2 * axsxz + gyc2 * aysyz + gze2 * azszz) * dt;
    if (lx) stA<S16>(a, 9, c, gxe3 * axsxx + gxe4 * dxSxx * dt);
    if (ly) stA<S16>(a, 10, c, gyc3 * aysxy + gyc4 * dySxy * dt);
    if (lz) stA<S16>(a, 11, c, gzc3 * azsxz + gzc4 * dzSxz * dt);
    if (lx) stA<S16>(a, 12, c, gxc3 * axsxy + gxc4 * dxSxy * dt);
    if (ly) stA<S16>(a, 13, c, gye3 * aysyy + gye4 * dySyy * dt);
    if (lz) stA<S16>(a, 14, c, gzc3 * azsyz + gzc4 * dzSyz * dt);
    if (lx) stA<S16>(a, 15, c, gxc3 * axsxz + gxc4 * dxSxz * dt);
    if (ly) stA<S16>(a, 16, c, gyc3 * aysyz + gyc4 * dySyz * dt);
    if (lz) stA<S16>(a, 17, c, gze3 * azszz + gze4 * dzSzz * dt);
}

template <typename T, bool S16>
__global__ void __launch_bounds__(256) k3_vel_pml(const K3<T> a, unsigned per)
{
    int i, j, k;
    if (pml_cell(a.pml, pml_block(per) * 256u + threadIdx.x, i, j, k)) vel_pml_cell<T, S16>(a, i, j, k);
}

template <typename T, bool S16>
__global__ void __launch_bounds__(256) k3_vel_pml_rows(const K3<T> a, const Rows3 R, unsigned per)
{
    int i, j, k;
    if (pml_row(R, pml_block(per), a.nx, i, j, k)) vel_pml_cell<T, S16>(a, i, j, k);
}

// ---------------------------------------------------------------------------
// 4-wide PML kernels (FP32 arithmetic, FP32 or reduced storage)
// ---------------------------------------------------------------------------
// The absorber boxes are enumerated in groups of 4 x-cells aligned to
// i = 1 + 4m (16-byte aligned rows); cells of a group outside its box are
// masked.  Profiles along x are read as float4, along y and z as scalars.
struct Box4 {
    int g0, ngx, i0, i1, j0, nj, k0;
    unsigned start;
};

struct Boxes4 {
    Box4 b[5];
    int n;
    unsigned total;
};

typedef unsigned short u4v __attribute__((ext_vector_type(4)));

template <bool S16>
__device__ __forceinline__ f4v ldA4(const K3<float> &a, int q, int c)
{
    if constexpr (S16) {
        const u4v r = __builtin_nontemporal_load((const u4v *)(a.hA[q] + c));
        f4v v;
        v.x = __uint_as_float((unsigned)r.x << 16);
        v.y = __uint_as_float((unsigned)r.y << 16);
        v.z = __uint_as_float((unsigned)r.z << 16);
        v.w = __uint_as_float((unsigned)r.w << 16);
        return v;
    } else {
        return ld4nt(a.A[q] + c);
    }
}

__device__ __forceinline__ unsigned short bf16_rne(float f)
{
    unsigned u = __float_as_uint(f);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (unsigned short)(u >> 16);
}

template <bool S16>
__device__ __forceinline__ void stA4(const K3<float> &a, int q, int c, f4v v, unsigned m)
{
    if constexpr (S16) {
        u4v r;
        if (a.sr) {
            r.x = f2bf_sr(v.x, (unsigned)c, 32u + (unsigned)q);
            r.y = f2bf_sr(v.y, (unsigned)c + 1u, 32u + (unsigned)q);
            r.z = f2bf_sr(v.z, (unsigned)c + 2u, 32u + (unsigned)q);
            r.w = f2bf_sr(v.w, (unsigned)c + 3u, 32u + (unsigned)q);
        } else {
            r.x = bf16_rne(v.x);
            r.y = bf16_rne(v.y);
            r.z = bf16_rne(v.z);
            r.w = bf16_rne(v.w);
        }
        unsigned short *p = (unsigned short *)(a.hA[q] + c);
        if (m == 15u) {
            __builtin_nontemporal_store(r, (u4v *)p);
        } else {
            if (m & 1u) p[0] = r.x;
            if (m & 2u) p[1] = r.y;
            if (m & 4u) p[2] = r.z;
            if (m & 8u) p[3] = r.w;
        }
    } else {
        st4m(a.A[q] + c, v, m);
    }
}

__device__ __forceinline__ bool pml4_cell(const Boxes4 &B, unsigned id, int &i0, int &j, int &k, unsigned &m)
{
    if (id >= B.total) return false;
    int b = 0;
#pragma unroll
    for (int q = 1; q < 5; q++)
        if (q < B.n && id >= B.b[q].start) b = q;
    const Box4 &x = B.b[b];
    const unsigned l = id - x.start;
    const unsigned gx = l % (unsigned)x.ngx;
    const unsigned t = l / (unsigned)x.ngx;
    j = x.j0 + (int)(t % (unsigned)x.nj);
    k = x.k0 + (int)(t / (unsigned)x.nj);
    i0 = 1 + 4 * (x.g0 + (int)gx);
    m = 0;
#pragma unroll
    for (int q = 0; q < 4; q++)
        if (i0 + q >= x.i0 && i0 + q <= x.i1) m |= 1u << q;
    return true;
}

// profile row q (0..7) of x at cells i0..i0+3 (profiles padded to 4)
__device__ __forceinline__ f4v gx4(const K3<float> &a, int q, int i0) { return ld4(a.gxv + q * a.gxp + (i0 - 1)); }

template <bool S16>
__global__ void __launch_bounds__(256) k3x_stress_pml(const K3<float> a, const Boxes4 B)
{
    int i0, j, k;
    unsigned msk;
    if (!pml4_cell(B, blockIdx.x * 256u + threadIdx.x, i0, j, k, msk)) return;
    const int px = a.px, sxy = a.sxy, ny = a.ny, nz = a.nz;
    const int c = at3(px, sxy, i0, j, k);
    const float *__restrict__ Vx = a.V[0];
    const float *__restrict__ Vy = a.V[1];
    const float *__restrict__ Vz = a.V[2];
    const f4v gxc1 = gx4(a, 0, i0), gxc2 = gx4(a, 1, i0), gxc3 = gx4(a, 2, i0), gxc4 = gx4(a, 3, i0);
    const f4v gxe1 = gx4(a, 4, i0), gxe2 = gx4(a, 5, i0), gxe3 = gx4(a, 6, i0), gxe4 = gx4(a, 7, i0);
    const float gyc1 = a.gy[0 * ny + j - 1], gyc2 = a.gy[1 * ny + j - 1], gyc3 = a.gy[2 * ny + j - 1], gyc4 = a.gy[3 * ny + j - 1];
    const float gye1 = a.gy[4 * ny + j - 1], gye2 = a.gy[5 * ny + j - 1], gye3 = a.gy[6 * ny + j - 1], gye4 = a.gy[7 * ny + j - 1];
    const float gzc1 = a.gz[0 * nz + k - 1], gzc2 = a.gz[1 * nz + k - 1], gzc3 = a.gz[2 * nz + k - 1], gzc4 = a.gz[3 * nz + k - 1];
    const float gze1 = a.gz[4 * nz + k - 1], gze2 = a.gz[5 * nz + k - 1], gze3 = a.gz[6 * nz + k - 1], gze4 = a.gz[7 * nz + k - 1];
    const float r20x = a.r2[0], r20y = a.r2[1], r20z = a.r2[2], dt = a.dt;

    const f4v vx = ld4(Vx + c), vxyp = ld4(Vx + c + px), vxzp = ld4(Vx + c + sxy);
    const float vxl = Vx[c - 1];
    const f4v vy = ld4(Vy + c), vyym = ld4(Vy + c - px), vyzp = ld4(Vy + c + sxy);
    const float vyr = Vy[c + 4];
    const f4v vz = ld4(Vz + c), vzzm = ld4(Vz + c - sxy), vzyp = ld4(Vz + c + px);
    const float vzr = Vz[c + 4];
    const f4v lam = med4<S16>(a, MLAM, c, k), mu = med4<S16>(a, MMU, c, k);
    const f4v myz = med4<S16>(a, MMUYZ, c, k), mxz = med4<S16>(a, MMUXZ, c, k), mxy = med4<S16>(a, MMUXY, c, k);
    f4v s0 = ld4nt(a.S[0] + c), s1 = ld4nt(a.S[1] + c), s2 = ld4nt(a.S[2] + c);
    f4v s3 = ld4nt(a.S[3] + c), s4 = ld4nt(a.S[4] + c), s5 = ld4nt(a.S[5] + c);
    f4v A[9];
#pragma unroll
    for (int q = 0; q < 9; q++) A[q] = ldA4<S16>(a, q, c);
    // x rows: Vx at i0-1..i0+3, Vy and Vz at i0..i0+4
    const float xvx[5] = {vxl, vx.x, vx.y, vx.z, vx.w};
    const float xvy[5] = {vy.x, vy.y, vy.z, vy.w, vyr};
    const float xvz[5] = {vz.x, vz.y, vz.z, vz.w, vzr};
#pragma unroll
    for (int q = 0; q < 4; q++) {
        // normal (m_absorb_p.f90:445-462)
        const float dxVx = (xvx[q + 1] - xvx[q]) * r20x;
        const float dyVy = (vy[q] - vyym[q]) * r20y;
        const float dzVz = (vz[q] - vzzm[q]) * r20z;
        float lam2mu_R;
        if constexpr (S16) lam2mu_R = lam[q] + 2 * mu[q];
        else lam2mu_R = lam[q] + 2 * mu[q];
        const float lam_R = lam2mu_R - 2 * mu[q];
        const float axvx = A[0][q], ayvy = A[4][q], azvz = A[8][q];
        const float dxVx_ade = gxc1[q] * dxVx + gxc2[q] * axvx;
        const float dyVy_ade = gyc1 * dyVy + gyc2 * ayvy;
        const float dzVz_ade = gzc1 * dzVz + gzc2 * azvz;
        s0[q] = s0[q] + (lam2mu_R * dxVx_ade + lam_R * (dyVy_ade + dzVz_ade)) * dt;
        s1[q] = s1[q] + (lam2mu_R * dyVy_ade + lam_R * (dxVx_ade + dzVz_ade)) * dt;
        s2[q] = s2[q] + (lam2mu_R * dzVz_ade + lam_R * (dxVx_ade + dyVy_ade)) * dt;
        A[0][q] = gxc3[q] * axvx + gxc4[q] * dxVx * dt;
        A[4][q] = gyc3 * ayvy + gyc4 * dyVy * dt;
        A[8][q] = gzc3 * azvz + gzc4 * dzVz * dt;
        // shear (m_absorb_p.f90:469-491)
        const float dxVy = (xvy[q + 1] - xvy[q]) * r20x;
        const float dxVz = (xvz[q + 1] - xvz[q]) * r20x;
        const float dyVx = (vxyp[q] - vx[q]) * r20y;
        const float dyVz = (vzyp[q] - vz[q]) * r20y;
        const float dzVx = (vxzp[q] - vx[q]) * r20z;
        const float dzVy = (vyzp[q] - vy[q]) * r20z;
        const float ayvx = A[1][q], azvx = A[2][q], axvy = A[3][q], azvy = A[5][q], axvz = A[6][q], ayvz = A[7][q];
        s3[q] = s3[q] + myz[q] * (gye1 * dyVz + gze1 * dzVy + gye2 * ayvz + gze2 * azvy) * dt;
        s4[q] = s4[q] + mxz[q] * (gxe1[q] * dxVz + gze1 * dzVx + gxe2[q] * axvz + gze2 * azvx) * dt;
        s5[q] = s5[q] + mxy[q] * (gxe1[q] * dxVy + gye1 * dyVx + gxe2[q] * axvy + gye2 * ayvx) * dt;
        A[1][q] = gye3 * ayvx + gye4 * dyVx * dt;
        A[2][q] = gze3 * azvx + gze4 * dzVx * dt;
        A[3][q] = gxe3[q] * axvy + gxe4[q] * dxVy * dt;
        A[5][q] = gze3 * azvy + gze4 * dzVy * dt;
        A[6][q] = gxe3[q] * axvz + gxe4[q] * dxVz * dt;
        A[7][q] = gye3 * ayvz + gye4 * dyVz * dt;
    }
    st4m(a.S[0] + c, s0, msk);
    st4m(a.S[1] + c, s1, msk);
    st4m(a.S[2] + c, s2, msk);
    st4m(a.S[3] + c, s3, msk);
    st4m(a.S[4] + c, s4, msk);
    st4m(a.S[5] + c, s5, msk);
#pragma unroll
    for (int q = 0; q < 9; q++) stA4<S16>(a, q, c, A[q], msk);
}

template <bool S16>
__global__ void __launch_bounds__(256) k3x_vel_pml(const K3<float> a, const Boxes4 B)
{
    int i0, j, k;
    unsigned msk;
    if (!pml4_cell(B, blockIdx.x * 256u + threadIdx.x, i0, j, k, msk)) return;
    const int px = a.px, sxy = a.sxy, ny = a.ny, nz = a.nz;
    const int c = at3(px, sxy, i0, j, k);
    const STile<float> Sxx = a.S[0];
    const STile<float> Syy = a.S[1];
    const STile<float> Szz = a.S[2];
    const STile<float> Syz = a.S[3];
    const STile<float> Sxz = a.S[4];
    const STile<float> Sxy = a.S[5];
    const f4v gxc1 = gx4(a, 0, i0), gxc2 = gx4(a, 1, i0), gxc3 = gx4(a, 2, i0), gxc4 = gx4(a, 3, i0);
    const f4v gxe1 = gx4(a, 4, i0), gxe2 = gx4(a, 5, i0), gxe3 = gx4(a, 6, i0), gxe4 = gx4(a, 7, i0);
    const float gyc1 = a.gy[0 * ny + j - 1], gyc2 = a.gy[1 * ny + j - 1], gyc3 = a.gy[2 * ny + j - 1], gyc4 = a.gy[3 * ny + j - 1];
    const float gye1 = a.gy[4 * ny + j - 1], gye2 = a.gy[5 * ny + j - 1], gye3 = a.gy[6 * ny + j - 1], gye4 = a.gy[7 * ny + j - 1];
    const float gzc1 = a.gz[0 * nz + k - 1], gzc2 = a.gz[1 * nz + k - 1], gzc3 = a.gz[2 * nz + k - 1], gzc4 = a.gz[3 * nz + k - 1];
    const float gze1 = a.gz[4 * nz + k - 1], gze2 = a.gz[5 * nz + k - 1], gze3 = a.gz[6 * nz + k - 1], gze4 = a.gz[7 * nz + k - 1];
    const float r20x = a.r2[0], r20y = a.r2[1], r20z = a.r2[2], dt = a.dt;

    const f4v sxx = ld4(Sxx + c);
    const float sxxr = Sxx[c + 4];
    const f4v syy = ld4(Syy + c), syyp = ld4(Syy + c + px);
    const f4v szz = ld4(Szz + c), szzp = ld4(Szz + c + sxy);
    const f4v syz = ld4(Syz + c), syzym = ld4(Syz + c - px), syzzm = ld4(Syz + c - sxy);
    const f4v sxz = ld4(Sxz + c), sxzzm = ld4(Sxz + c - sxy);
    const float sxzl = Sxz[c - 1];
    const f4v sxy_ = ld4(Sxy + c), sxyym = ld4(Sxy + c - px);
    const float sxyl = Sxy[c - 1];
    const f4v bx = med4<S16>(a, MBX, c, k), by = med4<S16>(a, MBY, c, k), bz = med4<S16>(a, MBZ, c, k);
    f4v vx = ld4nt(a.V[0] + c), vy = ld4nt(a.V[1] + c), vz = ld4nt(a.V[2] + c);
    f4v A[9];
#pragma unroll
    for (int q = 0; q < 9; q++) A[q] = ldA4<S16>(a, 9 + q, c);
    const float xxx[5] = {sxx.x, sxx.y, sxx.z, sxx.w, sxxr};
    const float xxz[5] = {sxzl, sxz.x, sxz.y, sxz.z, sxz.w};
    const float xxy[5] = {sxyl, sxy_.x, sxy_.y, sxy_.z, sxy_.w};
#pragma unroll
    for (int q = 0; q < 4; q++) {
        // m_absorb_p.f90:266-298
        const float dxSxx = (xxx[q + 1] - xxx[q]) * r20x;
        const float dySyy = (syyp[q] - syy[q]) * r20y;
        const float dzSzz = (szzp[q] - szz[q]) * r20z;
        const float dySyz = (syz[q] - syzym[q]) * r20y;
        const float dzSyz = (syz[q] - syzzm[q]) * r20z;
        const float dxSxz = (xxz[q + 1] - xxz[q]) * r20x;
        const float dzSxz = (sxz[q] - sxzzm[q]) * r20z;
        const float dxSxy = (xxy[q + 1] - xxy[q]) * r20x;
        const float dySxy = (sxy_[q] - sxyym[q]) * r20y;
        const float axsxx = A[0][q], aysxy = A[1][q], azsxz = A[2][q];
        const float axsxy = A[3][q], aysyy = A[4][q], azsyz = A[5][q];
        const float axsxz = A[6][q], aysyz = A[7][q], azszz = A[8][q];
        vx[q] = vx[q] + bx[q] * (gxe1[q] * dxSxx + gyc1 * dySxy + gzc1 * dzSxz
                                 + gxe2[q] * axsxx + gyc2 * aysxy + gzc2 * azsxz) * dt;
        vy[q] = vy[q] + by[q] * (gxc1[q] * dxSxy + gye1 * dySyy + gzc1 * dzSyz
                                 + gxc2[q] * axsxy + gye2 * aysyy + gzc2 * azsyz) * dt;
        vz[q] = vz[q] + bz[q] * (gxc1[q] * dxSxz + gyc1 * dySyz + gze1 * dzSzz
                                 + gxc2[q] * axsxz + gyc2 * aysyz + gze2 * azszz) * dt;
        A[0][q] = gxe3[q] * axsxx + gxe4[q] * dxSxx * dt;
        A[1][q] = gyc3 * aysxy + gyc4 * dySxy * dt;
        A[2][q] = gzc3 * azsxz + gzc4 * dzSxz * dt;
        A[3][q] = gxc3[q] * axsxy + gxc4[q] * dxSxy * dt;
        A[4][q] = gye3 * aysyy + gye4 * dySyy * dt;
        A[5][q] = gzc3 * azsyz + gzc4 * dzSyz * dt;
        A[6][q] = gxc3[q] * axsxz + gxc4[q] * dxSxz * dt;
        A[7][q] = gyc3 * aysyz + gyc4 * dySyz * dt;
        A[8][q] = gze3 * azszz + gze4 * dzSzz * dt;
    }
    st4m(a.V[0] + c, vx, msk);
    st4m(a.V[1] + c, vy, msk);
    st4m(a.V[2] + c, vz, msk);
#pragma unroll
    for (int q = 0; q < 9; q++) stA4<S16>(a, 9 + q, c, A[q], msk);
}

// ---------------------------------------------------------------------------
// sources, receivers, vmax
// ---------------------------------------------------------------------------
// Sources are applied in the reference's sequential order by one thread
// (m_source.f90:796 `loop seq`): several sources may share a cell.
template <typename T>
__device__ __forceinline__ void stressglut_one(const K3<T> &a, int nsrc, const int *__restrict__ src,
                                               const double *__restrict__ m6, const double *__restrict__ sdrop_all, int nt,
                                               const int *__restrict__ dit)
{
    const int it = *dit;
    if (it < 1 || it > nt) return;
    const double *__restrict__ sdrop = sdrop_all + (size_t)(it - 1) * nsrc;
    const int px = a.px, sxy = a.sxy;
    for (int n = 0; n < nsrc; n++) {
        const int ii = src[3 * n], jj = src[3 * n + 1], kk = src[3 * n + 2];
        const T sd = (T)sdrop[n];
        const T mxx = (T)m6[6 * n + 0], myy = (T)m6[6 * n + 1], mzz = (T)m6[6 * n + 2];
        const T myz = (T)m6[6 * n + 3], mxz = (T)m6[6 * n + 4], mxy = (T)m6[6 * n + 5];
        const int c = at3(px, sxy, ii, jj, kk);
        a.S[0][c] -= mxx * sd;
        a.S[1][c] -= myy * sd;
        a.S[2][c] -= mzz * sd;
        a.S[5][c] -= mxy * sd / 4;
        a.S[5][c - px] -= mxy * sd / 4;
        a.S[5][c - 1] -= mxy * sd / 4;
        a.S[5][c - 1 - px] -= mxy * sd / 4;
        a.S[4][c] -= mxz * sd / 4;
        a.S[4][c - sxy] -= mxz * sd / 4;
        a.S[4][c - 1] -= mxz * sd / 4;
        a.S[4][c - 1 - sxy] -= mxz * sd / 4;
        a.S[3][c] -= myz * sd / 4;
        a.S[3][c - sxy] -= myz * sd / 4;
        a.S[3][c - px] -= myz * sd / 4;
        a.S[3][c - px - sxy] -= myz * sd / 4;
    }
}

template <typename T>
__global__ void k3_stressglut(const K3<T> a, int nsrc, const int *__restrict__ src, const double *__restrict__ m6,
                              const double *__restrict__ sdrop_all, int nt, const int *__restrict__ dit)
{
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    stressglut_one<T>(a, nsrc, src, m6, sdrop_all, nt, dit);
}

template <typename T, bool S16>
__global__ void k3_bodyforce(const K3<T> a, int nsrc, const int *__restrict__ src, const double *__restrict__ f3,
                             const double *__restrict__ fdt_all, int nt, const int *__restrict__ dit)
{
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    const int it = *dit;
    if (it < 1 || it > nt) return;
    const double *__restrict__ fdt = fdt_all + (size_t)(it - 1) * nsrc;
    const int px = a.px, sxy = a.sxy;
    for (int n = 0; n < nsrc; n++) {
        const int ii = src[3 * n], jj = src[3 * n + 1], kk = src[3 * n + 2];
        const T f = (T)fdt[n];
        const T fx = (T)f3[3 * n], fy = (T)f3[3 * n + 1], fz = (T)f3[3 * n + 2];
        const int c = at3(px, sxy, ii, jj, kk);
        a.V[0][c] += medv<S16>(a, MBX, c, kk) * fx * f;
        a.V[0][c - 1] += medv<S16>(a, MBX, c - 1, kk) * fx * f;
        a.V[1][c] += medv<S16>(a, MBY, c, kk) * fy * f;
        a.V[1][c - px] += medv<S16>(a, MBY, c - px, kk) * fy * f;
        a.V[2][c] += medv<S16>(a, MBZ, c, kk) * fz * f;
        a.V[2][c - sxy] += medv<S16>(a, MBZ, c - sxy, kk - 1) * fz * f;
    }
}

// wav__store (m_wav.f90:436-447, 524-535): disp accumulates every step in
// FP64; at it with mod(it-1, ntdec_w) = 0 the sample itw = (it-1)/ntdec_w
// is written (while itw < ntw).  The step number comes from the device
// counter dit, so a captured step replays unchanged.
template <typename T>
__device__ __forceinline__ void receiver_one(const K3<T> &a, int n, int nst, const int *__restrict__ st,
                                             double *__restrict__ uacc, float *__restrict__ wvel,
                                             double *__restrict__ wdisp, int ntdec_w, int ntw, const int *__restrict__ dit)
{
    const int it = *dit;
    const int itw = (it - 1) / ntdec_w;
    const int store = ((it - 1) % ntdec_w) == 0 && itw < ntw;
    float *__restrict__ vel_out = wvel + (size_t)(store ? itw : 0) * nst * 3;
    double *__restrict__ disp_out = wdisp + (size_t)(store ? itw : 0) * nst * 3;
    const int px = a.px, sxy = a.sxy;
    const int c = at3(px, sxy, st[3 * n], st[3 * n + 1], st[3 * n + 2]);
    const T vx = a.V[0][c] + a.V[0][c - 1];
    const T vy = a.V[1][c] + a.V[1][c - px];
    const T vz = a.V[2][c] + a.V[2][c - sxy];
    const double hdt = 0.5 * (double)a.dt;
    double ux = uacc[3 * n] + (double)vx * hdt;
    double uy = uacc[3 * n + 1] + (double)vy * hdt;
    double uz = uacc[3 * n + 2] - (double)vz * hdt;
    uacc[3 * n] = ux;
    uacc[3 * n + 1] = uy;
    uacc[3 * n + 2] = uz;
    if (store) {
        vel_out[3 * n] = (float)(vx / 2);
        vel_out[3 * n + 1] = (float)(vy / 2);
        vel_out[3 * n + 2] = (float)(-vz / 2);
        disp_out[3 * n] = ux;
        disp_out[3 * n + 1] = uy;
        disp_out[3 * n + 2] = uz;
    }
}

template <typename T>
__global__ void k3_receivers(const K3<T> a, int nst, const int *__restrict__ st, double *__restrict__ uacc,
                             float *__restrict__ wvel, double *__restrict__ wdisp, int ntdec_w, int ntw,
                             const int *__restrict__ dit)
{
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= nst) return;
    receiver_one<T>(a, n, nst, st, uacc, wvel, wdisp, ntdec_w, ntw, dit);
}

__global__ void k_fill_u32(unsigned *p, int n, unsigned v)
{
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) p[t] = v;
}

// Device-side step number of the time loop: set once per run, advanced by
// one at the end of every step, read by the kernels whose work depends on
// it (sources, receivers, snapshots).  A captured step graph therefore
// replays unchanged for any it.
__global__ void k_step_set(int *dit, int it)
{
    if (blockIdx.x == 0 && threadIdx.x == 0) *dit = it;
}

__global__ void k_step_advance(int *dit)
{
    if (blockIdx.x == 0 && threadIdx.x == 0) *dit = *dit + 1;
}

template <typename T>
__global__ void k3_vmax(const K3<T> a, const int *__restrict__ kob, unsigned *__restrict__ out)
{
    const int margin = 5;
    const int i = a.na + margin + 1 + (int)(blockIdx.x * blockDim.x + threadIdx.x);
    const int j = a.na + margin + 1 + (int)blockIdx.y;
    float m0 = 0, m1 = 0, m2 = 0;
    if (i <= a.nx - a.na - margin && j <= a.ny - a.na - margin) {
        const int k = kob[(j - 1) * a.nx + (i - 1)] + 1;
        const int c = at3(a.px, a.sxy, i, j, k);
        m0 = fabsf((float)a.V[0][c]);
        m1 = fabsf((float)a.V[1][c]);
        m2 = fabsf((float)a.V[2][c]);
    }
    // wavefront max, then one atomic per wave (non-negative floats order as integers)
    for (int off = 32; off > 0; off >>= 1) {
        m0 = fmaxf(m0, __shfl_xor(m0, off));
        m1 = fmaxf(m1, __shfl_xor(m1, off));
        m2 = fmaxf(m2, __shfl_xor(m2, off));
    }
    if ((threadIdx.x & 63) == 0) {
        atomicMax(out + 0, __float_as_uint(m0));
        atomicMax(out + 1, __float_as_uint(m1));
        atomicMax(out + 2, __float_as_uint(m2));
    }
}

// Snapshot products at a list of cells (snap__write and its wbuf_* loops,
// m_snap.f90:911-2185).  kind 0 ('ps'): divergence and rotation from
// one-sided 2nd-order differences, masked by lam / mu_ij and scaled by
// UC*M0*1e-3 (m_snap.f90:1007-1027); kind 1 ('v'): V*UC*M0 (1469-1472);
// kind 2 ('u'): running sum of V*UC*M0*dt, advanced at every call
// (2046-2048).  Output planes are variable-major: out[n*npts + p].
// With peak != nullptr a 'v' or 'u' stream also keeps, at every call, the
// running maxima of wbuf_fs_v/wbuf_ob_v/wbuf_*_u (m_snap.f90:1690-1703):
// peak[p] = max |Z|, peak[npts+p] = max sqrt(X^2+Y^2) and
// peak[2*npts+p] = sqrt of the sum of their squares.  A closing call
// (snap__closefiles' wbuf_*(nt+1), m_snap.f90:2215-2231) advances the
// running sums and maxima but stores no frame.
// Evaluated in the reference's order without contraction.
__device__ __forceinline__ void snap_peak(float *__restrict__ peak, int p, int npts, float x, float y, float z)
{
#pragma clang fp contract(off)
    const float pv = fmaxf(peak[p], fabsf(z));
    const float ph = fmaxf(peak[npts + p], sqrtf(x * x + y * y));
    peak[p] = pv;
    peak[npts + p] = ph;
    peak[2 * npts + p] = sqrtf(pv * pv + ph * ph);
}

template <typename T, bool S16>
__global__ void k3_snap(const K3<T> a, int kind, int npts, const int4 *__restrict__ pts, float uc, float m0, float dt,
                        float *__restrict__ acc, float *__restrict__ frames, float *__restrict__ peak, int ntdec,
                        int nbuf, int closing, const int *__restrict__ dit)
{
#pragma clang fp contract(off)
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= npts) return;
    // frame (it-1)/ntdec at mod(it-1, ntdec) = 0, kept in slot frame % nbuf
    const int it = *dit;
    const int store = !closing && ((it - 1) % ntdec) == 0;
    if (!store && kind != 2 && !(peak && kind == 1)) return;
    const int nvar = kind == 0 ? 4 : 3;
    float *__restrict__ out = frames + (size_t)(store ? ((it - 1) / ntdec) % nbuf : 0) * nvar * npts;
    const int4 q = pts[p];
    const int px = a.px, sxy = a.sxy;
    const int c = at3(px, sxy, q.x, q.y, q.z);
    const T *vx = a.V[0], *vy = a.V[1], *vz = a.V[2];
    if (kind == 0) {
        const T rx = a.r2[0], ry = a.r2[1], rz = a.r2[2];
        float dv = (float)((vx[c] - vx[c - 1]) * rx + (vy[c] - vy[c - px]) * ry + (vz[c] - vz[c - sxy]) * rz);
        float r1 = (float)((vz[c + px] - vz[c]) * ry - (vy[c + sxy] - vy[c]) * rz);
        float r2 = (float)((vx[c + sxy] - vx[c]) * rz - (vz[c + 1] - vz[c]) * rx);
        float r3 = (float)((vy[c + 1] - vy[c]) * rx - (vx[c + px] - vx[c]) * ry);
        const float eps = 1.1920929e-07f;  // epsilon(1.0)
        const float lam = (float)medv<S16>(a, MLAM, c, q.z), myz = (float)medv<S16>(a, MMUYZ, c, q.z);
        const float mxz = (float)medv<S16>(a, MMUXZ, c, q.z), mxy = (float)medv<S16>(a, MMUXY, c, q.z);
        dv = dv * lam / fabsf(lam + eps);
        r1 = r1 * myz / fabsf(myz + eps);
        r2 = r2 * mxz / fabsf(mxz + eps);
        r3 = r3 * mxy / fabsf(mxy + eps);
        out[p] = dv * uc * m0 * 1e-3f;
        out[npts + p] = r1 * uc * m0 * 1e-3f;
        out[2 * npts + p] = r2 * uc * m0 * 1e-3f;
        out[3 * npts + p] = r3 * uc * m0 * 1e-3f;
    } else if (kind == 1) {
        const float bx = (float)(vx[c] * (T)uc * (T)m0);
        const float by = (float)(vy[c] * (T)uc * (T)m0);
        const float bz = (float)(vz[c] * (T)uc * (T)m0);
        if (store) {
            out[p] = bx;
            out[npts + p] = by;
            out[2 * npts + p] = bz;
        }
        if (peak) snap_peak(peak, p, npts, bx, by, bz);
    } else {
        const T *v[3] = {vx, vy, vz};
        float u[3];
#pragma unroll
        for (int n = 0; n < 3; n++) {
            u[n] = (float)((T)acc[n * npts + p] + v[n][c] * (T)uc * (T)m0 * (T)dt);
            acc[n * npts + p] = u[n];
            if (store) out[n * npts + p] = u[n];
        }
        if (peak) snap_peak(peak, p, npts, u[0], u[1], u[2]);
    }
}

// ---------------------------------------------------------------------------
// context
// ---------------------------------------------------------------------------
struct Ctx3Base {
    virtual ~Ctx3Base() {}
};

}  // namespace swpc

struct swpc3d_ctx {
    swpc3d_desc d;
    swpc::Lay3 L;
    size_t esz;  // bytes per state element (4 or 8)
    hipStream_t stream = nullptr;
    long long bytes = 0;
    swpc::DBuf V[3], A[18];
    swpc::DBuf Rall;        // the 18 memory variables, interleaved by 256-cell tiles (rix)
    swpc::DBuf Sall;        // the 6 stresses, interleaved by 256-cell tiles (STile)
    long long rtiles = 0;   // tiles of Rall and of Sall
    swpc::DBuf med[10];     // bx by bz lam mu taup taus muyz muxz muxy
    swpc::DBuf mbs[10];     // reduced storage: per-plane (base, scale)
    swpc::DBuf cpl[10];     // FP32 medium: per-plane constant values (K3::cpl)
    bool use_cpl[10] = {};
    swpc::DBuf rho;         // optional density (swpc3d_set_density)
    bool lab = false;       // learned absorbing boundary on (swpc3d_set_lab)
    swpc::DBuf lab_w;       // its ten nets as the fill kernel's LDS images (BF16 values, nearest even)
    swpc::DBuf lab_colmed;  // per-column medium features
    swpc::LabDomain lab_dom;
    // strip cache of the fill (9 fields x 6 strip cells x the x faces'
    // columns): a phase reads the other phase's strips from it while
    // lab_ok[phase]
    swpc::DBuf lab_cache;
    bool lab_ok[2] = {false, false};
    bool lab_cache_on = true;   // tuning entry 23
    float lab_sponge = swpc::LAB_SPONGE;  // strip sponge strength (swpc3d_set_lab_sponge)
    bool s16 = false;       // memory variables in 16 bit (SWPC_FP32_S16, SWPC_FP32_R16)
    bool m16 = false;       // medium in 16 bit too (SWPC_FP32_S16)
    bool a16 = false;       // ADE arrays in BF16 (SWPC_FP32_S16)
    float r_scale = 1.0f;   // memory-variable scale of the reduced storage
    float r_scale0 = 1.0f;  // ... as derived from the sources, before r_shift
    int r_shift = 0;        // log2 adjustment of r_scale (tuning entry 18)
    swpc::DBuf surf, kob, gx, gy, gz, gxv;
    swpc::DBuf src, m6, sdrop, bsrc, f3, fdt;
    swpc::DBuf st, uacc, wvel, wdisp, vmax;
    int nsrc = 0, nt_src = 0, nbf = 0, nt_bf = 0;
    int nst = 0, ntdec_w = 1, ntw = 0;
    bool have_medium = false, have_surface = false, have_pml = false;
    swpc::Boxes3 boxes;
    swpc::Boxes3 boxes_x;   // the boxes not walked by rows (the x faces)
    swpc::Rows3 rows;       // the boxes of whole rows i in [1, nx]
    swpc::Boxes4 boxes4;
    int gxp = 0;
    int ti0 = 1;
    // interior launch configuration (swpc3d_set_tuning)
    int s_by = 4, s_wpe = 3, s_split = 0, v_by = 4, v_wpe = 5, xcd = 2, s_kc = 32, v_kc = 0;
    int x4 = 1, x4_by = 2;  // 4-wide FP32 / reduced-storage interior kernels
    int x4_form = 0;        // 4-wide stress: 0 normal + shear launches, 1 one pass (staged), 2 one pass (free)
    int x4_vby = 2;         // tile rows of the z-marching velocity kernel (1, 2, 4, 8)
    int x4_vz = 0;          // 4-wide velocity: 0 one plane per workgroup row, > 0 z-marching chunk length
    int x4_sz = 0;          // 4-wide stress: > 0 one-pass z-marching kernel with this chunk length
    int pml4 = 0;           // 4-wide PML kernels (fewer waves per SIMD; off by default)
    int pml_rows = 1;       // whole-row absorber boxes by row segments (tuning entry 19)
    int pml_xcd = 0;        // PML workgroups XCD-contiguous (tuning entry 20)
    int x4_vlds = 0;        // z-marching velocity with LDS-staged y rows (tuning entry 21)
    int xs = 2;             // x neighbours by wavefront shifts: bit 0 stress, bit 1 velocity (tuning entry 22)
    bool ade_sparse = true; // no ADE array has been set from outside since creation
    // The PML kernels run on a second stream, concurrently with the
    // interior kernels of the same half-step (they write disjoint cells and
    // read the same inputs); events fork and join the two streams.
    int conc = 0;
    int graph_k = 16;       // steps per captured step graph (0 = launch step by step)
    int sr = 0;             // stochastic rounding of the 16-bit stores (reduced storage)
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    unsigned gen = 0;       // bumped by every graph drop (batches recapture on change)
    swpc::DBuf dit;         // device step number (int)
    hipStream_t side = nullptr;
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
    swpc::Prof prof;
    struct Snap {
        int kind, npts, ntdec, nbuf;
        float uc, m0, dt;
        swpc::DBuf pts, acc, frames, peak;
        long long last = -1;  // latest frame index written
    };
    std::vector<Snap> snaps;
};

namespace swpc {

template <typename T>
static K3<T> make_args(swpc3d_ctx *c)
{
    K3<T> a;
    memset(&a, 0, sizeof a);
    const swpc3d_desc &d = c->d;
    a.px = c->L.px; a.sxy = c->L.sxy; a.nx = d.nx; a.ny = d.ny; a.nz = d.nz; a.na = d.na;
    for (int q = 0; q < 3; q++) a.V[q] = (T *)c->V[q].p;
    for (int q = 0; q < 6; q++) a.S[q].p = (T *)c->Sall.p + (size_t)q * RTILE;
    for (int q = 0; q < 18; q++) a.R[q] = c->s16 ? nullptr : (T *)c->Rall.p + (size_t)q * RTILE;
    for (int q = 0; q < 18; q++) a.A[q] = (T *)c->A[q].p;
    for (int q = 0; q < 10; q++) a.med[q] = (const float *)c->med[q].p;
    for (int q = 0; q < 10; q++) a.cpl[q] = !c->m16 && c->use_cpl[q] ? (const float *)c->cpl[q].p : nullptr;
    a.rho = c->m16 ? nullptr : (const float *)c->rho.p;
    a.m16 = c->m16 ? 1 : 0;
    if (c->s16) {
        for (int q = 0; q < 10; q++) {
            a.hmed[q] = (const __half *)c->med[q].p;
            a.mbs[q] = (const float2 *)c->mbs[q].p;
        }
        a.hR = (__half *)c->Rall.p;
        for (int q = 0; q < 18; q++) a.hA[q] = (__hip_bfloat16 *)c->A[q].p;
        a.r_scale = c->r_scale;
        a.r_inv = 1.0f / c->r_scale;
        a.sr = c->sr;
    }
    a.surf = (const int4 *)c->surf.p;
    a.gx = (const T *)c->gx.p; a.gy = (const T *)c->gy.p; a.gz = (const T *)c->gz.p;
    a.gxv = (const float *)c->gxv.p;
    a.gxp = c->gxp;
    // m_kernel.f90:43-54 evaluated in T, then re40 = rc40 + isign*rd40
    const double h[3] = {d.dx, d.dy, d.dz};
    for (int q = 0; q < 3; q++) {
        T hh = (T)h[q];
        T rc40 = (T)17.0 / (T)16.0 / hh, rc41 = (T)1.0 / (T)48.0 / hh;
        T rd40 = (T)(-1.0) / (T)16.0 / hh, rd41 = (T)(-1.0) / (T)48.0 / hh;
        a.e4[q][0] = rc40 + (-1) * rd40; a.e4[q][1] = rc41 + (-1) * rd41;
        a.e2[q][0] = rc40 + (+1) * rd40; a.e2[q][1] = rc41 + (+1) * rd41;
        a.r2[q] = (T)1.0 / hh;
    }
    // m_kernel.f90:58-67
    T t = (T)(float)d.dt, s = 0;
    for (int m = 0; m < NM3; m++) {
        T tm = (T)d.ts[m];
        a.c1[m] = (2 * tm - t) / (2 * tm + t);
        a.c2[m] = (T)2 / (2 * tm + t) / NM3;
        s += t / (2 * tm - t);
        a.d1[m] = 2 * tm / (2 * tm - t);
    }
    a.d2 = s / NM3;
    a.dt = t;
    a.ti0 = c->ti0;
    a.pml = c->boxes;
    a.xcd = c->xcd == 1;  // the y-group orders (>= 2) are the z-marching velocity update's (x4_vel)
    a.ade_sparse = c->ade_sparse ? 1 : 0;
    return a;
}


// Interior tiling for a 64 x by tile and a z-chunk length kc (0 = chosen
// so that the launch holds at least ~4096 workgroups, 16 per CU).
template <typename T>
static void set_tiles(const swpc3d_ctx *c, K3<T> &a, int by, int kc)
{
    const swpc3d_desc &d = c->d;
    const int ihi = d.nx - d.na;
    a.ntx = ihi >= d.na + 1 ? (ihi - c->ti0) / 64 + 1 : 0;
    const int nj = d.ny - 2 * d.na;
    a.nty = nj > 0 ? (nj + by - 1) / by : 0;
    const int nk = d.nz - d.na;
    if (kc <= 0) {
        long long cols = std::max(1LL, (long long)a.ntx * a.nty);
        int want = (int)std::max(1LL, std::min((long long)std::max(1, nk / 16), (4096 + cols - 1) / cols));
        kc = (nk + want - 1) / want;
    }
    a.kc = std::max(1, kc);
    a.nkc = nk > 0 ? (nk + a.kc - 1) / a.kc : 0;
    a.nblk = (unsigned)a.ntx * (unsigned)a.nty * (unsigned)a.nkc;
    a.nblk_pad = (a.nblk + 7u) & ~7u;
}

template <typename T, int BY, int PART, int WPE, bool S16 = false>
static void go_stress(const K3<T> &a, hipStream_t s)
{
    hipLaunchKernelGGL((k3_stress_int<T, BY, PART, WPE, S16>), dim3(a.xcd ? a.nblk_pad : a.nblk), dim3(64, BY), 0, s, a);
}

template <typename T, int BY, int WPE, bool S16 = false>
static void go_vel(const K3<T> &a, hipStream_t s)
{
    hipLaunchKernelGGL((k3_vel_int<T, BY, WPE, S16>), dim3(a.xcd ? a.nblk_pad : a.nblk), dim3(64, BY), 0, s, a);
}

// Reduced-storage interior variants (FP32 arithmetic).
template <int PART>
static bool pick_stress16(int by, int wpe, const K3<float> &a, hipStream_t s)
{
#define SWPC_S(BY_, W_) if (by == BY_ && wpe == W_) { go_stress<float, BY_, PART, W_, true>(a, s); return true; }
    SWPC_S(2, 0) SWPC_S(4, 0) SWPC_S(4, 4) SWPC_S(8, 0)
#undef SWPC_S
    return false;
}

static bool pick_vel16(int by, int wpe, const K3<float> &a, hipStream_t s)
{
#define SWPC_V(BY_, W_) if (by == BY_ && wpe == W_) { go_vel<float, BY_, W_, true>(a, s); return true; }
    SWPC_V(2, 0) SWPC_V(4, 0) SWPC_V(8, 0)
#undef SWPC_V
    return false;
}

// Instantiated interior variants.  FP64 always runs the 64 x 4 fused form
// (the FP64 scheme is the accuracy reference, not a tuned path).
template <typename T, int PART>
static bool pick_stress(int by, int wpe, const K3<T> &a, hipStream_t s)
{
    if constexpr (sizeof(T) == 8) {
        go_stress<T, 4, PART, 0>(a, s);
        return true;
    } else {
#define SWPC_S(BY_, W_) if (by == BY_ && wpe == W_) { go_stress<T, BY_, PART, W_>(a, s); return true; }
        SWPC_S(2, 0) SWPC_S(2, 4) SWPC_S(4, 0) SWPC_S(4, 3) SWPC_S(4, 4) SWPC_S(8, 0) SWPC_S(8, 4)
#undef SWPC_S
        return false;
    }
}

template <typename T>
static bool pick_vel(int by, int wpe, const K3<T> &a, hipStream_t s)
{
    if constexpr (sizeof(T) == 8) {
        go_vel<T, 4, 0>(a, s);
        return true;
    } else {
#define SWPC_V(BY_, W_) if (by == BY_ && wpe == W_) { go_vel<T, BY_, W_>(a, s); return true; }
        SWPC_V(2, 0) SWPC_V(4, 0) SWPC_V(4, 5) SWPC_V(8, 0) SWPC_V(8, 5)
#undef SWPC_V
        return false;
    }
}

// Tiling of the 4-wide kernels (256-cell x tiles, by rows, one plane per
// workgroup row); returns the number of workgroups to launch.
static unsigned x4_tiles(const swpc3d_ctx *c, K3<float> &a, int by)
{
    const swpc3d_desc &d = c->d;
    a.ti0 = 1 + (d.na / 256) * 256;
    const int ihi = d.nx - d.na;
    a.ntx = ihi >= d.na + 1 ? (ihi - a.ti0) / 256 + 1 : 0;
    const int nj = d.ny - 2 * d.na;
    a.nty = nj > 0 ? (nj + by - 1) / by : 0;
    const int nk = d.nz - d.na;
    a.nkc = nk;
    a.xcd = c->xcd == 1;
    if (a.ntx <= 0 || a.nty <= 0 || nk <= 0) return 0;
    if (a.xcd) return 8u * (unsigned)a.ntx * (((unsigned)a.nty + 7u) / 8u) * (unsigned)nk;
    return (unsigned)a.ntx * (unsigned)a.nty * (unsigned)nk;
}

template <int BY, int BAR>
static void go_x4_fused(const K3<float> &a, unsigned nb, bool s16, bool xs, hipStream_t s)
{
    if (xs) {
        if (s16) hipLaunchKernelGGL((k3x_stress<BY, true, BAR, true>), dim3(nb), dim3(64, BY), 0, s, a);
        else hipLaunchKernelGGL((k3x_stress<BY, false, BAR, true>), dim3(nb), dim3(64, BY), 0, s, a);
        return;
    }
    if (s16) hipLaunchKernelGGL((k3x_stress<BY, true, BAR, false>), dim3(nb), dim3(64, BY), 0, s, a);
    else hipLaunchKernelGGL((k3x_stress<BY, false, BAR, false>), dim3(nb), dim3(64, BY), 0, s, a);
}

template <int BY>
static void go_x4_velz(const K3<float> &a, unsigned nb, bool s16, bool xs, hipStream_t s)
{
    // V and S are FP32 in every 4-wide mode: with a density the kernel
    // reads no medium array at all
    if (xs) {
        if (a.rho) hipLaunchKernelGGL((k3x_velz<BY, false, true, true>), dim3(nb), dim3(64, BY), 0, s, a);
        else if (s16) hipLaunchKernelGGL((k3x_velz<BY, true, false, true>), dim3(nb), dim3(64, BY), 0, s, a);
        else hipLaunchKernelGGL((k3x_velz<BY, false, false, true>), dim3(nb), dim3(64, BY), 0, s, a);
        return;
    }
    if (a.rho) hipLaunchKernelGGL((k3x_velz<BY, false, true, false>), dim3(nb), dim3(64, BY), 0, s, a);
    else if (s16) hipLaunchKernelGGL((k3x_velz<BY, true, false, false>), dim3(nb), dim3(64, BY), 0, s, a);
    else hipLaunchKernelGGL((k3x_velz<BY, false, false, false>), dim3(nb), dim3(64, BY), 0, s, a);
}

template <int BY, int NB>
static void go_x4_velzl(const K3<float> &a, unsigned nb, bool s16, hipStream_t s)
{
    if (a.rho) hipLaunchKernelGGL((k3x_velzl<BY, NB, false, true>), dim3(nb), dim3(64, BY), 0, s, a);
    else if (s16) hipLaunchKernelGGL((k3x_velzl<BY, NB, true, false>), dim3(nb), dim3(64, BY), 0, s, a);
    else hipLaunchKernelGGL((k3x_velzl<BY, NB, false, false>), dim3(nb), dim3(64, BY), 0, s, a);
}

template <int BY>
static void go_x4_stress(const K3<float> &a, unsigned nb, bool s16, hipStream_t s)
{
    if (s16) {
        hipLaunchKernelGGL((k3x_normal<BY, true>), dim3(nb), dim3(64, BY), 0, s, a);
        hipLaunchKernelGGL((k3x_shear<BY, true>), dim3(nb), dim3(64, BY), 0, s, a);
    } else {
        hipLaunchKernelGGL((k3x_normal<BY, false>), dim3(nb), dim3(64, BY), 0, s, a);
        hipLaunchKernelGGL((k3x_shear<BY, false>), dim3(nb), dim3(64, BY), 0, s, a);
    }
}

template <int BY>
static void go_x4_vel(const K3<float> &a, unsigned nb, bool s16, hipStream_t s)
{
    if (s16) hipLaunchKernelGGL((k3x_vel<BY, true>), dim3(nb), dim3(64, BY), 0, s, a);
    else hipLaunchKernelGGL((k3x_vel<BY, false>), dim3(nb), dim3(64, BY), 0, s, a);
}

static bool x4_stress(const swpc3d_ctx *c, K3<float> a, hipStream_t s)
{
    if (c->x4_sz > 0) {
        // z-marching one-pass stress: 64 x 2 tiles, chunks of x4_sz planes
        x4_tiles(c, a, 2);
        const int nk = c->d.nz - c->d.na;
        a.kc = c->x4_sz;
        a.nkc = (nk + a.kc - 1) / a.kc;
        if (a.ntx <= 0 || a.nty <= 0 || nk <= 0) return true;
        const unsigned nb = (unsigned)a.ntx * (unsigned)a.nty * (unsigned)a.nkc;
        if (c->s16) hipLaunchKernelGGL((k3x_stressz<2, true>), dim3(nb), dim3(64, 2), 0, s, a);
        else hipLaunchKernelGGL((k3x_stressz<2, false>), dim3(nb), dim3(64, 2), 0, s, a);
        return true;
    }
    if (c->x4_form != 0) {
        // one-pass normal + shear kernel (64 x by tiles)
        const unsigned nb = x4_tiles(c, a, c->x4_by);
        if (nb == 0) return true;
        const bool bar = c->x4_form == 1;
        const bool xs = (c->xs & 1) != 0;
        switch (c->x4_by) {
        case 1: bar ? go_x4_fused<1, 1>(a, nb, c->s16, xs, s) : go_x4_fused<1, 0>(a, nb, c->s16, xs, s); return true;
        case 2: bar ? go_x4_fused<2, 1>(a, nb, c->s16, xs, s) : go_x4_fused<2, 0>(a, nb, c->s16, xs, s); return true;
        case 4: bar ? go_x4_fused<4, 1>(a, nb, c->s16, xs, s) : go_x4_fused<4, 0>(a, nb, c->s16, xs, s); return true;
        default: return false;
        }
    }
    const unsigned nb = x4_tiles(c, a, c->x4_by);
    if (nb == 0) return true;
    switch (c->x4_by) {
    case 1: go_x4_stress<1>(a, nb, c->s16, s); return true;
    case 2: go_x4_stress<2>(a, nb, c->s16, s); return true;
    case 4: go_x4_stress<4>(a, nb, c->s16, s); return true;
    default: return false;
    }
}

static bool x4_vel(const swpc3d_ctx *c, K3<float> a, hipStream_t s)
{
    if (c->x4_vz > 0) {
        // z-marching 4-wide velocity: 64 x vby tiles, chunks of x4_vz planes
        const int vby = c->x4_vby;
        x4_tiles(c, a, vby);
        const int nk = c->d.nz - c->d.na;
        a.kc = c->x4_vz;
        a.nkc = (nk + a.kc - 1) / a.kc;
        if (a.ntx <= 0 || a.nty <= 0 || nk <= 0) return true;
        a.nblk = (unsigned)a.ntx * (unsigned)a.nty * (unsigned)a.nkc;
        a.nblk_pad = (a.nblk + 7u) & ~7u;
        a.xcd = c->xcd;
        unsigned nb = a.xcd ? a.nblk_pad : a.nblk;
        if (a.xcd >= 2) {
            const unsigned G = (unsigned)a.xcd, ngrp = ((unsigned)a.nty + G - 1u) / G * (unsigned)a.nkc;
            nb = 8u * G * (unsigned)a.ntx * ((ngrp + 7u) / 8u);
        }
        if (c->x4_vlds) {
            // the LDS-staged form walks the tiles in order (no XCD remap)
            switch (vby) {
            case 2: go_x4_velzl<2, 2>(a, a.nblk, c->s16, s); return true;
            case 4: go_x4_velzl<4, 2>(a, a.nblk, c->s16, s); return true;
            case 8: go_x4_velzl<8, 1>(a, a.nblk, c->s16, s); return true;
            default: return false;
            }
        }
        switch (vby) {
        case 1: go_x4_velz<1>(a, nb, c->s16, (c->xs & 2) != 0, s); return true;
        case 2: go_x4_velz<2>(a, nb, c->s16, (c->xs & 2) != 0, s); return true;
        case 4: go_x4_velz<4>(a, nb, c->s16, (c->xs & 2) != 0, s); return true;
        case 8: go_x4_velz<8>(a, nb, c->s16, (c->xs & 2) != 0, s); return true;
        default: return false;
        }
    }
    const unsigned nb = x4_tiles(c, a, c->x4_by);
    if (nb == 0) return true;
    switch (c->x4_by) {
    case 1: go_x4_vel<1>(a, nb, c->s16, s); return true;
    case 2: go_x4_vel<2>(a, nb, c->s16, s); return true;
    case 4: go_x4_vel<4>(a, nb, c->s16, s); return true;
    default: return false;
    }
}

// Second stream for the PML kernels: it waits for everything enqueued on
// the main stream so far; join() makes the main stream wait for it.
static int fork_side(swpc3d_ctx *c, hipStream_t &ps)
{
    ps = c->stream;
    if (!c->conc || c->boxes.total == 0) return SWPC_OK;
    SWPC_HIP(hipEventRecord(c->ev_fork, c->stream));
    SWPC_HIP(hipStreamWaitEvent(c->side, c->ev_fork, 0));
    ps = c->side;
    return SWPC_OK;
}

static int join_side(swpc3d_ctx *c, hipStream_t ps)
{
    if (ps == c->stream) return SWPC_OK;
    SWPC_HIP(hipEventRecord(c->ev_join, ps));
    SWPC_HIP(hipStreamWaitEvent(c->stream, c->ev_join, 0));
    return SWPC_OK;
}

// What the learned-boundary kernels need of the context.
static LabGrid lab_grid(const swpc3d_ctx *c)
{
    LabGrid g;
    g.px = c->L.px;
    g.sxy = c->L.sxy;
    for (int q = 0; q < 3; q++) g.F[q] = c->V[q].p;
    for (int q = 0; q < 6; q++) g.F[3 + q] = (char *)c->Sall.p + (size_t)q * RTILE * c->esz;
    g.f64 = c->d.precision == SWPC_FP64 ? 1 : 0;
    g.rho = (const float *)c->rho.p;
    g.lam = (const float *)c->med[MLAM].p;
    g.mu = (const float *)c->med[MMU].p;
    g.dt = (float)c->d.dt;
    g.h[0] = (float)c->d.dx; g.h[1] = (float)c->d.dy; g.h[2] = (float)c->d.dz;
    return g;
}

// Ghost fill of the learned boundary before a half-step (phase 0: the
// velocity ghosts the stress update reads; 1: the stress ghosts), timed
// under the absorber class of that half-step.
static int launch_lab(swpc3d_ctx *c, int phase)
{
    if (!c->lab) return SWPC_OK;
    hipEvent_t e0;
    const int cls = phase == 0 ? 1 : 3;
    if (int r = prof_begin(c->prof, cls, c->stream, &e0)) return r;
    if (int r = lab_fill(lab_grid(c), c->lab_dom, c->lab_w.p, (const float4 *)c->lab_colmed.p, phase, c->lab_sponge,
                         c->lab_cache_on ? c->lab_cache.p : nullptr, c->lab_ok[phase], c->stream))
        return r;
    // this phase's strips are now cached for the other phase; the other
    // phase's cached strips go stale with the coming half-step
    c->lab_ok[1 - phase] = c->lab_cache_on && c->lab_cache.p != nullptr;
    c->lab_ok[phase] = false;
    return prof_end(c->prof, cls, c->stream, e0);
}

// Workgroups of a PML launch of nblk logical workgroups: with xcd, 8 runs
// of per = ceil(nblk / 8) (pml_block), else nblk in order (per = 0).
static unsigned pml_launch(unsigned nblk, int xcd, unsigned &per)
{
    per = xcd ? (nblk + 7u) / 8u : 0u;
    return xcd ? 8u * per : nblk;
}

template <typename T>
static int launch_stress(swpc3d_ctx *c)
{
    if (int r = launch_lab(c, 0)) return r;
    K3<T> a = make_args<T>(c);
    const bool f64 = sizeof(T) == 8;
    const int by = f64 ? 4 : c->s_by, wpe = f64 ? 0 : c->s_wpe;
    set_tiles(c, a, by, c->s_kc);
    hipStream_t ps;
    if (int r = fork_side(c, ps)) return r;
    hipEvent_t e0;
    if (c->boxes.total > 0) {
        if (int r = prof_begin(c->prof, 1, ps, &e0)) return r;
        bool done = false;
        if constexpr (sizeof(T) == 4) {
            if (c->pml4) {
                const unsigned nb = (c->boxes4.total + 255) / 256;
                if (c->a16) hipLaunchKernelGGL((k3x_stress_pml<true>), dim3(nb), dim3(256), 0, ps, a, c->boxes4);
                else hipLaunchKernelGGL((k3x_stress_pml<false>), dim3(nb), dim3(256), 0, ps, a, c->boxes4);
                done = true;
            }
        }
        if (!done) {
            K3<T> b = a;
            if (c->pml_rows && c->rows.total > 0) {
                unsigned per, nb = pml_launch(c->rows.total * (unsigned)c->rows.nseg, c->pml_xcd, per);
                if (c->a16) hipLaunchKernelGGL((k3_stress_pml_rows<T, true>), dim3(nb), dim3(256), 0, ps, a, c->rows, per);
                else hipLaunchKernelGGL((k3_stress_pml_rows<T, false>), dim3(nb), dim3(256), 0, ps, a, c->rows, per);
                b.pml = c->boxes_x;
            }
            if (b.pml.total > 0) {
                unsigned per, nb = pml_launch((b.pml.total + 255) / 256, c->pml_xcd, per);
                if (c->a16) hipLaunchKernelGGL((k3_stress_pml<T, true>), dim3(nb), dim3(256), 0, ps, b, per);
                else hipLaunchKernelGGL((k3_stress_pml<T, false>), dim3(nb), dim3(256), 0, ps, b, per);
            }
        }
        SWPC_CHECK_LAUNCH();
        if (int r = prof_end(c->prof, 1, ps, e0)) return r;
    }
    if (a.nblk > 0) {
        if (int r = prof_begin(c->prof, 0, c->stream, &e0)) return r;
        bool ok;
        if constexpr (sizeof(T) == 4) {
            if (c->x4) {
                ok = x4_stress(c, a, c->stream);
            } else if (c->s16) {
                if (c->s_split) {
                    ok = pick_stress16<1>(by, wpe, a, c->stream);
                    if (ok) ok = pick_stress16<2>(by, wpe, a, c->stream);
                } else {
                    ok = pick_stress16<0>(by, wpe, a, c->stream);
                }
            } else if (c->s_split) {
                ok = pick_stress<T, 1>(by, wpe, a, c->stream);
                if (ok) ok = pick_stress<T, 2>(by, wpe, a, c->stream);
            } else {
                ok = pick_stress<T, 0>(by, wpe, a, c->stream);
            }
        } else {
            ok = pick_stress<T, 0>(by, wpe, a, c->stream);
        }
        if (!ok) return fail(SWPC_EINVAL, "stress variant by=%d wpe=%d not built", c->s_by, c->s_wpe);
        SWPC_CHECK_LAUNCH();
        if (int r = prof_end(c->prof, 0, c->stream, e0)) return r;
    }
    return join_side(c, ps);
}

template <typename T>
static int launch_vel(swpc3d_ctx *c)
{
    if (int r = launch_lab(c, 1)) return r;
    K3<T> a = make_args<T>(c);
    const bool f64 = sizeof(T) == 8;
    const int by = f64 ? 4 : c->v_by, wpe = f64 ? 0 : c->v_wpe;
    set_tiles(c, a, by, c->v_kc);
    hipStream_t ps;
    if (int r = fork_side(c, ps)) return r;
    hipEvent_t e0;
    if (c->boxes.total > 0) {
        if (int r = prof_begin(c->prof, 3, ps, &e0)) return r;
        bool done = false;
        if constexpr (sizeof(T) == 4) {
            if (c->pml4) {
                const unsigned nb = (c->boxes4.total + 255) / 256;
                if (c->a16) hipLaunchKernelGGL((k3x_vel_pml<true>), dim3(nb), dim3(256), 0, ps, a, c->boxes4);
                else hipLaunchKernelGGL((k3x_vel_pml<false>), dim3(nb), dim3(256), 0, ps, a, c->boxes4);
                done = true;
            }
        }
        if (!done) {
            K3<T> b = a;
            if (c->pml_rows && c->rows.total > 0) {
                unsigned per, nb = pml_launch(c->rows.total * (unsigned)c->rows.nseg, c->pml_xcd, per);
                if (c->a16) hipLaunchKernelGGL((k3_vel_pml_rows<T, true>), dim3(nb), dim3(256), 0, ps, a, c->rows, per);
                else hipLaunchKernelGGL((k3_vel_pml_rows<T, false>), dim3(nb), dim3(256), 0, ps, a, c->rows, per);
                b.pml = c->boxes_x;
            }
            if (b.pml.total > 0) {
                unsigned per, nb = pml_launch((b.pml.total + 255) / 256, c->pml_xcd, per);
                if (c->a16) hipLaunchKernelGGL((k3_vel_pml<T, true>), dim3(nb), dim3(256), 0, ps, b, per);
                else hipLaunchKernelGGL((k3_vel_pml<T, false>), dim3(nb), dim3(256), 0, ps, b, per);
            }
        }
        SWPC_CHECK_LAUNCH();
        if (int r = prof_end(c->prof, 3, ps, e0)) return r;
    }
    if (a.nblk > 0) {
        if (int r = prof_begin(c->prof, 2, c->stream, &e0)) return r;
        bool ok;
        if constexpr (sizeof(T) == 4) {
            if (c->x4) ok = x4_vel(c, a, c->stream);
            else ok = c->s16 ? pick_vel16(by, wpe, a, c->stream) : pick_vel<T>(by, wpe, a, c->stream);
        }
        else ok = pick_vel<T>(by, wpe, a, c->stream);
        if (!ok)
            return fail(SWPC_EINVAL, "velocity variant by=%d wpe=%d not built", c->v_by, c->v_wpe);
        SWPC_CHECK_LAUNCH();
        if (int r = prof_end(c->prof, 2, c->stream, e0)) return r;
    }
    return join_side(c, ps);
}

template <typename T>
static int launch_glut(swpc3d_ctx *c)
{
    if (c->nsrc == 0) return SWPC_OK;
    K3<T> a = make_args<T>(c);
    hipLaunchKernelGGL((k3_stressglut<T>), dim3(1), dim3(64), 0, c->stream, a, c->nsrc, (const int *)c->src.p,
                       (const double *)c->m6.p, (const double *)c->sdrop.p, c->nt_src, (const int *)c->dit.p);
    SWPC_CHECK_LAUNCH();
    return SWPC_OK;
}

template <typename T>
static int launch_bf(swpc3d_ctx *c)
{
    if (c->nbf == 0) return SWPC_OK;
    K3<T> a = make_args<T>(c);
    if (c->s16)
        hipLaunchKernelGGL((k3_bodyforce<T, true>), dim3(1), dim3(64), 0, c->stream, a, c->nbf, (const int *)c->bsrc.p,
                           (const double *)c->f3.p, (const double *)c->fdt.p, c->nt_bf, (const int *)c->dit.p);
    else
        hipLaunchKernelGGL((k3_bodyforce<T, false>), dim3(1), dim3(64), 0, c->stream, a, c->nbf, (const int *)c->bsrc.p,
                           (const double *)c->f3.p, (const double *)c->fdt.p, c->nt_bf, (const int *)c->dit.p);
    SWPC_CHECK_LAUNCH();
    return SWPC_OK;
}

template <typename T>
static int launch_recv(swpc3d_ctx *c)
{
    if (c->nst == 0) return SWPC_OK;
    K3<T> a = make_args<T>(c);
    hipLaunchKernelGGL((k3_receivers<T>), dim3((c->nst + 63) / 64), dim3(64), 0, c->stream, a, c->nst,
                       (const int *)c->st.p, (double *)c->uacc.p, (float *)c->wvel.p, (double *)c->wdisp.p,
                       c->ntdec_w, c->ntw, (const int *)c->dit.p);
    SWPC_CHECK_LAUNCH();
    return SWPC_OK;
}

template <typename T>
static int launch_vmax(swpc3d_ctx *c)
{
    K3<T> a = make_args<T>(c);
    hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, c->stream, (unsigned *)c->vmax.p, 3, 0u);
    SWPC_CHECK_LAUNCH();
    int ni = c->d.nx - 2 * (c->d.na + 5), nj = c->d.ny - 2 * (c->d.na + 5);
    if (ni > 0 && nj > 0) {
        hipLaunchKernelGGL((k3_vmax<T>), dim3((ni + 255) / 256, nj), dim3(256), 0, c->stream, a,
                           (const int *)c->kob.p, (unsigned *)c->vmax.p);
        SWPC_CHECK_LAUNCH();
    }
    return SWPC_OK;
}

template <typename T>
static int launch_snap(swpc3d_ctx *c, int closing = 0)
{
    if (c->snaps.empty()) return SWPC_OK;
    K3<T> a = make_args<T>(c);
    for (auto &sn : c->snaps) {
        if (closing && sn.kind == SWPC_SNAP_PS) continue;
        if (c->s16)
            hipLaunchKernelGGL((k3_snap<T, true>), dim3((sn.npts + 255) / 256), dim3(256), 0, c->stream, a, sn.kind,
                               sn.npts, (const int4 *)sn.pts.p, sn.uc, sn.m0, sn.dt, (float *)sn.acc.p,
                               (float *)sn.frames.p, (float *)sn.peak.p, sn.ntdec, sn.nbuf, closing,
                               (const int *)c->dit.p);
        else
            hipLaunchKernelGGL((k3_snap<T, false>), dim3((sn.npts + 255) / 256), dim3(256), 0, c->stream, a, sn.kind,
                               sn.npts, (const int4 *)sn.pts.p, sn.uc, sn.m0, sn.dt, (float *)sn.acc.p,
                               (float *)sn.frames.p, (float *)sn.peak.p, sn.ntdec, sn.nbuf, closing,
                               (const int *)c->dit.p);
        SWPC_CHECK_LAUNCH();
    }
    return SWPC_OK;
}

// host record of the snapshot frames that steps it0 .. it1 produce
static void note_snapshots(swpc3d_ctx *c, int it0, int it1)
{
    for (auto &sn : c->snaps) {
        const long long last_due = (long long)(it1 - 1) / sn.ntdec * sn.ntdec + 1;  // largest due it <= it1
        if (last_due >= it0) sn.last = std::max(sn.last, (last_due - 1) / sn.ntdec);
    }
}

static int set_step(swpc3d_ctx *c, int it)
{
    hipLaunchKernelGGL(k_step_set, dim3(1), dim3(64), 0, c->stream, (int *)c->dit.p, it);
    SWPC_CHECK_LAUNCH();
    return SWPC_OK;
}

static bool fp64(const swpc3d_ctx *c) { return c->d.precision == SWPC_FP64; }

static int ready(const swpc3d_ctx *c)
{
    if (!c) return fail(SWPC_EINVAL, "null context");
    if (!c->have_medium || !c->have_surface || !c->have_pml)
        return fail(SWPC_ESTATE, "set_medium, set_surface and set_pml must precede time stepping");
    return SWPC_OK;
}

// One time step at the device step number (main.f90:125-144 order), then
// the counter advances.
template <typename T>
static int step_dev(swpc3d_ctx *c)
{
    int r;
    if ((r = launch_recv<T>(c))) return r;
    if ((r = launch_snap<T>(c))) return r;
    if ((r = launch_stress<T>(c))) return r;
    if ((r = launch_glut<T>(c))) return r;
    if ((r = launch_vel<T>(c))) return r;
    if ((r = launch_bf<T>(c))) return r;
    hipLaunchKernelGGL(k_step_advance, dim3(1), dim3(64), 0, c->stream, (int *)c->dit.p);
    SWPC_CHECK_LAUNCH();
    return SWPC_OK;
}

static int step_dev_any(swpc3d_ctx *c) { return fp64(c) ? step_dev<double>(c) : step_dev<float>(c); }

static void graph_drop(swpc3d_ctx *c)
{
    c->gen++;
    if (c->gexec) (void)hipGraphExecDestroy(c->gexec);
    if (c->graph) (void)hipGraphDestroy(c->graph);
    c->gexec = nullptr;
    c->graph = nullptr;
}

// Fields changed, or may have, outside the fixed order of a step: the
// learned-boundary fill's strip cache is not to be read, and a captured
// step sequence that would read it is dropped.
static void lab_stale(swpc3d_ctx *c)
{
    if (c->lab_ok[0] || c->lab_ok[1]) graph_drop(c);
    c->lab_ok[0] = c->lab_ok[1] = false;
}

// Capture graph_k steps (kernels and counter advances) once; the graph is
// dropped whenever a setter changes what the steps launch.
static int graph_ready(swpc3d_ctx *c)
{
    if (c->gexec) return SWPC_OK;
    SWPC_HIP(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    int r = SWPC_OK;
    for (int q = 0; q < c->graph_k && r == SWPC_OK; q++) r = step_dev_any(c);
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(c->stream, &g);
    if (r != SWPC_OK || e != hipSuccess || !g) c->lab_ok[0] = c->lab_ok[1] = false;
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

// Steps it0 .. it0+n-1: replays of the captured graph where they fit, one
// step at a time otherwise (and always while profiling, whose events are
// per launch, or with the PML on a second stream).
static int run_steps(swpc3d_ctx *c, int it0, int n)
{
    if (n <= 0) return SWPC_OK;
    if (c->nsrc > 0 && it0 + n - 1 > c->nt_src)
        return fail(SWPC_EINVAL, "stressglut: steps %d..%d outside 1..%d", it0, it0 + n - 1, c->nt_src);
    if (c->nbf > 0 && it0 + n - 1 > c->nt_bf)
        return fail(SWPC_EINVAL, "bodyforce: steps %d..%d outside 1..%d", it0, it0 + n - 1, c->nt_bf);
    int r;
    if ((r = set_step(c, it0))) return r;
    int done = 0;
    const bool graphs = c->graph_k > 0 && !c->prof.on && !c->conc;
    if (graphs && n >= c->graph_k) {
        if ((r = graph_ready(c))) return r;
        for (; done + c->graph_k <= n; done += c->graph_k) SWPC_HIP(hipGraphLaunch(c->gexec, c->stream));
    }
    for (; done < n; done++)
        if ((r = step_dev_any(c))) return r;
    note_snapshots(c, it0, it0 + n - 1);
    return SWPC_OK;
}

// Hooks of swpc_batch.hip.
int batch_member3(swpc3d_ctx *c, BatchMember &m)
{
    if (int r = ready(c)) return r;
    if (c->prof.on || c->conc) return fail(SWPC_ESTATE, "a batch member may not profile or run its PML on a second stream");
    m.stream = c->stream;
    m.dit = (int *)c->dit.p;
    m.gen = c->gen;
    return SWPC_OK;
}

int batch_check3(swpc3d_ctx *c, int it0, int n)
{
    if (c->nsrc > 0 && it0 + n - 1 > c->nt_src)
        return fail(SWPC_EINVAL, "stressglut: steps %d..%d outside 1..%d", it0, it0 + n - 1, c->nt_src);
    if (c->nbf > 0 && it0 + n - 1 > c->nt_bf)
        return fail(SWPC_EINVAL, "bodyforce: steps %d..%d outside 1..%d", it0, it0 + n - 1, c->nt_bf);
    return SWPC_OK;
}

int batch_step3(swpc3d_ctx *c) { return step_dev_any(c); }

void batch_done3(swpc3d_ctx *c, int it0, int it1) { note_snapshots(c, it0, it1); }

// The learned-boundary fill's strip cache flags were advanced by steps that
// were captured but will not run.
void batch_abort3(swpc3d_ctx *c) { lab_stale(c); }

// ---------------------------------------------------------------------------
// Groups of 3-D runs in one launch per phase (batches)
// ---------------------------------------------------------------------------
// The workgroups of every member's kernels of one phase, end to end: a
// workgroup finds its member by bisection of the members' first
// workgroups and runs that member's block with the member's own arguments
// (read from a device table), sources, receivers and step counter.  The
// group uses its own launch shapes (interior stress over 256 x 4 tiles,
// z-marching velocity over 256 x 4 tiles and 8-plane chunks, absorber
// rows by 256-cell segments and the x faces by cells); each cell sees the
// operations of the member's own run whatever its tuning, so results equal
// the members' own steps bit for bit.
constexpr int G3_SBY = 4, G3_VBY = 2, G3_VKC = 16;

struct Member3 {
    K3<float> s, v;          // stress- and velocity-phase arguments (group tilings; pml = the x faces)
    Rows3 rows;              // the whole-row absorber boxes
    unsigned nrow, nxc;      // absorber workgroups: row segments, then 256-cell blocks of the x faces
    int nsrc, nt_src;
    const int *src;
    const double *m6, *sdrop;
    int nst, ntdec_w, ntw;
    const int *st;
    double *uacc;
    float *wvel;
    double *wdisp;
    int *dit;
};

__device__ __forceinline__ int member3_of(const unsigned *__restrict__ first, int nm, unsigned b)
{
    int lo = 0, hi = nm - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (first[mid] <= b) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

template <bool VEL>
__global__ void __launch_bounds__(256) k3_pml_group(const Member3 *__restrict__ mem, const unsigned *__restrict__ first,
                                                    int nm)
{
    const int m = member3_of(first, nm, blockIdx.x);
    const Member3 &g = mem[m];
    const unsigned l = blockIdx.x - first[m];
    const K3<float> &a = VEL ? g.v : g.s;
    int i, j, k;
    const bool on = l < g.nrow ? pml_row(g.rows, l, a.nx, i, j, k)
                               : pml_cell(a.pml, (l - g.nrow) * 256u + threadIdx.x, i, j, k);
    if (!on) return;
    if constexpr (VEL) vel_pml_cell<float, false>(a, i, j, k);
    else stress_pml_cell<float, false>(a, i, j, k);
}

__global__ void __launch_bounds__(64 * G3_SBY) k3x_stress_group(const Member3 *__restrict__ mem,
                                                                const unsigned *__restrict__ first, int nm)
{
    const int m = member3_of(first, nm, blockIdx.x);
    stress4_blk<G3_SBY, false, 0>(mem[m].s, blockIdx.x - first[m]);
}

__global__ void __launch_bounds__(64 * G3_VBY) k3x_velz_group(const Member3 *__restrict__ mem,
                                                              const unsigned *__restrict__ first, int nm)
{
    const int m = member3_of(first, nm, blockIdx.x);
    velz_blk<G3_VBY, false, true>(mem[m].v, blockIdx.x - first[m]);
}

// receivers of every member (before the step's updates): one workgroup each
__global__ void __launch_bounds__(64) k3_recv_group(const Member3 *__restrict__ mem)
{
    const Member3 &g = mem[blockIdx.x];
    for (int n = (int)threadIdx.x; n < g.nst; n += 64)
        receiver_one<float>(g.s, n, g.nst, g.st, g.uacc, g.wvel, g.wdisp, g.ntdec_w, g.ntw, g.dit);
}

__global__ void __launch_bounds__(64) k3_glut_group(const Member3 *__restrict__ mem)
{
    const Member3 &g = mem[blockIdx.x];
    if (threadIdx.x == 0 && g.nsrc > 0) stressglut_one<float>(g.s, g.nsrc, g.src, g.m6, g.sdrop, g.nt_src, g.dit);
}

__global__ void k3_group_advance(const Member3 *__restrict__ mem, int n)
{
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q < n) *mem[q].dit = *mem[q].dit + 1;
}

struct Group3 {
    DBuf mem, first_p, first_s, first_v;
    unsigned np = 0, ns = 0, nv = 0;
    int nm = 0;
};

long long group3_cells(const swpc3d_ctx *c) { return (long long)c->d.nx * c->d.ny * c->d.nz; }

bool group3_eligible(const swpc3d_ctx *c)
{
    return c && ready(c) == SWPC_OK && !fp64(c) && !c->s16 && !c->m16 && c->rho.p && c->x4 &&
           (c->x4_form == 1 || c->x4_form == 2) && c->x4_sz == 0 && c->x4_vz > 0 && !c->pml4 && !c->lab && !c->conc &&
           !c->prof.on && c->snaps.empty() && c->nbf == 0;
}

// The member's table entry and its workgroup counts per phase (absorber,
// interior stress, interior velocity).
static void member3_args(swpc3d_ctx *c, Member3 &g, unsigned &np, unsigned &ns, unsigned &nv)
{
    memset(&g, 0, sizeof g);
    const swpc3d_desc &d = c->d;
    K3<float> a = make_args<float>(c);
    a.pml = c->boxes_x;
    const int nk = d.nz - d.na;
    g.s = a;
    (void)x4_tiles(c, g.s, G3_SBY);
    g.s.xcd = 0;
    ns = (g.s.ntx > 0 && g.s.nty > 0 && nk > 0) ? (unsigned)g.s.ntx * (unsigned)g.s.nty * (unsigned)nk : 0u;
    g.v = a;
    (void)x4_tiles(c, g.v, G3_VBY);
    g.v.xcd = 0;
    g.v.kc = G3_VKC;
    g.v.nkc = nk > 0 ? (nk + G3_VKC - 1) / G3_VKC : 0;
    g.v.nblk = (unsigned)std::max(g.v.ntx, 0) * (unsigned)std::max(g.v.nty, 0) * (unsigned)g.v.nkc;
    g.v.nblk_pad = g.v.nblk;
    nv = g.v.nblk;
    g.rows = c->rows;
    g.nrow = c->boxes.total > 0 ? c->rows.total * (unsigned)c->rows.nseg : 0u;
    g.nxc = c->boxes.total > 0 ? (c->boxes_x.total + 255u) / 256u : 0u;
    np = g.nrow + g.nxc;
    g.nsrc = c->nsrc;
    g.nt_src = c->nt_src;
    g.src = (const int *)c->src.p;
    g.m6 = (const double *)c->m6.p;
    g.sdrop = (const double *)c->sdrop.p;
    g.nst = c->nst;
    g.ntdec_w = c->ntdec_w;
    g.ntw = c->ntw;
    g.st = (const int *)c->st.p;
    g.uacc = (double *)c->uacc.p;
    g.wvel = (float *)c->wvel.p;
    g.wdisp = (double *)c->wdisp.p;
    g.dit = (int *)c->dit.p;
}

void group3_free(Group3 *g)
{
    if (!g) return;
    dfree(g->mem);
    dfree(g->first_p);
    dfree(g->first_s);
    dfree(g->first_v);
    delete g;
}

int group3_build(const std::vector<swpc3d_ctx *> &cs, Group3 **out, long long &bytes)
{
    *out = nullptr;
    if (cs.empty()) return SWPC_OK;
    std::vector<Member3> mem(cs.size());
    std::vector<unsigned> fp(cs.size()), fs(cs.size()), fv(cs.size());
    unsigned long long np = 0, ns = 0, nv = 0;
    for (size_t q = 0; q < cs.size(); q++) {
        if (!group3_eligible(cs[q])) return fail(SWPC_EINVAL, "3-D member %zu cannot join a group", q);
        unsigned a, b, v;
        member3_args(cs[q], mem[q], a, b, v);
        fp[q] = (unsigned)np;
        fs[q] = (unsigned)ns;
        fv[q] = (unsigned)nv;
        np += a;
        ns += b;
        nv += v;
    }
    if (np >= (1ULL << 31) || ns >= (1ULL << 31) || nv >= (1ULL << 31))
        return fail(SWPC_EINVAL, "3-D group too large for one launch");
    auto *g = new Group3();
    g->nm = (int)cs.size();
    g->np = (unsigned)np;
    g->ns = (unsigned)ns;
    g->nv = (unsigned)nv;
    int r = dalloc(g->mem, mem.size() * sizeof(Member3), bytes);
    if (!r) r = dalloc(g->first_p, fp.size() * 4, bytes);
    if (!r) r = dalloc(g->first_s, fs.size() * 4, bytes);
    if (!r) r = dalloc(g->first_v, fv.size() * 4, bytes);
    hipError_t e = hipSuccess;
    if (!r) e = hipMemcpy(g->mem.p, mem.data(), mem.size() * sizeof(Member3), hipMemcpyHostToDevice);
    if (!r && e == hipSuccess) e = hipMemcpy(g->first_p.p, fp.data(), fp.size() * 4, hipMemcpyHostToDevice);
    if (!r && e == hipSuccess) e = hipMemcpy(g->first_s.p, fs.data(), fs.size() * 4, hipMemcpyHostToDevice);
    if (!r && e == hipSuccess) e = hipMemcpy(g->first_v.p, fv.data(), fv.size() * 4, hipMemcpyHostToDevice);
    if (r || e != hipSuccess) {
        group3_free(g);
        return r ? r : fail(SWPC_EHIP, "3-D group tables: %s", hipGetErrorString(e));
    }
    *out = g;
    return SWPC_OK;
}

// One step of every member, in the order of step_dev: receivers, absorber
// and interior stress, stress glut, absorber and interior velocity, step
// counters.
int group3_step(Group3 *g, hipStream_t s)
{
    const Member3 *m = (const Member3 *)g->mem.p;
    hipLaunchKernelGGL(k3_recv_group, dim3(g->nm), dim3(64), 0, s, m);
    if (g->np > 0)
        hipLaunchKernelGGL(k3_pml_group<false>, dim3(g->np), dim3(256), 0, s, m, (const unsigned *)g->first_p.p, g->nm);
    if (g->ns > 0)
        hipLaunchKernelGGL(k3x_stress_group, dim3(g->ns), dim3(64, G3_SBY), 0, s, m, (const unsigned *)g->first_s.p,
                           g->nm);
    hipLaunchKernelGGL(k3_glut_group, dim3(g->nm), dim3(64), 0, s, m);
    if (g->np > 0)
        hipLaunchKernelGGL(k3_pml_group<true>, dim3(g->np), dim3(256), 0, s, m, (const unsigned *)g->first_p.p, g->nm);
    if (g->nv > 0)
        hipLaunchKernelGGL(k3x_velz_group, dim3(g->nv), dim3(64, G3_VBY), 0, s, m, (const unsigned *)g->first_v.p,
                           g->nm);
    hipLaunchKernelGGL(k3_group_advance, dim3((g->nm + 63) / 64), dim3(64), 0, s, m, g->nm);
    SWPC_CHECK_LAUNCH();
    return SWPC_OK;
}

// Fortran (k,i,j) memory box -> device layout (row-blocked transposition).
template <typename S, typename D>
static void box_to_dev(const Lay3 &L, const S *src, D *dst)
{
    const int nkb = L.nz + 6, nib = L.nx + 6;
    for (int j = -2; j <= L.ny + 3; j++)
        for (int k0 = -2; k0 <= L.nz + 3; k0 += 16)
            for (int i = -2; i <= L.nx + 3; i++) {
                const S *s = src + (size_t)(k0 + 2) + (size_t)nkb * ((size_t)(i + 2) + (size_t)nib * (size_t)(j + 2));
                int k1 = std::min(k0 + 15, L.nz + 3);
                for (int k = k0; k <= k1; k++) dst[at3(L.px, L.sxy, i, j, k)] = (D)s[k - k0];
            }
}

template <typename S, typename D>
static void dev_to_box(const Lay3 &L, const S *src, D *dst)
{
    const int nkb = L.nz + 6, nib = L.nx + 6;
    for (int j = -2; j <= L.ny + 3; j++)
        for (int k0 = -2; k0 <= L.nz + 3; k0 += 16)
            for (int i = -2; i <= L.nx + 3; i++) {
                D *o = dst + (size_t)(k0 + 2) + (size_t)nkb * ((size_t)(i + 2) + (size_t)nib * (size_t)(j + 2));
                int k1 = std::min(k0 + 15, L.nz + 3);
                for (int k = k0; k <= k1; k++) o[k - k0] = (D)src[at3(L.px, L.sxy, i, j, k)];
            }
}

static DBuf *field_buf(swpc3d_ctx *c, int f)
{
    if (f >= SWPC3D_VX && f <= SWPC3D_VZ) return &c->V[f];
    if (f >= SWPC3D_SXX && f <= SWPC3D_SXY) return &c->Sall;
    if (f >= SWPC3D_R0 && f < SWPC3D_ADE0) return &c->Rall;
    if (f >= SWPC3D_ADE0 && f < SWPC3D_NFIELD) return &c->A[f - SWPC3D_ADE0];
    return nullptr;
}

// A tiled field's region, its component slot and the components per tile;
// null for a field stored as a plain array.
static DBuf *field_tiles(swpc3d_ctx *c, int f, size_t &q, size_t &ns)
{
    if (f >= SWPC3D_SXX && f <= SWPC3D_SXY) {
        q = (size_t)(f - SWPC3D_SXX);
        ns = NS3;
        return &c->Sall;
    }
    if (f >= SWPC3D_R0 && f < SWPC3D_ADE0) {
        q = (size_t)(f - SWPC3D_R0);
        ns = NR3;
        return &c->Rall;
    }
    return nullptr;
}

// Component q of a tiled region (ns components per tile, or the FP16
// memory variables' pair layout r16ix) to and from a plain array of n
// cells, elements of type E (the bits only).
__device__ __forceinline__ long long tile_elem(int ns, int pair16, int q, long long e)
{
    if (pair16) return r16ix(q, (int)e);
    const long long u = e + RB;
    return ((u >> 8) * ns + q) * RTILE + (u & (RTILE - 1));
}

template <typename E>
__global__ void k_tile_gather(const E *__restrict__ t, int ns, int pair16, int q, E *__restrict__ out, long long n)
{
    for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (long long)gridDim.x * blockDim.x)
        out[e] = t[tile_elem(ns, pair16, q, e)];
}

template <typename E>
__global__ void k_tile_scatter(E *__restrict__ t, int ns, int pair16, int q, const E *__restrict__ in, long long n)
{
    for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (long long)gridDim.x * blockDim.x)
        t[tile_elem(ns, pair16, q, e)] = in[e];
}

template <typename E>
static void tile_copy(bool gather, void *tiles, int ns, int pair16, int q, void *plain, long long n, hipStream_t s)
{
    const unsigned nb = (unsigned)std::min<long long>((n + 255) / 256, 65536);
    if (gather)
        hipLaunchKernelGGL(k_tile_gather<E>, dim3(nb), dim3(256), 0, s, (const E *)tiles, ns, pair16, q, (E *)plain, n);
    else
        hipLaunchKernelGGL(k_tile_scatter<E>, dim3(nb), dim3(256), 0, s, (E *)tiles, ns, pair16, q, (const E *)plain, n);
}

// Gather (scatter) a tiled field through a temporary plain device array:
// one kernel and one contiguous copy.  False when the temporary cannot be
// allocated (the caller then copies tile by tile).
static bool tile_copy_dev(swpc3d_ctx *c, bool gather, void *tiles, int ns, int pair16, int q, void *host, size_t el,
                          int &err)
{
    const long long n = c->L.vol;
    void *tmp = nullptr;
    err = SWPC_OK;
    if (hipMalloc(&tmp, (size_t)n * el) != hipSuccess) {
        (void)hipGetLastError();
        return false;
    }
    hipError_t e = hipSuccess;
    if (!gather) e = hipMemcpy(tmp, host, (size_t)n * el, hipMemcpyHostToDevice);
    if (e == hipSuccess) {
        if (el == 8) tile_copy<unsigned long long>(gather, tiles, ns, pair16, q, tmp, n, c->stream);
        else if (el == 4) tile_copy<unsigned>(gather, tiles, ns, pair16, q, tmp, n, c->stream);
        else tile_copy<unsigned short>(gather, tiles, ns, pair16, q, tmp, n, c->stream);
        e = hipGetLastError();
        if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    }
    if (e == hipSuccess && gather) e = hipMemcpy(host, tmp, (size_t)n * el, hipMemcpyDeviceToHost);
    (void)hipFree(tmp);
    if (e != hipSuccess) err = fail(SWPC_EHIP, "tiled field copy failed: %s", hipGetErrorString(e));
    return true;
}

// Field f's vol elements of el bytes between the device and a contiguous
// host array.  A tiled field is gathered from (scattered to) its slots of
// the tiles on the device; if no temporary array fits, one strided copy
// moves the tiles' slots (the cells before the first and after the last of
// the box left out, zero on the device).
static int field_read(swpc3d_ctx *c, int f, void *dst, size_t el)
{
    const size_t n = (size_t)c->L.vol;
    size_t q = 0, ns = 0;
    DBuf *t = field_tiles(c, f, q, ns);
    if (!t) {
        SWPC_HIP(hipMemcpy(dst, field_buf(c, f)->p, n * el, hipMemcpyDeviceToHost));
        return SWPC_OK;
    }
    int err;
    const int pair16 = t == &c->Rall && el == 2;
    if (tile_copy_dev(c, true, t->p, (int)ns, pair16, (int)q, dst, el, err)) return err;
    if (pair16) return fail(SWPC_ENOMEM, "no device memory for a temporary copy of the field");
    const size_t nt = (size_t)c->rtiles;
    std::vector<char> tl(nt * RTILE * el);
    SWPC_HIP(hipMemcpy2D(tl.data(), RTILE * el, (const char *)t->p + q * RTILE * el, ns * RTILE * el, RTILE * el, nt,
                         hipMemcpyDeviceToHost));
    memcpy(dst, tl.data() + (size_t)RB * el, n * el);
    return SWPC_OK;
}

static int field_write(swpc3d_ctx *c, int f, const void *src, size_t el)
{
    const size_t n = (size_t)c->L.vol;
    size_t q = 0, ns = 0;
    DBuf *t = field_tiles(c, f, q, ns);
    if (!t) {
        SWPC_HIP(hipMemcpy(field_buf(c, f)->p, src, n * el, hipMemcpyHostToDevice));
        return SWPC_OK;
    }
    int err;
    const int pair16 = t == &c->Rall && el == 2;
    if (tile_copy_dev(c, false, t->p, (int)ns, pair16, (int)q, const_cast<void *>(src), el, err)) return err;
    if (pair16) return fail(SWPC_ENOMEM, "no device memory for a temporary copy of the field");
    const size_t nt = (size_t)c->rtiles;
    std::vector<char> tl(nt * RTILE * el, 0);
    memcpy(tl.data() + (size_t)RB * el, src, n * el);
    SWPC_HIP(hipMemcpy2D((char *)t->p + q * RTILE * el, ns * RTILE * el, tl.data(), RTILE * el, RTILE * el, nt,
                         hipMemcpyHostToDevice));
    return SWPC_OK;
}

static int upload_profile(swpc3d_ctx *c, DBuf &dst, const float *gc, const float *ge, int n)
{
    // host (4, 1:n) pairs -> device [8][n] in T
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

int swpc3d_create(const swpc3d_desc *desc, swpc3d_ctx **out)
{
    clear_error();
    if (!desc || !out) return fail(SWPC_EINVAL, "null argument");
    *out = nullptr;
    const swpc3d_desc &d = *desc;
    if (d.nx < 8 || d.ny < 8 || d.nz < 8) return fail(SWPC_EINVAL, "grid %dx%dx%d too small", d.nx, d.ny, d.nz);
    if (d.na < 0 || 2 * d.na >= d.nx || 2 * d.na >= d.ny || d.na >= d.nz)
        return fail(SWPC_EINVAL, "absorber thickness na=%d does not fit the grid", d.na);
    if (d.nm != NM3) return fail(SWPC_EINVAL, "nm=%d: the kernels are built for NM = 3", d.nm);
    if (d.precision != SWPC_FP32 && d.precision != SWPC_FP64 && d.precision != SWPC_FP32_S16 &&
        d.precision != SWPC_FP32_R16)
        return fail(SWPC_EINVAL, "precision mode %d not supported by this build", d.precision);
    if (!(d.dx > 0 && d.dy > 0 && d.dz > 0 && d.dt > 0)) return fail(SWPC_EINVAL, "non-positive dx/dy/dz/dt");
    for (int m = 0; m < d.nm; m++)
        if (!(d.ts[m] > 0)) return fail(SWPC_EINVAL, "ts(%d) must be positive", m + 1);
    Lay3 L = make_lay3(d.nx, d.ny, d.nz);
    if (L.vol >= (1LL << 31)) return fail(SWPC_EINVAL, "grid too large for 32-bit element offsets");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return fail(SWPC_EHIP, "no HIP device available");

    std::unique_ptr<swpc3d_ctx> c(new swpc3d_ctx());
    c->d = d;
    c->L = L;
    c->esz = d.precision == SWPC_FP64 ? 8 : 4;
    c->s16 = d.precision == SWPC_FP32_S16 || d.precision == SWPC_FP32_R16;
    // ADE arrays in BF16 with the 16-bit medium only: the whole-row absorber
    // kernels run slower on 2-byte ADE accesses (stress 236 against 205,
    // velocity 191 against 170 us per step at 512 x 512 x 256) than the
    // halved bytes save, so reduced storage keeps them in FP32
    c->a16 = d.precision == SWPC_FP32_S16;
    c->m16 = d.precision == SWPC_FP32_S16;
    // measured defaults at 512 x 512 x 256 on MI355X: one-pass stress over
    // 256 x 2 tiles (the two-launch form for the 16-bit medium), z-marching
    // velocity
    c->x4_form = c->m16 ? 0 : 1;
    c->x4_vz = c->m16 ? 16 : 8;
    c->x4_vby = c->m16 ? 2 : 4;
    SWPC_HIP(hipStreamCreate(&c->stream));
    SWPC_HIP(hipStreamCreate(&c->side));
    SWPC_HIP(hipEventCreateWithFlags(&c->ev_fork, hipEventDisableTiming));
    SWPC_HIP(hipEventCreateWithFlags(&c->ev_join, hipEventDisableTiming));
    const size_t vb = (size_t)L.vol * c->esz;
    int r;
    for (auto &b : c->V) if ((r = dalloc(b, vb, c->bytes))) return r;
    // 16-bit arrays start 64 bytes into a line: cell i = 1 sits at element
    // IO + 1 = 32, so its rows then begin on 128-byte lines as the FP32
    // arrays' do, and a wave's 4-cell accesses cover whole lines
    const size_t vb16 = (size_t)L.vol * 2, lead16 = 128 - (size_t)(IO + 1) * 2;
    c->rtiles = ((long long)L.vol + RB + RTILE - 1) / RTILE;
    if ((r = dalloc(c->Sall, (size_t)c->rtiles * NS3 * RTILE * c->esz, c->bytes))) return r;
    if ((r = dalloc(c->Rall, (size_t)c->rtiles * NR3 * RTILE * (c->s16 ? 2 : c->esz), c->bytes))) return r;
    for (auto &b : c->A) if ((r = dalloc(b, c->a16 ? vb16 : vb, c->bytes, c->a16 ? lead16 : 0))) return r;
    for (auto &b : c->med)
        if ((r = dalloc(b, c->m16 ? vb16 : (size_t)L.vol * 4, c->bytes, c->m16 ? lead16 : 0))) return r;
    if (c->m16)
        for (auto &b : c->mbs) if ((r = dalloc(b, (size_t)L.pz * 8, c->bytes))) return r;
    if ((r = dalloc(c->surf, (size_t)d.nx * d.ny * 16, c->bytes))) return r;
    if ((r = dalloc(c->kob, (size_t)d.nx * d.ny * 4, c->bytes))) return r;
    if ((r = dalloc(c->vmax, 64, c->bytes))) return r;
    if ((r = dalloc(c->dit, 64, c->bytes))) return r;

    // interior tiles: 64 wide, aligned to i = 1 + 64 t (128-byte rows)
    c->ti0 = 1 + (d.na / 64) * 64;

    // absorber boxes (kbeg_a, m_global.f90:336-345)
    Boxes3 &B = c->boxes;
    memset(&B, 0, sizeof B);
    auto add = [&](int i0, int i1, int j0, int j1, int k0, int k1) {
        int ni = i1 - i0 + 1, nj = j1 - j0 + 1, nk = k1 - k0 + 1;
        if (ni <= 0 || nj <= 0 || nk <= 0) return;
        Box3 &x = B.b[B.n++];
        x.i0 = i0; x.j0 = j0; x.k0 = k0; x.ni = ni; x.nj = nj; x.nk = nk;
        x.start = B.total;
        B.total += (unsigned)ni * (unsigned)nj * (unsigned)nk;
    };
    // The y faces and the bottom take whole rows i in [1, nx] (their rows
    // stream as aligned 128-byte lines); the x faces keep only the rows of
    // the interior, na cells each.
    if (d.na > 0) {
        add(1, d.nx, 1, d.na, 1, d.nz);
        add(1, d.nx, d.ny - d.na + 1, d.ny, 1, d.nz);
        add(1, d.nx, d.na + 1, d.ny - d.na, d.nz - d.na + 1, d.nz);
        add(1, d.na, d.na + 1, d.ny - d.na, 1, d.nz - d.na);
        add(d.nx - d.na + 1, d.nx, d.na + 1, d.ny - d.na, 1, d.nz - d.na);
    }
    // whole-row boxes as rows, the others as they are
    Rows3 &RW = c->rows;
    Boxes3 &BX = c->boxes_x;
    memset(&RW, 0, sizeof RW);
    memset(&BX, 0, sizeof BX);
    RW.nseg = (d.nx + 255) / 256;
    for (int q = 0; q < B.n; q++) {
        const Box3 &x = B.b[q];
        if (x.i0 == 1 && x.ni == d.nx) {
            RW.j0[RW.n] = x.j0; RW.nj[RW.n] = x.nj; RW.k0[RW.n] = x.k0;
            RW.start[RW.n] = RW.total;
            RW.total += (unsigned)x.nj * (unsigned)x.nk;
            RW.n++;
        } else {
            BX.b[BX.n] = x;
            BX.b[BX.n].start = BX.total;
            BX.total += (unsigned)x.ni * (unsigned)x.nj * (unsigned)x.nk;
            BX.n++;
        }
    }
    // the same boxes in groups of 4 x-cells for the 4-wide PML kernels
    Boxes4 &B4 = c->boxes4;
    memset(&B4, 0, sizeof B4);
    for (int q = 0; q < B.n; q++) {
        const Box3 &x = B.b[q];
        Box4 &y = B4.b[B4.n++];
        y.i0 = x.i0;
        y.i1 = x.i0 + x.ni - 1;
        y.g0 = (y.i0 - 1) / 4;
        y.ngx = (y.i1 - 1) / 4 - y.g0 + 1;
        y.j0 = x.j0;
        y.nj = x.nj;
        y.k0 = x.k0;
        y.start = B4.total;
        B4.total += (unsigned)y.ngx * (unsigned)x.nj * (unsigned)x.nk;
    }
    *out = c.release();
    return SWPC_OK;
}

int swpc3d_destroy(swpc3d_ctx *c)
{
    if (!c) return SWPC_OK;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    for (auto &b : c->V) dfree(b);
    dfree(c->Sall);
    dfree(c->Rall);
    for (auto &b : c->A) dfree(b);
    for (auto &b : c->med) dfree(b);
    for (auto &b : c->mbs) dfree(b);
    for (auto &b : c->cpl) dfree(b);
    dfree(c->rho);
    dfree(c->lab_w);
    dfree(c->lab_cache);
    dfree(c->lab_colmed);
    DBuf *rest[] = {&c->surf, &c->kob, &c->gx, &c->gy, &c->gz, &c->gxv, &c->src, &c->m6, &c->sdrop, &c->bsrc, &c->f3,
                    &c->fdt, &c->st, &c->uacc, &c->wvel, &c->wdisp, &c->vmax};
    for (auto b : rest) dfree(*b);
    for (auto &sn : c->snaps) {
        dfree(sn.pts);
        dfree(sn.acc);
        dfree(sn.frames);
        dfree(sn.peak);
    }
    c->prof.reset();
    graph_drop(c);
    dfree(c->dit);
    if (c->side) (void)hipStreamSynchronize(c->side);
    if (c->ev_fork) (void)hipEventDestroy(c->ev_fork);
    if (c->ev_join) (void)hipEventDestroy(c->ev_join);
    if (c->side) (void)hipStreamDestroy(c->side);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return SWPC_OK;
}

int swpc3d_set_medium(swpc3d_ctx *c, const float *bx, const float *by, const float *bz, const float *lam,
                      const float *mu, const float *taup, const float *taus, const float *muyz, const float *muxz,
                      const float *muxy)
{
    clear_error();
    if (c) graph_drop(c);  // the captured steps would launch stale arguments
    if (!c) return fail(SWPC_EINVAL, "null context");
    const float *in[10] = {bx, by, bz, lam, mu, taup, taus, muyz, muxz, muxy};
    for (int q = 0; q < 10; q++)
        if (!in[q]) return fail(SWPC_EINVAL, "medium array %d is null", q);
    SWPC_HIP(hipStreamSynchronize(c->stream));
    std::vector<float> h((size_t)c->L.vol, 0.0f);
    for (int q = 0; q < 10; q++) {
        box_to_dev(c->L, in[q], h.data());
        if (!c->m16) {
            SWPC_HIP(hipMemcpy(c->med[q].p, h.data(), h.size() * 4, hipMemcpyHostToDevice));
            // planes whose cells i, j in [1, n] hold one value (bit for bit)
            // are read from a per-plane table instead of the array
            const Lay3 &L = c->L;
            std::vector<float> pc((size_t)L.pz, std::nanf(""));
            bool any = false;
            for (int k = 1; k <= L.nz; k++) {
                const float v0 = h[at3(L.px, L.sxy, 1, 1, k)];
                bool same = v0 == v0;
                for (int j = 1; j <= L.ny && same; j++) {
                    const float *row = h.data() + at3(L.px, L.sxy, 1, j, k);
                    for (int i = 0; i < L.nx && same; i++) same = memcmp(row + i, &v0, 4) == 0;
                }
                if (same) {
                    pc[k + KO] = v0;
                    any = true;
                }
            }
            c->use_cpl[q] = any;
            if (any) {
                if (!c->cpl[q].p)
                    if (int r = dalloc(c->cpl[q], pc.size() * 4, c->bytes)) return r;
                SWPC_HIP(hipMemcpy(c->cpl[q].p, pc.data(), pc.size() * 4, hipMemcpyHostToDevice));
            }
            continue;
        }
        // per-plane base (mean over the cells the kernels read, i, j in
        // [1, n]) and an FP16 residual scaled so that the largest maps to 2^14
        std::vector<__half> hq((size_t)c->L.vol, __float2half(0.0f));
        std::vector<float2> bs((size_t)c->L.pz, make_float2(0.0f, 0.0f));
        const Lay3 &L = c->L;
        for (int k = 1; k <= L.nz; k++) {
            double sum = 0;
            for (int j = 1; j <= L.ny; j++)
                for (int i = 1; i <= L.nx; i++) sum += h[at3(L.px, L.sxy, i, j, k)];
            const float base = (float)(sum / ((double)L.nx * L.ny));
            float mx = 0;
            for (int j = 1; j <= L.ny; j++)
                for (int i = 1; i <= L.nx; i++) mx = std::max(mx, std::fabs(h[at3(L.px, L.sxy, i, j, k)] - base));
            const float scale = mx > 0 ? mx / 16384.0f : 0.0f;
            bs[k + KO] = make_float2(base, scale);
            if (scale > 0)
                for (int j = -2; j <= L.ny + 3; j++)
                    for (int i = -2; i <= L.nx + 3; i++) {
                        const int e = at3(L.px, L.sxy, i, j, k);
                        float v = (h[e] - base) / scale;
                        v = std::min(std::max(v, -65504.0f), 65504.0f);
                        hq[e] = __float2half(v);
                    }
        }
        SWPC_HIP(hipMemcpy(c->med[q].p, hq.data(), hq.size() * 2, hipMemcpyHostToDevice));
        SWPC_HIP(hipMemcpy(c->mbs[q].p, bs.data(), bs.size() * 8, hipMemcpyHostToDevice));
    }
    c->have_medium = true;
    return SWPC_OK;
}

int swpc3d_set_density(swpc3d_ctx *c, const float *rho)
{
    clear_error();
    if (c) graph_drop(c);  // the captured steps would launch stale arguments
    if (!c || !rho) return fail(SWPC_EINVAL, "null argument");
    if (c->m16) return SWPC_OK;  // the 16-bit medium kernels stream their own buoyancy arrays
    SWPC_HIP(hipStreamSynchronize(c->stream));
    std::vector<float> h((size_t)c->L.vol, 0.0f);
    box_to_dev(c->L, rho, h.data());
    if (int r = dalloc(c->rho, h.size() * 4, c->bytes)) return r;
    SWPC_HIP(hipMemcpy(c->rho.p, h.data(), h.size() * 4, hipMemcpyHostToDevice));
    return SWPC_OK;
}

// Scratch device buffers of one call, freed on every return path.
struct DScratch {
    DBuf b[4];
    long long bytes = 0;
    ~DScratch()
    {
        for (auto &x : b) dfree(x);
    }
};

int swpc3d_setup_medium(swpc3d_ctx *c, const float *rho, const float *lam, const float *mu, const float *qp,
                        const float *qs, float zeta3, float cc_re, float cc_im, int *kfs, int *kob, float *vminmax)
{
    clear_error();
    if (c) graph_drop(c);  // the captured steps would launch stale arguments
    if (!c || !rho || !lam || !mu || !qp || !qs) return fail(SWPC_EINVAL, "null argument");
    if (c->m16) return fail(SWPC_EINVAL, "the device medium setup needs an FP32 medium (not SWPC_FP32_S16)");
    const Lay3 &L = c->L;
    if (c->d.na < 0 || 2 * c->d.na >= std::min(L.nx, L.ny) || c->d.na >= L.nz)
        return fail(SWPC_EINVAL, "absorber of %d cells does not fit the grid", c->d.na);
    SWPC_HIP(hipStreamSynchronize(c->stream));
    if (!c->rho.p)
        if (int r = dalloc(c->rho, (size_t)L.vol * 4, c->bytes)) return r;
    MedSetup m;
    m.nx = L.nx; m.ny = L.ny; m.nz = L.nz; m.na = c->d.na; m.px = L.px; m.sxy = L.sxy;
    m.rho = (float *)c->rho.p;
    m.bx = (float *)c->med[MBX].p; m.by = (float *)c->med[MBY].p; m.bz = (float *)c->med[MBZ].p;
    m.lam = (float *)c->med[MLAM].p; m.mu = (float *)c->med[MMU].p;
    m.taup = (float *)c->med[MTAUP].p; m.taus = (float *)c->med[MTAUS].p;
    m.muyz = (float *)c->med[MMUYZ].p; m.muxz = (float *)c->med[MMUXZ].p; m.muxy = (float *)c->med[MMUXY].p;
    m.zeta3 = zeta3; m.cc_re = cc_re; m.cc_im = cc_im;

    DScratch t;
    const size_t nbox = (size_t)(L.nx + 6) * (L.ny + 6) * (L.nz + 6), nmap = (size_t)(L.nx + 6) * (L.ny + 6);
    if (int r = dalloc(t.b[0], nbox * 4, t.bytes)) return r;
    const float *in[5] = {rho, lam, mu, qp, qs};
    float *out[5] = {m.rho, m.lam, m.mu, m.taup, m.taus};
    for (int q = 0; q < 5; q++) {
        SWPC_HIP(hipMemcpy(t.b[0].p, in[q], nbox * 4, hipMemcpyHostToDevice));
        if (int r = med_ingest(m, (const float *)t.b[0].p, out[q], c->stream)) return r;
        SWPC_HIP(hipStreamSynchronize(c->stream));   // the next upload reuses the staging buffer
    }
    dfree(t.b[0]);
    if (int r = med_relax(m, c->stream)) return r;
    if (int r = dalloc(t.b[1], nmap * 4, t.bytes)) return r;
    if (int r = dalloc(t.b[2], nmap * 4, t.bytes)) return r;
    int *dkfs = (int *)t.b[1].p, *dkob = (int *)t.b[2].p;
    if (int r = med_surface(m, dkfs, dkob, (int4 *)c->surf.p, (int *)c->kob.p, c->stream)) return r;
    if (int r = med_average(m, c->stream)) return r;
    float mm[2];
    if (int r = med_minmax(m, dkfs, mm, c->stream)) return r;
    if (vminmax) {
        vminmax[0] = mm[0];
        vminmax[1] = mm[1];
    }
    if (kfs) SWPC_HIP(hipMemcpy(kfs, dkfs, nmap * 4, hipMemcpyDeviceToHost));
    if (kob) SWPC_HIP(hipMemcpy(kob, dkob, nmap * 4, hipMemcpyDeviceToHost));
    // planes on which an array is constant (K3::cpl)
    if (int r = dalloc(t.b[3], (size_t)L.pz * 4, t.bytes)) return r;
    std::vector<int> fl((size_t)L.pz);
    for (int q = 0; q < 10; q++) {
        if (!c->cpl[q].p)
            if (int r = dalloc(c->cpl[q], (size_t)L.pz * 4, c->bytes)) return r;
        if (int r = med_planes(m, (const float *)c->med[q].p, (int *)t.b[3].p, (float *)c->cpl[q].p, L.pz, c->stream))
            return r;
        SWPC_HIP(hipMemcpyAsync(fl.data(), t.b[3].p, fl.size() * 4, hipMemcpyDeviceToHost, c->stream));
        SWPC_HIP(hipStreamSynchronize(c->stream));
        bool any = false;
        for (int v : fl) any |= v != 0;
        c->use_cpl[q] = any;
    }
    c->have_medium = true;
    c->have_surface = true;
    return SWPC_OK;
}

int swpc3d_get_medium(swpc3d_ctx *c, int which, float *out)
{
    clear_error();
    if (!c || !out) return fail(SWPC_EINVAL, "null argument");
    if (which < 0 || which > 10) return fail(SWPC_EINVAL, "medium array %d: expected 0..9 (bx..muxy) or 10 (rho)", which);
    if (c->m16) return fail(SWPC_EINVAL, "the medium is stored in 16 bit (SWPC_FP32_S16)");
    const DBuf &b = which == 10 ? c->rho : c->med[which];
    if (!b.p) return fail(SWPC_ESTATE, "medium array %d is not set", which);
    SWPC_HIP(hipStreamSynchronize(c->stream));
    std::vector<float> h((size_t)c->L.vol);
    SWPC_HIP(hipMemcpy(h.data(), b.p, h.size() * 4, hipMemcpyDeviceToHost));
    dev_to_box(c->L, h.data(), out);
    return SWPC_OK;
}

// FP32 -> BF16, round to nearest even (the rounding the tests apply too)
static uint16_t bf16_rne(float f)
{
    uint32_t u;
    memcpy(&u, &f, 4);
    if ((u & 0x7f800000u) == 0x7f800000u) return (uint16_t)((u >> 16) | ((u & 0xffffu) ? 0x40u : 0u));
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

static int lab_check(const swpc3d_ctx *c, bool need_na0)
{
    if (!c) return fail(SWPC_EINVAL, "null context");
    if (c->m16) return fail(SWPC_EINVAL, "the learned boundary needs an FP32 medium (not SWPC_FP32_S16)");
    if (need_na0 && c->d.na != 0) return fail(SWPC_EINVAL, "the learned boundary replaces the PML: create the context with na = 0");
    if (!c->have_medium || !c->rho.p) return fail(SWPC_ESTATE, "set_medium and set_density must precede the learned boundary");
    return SWPC_OK;
}

int swpc3d_set_lab(swpc3d_ctx *c, const float *params)
{
    clear_error();
    if (c) graph_drop(c);  // the captured steps would launch stale arguments
    if (c) lab_stale(c);
    if (c && !params) {
        c->lab = false;
        return SWPC_OK;
    }
    if (int r = lab_check(c, true)) return r;
    SWPC_HIP(hipStreamSynchronize(c->stream));
    // the nets are rounded to BF16 and laid out as the fill kernel's LDS
    // images (lab_fill_image)
    std::vector<float> w((size_t)LAB_NNET * LAB_NP);
    for (size_t q = 0; q < w.size(); q++) {
        if (!std::isfinite(params[q])) return fail(SWPC_EINVAL, "learned-boundary parameter %zu is not finite", q);
        const uint32_t b = (uint32_t)bf16_rne(params[q]) << 16;
        memcpy(&w[q], &b, 4);
    }
    std::vector<unsigned char> img((size_t)LAB_NNET * LAB_IMG_BYTES);
    lab_fill_image(w.data(), img.data());
    if (!c->lab_w.p)
        if (int r = dalloc(c->lab_w, img.size(), c->bytes)) return r;
    SWPC_HIP(hipMemcpy(c->lab_w.p, img.data(), img.size(), hipMemcpyHostToDevice));
    c->lab_dom = lab_domain(0, 0, c->d.nx, c->d.ny, c->d.nz);
    if (int r = lab_colmed(lab_grid(c), c->lab_dom, c->lab_colmed, c->bytes, c->stream)) return r;
    const size_t cache_bytes = (size_t)LAB_NF * (size_t)c->lab_dom.col0[2] * c->esz;
    if (c->lab_cache.bytes != cache_bytes)
        if (int r = dalloc(c->lab_cache, cache_bytes, c->bytes)) return r;
    SWPC_HIP(hipStreamSynchronize(c->stream));
    c->lab = true;
    return SWPC_OK;
}

int swpc3d_set_lab_sponge(swpc3d_ctx *c, float sponge)
{
    clear_error();
    if (c) graph_drop(c);  // the captured steps would launch stale arguments
    if (!c) return fail(SWPC_EINVAL, "null context");
    if (!(sponge >= 0.0f && sponge < 1.0f)) return fail(SWPC_EINVAL, "sponge must be in [0, 1)");
    c->lab_sponge = sponge;
    return SWPC_OK;
}

int swpc_lab_begin_run(swpc_lab_trainer *tr, swpc3d_ctx *c, int oi, int oj, int nxs, int nys, int nzs, const int *quota)
{
    clear_error();
    if (!tr) return fail(SWPC_EINVAL, "null trainer");
    if (int r = lab_check(c, false)) return r;
    if (oi < 2 || oj < 2 || nxs < LAB_NS || nys < LAB_NS || nzs < LAB_NS || oi + nxs + 2 > c->d.nx ||
        oj + nys + 2 > c->d.ny || nzs + 2 > c->d.nz)
        return fail(SWPC_EINVAL, "gather domain %dx%dx%d at (%d,%d) does not fit inside the grid with its ghosts", nxs,
                    nys, nzs, oi, oj);
    SWPC_HIP(hipStreamSynchronize(c->stream));
    return lab_begin(tr, lab_grid(c), lab_domain(oi, oj, nxs, nys, nzs), quota, c->stream);
}

int swpc3d_lab_gather(swpc3d_ctx *c, swpc_lab_trainer *tr, int phase, float keep, unsigned seed, float thr)
{
    clear_error();
    if (!tr || (phase != 0 && phase != 1)) return fail(SWPC_EINVAL, "bad argument");
    if (int r = lab_check(c, false)) return r;
    return lab_gather(tr, lab_grid(c), phase, keep, seed, thr, c->stream);
}

int swpc3d_set_surface(swpc3d_ctx *c, const int *kfs_top, const int *kfs_bot, const int *kob_top, const int *kob_bot,
                       const int *kob)
{
    clear_error();
    if (c) graph_drop(c);  // the captured steps would launch stale arguments
    if (!c || !kfs_top || !kfs_bot || !kob_top || !kob_bot || !kob) return fail(SWPC_EINVAL, "null argument");
    const int nx = c->d.nx, ny = c->d.ny, nib = nx + 6;
    std::vector<int> s4((size_t)nx * ny * 4), kb((size_t)nx * ny);
    for (int j = 1; j <= ny; j++)
        for (int i = 1; i <= nx; i++) {
            size_t h = (size_t)(i + 2) + (size_t)nib * (size_t)(j + 2), o = (size_t)(j - 1) * nx + (i - 1);
            s4[4 * o + 0] = kfs_top[h];
            s4[4 * o + 1] = kfs_bot[h];
            s4[4 * o + 2] = kob_top[h];
            s4[4 * o + 3] = kob_bot[h];
            kb[o] = kob[h];
            if (kob[h] + 1 < -2 || kob[h] + 1 > c->d.nz + 3)
                return fail(SWPC_EINVAL, "kob(%d,%d)=%d outside the memory box", i, j, kob[h]);
        }
    SWPC_HIP(hipStreamSynchronize(c->stream));
    SWPC_HIP(hipMemcpy(c->surf.p, s4.data(), s4.size() * 4, hipMemcpyHostToDevice));
    SWPC_HIP(hipMemcpy(c->kob.p, kb.data(), kb.size() * 4, hipMemcpyHostToDevice));
    c->have_surface = true;
    return SWPC_OK;
}

int swpc3d_set_pml(swpc3d_ctx *c, const float *gxc, const float *gxe, const float *gyc, const float *gye,
                   const float *gzc, const float *gze)
{
    clear_error();
    if (c) graph_drop(c);  // the captured steps would launch stale arguments
    if (!c || !gxc || !gxe || !gyc || !gye || !gzc || !gze) return fail(SWPC_EINVAL, "null argument");
    SWPC_HIP(hipStreamSynchronize(c->stream));
    int r;
    if ((r = upload_profile(c, c->gx, gxc, gxe, c->d.nx))) return r;
    {
        // padded FP32 copy of the x profiles for the 4-wide PML kernels
        const int n = c->d.nx;
        c->gxp = (n + 3) / 4 * 4 + 4;
        std::vector<float> h((size_t)8 * c->gxp, 0.0f);
        for (int q = 0; q < 4; q++)
            for (int i = 0; i < n; i++) {
                h[(size_t)q * c->gxp + i] = gxc[q + 4 * i];
                h[(size_t)(4 + q) * c->gxp + i] = gxe[q + 4 * i];
            }
        if ((r = dalloc(c->gxv, h.size() * 4, c->bytes))) return r;
        SWPC_HIP(hipMemcpy(c->gxv.p, h.data(), h.size() * 4, hipMemcpyHostToDevice));
    }
    if ((r = upload_profile(c, c->gy, gyc, gye, c->d.ny))) return r;
    if ((r = upload_profile(c, c->gz, gzc, gze, c->d.nz))) return r;
    c->have_pml = true;
    return SWPC_OK;
}

static int check_src(swpc3d_ctx *c, int nsrc, const int *is, const int *js, const int *ks, std::vector<int> &pack)
{
    pack.resize((size_t)3 * nsrc);
    for (int n = 0; n < nsrc; n++) {
        // the stress glut touches i-1, j-1, k-1; the reference accepts
        // sources over [beg-2, end+3] (m_source.f90:233-235)
        if (is[n] < -1 || is[n] > c->d.nx + 3 || js[n] < -1 || js[n] > c->d.ny + 3 || ks[n] < -1 ||
            ks[n] > c->d.nz + 3)
            return fail(SWPC_EINVAL, "source %d at (%d,%d,%d) outside the memory box", n + 1, is[n], js[n], ks[n]);
        pack[3 * n] = is[n];
        pack[3 * n + 1] = js[n];
        pack[3 * n + 2] = ks[n];
    }
    return SWPC_OK;
}

int swpc3d_set_sources(swpc3d_ctx *c, int nsrc, const int *isrc, const int *jsrc, const int *ksrc, const double *m6,
                       int nt, const double *sdrop)
{
    clear_error();
    if (c) graph_drop(c);  // the captured steps would launch stale arguments
    if (!c || nsrc < 0 || nt < 0) return fail(SWPC_EINVAL, "bad argument");
    if (nsrc > 0 && (!isrc || !jsrc || !ksrc || !m6 || !sdrop || nt < 1)) return fail(SWPC_EINVAL, "null source array");
    std::vector<int> pack;
    if (int r = check_src(c, nsrc, isrc, jsrc, ksrc, pack)) return r;
    SWPC_HIP(hipStreamSynchronize(c->stream));
    int r;
    if ((r = dalloc(c->src, pack.size() * 4, c->bytes))) return r;
    if ((r = dalloc(c->m6, (size_t)6 * nsrc * 8, c->bytes))) return r;
    if ((r = dalloc(c->sdrop, (size_t)nt * nsrc * 8, c->bytes))) return r;
    if (nsrc > 0) {
        SWPC_HIP(hipMemcpy(c->src.p, pack.data(), pack.size() * 4, hipMemcpyHostToDevice));
        SWPC_HIP(hipMemcpy(c->m6.p, m6, (size_t)6 * nsrc * 8, hipMemcpyHostToDevice));
        SWPC_HIP(hipMemcpy(c->sdrop.p, sdrop, (size_t)nt * nsrc * 8, hipMemcpyHostToDevice));
    }
    c->nsrc = nsrc;
    c->nt_src = nt;
    if (c->s16 && nsrc > 0) {
        // B = sum_t |sdrop| max|m| is the total stress glut injected; the
        // memory variables (of order tau w S) stay far below it.  B maps to
        // 2^17, so values up to B/2 fit the FP16 range and the far field,
        // many decades below B, stays clear of the FP16 subnormals, which
        // would otherwise cost it its relative precision (a mapping of 32 B
        // to 2^14 put the coda there and gave 8e-3 trace misfit at
        // dx = 125 m).  A power of two, so the scaling itself is exact.
        double bound = 0, mmax = 0;
        for (int it = 0; it < nt; it++) {
            double mx = 0;
            for (int n = 0; n < nsrc; n++) mx = std::max(mx, std::fabs(sdrop[(size_t)it * nsrc + n]));
            bound += mx;
        }
        for (int q = 0; q < 6 * nsrc; q++) mmax = std::max(mmax, std::fabs(m6[q]));
        bound *= mmax;
        if (bound > 0) c->r_scale0 = (float)std::ldexp(1.0, (int)std::ceil(std::log2(bound / 131072.0)));
        c->r_scale = std::ldexp(c->r_scale0, c->r_shift);
    }
    return SWPC_OK;
}

int swpc3d_set_bodyforce(swpc3d_ctx *c, int nsrc, const int *isrc, const int *jsrc, const int *ksrc, const double *f3,
                         int nt, const double *fdt)
{
    clear_error();
    if (c) graph_drop(c);  // the captured steps would launch stale arguments
    if (!c || nsrc < 0 || nt < 0) return fail(SWPC_EINVAL, "bad argument");
    if (nsrc > 0 && (!isrc || !jsrc || !ksrc || !f3 || !fdt || nt < 1)) return fail(SWPC_EINVAL, "null source array");
    std::vector<int> pack;
    if (int r = check_src(c, nsrc, isrc, jsrc, ksrc, pack)) return r;
    SWPC_HIP(hipStreamSynchronize(c->stream));
    int r;
    if ((r = dalloc(c->bsrc, pack.size() * 4, c->bytes))) return r;
    if ((r = dalloc(c->f3, (size_t)3 * nsrc * 8, c->bytes))) return r;
    if ((r = dalloc(c->fdt, (size_t)nt * nsrc * 8, c->bytes))) return r;
    if (nsrc > 0) {
        SWPC_HIP(hipMemcpy(c->bsrc.p, pack.data(), pack.size() * 4, hipMemcpyHostToDevice));
        SWPC_HIP(hipMemcpy(c->f3.p, f3, (size_t)3 * nsrc * 8, hipMemcpyHostToDevice));
        SWPC_HIP(hipMemcpy(c->fdt.p, fdt, (size_t)nt * nsrc * 8, hipMemcpyHostToDevice));
    }
    c->nbf = nsrc;
    c->nt_bf = nt;
    return SWPC_OK;
}

int swpc3d_set_receivers(swpc3d_ctx *c, int nst, const int *ist, const int *jst, const int *kst, int ntdec_w, int ntw)
{
    clear_error();
    if (c) graph_drop(c);  // the captured steps would launch stale arguments
    if (!c || nst < 0 || ntdec_w < 1 || ntw < 0) return fail(SWPC_EINVAL, "bad argument");
    if (nst > 0 && (!ist || !jst || !kst)) return fail(SWPC_EINVAL, "null station array");
    std::vector<int> pack((size_t)3 * nst);
    for (int n = 0; n < nst; n++) {
        if (ist[n] < 1 || ist[n] > c->d.nx || jst[n] < 1 || jst[n] > c->d.ny || kst[n] < 1 || kst[n] > c->d.nz)
            return fail(SWPC_EINVAL, "station %d at (%d,%d,%d) outside the grid", n + 1, ist[n], jst[n], kst[n]);
        pack[3 * n] = ist[n];
        pack[3 * n + 1] = jst[n];
        pack[3 * n + 2] = kst[n];
    }
    SWPC_HIP(hipStreamSynchronize(c->stream));
    int r;
    if ((r = dalloc(c->st, pack.size() * 4, c->bytes))) return r;
    if ((r = dalloc(c->uacc, (size_t)3 * nst * 8, c->bytes))) return r;
    if ((r = dalloc(c->wvel, (size_t)3 * nst * std::max(ntw, 1) * 4, c->bytes))) return r;
    if ((r = dalloc(c->wdisp, (size_t)3 * nst * std::max(ntw, 1) * 8, c->bytes))) return r;
    if (nst > 0) SWPC_HIP(hipMemcpy(c->st.p, pack.data(), pack.size() * 4, hipMemcpyHostToDevice));
    c->nst = nst;
    c->ntdec_w = ntdec_w;
    c->ntw = ntw;
    return SWPC_OK;
}

int swpc3d_update_stress(swpc3d_ctx *c)
{
    clear_error();
    if (int r = ready(c)) return r;
    lab_stale(c);
    const int r = fp64(c) ? launch_stress<double>(c) : launch_stress<float>(c);
    lab_stale(c);
    return r;
}

int swpc3d_update_vel(swpc3d_ctx *c)
{
    clear_error();
    if (int r = ready(c)) return r;
    lab_stale(c);
    const int r = fp64(c) ? launch_vel<double>(c) : launch_vel<float>(c);
    lab_stale(c);
    return r;
}

int swpc3d_stressglut(swpc3d_ctx *c, int it)
{
    clear_error();
    if (int r = ready(c)) return r;
    if (c->nsrc > 0 && (it < 1 || it > c->nt_src)) return fail(SWPC_EINVAL, "stressglut: it=%d outside 1..%d", it, c->nt_src);
    if (int r = set_step(c, it)) return r;
    lab_stale(c);
    const int r = fp64(c) ? launch_glut<double>(c) : launch_glut<float>(c);
    lab_stale(c);
    return r;
}

int swpc3d_bodyforce(swpc3d_ctx *c, int it)
{
    clear_error();
    if (int r = ready(c)) return r;
    if (c->nbf > 0 && (it < 1 || it > c->nt_bf)) return fail(SWPC_EINVAL, "bodyforce: it=%d outside 1..%d", it, c->nt_bf);
    if (int r = set_step(c, it)) return r;
    lab_stale(c);
    const int r = fp64(c) ? launch_bf<double>(c) : launch_bf<float>(c);
    lab_stale(c);
    return r;
}

int swpc3d_store_receivers(swpc3d_ctx *c, int it)
{
    clear_error();
    if (int r = ready(c)) return r;
    if (it < 1) return fail(SWPC_EINVAL, "it must be >= 1");
    if (int r = set_step(c, it)) return r;
    return fp64(c) ? launch_recv<double>(c) : launch_recv<float>(c);
}

int swpc3d_add_snapshot(swpc3d_ctx *c, int kind, int npts, const int *ijk, int ntdec_s, int nbuf, float uc, float m0,
                        float dt, int *id)
{
    clear_error();
    if (!c || !ijk || !id) return fail(SWPC_EINVAL, "null argument");
    if (kind < SWPC_SNAP_PS || kind > SWPC_SNAP_U) return fail(SWPC_EINVAL, "unknown snapshot kind %d", kind);
    if (npts < 1 || ntdec_s < 1 || nbuf < 1) return fail(SWPC_EINVAL, "npts, ntdec_s and nbuf must be >= 1");
    const int nvar = kind == SWPC_SNAP_PS ? 4 : 3;
    if ((long long)nvar * npts * nbuf >= (1LL << 31)) return fail(SWPC_EINVAL, "snapshot store too large");
    std::vector<int> p4((size_t)4 * npts);
    for (int n = 0; n < npts; n++) {
        const int i = ijk[3 * n], j = ijk[3 * n + 1], k = ijk[3 * n + 2];
        // the differences reach one cell either side
        if (i < 1 || i > c->d.nx || j < 1 || j > c->d.ny || k < -1 || k > c->d.nz + 2)
            return fail(SWPC_EINVAL, "snapshot point %d at (%d,%d,%d) outside the grid", n, i, j, k);
        p4[4 * n] = i; p4[4 * n + 1] = j; p4[4 * n + 2] = k; p4[4 * n + 3] = 0;
    }
    SWPC_HIP(hipStreamSynchronize(c->stream));
    swpc3d_ctx::Snap sn;
    sn.kind = kind; sn.npts = npts; sn.ntdec = ntdec_s; sn.nbuf = nbuf; sn.uc = uc; sn.m0 = m0; sn.dt = dt;
    int r;
    if ((r = dalloc(sn.pts, p4.size() * 4, c->bytes))) return r;
    if ((r = dalloc(sn.frames, (size_t)nvar * npts * nbuf * 4, c->bytes))) { dfree(sn.pts); return r; }
    if (kind == SWPC_SNAP_U && (r = dalloc(sn.acc, (size_t)3 * npts * 4, c->bytes))) {
        dfree(sn.pts);
        dfree(sn.frames);
        return r;
    }
    SWPC_HIP(hipMemcpy(sn.pts.p, p4.data(), p4.size() * 4, hipMemcpyHostToDevice));
    c->snaps.push_back(sn);
    *id = (int)c->snaps.size() - 1;
    graph_drop(c);
    return SWPC_OK;
}

int swpc3d_snapshots(swpc3d_ctx *c, int it)
{
    clear_error();
    if (int r = ready(c)) return r;
    if (it < 1) return fail(SWPC_EINVAL, "it must be >= 1");
    if (int r = set_step(c, it)) return r;
    if (int r = fp64(c) ? launch_snap<double>(c) : launch_snap<float>(c)) return r;
    note_snapshots(c, it, it);
    return SWPC_OK;
}

int swpc3d_snapshot_peaks(swpc3d_ctx *c, int id)
{
    clear_error();
    if (!c) return fail(SWPC_EINVAL, "null context");
    if (id < 0 || id >= (int)c->snaps.size()) return fail(SWPC_EINVAL, "unknown snapshot id %d", id);
    auto &sn = c->snaps[id];
    if (sn.kind == SWPC_SNAP_PS)
        return fail(SWPC_EINVAL, "peak amplitudes are kept for velocity and displacement snapshots only");
    if (sn.peak.p) return SWPC_OK;
    SWPC_HIP(hipStreamSynchronize(c->stream));
    if (int r = dalloc(sn.peak, (size_t)3 * sn.npts * 4, c->bytes)) return r;
    graph_drop(c);
    return SWPC_OK;
}

int swpc3d_finish_snapshots(swpc3d_ctx *c, int it)
{
    clear_error();
    if (int r = ready(c)) return r;
    if (it < 1) return fail(SWPC_EINVAL, "it must be >= 1");
    if (int r = set_step(c, it)) return r;
    return fp64(c) ? launch_snap<double>(c, 1) : launch_snap<float>(c, 1);
}

int swpc3d_get_peaks(swpc3d_ctx *c, int id, float *out)
{
    clear_error();
    if (!c || !out) return fail(SWPC_EINVAL, "null argument");
    if (id < 0 || id >= (int)c->snaps.size()) return fail(SWPC_EINVAL, "unknown snapshot id %d", id);
    const auto &sn = c->snaps[id];
    if (!sn.peak.p) return fail(SWPC_ESTATE, "snapshot %d keeps no peak amplitudes (swpc3d_snapshot_peaks)", id);
    SWPC_HIP(hipStreamSynchronize(c->stream));
    SWPC_HIP(hipMemcpy(out, sn.peak.p, (size_t)3 * sn.npts * 4, hipMemcpyDeviceToHost));
    return SWPC_OK;
}

int swpc3d_get_snapshot(swpc3d_ctx *c, int id, int frame, float *out)
{
    clear_error();
    if (!c || !out) return fail(SWPC_EINVAL, "null argument");
    if (id < 0 || id >= (int)c->snaps.size()) return fail(SWPC_EINVAL, "unknown snapshot id %d", id);
    const auto &sn = c->snaps[id];
    if (frame < 0 || frame > sn.last || frame <= sn.last - sn.nbuf)
        return fail(SWPC_EINVAL, "snapshot %d frame %d not held (latest %lld, %d buffered)", id, frame, sn.last,
                    sn.nbuf);
    const int nvar = sn.kind == SWPC_SNAP_PS ? 4 : 3;
    const size_t n = (size_t)nvar * sn.npts;
    SWPC_HIP(hipStreamSynchronize(c->stream));
    SWPC_HIP(hipMemcpy(out, (const float *)sn.frames.p + (size_t)(frame % sn.nbuf) * n, n * 4, hipMemcpyDeviceToHost));
    return SWPC_OK;
}

int swpc3d_step(swpc3d_ctx *c, int it)
{
    clear_error();
    if (int r = ready(c)) return r;
    if (it < 1) return fail(SWPC_EINVAL, "it must be >= 1");
    return run_steps(c, it, 1);
}

int swpc3d_run(swpc3d_ctx *c, int it0, int nsteps)
{
    clear_error();
    if (int r = ready(c)) return r;
    if (it0 < 1 || nsteps < 0) return fail(SWPC_EINVAL, "bad step range");
    return run_steps(c, it0, nsteps);
}

int swpc3d_get_receivers(swpc3d_ctx *c, float *vel, double *disp)
{
    clear_error();
    if (!c) return fail(SWPC_EINVAL, "null context");
    SWPC_HIP(hipStreamSynchronize(c->stream));
    size_t n = (size_t)3 * c->nst * c->ntw;
    if (n == 0) return SWPC_OK;
    if (vel) SWPC_HIP(hipMemcpy(vel, c->wvel.p, n * 4, hipMemcpyDeviceToHost));
    if (disp) SWPC_HIP(hipMemcpy(disp, c->wdisp.p, n * 8, hipMemcpyDeviceToHost));
    return SWPC_OK;
}

int swpc3d_get_field(swpc3d_ctx *c, int f, double *out)
{
    clear_error();
    if (!c || !out) return fail(SWPC_EINVAL, "null argument");
    DBuf *b = field_buf(c, f);
    if (!b) return fail(SWPC_EINVAL, "unknown field %d", f);
    SWPC_HIP(hipStreamSynchronize(c->stream));
    const size_t n = (size_t)c->L.vol;
    if (fp64(c)) {
        std::vector<double> h(n);
        if (int r = field_read(c, f, h.data(), 8)) return r;
        dev_to_box(c->L, h.data(), out);
    } else if (c->s16 && f >= SWPC3D_R0 && (f < SWPC3D_ADE0 || c->a16)) {
        std::vector<uint16_t> raw(n);
        if (int r = field_read(c, f, raw.data(), 2)) return r;
        std::vector<float> h(n);
        for (size_t e = 0; e < n; e++) {
            if (f < SWPC3D_ADE0) {
                __half_raw hr;
                hr.x = raw[e];
                h[e] = __half2float(__half(hr)) * c->r_scale;
            } else {
                uint32_t u = (uint32_t)raw[e] << 16;
                memcpy(&h[e], &u, 4);
            }
        }
        dev_to_box(c->L, h.data(), out);
    } else {
        std::vector<float> h(n);
        if (int r = field_read(c, f, h.data(), 4)) return r;
        dev_to_box(c->L, h.data(), out);
    }
    return SWPC_OK;
}

int swpc3d_set_field(swpc3d_ctx *c, int f, const double *in)
{
    clear_error();
    if (c) graph_drop(c);  // the captured steps would launch stale arguments
    if (c) lab_stale(c);   // the fill's strip cache no longer holds the fields
    if (!c || !in) return fail(SWPC_EINVAL, "null argument");
    DBuf *b = field_buf(c, f);
    if (!b) return fail(SWPC_EINVAL, "unknown field %d", f);
    // an ADE array set from outside may be non-zero where its axis is
    // undamped: the PML kernels then read and write every ADE value
    if (f >= SWPC3D_ADE0) c->ade_sparse = false;
    SWPC_HIP(hipStreamSynchronize(c->stream));
    const size_t n = (size_t)c->L.vol;
    if (fp64(c)) {
        std::vector<double> h(n, 0.0);
        box_to_dev(c->L, in, h.data());
        if (int r = field_write(c, f, h.data(), 8)) return r;
    } else if (c->s16 && f >= SWPC3D_R0 && (f < SWPC3D_ADE0 || c->a16)) {
        std::vector<float> h(n, 0.0f);
        box_to_dev(c->L, in, h.data());
        std::vector<uint16_t> raw(n);
        for (size_t e = 0; e < n; e++) {
            if (f < SWPC3D_ADE0) {
                __half_raw hr = __half_raw(__float2half(std::fmin(std::fmax(h[e] / c->r_scale, -65504.0), 65504.0)));
                raw[e] = hr.x;
            } else {
                // round to nearest even, as __float2bfloat16
                uint32_t u;
                memcpy(&u, &h[e], 4);
                u += 0x7FFFu + ((u >> 16) & 1u);
                raw[e] = (uint16_t)(u >> 16);
            }
        }
        if (int r = field_write(c, f, raw.data(), 2)) return r;
    } else {
        std::vector<float> h(n, 0.0f);
        box_to_dev(c->L, in, h.data());
        if (int r = field_write(c, f, h.data(), 4)) return r;
    }
    return SWPC_OK;
}

int swpc3d_vmax(swpc3d_ctx *c, float *out3)
{
    clear_error();
    if (int r = ready(c)) return r;
    if (!out3) return fail(SWPC_EINVAL, "null argument");
    int r = fp64(c) ? launch_vmax<double>(c) : launch_vmax<float>(c);
    if (r) return r;
    unsigned u[3];
    SWPC_HIP(hipMemcpyAsync(u, c->vmax.p, 12, hipMemcpyDeviceToHost, c->stream));
    SWPC_HIP(hipStreamSynchronize(c->stream));
    for (int q = 0; q < 3; q++) memcpy(&out3[q], &u[q], 4);
    return SWPC_OK;
}

int swpc3d_sync(swpc3d_ctx *c)
{
    clear_error();
    if (!c) return fail(SWPC_EINVAL, "null context");
    SWPC_HIP(hipStreamSynchronize(c->stream));
    return SWPC_OK;
}

int swpc3d_profile(swpc3d_ctx *c, int enable)
{
    clear_error();
    if (!c) return fail(SWPC_EINVAL, "null context");
    SWPC_HIP(hipStreamSynchronize(c->stream));
    c->prof.reset();
    c->prof.on = enable != 0;
    return SWPC_OK;
}

int swpc3d_profile_read(swpc3d_ctx *c, double *ms5, long long *count5)
{
    clear_error();
    if (!c) return fail(SWPC_EINVAL, "null context");
    return prof_read(c->prof, ms5, count5);
}

__global__ void k_rescale_h(__half *p, long long n, float f)
{
    const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q < n) p[q] = __float2half_rn(fminf(fmaxf(__half2float(p[q]) * f, -65504.0f), 65504.0f));
}

// Multiply the stored 16-bit memory variables by f (a power of two).
static int rescale_r(swpc3d_ctx *c, float f)
{
    if (!c->s16 || f == 1.0f) return SWPC_OK;
    const long long n = (long long)(c->Rall.bytes / 2);
    if (n == 0) return SWPC_OK;
    hipLaunchKernelGGL(k_rescale_h, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->stream, (__half *)c->Rall.p, n,
                       f);
    SWPC_CHECK_LAUNCH();
    return SWPC_OK;
}

int swpc3d_set_tuning(swpc3d_ctx *c, const int *p, int n)
{
    clear_error();
    if (c) graph_drop(c);  // the captured steps would launch stale arguments
    if (!c || (n > 0 && !p)) return fail(SWPC_EINVAL, "null argument");
    int v[24] = {c->s_by, c->s_wpe, c->s_split, c->v_by, c->v_wpe, c->xcd, c->s_kc, c->v_kc, c->x4, c->x4_by, c->pml4,
                 c->x4_form, c->x4_vby, c->conc, c->x4_vz, c->x4_sz, c->graph_k, c->sr, c->r_shift, c->pml_rows,
                 c->pml_xcd, c->x4_vlds, c->xs, c->lab_cache_on ? 1 : 0};
    for (int q = 0; q < n && q < 24; q++) v[q] = p[q];
    if (v[18] < -24 || v[18] > 24) return fail(SWPC_EINVAL, "memory-variable scale shift must be in -24..24");
    if (v[16] < 0 || v[16] > 1024) return fail(SWPC_EINVAL, "steps per graph must be in 0..1024");
    if (v[5] < 0 || v[5] > 64) return fail(SWPC_EINVAL, "XCD order must be in 0..64");
    if (v[14] < 0 || v[15] < 0) return fail(SWPC_EINVAL, "negative z-chunk");
    if (v[6] < 0 || v[7] < 0) return fail(SWPC_EINVAL, "negative z-chunk");
    if (v[9] != 1 && v[9] != 2 && v[9] != 4) return fail(SWPC_EINVAL, "4-wide tile rows must be 1, 2 or 4");
    if (v[11] < 0 || v[11] > 2) return fail(SWPC_EINVAL, "4-wide stress form must be 0, 1 or 2");
    if (v[12] != 1 && v[12] != 2 && v[12] != 4 && v[12] != 8)
        return fail(SWPC_EINVAL, "z-marching velocity tile rows must be 1, 2, 4 or 8");
    c->s_by = v[0]; c->s_wpe = v[1]; c->s_split = v[2] != 0; c->v_by = v[3]; c->v_wpe = v[4];
    c->xcd = v[5]; c->s_kc = v[6]; c->v_kc = v[7]; c->x4 = v[8] != 0; c->x4_by = v[9]; c->pml4 = v[10] != 0;
    c->x4_form = v[11]; c->x4_vby = v[12]; c->conc = v[13] != 0; c->x4_vz = v[14]; c->x4_sz = v[15];
    c->graph_k = v[16];
    c->sr = v[17] != 0;
    c->pml_rows = v[19] != 0;
    c->pml_xcd = v[20] != 0;
    c->x4_vlds = v[21] != 0;
    c->xs = v[22] & 3;
    if (c->lab_cache_on != (v[23] != 0)) lab_stale(c);
    c->lab_cache_on = v[23] != 0;
    if (v[18] != c->r_shift) {
        // the stored memory variables are rescaled with the new scale
        if (int r = rescale_r(c, std::ldexp(1.0f, c->r_shift - v[18]))) return r;
        c->r_shift = v[18];
        c->r_scale = std::ldexp(c->r_scale0, c->r_shift);
    }
    return SWPC_OK;
}

int swpc3d_memory(swpc3d_ctx *c, long long *bytes)
{
    if (!c || !bytes) return fail(SWPC_EINVAL, "null argument");
    *bytes = c->bytes;
    return SWPC_OK;
}

}  // extern "C"
