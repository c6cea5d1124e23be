// 2-D SH viscoelastic velocity-stress core (swpc_sh) for MI355X.
//
//   ksh_stress_int  interior Syz, Sxy + memory variables
//                   (src/swpc_sh/m_kernel.f90:141-223); the two-point
//                   harmonic mu_xy, mu_yz are formed in the kernel
//   ksh_vel_int     interior Vy with buoyancy 1/rho (m_kernel.f90:88-139)
//   ksh_stress_pml / ksh_vel_pml   ADE-CFS PML (src/swpc_sh/m_absorb_p.f90:102-289)
//   ksh_stressglut / ksh_bodyforce / ksh_receivers / ksh_vmax
//
// Same layout and launch shapes as the P-SV core: x-fastest (k, i) with
// the memory box [-2, n+3], one thread per cell, the step number in device
// memory so that steps can be captured once and replayed.
#include "swpc_common.h"

#include <algorithm>
#include <memory>

namespace swpc {

// Every SH kernel rounds operation by operation in the reference's
// expression order (no contraction into FMA), so that a run matches the
// reference loops evaluated in IEEE single precision bit for bit.
#pragma clang fp contract(off)

constexpr int NMS = 3;

struct BoxS {
    int i0, k0, ni, nk;
    unsigned start;
};

struct BoxesS {
    BoxS b[3];
    int n;
    unsigned total;
};

template <typename T>
struct KS {
    int px, nx, nz, na;
    T *Vy, *Syz, *Sxy;
    T *R[6];    // Ryz(m) Rxy(m)
    T *A[4];    // axVy azVy axSxy azSyz
    const float *rho, *mu, *taus;
    const int4 *surf;   // per column i in [1, nx]
    const T *gx, *gz;   // [8][n]: c1..c4 at centres, e1..e4 at edges
    T e4[2][2], e2[2][2];
    T r2[2];
    T c1[NMS], c2[NMS], d1[NMS], d2;
    T dt;
    int ti0;
    BoxesS pml;
};

// 2 a b / (a + b + epsilon(1.0)) in IEEE single precision operation by
// operation, as the reference's real(SP) expression (no contraction)
__device__ inline float harm2(float a, float b)
{
#pragma clang fp contract(off)
    const float epsl = 1.1920928955078125e-07f;
    return 2 * a * b / (a + b + epsl);
}

template <typename T>
__device__ inline void order_sh(const KS<T> &a, int i, int k, T &e0x, T &e1x, T &e0z, T &e1z)
{
    const int4 fs = a.surf[i - 1];
    const int o1 = (k - fs.x) * (fs.y - k), o2 = (k - fs.z) * (fs.w - k);
    const bool lo = max(o1, o2) >= 0;
    e0x = lo ? a.e2[0][0] : a.e4[0][0];
    e1x = lo ? a.e2[0][1] : a.e4[0][1];
    e0z = lo ? a.e2[1][0] : a.e4[1][0];
    e1z = lo ? a.e2[1][1] : a.e4[1][1];
}

template <typename T>
__global__ void __launch_bounds__(256) ksh_stress_int(const KS<T> a)
{
    const int i = a.ti0 + (int)(blockIdx.x * 64 + threadIdx.x);
    const int k = 1 + (int)(blockIdx.y * 4 + threadIdx.y);
    if (i <= a.na || i > a.nx - a.na || k > a.nz - a.na) return;
    const int px = a.px;
    const int c = at2(px, i, k);
    T e0x, e1x, e0z, e1z;
    order_sh(a, i, k, e0x, e1x, e0z, e1z);
    const T *__restrict__ Vy = a.Vy;
    const T dt = a.dt;
    const T vy0 = Vy[c];
    // m_kernel.f90:197-220
    const T dxVy = (Vy[c + 1] - vy0) * e0x - (Vy[c + 2] - Vy[c - 1]) * e1x;
    const T dzVy = (Vy[c + px] - vy0) * e0z - (Vy[c + 2 * px] - Vy[c - px]) * e1z;
    const float nnn = a.mu[c], pnn = a.mu[c + px], npn = a.mu[c + 1];
    const T mu_xy = (T)harm2(nnn, npn), mu_yz = (T)harm2(nnn, pnn);
    const T taus1 = (T)a.taus[c];
    const T f_Ryz = mu_yz * taus1 * dzVy;
    const T f_Rxy = mu_xy * taus1 * dxVy;
    T Ryz_n = 0, Rxy_n = 0;
#pragma unroll
    for (int m = 0; m < NMS; m++) {
        const T ryz = a.c1[m] * a.R[m][c] - a.c2[m] * f_Ryz * dt;
        const T rxy = a.c1[m] * a.R[3 + m][c] - a.c2[m] * f_Rxy * dt;
        a.R[m][c] = ryz;
        a.R[3 + m][c] = rxy;
        Ryz_n += a.d1[m] * ryz;
        Rxy_n += a.d1[m] * rxy;
    }
    const T taus_plus1 = 1 + taus1 * (1 + a.d2);
    a.Syz[c] = a.Syz[c] + (mu_yz * taus_plus1 * dzVy + Ryz_n) * dt;
    a.Sxy[c] = a.Sxy[c] + (mu_xy * taus_plus1 * dxVy + Rxy_n) * dt;
}

template <typename T>
__global__ void __launch_bounds__(256) ksh_vel_int(const KS<T> a)
{
    const int i = a.ti0 + (int)(blockIdx.x * 64 + threadIdx.x);
    const int k = 1 + (int)(blockIdx.y * 4 + threadIdx.y);
    if (i <= a.na || i > a.nx - a.na || k > a.nz - a.na) return;
    const int px = a.px;
    const int c = at2(px, i, k);
    T e0x, e1x, e0z, e1z;
    order_sh(a, i, k, e0x, e1x, e0z, e1z);
    // m_kernel.f90:127-133
    const T *__restrict__ Syz = a.Syz;
    const T *__restrict__ Sxy = a.Sxy;
    const T dzSyz = (Syz[c] - Syz[c - px]) * e0z - (Syz[c + px] - Syz[c - 2 * px]) * e1z;
    const T dxSxy = (Sxy[c] - Sxy[c - 1]) * e0x - (Sxy[c + 1] - Sxy[c - 2]) * e1x;
    const float by = 1.0f / a.rho[c];
    a.Vy[c] = a.Vy[c] + (T)by * (dxSxy + dzSyz) * a.dt;
}

__device__ inline bool pml_cell_sh(const BoxesS &B, unsigned id, int &i, int &k)
{
    if (id >= B.total) return false;
    int b = 0;
#pragma unroll
    for (int q = 1; q < 3; q++)
        if (q < B.n && id >= B.b[q].start) b = q;
    const BoxS &x = B.b[b];
    const unsigned l = id - x.start;
    i = x.i0 + (int)(l % (unsigned)x.ni);
    k = x.k0 + (int)(l / (unsigned)x.ni);
    return true;
}

template <typename T>
__global__ void __launch_bounds__(256) ksh_stress_pml(const KS<T> a)
{
    int i, k;
    if (!pml_cell_sh(a.pml, blockIdx.x * 256u + threadIdx.x, i, k)) return;
    const int px = a.px, nx = a.nx, nz = a.nz;
    const int c = at2(px, i, k);
    const T gxe1 = a.gx[4 * nx + i - 1], gxe2 = a.gx[5 * nx + i - 1], gxe3 = a.gx[6 * nx + i - 1], gxe4 = a.gx[7 * nx + i - 1];
    const T gze1 = a.gz[4 * nz + k - 1], gze2 = a.gz[5 * nz + k - 1], gze3 = a.gz[6 * nz + k - 1], gze4 = a.gz[7 * nz + k - 1];
    const T r20x = a.r2[0], r20z = a.r2[1], dt = a.dt;
    // src/swpc_sh/m_absorb_p.f90:262-275: the ADE values are advanced
    // first and the stresses use the advanced values
    const T vy0 = a.Vy[c];
    const T dxVy = (a.Vy[c + 1] - vy0) * r20x;
    const T dzVy = (a.Vy[c + px] - vy0) * r20z;
    const float nnn = a.mu[c], pnn = a.mu[c + px], npn = a.mu[c + 1];
    const T muxy = (T)harm2(nnn, npn), muyz = (T)harm2(nnn, pnn);
    const T axvy = gxe3 * a.A[0][c] + gxe4 * dxVy * dt;
    const T azvy = gze3 * a.A[1][c] + gze4 * dzVy * dt;
    a.A[0][c] = axvy;
    a.A[1][c] = azvy;
    a.Syz[c] = a.Syz[c] + muyz * (gze1 * dzVy + gze2 * azvy) * dt;
    a.Sxy[c] = a.Sxy[c] + muxy * (gxe1 * dxVy + gxe2 * axvy) * dt;
}

template <typename T>
__global__ void __launch_bounds__(256) ksh_vel_pml(const KS<T> a)
{
    int i, k;
    if (!pml_cell_sh(a.pml, blockIdx.x * 256u + threadIdx.x, i, k)) return;
    const int px = a.px, nx = a.nx, nz = a.nz;
    const int c = at2(px, i, k);
    const T gxc1 = a.gx[0 * nx + i - 1], gxc2 = a.gx[1 * nx + i - 1], gxc3 = a.gx[2 * nx + i - 1], gxc4 = a.gx[3 * nx + i - 1];
    // src/swpc_sh/m_absorb_p.f90:174-187: the velocity update takes
    // gzc(1:2, i), the ADE update gzc(3:4, k).  Where nx > nz the index i
    // runs past the profile's bounds (no defined result); there the
    // profile is taken at k.
    const int kz = nx <= nz ? i : k;
    const T gzc1 = a.gz[0 * nz + kz - 1], gzc2 = a.gz[1 * nz + kz - 1];
    const T gzc3 = a.gz[2 * nz + k - 1], gzc4 = a.gz[3 * nz + k - 1];
    const T r20x = a.r2[0], r20z = a.r2[1], dt = a.dt;
    const T dzSyz = (a.Syz[c] - a.Syz[c - px]) * r20z;
    const T dxSxy = (a.Sxy[c] - a.Sxy[c - 1]) * r20x;
    const float by = 1.0f / a.rho[c];
    const T axsxy = a.A[2][c], azsyz = a.A[3][c];
    a.Vy[c] = a.Vy[c] + (T)by * (gxc1 * dxSxy + gzc1 * dzSyz + gxc2 * axsxy + gzc2 * azsyz) * dt;
    a.A[2][c] = gxc3 * axsxy + gxc4 * dxSxy * dt;
    a.A[3][c] = gzc3 * azsyz + gzc4 * dzSyz * dt;
}

// m_source.f90:554-594: Sxy at (k, i) and (k, i-1), Syz at (k, i) and
// (k-1, i), each with half the moment
template <typename T>
__global__ void ksh_stressglut(const KS<T> a, int nsrc, const int *__restrict__ src, const double *__restrict__ m2,
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
        const T myz = (T)m2[2 * n], mxy = (T)m2[2 * n + 1];
        a.Sxy[c] -= mxy * sd / 2;
        a.Sxy[c - 1] -= mxy * sd / 2;
        a.Syz[c] -= myz * sd / 2;
        a.Syz[c - px] -= myz * sd / 2;
    }
}

// m_source.f90:596-629: Vy(k, i) += fy / rho(k, i) * stime * dt/(dx dz)
template <typename T>
__global__ void ksh_bodyforce(const KS<T> a, int nsrc, const int *__restrict__ src, const double *__restrict__ fy,
                              const double *__restrict__ fdt_all, int nt, const int *__restrict__ dit)
{
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    const int it = *dit;
    if (it < 1 || it > nt) return;
    const double *__restrict__ fdt = fdt_all + (size_t)(it - 1) * nsrc;
    for (int n = 0; n < nsrc; n++) {
        const int c = at2(a.px, src[2 * n], src[2 * n + 1]);
        a.Vy[c] += (T)fy[n] / (T)a.rho[c] * (T)fdt[n];
    }
}

// m_wav.f90:139-302 (velocity and displacement): Vy at the station cell;
// the displacement sums Vy dt before the step's updates, in FP64
template <typename T>
__global__ void ksh_receivers(const KS<T> a, int nst, const int *__restrict__ st, double *__restrict__ uacc,
                              float *__restrict__ wvel, double *__restrict__ wdisp, int ntdec_w, int ntw,
                              const int *__restrict__ dit)
{
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= nst) return;
    const int it = *dit;
    const int itw = (it - 1) / ntdec_w;
    const bool store = ((it - 1) % ntdec_w) == 0 && itw < ntw;
    const T vy = a.Vy[at2(a.px, st[2 * n], st[2 * n + 1])];
    const double uy = uacc[n] + (double)vy * (double)a.dt;
    uacc[n] = uy;
    if (store) {
        wvel[(size_t)itw * nst + n] = (float)vy;
        wdisp[(size_t)itw * nst + n] = uy;
    }
}

// SH snapshot products at a list of (i,k) cells (src/swpc_sh/m_snap.f90):
// kind 1 'v' follows wbuf_xz_v (Vy x UC x M0), kind 2 'u' wbuf_xz_u (the
// real(SP) running sum of Vy x UC x M0 x dt, advanced at every call).  One
// variable per frame, frames of npts floats in a ring of nbuf.
template <typename T>
__global__ void ksh_snap(const KS<T> a, int kind, int npts, const int2 *__restrict__ pts, float uc, float m0,
                         float dt, float *__restrict__ acc, float *__restrict__ frames, int ntdec, int nbuf,
                         const int *__restrict__ dit)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= npts) return;
    const int it = *dit;
    const bool store = ((it - 1) % ntdec) == 0;
    if (!store && kind != 2) return;
    float *__restrict__ out = frames + (size_t)(((it - 1) / ntdec) % nbuf) * npts;
    const int2 q = pts[p];
    const T vy = a.Vy[at2(a.px, q.x, q.y)];
    if (kind == 1) {
        out[p] = (float)(vy * (T)uc * (T)m0);
    } else {
        const float uy = (float)((T)acc[p] + vy * (T)uc * (T)m0 * (T)dt);
        acc[p] = uy;
        if (store) out[p] = uy;
    }
}

__global__ void ksh_step_set(int *dit, int it)
{
    if (blockIdx.x == 0 && threadIdx.x == 0) *dit = it;
}

__global__ void ksh_step_advance(int *dit)
{
    if (blockIdx.x == 0 && threadIdx.x == 0) *dit = *dit + 1;
}

__global__ void ksh_fill_u32(unsigned *p, int n, unsigned v)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) p[t] = v;
}

// kernel__vmax (m_kernel.f90:225-240): max |Vy(kob+1, i)| over the interior columns
template <typename T>
__global__ void ksh_vmax(const KS<T> a, const int *__restrict__ kob, unsigned *__restrict__ out)
{
    const int i = a.na + 1 + (int)(blockIdx.x * blockDim.x + threadIdx.x);
    float m = 0;
    if (i <= a.nx - a.na) m = fabsf((float)a.Vy[at2(a.px, i, kob[i - 1] + 1)]);
    for (int off = 32; off > 0; off >>= 1) m = fmaxf(m, __shfl_xor(m, off));
    if ((threadIdx.x & 63) == 0) atomicMax(out, __float_as_uint(m));
}

}  // namespace swpc

struct swpcsh_ctx {
    swpc2d_desc d;
    swpc::Lay2 L;
    size_t esz;
    hipStream_t stream = nullptr;
    long long bytes = 0;
    swpc::DBuf V, S[2], R[6], A[4];
    swpc::DBuf med[3];  // rho mu taus
    swpc::DBuf surf, kob, gx, gz;
    swpc::DBuf src, m2, sdrop, bsrc, fy, fdt;
    swpc::DBuf st, uacc, wvel, wdisp, vmax, dit;
    int nsrc = 0, nt_src = 0, nbf = 0, nt_bf = 0;
    int nst = 0, ntdec_w = 1, ntw = 0;
    bool have_medium = false, have_surface = false, have_pml = false;
    swpc::BoxesS boxes;
    int ti0 = 1, ntx = 0, ntz = 0;
    int graph_k = 16;
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
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
static KS<T> make_args_sh(swpcsh_ctx *c)
{
    KS<T> a;
    memset(&a, 0, sizeof a);
    const swpc2d_desc &d = c->d;
    a.px = c->L.px; a.nx = d.nx; a.nz = d.nz; a.na = d.na;
    a.Vy = (T *)c->V.p; a.Syz = (T *)c->S[0].p; a.Sxy = (T *)c->S[1].p;
    for (int q = 0; q < 6; q++) a.R[q] = (T *)c->R[q].p;
    for (int q = 0; q < 4; q++) a.A[q] = (T *)c->A[q].p;
    a.rho = (const float *)c->med[0].p; a.mu = (const float *)c->med[1].p; a.taus = (const float *)c->med[2].p;
    a.surf = (const int4 *)c->surf.p;
    a.gx = (const T *)c->gx.p; a.gz = (const T *)c->gz.p;
    // kernel__setup (m_kernel.f90:36-71)
    const double h[2] = {d.dx, d.dz};
    for (int q = 0; q < 2; q++) {
        const T hh = (T)h[q];
        const T rc40 = (T)17.0 / (T)16.0 / hh, rc41 = (T)1.0 / (T)48.0 / hh;
        const T rd40 = (T)(-1.0) / (T)16.0 / hh, rd41 = (T)(-1.0) / (T)48.0 / hh;
        a.e4[q][0] = rc40 - rd40; a.e4[q][1] = rc41 - rd41;
        a.e2[q][0] = rc40 + rd40; a.e2[q][1] = rc41 + rd41;
    }
    // r20 is real(SP) in the SH absorber (m_absorb_p.f90:58-59)
    a.r2[0] = (T)(float)(1.0 / d.dx);
    a.r2[1] = (T)(float)(1.0 / d.dz);
    const T t = (T)(float)d.dt;
    T s = 0;
    for (int m = 0; m < NMS; m++) {
        const T tm = (T)d.ts[m];
        a.c1[m] = (2 * tm - t) / (2 * tm + t);
        a.c2[m] = (T)2 / (2 * tm + t) / NMS;
        s += t / (2 * tm - t);
        a.d1[m] = 2 * tm / (2 * tm - t);
    }
    a.d2 = s / NMS;
    a.dt = t;
    a.ti0 = c->ti0;
    a.pml = c->boxes;
    return a;
}

static bool fp64_sh(const swpcsh_ctx *c) { return c->d.precision == SWPC_FP64; }

static int ready_sh(const swpcsh_ctx *c)
{
    if (!c) return fail(SWPC_EINVAL, "null context");
    if (!c->have_medium || !c->have_surface || !c->have_pml)
        return fail(SWPC_ESTATE, "set_medium, set_surface and set_pml must precede time stepping");
    return SWPC_OK;
}

template <typename T>
static int lsh_stress(swpcsh_ctx *c)
{
    const KS<T> a = make_args_sh<T>(c);
    if (c->ntx > 0 && c->ntz > 0) {
        hipLaunchKernelGGL((ksh_stress_int<T>), dim3(c->ntx, c->ntz), dim3(64, 4), 0, c->stream, a);
        SWPC_CHECK_LAUNCH();
    }
    if (c->boxes.total > 0) {
        hipLaunchKernelGGL((ksh_stress_pml<T>), dim3((c->boxes.total + 255) / 256), dim3(256), 0, c->stream, a);
        SWPC_CHECK_LAUNCH();
    }
    return SWPC_OK;
}

template <typename T>
static int lsh_vel(swpcsh_ctx *c)
{
    const KS<T> a = make_args_sh<T>(c);
    if (c->ntx > 0 && c->ntz > 0) {
        hipLaunchKernelGGL((ksh_vel_int<T>), dim3(c->ntx, c->ntz), dim3(64, 4), 0, c->stream, a);
        SWPC_CHECK_LAUNCH();
    }
    if (c->boxes.total > 0) {
        hipLaunchKernelGGL((ksh_vel_pml<T>), dim3((c->boxes.total + 255) / 256), dim3(256), 0, c->stream, a);
        SWPC_CHECK_LAUNCH();
    }
    return SWPC_OK;
}

template <typename T>
static int lsh_glut(swpcsh_ctx *c)
{
    if (c->nsrc == 0) return SWPC_OK;
    const KS<T> a = make_args_sh<T>(c);
    hipLaunchKernelGGL((ksh_stressglut<T>), dim3(1), dim3(64), 0, c->stream, a, c->nsrc, (const int *)c->src.p,
                       (const double *)c->m2.p, (const double *)c->sdrop.p, c->nt_src, (const int *)c->dit.p);
    SWPC_CHECK_LAUNCH();
    return SWPC_OK;
}

template <typename T>
static int lsh_bf(swpcsh_ctx *c)
{
    if (c->nbf == 0) return SWPC_OK;
    const KS<T> a = make_args_sh<T>(c);
    hipLaunchKernelGGL((ksh_bodyforce<T>), dim3(1), dim3(64), 0, c->stream, a, c->nbf, (const int *)c->bsrc.p,
                       (const double *)c->fy.p, (const double *)c->fdt.p, c->nt_bf, (const int *)c->dit.p);
    SWPC_CHECK_LAUNCH();
    return SWPC_OK;
}

template <typename T>
static int lsh_recv(swpcsh_ctx *c)
{
    if (c->nst == 0) return SWPC_OK;
    const KS<T> a = make_args_sh<T>(c);
    hipLaunchKernelGGL((ksh_receivers<T>), dim3((c->nst + 63) / 64), dim3(64), 0, c->stream, a, c->nst,
                       (const int *)c->st.p, (double *)c->uacc.p, (float *)c->wvel.p, (double *)c->wdisp.p,
                       c->ntdec_w, c->ntw, (const int *)c->dit.p);
    SWPC_CHECK_LAUNCH();
    return SWPC_OK;
}

template <typename T>
static int lsh_snap(swpcsh_ctx *c)
{
    if (c->snaps.empty()) return SWPC_OK;
    const KS<T> a = make_args_sh<T>(c);
    for (auto &sn : c->snaps) {
        hipLaunchKernelGGL((ksh_snap<T>), dim3((sn.npts + 255) / 256), dim3(256), 0, c->stream, a, sn.kind,
                           sn.npts, (const int2 *)sn.pts.p, sn.uc, sn.m0, sn.dt, (float *)sn.acc.p,
                           (float *)sn.frames.p, sn.ntdec, sn.nbuf, (const int *)c->dit.p);
        SWPC_CHECK_LAUNCH();
    }
    return SWPC_OK;
}

// the newest frame each stream holds after steps it0..it1
static void note_snapshots_sh(swpcsh_ctx *c, int it0, int it1)
{
    for (auto &sn : c->snaps) {
        const long long last_due = (long long)(it1 - 1) / sn.ntdec * sn.ntdec + 1;
        if (last_due >= it0) sn.last = std::max(sn.last, (last_due - 1) / sn.ntdec);
    }
}

static int set_step_sh(swpcsh_ctx *c, int it)
{
    hipLaunchKernelGGL(ksh_step_set, dim3(1), dim3(64), 0, c->stream, (int *)c->dit.p, it);
    SWPC_CHECK_LAUNCH();
    return SWPC_OK;
}

// SH step order (src/swpc_sh/main.f90:97-114): receivers, snapshots,
// stress, absorber, stress glut, velocity, absorber, body force
template <typename T>
static int step_sh_dev(swpcsh_ctx *c)
{
    int r;
    if ((r = lsh_recv<T>(c))) return r;
    if ((r = lsh_snap<T>(c))) return r;
    if ((r = lsh_stress<T>(c))) return r;
    if ((r = lsh_glut<T>(c))) return r;
    if ((r = lsh_vel<T>(c))) return r;
    if ((r = lsh_bf<T>(c))) return r;
    hipLaunchKernelGGL(ksh_step_advance, dim3(1), dim3(64), 0, c->stream, (int *)c->dit.p);
    SWPC_CHECK_LAUNCH();
    return SWPC_OK;
}

static int step_sh_any(swpcsh_ctx *c) { return fp64_sh(c) ? step_sh_dev<double>(c) : step_sh_dev<float>(c); }

static void graph_drop_sh(swpcsh_ctx *c)
{
    if (c->gexec) (void)hipGraphExecDestroy(c->gexec);
    if (c->graph) (void)hipGraphDestroy(c->graph);
    c->gexec = nullptr;
    c->graph = nullptr;
}

static int graph_ready_sh(swpcsh_ctx *c)
{
    if (c->gexec) return SWPC_OK;
    SWPC_HIP(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    int r = SWPC_OK;
    for (int q = 0; q < c->graph_k && r == SWPC_OK; q++) r = step_sh_any(c);
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

static int check_steps_sh(swpcsh_ctx *c, int it0, int n)
{
    if (c->nsrc > 0 && it0 + n - 1 > c->nt_src)
        return fail(SWPC_EINVAL, "stressglut: steps %d..%d outside 1..%d", it0, it0 + n - 1, c->nt_src);
    if (c->nbf > 0 && it0 + n - 1 > c->nt_bf)
        return fail(SWPC_EINVAL, "bodyforce: steps %d..%d outside 1..%d", it0, it0 + n - 1, c->nt_bf);
    return SWPC_OK;
}

template <typename S, typename D>
static void box_to_dev_sh(const Lay2 &L, const S *src, D *dst)
{
    const int nkb = L.nz + 6;
    for (int k = -2; k <= L.nz + 3; k++)
        for (int i = -2; i <= L.nx + 3; i++) dst[at2(L.px, i, k)] = (D)src[(size_t)(k + 2) + (size_t)nkb * (size_t)(i + 2)];
}

template <typename S, typename D>
static void dev_to_box_sh(const Lay2 &L, const S *src, D *dst)
{
    const int nkb = L.nz + 6;
    for (int k = -2; k <= L.nz + 3; k++)
        for (int i = -2; i <= L.nx + 3; i++) dst[(size_t)(k + 2) + (size_t)nkb * (size_t)(i + 2)] = (D)src[at2(L.px, i, k)];
}

static DBuf *field_buf_sh(swpcsh_ctx *c, int f)
{
    if (f == SWPCSH_VY) return &c->V;
    if (f == SWPCSH_SYZ || f == SWPCSH_SXY) return &c->S[f - SWPCSH_SYZ];
    if (f >= SWPCSH_R0 && f < SWPCSH_ADE0) return &c->R[f - SWPCSH_R0];
    if (f >= SWPCSH_ADE0 && f < SWPCSH_NFIELD) return &c->A[f - SWPCSH_ADE0];
    return nullptr;
}

static int upload_profile_sh(swpcsh_ctx *c, DBuf &dst, const float *gc, const float *ge, int n)
{
    std::vector<double> h((size_t)8 * n);
    for (int q = 0; q < 4; q++)
        for (int i = 0; i < n; i++) {
            h[(size_t)q * n + i] = gc[q + 4 * i];
            h[(size_t)(4 + q) * n + i] = ge[q + 4 * i];
        }
    if (int r = dalloc(dst, (size_t)8 * n * c->esz, c->bytes)) return r;
    if (fp64_sh(c)) {
        SWPC_HIP(hipMemcpy(dst.p, h.data(), h.size() * 8, hipMemcpyHostToDevice));
    } else {
        std::vector<float> f(h.begin(), h.end());
        SWPC_HIP(hipMemcpy(dst.p, f.data(), f.size() * 4, hipMemcpyHostToDevice));
    }
    return SWPC_OK;
}

static int pack_src_sh(swpcsh_ctx *c, int nsrc, const int *is, const int *ks, std::vector<int> &pack)
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

}  // namespace swpc

using namespace swpc;

extern "C" {

int swpcsh_create(const swpc2d_desc *desc, swpcsh_ctx **out)
{
    clear_error();
    if (!desc || !out) return fail(SWPC_EINVAL, "null argument");
    *out = nullptr;
    const swpc2d_desc &d = *desc;
    if (d.nx < 8 || d.nz < 8) return fail(SWPC_EINVAL, "grid %dx%d too small", d.nx, d.nz);
    if (d.na < 0 || 2 * d.na >= d.nx || d.na >= d.nz) return fail(SWPC_EINVAL, "absorber thickness na=%d does not fit", d.na);
    if (d.nm != NMS) return fail(SWPC_EINVAL, "nm=%d: the kernels are built for NM = 3", d.nm);
    if (d.precision != SWPC_FP32 && d.precision != SWPC_FP64)
        return fail(SWPC_EINVAL, "precision mode %d not supported by the SH core", d.precision);
    if (!(d.dx > 0 && d.dz > 0 && d.dt > 0)) return fail(SWPC_EINVAL, "non-positive dx/dz/dt");
    for (int m = 0; m < d.nm; m++)
        if (!(d.ts[m] > 0)) return fail(SWPC_EINVAL, "ts(%d) must be positive", m + 1);
    const Lay2 L = make_lay2(d.nx, d.nz);
    if (L.vol >= (1LL << 31)) return fail(SWPC_EINVAL, "grid too large");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return fail(SWPC_EHIP, "no HIP device available");
    std::unique_ptr<swpcsh_ctx> c(new swpcsh_ctx());
    c->d = d;
    c->L = L;
    c->esz = d.precision == SWPC_FP64 ? 8 : 4;
    SWPC_HIP(hipStreamCreate(&c->stream));
    const size_t vb = (size_t)L.vol * c->esz;
    int r;
    if ((r = dalloc(c->V, vb, c->bytes))) return r;
    for (auto &b : c->S) if ((r = dalloc(b, vb, c->bytes))) return r;
    for (auto &b : c->R) if ((r = dalloc(b, vb, c->bytes))) return r;
    for (auto &b : c->A) if ((r = dalloc(b, vb, c->bytes))) return r;
    for (auto &b : c->med) if ((r = dalloc(b, (size_t)L.vol * 4, c->bytes))) return r;
    if ((r = dalloc(c->surf, (size_t)d.nx * 16, c->bytes))) return r;
    if ((r = dalloc(c->kob, (size_t)d.nx * 4, c->bytes))) return r;
    if ((r = dalloc(c->vmax, 64, c->bytes))) return r;
    if ((r = dalloc(c->dit, 64, c->bytes))) return r;
    // interior: i in [na+1, nx-na], k in [1, nz-na] (m_global.f90:277-292)
    const int ilo = d.na + 1, ihi = d.nx - d.na;
    c->ti0 = 1 + ((ilo - 1) / 64) * 64;
    c->ntx = ihi >= ilo ? (ihi - c->ti0) / 64 + 1 : 0;
    const int nk = d.nz - d.na;
    c->ntz = nk > 0 ? (nk + 3) / 4 : 0;
    BoxesS &B = c->boxes;
    memset(&B, 0, sizeof B);
    auto add = [&](int i0, int i1, int k0, int k1) {
        const int ni = i1 - i0 + 1, nkk = k1 - k0 + 1;
        if (ni <= 0 || nkk <= 0) return;
        BoxS &x = B.b[B.n++];
        x.i0 = i0; x.k0 = k0; x.ni = ni; x.nk = nkk;
        x.start = B.total;
        B.total += (unsigned)ni * (unsigned)nkk;
    };
    if (d.na > 0) {   // kbeg_a: the x-absorber columns from the top, the bottom below the interior
        add(1, d.na, 1, d.nz);
        add(d.nx - d.na + 1, d.nx, 1, d.nz);
        add(d.na + 1, d.nx - d.na, d.nz - d.na + 1, d.nz);
    }
    *out = c.release();
    return SWPC_OK;
}

int swpcsh_destroy(swpcsh_ctx *c)
{
    if (!c) return SWPC_OK;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    graph_drop_sh(c);
    dfree(c->V);
    for (auto &b : c->S) dfree(b);
    for (auto &b : c->R) dfree(b);
    for (auto &b : c->A) dfree(b);
    for (auto &b : c->med) dfree(b);
    DBuf *rest[] = {&c->surf, &c->kob, &c->gx, &c->gz, &c->src, &c->m2, &c->sdrop, &c->bsrc, &c->fy,
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

int swpcsh_set_medium(swpcsh_ctx *c, const float *rho, const float *mu, const float *taus)
{
    clear_error();
    if (c) graph_drop_sh(c);  // the captured steps would launch stale arguments
    if (!c || !rho || !mu || !taus) return fail(SWPC_EINVAL, "null argument");
    const float *in[3] = {rho, mu, taus};
    SWPC_HIP(hipStreamSynchronize(c->stream));
    std::vector<float> h((size_t)c->L.vol, 0.0f);
    for (int q = 0; q < 3; q++) {
        box_to_dev_sh(c->L, in[q], h.data());
        SWPC_HIP(hipMemcpy(c->med[q].p, h.data(), h.size() * 4, hipMemcpyHostToDevice));
    }
    c->have_medium = true;
    return SWPC_OK;
}

int swpcsh_set_surface(swpcsh_ctx *c, const int *kfs_top, const int *kfs_bot, const int *kob_top, const int *kob_bot,
                       const int *kob)
{
    clear_error();
    if (c) graph_drop_sh(c);
    if (!c || !kfs_top || !kfs_bot || !kob_top || !kob_bot || !kob) return fail(SWPC_EINVAL, "null argument");
    const int nx = c->d.nx;
    std::vector<int> s4((size_t)nx * 4), kb(nx);
    for (int i = 1; i <= nx; i++) {
        const size_t h = (size_t)(i + 2);
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

int swpcsh_set_pml(swpcsh_ctx *c, const float *gxc, const float *gxe, const float *gzc, const float *gze)
{
    clear_error();
    if (c) graph_drop_sh(c);
    if (!c || !gxc || !gxe || !gzc || !gze) return fail(SWPC_EINVAL, "null argument");
    SWPC_HIP(hipStreamSynchronize(c->stream));
    int r;
    if ((r = upload_profile_sh(c, c->gx, gxc, gxe, c->d.nx))) return r;
    if ((r = upload_profile_sh(c, c->gz, gzc, gze, c->d.nz))) return r;
    c->have_pml = true;
    return SWPC_OK;
}

int swpcsh_set_sources(swpcsh_ctx *c, int nsrc, const int *isrc, const int *ksrc, const double *m2, int nt,
                       const double *sdrop)
{
    clear_error();
    if (c) graph_drop_sh(c);
    if (!c || nsrc < 0 || nt < 0) return fail(SWPC_EINVAL, "bad argument");
    if (nsrc > 0 && (!isrc || !ksrc || !m2 || !sdrop || nt < 1)) return fail(SWPC_EINVAL, "null source array");
    std::vector<int> pack;
    if (int r = pack_src_sh(c, nsrc, isrc, ksrc, pack)) return r;
    SWPC_HIP(hipStreamSynchronize(c->stream));
    int r;
    if ((r = dalloc(c->src, pack.size() * 4, c->bytes))) return r;
    if ((r = dalloc(c->m2, (size_t)2 * nsrc * 8, c->bytes))) return r;
    if ((r = dalloc(c->sdrop, (size_t)nt * nsrc * 8, c->bytes))) return r;
    if (nsrc > 0) {
        SWPC_HIP(hipMemcpy(c->src.p, pack.data(), pack.size() * 4, hipMemcpyHostToDevice));
        SWPC_HIP(hipMemcpy(c->m2.p, m2, (size_t)2 * nsrc * 8, hipMemcpyHostToDevice));
        SWPC_HIP(hipMemcpy(c->sdrop.p, sdrop, (size_t)nt * nsrc * 8, hipMemcpyHostToDevice));
    }
    c->nsrc = nsrc;
    c->nt_src = nt;
    return SWPC_OK;
}

int swpcsh_set_bodyforce(swpcsh_ctx *c, int nsrc, const int *isrc, const int *ksrc, const double *fy, int nt,
                         const double *fdt)
{
    clear_error();
    if (c) graph_drop_sh(c);
    if (!c || nsrc < 0 || nt < 0) return fail(SWPC_EINVAL, "bad argument");
    if (nsrc > 0 && (!isrc || !ksrc || !fy || !fdt || nt < 1)) return fail(SWPC_EINVAL, "null source array");
    std::vector<int> pack;
    if (int r = pack_src_sh(c, nsrc, isrc, ksrc, pack)) return r;
    SWPC_HIP(hipStreamSynchronize(c->stream));
    int r;
    if ((r = dalloc(c->bsrc, pack.size() * 4, c->bytes))) return r;
    if ((r = dalloc(c->fy, (size_t)nsrc * 8, c->bytes))) return r;
    if ((r = dalloc(c->fdt, (size_t)nt * nsrc * 8, c->bytes))) return r;
    if (nsrc > 0) {
        SWPC_HIP(hipMemcpy(c->bsrc.p, pack.data(), pack.size() * 4, hipMemcpyHostToDevice));
        SWPC_HIP(hipMemcpy(c->fy.p, fy, (size_t)nsrc * 8, hipMemcpyHostToDevice));
        SWPC_HIP(hipMemcpy(c->fdt.p, fdt, (size_t)nt * nsrc * 8, hipMemcpyHostToDevice));
    }
    c->nbf = nsrc;
    c->nt_bf = nt;
    return SWPC_OK;
}

int swpcsh_set_receivers(swpcsh_ctx *c, int nst, const int *ist, const int *kst, int ntdec_w, int ntw)
{
    clear_error();
    if (c) graph_drop_sh(c);
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
    if ((r = dalloc(c->uacc, (size_t)std::max(nst, 1) * 8, c->bytes))) return r;
    if ((r = dalloc(c->wvel, (size_t)std::max(nst, 1) * std::max(ntw, 1) * 4, c->bytes))) return r;
    if ((r = dalloc(c->wdisp, (size_t)std::max(nst, 1) * std::max(ntw, 1) * 8, c->bytes))) return r;
    if (nst > 0) SWPC_HIP(hipMemcpy(c->st.p, pack.data(), pack.size() * 4, hipMemcpyHostToDevice));
    c->nst = nst;
    c->ntdec_w = ntdec_w;
    c->ntw = ntw;
    return SWPC_OK;
}

int swpcsh_update_stress(swpcsh_ctx *c)
{
    clear_error();
    if (int r = ready_sh(c)) return r;
    return fp64_sh(c) ? lsh_stress<double>(c) : lsh_stress<float>(c);
}

int swpcsh_update_vel(swpcsh_ctx *c)
{
    clear_error();
    if (int r = ready_sh(c)) return r;
    return fp64_sh(c) ? lsh_vel<double>(c) : lsh_vel<float>(c);
}

int swpcsh_stressglut(swpcsh_ctx *c, int it)
{
    clear_error();
    if (int r = ready_sh(c)) return r;
    if (c->nsrc > 0 && (it < 1 || it > c->nt_src)) return fail(SWPC_EINVAL, "stressglut: it=%d outside 1..%d", it, c->nt_src);
    if (int r = set_step_sh(c, it)) return r;
    return fp64_sh(c) ? lsh_glut<double>(c) : lsh_glut<float>(c);
}

int swpcsh_bodyforce(swpcsh_ctx *c, int it)
{
    clear_error();
    if (int r = ready_sh(c)) return r;
    if (c->nbf > 0 && (it < 1 || it > c->nt_bf)) return fail(SWPC_EINVAL, "bodyforce: it=%d outside 1..%d", it, c->nt_bf);
    if (int r = set_step_sh(c, it)) return r;
    return fp64_sh(c) ? lsh_bf<double>(c) : lsh_bf<float>(c);
}

int swpcsh_store_receivers(swpcsh_ctx *c, int it)
{
    clear_error();
    if (int r = ready_sh(c)) return r;
    if (it < 1) return fail(SWPC_EINVAL, "it must be >= 1");
    if (int r = set_step_sh(c, it)) return r;
    return fp64_sh(c) ? lsh_recv<double>(c) : lsh_recv<float>(c);
}

int swpcsh_step(swpcsh_ctx *c, int it)
{
    clear_error();
    if (int r = ready_sh(c)) return r;
    if (it < 1) return fail(SWPC_EINVAL, "it must be >= 1");
    if (int r = check_steps_sh(c, it, 1)) return r;
    if (int r = set_step_sh(c, it)) return r;
    if (int r = step_sh_any(c)) return r;
    note_snapshots_sh(c, it, it);
    return SWPC_OK;
}

int swpcsh_set_graph(swpcsh_ctx *c, int steps)
{
    clear_error();
    if (!c || steps < 0 || steps > 1024) return fail(SWPC_EINVAL, "steps per graph must be in 0..1024");
    graph_drop_sh(c);
    c->graph_k = steps;
    return SWPC_OK;
}

int swpcsh_run(swpcsh_ctx *c, int it0, int nsteps)
{
    clear_error();
    if (int r = ready_sh(c)) return r;
    if (it0 < 1 || nsteps < 0) return fail(SWPC_EINVAL, "bad step range");
    if (nsteps == 0) return SWPC_OK;
    if (int r = check_steps_sh(c, it0, nsteps)) return r;
    if (int r = set_step_sh(c, it0)) return r;
    int done = 0;
    if (c->graph_k > 0 && nsteps >= c->graph_k) {
        if (int r = graph_ready_sh(c)) return r;
        for (; done + c->graph_k <= nsteps; done += c->graph_k) SWPC_HIP(hipGraphLaunch(c->gexec, c->stream));
    }
    for (; done < nsteps; done++)
        if (int r = step_sh_any(c)) return r;
    note_snapshots_sh(c, it0, it0 + nsteps - 1);
    return SWPC_OK;
}

int swpcsh_add_snapshot(swpcsh_ctx *c, int kind, int npts, const int *ik, int ntdec_s, int nbuf, float uc, float m0,
                        float dt, int *id)
{
    clear_error();
    if (!c || !ik || !id) return fail(SWPC_EINVAL, "null argument");
    if (kind != SWPC_SNAP_V && kind != SWPC_SNAP_U)
        return fail(SWPC_EINVAL, "SH snapshot kind %d: only 'v' (%d) and 'u' (%d) exist", kind, SWPC_SNAP_V,
                    SWPC_SNAP_U);
    if (npts < 1 || ntdec_s < 1 || nbuf < 1) return fail(SWPC_EINVAL, "npts, ntdec_s and nbuf must be >= 1");
    if ((long long)npts * nbuf >= (1LL << 31)) return fail(SWPC_EINVAL, "snapshot store too large");
    std::vector<int> p2((size_t)2 * npts);
    for (int n = 0; n < npts; n++) {
        const int i = ik[2 * n], k = ik[2 * n + 1];
        if (i < -1 || i > c->d.nx + 2 || k < -1 || k > c->d.nz + 2)
            return fail(SWPC_EINVAL, "snapshot point %d at (%d,%d) outside the grid", n, i, k);
        p2[2 * n] = i;
        p2[2 * n + 1] = k;
    }
    SWPC_HIP(hipStreamSynchronize(c->stream));
    swpcsh_ctx::Snap sn;
    sn.kind = kind; sn.npts = npts; sn.ntdec = ntdec_s; sn.nbuf = nbuf; sn.uc = uc; sn.m0 = m0; sn.dt = dt;
    int r;
    if ((r = dalloc(sn.pts, p2.size() * 4, c->bytes))) return r;
    if ((r = dalloc(sn.frames, (size_t)npts * nbuf * 4, c->bytes))) {
        dfree(sn.pts);
        return r;
    }
    if (kind == SWPC_SNAP_U && (r = dalloc(sn.acc, (size_t)npts * 4, c->bytes))) {
        dfree(sn.pts);
        dfree(sn.frames);
        return r;
    }
    SWPC_HIP(hipMemcpy(sn.pts.p, p2.data(), p2.size() * 4, hipMemcpyHostToDevice));
    c->snaps.push_back(sn);
    *id = (int)c->snaps.size() - 1;
    graph_drop_sh(c);
    return SWPC_OK;
}

int swpcsh_snapshots(swpcsh_ctx *c, int it)
{
    clear_error();
    if (int r = ready_sh(c)) return r;
    if (it < 1) return fail(SWPC_EINVAL, "it must be >= 1");
    if (int r = set_step_sh(c, it)) return r;
    if (int r = fp64_sh(c) ? lsh_snap<double>(c) : lsh_snap<float>(c)) return r;
    note_snapshots_sh(c, it, it);
    return SWPC_OK;
}

int swpcsh_get_snapshot(swpcsh_ctx *c, int id, int frame, float *out)
{
    clear_error();
    if (!c || !out) return fail(SWPC_EINVAL, "null argument");
    if (id < 0 || id >= (int)c->snaps.size()) return fail(SWPC_EINVAL, "unknown snapshot id %d", id);
    const auto &sn = c->snaps[id];
    if (frame < 0 || frame > sn.last || frame <= sn.last - sn.nbuf)
        return fail(SWPC_EINVAL, "snapshot %d frame %d not held (latest %lld, %d buffered)", id, frame, sn.last,
                    sn.nbuf);
    SWPC_HIP(hipStreamSynchronize(c->stream));
    SWPC_HIP(hipMemcpy(out, (const float *)sn.frames.p + (size_t)(frame % sn.nbuf) * sn.npts, (size_t)sn.npts * 4,
                       hipMemcpyDeviceToHost));
    return SWPC_OK;
}

int swpcsh_get_receivers(swpcsh_ctx *c, float *vel, double *disp)
{
    clear_error();
    if (!c) return fail(SWPC_EINVAL, "null context");
    SWPC_HIP(hipStreamSynchronize(c->stream));
    const size_t n = (size_t)c->nst * c->ntw;
    if (n == 0) return SWPC_OK;
    if (vel) SWPC_HIP(hipMemcpy(vel, c->wvel.p, n * 4, hipMemcpyDeviceToHost));
    if (disp) SWPC_HIP(hipMemcpy(disp, c->wdisp.p, n * 8, hipMemcpyDeviceToHost));
    return SWPC_OK;
}

int swpcsh_get_field(swpcsh_ctx *c, int f, double *out)
{
    clear_error();
    if (!c || !out) return fail(SWPC_EINVAL, "null argument");
    DBuf *b = field_buf_sh(c, f);
    if (!b) return fail(SWPC_EINVAL, "unknown SH field %d", f);
    SWPC_HIP(hipStreamSynchronize(c->stream));
    if (fp64_sh(c)) {
        std::vector<double> h((size_t)c->L.vol);
        SWPC_HIP(hipMemcpy(h.data(), b->p, h.size() * 8, hipMemcpyDeviceToHost));
        dev_to_box_sh(c->L, h.data(), out);
    } else {
        std::vector<float> h((size_t)c->L.vol);
        SWPC_HIP(hipMemcpy(h.data(), b->p, h.size() * 4, hipMemcpyDeviceToHost));
        dev_to_box_sh(c->L, h.data(), out);
    }
    return SWPC_OK;
}

int swpcsh_set_field(swpcsh_ctx *c, int f, const double *in)
{
    clear_error();
    if (!c || !in) return fail(SWPC_EINVAL, "null argument");
    DBuf *b = field_buf_sh(c, f);
    if (!b) return fail(SWPC_EINVAL, "unknown SH field %d", f);
    SWPC_HIP(hipStreamSynchronize(c->stream));
    if (fp64_sh(c)) {
        std::vector<double> h((size_t)c->L.vol, 0.0);
        box_to_dev_sh(c->L, in, h.data());
        SWPC_HIP(hipMemcpy(b->p, h.data(), h.size() * 8, hipMemcpyHostToDevice));
    } else {
        std::vector<float> h((size_t)c->L.vol, 0.0f);
        box_to_dev_sh(c->L, in, h.data());
        SWPC_HIP(hipMemcpy(b->p, h.data(), h.size() * 4, hipMemcpyHostToDevice));
    }
    return SWPC_OK;
}

int swpcsh_vmax(swpcsh_ctx *c, float *out)
{
    clear_error();
    if (int r = ready_sh(c)) return r;
    if (!out) return fail(SWPC_EINVAL, "null argument");
    const int n = c->d.nx - 2 * c->d.na;
    hipLaunchKernelGGL(ksh_fill_u32, dim3(1), dim3(64), 0, c->stream, (unsigned *)c->vmax.p, 1, 0u);
    if (n > 0) {
        if (fp64_sh(c)) {
            const KS<double> a = make_args_sh<double>(c);
            hipLaunchKernelGGL((ksh_vmax<double>), dim3((n + 255) / 256), dim3(256), 0, c->stream, a,
                               (const int *)c->kob.p, (unsigned *)c->vmax.p);
        } else {
            const KS<float> a = make_args_sh<float>(c);
            hipLaunchKernelGGL((ksh_vmax<float>), dim3((n + 255) / 256), dim3(256), 0, c->stream, a,
                               (const int *)c->kob.p, (unsigned *)c->vmax.p);
        }
        SWPC_CHECK_LAUNCH();
    }
    unsigned u = 0;
    SWPC_HIP(hipMemcpyAsync(&u, c->vmax.p, 4, hipMemcpyDeviceToHost, c->stream));
    SWPC_HIP(hipStreamSynchronize(c->stream));
    memcpy(out, &u, 4);
    return SWPC_OK;
}

int swpcsh_sync(swpcsh_ctx *c)
{
    clear_error();
    if (!c) return fail(SWPC_EINVAL, "null context");
    SWPC_HIP(hipStreamSynchronize(c->stream));
    return SWPC_OK;
}

}  // extern "C"
