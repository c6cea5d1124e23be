// Learned absorbing boundary for MI355X: the ghost-fill networks of the
// per-face MLP boundary (inference inside the time step), the gather of
// training samples from wide-PML runs, and the training kernels (forward,
// backward, Adam) that fit the networks on the GPU.
//
// The definition of the boundary (features, network, ghost cells) is
// stated in swpc_lab.h.  Everything here is FP32 arithmetic; the nets used
// in the time step are stored in BF16.
//
// Kernels
//   k_lab_colmed  per-column medium features (vs/vp, vp*dt/h, rho*vp)
//   k_lab_fill_f32 / k_lab_fill_f64   ghost fill of one phase: one lane
//                 per column; the hidden layers on the matrix cores
//                 (32x32x16 BF16 MFMA, FP32 accumulation) over each
//                 wave's 64 columns, the output layer in FP32
//   k_lab_gather  training samples of one phase from a wide run's fields
//   k_lab_fwdbwd  per-workgroup forward + backward of 128 samples; the
//                 weight gradients are formed as LDS-resident outer-product
//                 sums and written as one partial per workgroup
//   k_lab_adam    deterministic reduction of the partials + Adam update
// The training kernels stay on the vector units: they are bound by launch
// latency, not arithmetic.
#include "swpc_lab.h"

#include <mutex>

#include <hip/hip_bf16.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>

namespace swpc {

// ghost targets per net (2*face + phase): field id (0..8 = Vx Vy Vz Sxx Syy
// Szz Syz Sxz Sxy) and ghost depth (1 = the cell next to the face).  These
// are exactly the ghost cells that the 4th-order stress (phase 0) and
// velocity (phase 1) stencils read beyond each face.
__constant__ int c_tgt[LAB_NNET][LAB_NOUT][2] = {
    {{0, 2}, {0, 1}, {1, 1}, {2, 1}, {0, 0}},   // x- V
    {{3, 1}, {8, 2}, {8, 1}, {7, 2}, {7, 1}},   // x- S
    {{0, 1}, {1, 1}, {1, 2}, {2, 1}, {2, 2}},   // x+ V
    {{3, 1}, {3, 2}, {8, 1}, {7, 1}, {0, 0}},   // x+ S
    {{1, 2}, {1, 1}, {0, 1}, {2, 1}, {0, 0}},   // y- V
    {{4, 1}, {8, 2}, {8, 1}, {6, 2}, {6, 1}},   // y- S
    {{1, 1}, {0, 1}, {0, 2}, {2, 1}, {2, 2}},   // y+ V
    {{4, 1}, {4, 2}, {8, 1}, {6, 1}, {0, 0}},   // y+ S
    {{2, 1}, {0, 1}, {0, 2}, {1, 1}, {1, 2}},   // z+ V
    {{5, 1}, {5, 2}, {7, 1}, {6, 1}, {0, 0}},   // z+ S
};
static const int h_nout[LAB_NNET] = {4, 5, 5, 4, 4, 5, 5, 4, 5, 4};
__constant__ int c_nout[LAB_NNET] = {4, 5, 5, 4, 4, 5, 5, 4, 5, 4};

LabDomain lab_domain(int oi, int oj, int nx, int ny, int nz)
{
    LabDomain d;
    d.oi = oi; d.oj = oj; d.nx = nx; d.ny = ny; d.nz = nz;
    d.ncol[0] = d.ncol[1] = ny * nz;
    d.ncol[2] = d.ncol[3] = nx * nz;
    d.ncol[4] = nx * ny;
    d.col0[0] = 0;
    for (int f = 0; f < 5; f++) d.col0[f + 1] = d.col0[f] + d.ncol[f];
    return d;
}

struct LabDev {
    int px, sxy;
    void *F[9];             // FP32 or FP64 fields (f64)
    int f64;
    const float *rho, *lam, *mu;
    float dt, h[3];
    int oi, oj, nx, ny, nz;
    int ncol[5];
    long long col0[6];
};

static LabDev make_dev(const LabGrid &g, const LabDomain &d)
{
    LabDev a;
    a.px = g.px; a.sxy = g.sxy;
    for (int q = 0; q < 9; q++) a.F[q] = g.F[q];
    a.f64 = g.f64;
    a.rho = g.rho; a.lam = g.lam; a.mu = g.mu;
    a.dt = g.dt;
    for (int q = 0; q < 3; q++) a.h[q] = g.h[q];
    a.oi = d.oi; a.oj = d.oj; a.nx = d.nx; a.ny = d.ny; a.nz = d.nz;
    for (int f = 0; f < 5; f++) a.ncol[f] = d.ncol[f];
    for (int f = 0; f < 6; f++) a.col0[f] = d.col0[f];
    return a;
}

// Element e of field fi in its array: the stresses (fi >= 3) are stored
// in cell tiles of NS3 components (tix), the velocities as plain arrays.
__device__ __forceinline__ long long lab_ix(int fi, int e) { return fi >= 3 ? tix<NS3>(e) : (long long)e; }

// Field value at an element, FP32 or FP64 state (T).
template <typename T>
__device__ __forceinline__ float lab_ld(const LabDev &a, int fi, int e)
{
    return (float)((const T *)a.F[fi])[lab_ix(fi, e)];
}

template <typename T>
__device__ __forceinline__ void lab_st(const LabDev &a, int fi, int e, float v)
{
    ((T *)a.F[fi])[lab_ix(fi, e)] = (T)v;
}

// Address of strip cell d = 1 of column col of face f, and the address
// step from cell d to d+1 (inward); ghost depth h sits at d = 1 - h.
__device__ __forceinline__ void lab_col(const LabDev &a, int f, int col, int &base, int &step)
{
    int i, j, k;
    if (f < 2) {
        j = col % a.ny + 1; k = col / a.ny + 1;
        i = f == 0 ? 1 : a.nx;
        step = f == 0 ? 1 : -1;
    } else if (f < 4) {
        i = col % a.nx + 1; k = col / a.nx + 1;
        j = f == 2 ? 1 : a.ny;
        step = f == 2 ? a.px : -a.px;
    } else {
        i = col % a.nx + 1; j = col / a.nx + 1;
        k = a.nz;
        step = -a.sxy;
    }
    base = at3(a.px, a.sxy, i + a.oi, j + a.oj, k);
}

__device__ __forceinline__ int lab_face(const LabDev &a, long long g, int &col)
{
    int f = 0;
#pragma unroll
    for (int q = 1; q < 5; q++) f += g >= a.col0[q] ? 1 : 0;
    col = (int)(g - a.col0[f]);
    return f;
}

__global__ void __launch_bounds__(256) k_lab_colmed(const LabDev a, float4 *__restrict__ out)
{
    const long long g = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= a.col0[5]) return;
    int col;
    const int f = lab_face(a, g, col);
    int base, step;
    lab_col(a, f, col, base, step);
    const float rho = a.rho[base], lam = a.lam[base], mu = a.mu[base];
    const float l2m = lam + 2.0f * mu;
    const float h = a.h[f < 2 ? 0 : (f < 4 ? 1 : 2)];
    float4 m = make_float4(0.0f, 0.0f, 1.0f, 0.0f);
    if (l2m > 0.0f) {
        const float vp = sqrtf(l2m / rho);
        const float vs = sqrtf(fmaxf(mu, 0.0f) / rho);
        m.x = vs / vp;
        m.y = vp * a.dt / h;
        m.z = rho * vp;
    }
    out[g] = m;
}

// The six strip values (cells d = 1..6) of one field of a column.  The
// x faces' columns lie along a row, one row per lane: with FP32 state and
// a 16-byte or 8-byte aligned first cell, one 16-byte and one 8-byte
// access cover the six cells (two cache-line requests per lane instead of
// six).  The alignment is the same for every column of a face (rows start
// at multiples of 32 elements).  A velocity component (a plain array).
template <typename T>
__device__ __forceinline__ void lab_strip_load(const T *__restrict__ F, int base, int step, bool xrow, T *v)
{
    if constexpr (sizeof(T) == 4) {
        const int lo = step > 0 ? base : base - (LAB_NS - 1);
        if (xrow && (lo & 1) == 0) {
            float w[LAB_NS];
            if ((lo & 3) == 0) {
                const float4 A = *(const float4 *)(F + lo);
                const float2 B = *(const float2 *)(F + lo + 4);
                w[0] = A.x; w[1] = A.y; w[2] = A.z; w[3] = A.w; w[4] = B.x; w[5] = B.y;
            } else {
                const float2 A = *(const float2 *)(F + lo);
                const float4 B = *(const float4 *)(F + lo + 2);
                w[0] = A.x; w[1] = A.y; w[2] = B.x; w[3] = B.y; w[4] = B.z; w[5] = B.w;
            }
#pragma unroll
            for (int d = 0; d < LAB_NS; d++) v[d] = step > 0 ? w[d] : w[LAB_NS - 1 - d];
            return;
        }
    }
#pragma unroll
    for (int d = 0; d < LAB_NS; d++) v[d] = F[base + d * step];
}

// The six strip values of each stress component, into v[q * LAB_NS + d]
// (q = 0..5).  The components are tiled (tix): component q lies q * RTILE
// elements after component 0 (S0) in every tile, so one set of cell
// offsets serves all six; an aligned group of 2 or 4 cells lies in one
// tile.
template <typename T>
__device__ __forceinline__ void lab_strip_load_s(const T *__restrict__ S0, int base, int step, bool xrow, T *v)
{
    if constexpr (sizeof(T) == 4) {
        const int lo = step > 0 ? base : base - (LAB_NS - 1);
        if (xrow && (lo & 1) == 0) {
            const bool a4 = (lo & 3) == 0;
            const float *pa = S0 + tix<NS3>(lo), *pb = S0 + tix<NS3>(lo + (a4 ? 4 : 2));
#pragma unroll
            for (int q = 0; q < NS3; q++) {
                float w[LAB_NS];
                if (a4) {
                    const float4 A = *(const float4 *)(pa + q * RTILE);
                    const float2 B = *(const float2 *)(pb + q * RTILE);
                    w[0] = A.x; w[1] = A.y; w[2] = A.z; w[3] = A.w; w[4] = B.x; w[5] = B.y;
                } else {
                    const float2 A = *(const float2 *)(pa + q * RTILE);
                    const float4 B = *(const float4 *)(pb + q * RTILE);
                    w[0] = A.x; w[1] = A.y; w[2] = B.x; w[3] = B.y; w[4] = B.z; w[5] = B.w;
                }
#pragma unroll
                for (int d = 0; d < LAB_NS; d++) v[q * LAB_NS + d] = step > 0 ? w[d] : w[LAB_NS - 1 - d];
            }
            return;
        }
    }
#pragma unroll
    for (int d = 0; d < LAB_NS; d++) {
        const T *p = S0 + tix<NS3>(base + d * step);
#pragma unroll
        for (int q = 0; q < NS3; q++) v[q * LAB_NS + d] = p[q * RTILE];
    }
}

// Store v[d] * sp[d] back to the six strip cells (the sponge).
template <typename T>
__device__ __forceinline__ void lab_strip_store(T *__restrict__ F, int base, int step, bool xrow, const float *sp,
                                                const T *v)
{
    T u[LAB_NS];
#pragma unroll
    for (int d = 0; d < LAB_NS; d++) u[d] = v[d] * (T)sp[d];
    if constexpr (sizeof(T) == 4) {
        const int lo = step > 0 ? base : base - (LAB_NS - 1);
        if (xrow && (lo & 1) == 0) {
            float w[LAB_NS];
#pragma unroll
            for (int d = 0; d < LAB_NS; d++) w[d] = step > 0 ? u[d] : u[LAB_NS - 1 - d];
            if ((lo & 3) == 0) {
                *(float4 *)(F + lo) = make_float4(w[0], w[1], w[2], w[3]);
                *(float2 *)(F + lo + 4) = make_float2(w[4], w[5]);
            } else {
                *(float2 *)(F + lo) = make_float2(w[0], w[1]);
                *(float4 *)(F + lo + 2) = make_float4(w[2], w[3], w[4], w[5]);
            }
            return;
        }
    }
#pragma unroll
    for (int d = 0; d < LAB_NS; d++) F[base + d * step] = u[d];
}

// v[q * LAB_NS + d] * sp[d] back to the stress components' strip cells
// (tiled, as lab_strip_load_s).
template <typename T>
__device__ __forceinline__ void lab_strip_store_s(T *__restrict__ S0, int base, int step, bool xrow, const float *sp,
                                                  const T *v)
{
    if constexpr (sizeof(T) == 4) {
        const int lo = step > 0 ? base : base - (LAB_NS - 1);
        if (xrow && (lo & 1) == 0) {
            const bool a4 = (lo & 3) == 0;
            float *pa = S0 + tix<NS3>(lo), *pb = S0 + tix<NS3>(lo + (a4 ? 4 : 2));
#pragma unroll
            for (int q = 0; q < NS3; q++) {
                float u[LAB_NS], w[LAB_NS];
#pragma unroll
                for (int d = 0; d < LAB_NS; d++) u[d] = v[q * LAB_NS + d] * sp[d];
#pragma unroll
                for (int d = 0; d < LAB_NS; d++) w[d] = step > 0 ? u[d] : u[LAB_NS - 1 - d];
                if (a4) {
                    *(float4 *)(pa + q * RTILE) = make_float4(w[0], w[1], w[2], w[3]);
                    *(float2 *)(pb + q * RTILE) = make_float2(w[4], w[5]);
                } else {
                    *(float2 *)(pa + q * RTILE) = make_float2(w[0], w[1]);
                    *(float4 *)(pb + q * RTILE) = make_float4(w[2], w[3], w[4], w[5]);
                }
            }
            return;
        }
    }
#pragma unroll
    for (int d = 0; d < LAB_NS; d++) {
        T *p = S0 + tix<NS3>(base + d * step);
#pragma unroll
        for (int q = 0; q < NS3; q++) p[q * RTILE] = v[q * LAB_NS + d] * (T)sp[d];
    }
}

// Features of one column into x[LAB_NIN]; prev holds the previous ghost
// values (stresses already divided by rho*vp).  raw[LAB_NF] receives the
// strip values as stored (field-major).  Returns N (0: no wave).
// Field fi's strip value d of global column g in the strip cache
// ([field][cell][column], so a wave's 64 columns are 64 consecutive values).
template <typename T>
__device__ __forceinline__ T *lab_cache_at(T *cache, long long ncol, int fi, int d, long long g)
{
    return cache + (long long)(fi * LAB_NS + d) * ncol + g;
}

// The strips of the velocities (vc) or the stresses (sc) come from the
// strip cache of column g instead of the fields.
template <typename T>
__device__ __forceinline__ float lab_features(const LabDev &a, int base, int step, bool xrow, float4 m,
                                              const float *prev, float *x, T *raw, const T *cache, long long ncol,
                                              long long g, bool vc, bool sc)
{
    const float inv_imp = 1.0f / m.z;
    float amax = 0.0f;
    if (vc) {
#pragma unroll
        for (int q = 0; q < 3 * LAB_NS; q++) raw[q] = *lab_cache_at(cache, ncol, q / LAB_NS, q % LAB_NS, g);
    } else {
#pragma unroll
        for (int fi = 0; fi < 3; fi++) lab_strip_load<T>((const T *)a.F[fi], base, step, xrow, raw + fi * LAB_NS);
    }
    if (sc) {
#pragma unroll
        for (int q = 3 * LAB_NS; q < LAB_NF; q++) raw[q] = *lab_cache_at(cache, ncol, q / LAB_NS, q % LAB_NS, g);
    } else {
        lab_strip_load_s<T>((const T *)a.F[3], base, step, xrow, raw + 3 * LAB_NS);
    }
#pragma unroll
    for (int fi = 0; fi < 9; fi++) {
#pragma unroll
        for (int d = 0; d < LAB_NS; d++) {
            float v = (float)raw[fi * LAB_NS + d];
            if (fi >= 3) v = v * inv_imp;
            x[fi * LAB_NS + d] = v;
            amax = fmaxf(amax, fabsf(v));
        }
    }
#pragma unroll
    for (int q = 0; q < LAB_NG; q++) {
        x[LAB_NF + q] = prev[q];
        amax = fmaxf(amax, fabsf(prev[q]));
    }
    // N = sqrt(mean square), formed on values scaled by their largest
    // magnitude so that the far-field precursors (|v| ~ 1e-25) do not
    // underflow; columns below LAB_TINY count as empty (N = 0), which keeps
    // 1/amax finite and denormals (flushed or not) out of the result
    float N = 0.0f;
    if (amax > LAB_TINY) {
        const float ia = 1.0f / amax;
        float s2 = 0.0f;
#pragma unroll
        for (int q = 0; q < LAB_NSC; q++) {
            const float u = x[q] * ia;
            s2 = fmaf(u, u, s2);
        }
        N = amax * sqrtf(s2 * (1.0f / (float)LAB_NSC));
        const float inv = 1.0f / N;
#pragma unroll
        for (int q = 0; q < LAB_NSC; q++) x[q] *= inv;
    }
    x[LAB_NSC] = m.x;
    x[LAB_NSC + 1] = m.y;
    return N;
}

// y = Ws x + W3 tanh(W2 tanh(W1 x + b1) + b2) + b3 with the net in LDS
// (tanhf of the device library).
__device__ __forceinline__ void lab_forward(const float *__restrict__ W, const float *x, float *y)
{
    float h1[LAB_NH], h2[LAB_NH];
#pragma unroll 4
    for (int h = 0; h < LAB_NH; h++) {
        float z = W[LAB_B1 + h];
#pragma unroll
        for (int q = 0; q < LAB_NIN; q++) z = fmaf(W[LAB_W1 + h * LAB_NIN + q], x[q], z);
        h1[h] = tanhf(z);
    }
#pragma unroll 4
    for (int h = 0; h < LAB_NH; h++) {
        float z = W[LAB_B2 + h];
#pragma unroll
        for (int g = 0; g < LAB_NH; g++) z = fmaf(W[LAB_W2 + h * LAB_NH + g], h1[g], z);
        h2[h] = tanhf(z);
    }
#pragma unroll
    for (int o = 0; o < LAB_NOUT; o++) {
        float z = W[LAB_B3 + o];
#pragma unroll
        for (int h = 0; h < LAB_NH; h++) z = fmaf(W[LAB_W3 + o * LAB_NH + h], h2[h], z);
#pragma unroll
        for (int q = 0; q < LAB_NIN; q++) z = fmaf(W[LAB_WS + o * LAB_NIN + q], x[q], z);
        y[o] = z;
    }
}

struct LabFillArgs {
    LabDev a;
    const uint4 *img;          // ten nets in their LDS form (lab_fill_image)
    const float4 *colmed;
    int phase;
    int face[2];               // blocks [0, nb0) fill face[0], the rest face[1]
    int nb0;
    float sp[LAB_NS];          // strip sponge factor of cell d = 1..6 (all 1: no sponge)
    int sponge;                // any factor below 1
    void *cache;               // strip cache (T [9][LAB_NS][ncol]) or null
    long long ncol;            // columns of the x faces (the cache's length)
    int use_cache;             // the other phase's strips come from the cache
};

// tanh(z) = 1 - 2/(1 + e^(2z)) with the hardware exponential and a fast
// reciprocal: absolute error ~1e-7; saturates cleanly (e^(2z) = inf gives
// 1, 0 gives -1).
__device__ __forceinline__ float lab_tanh(float z)
{
    const float t = __expf(2.0f * z);
    return 1.0f - __fdividef(2.0f, 1.0f + t);
}

// BF16 of an FP32 value, nearest even (finite inputs)
__device__ __forceinline__ __bf16 lab_bf16(float f)
{
    unsigned u = __float_as_uint(f);
    u += 0x7fffu + ((u >> 16) & 1u);
    const unsigned short h = (unsigned short)(u >> 16);
    __bf16 r;
    __builtin_memcpy(&r, &h, 2);
    return r;
}

typedef __bf16 lab_bf16x8 __attribute__((ext_vector_type(8)));
typedef float lab_f32x16 __attribute__((ext_vector_type(16)));

// LDS row strides (elements): 8 extra BF16 per row spread the 16-byte
// operand reads of one MFMA over the banks.
constexpr int LX = 72;     // features, BF16 [64 columns][72]
constexpr int LH = 40;     // hidden layer 1, BF16 [64][40]
constexpr int LH2 = 33;    // hidden layer 2, FP32 [64][33] (in the feature rows' space)
constexpr int LW1 = 72;    // W1 as BF16 [32][72]
constexpr int LW2 = 40;    // W2 as BF16 [32][40]
constexpr int LO = 96;     // output layer as FP32 [5][96]: W3 row, Ws row, b3
constexpr int FILL_WAVES = 4;
// Per-wave LDS region, shared by three arrays that are each written only
// after the rows they overwrite have been consumed by the same wave:
//   features X  BF16 [64][LX] at 0 (9216 B), read by layer 1;
//   layer 1 H1  BF16 [64][LH] at 0 (5120 B): tile t writes its rows after
//               its MFMAs have read X rows 32t..32t+31, and rows 32..63 of
//               X start at byte 4608, past H1's tile-0 rows;
//   layer 2 H2  FP32 rows of LH2: tile 0's rows at H2_T0 = 5120 (past all
//               of H1), tile 1's at 0 (H1 rows 0..31 are dead by then; tile
//               1 reads its own H1 rows before it writes).
// 9344 B per wave instead of 14336 lets three 4-wave blocks share a CU.
constexpr int H2_T0 = 64 * LH * 2;
constexpr int WAVE_LDS = H2_T0 + 32 * LH2 * 4;   // 9344 bytes per wave
static_assert(WAVE_LDS >= 64 * LX * 2 && WAVE_LDS % 16 == 0, "fill LDS region");
constexpr int IMG_W2 = 32 * LW1 * 2, IMG_O = IMG_W2 + 32 * LW2 * 2;      // byte offsets in a net's image
constexpr int IMG_B1 = IMG_O + LAB_NOUT * LO * 4, IMG_B2 = IMG_B1 + 32 * 4;
constexpr int FILL_LDS0 = IMG_B2 + 32 * 4;   // one net's image: bytes before the waves' rows
static_assert(FILL_LDS0 == LAB_IMG_BYTES && FILL_LDS0 % 16 == 0, "net image layout");

// The ghost fill on the matrix cores.  Each wave owns 64 columns (one per
// lane).  A lane forms its column's 61 features in FP32; the two hidden
// layers then run as 32x32x16 BF16 MFMAs with FP32 accumulation over the
// wave's 64 columns (features and layer-1 activations rounded to BF16 on
// the way in, as the weights are), and the output layer Ws x + W3 h2 + b3
// stays in FP32 per lane, its weights read from LDS at wave-uniform
// addresses.  The strip loads of all nine fields go out together, and the
// sponge stores follow from the loaded values, so a column costs one
// memory round trip.  Opposite faces (disjoint strips) share a launch.  MFMA operand layouts (gfx950,
// v_mfma_f32_32x32x16_bf16): lane l holds A[l % 32][8 (l / 32) + 0..7] and
// B[8 (l / 32) + 0..7][l % 32]; the accumulator's register r holds
// D[8 (r / 4) + 4 (l / 32) + r % 4][l % 32].
template <typename T>
__device__ __forceinline__ void lab_fill_body(const LabFillArgs &p)
{
    extern __shared__ __align__(16) unsigned char lds[];
    const __bf16 *sW1 = (const __bf16 *)lds;                 // [32][LW1]
    const __bf16 *sW2 = (const __bf16 *)(lds + IMG_W2);      // [32][LW2]
    const float *sB1 = (const float *)(lds + IMG_B1), *sB2 = (const float *)(lds + IMG_B2);
    const int wave = threadIdx.x >> 6, l = threadIdx.x & 63;
    const float *sO = (const float *)(lds + IMG_O);          // [LAB_NOUT][LO]: W3 row, Ws row, b3
    unsigned char *wl = lds + FILL_LDS0 + wave * WAVE_LDS;
    __bf16 *sX = (__bf16 *)wl;                         // [64][LX]
    __bf16 *sH1 = (__bf16 *)wl;                        // [64][LH], over X (see WAVE_LDS)
    float *sH2t[2] = {(float *)(wl + H2_T0), (float *)wl};   // rows 0..31, rows 32..63

    const bool second = (int)blockIdx.x >= p.nb0;
    const int f = p.face[second ? 1 : 0], net = 2 * f + p.phase;
    // the net's LDS image, 16 bytes per lane per round: loaded first,
    // written to LDS after the column's loads have been issued
    constexpr int NV = FILL_LDS0 / 16, NT = 64 * FILL_WAVES, NR = (NV + NT - 1) / NT;
    uint4 imgr[NR];
    {
        const uint4 *__restrict__ src = p.img + (size_t)net * NV;
#pragma unroll
        for (int r = 0; r < NR; r++) {
            const int q = threadIdx.x + r * NT;
            imgr[r] = q < NV ? src[q] : make_uint4(0, 0, 0, 0);
        }
    }

    const LabDev &a = p.a;
    const int col = ((int)blockIdx.x - (second ? p.nb0 : 0)) * blockDim.x + threadIdx.x;
    const bool have = col < a.ncol[f];
    const bool xrow = f < 2;
    int base = 0, step = 0;
    float4 m = make_float4(0.0f, 0.0f, 1.0f, 0.0f);
    const int nout = c_nout[net];
    int gaddr[LAB_NOUT];
    float prev[LAB_NG], x[LAB_NIN];
    float N = 0.0f;
    if (have) {
        lab_col(a, f, col, base, step);
        m = p.colmed[a.col0[f] + col];
#pragma unroll
        for (int q = 0; q < LAB_NOUT; q++) {
            const int fi = c_tgt[net][q][0], h = c_tgt[net][q][1];
            gaddr[q] = base - h * step;
            const float v = q < nout ? lab_ld<T>(a, fi, gaddr[q]) : 0.0f;
            prev[q] = fi >= 3 ? v / m.z : v;
        }
        T raw[LAB_NF];
        // Only the x faces' columns are cached: their strips lie along rows,
        // so the fields' strips cost a scattered line per field, where the
        // other faces read theirs in coalesced rows.  And a column is cached
        // only if no later face of the phase (faces run x pair, y pair,
        // bottom) rescales its strip cells with its sponge: away from the y
        // faces' and the bottom's strips.
        bool own = false;
        if (f < 2) {
            const int jj = col % a.ny + 1, kk = col / a.ny + 1;
            own = jj > LAB_NS && jj <= a.ny - LAB_NS && kk <= a.nz - LAB_NS;
        }
        T *cache = own ? (T *)p.cache : nullptr;
        const long long g = a.col0[f] + col;
        const bool use = cache && p.use_cache;
        N = lab_features<T>(a, base, step, xrow, m, prev, x, raw, cache, p.ncol, g, use && p.phase == 1,
                            use && p.phase == 0);
        // strip sponge on this phase's own fields, from the values read
        // above (the column owns its strip cells within the launch)
        if (p.sponge) {
            if (p.phase == 0) {
#pragma unroll
                for (int fi = 0; fi < 3; fi++) lab_strip_store<T>((T *)a.F[fi], base, step, xrow, p.sp, raw + fi * LAB_NS);
            } else {
                lab_strip_store_s<T>((T *)a.F[3], base, step, xrow, p.sp, raw + 3 * LAB_NS);
            }
        }
        // this phase's strips as they now stand, for the other phase: the
        // half-step between them leaves these fields alone
        if (cache) {
            const int q0 = p.phase == 0 ? 0 : 3 * LAB_NS, q1 = p.phase == 0 ? 3 * LAB_NS : LAB_NF;
#pragma unroll
            for (int q = 0; q < LAB_NF; q++)
                if (q >= q0 && q < q1)
                    *lab_cache_at(cache, p.ncol, q / LAB_NS, q % LAB_NS, g) =
                        p.sponge ? raw[q] * (T)p.sp[q % LAB_NS] : raw[q];
        }
    } else {
#pragma unroll
        for (int q = 0; q < LAB_NIN; q++) x[q] = 0.0f;
    }
#pragma unroll
    for (int r = 0; r < NR; r++) {
        const int q = threadIdx.x + r * NT;
        if (q < NV) ((uint4 *)lds)[q] = imgr[r];
    }
    const bool live = have && N > 0.0f && m.y > 0.0f;   // air columns keep zero ghosts
#pragma unroll
    for (int q = 0; q < LX; q++) sX[l * LX + q] = lab_bf16(q < LAB_NIN && live ? x[q] : 0.0f);
    __syncthreads();

    // hidden layer 1: H1 = tanh(X W1^T + b1), two 32-row tiles of the wave
    const int ln = l & 31, lk = 8 * (l >> 5);
    const float b1 = sB1[ln], b2 = sB2[ln];
#pragma unroll
    for (int t = 0; t < 2; t++) {
        lab_f32x16 acc = {};
#pragma unroll
        for (int kc = 0; kc < 4; kc++) {
            const lab_bf16x8 av = *(const lab_bf16x8 *)(sX + (32 * t + ln) * LX + 16 * kc + lk);
            const lab_bf16x8 bv = *(const lab_bf16x8 *)(sW1 + ln * LW1 + 16 * kc + lk);
            acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(av, bv, acc, 0, 0, 0);
        }
#pragma unroll
        for (int r = 0; r < 16; r++) {
            const int row = 32 * t + 8 * (r >> 2) + 4 * (l >> 5) + (r & 3);
            sH1[row * LH + ln] = lab_bf16(lab_tanh(acc[r] + b1));
        }
    }
    __syncthreads();   // layer 1 done: the feature rows are free for H2

    // hidden layer 2: H2 = tanh(H1 W2^T + b2), kept in FP32
#pragma unroll
    for (int t = 0; t < 2; t++) {
        lab_f32x16 acc = {};
#pragma unroll
        for (int kc = 0; kc < 2; kc++) {
            const lab_bf16x8 av = *(const lab_bf16x8 *)(sH1 + (32 * t + ln) * LH + 16 * kc + lk);
            const lab_bf16x8 bv = *(const lab_bf16x8 *)(sW2 + ln * LW2 + 16 * kc + lk);
            acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(av, bv, acc, 0, 0, 0);
        }
#pragma unroll
        for (int r = 0; r < 16; r++) {
            const int row = 32 * t + 8 * (r >> 2) + 4 * (l >> 5) + (r & 3);
            sH2t[t][(row - 32 * t) * LH2 + ln] = lab_tanh(acc[r] + b2);
        }
    }
    __syncthreads();

    if (!have) return;
    // output layer in FP32: y = b3 + W3 h2 + Ws x
    float h2[LAB_NH];
#pragma unroll
    for (int h = 0; h < LAB_NH; h++) h2[h] = sH2t[l >> 5][(l & 31) * LH2 + h];
    // the ghosts are stored once all are formed: a stress ghost's 64-bit
    // offset in its tile (lab_ix) then needs no registers beside the
    // features and hidden values
    float out[LAB_NOUT];
#pragma unroll
    for (int o = 0; o < LAB_NOUT; o++) {
        out[o] = 0.0f;
        if (o < nout) {
            const float *wo = sO + o * LO;
            float z = wo[LAB_NH + LAB_NIN];
#pragma unroll
            for (int h = 0; h < LAB_NH; h++) z = fmaf(wo[h], h2[h], z);
#pragma unroll
            for (int q = 0; q < LAB_NIN; q++) z = fmaf(wo[LAB_NH + q], x[q], z);
            const int fi = c_tgt[net][o][0];
            float v = live ? z * N : 0.0f;
            if (fi >= 3) v = v * m.z;
            out[o] = v;
        }
    }
#pragma unroll
    for (int o = 0; o < LAB_NOUT; o++)
        if (o < nout) lab_st<T>(a, c_tgt[net][o][0], gaddr[o], out[o]);
}

// FP32 state: 168 VGPRs and three 4-wave blocks per CU (the LDS of
// lab_fill_body fits three); FP64 state: two, without spills.
__global__ void __launch_bounds__(64 * FILL_WAVES) __attribute__((amdgpu_waves_per_eu(3, 3)))
k_lab_fill_f32(const LabFillArgs p)
{
    lab_fill_body<float>(p);
}

__global__ void __launch_bounds__(64 * FILL_WAVES) __attribute__((amdgpu_waves_per_eu(2, 2))) k_lab_fill_f64(const LabFillArgs p)
{
    lab_fill_body<double>(p);
}

int lab_colmed(const LabGrid &g, const LabDomain &dom, DBuf &out, long long &bytes, hipStream_t s)
{
    const long long n = dom.col0[5];
    if (int r = dalloc(out, (size_t)n * sizeof(float4), bytes)) return r;
    if (n == 0) return SWPC_OK;
    const LabDev a = make_dev(g, dom);
    hipLaunchKernelGGL(k_lab_colmed, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, a, (float4 *)out.p);
    SWPC_CHECK_LAUNCH();
    return SWPC_OK;
}

void lab_sponge_factors(float sponge, float *sp)
{
    for (int d = 0; d < LAB_NS; d++) {
        const double r = (double)(LAB_NS - d) / LAB_NS;
        sp[d] = (float)(1.0 - (double)sponge * r * r);
    }
}

void lab_fill_image(const float *w, unsigned char *img)
{
    for (int n = 0; n < LAB_NNET; n++) {
        const float *W = w + (size_t)n * LAB_NP;
        unsigned char *o = img + (size_t)n * LAB_IMG_BYTES;
        memset(o, 0, LAB_IMG_BYTES);
        auto bf = [](float v) {   // the upper half of an FP32 value already rounded to BF16
            uint32_t u;
            memcpy(&u, &v, 4);
            return (uint16_t)(u >> 16);
        };
        uint16_t *w1 = (uint16_t *)o, *w2 = (uint16_t *)(o + IMG_W2);
        for (int h = 0; h < 32; h++) {
            for (int k = 0; k < LAB_NIN; k++) w1[h * LW1 + k] = bf(W[LAB_W1 + h * LAB_NIN + k]);
            for (int k = 0; k < LAB_NH; k++) w2[h * LW2 + k] = bf(W[LAB_W2 + h * LAB_NH + k]);
        }
        float *ow = (float *)(o + IMG_O);
        for (int q = 0; q < LAB_NOUT; q++) {
            for (int h = 0; h < LAB_NH; h++) ow[q * LO + h] = W[LAB_W3 + q * LAB_NH + h];
            for (int k = 0; k < LAB_NIN; k++) ow[q * LO + LAB_NH + k] = W[LAB_WS + q * LAB_NIN + k];
            ow[q * LO + LAB_NH + LAB_NIN] = W[LAB_B3 + q];
        }
        float *b1 = (float *)(o + IMG_B1), *b2 = (float *)(o + IMG_B2);
        for (int h = 0; h < 32; h++) {
            b1[h] = W[LAB_B1 + h];
            b2[h] = W[LAB_B2 + h];
        }
    }
}

int lab_fill(const LabGrid &g, const LabDomain &dom, const void *img, const float4 *colmed, int phase, float sponge,
             void *cache, bool use_cache, hipStream_t s)
{
    LabFillArgs p;
    lab_sponge_factors(sponge, p.sp);
    p.sponge = sponge > 0.0f;
    p.cache = cache;
    p.ncol = dom.col0[2];   // the x faces' columns (the cached ones) come first
    p.use_cache = use_cache ? 1 : 0;
    p.a = make_dev(g, dom);
    p.img = (const uint4 *)img;
    p.colmed = colmed;
    p.phase = phase;
    const size_t lds = (size_t)FILL_LDS0 + (size_t)FILL_WAVES * WAVE_LDS;
    const int cols = 64 * FILL_WAVES;
    // faces in order x-, x+, y-, y+, bottom; opposite faces share a launch
    // when their strips are disjoint (the sponge of one does not touch the
    // other's features)
    int f = 0;
    while (f < 5) {
        const bool pair = f < 4 && (f < 2 ? dom.nx : dom.ny) >= 2 * LAB_NS;
        p.face[0] = f;
        p.face[1] = pair ? f + 1 : f;
        p.nb0 = (dom.ncol[f] + cols - 1) / cols;
        const int nb = p.nb0 + (pair ? (dom.ncol[f + 1] + cols - 1) / cols : 0);
        f += pair ? 2 : 1;
        if (nb == 0) continue;
        if (g.f64) hipLaunchKernelGGL(k_lab_fill_f64, dim3(nb), dim3(cols), lds, s, p);
        else hipLaunchKernelGGL(k_lab_fill_f32, dim3(nb), dim3(cols), lds, s, p);
        SWPC_CHECK_LAUNCH();
    }
    return SWPC_OK;
}

__host__ __device__ __forceinline__ unsigned lab_hash(unsigned x)
{
    x ^= x >> 16;
    x *= 0x7feb352du;
    x ^= x >> 15;
    x *= 0x846ca68bu;
    x ^= x >> 16;
    return x;
}

}  // namespace swpc

// ---------------------------------------------------------------------------
// trainer
// ---------------------------------------------------------------------------
struct swpc_lab_trainer {
    int cap = 0;                 // samples per net
    long long bytes = 0;
    hipStream_t stream = nullptr;
    swpc::DBuf samples;          // [10][cap][LAB_REC]
    swpc::DBuf count;            // int[10], may run past the limit until commit
    swpc::DBuf runmax;           // float[10] (as uint for atomicMax)
    swpc::DBuf prev;             // gather domain: previous ghost values [10][prev_cols][5]
    long long prev_cols = 0;
    swpc::LabDomain dom;         // gather domain inside the wide grid
    swpc::DBuf colmed;           // its column medium
    int limit[swpc::LAB_NNET] = {0};
    swpc::DBuf P, M, V, partial, ploss, loss, X;   // optimiser state
};

extern "C" int swpc_lab_count(swpc_lab_trainer *tr, int *counts);

namespace swpc {

struct LabGatherArgs {
    LabDev a;
    const float4 *colmed;
    float *samples;
    int *count;
    unsigned *runmax;
    float *prev;                 // [10][prev_cols][5]
    long long prev_cols;
    int cap;
    int limit[5];
    int phase;
    float keep, thr;
    unsigned seed;
};

template <typename T>
__global__ void __launch_bounds__(256) k_lab_gather(const LabGatherArgs p)
{
    const LabDev &a = p.a;
    const long long g = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= a.col0[5]) return;
    int col;
    const int f = lab_face(a, g, col);
    const int net = 2 * f + p.phase;
    int base, step;
    lab_col(a, f, col, base, step);
    const float4 m = p.colmed[g];
    const int nout = c_nout[net];
    float *pv = p.prev + ((size_t)net * p.prev_cols + col) * LAB_NG;
    float prev[LAB_NG], t[LAB_NOUT];
#pragma unroll
    for (int q = 0; q < LAB_NOUT; q++) {
        prev[q] = pv[q];
        const int fi = c_tgt[net][q][0], h = c_tgt[net][q][1];
        float v = q < nout ? lab_ld<T>(a, fi, base - h * step) : 0.0f;
        t[q] = fi >= 3 ? v / m.z : v;
    }
    float x[LAB_NIN];
    T raw[LAB_NF];
    const float N = lab_features<T>(a, base, step, false, m, prev, x, raw, (const T *)nullptr, 0, 0, false, false);
#pragma unroll
    for (int q = 0; q < LAB_NOUT; q++) pv[q] = t[q];
    if (!(N > 0.0f) || !(m.y > 0.0f)) return;
    const float rmax = __uint_as_float(p.runmax[net]);
    atomicMax(p.runmax + net, __float_as_uint(N));
    if (N < p.thr * rmax) return;
    const unsigned u = lab_hash(p.seed ^ lab_hash((unsigned)g * 2u + (unsigned)p.phase + 0x9e3779b9u));
    if ((float)(u >> 8) * (1.0f / 16777216.0f) >= p.keep) return;
    const int idx = atomicAdd(p.count + net, 1);
    if (idx >= p.limit[f] || idx >= p.cap) return;
    float *o = p.samples + ((size_t)net * p.cap + idx) * LAB_REC;
    const float inv = 1.0f / N;
#pragma unroll
    for (int q = 0; q < LAB_NIN; q++) o[q] = x[q];
#pragma unroll
    for (int q = 0; q < LAB_NOUT; q++) o[LAB_T0 + q] = t[q] * inv;
    o[LAB_T0 + 5] = (float)col;
    o[LAB_T0 + 6] = N;
    o[LAB_T0 + 7] = m.z;
#pragma unroll
    for (int q = LAB_T0 + 8; q < LAB_REC; q++) o[q] = 0.0f;
}

int lab_begin(swpc_lab_trainer *tr, const LabGrid &g, const LabDomain &dom, const int *quota, hipStream_t s)
{
    int c[LAB_NNET];
    if (int r = swpc_lab_count(tr, c)) return r;
    long long maxcol = 0;
    for (int f = 0; f < 5; f++) maxcol = std::max(maxcol, (long long)dom.ncol[f]);
    if (int r = dalloc(tr->prev, (size_t)LAB_NNET * maxcol * LAB_NG * sizeof(float), tr->bytes)) return r;
    tr->prev_cols = maxcol;
    tr->dom = dom;
    if (int r = lab_colmed(g, dom, tr->colmed, tr->bytes, s)) return r;
    SWPC_HIP(hipMemsetAsync(tr->runmax.p, 0, LAB_NNET * sizeof(unsigned), s));
    for (int n = 0; n < LAB_NNET; n++) tr->limit[n] = std::min(tr->cap, c[n] + std::max(0, quota ? quota[n] : tr->cap));
    SWPC_HIP(hipStreamSynchronize(s));
    return SWPC_OK;
}

int lab_gather(swpc_lab_trainer *tr, const LabGrid &g, int phase, float keep, unsigned seed, float thr, hipStream_t s)
{
    const LabDomain &dom = tr->dom;
    const long long n = dom.col0[5];
    if (!tr->prev.p) return fail(SWPC_ESTATE, "swpc_lab_begin_run must precede swpc3d_lab_gather");
    LabGatherArgs p;
    p.a = make_dev(g, dom);
    p.colmed = (const float4 *)tr->colmed.p;
    p.samples = (float *)tr->samples.p;
    p.count = (int *)tr->count.p;
    p.runmax = (unsigned *)tr->runmax.p;
    p.prev = (float *)tr->prev.p;
    p.prev_cols = tr->prev_cols;
    p.cap = tr->cap;
    for (int f = 0; f < 5; f++) p.limit[f] = tr->limit[2 * f + phase];
    p.phase = phase;
    p.keep = keep;
    p.thr = thr;
    p.seed = seed;
    if (n == 0) return SWPC_OK;
    if (g.f64) hipLaunchKernelGGL(k_lab_gather<double>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, p);
    else hipLaunchKernelGGL(k_lab_gather<float>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, p);
    SWPC_CHECK_LAUNCH();
    return SWPC_OK;
}

// ---------------------------------------------------------------------------
// training
// ---------------------------------------------------------------------------
constexpr int TB = 128;                 // samples (threads) per workgroup
constexpr int XP = LAB_NIN + 1;         // padded LDS rows (odd strides)
constexpr int HP = LAB_NH + 1;

struct FwdBwdArgs {
    const float *P;        // FP32 parameters of one net
    const float *samples;  // records of this net
    int count;             // records available
    int nout;
    int mode;              // 0: random records with random sign; 1: records 0..count-1 as given
    unsigned seed;
    float noise;           // mode 0: N(0, noise^2) added to the non-zero previous-ghost features
    float scale;           // 2 / (batch * nout)
    float *partial;        // [nblk][LAB_NP]
    float *ploss;          // [nblk]
};

// Forward and backward of TB samples per workgroup: one thread per sample
// keeps its activations in registers and parks x, h1, h2 and the deltas in
// LDS; the weight gradients are then summed over the workgroup's samples
// from LDS, each thread owning a set of parameters.  The loss is the mean
// square error over the batch and the net's outputs; with mode 0 a random
// half of the samples enter with the sign of every field feature and
// target flipped (the wave equation is odd in the fields).
__global__ void __launch_bounds__(TB) k_lab_fwdbwd(const FwdBwdArgs p)
{
    __shared__ float W[LAB_NP];
    __shared__ float sX[TB][XP];
    __shared__ float sH1[TB][HP], sD1[TB][HP], sH2[TB][HP], sD2[TB][HP];
    __shared__ float sD3[TB][LAB_NOUT];
    __shared__ float red[TB];
    const int tid = threadIdx.x;
    for (int q = tid; q < LAB_NP; q += TB) W[q] = p.P[q];
    __syncthreads();
    const int b = blockIdx.x * TB + tid;
    int idx;
    float sg = 1.0f;
    if (p.mode == 0) {
        const unsigned u = lab_hash(p.seed ^ lab_hash((unsigned)b + 0x632be5abu));
        idx = p.count > 0 ? (int)(u % (unsigned)p.count) : -1;
        sg = (lab_hash(u) & 1u) ? -1.0f : 1.0f;
    } else {
        idx = b < p.count ? b : -1;
    }
    float x[LAB_NIN], t[LAB_NOUT];
    if (idx >= 0) {
        const float *r = p.samples + (size_t)idx * LAB_REC;
#pragma unroll
        for (int q = 0; q < LAB_NIN; q++) x[q] = q < LAB_NSC ? sg * r[q] : r[q];
#pragma unroll
        for (int q = 0; q < LAB_NOUT; q++) t[q] = sg * r[LAB_T0 + q];
        if (p.mode == 0 && p.noise > 0.0f) {
            // Box-Muller pairs from hashed uniforms: the recurrent ghost
            // inputs are trained against perturbations of their own
            // previous outputs, which keeps the closed loop stable
#pragma unroll
            for (int q = 0; q < LAB_NG; q += 2) {
                const unsigned h1 = lab_hash(p.seed ^ lab_hash((unsigned)b * 8u + (unsigned)q + 0x51ed27u));
                const unsigned h2 = lab_hash(h1 + 0x2545f491u);
                const float u1 = ((float)(h1 >> 8) + 0.5f) * (1.0f / 16777216.0f);
                const float u2 = (float)(h2 >> 8) * (1.0f / 16777216.0f);
                const float rad = sqrtf(-2.0f * logf(u1));
                const float g0 = rad * cosf(6.2831853f * u2), g1 = rad * sinf(6.2831853f * u2);
                if (x[LAB_NF + q] != 0.0f) x[LAB_NF + q] += p.noise * g0;
                if (q + 1 < LAB_NG && x[LAB_NF + q + 1] != 0.0f) x[LAB_NF + q + 1] += p.noise * g1;
            }
        }
    } else {
#pragma unroll
        for (int q = 0; q < LAB_NIN; q++) x[q] = 0.0f;
#pragma unroll
        for (int q = 0; q < LAB_NOUT; q++) t[q] = 0.0f;
    }
    float h1[LAB_NH], h2[LAB_NH], y[LAB_NOUT];
    for (int h = 0; h < LAB_NH; h++) {
        float z = W[LAB_B1 + h];
#pragma unroll
        for (int q = 0; q < LAB_NIN; q++) z = fmaf(W[LAB_W1 + h * LAB_NIN + q], x[q], z);
        h1[h] = tanhf(z);
    }
    for (int h = 0; h < LAB_NH; h++) {
        float z = W[LAB_B2 + h];
#pragma unroll
        for (int g = 0; g < LAB_NH; g++) z = fmaf(W[LAB_W2 + h * LAB_NH + g], h1[g], z);
        h2[h] = tanhf(z);
    }
    float l = 0.0f;
    float d3[LAB_NOUT];
#pragma unroll
    for (int o = 0; o < LAB_NOUT; o++) {
        float z = W[LAB_B3 + o];
#pragma unroll
        for (int h = 0; h < LAB_NH; h++) z = fmaf(W[LAB_W3 + o * LAB_NH + h], h2[h], z);
#pragma unroll
        for (int q = 0; q < LAB_NIN; q++) z = fmaf(W[LAB_WS + o * LAB_NIN + q], x[q], z);
        y[o] = z;
        const float e = (idx >= 0 && o < p.nout) ? y[o] - t[o] : 0.0f;
        l = fmaf(e, e, l);
        d3[o] = p.scale * e;
    }
    float d2[LAB_NH];
    for (int h = 0; h < LAB_NH; h++) {
        float s = 0.0f;
#pragma unroll
        for (int o = 0; o < LAB_NOUT; o++) s = fmaf(W[LAB_W3 + o * LAB_NH + h], d3[o], s);
        d2[h] = s * (1.0f - h2[h] * h2[h]);
    }
    for (int g = 0; g < LAB_NH; g++) {
        float s = 0.0f;
#pragma unroll
        for (int h = 0; h < LAB_NH; h++) s = fmaf(W[LAB_W2 + h * LAB_NH + g], d2[h], s);
        sD1[tid][g] = s * (1.0f - h1[g] * h1[g]);
    }
#pragma unroll
    for (int q = 0; q < LAB_NIN; q++) sX[tid][q] = x[q];
#pragma unroll
    for (int h = 0; h < LAB_NH; h++) {
        sH1[tid][h] = h1[h];
        sH2[tid][h] = h2[h];
        sD2[tid][h] = d2[h];
    }
#pragma unroll
    for (int o = 0; o < LAB_NOUT; o++) sD3[tid][o] = d3[o];
    red[tid] = l;
    __syncthreads();
    float *out = p.partial + (size_t)blockIdx.x * LAB_NP;
    for (int q = tid; q < LAB_NP; q += TB) {
        float s = 0.0f;
        if (q < LAB_B1) {
            const int h = q / LAB_NIN, f = q % LAB_NIN;
            for (int n = 0; n < TB; n++) s = fmaf(sD1[n][h], sX[n][f], s);
        } else if (q < LAB_W2) {
            const int h = q - LAB_B1;
            for (int n = 0; n < TB; n++) s += sD1[n][h];
        } else if (q < LAB_B2) {
            const int h = (q - LAB_W2) / LAB_NH, g = (q - LAB_W2) % LAB_NH;
            for (int n = 0; n < TB; n++) s = fmaf(sD2[n][h], sH1[n][g], s);
        } else if (q < LAB_W3) {
            const int h = q - LAB_B2;
            for (int n = 0; n < TB; n++) s += sD2[n][h];
        } else if (q < LAB_B3) {
            const int o = (q - LAB_W3) / LAB_NH, h = (q - LAB_W3) % LAB_NH;
            for (int n = 0; n < TB; n++) s = fmaf(sD3[n][o], sH2[n][h], s);
        } else if (q < LAB_WS) {
            const int o = q - LAB_B3;
            for (int n = 0; n < TB; n++) s += sD3[n][o];
        } else {
            const int o = (q - LAB_WS) / LAB_NIN, f = (q - LAB_WS) % LAB_NIN;
            for (int n = 0; n < TB; n++) s = fmaf(sD3[n][o], sX[n][f], s);
        }
        out[q] = s;
    }
    if (tid == 0) {
        float s = 0.0f;
        for (int n = 0; n < TB; n++) s += red[n];
        p.ploss[blockIdx.x] = s;
    }
}

// Sum of the partials in workgroup order (deterministic), then Adam with
// bias correction; thread 0 of workgroup 0 also records the step's loss.
__global__ void __launch_bounds__(256) k_lab_adam(float *__restrict__ P, float *__restrict__ M, float *__restrict__ V,
                                                  const float *__restrict__ partial, const float *__restrict__ ploss,
                                                  int nblk, float lr, float b1, float b2, float eps, float c1, float c2,
                                                  float lnorm, float *__restrict__ loss_out, float *__restrict__ grad)
{
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q == 0 && loss_out) {
        float s = 0.0f;
        for (int n = 0; n < nblk; n++) s += ploss[n];
        *loss_out = s * lnorm;
    }
    if (q >= LAB_NP) return;
    float g = 0.0f;
    for (int n = 0; n < nblk; n++) g += partial[(size_t)n * LAB_NP + q];
    if (grad) {
        grad[q] = g;
        return;
    }
    const float m = b1 * M[q] + (1.0f - b1) * g;
    const float v = b2 * V[q] + (1.0f - b2) * g * g;
    M[q] = m;
    V[q] = v;
    P[q] -= lr * (m * c1) / (sqrtf(v * c2) + eps);
}

static int tr_ensure(swpc_lab_trainer *tr, int nblk)
{
    if (!tr->P.p) {
        if (int r = dalloc(tr->P, LAB_NP * sizeof(float), tr->bytes)) return r;
        if (int r = dalloc(tr->M, LAB_NP * sizeof(float), tr->bytes)) return r;
        if (int r = dalloc(tr->V, LAB_NP * sizeof(float), tr->bytes)) return r;
    }
    if (tr->partial.bytes < (size_t)nblk * LAB_NP * sizeof(float)) {
        if (int r = dalloc(tr->partial, (size_t)nblk * LAB_NP * sizeof(float), tr->bytes)) return r;
        if (int r = dalloc(tr->ploss, (size_t)nblk * sizeof(float), tr->bytes)) return r;
    }
    return SWPC_OK;
}

}  // namespace swpc

using namespace swpc;

extern "C" {

int swpc_lab_nparam(void) { return LAB_NP; }

int swpc_lab_trainer_create(int cap, swpc_lab_trainer **out)
{
    clear_error();
    if (!out || cap <= 0) return fail(SWPC_EINVAL, "bad trainer capacity");
    *out = nullptr;
    auto *tr = new swpc_lab_trainer();
    tr->cap = cap;
    int r;
    if (hipStreamCreate(&tr->stream) != hipSuccess) r = fail(SWPC_EHIP, "hipStreamCreate failed");
    else if ((r = dalloc(tr->samples, (size_t)LAB_NNET * cap * LAB_REC * sizeof(float), tr->bytes)) == 0 &&
             (r = dalloc(tr->count, LAB_NNET * sizeof(int), tr->bytes)) == 0)
        r = dalloc(tr->runmax, LAB_NNET * sizeof(unsigned), tr->bytes);
    if (r) {
        swpc_lab_trainer_destroy(tr);
        return r;
    }
    *out = tr;
    return SWPC_OK;
}

int swpc_lab_trainer_destroy(swpc_lab_trainer *tr)
{
    if (!tr) return SWPC_OK;
    if (tr->stream) (void)hipStreamSynchronize(tr->stream);
    for (DBuf *b : {&tr->samples, &tr->count, &tr->runmax, &tr->prev, &tr->P, &tr->M, &tr->V, &tr->partial,
                    &tr->ploss, &tr->loss, &tr->X})
        dfree(*b);
    if (tr->stream) (void)hipStreamDestroy(tr->stream);
    delete tr;
    return SWPC_OK;
}

// Clamp the sample counts to what was stored and copy them to counts[10].
int swpc_lab_count(swpc_lab_trainer *tr, int *counts)
{
    clear_error();
    if (!tr || !counts) return fail(SWPC_EINVAL, "null argument");
    SWPC_HIP(hipDeviceSynchronize());
    int c[LAB_NNET];
    SWPC_HIP(hipMemcpy(c, tr->count.p, sizeof c, hipMemcpyDeviceToHost));
    for (int n = 0; n < LAB_NNET; n++) c[n] = std::min(c[n], std::min(tr->cap, tr->limit[n]));
    SWPC_HIP(hipMemcpy(tr->count.p, c, sizeof c, hipMemcpyHostToDevice));
    for (int n = 0; n < LAB_NNET; n++) counts[n] = c[n];
    return SWPC_OK;
}

int swpc_lab_add(swpc_lab_trainer *tr, int net, int n, const float *rec)
{
    clear_error();
    if (!tr || !rec || net < 0 || net >= LAB_NNET || n < 0) return fail(SWPC_EINVAL, "bad argument");
    int c[LAB_NNET];
    if (int r = swpc_lab_count(tr, c)) return r;
    if (c[net] + n > tr->cap) return fail(SWPC_EINVAL, "trainer capacity %d exceeded", tr->cap);
    float *dst = (float *)tr->samples.p + ((size_t)net * tr->cap + c[net]) * LAB_REC;
    SWPC_HIP(hipMemcpy(dst, rec, (size_t)n * LAB_REC * sizeof(float), hipMemcpyHostToDevice));
    c[net] += n;
    tr->limit[net] = std::max(tr->limit[net], c[net]);
    SWPC_HIP(hipMemcpy(tr->count.p, c, sizeof c, hipMemcpyHostToDevice));
    return SWPC_OK;
}

int swpc_lab_samples(swpc_lab_trainer *tr, int net, int first, int n, float *rec)
{
    clear_error();
    if (!tr || !rec || net < 0 || net >= LAB_NNET || first < 0 || n < 0 || first + n > tr->cap)
        return fail(SWPC_EINVAL, "bad argument");
    SWPC_HIP(hipDeviceSynchronize());
    const float *src = (const float *)tr->samples.p + ((size_t)net * tr->cap + first) * LAB_REC;
    SWPC_HIP(hipMemcpy(rec, src, (size_t)n * LAB_REC * sizeof(float), hipMemcpyDeviceToHost));
    return SWPC_OK;
}

// Adam on net `net` of the sample store: params (FP32, LAB_NP) in and out;
// the learning rate follows a cosine from lr to lr/100 over `steps`;
// noise perturbs the previous-ghost features; loss[steps] (optional)
// receives the batch loss before each update.
int swpc_lab_train(swpc_lab_trainer *tr, int net, float *params, int steps, int batch, float lr, float noise,
                   unsigned seed, float *loss)
{
    clear_error();
    if (!tr || !params || net < 0 || net >= LAB_NNET || steps < 0 || batch < TB || batch % TB)
        return fail(SWPC_EINVAL, "bad argument (batch must be a positive multiple of %d)", TB);
    int c[LAB_NNET];
    if (int r = swpc_lab_count(tr, c)) return r;
    if (c[net] == 0) return fail(SWPC_ESTATE, "no samples for net %d", net);
    const int nblk = batch / TB;
    if (int r = tr_ensure(tr, nblk)) return r;
    if (int r = dalloc(tr->loss, (size_t)std::max(steps, 1) * sizeof(float), tr->bytes)) return r;
    hipStream_t s = tr->stream;
    SWPC_HIP(hipMemcpyAsync(tr->P.p, params, LAB_NP * sizeof(float), hipMemcpyHostToDevice, s));
    SWPC_HIP(hipMemsetAsync(tr->M.p, 0, LAB_NP * sizeof(float), s));
    SWPC_HIP(hipMemsetAsync(tr->V.p, 0, LAB_NP * sizeof(float), s));
    FwdBwdArgs a;
    a.P = (const float *)tr->P.p;
    a.samples = (const float *)tr->samples.p + (size_t)net * tr->cap * LAB_REC;
    a.count = c[net];
    a.nout = h_nout[net];
    a.mode = 0;
    a.noise = noise;
    a.scale = 2.0f / ((float)batch * (float)a.nout);
    a.partial = (float *)tr->partial.p;
    a.ploss = (float *)tr->ploss.p;
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    double p1 = 1.0, p2 = 1.0;
    for (int t = 0; t < steps; t++) {
        a.seed = lab_hash(seed * 0x9e3779b1u + (unsigned)t * 0x85ebca77u + 1u);
        hipLaunchKernelGGL(k_lab_fwdbwd, dim3(nblk), dim3(TB), 0, s, a);
        SWPC_CHECK_LAUNCH();
        p1 *= b1;
        p2 *= b2;
        const double frac = steps > 1 ? (double)t / (double)(steps - 1) : 0.0;
        const float lrt = (float)(lr * (0.01 + 0.99 * 0.5 * (1.0 + cos(M_PI * frac))));
        hipLaunchKernelGGL(k_lab_adam, dim3((LAB_NP + 255) / 256), dim3(256), 0, s, (float *)tr->P.p, (float *)tr->M.p,
                           (float *)tr->V.p, (const float *)tr->partial.p, (const float *)tr->ploss.p, nblk, lrt, b1,
                           b2, eps, (float)(1.0 / (1.0 - p1)), (float)(1.0 / (1.0 - p2)),
                           1.0f / ((float)batch * (float)a.nout), (float *)tr->loss.p + t, (float *)nullptr);
        SWPC_CHECK_LAUNCH();
    }
    SWPC_HIP(hipMemcpyAsync(params, tr->P.p, LAB_NP * sizeof(float), hipMemcpyDeviceToHost, s));
    if (loss && steps > 0)
        SWPC_HIP(hipMemcpyAsync(loss, tr->loss.p, (size_t)steps * sizeof(float), hipMemcpyDeviceToHost, s));
    SWPC_HIP(hipStreamSynchronize(s));
    return SWPC_OK;
}

// Loss and gradient of one net over n host records (no sign flips): the
// check of the training kernels.  nout selects the outputs in the loss.
int swpc_lab_grad(swpc_lab_trainer *tr, const float *params, int nout, int n, const float *rec, double *loss,
                  float *grad)
{
    clear_error();
    if (!tr || !params || !rec || !grad || n <= 0 || nout < 1 || nout > LAB_NOUT) return fail(SWPC_EINVAL, "bad argument");
    const int nblk = (n + TB - 1) / TB;
    if (int r = tr_ensure(tr, nblk)) return r;
    if (tr->X.bytes < (size_t)n * LAB_REC * sizeof(float))
        if (int r = dalloc(tr->X, (size_t)n * LAB_REC * sizeof(float), tr->bytes)) return r;
    if (!tr->loss.p)
        if (int r = dalloc(tr->loss, sizeof(float), tr->bytes)) return r;
    hipStream_t s = tr->stream;
    SWPC_HIP(hipMemcpyAsync(tr->P.p, params, LAB_NP * sizeof(float), hipMemcpyHostToDevice, s));
    SWPC_HIP(hipMemcpyAsync(tr->X.p, rec, (size_t)n * LAB_REC * sizeof(float), hipMemcpyHostToDevice, s));
    FwdBwdArgs a;
    a.P = (const float *)tr->P.p;
    a.samples = (const float *)tr->X.p;
    a.count = n;
    a.nout = nout;
    a.mode = 1;
    a.seed = 0;
    a.noise = 0.0f;
    a.scale = 2.0f / ((float)n * (float)nout);
    a.partial = (float *)tr->partial.p;
    a.ploss = (float *)tr->ploss.p;
    hipLaunchKernelGGL(k_lab_fwdbwd, dim3(nblk), dim3(TB), 0, s, a);
    SWPC_CHECK_LAUNCH();
    // M doubles as the gradient output buffer here
    hipLaunchKernelGGL(k_lab_adam, dim3((LAB_NP + 255) / 256), dim3(256), 0, s, (float *)tr->P.p, (float *)nullptr,
                       (float *)nullptr, (const float *)tr->partial.p, (const float *)tr->ploss.p, nblk, 0.0f, 0.0f,
                       0.0f, 0.0f, 0.0f, 0.0f, 1.0f / ((float)n * (float)nout), (float *)tr->loss.p,
                       (float *)tr->M.p);
    SWPC_CHECK_LAUNCH();
    float l;
    SWPC_HIP(hipMemcpyAsync(grad, tr->M.p, LAB_NP * sizeof(float), hipMemcpyDeviceToHost, s));
    SWPC_HIP(hipMemcpyAsync(&l, tr->loss.p, sizeof(float), hipMemcpyDeviceToHost, s));
    SWPC_HIP(hipStreamSynchronize(s));
    if (loss) *loss = l;
    return SWPC_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// The same loss and gradient with no device allocation: the parameters,
// the records, the per-workgroup partial sums and the outputs are staged
// in `scratch`, a file-scope device array that each GPU holds for the life
// of the process.  On the host the array's name is not a device address:
// its address on each device comes from hipGetSymbolAddress, taken once per
// device at that device's first call, and only that pointer is handed to
// the copies and kernels.  Every call on one GPU shares the array; calls
// that overlap on one GPU race.  Nothing in it needs to start at zero: the
// kernels write every partial sum and output they later read.
// ---------------------------------------------------------------------------
namespace swpc {

constexpr int NOALLOC_MAXBLK = SWPC_LAB_GRAD_NOALLOC_MAX / TB;
constexpr size_t NA_P = 0;                                   // parameters [LAB_NP]
constexpr size_t NA_G = NA_P + 3584;                         // gradient [LAB_NP]
constexpr size_t NA_LOSS = NA_G + 3584;                      // loss [1]
constexpr size_t NA_PLOSS = NA_LOSS + 64;                    // per-workgroup loss [MAXBLK]
constexpr size_t NA_PART = NA_PLOSS + NOALLOC_MAXBLK;        // partials [MAXBLK][LAB_NP]
constexpr size_t NA_X = NA_PART + (size_t)NOALLOC_MAXBLK * LAB_NP;   // records [MAX][LAB_REC]
constexpr size_t NA_TOTAL = NA_X + (size_t)SWPC_LAB_GRAD_NOALLOC_MAX * LAB_REC;
static_assert(LAB_NP <= 3584, "parameter block larger than its scratch slot");
static_assert(SWPC_LAB_GRAD_NOALLOC_MAX % TB == 0, "record limit must be a multiple of the workgroup size");
static_assert(NA_PART % 64 == 0 && NA_X % 64 == 0, "scratch slots keep 256-byte alignment");

static __device__ __align__(256) float scratch[NA_TOTAL];

constexpr int NOALLOC_MAXDEV = 64;
static std::mutex g_scratch_mu;
static float *g_scratch_ptr[NOALLOC_MAXDEV];

// The scratch array's device address on the current device (looked up once
// per device, with that device current).
static int scratch_ptr(float **out)
{
    int dev = -1;
    SWPC_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= NOALLOC_MAXDEV) return fail(SWPC_EINVAL, "device %d beyond the %d the library tracks", dev,
                                                      NOALLOC_MAXDEV);
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail(SWPC_EHIP, "hipGetSymbolAddress(scratch) failed on device %d: %s", dev, hipGetErrorString(e));
        g_scratch_ptr[dev] = (float *)p;
    }
    *out = g_scratch_ptr[dev];
    return SWPC_OK;
}

}  // namespace swpc

extern "C" {

int swpc_lab_grad_noalloc(const float *params, int nout, int n, const float *rec, double *loss, float *grad)
{
    clear_error();
    if (!params || !rec || !grad || n <= 0 || nout < 1 || nout > LAB_NOUT) return fail(SWPC_EINVAL, "bad argument");
    if (n > SWPC_LAB_GRAD_NOALLOC_MAX)
        return fail(SWPC_EINVAL, "swpc_lab_grad_noalloc: %d records exceed its fixed scratch of %d records "
                                 "(use swpc_lab_grad)", n, SWPC_LAB_GRAD_NOALLOC_MAX);
    float *base = nullptr;
    if (int r = scratch_ptr(&base)) return r;
    const int nblk = (n + TB - 1) / TB;
    hipStream_t s = nullptr;   // the device's null stream
    SWPC_HIP(hipMemcpyAsync(base + NA_P, params, LAB_NP * sizeof(float), hipMemcpyHostToDevice, s));
    SWPC_HIP(hipMemcpyAsync(base + NA_X, rec, (size_t)n * LAB_REC * sizeof(float), hipMemcpyHostToDevice, s));
    FwdBwdArgs a;
    a.P = base + NA_P;
    a.samples = base + NA_X;
    a.count = n;
    a.nout = nout;
    a.mode = 1;
    a.seed = 0;
    a.noise = 0.0f;
    a.scale = 2.0f / ((float)n * (float)nout);
    a.partial = base + NA_PART;
    a.ploss = base + NA_PLOSS;
    hipLaunchKernelGGL(k_lab_fwdbwd, dim3(nblk), dim3(TB), 0, s, a);
    SWPC_CHECK_LAUNCH();
    hipLaunchKernelGGL(k_lab_adam, dim3((LAB_NP + 255) / 256), dim3(256), 0, s, base + NA_P, (float *)nullptr,
                       (float *)nullptr, (const float *)(base + NA_PART), (const float *)(base + NA_PLOSS), nblk, 0.0f,
                       0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 1.0f / ((float)n * (float)nout), base + NA_LOSS, base + NA_G);
    SWPC_CHECK_LAUNCH();
    float l;
    SWPC_HIP(hipMemcpyAsync(grad, base + NA_G, LAB_NP * sizeof(float), hipMemcpyDeviceToHost, s));
    SWPC_HIP(hipMemcpyAsync(&l, base + NA_LOSS, sizeof(float), hipMemcpyDeviceToHost, s));
    SWPC_HIP(hipStreamSynchronize(s));
    if (loss) *loss = l;
    return SWPC_OK;
}

}  // extern "C"
