// Shared host/device helpers of the swpc HIP library: error reporting,
// device layout and host<->device transposition of the reference's
// Fortran-ordered arrays.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/swpc_hip.h"

namespace swpc {

int fail(int code, const char *fmt, ...);
void clear_error();

#define SWPC_HIP(x)                                                                                   \
    do {                                                                                              \
        hipError_t e_ = (x);                                                                          \
        if (e_ != hipSuccess)                                                                         \
            return ::swpc::fail(SWPC_EHIP, "%s failed: %s (%s:%d)", #x, hipGetErrorString(e_),        \
                                __FILE__, __LINE__);                                                  \
    } while (0)

#define SWPC_CHECK_LAUNCH()                                                                           \
    do {                                                                                              \
        hipError_t e_ = hipGetLastError();                                                            \
        if (e_ != hipSuccess)                                                                         \
            return ::swpc::fail(SWPC_EHIP, "kernel launch failed: %s (%s:%d)", hipGetErrorString(e_), \
                                __FILE__, __LINE__);                                                  \
    } while (0)

// Device layout.  Cells carry the reference's 1-based indices; x (i) is
// the fastest axis.  Cell i = 1 sits at element offset IO + 1 = 32 of its
// row so that rows of 64 interior cells start on a 128-byte boundary for
// FP32.  The memory box of the reference, [-2, n+3] per axis, is kept so
// that the ghost cells that replace the MPI halo stay zero in HBM.
constexpr int IO = 31;  // element offset of i = 0
constexpr int JO = 2;   // row offset of j = 0 (j = -2 is row 0)
constexpr int KO = 2;   // plane offset of k = 0

struct Lay3 {
    int nx, ny, nz;
    int px;       // elements per row
    int py;       // rows per plane (ny + 6)
    int pz;       // planes (nz + 6)
    int sxy;      // px * py
    long long vol;
};

inline Lay3 make_lay3(int nx, int ny, int nz)
{
    Lay3 L;
    L.nx = nx; L.ny = ny; L.nz = nz;
    int need = IO + nx + 4;                // i = nx + 3 is the last element
    // whole 128-byte lines per row for FP32 and for 16-bit arrays
    L.px = (need + 63) / 64 * 64;
    L.py = ny + 6;
    L.pz = nz + 6;
    L.sxy = L.px * L.py;
    L.vol = (long long)L.sxy * L.pz;
    return L;
}

__host__ __device__ inline int at3(int px, int sxy, int i, int j, int k)
{
    return (k + KO) * sxy + (j + JO) * px + (i + IO);
}

// Cell tiles of the 3-D state: arrays of per-cell values interleaved by
// tiles of RTILE cells, N components per tile.  Cell c (the at3 offset)
// lives in tile (c + RB) / RTILE at slot (c + RB) % RTILE; component q of
// tile t starts at element (t * N + q) * RTILE.  RB = 32 puts the first
// cell of every row (element IO + 1 = 32 mod 64) at a slot that is a
// multiple of 64, so rows start on 128-byte lines, and aligned groups of
// 2 or 4 cells never straddle two tiles.  64-bit: a tiled region passes
// 2^32 elements at 1024 x 1024 x 512.
constexpr int RB = 32;
constexpr int RTILE = 256;
constexpr int NS3 = 6;    // the 3-D stresses Sxx Syy Szz Syz Sxz Sxy, tiled together
constexpr int NR3 = 18;   // the 3-D memory variables, tiled together

template <int N>
__host__ __device__ __forceinline__ long long tix(int c)
{
    const long long t = (long long)((unsigned)(c + RB) >> 8);
    return t * (N * RTILE) + ((c + RB) & (RTILE - 1));
}

struct Lay2 {
    int nx, nz;
    int px;       // elements per row (x)
    int pz;       // rows (nz + 6)
    long long vol;
};

inline Lay2 make_lay2(int nx, int nz)
{
    Lay2 L;
    L.nx = nx; L.nz = nz;
    int need = IO + nx + 4;
    L.px = (need + 31) / 32 * 32;
    L.pz = nz + 6;
    L.vol = (long long)L.px * L.pz;
    return L;
}

__host__ __device__ inline int at2(int px, int i, int k) { return (k + KO) * px + (i + IO); }

// Device buffer owned by a context.
struct DBuf {
    void *p = nullptr;       // the buffer
    size_t bytes = 0;
    void *base = nullptr;    // the allocation (p may start past it)
};

// Device buffer of `bytes`, zeroed; b.p starts `lead` bytes past a
// 128-byte line boundary (plus the stagger of large buffers).
int dalloc(DBuf &b, size_t bytes, long long &total, size_t lead = 0);
void dfree(DBuf &b);

// Event-pair timing per kernel class.
struct Prof {
    bool on = false;
    std::vector<hipEvent_t> ev[5];
    void reset();
    ~Prof() { reset(); }
};

int prof_begin(Prof &p, int cls, hipStream_t s, hipEvent_t *out);
int prof_end(Prof &p, int cls, hipStream_t s, hipEvent_t start);
int prof_read(Prof &p, double *ms5, long long *count5);

}  // namespace swpc
