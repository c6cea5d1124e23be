// Learned absorbing boundary: shared definitions of the per-face ghost-fill
// networks (kernels in swpc_lab.hip, wired into the 3-D step in swpc3d.hip).
//
// On a grid without PML (na = 0) the interior update runs over every cell,
// and before each half-step the ghost cells beyond the five absorbing faces
// (x-, x+, y-, y+ and the bottom z+; the top is the free surface) are
// written by a small MLP per face and half-step.  Phase 0 runs before the
// stress update and writes the velocity ghosts that the 4th-order stress
// stencils read; phase 1 runs before the velocity update and writes the
// stress ghosts.  Net index = 2*face + phase.
//
// One column = the line of cells normal to a face through one face cell.
// Features (FP32, computed per column):
//   [0, 54)   Vx Vy Vz Sxx Syy Szz Syz Sxz Sxy at strip cells d = 1..6
//             (d = 1 outermost), index field*6 + d-1, stresses divided by
//             the impedance rho*vp of cell d = 1 (1 where lam+2mu = 0)
//   [54, 59)  the ghost values the same net wrote one step earlier (its
//             outputs, read back from the ghost cells), zero-padded
//   [59]      vs/vp of cell d = 1, [60] vp*dt/h (h the spacing normal to
//             the face)
// N = sqrt(sum of squares of [0, 59) / 59), or 0 when their largest
// magnitude is below 1e-30; features [0, 59) are divided by N.  The net y = Ws x + W3 tanh(W2 tanh(W1 x + b1) + b2) + b3 gives the
// ghost values y_q * N (times rho*vp for a stress).  N = 0 and air columns
// (lam + 2 mu = 0 at d = 1) write zeros.
#pragma once
#include "swpc_common.h"

namespace swpc {

constexpr int LAB_NS = 6;       // strip cells per column
constexpr int LAB_NF = 54;      // strip features (9 fields x 6 cells)
constexpr int LAB_NG = 5;       // previous-ghost features (max outputs)
constexpr int LAB_NSC = LAB_NF + LAB_NG;   // features scaled by 1/N
constexpr int LAB_NIN = LAB_NSC + 2;       // + vs/vp, vp*dt/h
constexpr int LAB_NH = 32;
constexpr int LAB_NOUT = 5;
constexpr int LAB_W1 = 0;
constexpr int LAB_B1 = LAB_W1 + LAB_NH * LAB_NIN;
constexpr int LAB_W2 = LAB_B1 + LAB_NH;
constexpr int LAB_B2 = LAB_W2 + LAB_NH * LAB_NH;
constexpr int LAB_W3 = LAB_B2 + LAB_NH;
constexpr int LAB_B3 = LAB_W3 + LAB_NOUT * LAB_NH;
constexpr int LAB_WS = LAB_B3 + LAB_NOUT;
constexpr int LAB_NP = LAB_WS + LAB_NOUT * LAB_NIN;   // parameters per net
constexpr int LAB_NNET = 10;
constexpr float LAB_SPONGE = 0.02f;  // default strip sponge strength
constexpr float LAB_TINY = 1e-30f;  // columns whose largest |feature| is below this count as empty
// Training sample record (floats): x[0, 61) | targets / N (/ rho*vp for a
// stress) [61, 66) | column id, N, rho*vp [66, 69) | zero pad to 72.
constexpr int LAB_REC = 72;
constexpr int LAB_T0 = LAB_NIN;

// The absorbing domain: nx x ny x nz cells whose cell (1,1,1) sits at cell
// (oi+1, oj+1, 1) of the context's grid (offset 0 for the ghost fill, the
// inner box of a wider grid for the training gather).
struct LabDomain {
    int oi = 0, oj = 0, nx = 0, ny = 0, nz = 0;
    int ncol[5] = {0, 0, 0, 0, 0};
    long long col0[6] = {0, 0, 0, 0, 0, 0};   // prefix sums of ncol
};

LabDomain lab_domain(int oi, int oj, int nx, int ny, int nz);

// What the lab kernels need of a 3-D context.
struct LabGrid {
    int px, sxy;
    void *F[9];             // Vx Vy Vz Sxx Syy Szz Syz Sxz Sxy (FP32, or FP64 with f64); the
                            // stresses in cell tiles of NS3 components (tix), F[3 + q] at slot q of tile 0
    int f64 = 0;
    const float *rho, *lam, *mu;
    float dt, h[3];
};

// Column medium (vs/vp, vp*dt/h, rho*vp) of every column of dom.
int lab_colmed(const LabGrid &g, const LabDomain &dom, DBuf &out, long long &bytes, hipStream_t s);
// Bytes of one net's image: its weights in the fill kernel's LDS form
// (W1 and W2 in BF16, zero-padded rows; the output layer and the hidden
// biases in FP32).
constexpr int LAB_IMG_BYTES = 9344;
// The ten nets' images from w[10][LAB_NP] (values already rounded to BF16).
void lab_fill_image(const float *w, unsigned char *img);
// Ghost fill of one phase; img holds the ten nets' images.  With sponge > 0 the phase's own fields (V at phase 0, S at
// phase 1) at strip cell d = 1..6 of every column are then multiplied by
// 1 - sponge ((7 - d)/6)^2 (lab_sponge_factors, formed in FP64 and
// rounded to FP32).
// With a strip cache (cache: the state's type, [9][LAB_NS][dom.col0[2]], the
// x faces' columns),
// the phase stores its own fields' strips there as they stand after the
// sponge, and with use_cache reads the other phase's strips from it
// instead of the fields (valid while nothing but the half-step between
// the phases has run since the other phase stored them).
int lab_fill(const LabGrid &g, const LabDomain &dom, const void *img, const float4 *colmed, int phase, float sponge,
             void *cache, bool use_cache, hipStream_t s);
void lab_sponge_factors(float sponge, float *sp);

}  // namespace swpc

// Trainer: sample store + optimiser state (opaque in the C ABI).
struct swpc_lab_trainer;
namespace swpc {
// Start a wide run: gather domain, its column medium, zero previous ghosts
// and running maxima, per-net sample quota for the run.
int lab_begin(swpc_lab_trainer *tr, const LabGrid &g, const LabDomain &dom, const int *quota, hipStream_t s);
// Append the training samples of one phase of the wide context's current
// fields (see swpc3d_lab_gather).
int lab_gather(swpc_lab_trainer *tr, const LabGrid &g, int phase, float keep, unsigned seed, float thr, hipStream_t s);
}  // namespace swpc
