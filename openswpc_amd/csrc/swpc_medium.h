// medium__setup on the GPU (src/swpc_3d/m_medium.f90:36-467): from the
// velocity model's memory-box arrays rho, lam, mu, qp, qs to everything the
// time step reads, without a round trip through the host.
#pragma once

#include "swpc_common.h"

namespace swpc {

struct MedSetup {
    int nx, ny, nz, na, px, sxy;
    float *rho, *lam, *mu, *taup, *taus;      // device layout (at3); taup/taus hold qp/qs until med_relax
    float *bx, *by, *bz, *muyz, *muxz, *muxy;
    float zeta3;                              // NM * zeta, the FP32 product
    float cc_re, cc_im;                       // mean of i w ts/(1 - i w ts) at fq_ref (complex FP32)
};

// src in the reference's (k,i,j) memory-box order -> dst in the device
// layout, with the absorber homogenisation (m_medium.f90:124-190) applied
// as a clamp of each coordinate: i, j into [na+1, n-na], k to <= nz-na.
int med_ingest(const MedSetup &m, const float *src, float *dst, hipStream_t s);
// taup = zeta3/qp, taus = zeta3/qs, then the unrelaxed lam and mu at
// fq_ref (relaxed_medium, m_medium.f90:234-274), over the memory box.
int med_relax(const MedSetup &m, hipStream_t s);
// surface detection (m_medium.f90:342-397): kfs and kob maps over
// (ny+6) x (nx+6) (box index (j+2, i+2)), columns i, j in [0, n+2];
// surf[(j-1) nx + (i-1)] = (kfs_top, kfs_bot, kob_top, kob_bot) and
// kobc[(j-1) nx + (i-1)] = kob for i, j in [1, n].
int med_surface(const MedSetup &m, int *kfs, int *kob, int4 *surf, int *kobc, hipStream_t s);
// averaged_medium (m_medium.f90:432-465): bx, by, bz and the harmonic
// shear means over [1, n]^3, zero elsewhere in the box.
int med_average(const MedSetup &m, hipStream_t s);
// velocity_minmax (m_medium.f90:399-430) into out[0] = vmin, out[1] = vmax.
int med_minmax(const MedSetup &m, const int *kfs, float *out, hipStream_t s);
// Per plane k in [1, nz] of a: 1 when every cell i, j in [1, n] holds the
// bits of cell (1, 1, k) (and it is not NaN), else 0; tab[k + KO] = that
// value or NaN (NaN on the other planes).  flags has pz ints.
int med_planes(const MedSetup &m, const float *a, int *flags, float *tab, int pz, hipStream_t s);

}  // namespace swpc
