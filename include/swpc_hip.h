/*
 * swpc_hip.h -- C ABI of the MI355X-native OpenSWPC time-stepping core.
 *
 * The reference drives its hot path through parameterless Fortran module
 * subroutines that read and write module-global arrays.  Each entry point
 * below replaces one of them; the Fortran host reaches them through the
 * bind(C) interfaces in include/swpc_hip.f90.
 *
 * Conventions
 *   - Every call returns 0 on success and a negative code on error;
 *     swpc_last_error() returns the text of the calling thread's last
 *     error.  No exception crosses the ABI.
 *   - Host arrays use the reference's Fortran layout: 3-D fields are
 *     (k,i,j) with k fastest over the memory box [-2:n+3] in each axis
 *     (src/swpc_3d/m_global.f90:297-302); P-SV fields are (k,i) over
 *     [-2:nz+3] x [-2:nx+3] (src/swpc_psv/m_global.f90).  The library
 *     copies them into its own x-fastest device layout; the caller keeps
 *     ownership of host memory.
 *   - Indices (isrc, ist, kfs, ...) are the reference's 1-based cell
 *     numbers.
 *   - All work of one context is ordered on one HIP stream.  Only the
 *     get/vmax/sync calls wait for it.  A context is not thread-safe;
 *     several contexts may coexist.
 */
#ifndef SWPC_HIP_H
#define SWPC_HIP_H

#ifdef __cplusplus
extern "C" {
#endif

enum swpc_precision {
    SWPC_FP32 = 0,        /* state, memory variables and arithmetic in FP32 (reference built with MP = SP) */
    SWPC_FP64 = 1,        /* the same scheme in FP64 */
    SWPC_FP32_S16 = 2,    /* FP32 arithmetic and V/S; memory variables, ADE arrays (BF16) and medium stored in 16 bit */
    SWPC_FP32_R16 = 3     /* FP32 arithmetic, V/S, medium and ADE arrays; memory variables in 16 bit (FP16, scaled) */
};

enum swpc_error {
    SWPC_OK = 0,
    SWPC_EINVAL = -1,     /* bad argument */
    SWPC_EHIP = -2,       /* HIP runtime error */
    SWPC_ESTATE = -3,     /* call out of order (e.g. step before set_medium) */
    SWPC_ENOMEM = -4
};

/* Field identifiers for get_field / set_field. */
enum swpc3d_field {
    SWPC3D_VX = 0, SWPC3D_VY, SWPC3D_VZ,
    SWPC3D_SXX, SWPC3D_SYY, SWPC3D_SZZ, SWPC3D_SYZ, SWPC3D_SXZ, SWPC3D_SXY,
    /* memory variables, component-major: R{xx,yy,zz,yz,xz,xy} x m = 1..3 */
    SWPC3D_R0 = 9,
    /* ADE arrays of the PML, in the order axVx ayVx azVx axVy ayVy azVy
       axVz ayVz azVz axSxx aySxy azSxz axSxy aySyy azSyz axSxz aySyz azSzz */
    SWPC3D_ADE0 = 27,
    SWPC3D_NFIELD = 45
};

/* Snapshot products of snap__write (src/swpc_3d/m_snap.f90:911-2185;
   P-SV src/swpc_psv/m_snap.f90:419-650). */
enum swpc_snap_kind {
    SWPC_SNAP_PS = 0,     /* div, rot_x, rot_y, rot_z (3-D) or div, rot (P-SV), x UC*M0*1e-3 */
    SWPC_SNAP_V = 1,      /* velocity x UC*M0 */
    SWPC_SNAP_U = 2       /* running sum of velocity x UC*M0*dt, advanced every step */
};

enum swpc2d_field {
    SWPC2D_VX = 0, SWPC2D_VZ, SWPC2D_SXX, SWPC2D_SZZ, SWPC2D_SXZ,
    SWPC2D_R0 = 5,        /* R{xx,zz,xz} x m = 1..3 */
    SWPC2D_ADE0 = 14,     /* axVx azVx axVz azVz axSxx azSxz axSxz azSzz */
    SWPC2D_NFIELD = 22
};

enum swpcsh_field {
    SWPCSH_VY = 0, SWPCSH_SYZ, SWPCSH_SXY,
    SWPCSH_R0 = 3,        /* Ryz x m = 1..3, then Rxy x m = 1..3 */
    SWPCSH_ADE0 = 9,      /* axVy azVy axSxy azSyz */
    SWPCSH_NFIELD = 13
};

const char *swpc_last_error(void);
int swpc_version(void);
/* Number of HIP devices visible (0 when no GPU); never fails. */
int swpc_device_count(void);
/* Measured device copy bandwidth: `iters` copies of `bytes` bytes with a
   16-byte non-temporal copy kernel; *gbs = (read + write bytes) / time in GB/s.
   A diagnostic with no counterpart in the reference: it is the measured
   ceiling the stencil kernels are reported against. */
int swpc_membw(long long bytes, int iters, double *gbs);

/* Random-media synthesis: replaces rmedia__3dgen_ani / rmedia__2dgen_ani
   (src/shared/m_rmedia.f90:40-173; its PSDFs m_rmedia.f90:548-588 in 2-D,
   682-732 in 3-D).
   nd = 2 or 3; shape[nd] with the LAST axis fastest (C order); d[nd] grid
   widths and a[nd] correlation lengths (km); rms epsil; kappa (von Karman
   only); ptype 0 von Karman, 1 Gaussian, 2 exponential.  noise[N] is the
   uniform noise the realisation is made of (the caller's random stream, as
   the reference draws it inside the routine); out[N] receives the field:
   FFT(noise) whitened to unit modulus, times sqrt(PSDF(k) N prod(d)), DC
   zeroed, inverse FFT / N, demeaned and scaled to rms epsil.  Transforms in
   FP32 (hipFFT), PSDF and mean/rms in FP64.  Host buffers; N < 2^31;
   allocates its own device memory and stream and synchronises before
   returning. */
int swpc_rmedia_generate(int nd, const int *shape, const double *d, const double *a, double epsil, double kappa,
                         int ptype, const float *noise, float *out);

/* ------------------------------------------------------------------ */
/* 3-D (swpc_3d)                                                       */
/* ------------------------------------------------------------------ */

typedef struct swpc3d_ctx swpc3d_ctx;

typedef struct {
    int nx, ny, nz;       /* grid cells (global; one GPU holds the domain) */
    int na;               /* absorber thickness in cells (m_global.f90:166) */
    int nm;               /* Zener bodies; the reference fixes NM = 3 (m_global.f90:31) */
    int precision;        /* enum swpc_precision */
    double dx, dy, dz;    /* grid widths [km] */
    double dt;            /* time step [s] */
    float ts[8];          /* relaxation times ts(1:nm) (m_fdtool.f90:691-727) */
} swpc3d_desc;

/* Replaces kernel__setup + memory_allocate (m_kernel.f90:30-73, 354-378)
   and the absorber allocation of absorb_p__setup (m_absorb_p.f90:95-116). */
int swpc3d_create(const swpc3d_desc *desc, swpc3d_ctx **out);
int swpc3d_destroy(swpc3d_ctx *ctx);

/* Medium arrays produced by medium__setup (m_medium.f90:36-467), all
   real(SP) over the memory box (m_global.f90:39-42); replaces the
   `!$acc enter data copyin` of main.f90:97-109. */
int swpc3d_set_medium(swpc3d_ctx *ctx, const float *bx, const float *by, const float *bz,
                      const float *lam, const float *mu, const float *taup, const float *taus,
                      const float *muyz, const float *muxz, const float *muxy);

/* Density rho over the memory box (the reference's real(SP) rho,
   m_global.f90:39).  Optional: with it the FP32 interior velocity kernel
   forms bx, by, bz = 2/(rho + rho') itself (m_medium.f90:442-444) from
   one density stream instead of reading three buoyancy arrays; results are
   identical.  Ignored in SWPC_FP32_S16 mode. */
int swpc3d_set_density(swpc3d_ctx *ctx, const float *rho);
/* medium__setup (m_medium.f90:36-467) on the GPU, in place of the host
   assembly followed by set_medium, set_density and set_surface.  In: the
   velocity model's memory-box arrays rho, lam, mu, qp, qs (reference
   (k,i,j) order, as vmodel_* leaves them).  The library applies the
   absorber homogenisation (124-190), taup = zeta3/qp and taus = zeta3/qs
   with zeta3 = NM*zeta (the FP32 product; 0 in benchmark mode), the
   unrelaxed moduli at fq_ref through cc = (cc_re, cc_im), the mean of
   i w ts/(1 - i w ts) (relaxed_medium, 234-274), surface detection and
   its 6 x 6 window extremes (342-397), buoyancy and the harmonic shear
   means (432-465), and velocity_minmax (399-430).  Out (each optional):
   kfs and kob over the (ny+6) x (nx+6) columns of the memory box (box
   index (j+2, i+2); the receivers are placed from them) and vminmax[2] =
   (vmin, vmax).  Results equal the host assembly's except the
   relaxation factor, formed in FP64 and rounded once (FP32 ulps).  Not
   for SWPC_FP32_S16. */
int swpc3d_setup_medium(swpc3d_ctx *ctx, const float *rho, const float *lam, const float *mu, const float *qp,
                        const float *qs, float zeta3, float cc_re, float cc_im, int *kfs, int *kob, float *vminmax);
/* Read back one medium array over the memory box in the reference's
   (k,i,j) order: which = 0..9 for bx by bz lam mu taup taus muyz muxz
   muxy (the set_medium order), 10 for rho.  Not for SWPC_FP32_S16. */
int swpc3d_get_medium(swpc3d_ctx *ctx, int which, float *out);

/* Order-switch maps kfs_top/bot, kob_top/bot and kob over the (i,j)
   memory box (m_medium.f90:342-397; used by m_kernel.f90:103-104 and
   kernel__vmax m_kernel.f90:345). */
int swpc3d_set_surface(swpc3d_ctx *ctx, const int *kfs_top, const int *kfs_bot,
                       const int *kob_top, const int *kob_bot, const int *kob);

/* Damping profiles g(4, 1:n) of absorb_p__setup (m_absorb_p.f90:75-93). */
int swpc3d_set_pml(swpc3d_ctx *ctx, const float *gxc, const float *gxe, const float *gyc,
                   const float *gye, const float *gzc, const float *gze);

/* Stress-glut sources (source__setup m_source.f90:41-313): m6 =
   (mxx,myy,mzz,myz,mxz,mxy) per source; sdrop[it-1][n] =
   mo*momentrate(tbeg+(it-1/2)dt)*dt/(dx dy dz) for it = 1..nt, evaluated
   on the host in FP64 (replaces the per-step evaluation at
   m_source.f90:800-801). */
int swpc3d_set_sources(swpc3d_ctx *ctx, int nsrc, const int *isrc, const int *jsrc, const int *ksrc,
                       const double *m6, int nt, const double *sdrop);

/* Body-force sources (m_source.f90:835-876): f3 = (fx,fy,fz)/M0 per
   source; fdt[it-1][n] = momentrate(tbeg+it*dt)*dt/(dx dy dz)/2. */
int swpc3d_set_bodyforce(swpc3d_ctx *ctx, int nsrc, const int *isrc, const int *jsrc, const int *ksrc,
                         const double *f3, int nt, const double *fdt);

/* Receivers of wav__setup/set_stinfo (m_wav.f90:54-274): 1-based cell
   indices, decimation ntdec_w, ntw samples. */
int swpc3d_set_receivers(swpc3d_ctx *ctx, int nst, const int *ist, const int *jst, const int *kst,
                         int ntdec_w, int ntw);

/* kernel__update_stress + absorb__update_stress (m_kernel.f90:142-326,
   m_absorb_p.f90:313-505). */
int swpc3d_update_stress(swpc3d_ctx *ctx);
/* kernel__update_vel + absorb__update_vel (m_kernel.f90:75-140,
   m_absorb_p.f90:126-311). */
int swpc3d_update_vel(swpc3d_ctx *ctx);
/* source__stressglut(it) (m_source.f90:775-833). */
int swpc3d_stressglut(swpc3d_ctx *ctx, int it);
/* source__bodyforce(it) (m_source.f90:835-876). */
int swpc3d_bodyforce(swpc3d_ctx *ctx, int it);
/* wav__store(it) for velocity and displacement (m_wav.f90:400-628);
   displacement is accumulated in FP64. */
int swpc3d_store_receivers(swpc3d_ctx *ctx, int it);
/* A snapshot stream of snap__setup/snap__write (m_snap.f90:97-323,
   911-2185) over an explicit list of cells: ijk[3*npts] holds 1-based
   (i,j,k) per output point, with 1 <= i <= nx, 1 <= j <= ny,
   -1 <= k <= nz+2 (the decimated yz/xz/xy section, or the kfs+1 / kob+1
   surfaces for fs/ob).  A frame of nvar x npts floats (nvar = 4 for 'ps',
   3 for 'v' and 'u', variable-major) is produced at every it with
   mod(it-1, ntdec_s) = 0 and kept in a device ring of nbuf frames;
   'u' accumulates at every step.  uc, m0 and dt are the reference's
   real(SP) UC, M0 and dt.  *id receives the stream's handle. */
int swpc3d_add_snapshot(swpc3d_ctx *ctx, int kind, int npts, const int *ijk, int ntdec_s, int nbuf,
                        float uc, float m0, float dt, int *id);
/* snap__write(it) alone (step() already calls it). */
int swpc3d_snapshots(swpc3d_ctx *ctx, int it);
/* Copy frame (0-based: it = frame*ntdec_s + 1) of stream id to out
   (synchronises); fails if the frame is not yet written or has been
   overwritten in the ring. */
int swpc3d_get_snapshot(swpc3d_ctx *ctx, int id, int frame, float *out);
/* Keep the running maximum amplitudes of a 'v' or 'u' stream from the
   next step on: the max_fs_v / max_ob_v / max_fs_u / max_ob_u arrays of
   wbuf_fs_v and its siblings (m_snap.f90:1690-1703), written to the
   netCDF file as max-V, max-H, max-A by output__put_maxval
   (m_snap.f90:2291-2344). */
int swpc3d_snapshot_peaks(swpc3d_ctx *ctx, int id);
/* The closing update of snap__closefiles in netCDF mode
   (m_snap.f90:2215-2231, wbuf_*(nt+1)): advances the 'u' running sums
   and the maxima at step number it, stores no frame. */
int swpc3d_finish_snapshots(swpc3d_ctx *ctx, int it);
/* Copy the 3 x npts maxima (max-V, max-H, max-A, variable-major) of
   stream id to out (synchronises). */
int swpc3d_get_peaks(swpc3d_ctx *ctx, int id, float *out);
/* One time step it in the order of main.f90:125-144: receivers,
   snapshots, stress (interior + PML), stress glut, velocity (interior +
   PML), body force.  The halo exchange is absent: ghost cells stay in
   HBM. */
int swpc3d_step(swpc3d_ctx *ctx, int it);
/* Steps it0 .. it0+nsteps-1, enqueued without host synchronisation. */
int swpc3d_run(swpc3d_ctx *ctx, int it0, int nsteps);

/* Gathers (device -> host; these synchronise the context's stream).
   vel[itw][n][3] are two-node averaged velocities in solver units
   (z positive up), disp[itw][n][3] the FP64 displacement; either may be
   NULL.  Multiply by M0*UC*1e9 for the reference's output units. */
int swpc3d_get_receivers(swpc3d_ctx *ctx, float *vel, double *disp);
int swpc3d_get_field(swpc3d_ctx *ctx, int field, double *out);
int swpc3d_set_field(swpc3d_ctx *ctx, int field, const double *in);
/* kernel__vmax (m_kernel.f90:328-352). */
int swpc3d_vmax(swpc3d_ctx *ctx, float *out3);
int swpc3d_sync(swpc3d_ctx *ctx);

/* Per-kernel timing with HIP events recorded on the context's stream.
   enable != 0 starts collection (and clears it); read returns, for each
   kernel class (0 stress interior, 1 stress PML, 2 vel interior, 3 vel
   PML, 4 other), the summed milliseconds and launch counts. */
int swpc3d_profile(swpc3d_ctx *ctx, int enable);
int swpc3d_profile_read(swpc3d_ctx *ctx, double *ms5, long long *count5);
/* Device bytes held by the context. */
int swpc3d_memory(swpc3d_ctx *ctx, long long *bytes);
/* Launch configuration of the kernels (no counterpart in the reference;
   results do not depend on it beyond FP32 rounding order).  p[0..n-1]
   overrides, in order:
     [0..7]  one-cell-per-thread kernels (FP64 always uses them): stress
             tile rows (2/4/8), stress waves-per-SIMD hint (0/3/4), split
             normal/shear stress (0/1), velocity tile rows, velocity waves
             hint (0/5), workgroup order [5], stress z-chunk, velocity
             z-chunk (0 = automatic).  [5]: 0 as launched, 1 XCD-contiguous
             (one-cell and 4-wide kernels), G = 2..64 the z-marching
             4-wide velocity update's y tiles in groups of G, group g on
             XCD g mod 8 (default 2: 593 against 603 us at 512 x 512 x 256;
             the other kernels as launched);
     [8]     4-cell-per-thread kernels for FP32 and reduced storage (0/1,
             default 1); [9] their tile rows (1/2/4, default 2);
     [10]    4-cell-per-thread PML kernels (0/1, default 0);
     [11]    4-wide stress form: 0 normal + shear launches, 1 one pass with
             staged loads, 2 one pass (default 1; 0 for the 16-bit
             medium); [12] tile rows of the
             z-marching velocity kernel (1/2/4/8, default 4; 2 for the
             16-bit medium);
     [13]    PML kernels on a second stream, concurrent with the interior
             (0/1, default 0);
     [14]    4-wide velocity z-marching chunk (0 = one plane per workgroup;
             default 8 for FP32, 16 for 16-bit storage);
     [15]    4-wide one-pass z-marching stress chunk (0 = off, default);
     [16]    steps per captured HIP graph in swpc3d_run (0 = launch step
             by step; default 16).  The step number lives in device memory,
             so one captured sequence replays for any it; setters drop the
             graph, and profiling runs step by step;
     [17]    reduced storage: stochastic rounding of the 16-bit memory
             variable and ADE stores (0/1, default 0), hashed from cell,
             array and value (deterministic);
     [18]    reduced storage: log2 shift of the memory-variable scale
             (-24..24, default 0; negative stores larger values, away from
             the FP16 subnormals, at the cost of headroom before the
             +-65504 clamp).  Stored values are rescaled with it;
     [19]    absorber boxes of whole rows (the y faces and the bottom)
             walked one 256-cell row segment per workgroup (0/1, default 1;
             0 walks every box cell by cell);
     [20]    PML workgroups ordered so that each XCD takes one contiguous
             run of them (0/1, default 0);
     [21]    z-marching velocity with the y-stencil rows of Syy, Sxy and
             Syz staged in LDS per plane (0/1, default 0; tile rows 2, 4
             or 8 from entry [12]);
     [22]    the 4-wide one-pass stress (bit 0) and z-marching velocity
             (bit 1) kernels take the x neighbours of a thread's four
             cells from the adjacent lanes by DPP wavefront shifts instead
             of loads (default 2: the velocity update, 1.6 % faster; the
             stress update is 0.4 % slower with them, 2 % with reduced
             storage);
     [23]    learned boundary: strip cache of the ghost fill (0/1, default
             1).  Each phase stores its own fields' strips, as they stand
             after its sponge, in a compact array, and the other phase
             reads them from there instead of the fields (the half-step
             between them leaves those fields alone); dropped whenever a
             field is set or a single update is launched outside
             swpc3d_step/swpc3d_run.  Results are the same bit for bit. */
int swpc3d_set_tuning(swpc3d_ctx *ctx, const int *p, int n);

/* ------------------------------------------------------------------ */
/* Learned absorbing boundary (3-D)                                    */
/* ------------------------------------------------------------------ */
/* No counterpart in the reference: a thin absorbing boundary that takes
   the place of the ADE-PML of absorb_p__update_stress/vel
   (m_absorb_p.f90:126-311, 313-505) on a grid created with na = 0.  The
   interior update then covers every cell, and before each half-step the
   ghost cells beyond the x-, x+, y-, y+ and bottom faces are written by a
   small MLP per face and half-step (ten nets, 2*face + phase, phase 0
   before the stress update, 1 before the velocity update) from the six
   outermost cells of each column and the net's own previous ghost values;
   columns whose outermost cell is air (lam + 2 mu = 0) keep zero ghosts.
   Each net has swpc_lab_nparam() FP32 parameters: W1[32][61] b1[32]
   W2[32][32] b2[32] W3[5][32] b3[5] Ws[5][61]; the time step uses them
   rounded to BF16 (nearest even).  Its two hidden layers run on the
   matrix cores with BF16 inputs (features, layer-1 activations) and FP32
   accumulation; the output layer Ws x + W3 h2 + b3 is FP32.  After the
   ghosts of a phase are written, a strip sponge multiplies that phase's
   own fields (velocities at phase 0, stresses at phase 1) at strip cell
   d = 1..6 (d = 1 outermost) by 1 - s ((7 - d)/6)^2, the factor formed in
   FP64 and rounded to FP32.  The nets alone are not a stable closed loop
   over thousands of steps (a mode confined to the strip grows); the
   sponge, default s = 0.02, removes it (the 64 x 64 x 48 test grid grows
   without bound at s <= 0.005 and decays at s >= 0.01). */
int swpc_lab_nparam(void);
/* Switch the learned boundary on with params[10][swpc_lab_nparam()], or
   off with params = NULL.  Needs na = 0, an FP32 medium (any precision
   but SWPC_FP32_S16) and set_medium + set_density first. */
int swpc3d_set_lab(swpc3d_ctx *ctx, const float *params);
/* Strip sponge strength s of the learned boundary (0 <= s < 1; 0 = none;
   default 0.02). */
int swpc3d_set_lab_sponge(swpc3d_ctx *ctx, float sponge);

/* Training (no counterpart in the reference).  A trainer holds up to cap
   samples per net on the device.  Samples come from wide-PML runs: the
   domain of nxs x nys x nzs cells at offset (oi, oj, 0) of a wider grid
   stands for a grid with a learned boundary, and the wide run's true
   fields in its ghost cells are the targets.  swpc_lab_begin_run starts
   such a run (quota[10] more samples per net, NULL: up to capacity);
   swpc3d_lab_gather must then be called at every step before
   swpc3d_update_stress (phase 0) and before swpc3d_update_vel (phase 1).
   A column enters with probability keep (hashed from seed and the column)
   once its scale N reaches thr times the largest N seen so far in the
   run.  A sample is 72 floats: features[61] | targets[5] | column, N,
   rho*vp | pad. */
typedef struct swpc_lab_trainer swpc_lab_trainer;
int swpc_lab_trainer_create(int cap, swpc_lab_trainer **out);
int swpc_lab_trainer_destroy(swpc_lab_trainer *tr);
int swpc_lab_begin_run(swpc_lab_trainer *tr, swpc3d_ctx *wide, int oi, int oj, int nxs, int nys, int nzs,
                       const int *quota);
int swpc3d_lab_gather(swpc3d_ctx *wide, swpc_lab_trainer *tr, int phase, float keep, unsigned seed, float thr);
/* counts[10]: samples held per net. */
int swpc_lab_count(swpc_lab_trainer *tr, int *counts);
/* Append n host samples to net `net`; copy samples first..first+n-1 out. */
int swpc_lab_add(swpc_lab_trainer *tr, int net, int n, const float *rec);
int swpc_lab_samples(swpc_lab_trainer *tr, int net, int first, int n, float *rec);
/* Adam (0.9, 0.999, 1e-8) on net `net` for `steps` batches of `batch`
   samples (a multiple of 128) drawn at random from the store, half of them
   with the sign of every field feature and target flipped, and Gaussian
   noise of standard deviation `noise` added to the non-zero
   previous-ghost features (robustness of the closed loop); the learning
   rate follows a cosine from lr to lr/100.  params (FP32) in and out;
   loss[steps] (or NULL) receives each batch's mean square error. */
int swpc_lab_train(swpc_lab_trainer *tr, int net, float *params, int steps, int batch, float lr, float noise,
                   unsigned seed, float *loss);
/* Mean square error over the first nout outputs of n host samples and its
   gradient grad[swpc_lab_nparam()], computed by the training kernels. */
int swpc_lab_grad(swpc_lab_trainer *tr, const float *params, int nout, int n, const float *rec, double *loss,
                  float *grad);
/* No counterpart in the reference.  For callers that cannot allocate
   device memory (code under stream capture, tight serving or training
   loops, harnesses that forbid hipMalloc): the loss and gradient of
   swpc_lab_grad over n <= SWPC_LAB_GRAD_NOALLOC_MAX host records with no
   trainer and no allocation.  Its staging and partial sums live in a
   fixed device array of about 6.5 MB that each GPU holds until the process
   ends; larger n is rejected with SWPC_EINVAL.  Work runs on the current
   device's null stream and the call returns when it is done; calls that
   overlap on one GPU share the array and race. */
#define SWPC_LAB_GRAD_NOALLOC_MAX 16384
int swpc_lab_grad_noalloc(const float *params, int nout, int n, const float *rec, double *loss, float *grad);

/* ------------------------------------------------------------------ */
/* P-SV (swpc_psv)                                                     */
/* ------------------------------------------------------------------ */

typedef struct swpc2d_ctx swpc2d_ctx;

typedef struct {
    int nx, nz;
    int na;
    int nm;
    int precision;
    double dx, dz;
    double dt;
    float ts[8];
} swpc2d_desc;

int swpc2d_create(const swpc2d_desc *desc, swpc2d_ctx **out);
int swpc2d_destroy(swpc2d_ctx *ctx);
/* rho, lam, mu, taup, taus over the (k,i) memory box; the P-SV kernel
   forms buoyancy and the harmonic mu_xz itself (src/swpc_psv/m_kernel.f90:118-119, 270-274). */
int swpc2d_set_medium(swpc2d_ctx *ctx, const float *rho, const float *lam, const float *mu,
                      const float *taup, const float *taus);
/* maps over i in [-2:nx+3] */
int swpc2d_set_surface(swpc2d_ctx *ctx, const int *kfs_top, const int *kfs_bot,
                       const int *kob_top, const int *kob_bot, const int *kob);
int swpc2d_set_pml(swpc2d_ctx *ctx, const float *gxc, const float *gxe, const float *gzc, const float *gze);
/* m3 = (mxx, mzz, mxz); sdrop[it-1][n] = mo*momentrate(tbeg+(it-1/2)dt)*dt/(dx dz) */
int swpc2d_set_sources(swpc2d_ctx *ctx, int nsrc, const int *isrc, const int *ksrc,
                       const double *m3, int nt, const double *sdrop);
/* f2 = (fx, fz)/M0; fdt[it-1][n] = momentrate(tbeg+it*dt)*dt/(dx dz)/2 */
int swpc2d_set_bodyforce(swpc2d_ctx *ctx, int nsrc, const int *isrc, const int *ksrc,
                         const double *f2, int nt, const double *fdt);
int swpc2d_set_receivers(swpc2d_ctx *ctx, int nst, const int *ist, const int *kst, int ntdec_w, int ntw);
int swpc2d_update_stress(swpc2d_ctx *ctx);
int swpc2d_update_vel(swpc2d_ctx *ctx);
int swpc2d_stressglut(swpc2d_ctx *ctx, int it);
int swpc2d_bodyforce(swpc2d_ctx *ctx, int it);
int swpc2d_store_receivers(swpc2d_ctx *ctx, int it);
/* one step in the order of src/swpc_psv/main.f90:95-114 */
/* P-SV snapshot streams (src/swpc_psv/m_snap.f90:80-160, 419-650) over
   1-based (i,k) cells ik[2*npts]; frames of 2 x npts floats (div, rot /
   Vx, Vz / Ux, Uz).  Same conventions as swpc3d_add_snapshot. */
int swpc2d_add_snapshot(swpc2d_ctx *ctx, int kind, int npts, const int *ik, int ntdec_s, int nbuf,
                        float uc, float m0, float dt, int *id);
int swpc2d_snapshots(swpc2d_ctx *ctx, int it);
int swpc2d_get_snapshot(swpc2d_ctx *ctx, int id, int frame, float *out);
int swpc2d_step(swpc2d_ctx *ctx, int it);
/* Steps per captured HIP graph in swpc2d_run (0 launches step by step;
   default 16).  No counterpart in the reference: the per-step work of
   small P-SV grids is shorter than its launch overhead. */
int swpc2d_set_graph(swpc2d_ctx *ctx, int steps);
/* Steps of swpc2d_step/swpc2d_run in two fused kernels (on != 0, the
   default) or in the separate kernels of the per-routine entry points
   (0).  No counterpart in the reference; both give the same results bit
   for bit.  The fused kernels also take the receivers, and the stress glut
   and body force when there are at most 8 of them with their cells and
   the cells' -x and -z neighbours inside [1, nx] x [1, nz]. */
int swpc2d_set_fusion(swpc2d_ctx *ctx, int on);
int swpc2d_run(swpc2d_ctx *ctx, int it0, int nsteps);
/* vel[itw][n][2], disp[itw][n][2] (x, z-up) */
int swpc2d_get_receivers(swpc2d_ctx *ctx, float *vel, double *disp);
int swpc2d_get_field(swpc2d_ctx *ctx, int field, double *out);
int swpc2d_set_field(swpc2d_ctx *ctx, int field, const double *in);
int swpc2d_vmax(swpc2d_ctx *ctx, float *out2);
int swpc2d_sync(swpc2d_ctx *ctx);

/* ------------------------------------------------------------------ */
/* SH (swpc_sh)                                                        */
/* ------------------------------------------------------------------ */
/* The 2-D SH core (src/swpc_sh): Vy, Syz, Sxy with NM = 3 memory
   variables each, the ADE-CFS PML, stress glut and body force, receivers.
   Same descriptor, layouts and conventions as the P-SV context.  Step
   order of src/swpc_sh/main.f90:97-114.  The SH PML velocity update of
   the reference indexes the z damping profile by i (m_absorb_p.f90:
   181-182), which reads past the profile when nx > nz; this core takes it
   at k. */
typedef struct swpcsh_ctx swpcsh_ctx;

int swpcsh_create(const swpc2d_desc *desc, swpcsh_ctx **out);
int swpcsh_destroy(swpcsh_ctx *ctx);
/* rho, mu, taus over the (k,i) memory box (after medium__setup); the
   kernels form 1/rho and the two-point harmonic mu_xy, mu_yz themselves
   (src/swpc_sh/m_kernel.f90:127, 206-209). */
int swpcsh_set_medium(swpcsh_ctx *ctx, const float *rho, const float *mu, const float *taus);
/* maps over i in [-2:nx+3] */
int swpcsh_set_surface(swpcsh_ctx *ctx, const int *kfs_top, const int *kfs_bot,
                       const int *kob_top, const int *kob_bot, const int *kob);
/* profiles g(1..4) per cell (pa = 2 in the SH damping profile,
   src/swpc_sh/m_absorb_p.f90:393-425) */
int swpcsh_set_pml(swpcsh_ctx *ctx, const float *gxc, const float *gxe, const float *gzc, const float *gze);
/* m2 = (myz, mxy); sdrop[it-1][n] = mo/M0 * momentrate(tbeg+(it-1/2)dt) * dt/(dx dz)
   (m_source.f90:554-594: each component split over two nodes) */
int swpcsh_set_sources(swpcsh_ctx *ctx, int nsrc, const int *isrc, const int *ksrc,
                       const double *m2, int nt, const double *sdrop);
/* fy/M0; fdt[it-1][n] = momentrate(tbeg+it dt) * dt/(dx dz) (m_source.f90:596-629) */
int swpcsh_set_bodyforce(swpcsh_ctx *ctx, int nsrc, const int *isrc, const int *ksrc,
                         const double *fy, int nt, const double *fdt);
int swpcsh_set_receivers(swpcsh_ctx *ctx, int nst, const int *ist, const int *kst, int ntdec_w, int ntw);
int swpcsh_update_stress(swpcsh_ctx *ctx);
int swpcsh_update_vel(swpcsh_ctx *ctx);
int swpcsh_stressglut(swpcsh_ctx *ctx, int it);
int swpcsh_bodyforce(swpcsh_ctx *ctx, int it);
int swpcsh_store_receivers(swpcsh_ctx *ctx, int it);
/* SH snapshot streams (snap__write, src/swpc_sh/m_snap.f90:400-413, with
   wbuf_xz_v and wbuf_xz_u, 415-520) over 1-based (i,k) cells ik[2*npts]:
   kind SWPC_SNAP_V gives Vy x UC x M0, SWPC_SNAP_U the running sum of
   Vy x UC x M0 x dt (advanced at every step); frames of 1 x npts floats.
   Otherwise the conventions of swpc2d_add_snapshot. */
int swpcsh_add_snapshot(swpcsh_ctx *ctx, int kind, int npts, const int *ik, int ntdec_s, int nbuf,
                        float uc, float m0, float dt, int *id);
int swpcsh_snapshots(swpcsh_ctx *ctx, int it);
int swpcsh_get_snapshot(swpcsh_ctx *ctx, int id, int frame, float *out);
int swpcsh_step(swpcsh_ctx *ctx, int it);
int swpcsh_set_graph(swpcsh_ctx *ctx, int steps);
int swpcsh_run(swpcsh_ctx *ctx, int it0, int nsteps);
/* vel[itw][n] = Vy (x M0 UC 1e9 gives the reference's output), disp[itw][n] = sum Vy dt */
int swpcsh_get_receivers(swpcsh_ctx *ctx, float *vel, double *disp);
int swpcsh_get_field(swpcsh_ctx *ctx, int field, double *out);
int swpcsh_set_field(swpcsh_ctx *ctx, int field, const double *in);
/* kernel__vmax (m_kernel.f90:225-240): max |Vy(kob+1, i)| */
int swpcsh_vmax(swpcsh_ctx *ctx, float *out1);
int swpcsh_sync(swpcsh_ctx *ctx);

/* ------------------------------------------------------------------ */
/* Batches of independent runs                                         */
/* ------------------------------------------------------------------ */
/* No counterpart in the reference (one run per MPI job there): many 3-D
   and P-SV contexts of uneven size on one GPU advanced together.
   swpc_batch_run(b, it0, n, k) advances every member through steps
   it0..it0+n-1 (each in its own step order, as swpc3d_run/swpc2d_run
   would); k steps of all members are captured once into one HIP graph,
   one branch per member, and replayed with one launch per k steps (k = 0:
   step by step).  A member's setters invalidate the captured graph, which
   is then recaptured.  Members must outlive the batch; a member may not
   profile or use the concurrent PML stream. */
typedef struct swpc_batch swpc_batch;
int swpc_batch_create(swpc_batch **out);
int swpc_batch_destroy(swpc_batch *b);
int swpc_batch_add3d(swpc_batch *b, swpc3d_ctx *ctx);
int swpc_batch_add2d(swpc_batch *b, swpc2d_ctx *ctx);
int swpc_batch_run(swpc_batch *b, int it0, int nsteps, int graph_k);
/* Groups, one launch per phase for all their members instead of one graph
   branch each (mode is a bit mask; results are the same bit for bit in
   every mode):
     bit 0 (default on): FP32 P-SV members with fused steps, no snapshot
           streams and their sources inside the fused kernels;
     bit 1 (default on): FP32 3-D members of at most 2^21 cells with a
           density, the one-pass stress and z-marching velocity forms, no
           snapshots and no body force (larger members, which fill the GPU
           on their own, stay branches; measured 1.05x the branches on 64
           small 3-D grids, within 0.5 % of the autotuned branches on the
           48-run mix). */
int swpc_batch_set_grouping(swpc_batch *b, int mode);
int swpc_batch_sync(swpc_batch *b);

#ifdef __cplusplus
}
#endif

#endif /* SWPC_HIP_H */
