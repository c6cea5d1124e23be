/*
 * swpc_oracle.c -- CPU restatement of the OpenSWPC time-stepping core.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the oracle that the HIP product
 * path is checked against; only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it.  The product never links it.
 *
 * It restates, loop for loop and in the reference's own k-fastest (k,i,j)
 * layout, the hot-path routines of the reference:
 *   3-D  kernel__update_vel      src/swpc_3d/m_kernel.f90:75-140
 *   3-D  kernel__update_stress   src/swpc_3d/m_kernel.f90:142-326
 *   3-D  kernel__vmax            src/swpc_3d/m_kernel.f90:328-352
 *   3-D  absorb_p__update_vel    src/swpc_3d/m_absorb_p.f90:247-309
 *   3-D  absorb_p__update_stress src/swpc_3d/m_absorb_p.f90:423-503
 *   3-D  source__stressglut      src/swpc_3d/m_source.f90:775-833
 *   3-D  source__bodyforce       src/swpc_3d/m_source.f90:835-876
 *   3-D  wav__store (v, u)       src/swpc_3d/m_wav.f90:400-628
 *   P-SV kernel__update_vel      src/swpc_psv/m_kernel.f90:76-137
 *   P-SV kernel__update_stress   src/swpc_psv/m_kernel.f90:139-306
 *   P-SV absorb_p__update_vel    src/swpc_psv/m_absorb_p.f90:106-210
 *   P-SV absorb_p__update_stress src/swpc_psv/m_absorb_p.f90:277-479
 *   P-SV source__stressglut      src/swpc_psv/m_source.f90:572-614
 *   P-SV source__bodyforce       src/swpc_psv/m_source.f90:616-654
 *   P-SV wav__store (v, u)       src/swpc_psv/m_wav.f90:137-315
 *   P-SV snap__write products    src/swpc_psv/m_snap.f90:419-650
 *   SH   kernel, absorber, source, receivers  src/swpc_sh/m_kernel.f90,
 *        m_absorb_p.f90, m_source.f90, m_wav.f90 (cited per function)
 *   SH   snap__write products    src/swpc_sh/m_snap.f90:415-520
 *   kernel__setup coefficients   src/swpc_3d/m_kernel.f90:30-73
 *   damping_profile              src/swpc_3d/m_absorb_p.f90:507-547
 *
 * REAL is float (the reference built with MP = SP) or double (the same
 * scheme in FP64).  Medium arrays, like the reference's real(SP) arrays,
 * are always float.  The memory box is [-2, n+3] in every axis, as in
 * src/swpc_3d/m_global.f90:297-302; memory variables and ADE arrays are
 * kept over the same box (the reference keeps R over the interior box
 * only; the cells outside it are never read).
 *
 * Pinning: stencils on polynomial fields, the Zener recurrence in closed
 * form, the PML profile against the reference's own formula evaluated by
 * the amdflang-built m_fdtool, and whole-scheme physics checks live in
 * tests/test_oracle_*.py.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#ifndef REAL
#define REAL float
#endif

#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------ */
/* 3-D                                                                 */
/* ------------------------------------------------------------------ */

typedef struct {
    int nx, ny, nz, na, nm;
    double dx, dy, dz;
    float dt;
    /* fields over the memory box, (k,i,j) with k fastest */
    REAL *Vx, *Vy, *Vz, *Sxx, *Syy, *Szz, *Syz, *Sxz, *Sxy;
    /* memory variables: (m, box) with m fastest */
    REAL *Rxx, *Ryy, *Rzz, *Ryz, *Rxz, *Rxy;
    /* medium (reference real(SP)) */
    const float *bx, *by, *bz, *lam, *mu, *taup, *taus, *muyz, *muxz, *muxy;
    /* (i,j) over the memory box */
    const int *kfs_top, *kfs_bot, *kob_top, *kob_bot, *kob;
    /* damping profiles g(4, 1:n) */
    const float *gxc, *gxe, *gyc, *gye, *gzc, *gze;
    /* 18 ADE arrays over the memory box, in the order
       axVx ayVx azVx axVy ayVy azVy axVz ayVz azVz
       axSxx aySxy azSxz axSxy aySyy azSyz axSxz aySyz azSzz */
    REAL *ade[18];
    /* relaxation times ts(1:nm) */
    const float *ts;
} o3d_t;

#define NB(n) ((n) + 6)
#define I3(s, k, i, j) ((size_t)((k) + 2) + (size_t)NB((s)->nz) * ((size_t)((i) + 2) + (size_t)NB((s)->nx) * (size_t)((j) + 2)))
#define I2(s, i, j) ((size_t)((i) + 2) + (size_t)NB((s)->nx) * (size_t)((j) + 2))
#define G(a, c, n) ((a)[((c) - 1) + 4 * ((n) - 1)])

/* Stencil coefficients of kernel__setup (m_kernel.f90:43-54): with isign
   = -1 the pair is the 4th-order (9/8, 1/24)/dx, with +1 the 2nd-order
   (1, 0)/dx. */
typedef struct { REAL rc40, rc41, rd40, rd41; } coef_t;

static coef_t mkcoef(double d)
{
    coef_t c;
    REAL h = (REAL)d;
    c.rc40 = (REAL)17.0 / (REAL)16.0 / h;
    c.rc41 = (REAL)1.0 / (REAL)48.0 / h;
    c.rd40 = (REAL)(-1.0) / (REAL)16.0 / h;
    c.rd41 = (REAL)(-1.0) / (REAL)48.0 / h;
    return c;
}

/* Zener Crank-Nicolson constants (m_kernel.f90:58-67). */
typedef struct { REAL c1[8], c2[8], d1[8], d2; } zener_t;

static zener_t mkzener(int nm, const float *ts, float dt)
{
    zener_t z;
    memset(&z, 0, sizeof z);
    REAL s = 0, t = (REAL)dt;
    for (int m = 0; m < nm; m++) {
        REAL tm = (REAL)ts[m];
        z.c1[m] = (2 * tm - t) / (2 * tm + t);
        z.c2[m] = (REAL)2 / (2 * tm + t) / nm;
        s += t / (2 * tm - t);
        z.d1[m] = 2 * tm / (2 * tm - t);
    }
    z.d2 = nm > 0 ? s / nm : 0;
    return z;
}

/* Fortran sign(1, x) for integers: +1 when x >= 0. */
static inline int isgn(int a) { return a >= 0 ? 1 : -1; }

static inline int order_sign(int k, int kft, int kfb, int kot, int kobt)
{
    int a = (k - kft) * (kfb - k);
    int b = (k - kot) * (kobt - k);
    return isgn(a > b ? a : b);
}

void o3d_update_vel(o3d_t *s)
{
    const int na = s->na, nx = s->nx, ny = s->ny, nz = s->nz;
    const coef_t cx = mkcoef(s->dx), cy = mkcoef(s->dy), cz = mkcoef(s->dz);
    const REAL dt = (REAL)s->dt;
    /* interior: m_kernel.f90:99-129 */
#pragma omp parallel for schedule(static)
    for (int j = na + 1; j <= ny - na; j++)
        for (int i = na + 1; i <= nx - na; i++)
            for (int k = 1; k <= nz - na; k++) {
                size_t q = I2(s, i, j);
                int isign = order_sign(k, s->kfs_top[q], s->kfs_bot[q], s->kob_top[q], s->kob_bot[q]);
                REAL e40x = cx.rc40 + isign * cx.rd40, e41x = cx.rc41 + isign * cx.rd41;
                REAL e40y = cy.rc40 + isign * cy.rd40, e41y = cy.rc41 + isign * cy.rd41;
                REAL e40z = cz.rc40 + isign * cz.rd40, e41z = cz.rc41 + isign * cz.rd41;
#define A(F, dk, di, dj) s->F[I3(s, k + (dk), i + (di), j + (dj))]
                REAL d3Sx3 = (A(Sxx, 0, 1, 0) - A(Sxx, 0, 0, 0)) * e40x - (A(Sxx, 0, 2, 0) - A(Sxx, 0, -1, 0)) * e41x
                           + (A(Sxy, 0, 0, 0) - A(Sxy, 0, 0, -1)) * e40y - (A(Sxy, 0, 0, 1) - A(Sxy, 0, 0, -2)) * e41y
                           + (A(Sxz, 0, 0, 0) - A(Sxz, -1, 0, 0)) * e40z - (A(Sxz, 1, 0, 0) - A(Sxz, -2, 0, 0)) * e41z;
                REAL d3Sy3 = (A(Sxy, 0, 0, 0) - A(Sxy, 0, -1, 0)) * e40x - (A(Sxy, 0, 1, 0) - A(Sxy, 0, -2, 0)) * e41x
                           + (A(Syy, 0, 0, 1) - A(Syy, 0, 0, 0)) * e40y - (A(Syy, 0, 0, 2) - A(Syy, 0, 0, -1)) * e41y
                           + (A(Syz, 0, 0, 0) - A(Syz, -1, 0, 0)) * e40z - (A(Syz, 1, 0, 0) - A(Syz, -2, 0, 0)) * e41z;
                REAL d3Sz3 = (A(Sxz, 0, 0, 0) - A(Sxz, 0, -1, 0)) * e40x - (A(Sxz, 0, 1, 0) - A(Sxz, 0, -2, 0)) * e41x
                           + (A(Syz, 0, 0, 0) - A(Syz, 0, 0, -1)) * e40y - (A(Syz, 0, 0, 1) - A(Syz, 0, 0, -2)) * e41y
                           + (A(Szz, 1, 0, 0) - A(Szz, 0, 0, 0)) * e40z - (A(Szz, 2, 0, 0) - A(Szz, -1, 0, 0)) * e41z;
                size_t c = I3(s, k, i, j);
                s->Vx[c] = s->Vx[c] + s->bx[c] * d3Sx3 * dt;
                s->Vy[c] = s->Vy[c] + s->by[c] * d3Sy3 * dt;
                s->Vz[c] = s->Vz[c] + s->bz[c] * d3Sz3 * dt;
            }
}

void o3d_update_stress(o3d_t *s)
{
    const int na = s->na, nx = s->nx, ny = s->ny, nz = s->nz, nm = s->nm;
    const coef_t cx = mkcoef(s->dx), cy = mkcoef(s->dy), cz = mkcoef(s->dz);
    const zener_t z = mkzener(nm, s->ts, s->dt);
    const REAL dt = (REAL)s->dt;
    /* normal components: m_kernel.f90:176-239 */
#pragma omp parallel for schedule(static)
    for (int j = na + 1; j <= ny - na; j++)
        for (int i = na + 1; i <= nx - na; i++)
            for (int k = 1; k <= nz - na; k++) {
                size_t q = I2(s, i, j);
                int isign = order_sign(k, s->kfs_top[q], s->kfs_bot[q], s->kob_top[q], s->kob_bot[q]);
                REAL e40x = cx.rc40 + isign * cx.rd40, e41x = cx.rc41 + isign * cx.rd41;
                REAL e40y = cy.rc40 + isign * cy.rd40, e41y = cy.rc41 + isign * cy.rd41;
                REAL e40z = cz.rc40 + isign * cz.rd40, e41z = cz.rc41 + isign * cz.rd41;
                REAL dxVx = (A(Vx, 0, 0, 0) - A(Vx, 0, -1, 0)) * e40x - (A(Vx, 0, 1, 0) - A(Vx, 0, -2, 0)) * e41x;
                REAL dyVy = (A(Vy, 0, 0, 0) - A(Vy, 0, 0, -1)) * e40y - (A(Vy, 0, 0, 1) - A(Vy, 0, 0, -2)) * e41y;
                REAL dzVz = (A(Vz, 0, 0, 0) - A(Vz, -1, 0, 0)) * e40z - (A(Vz, 1, 0, 0) - A(Vz, -2, 0, 0)) * e41z;
                size_t c = I3(s, k, i, j);
                REAL mu2 = 2 * (REAL)s->mu[c];
                REAL lam2mu = (REAL)s->lam[c] + mu2;
                REAL taup1 = s->taup[c], taus1 = s->taus[c];
                REAL d3v3 = dxVx + dyVy + dzVz;
                REAL dyVy_dzVz = dyVy + dzVz, dxVx_dzVz = dxVx + dzVz, dxVx_dyVy = dxVx + dyVy;
                REAL Rxx_n = 0, Ryy_n = 0, Rzz_n = 0;
                for (int m = 0; m < nm; m++) {
                    size_t r = (size_t)m + (size_t)nm * c;
                    s->Rxx[r] = z.c1[m] * s->Rxx[r] - z.c2[m] * (lam2mu * taup1 * d3v3 - mu2 * taus1 * dyVy_dzVz) * dt;
                    s->Ryy[r] = z.c1[m] * s->Ryy[r] - z.c2[m] * (lam2mu * taup1 * d3v3 - mu2 * taus1 * dxVx_dzVz) * dt;
                    s->Rzz[r] = z.c1[m] * s->Rzz[r] - z.c2[m] * (lam2mu * taup1 * d3v3 - mu2 * taus1 * dxVx_dyVy) * dt;
                    Rxx_n += z.d1[m] * s->Rxx[r];
                    Ryy_n += z.d1[m] * s->Ryy[r];
                    Rzz_n += z.d1[m] * s->Rzz[r];
                }
                REAL taup_plus1 = 1 + taup1 * (1 + z.d2);
                REAL taus_plus1 = 1 + taus1 * (1 + z.d2);
                s->Sxx[c] = s->Sxx[c] + (lam2mu * taup_plus1 * d3v3 - mu2 * taus_plus1 * dyVy_dzVz + Rxx_n) * dt;
                s->Syy[c] = s->Syy[c] + (lam2mu * taup_plus1 * d3v3 - mu2 * taus_plus1 * dxVx_dzVz + Ryy_n) * dt;
                s->Szz[c] = s->Szz[c] + (lam2mu * taup_plus1 * d3v3 - mu2 * taus_plus1 * dxVx_dyVy + Rzz_n) * dt;
            }
    /* shear components: m_kernel.f90:263-315 */
#pragma omp parallel for schedule(static)
    for (int j = na + 1; j <= ny - na; j++)
        for (int i = na + 1; i <= nx - na; i++)
            for (int k = 1; k <= nz - na; k++) {
                size_t q = I2(s, i, j);
                int isign = order_sign(k, s->kfs_top[q], s->kfs_bot[q], s->kob_top[q], s->kob_bot[q]);
                REAL e40x = cx.rc40 + isign * cx.rd40, e41x = cx.rc41 + isign * cx.rd41;
                REAL e40y = cy.rc40 + isign * cy.rd40, e41y = cy.rc41 + isign * cy.rd41;
                REAL e40z = cz.rc40 + isign * cz.rd40, e41z = cz.rc41 + isign * cz.rd41;
                REAL dxVy_dyVx = (A(Vy, 0, 1, 0) - A(Vy, 0, 0, 0)) * e40x - (A(Vy, 0, 2, 0) - A(Vy, 0, -1, 0)) * e41x
                               + (A(Vx, 0, 0, 1) - A(Vx, 0, 0, 0)) * e40y - (A(Vx, 0, 0, 2) - A(Vx, 0, 0, -1)) * e41y;
                REAL dxVz_dzVx = (A(Vz, 0, 1, 0) - A(Vz, 0, 0, 0)) * e40x - (A(Vz, 0, 2, 0) - A(Vz, 0, -1, 0)) * e41x
                               + (A(Vx, 1, 0, 0) - A(Vx, 0, 0, 0)) * e40z - (A(Vx, 2, 0, 0) - A(Vx, -1, 0, 0)) * e41z;
                REAL dyVz_dzVy = (A(Vz, 0, 0, 1) - A(Vz, 0, 0, 0)) * e40y - (A(Vz, 0, 0, 2) - A(Vz, 0, 0, -1)) * e41y
                               + (A(Vy, 1, 0, 0) - A(Vy, 0, 0, 0)) * e40z - (A(Vy, 2, 0, 0) - A(Vy, -1, 0, 0)) * e41z;
                size_t c = I3(s, k, i, j);
                REAL taus1 = s->taus[c];
                REAL myz = s->muyz[c], mxz = s->muxz[c], mxy = s->muxy[c];
                REAL Ryz_n = 0, Rxz_n = 0, Rxy_n = 0;
                for (int m = 0; m < nm; m++) {
                    size_t r = (size_t)m + (size_t)nm * c;
                    s->Ryz[r] = z.c1[m] * s->Ryz[r] - z.c2[m] * myz * taus1 * dyVz_dzVy * dt;
                    s->Rxz[r] = z.c1[m] * s->Rxz[r] - z.c2[m] * mxz * taus1 * dxVz_dzVx * dt;
                    s->Rxy[r] = z.c1[m] * s->Rxy[r] - z.c2[m] * mxy * taus1 * dxVy_dyVx * dt;
                    Ryz_n += z.d1[m] * s->Ryz[r];
                    Rxz_n += z.d1[m] * s->Rxz[r];
                    Rxy_n += z.d1[m] * s->Rxy[r];
                }
                REAL taus_plus1 = 1 + taus1 * (1 + z.d2);
                s->Syz[c] = s->Syz[c] + (myz * taus_plus1 * dyVz_dzVy + Ryz_n) * dt;
                s->Sxz[c] = s->Sxz[c] + (mxz * taus_plus1 * dxVz_dzVx + Rxz_n) * dt;
                s->Sxy[c] = s->Sxy[c] + (mxy * taus_plus1 * dxVy_dyVx + Rxy_n) * dt;
            }
}

static inline int kbeg_a3(const o3d_t *s, int i, int j)
{
    /* m_global.f90:336-345 */
    if (i <= s->na || s->nx - s->na + 1 <= i || j <= s->na || s->ny - s->na + 1 <= j)
        return 1;
    return s->nz - s->na + 1;
}

void o3d_absorb_vel(o3d_t *s)
{
    const int nx = s->nx, ny = s->ny, nz = s->nz;
    const REAL r20x = (REAL)1.0 / (REAL)s->dx, r20y = (REAL)1.0 / (REAL)s->dy, r20z = (REAL)1.0 / (REAL)s->dz;
    const REAL dt = (REAL)s->dt;
    REAL *axSxx = s->ade[9], *aySxy = s->ade[10], *azSxz = s->ade[11];
    REAL *axSxy = s->ade[12], *aySyy = s->ade[13], *azSyz = s->ade[14];
    REAL *axSxz = s->ade[15], *aySyz = s->ade[16], *azSzz = s->ade[17];
    /* m_absorb_p.f90:260-302 */
#pragma omp parallel for schedule(dynamic)
    for (int j = 1; j <= ny; j++)
        for (int i = 1; i <= nx; i++)
            for (int k = kbeg_a3(s, i, j); k <= nz; k++) {
                REAL dxSxx = (A(Sxx, 0, 1, 0) - A(Sxx, 0, 0, 0)) * r20x;
                REAL dySyy = (A(Syy, 0, 0, 1) - A(Syy, 0, 0, 0)) * r20y;
                REAL dzSzz = (A(Szz, 1, 0, 0) - A(Szz, 0, 0, 0)) * r20z;
                REAL dySyz = (A(Syz, 0, 0, 0) - A(Syz, 0, 0, -1)) * r20y;
                REAL dzSyz = (A(Syz, 0, 0, 0) - A(Syz, -1, 0, 0)) * r20z;
                REAL dxSxz = (A(Sxz, 0, 0, 0) - A(Sxz, 0, -1, 0)) * r20x;
                REAL dzSxz = (A(Sxz, 0, 0, 0) - A(Sxz, -1, 0, 0)) * r20z;
                REAL dxSxy = (A(Sxy, 0, 0, 0) - A(Sxy, 0, -1, 0)) * r20x;
                REAL dySxy = (A(Sxy, 0, 0, 0) - A(Sxy, 0, 0, -1)) * r20y;
                size_t c = I3(s, k, i, j);
                REAL gxc1 = G(s->gxc, 1, i), gxc2 = G(s->gxc, 2, i), gxc3 = G(s->gxc, 3, i), gxc4 = G(s->gxc, 4, i);
                REAL gxe1 = G(s->gxe, 1, i), gxe2 = G(s->gxe, 2, i), gxe3 = G(s->gxe, 3, i), gxe4 = G(s->gxe, 4, i);
                REAL gyc1 = G(s->gyc, 1, j), gyc2 = G(s->gyc, 2, j), gyc3 = G(s->gyc, 3, j), gyc4 = G(s->gyc, 4, j);
                REAL gye1 = G(s->gye, 1, j), gye2 = G(s->gye, 2, j), gye3 = G(s->gye, 3, j), gye4 = G(s->gye, 4, j);
                REAL gzc1 = G(s->gzc, 1, k), gzc2 = G(s->gzc, 2, k), gzc3 = G(s->gzc, 3, k), gzc4 = G(s->gzc, 4, k);
                REAL gze1 = G(s->gze, 1, k), gze2 = G(s->gze, 2, k), gze3 = G(s->gze, 3, k), gze4 = G(s->gze, 4, k);
                s->Vx[c] = s->Vx[c] + s->bx[c] * (gxe1 * dxSxx + gyc1 * dySxy + gzc1 * dzSxz
                                                 + gxe2 * axSxx[c] + gyc2 * aySxy[c] + gzc2 * azSxz[c]) * dt;
                s->Vy[c] = s->Vy[c] + s->by[c] * (gxc1 * dxSxy + gye1 * dySyy + gzc1 * dzSyz
                                                 + gxc2 * axSxy[c] + gye2 * aySyy[c] + gzc2 * azSyz[c]) * dt;
                s->Vz[c] = s->Vz[c] + s->bz[c] * (gxc1 * dxSxz + gyc1 * dySyz + gze1 * dzSzz
                                                 + gxc2 * axSxz[c] + gyc2 * aySyz[c] + gze2 * azSzz[c]) * dt;
                axSxx[c] = gxe3 * axSxx[c] + gxe4 * dxSxx * dt;
                aySxy[c] = gyc3 * aySxy[c] + gyc4 * dySxy * dt;
                azSxz[c] = gzc3 * azSxz[c] + gzc4 * dzSxz * dt;
                axSxy[c] = gxc3 * axSxy[c] + gxc4 * dxSxy * dt;
                aySyy[c] = gye3 * aySyy[c] + gye4 * dySyy * dt;
                azSyz[c] = gzc3 * azSyz[c] + gzc4 * dzSyz * dt;
                axSxz[c] = gxc3 * axSxz[c] + gxc4 * dxSxz * dt;
                aySyz[c] = gyc3 * aySyz[c] + gyc4 * dySyz * dt;
                azSzz[c] = gze3 * azSzz[c] + gze4 * dzSzz * dt;
            }
}

void o3d_absorb_stress(o3d_t *s)
{
    const int nx = s->nx, ny = s->ny, nz = s->nz;
    const REAL r20x = (REAL)1.0 / (REAL)s->dx, r20y = (REAL)1.0 / (REAL)s->dy, r20z = (REAL)1.0 / (REAL)s->dz;
    const REAL dt = (REAL)s->dt;
    REAL *axVx = s->ade[0], *ayVx = s->ade[1], *azVx = s->ade[2];
    REAL *axVy = s->ade[3], *ayVy = s->ade[4], *azVy = s->ade[5];
    REAL *axVz = s->ade[6], *ayVz = s->ade[7], *azVz = s->ade[8];
    /* m_absorb_p.f90:438-496 */
#pragma omp parallel for schedule(dynamic)
    for (int j = 1; j <= ny; j++)
        for (int i = 1; i <= nx; i++) {
            REAL gxc1 = G(s->gxc, 1, i), gxc2 = G(s->gxc, 2, i), gxc3 = G(s->gxc, 3, i), gxc4 = G(s->gxc, 4, i);
            REAL gxe1 = G(s->gxe, 1, i), gxe2 = G(s->gxe, 2, i), gxe3 = G(s->gxe, 3, i), gxe4 = G(s->gxe, 4, i);
            REAL gyc1 = G(s->gyc, 1, j), gyc2 = G(s->gyc, 2, j), gyc3 = G(s->gyc, 3, j), gyc4 = G(s->gyc, 4, j);
            REAL gye1 = G(s->gye, 1, j), gye2 = G(s->gye, 2, j), gye3 = G(s->gye, 3, j), gye4 = G(s->gye, 4, j);
            for (int k = kbeg_a3(s, i, j); k <= nz; k++) {
                REAL dxVx = (A(Vx, 0, 0, 0) - A(Vx, 0, -1, 0)) * r20x;
                REAL dyVy = (A(Vy, 0, 0, 0) - A(Vy, 0, 0, -1)) * r20y;
                REAL dzVz = (A(Vz, 0, 0, 0) - A(Vz, -1, 0, 0)) * r20z;
                size_t c = I3(s, k, i, j);
                REAL lam2mu_R = (REAL)(s->lam[c] + 2 * s->mu[c]);
                REAL lam_R = lam2mu_R - 2 * (REAL)s->mu[c];
                REAL gzc1 = G(s->gzc, 1, k), gzc2 = G(s->gzc, 2, k), gzc3 = G(s->gzc, 3, k), gzc4 = G(s->gzc, 4, k);
                REAL dxVx_ade = gxc1 * dxVx + gxc2 * axVx[c];
                REAL dyVy_ade = gyc1 * dyVy + gyc2 * ayVy[c];
                REAL dzVz_ade = gzc1 * dzVz + gzc2 * azVz[c];
                s->Sxx[c] = s->Sxx[c] + (lam2mu_R * dxVx_ade + lam_R * (dyVy_ade + dzVz_ade)) * dt;
                s->Syy[c] = s->Syy[c] + (lam2mu_R * dyVy_ade + lam_R * (dxVx_ade + dzVz_ade)) * dt;
                s->Szz[c] = s->Szz[c] + (lam2mu_R * dzVz_ade + lam_R * (dxVx_ade + dyVy_ade)) * dt;
                axVx[c] = gxc3 * axVx[c] + gxc4 * dxVx * dt;
                ayVy[c] = gyc3 * ayVy[c] + gyc4 * dyVy * dt;
                azVz[c] = gzc3 * azVz[c] + gzc4 * dzVz * dt;
            }
            for (int k = kbeg_a3(s, i, j); k <= nz; k++) {
                REAL dxVy = (A(Vy, 0, 1, 0) - A(Vy, 0, 0, 0)) * r20x;
                REAL dxVz = (A(Vz, 0, 1, 0) - A(Vz, 0, 0, 0)) * r20x;
                REAL dyVx = (A(Vx, 0, 0, 1) - A(Vx, 0, 0, 0)) * r20y;
                REAL dyVz = (A(Vz, 0, 0, 1) - A(Vz, 0, 0, 0)) * r20y;
                REAL dzVx = (A(Vx, 1, 0, 0) - A(Vx, 0, 0, 0)) * r20z;
                REAL dzVy = (A(Vy, 1, 0, 0) - A(Vy, 0, 0, 0)) * r20z;
                size_t c = I3(s, k, i, j);
                REAL gze1 = G(s->gze, 1, k), gze2 = G(s->gze, 2, k), gze3 = G(s->gze, 3, k), gze4 = G(s->gze, 4, k);
                s->Syz[c] = s->Syz[c] + s->muyz[c] * (gye1 * dyVz + gze1 * dzVy + gye2 * ayVz[c] + gze2 * azVy[c]) * dt;
                s->Sxz[c] = s->Sxz[c] + s->muxz[c] * (gxe1 * dxVz + gze1 * dzVx + gxe2 * axVz[c] + gze2 * azVx[c]) * dt;
                s->Sxy[c] = s->Sxy[c] + s->muxy[c] * (gxe1 * dxVy + gye1 * dyVx + gxe2 * axVy[c] + gye2 * ayVx[c]) * dt;
                ayVx[c] = gye3 * ayVx[c] + gye4 * dyVx * dt;
                azVx[c] = gze3 * azVx[c] + gze4 * dzVx * dt;
                axVy[c] = gxe3 * axVy[c] + gxe4 * dxVy * dt;
                azVy[c] = gze3 * azVy[c] + gze4 * dzVy * dt;
                axVz[c] = gxe3 * axVz[c] + gxe4 * dxVz * dt;
                ayVz[c] = gye3 * ayVz[c] + gye4 * dyVz * dt;
            }
        }
}
#undef A

/* Stress-glut injection, m_source.f90:793-828.  sdrop is the per-source
   scalar mo*momentrate(tbeg+(it-1/2)dt)*dt/(dx dy dz), evaluated by the
   host; m6 holds (mxx, myy, mzz, myz, mxz, mxy) per source. */
void o3d_stressglut(o3d_t *s, int nsrc, const int *isrc, const int *jsrc, const int *ksrc,
                    const double *m6, const double *sdrop)
{
    for (int n = 0; n < nsrc; n++) {
        int ii = isrc[n], jj = jsrc[n], kk = ksrc[n];
        REAL sd = (REAL)sdrop[n];
        REAL mxx = (REAL)m6[6 * n + 0], myy = (REAL)m6[6 * n + 1], mzz = (REAL)m6[6 * n + 2];
        REAL myz = (REAL)m6[6 * n + 3], mxz = (REAL)m6[6 * n + 4], mxy = (REAL)m6[6 * n + 5];
        s->Sxx[I3(s, kk, ii, jj)] -= mxx * sd;
        s->Syy[I3(s, kk, ii, jj)] -= myy * sd;
        s->Szz[I3(s, kk, ii, jj)] -= mzz * sd;
        s->Sxy[I3(s, kk, ii, jj)] -= mxy * sd / 4;
        s->Sxy[I3(s, kk, ii, jj - 1)] -= mxy * sd / 4;
        s->Sxy[I3(s, kk, ii - 1, jj)] -= mxy * sd / 4;
        s->Sxy[I3(s, kk, ii - 1, jj - 1)] -= mxy * sd / 4;
        s->Sxz[I3(s, kk, ii, jj)] -= mxz * sd / 4;
        s->Sxz[I3(s, kk - 1, ii, jj)] -= mxz * sd / 4;
        s->Sxz[I3(s, kk, ii - 1, jj)] -= mxz * sd / 4;
        s->Sxz[I3(s, kk - 1, ii - 1, jj)] -= mxz * sd / 4;
        s->Syz[I3(s, kk, ii, jj)] -= myz * sd / 4;
        s->Syz[I3(s, kk - 1, ii, jj)] -= myz * sd / 4;
        s->Syz[I3(s, kk, ii, jj - 1)] -= myz * sd / 4;
        s->Syz[I3(s, kk - 1, ii, jj - 1)] -= myz * sd / 4;
    }
}

/* Body force, m_source.f90:850-873.  fdt[n] = stime*dt/(dx dy dz)/2 and
   f3 = (fx, fy, fz) per source, both host-evaluated. */
void o3d_bodyforce(o3d_t *s, int nsrc, const int *isrc, const int *jsrc, const int *ksrc,
                   const double *f3, const double *fdt)
{
    for (int n = 0; n < nsrc; n++) {
        int ii = isrc[n], jj = jsrc[n], kk = ksrc[n];
        REAL a = (REAL)fdt[n];
        REAL fx = (REAL)f3[3 * n], fy = (REAL)f3[3 * n + 1], fz = (REAL)f3[3 * n + 2];
        size_t c;
        c = I3(s, kk, ii, jj);     s->Vx[c] += s->bx[c] * fx * a;
        c = I3(s, kk, ii - 1, jj); s->Vx[c] += s->bx[c] * fx * a;
        c = I3(s, kk, ii, jj);     s->Vy[c] += s->by[c] * fy * a;
        c = I3(s, kk, ii, jj - 1); s->Vy[c] += s->by[c] * fy * a;
        c = I3(s, kk, ii, jj);     s->Vz[c] += s->bz[c] * fz * a;
        c = I3(s, kk - 1, ii, jj); s->Vz[c] += s->bz[c] * fz * a;
    }
}

/* Receiver gather, m_wav.f90:436-452 and 524-535.  vel[3*n+c] receives
   the two-node average of each velocity component (z positive upward)
   in solver units; the M0*UC*1e9 scale is applied by the caller.
   disp[3*n+c] accumulates the same averages times dt in FP64. */
void o3d_receivers(o3d_t *s, int nst, const int *ist, const int *jst, const int *kst,
                   float *vel, double *disp)
{
    for (int n = 0; n < nst; n++) {
        int i = ist[n], j = jst[n], k = kst[n];
        REAL vx = (s->Vx[I3(s, k, i, j)] + s->Vx[I3(s, k, i - 1, j)]);
        REAL vy = (s->Vy[I3(s, k, i, j)] + s->Vy[I3(s, k, i, j - 1)]);
        REAL vz = (s->Vz[I3(s, k, i, j)] + s->Vz[I3(s, k - 1, i, j)]);
        if (disp) {
            disp[3 * n + 0] += (double)vx * 0.5 * (double)s->dt;
            disp[3 * n + 1] += (double)vy * 0.5 * (double)s->dt;
            disp[3 * n + 2] -= (double)vz * 0.5 * (double)s->dt;
        }
        if (vel) {
            vel[3 * n + 0] = (float)(vx / 2);
            vel[3 * n + 1] = (float)(vy / 2);
            vel[3 * n + 2] = (float)(-vz / 2);
        }
    }
}

/* kernel__vmax, m_kernel.f90:328-352 (margin 5 off the absorber). */
void o3d_vmax(o3d_t *s, float *out3)
{
    const int margin = 5;
    float xm = 0, ym = 0, zm = 0;
    for (int j = s->na + margin + 1; j <= s->ny - s->na - margin; j++)
        for (int i = s->na + margin + 1; i <= s->nx - s->na - margin; i++) {
            int k = s->kob[I2(s, i, j)] + 1;
            size_t c = I3(s, k, i, j);
            float a = fabsf((float)s->Vx[c]), b = fabsf((float)s->Vy[c]), d = fabsf((float)s->Vz[c]);
            if (a > xm) xm = a;
            if (b > ym) ym = b;
            if (d > zm) zm = d;
        }
    out3[0] = xm; out3[1] = ym; out3[2] = zm;
}

/* Snapshot products on a cell list (snap__write, src/swpc_3d/m_snap.f90):
   kind 0 'ps' follows wbuf_xy_ps/wbuf_fs_ps (1183-1216, 1278-1304),
   kind 1 'v' wbuf_xy_v (1600-1606), kind 2 'u' wbuf_fs_u (2041-2049; the
   running sum acc[3*npts] is advanced at every call).  ijk holds 1-based
   (i,j,k) triplets; out is variable-major (n*npts + p). */
void o3d_snapshot(o3d_t *s, int kind, int npts, const int *ijk, float uc, float m0, float *acc, float *out)
{
    const REAL r20x = (REAL)1.0 / (REAL)s->dx, r20y = (REAL)1.0 / (REAL)s->dy, r20z = (REAL)1.0 / (REAL)s->dz;
    const float eps = 1.1920929e-07f; /* epsilon(1.0) */
    for (int p = 0; p < npts; p++) {
        const int i = ijk[3 * p], j = ijk[3 * p + 1], k = ijk[3 * p + 2];
        const size_t c = I3(s, k, i, j);
        if (kind == 0) {
            float div = (float)((s->Vx[c] - s->Vx[I3(s, k, i - 1, j)]) * r20x
                              + (s->Vy[c] - s->Vy[I3(s, k, i, j - 1)]) * r20y
                              + (s->Vz[c] - s->Vz[I3(s, k - 1, i, j)]) * r20z);
            float rot_x = (float)((s->Vz[I3(s, k, i, j + 1)] - s->Vz[c]) * r20y
                                - (s->Vy[I3(s, k + 1, i, j)] - s->Vy[c]) * r20z);
            float rot_y = (float)((s->Vx[I3(s, k + 1, i, j)] - s->Vx[c]) * r20z
                                - (s->Vz[I3(s, k, i + 1, j)] - s->Vz[c]) * r20x);
            float rot_z = (float)((s->Vy[I3(s, k, i + 1, j)] - s->Vy[c]) * r20x
                                - (s->Vx[I3(s, k, i, j + 1)] - s->Vx[c]) * r20y);
            div = div * s->lam[c] / fabsf(s->lam[c] + eps);
            rot_x = rot_x * s->muyz[c] / fabsf(s->muyz[c] + eps);
            rot_y = rot_y * s->muxz[c] / fabsf(s->muxz[c] + eps);
            rot_z = rot_z * s->muxy[c] / fabsf(s->muxy[c] + eps);
            out[p] = div * uc * m0 * 1e-3f;
            out[npts + p] = rot_x * uc * m0 * 1e-3f;
            out[2 * npts + p] = rot_y * uc * m0 * 1e-3f;
            out[3 * npts + p] = rot_z * uc * m0 * 1e-3f;
        } else {
            const REAL v[3] = {s->Vx[c], s->Vy[c], s->Vz[c]};
            for (int n = 0; n < 3; n++) {
                if (kind == 1) {
                    out[n * npts + p] = (float)(v[n] * (REAL)uc * (REAL)m0);
                } else {
                    acc[n * npts + p] = (float)((REAL)acc[n * npts + p] + v[n] * (REAL)uc * (REAL)m0 * (REAL)s->dt);
                    out[n * npts + p] = acc[n * npts + p];
                }
            }
        }
    }
}

/* One full step in the order of src/swpc_3d/main.f90:125-144 (receivers
   are sampled by the caller before the updates). */
void o3d_step(o3d_t *s, int nsrc, const int *isrc, const int *jsrc, const int *ksrc,
              const double *m6, const double *sdrop)
{
    o3d_update_stress(s);
    o3d_absorb_stress(s);
    if (nsrc > 0) o3d_stressglut(s, nsrc, isrc, jsrc, ksrc, m6, sdrop);
    o3d_update_vel(s);
    o3d_absorb_vel(s);
}

/* ------------------------------------------------------------------ */
/* P-SV                                                                */
/* ------------------------------------------------------------------ */

typedef struct {
    int nx, nz, na, nm;
    double dx, dz;
    float dt;
    REAL *Vx, *Vz, *Sxx, *Szz, *Sxz;
    REAL *Rxx, *Rzz, *Rxz;                 /* (m, box) */
    const float *rho, *lam, *mu, *taup, *taus;
    const int *kfs_top, *kfs_bot, *kob_top, *kob_bot, *kob;   /* over i in the memory box */
    const float *gxc, *gxe, *gzc, *gze;
    /* axVx azVx axVz azVz axSxx azSxz axSxz azSzz */
    REAL *ade[8];
    const float *ts;
} o2d_t;

#define J2(s, k, i) ((size_t)((k) + 2) + (size_t)NB((s)->nz) * (size_t)((i) + 2))
#define J1(i) ((size_t)((i) + 2))

void o2d_update_vel(o2d_t *s)
{
    const int na = s->na, nx = s->nx, nz = s->nz;
    const coef_t cx = mkcoef(s->dx), cz = mkcoef(s->dz);
    const REAL dt = (REAL)s->dt;
    /* src/swpc_psv/m_kernel.f90:101-126 */
#pragma omp parallel for schedule(static)
    for (int i = na + 1; i <= nx - na; i++)
        for (int k = 1; k <= nz - na; k++) {
            int isign = order_sign(k, s->kfs_top[J1(i)], s->kfs_bot[J1(i)], s->kob_top[J1(i)], s->kob_bot[J1(i)]);
            REAL e40x = cx.rc40 + isign * cx.rd40, e41x = cx.rc41 + isign * cx.rd41;
            REAL e40z = cz.rc40 + isign * cz.rd40, e41z = cz.rc41 + isign * cz.rd41;
#define B(F, dk, di) s->F[J2(s, k + (dk), i + (di))]
            REAL dxSxx = (B(Sxx, 0, 1) - B(Sxx, 0, 0)) * e40x - (B(Sxx, 0, 2) - B(Sxx, 0, -1)) * e41x;
            REAL dzSzz = (B(Szz, 1, 0) - B(Szz, 0, 0)) * e40z - (B(Szz, 2, 0) - B(Szz, -1, 0)) * e41z;
            REAL dxSxz = (B(Sxz, 0, 0) - B(Sxz, 0, -1)) * e40x - (B(Sxz, 0, 1) - B(Sxz, 0, -2)) * e41x;
            REAL dzSxz = (B(Sxz, 0, 0) - B(Sxz, -1, 0)) * e40z - (B(Sxz, 1, 0) - B(Sxz, -2, 0)) * e41z;
            size_t c = J2(s, k, i);
            float bx = 2.0f / (s->rho[c] + s->rho[J2(s, k, i + 1)]);
            float bz = 2.0f / (s->rho[c] + s->rho[J2(s, k + 1, i)]);
            s->Vx[c] = s->Vx[c] + bx * (dxSxx + dzSxz) * dt;
            s->Vz[c] = s->Vz[c] + bz * (dxSxz + dzSzz) * dt;
        }
}

static inline float harm4(float nnn, float pnn, float npn, float ppn)
{
    const float epsl = 1.1920928955078125e-07f; /* epsilon(1.0) */
    return 4 * nnn * pnn * npn * ppn / (nnn * pnn * npn + nnn * pnn * ppn + nnn * npn * ppn + pnn * npn * ppn + epsl);
}

void o2d_update_stress(o2d_t *s)
{
    const int na = s->na, nx = s->nx, nz = s->nz, nm = s->nm;
    const coef_t cx = mkcoef(s->dx), cz = mkcoef(s->dz);
    const zener_t z = mkzener(nm, s->ts, s->dt);
    const REAL dt = (REAL)s->dt;
    /* src/swpc_psv/m_kernel.f90:177-226 */
#pragma omp parallel for schedule(static)
    for (int i = na + 1; i <= nx - na; i++)
        for (int k = 1; k <= nz - na; k++) {
            int isign = order_sign(k, s->kfs_top[J1(i)], s->kfs_bot[J1(i)], s->kob_top[J1(i)], s->kob_bot[J1(i)]);
            REAL e40x = cx.rc40 + isign * cx.rd40, e41x = cx.rc41 + isign * cx.rd41;
            REAL e40z = cz.rc40 + isign * cz.rd40, e41z = cz.rc41 + isign * cz.rd41;
            REAL dxVx = (B(Vx, 0, 0) - B(Vx, 0, -1)) * e40x - (B(Vx, 0, 1) - B(Vx, 0, -2)) * e41x;
            REAL dzVz = (B(Vz, 0, 0) - B(Vz, -1, 0)) * e40z - (B(Vz, 1, 0) - B(Vz, -2, 0)) * e41z;
            size_t c = J2(s, k, i);
            REAL mu2 = 2 * (REAL)s->mu[c];
            REAL lam2mu = (REAL)s->lam[c] + mu2;
            REAL taup1 = s->taup[c], taus1 = s->taus[c];
            REAL d2v2 = dxVx + dzVz;
            REAL f_Rxx = lam2mu * taup1 * d2v2 - mu2 * taus1 * dzVz;
            REAL f_Rzz = lam2mu * taup1 * d2v2 - mu2 * taus1 * dxVx;
            REAL Rxx_n = 0, Rzz_n = 0;
            for (int m = 0; m < nm; m++) {
                size_t r = (size_t)m + (size_t)nm * c;
                s->Rxx[r] = z.c1[m] * s->Rxx[r] - z.c2[m] * f_Rxx * dt;
                s->Rzz[r] = z.c1[m] * s->Rzz[r] - z.c2[m] * f_Rzz * dt;
                Rxx_n += z.d1[m] * s->Rxx[r];
                Rzz_n += z.d1[m] * s->Rzz[r];
            }
            REAL taup_plus1 = 1 + taup1 * (1 + z.d2);
            REAL taus_plus1 = 1 + taus1 * (1 + z.d2);
            s->Sxx[c] = s->Sxx[c] + (lam2mu * taup_plus1 * d2v2 - mu2 * taus_plus1 * dzVz + Rxx_n) * dt;
            s->Szz[c] = s->Szz[c] + (lam2mu * taup_plus1 * d2v2 - mu2 * taus_plus1 * dxVx + Rzz_n) * dt;
        }
    /* src/swpc_psv/m_kernel.f90:251-295 */
#pragma omp parallel for schedule(static)
    for (int i = na + 1; i <= nx - na; i++)
        for (int k = 1; k <= nz - na; k++) {
            int isign = order_sign(k, s->kfs_top[J1(i)], s->kfs_bot[J1(i)], s->kob_top[J1(i)], s->kob_bot[J1(i)]);
            REAL e40x = cx.rc40 + isign * cx.rd40, e41x = cx.rc41 + isign * cx.rd41;
            REAL e40z = cz.rc40 + isign * cz.rd40, e41z = cz.rc41 + isign * cz.rd41;
            REAL dxVz = (B(Vz, 0, 1) - B(Vz, 0, 0)) * e40x - (B(Vz, 0, 2) - B(Vz, 0, -1)) * e41x;
            REAL dzVx = (B(Vx, 1, 0) - B(Vx, 0, 0)) * e40z - (B(Vx, 2, 0) - B(Vx, -1, 0)) * e41z;
            size_t c = J2(s, k, i);
            REAL taus1 = s->taus[c];
            REAL mu_xz = harm4(s->mu[c], s->mu[J2(s, k + 1, i)], s->mu[J2(s, k, i + 1)], s->mu[J2(s, k + 1, i + 1)]);
            REAL dxVz_dzVx = dxVz + dzVx;
            REAL f_Rxz = mu_xz * taus1 * dxVz_dzVx;
            REAL Rxz_n = 0;
            for (int m = 0; m < nm; m++) {
                size_t r = (size_t)m + (size_t)nm * c;
                s->Rxz[r] = z.c1[m] * s->Rxz[r] - z.c2[m] * f_Rxz * dt;
                Rxz_n += z.d1[m] * s->Rxz[r];
            }
            REAL taus_plus1 = 1 + taus1 * (1 + z.d2);
            s->Sxz[c] = s->Sxz[c] + (mu_xz * taus_plus1 * dxVz_dzVx + Rxz_n) * dt;
        }
}

static inline int kbeg_a2(const o2d_t *s, int i)
{
    /* src/swpc_psv/m_global.f90 kbeg_a */
    if (i <= s->na || s->nx - s->na + 1 <= i) return 1;
    return s->nz - s->na + 1;
}

void o2d_absorb_vel(o2d_t *s)
{
    const REAL r20x = (REAL)(1.0 / s->dx), r20z = (REAL)(1.0 / s->dz);
    const REAL dt = (REAL)s->dt;
    REAL *axSxx = s->ade[4], *azSxz = s->ade[5], *axSxz = s->ade[6], *azSzz = s->ade[7];
    /* src/swpc_psv/m_absorb_p.f90:171-203 */
#pragma omp parallel for schedule(dynamic)
    for (int i = 1; i <= s->nx; i++)
        for (int k = kbeg_a2(s, i); k <= s->nz; k++) {
            REAL dxSxx = (B(Sxx, 0, 1) - B(Sxx, 0, 0)) * r20x;
            REAL dzSzz = (B(Szz, 1, 0) - B(Szz, 0, 0)) * r20z;
            REAL dxSxz = (B(Sxz, 0, 0) - B(Sxz, 0, -1)) * r20x;
            REAL dzSxz = (B(Sxz, 0, 0) - B(Sxz, -1, 0)) * r20z;
            size_t c = J2(s, k, i);
            float bx = 2.0f / (s->rho[c] + s->rho[J2(s, k, i + 1)]);
            float bz = 2.0f / (s->rho[c] + s->rho[J2(s, k + 1, i)]);
            s->Vx[c] = s->Vx[c] + bx * (G(s->gxe, 1, i) * dxSxx + G(s->gzc, 1, k) * dzSxz
                                       + G(s->gxe, 2, i) * axSxx[c] + G(s->gzc, 2, k) * azSxz[c]) * dt;
            s->Vz[c] = s->Vz[c] + bz * (G(s->gxc, 1, i) * dxSxz + G(s->gze, 1, k) * dzSzz
                                       + G(s->gxc, 2, i) * axSxz[c] + G(s->gze, 2, k) * azSzz[c]) * dt;
            axSxx[c] = G(s->gxe, 3, i) * axSxx[c] + G(s->gxe, 4, i) * dxSxx * dt;
            azSxz[c] = G(s->gzc, 3, k) * azSxz[c] + G(s->gzc, 4, k) * dzSxz * dt;
            axSxz[c] = G(s->gxc, 3, i) * axSxz[c] + G(s->gxc, 4, i) * dxSxz * dt;
            azSzz[c] = G(s->gze, 3, k) * azSzz[c] + G(s->gze, 4, k) * dzSzz * dt;
        }
}

void o2d_absorb_stress(o2d_t *s)
{
    const REAL r20x = (REAL)(1.0 / s->dx), r20z = (REAL)(1.0 / s->dz);
    const REAL dt = (REAL)s->dt;
    REAL *axVx = s->ade[0], *azVx = s->ade[1], *axVz = s->ade[2], *azVz = s->ade[3];
    /* src/swpc_psv/m_absorb_p.f90:345-386 */
#pragma omp parallel for schedule(dynamic)
    for (int i = 1; i <= s->nx; i++) {
        for (int k = kbeg_a2(s, i); k <= s->nz; k++) {
            REAL dxVx = (B(Vx, 0, 0) - B(Vx, 0, -1)) * r20x;
            REAL dzVz = (B(Vz, 0, 0) - B(Vz, -1, 0)) * r20z;
            size_t c = J2(s, k, i);
            REAL lam2mu_R = (REAL)(s->lam[c] + 2 * s->mu[c]);
            REAL lam_R = lam2mu_R - 2 * (REAL)s->mu[c];
            REAL dxVx_ade = G(s->gxc, 1, i) * dxVx + G(s->gxc, 2, i) * axVx[c];
            REAL dzVz_ade = G(s->gzc, 1, k) * dzVz + G(s->gzc, 2, k) * azVz[c];
            s->Sxx[c] = s->Sxx[c] + (lam2mu_R * dxVx_ade + lam_R * dzVz_ade) * dt;
            s->Szz[c] = s->Szz[c] + (lam2mu_R * dzVz_ade + lam_R * dxVx_ade) * dt;
            axVx[c] = G(s->gxc, 3, i) * axVx[c] + G(s->gxc, 4, i) * dxVx * dt;
            azVz[c] = G(s->gzc, 3, k) * azVz[c] + G(s->gzc, 4, k) * dzVz * dt;
        }
        for (int k = kbeg_a2(s, i); k <= s->nz; k++) {
            REAL dzVx = (B(Vx, 1, 0) - B(Vx, 0, 0)) * r20z;
            REAL dxVz = (B(Vz, 0, 1) - B(Vz, 0, 0)) * r20x;
            size_t c = J2(s, k, i);
            REAL muxz = harm4(s->mu[c], s->mu[J2(s, k + 1, i)], s->mu[J2(s, k, i + 1)], s->mu[J2(s, k + 1, i + 1)]);
            s->Sxz[c] = s->Sxz[c] + muxz * (G(s->gxe, 1, i) * dxVz + G(s->gze, 1, k) * dzVx
                                           + G(s->gxe, 2, i) * axVz[c] + G(s->gze, 2, k) * azVx[c]) * dt;
            azVx[c] = G(s->gze, 3, k) * azVx[c] + G(s->gze, 4, k) * dzVx * dt;
            axVz[c] = G(s->gxe, 3, i) * axVz[c] + G(s->gxe, 4, i) * dxVz * dt;
        }
    }
}

/* src/swpc_psv/m_source.f90:590-607; m3 = (mxx, mzz, mxz). */
void o2d_stressglut(o2d_t *s, int nsrc, const int *isrc, const int *ksrc, const double *m3, const double *sdrop)
{
    for (int n = 0; n < nsrc; n++) {
        int ii = isrc[n], kk = ksrc[n];
        REAL sd = (REAL)sdrop[n];
        REAL mxx = (REAL)m3[3 * n], mzz = (REAL)m3[3 * n + 1], mxz = (REAL)m3[3 * n + 2];
        s->Sxx[J2(s, kk, ii)] -= mxx * sd;
        s->Szz[J2(s, kk, ii)] -= mzz * sd;
        s->Sxz[J2(s, kk, ii)] -= mxz * sd / 4;
        s->Sxz[J2(s, kk - 1, ii)] -= mxz * sd / 4;
        s->Sxz[J2(s, kk, ii - 1)] -= mxz * sd / 4;
        s->Sxz[J2(s, kk - 1, ii - 1)] -= mxz * sd / 4;
    }
}

/* src/swpc_psv/m_source.f90:636-646; fdt = stime*dt/(dx dz)/2, f2 = (fx, fz). */
void o2d_bodyforce(o2d_t *s, int nsrc, const int *isrc, const int *ksrc, const double *f2, const double *fdt)
{
    for (int n = 0; n < nsrc; n++) {
        int ii = isrc[n], kk = ksrc[n];
        REAL a = (REAL)fdt[n] / (REAL)s->rho[J2(s, kk, ii)];
        REAL fx = (REAL)f2[2 * n], fz = (REAL)f2[2 * n + 1];
        s->Vx[J2(s, kk, ii)] += fx * a;
        s->Vx[J2(s, kk, ii - 1)] += fx * a;
        s->Vz[J2(s, kk, ii)] += fz * a;
        s->Vz[J2(s, kk - 1, ii)] += fz * a;
    }
}

/* src/swpc_psv/m_wav.f90:170-240 (v and u). */
/* P-SV snapshot products on an (i,k) cell list (src/swpc_psv/m_snap.f90:
   wbuf_xz_ps 461-494, wbuf_xz_v 552-553, wbuf_xz_u 612-613). */
void o2d_snapshot(o2d_t *s, int kind, int npts, const int *ik, float uc, float m0, float *acc, float *out)
{
    const REAL r20x = (REAL)1.0 / (REAL)s->dx, r20z = (REAL)1.0 / (REAL)s->dz;
    const float eps = 1.1920929e-07f; /* epsilon(1.0) */
    for (int p = 0; p < npts; p++) {
        const int i = ik[2 * p], k = ik[2 * p + 1];
        const size_t c = J2(s, k, i);
        if (kind == 0) {
            float div = (float)((s->Vx[c] - s->Vx[J2(s, k, i - 1)]) * r20x + (s->Vz[c] - s->Vz[J2(s, k - 1, i)]) * r20z);
            float rot = (float)((s->Vx[J2(s, k + 1, i)] - s->Vx[c]) * r20z - (s->Vz[J2(s, k, i + 1)] - s->Vz[c]) * r20x);
            float nnn = s->mu[c], pnn = s->mu[J2(s, k + 1, i)], npn = s->mu[J2(s, k, i + 1)], ppn = s->mu[J2(s, k + 1, i + 1)];
            float mu_xz = 4 * nnn * pnn * npn * ppn
                        / (nnn * pnn * npn + nnn * pnn * ppn + nnn * npn * ppn + pnn * npn * ppn + eps);
            div = div * s->lam[c] / (fabsf(s->lam[c]) + eps);
            rot = rot * mu_xz / fabsf(mu_xz + eps);
            out[p] = div * uc * m0 * 1e-3f;
            out[npts + p] = rot * uc * m0 * 1e-3f;
        } else {
            const REAL v[2] = {s->Vx[c], s->Vz[c]};
            for (int n = 0; n < 2; n++) {
                if (kind == 1) {
                    out[n * npts + p] = (float)(v[n] * (REAL)uc * (REAL)m0);
                } else {
                    acc[n * npts + p] = (float)((REAL)acc[n * npts + p] + v[n] * (REAL)uc * (REAL)m0 * (REAL)s->dt);
                    out[n * npts + p] = acc[n * npts + p];
                }
            }
        }
    }
}

void o2d_receivers(o2d_t *s, int nst, const int *ist, const int *kst, float *vel, double *disp)
{
    for (int n = 0; n < nst; n++) {
        int i = ist[n], k = kst[n];
        REAL vx = s->Vx[J2(s, k, i)] + s->Vx[J2(s, k, i - 1)];
        REAL vz = s->Vz[J2(s, k, i)] + s->Vz[J2(s, k - 1, i)];
        if (disp) {
            disp[2 * n + 0] += (double)vx * 0.5 * (double)s->dt;
            disp[2 * n + 1] -= (double)vz * 0.5 * (double)s->dt;
        }
        if (vel) {
            vel[2 * n + 0] = (float)(vx / 2);
            vel[2 * n + 1] = (float)(-vz / 2);
        }
    }
}

void o2d_vmax(o2d_t *s, float *out2)
{
    float xm = 0, zm = 0;
    for (int i = s->na + 1; i <= s->nx - s->na; i++) {
        int k = s->kob[J1(i)] + 1;
        float a = fabsf((float)s->Vx[J2(s, k, i)]), b = fabsf((float)s->Vz[J2(s, k, i)]);
        if (a > xm) xm = a;
        if (b > zm) zm = b;
    }
    out2[0] = xm; out2[1] = zm;
}

void o2d_step(o2d_t *s, int nsrc, const int *isrc, const int *ksrc, const double *m3, const double *sdrop)
{
    o2d_update_stress(s);
    o2d_absorb_stress(s);
    if (nsrc > 0) o2d_stressglut(s, nsrc, isrc, ksrc, m3, sdrop);
    o2d_update_vel(s);
    o2d_absorb_vel(s);
}
#undef B

/* ------------------------------------------------------------------ */
/* SH (src/swpc_sh)                                                    */
/* ------------------------------------------------------------------ */
typedef struct {
    int nx, nz, na, nm;
    double dx, dz;
    float dt;
    REAL *Vy, *Syz, *Sxy;
    REAL *Ryz, *Rxy;                       /* (m, box) */
    const float *rho, *mu, *taus;
    const int *kfs_top, *kfs_bot, *kob_top, *kob_bot, *kob;   /* over i in the memory box */
    const float *gxc, *gxe, *gzc, *gze;
    REAL *ade[4];                          /* axVy azVy axSxy azSyz */
    const float *ts;
} osh_t;

/* 2 a b / (a + b + epsilon(1.0)) in real(SP) (src/swpc_sh/m_kernel.f90:206-209) */
static inline float harm2(float a, float b)
{
    const float epsl = 1.1920928955078125e-07f;
    return 2 * a * b / (a + b + epsl);
}

#define BS(F, dk, di) s->F[J2(s, k + (dk), i + (di))]

/* src/swpc_sh/m_kernel.f90:88-139 */
void osh_update_vel(osh_t *s)
{
    const int na = s->na, nx = s->nx, nz = s->nz;
    const coef_t cx = mkcoef(s->dx), cz = mkcoef(s->dz);
    const REAL dt = (REAL)s->dt;
#pragma omp parallel for schedule(static)
    for (int i = na + 1; i <= nx - na; i++)
        for (int k = 1; k <= nz - na; k++) {
            int isign = order_sign(k, s->kfs_top[J1(i)], s->kfs_bot[J1(i)], s->kob_top[J1(i)], s->kob_bot[J1(i)]);
            REAL e40x = cx.rc40 + isign * cx.rd40, e41x = cx.rc41 + isign * cx.rd41;
            REAL e40z = cz.rc40 + isign * cz.rd40, e41z = cz.rc41 + isign * cz.rd41;
            REAL dzSyz = (BS(Syz, 0, 0) - BS(Syz, -1, 0)) * e40z - (BS(Syz, 1, 0) - BS(Syz, -2, 0)) * e41z;
            REAL dxSxy = (BS(Sxy, 0, 0) - BS(Sxy, 0, -1)) * e40x - (BS(Sxy, 0, 1) - BS(Sxy, 0, -2)) * e41x;
            size_t c = J2(s, k, i);
            float by = 1.0f / s->rho[c];
            s->Vy[c] = s->Vy[c] + by * (dxSxy + dzSyz) * dt;
        }
}

/* src/swpc_sh/m_kernel.f90:141-223 */
void osh_update_stress(osh_t *s)
{
    const int na = s->na, nx = s->nx, nz = s->nz, nm = s->nm;
    const coef_t cx = mkcoef(s->dx), cz = mkcoef(s->dz);
    const zener_t z = mkzener(nm, s->ts, s->dt);
    const REAL dt = (REAL)s->dt;
#pragma omp parallel for schedule(static)
    for (int i = na + 1; i <= nx - na; i++)
        for (int k = 1; k <= nz - na; k++) {
            int isign = order_sign(k, s->kfs_top[J1(i)], s->kfs_bot[J1(i)], s->kob_top[J1(i)], s->kob_bot[J1(i)]);
            REAL e40x = cx.rc40 + isign * cx.rd40, e41x = cx.rc41 + isign * cx.rd41;
            REAL e40z = cz.rc40 + isign * cz.rd40, e41z = cz.rc41 + isign * cz.rd41;
            REAL dxVy = (BS(Vy, 0, 1) - BS(Vy, 0, 0)) * e40x - (BS(Vy, 0, 2) - BS(Vy, 0, -1)) * e41x;
            REAL dzVy = (BS(Vy, 1, 0) - BS(Vy, 0, 0)) * e40z - (BS(Vy, 2, 0) - BS(Vy, -1, 0)) * e41z;
            size_t c = J2(s, k, i);
            float nnn = s->mu[c], pnn = s->mu[J2(s, k + 1, i)], npn = s->mu[J2(s, k, i + 1)];
            REAL mu_xy = harm2(nnn, npn), mu_yz = harm2(nnn, pnn);
            REAL taus1 = s->taus[c];
            REAL f_Ryz = mu_yz * taus1 * dzVy;
            REAL f_Rxy = mu_xy * taus1 * dxVy;
            REAL Ryz_n = 0, Rxy_n = 0;
            for (int m = 0; m < nm; m++) {
                size_t r = (size_t)m + (size_t)nm * c;
                s->Ryz[r] = z.c1[m] * s->Ryz[r] - z.c2[m] * f_Ryz * dt;
                s->Rxy[r] = z.c1[m] * s->Rxy[r] - z.c2[m] * f_Rxy * dt;
                Ryz_n += z.d1[m] * s->Ryz[r];
                Rxy_n += z.d1[m] * s->Rxy[r];
            }
            REAL taus_plus1 = 1 + taus1 * (1 + z.d2);
            s->Syz[c] = s->Syz[c] + (mu_yz * taus_plus1 * dzVy + Ryz_n) * dt;
            s->Sxy[c] = s->Sxy[c] + (mu_xy * taus_plus1 * dxVy + Rxy_n) * dt;
        }
}

static inline int kbeg_ash(const osh_t *s, int i)
{
    if (i <= s->na || s->nx - s->na + 1 <= i) return 1;
    return s->nz - s->na + 1;
}

/* src/swpc_sh/m_absorb_p.f90:174-187.  The reference takes gzc(1:2, i) in
   the velocity update, reading past the profile when nx > nz; the z
   profile is taken at k here (its evident intent; gzc(3:4, k) in the same
   loop). */
void osh_absorb_vel(osh_t *s)
{
    const REAL r20x = (REAL)(float)(1.0 / s->dx), r20z = (REAL)(float)(1.0 / s->dz);
    const REAL dt = (REAL)s->dt;
    REAL *axSxy = s->ade[2], *azSyz = s->ade[3];
#pragma omp parallel for schedule(dynamic)
    for (int i = 1; i <= s->nx; i++)
        for (int k = kbeg_ash(s, i); k <= s->nz; k++) {
            REAL dzSyz = (BS(Syz, 0, 0) - BS(Syz, -1, 0)) * r20z;
            REAL dxSxy = (BS(Sxy, 0, 0) - BS(Sxy, 0, -1)) * r20x;
            size_t c = J2(s, k, i);
            float by = 1.0f / s->rho[c];
            /* src/swpc_sh/m_absorb_p.f90:180-181 take gzc(1:2, i); past the
               profile's bounds (nx > nz) the profile is taken at k */
            const int kz = s->nx <= s->nz ? i : k;
            s->Vy[c] = s->Vy[c] + by * (G(s->gxc, 1, i) * dxSxy + G(s->gzc, 1, kz) * dzSyz
                                       + G(s->gxc, 2, i) * axSxy[c] + G(s->gzc, 2, kz) * azSyz[c]) * dt;
            axSxy[c] = G(s->gxc, 3, i) * axSxy[c] + G(s->gxc, 4, i) * dxSxy * dt;
            azSyz[c] = G(s->gzc, 3, k) * azSyz[c] + G(s->gzc, 4, k) * dzSyz * dt;
        }
}

/* src/swpc_sh/m_absorb_p.f90:262-275: ADE values advanced before the
   stresses, which use the advanced values */
void osh_absorb_stress(osh_t *s)
{
    const REAL r20x = (REAL)(float)(1.0 / s->dx), r20z = (REAL)(float)(1.0 / s->dz);
    const REAL dt = (REAL)s->dt;
    REAL *axVy = s->ade[0], *azVy = s->ade[1];
#pragma omp parallel for schedule(dynamic)
    for (int i = 1; i <= s->nx; i++)
        for (int k = kbeg_ash(s, i); k <= s->nz; k++) {
            REAL dxVy = (BS(Vy, 0, 1) - BS(Vy, 0, 0)) * r20x;
            REAL dzVy = (BS(Vy, 1, 0) - BS(Vy, 0, 0)) * r20z;
            size_t c = J2(s, k, i);
            float nnn = s->mu[c], pnn = s->mu[J2(s, k + 1, i)], npn = s->mu[J2(s, k, i + 1)];
            REAL muxy = harm2(nnn, npn), muyz = harm2(nnn, pnn);
            axVy[c] = G(s->gxe, 3, i) * axVy[c] + G(s->gxe, 4, i) * dxVy * dt;
            azVy[c] = G(s->gze, 3, k) * azVy[c] + G(s->gze, 4, k) * dzVy * dt;
            s->Syz[c] = s->Syz[c] + muyz * (G(s->gze, 1, k) * dzVy + G(s->gze, 2, k) * azVy[c]) * dt;
            s->Sxy[c] = s->Sxy[c] + muxy * (G(s->gxe, 1, i) * dxVy + G(s->gxe, 2, i) * axVy[c]) * dt;
        }
}

/* src/swpc_sh/m_source.f90:575-588; m2 = (myz, mxy). */
void osh_stressglut(osh_t *s, int nsrc, const int *isrc, const int *ksrc, const double *m2, const double *sdrop)
{
    for (int n = 0; n < nsrc; n++) {
        int ii = isrc[n], kk = ksrc[n];
        REAL sd = (REAL)sdrop[n];
        REAL myz = (REAL)m2[2 * n], mxy = (REAL)m2[2 * n + 1];
        s->Sxy[J2(s, kk, ii)] -= mxy * sd / 2;
        s->Sxy[J2(s, kk, ii - 1)] -= mxy * sd / 2;
        s->Syz[J2(s, kk, ii)] -= myz * sd / 2;
        s->Syz[J2(s, kk - 1, ii)] -= myz * sd / 2;
    }
}

/* src/swpc_sh/m_source.f90:614-621; fdt = stime*dt/(dx dz). */
void osh_bodyforce(osh_t *s, int nsrc, const int *isrc, const int *ksrc, const double *fy, const double *fdt)
{
    for (int n = 0; n < nsrc; n++) {
        size_t c = J2(s, ksrc[n], isrc[n]);
        s->Vy[c] += (REAL)fy[n] / (REAL)s->rho[c] * (REAL)fdt[n];
    }
}

/* src/swpc_sh/m_wav.f90:163-200 (v and u): Vy at the station cell; u sums Vy dt */
void osh_receivers(osh_t *s, int nst, const int *ist, const int *kst, float *vel, double *disp)
{
    for (int n = 0; n < nst; n++) {
        REAL vy = s->Vy[J2(s, kst[n], ist[n])];
        if (disp) disp[n] += (double)vy * (double)s->dt;
        if (vel) vel[n] = (float)vy;
    }
}

/* snap__write products (src/swpc_sh/m_snap.f90:415-520): kind 1 'v' is
   Vy x UC x M0 (wbuf_xz_v), kind 2 'u' advances the real(SP) sum
   buf_u + Vy x UC x M0 x dt (wbuf_xz_u) and returns it. */
void osh_snapshot(osh_t *s, int kind, int npts, const int *ik, float uc, float m0, float *acc, float *out)
{
    for (int p = 0; p < npts; p++) {
        const REAL vy = s->Vy[J2(s, ik[2 * p + 1], ik[2 * p])];
        if (kind == 1) {
            out[p] = (float)(vy * (REAL)uc * (REAL)m0);
        } else {
            acc[p] = (float)((REAL)acc[p] + vy * (REAL)uc * (REAL)m0 * (REAL)s->dt);
            out[p] = acc[p];
        }
    }
}

/* src/swpc_sh/m_kernel.f90:225-240 */
float osh_vmax(osh_t *s)
{
    float m = 0;
    for (int i = s->na + 1; i <= s->nx - s->na; i++) {
        float a = fabsf((float)s->Vy[J2(s, s->kob[J1(i)] + 1, i)]);
        if (a > m) m = a;
    }
    return m;
}
#undef BS

/* ------------------------------------------------------------------ */
/* helpers                                                             */
/* ------------------------------------------------------------------ */

int oracle_real_bytes(void) { return (int)sizeof(REAL); }

int oracle_num_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* Zener constants for inspection by tests. */
void oracle_zener(int nm, const float *ts, float dt, double *c1, double *c2, double *d1, double *d2)
{
    zener_t z = mkzener(nm, ts, dt);
    for (int m = 0; m < nm; m++) { c1[m] = z.c1[m]; c2[m] = z.c2[m]; d1[m] = z.d1[m]; }
    *d2 = z.d2;
}

/* ------------------------------------------------------------------ */
/* Learned absorbing boundary                                          */
/* ------------------------------------------------------------------ */
/* This project's own absorbing boundary (the reference has none like it;
   it takes the place of absorb_p__update_* on a grid without PML, na = 0),
   restated here independently of the HIP kernels as their checker.
   Before each half-step the ghost cells beyond every absorbing face (x-,
   x+, y-, y+ and the bottom z+) are written with values that a small MLP
   per face and half-step predicts; the interior update then runs over the
   whole grid.  Net 2*face + phase: phase 0 fills the velocity ghosts read
   by the stress update, phase 1 the stress ghosts read by the velocity
   update.

   Column features (FP32): the nine fields Vx Vy Vz Sxx Syy Szz Syz Sxz
   Sxy at strip cells d = 1..6 (d = 1 outermost), index fi*6 + d-1, with
   the stresses divided by the impedance rho*vp of cell d = 1; then the
   net's previous ghost values (stresses / rho*vp, zero-padded to 5); N =
   sqrt(sum of the squares of these 59 / 59), formed on the values scaled
   by their largest magnitude (0 when that is below 1e-30); they are
   divided by N; then
   vs/vp and vp*dt/h (h the spacing normal to the face).  Net: y = Ws x +
   W3 tanh(W2 tanh(W1 x + b1) + b2) + b3; ghost value q = y_q * N (* rho*vp
   for a stress).  N = 0 and air columns (lam + 2 mu = 0 at d = 1) write
   zeros. */
#define LAB_NS 6
#define LAB_NF 54
#define LAB_NG 5
#define LAB_NSC 59
#define LAB_NIN 61
#define LAB_NH 32
#define LAB_NOUT 5
#define LAB_W1 0
#define LAB_B1 (LAB_W1 + LAB_NH * LAB_NIN)
#define LAB_W2 (LAB_B1 + LAB_NH)
#define LAB_B2 (LAB_W2 + LAB_NH * LAB_NH)
#define LAB_W3 (LAB_B2 + LAB_NH)
#define LAB_B3 (LAB_W3 + LAB_NOUT * LAB_NH)
#define LAB_WS (LAB_B3 + LAB_NOUT)
#define LAB_NP (LAB_WS + LAB_NOUT * LAB_NIN)
#define LAB_REC 72

/* ghost targets per net: (field id 0..8 = Vx Vy Vz Sxx Syy Szz Syz Sxz
   Sxy, ghost depth h); the sets are the ghost cells that the 4th-order
   stencils of o3d_update_stress (phase 0) and o3d_update_vel (phase 1)
   read beyond each face. */
static const int lab_tgt[10][LAB_NOUT][2] = {
    {{0, 2}, {0, 1}, {1, 1}, {2, 1}, {0, 0}},   /* x- V */
    {{3, 1}, {8, 2}, {8, 1}, {7, 2}, {7, 1}},   /* x- S */
    {{0, 1}, {1, 1}, {1, 2}, {2, 1}, {2, 2}},   /* x+ V */
    {{3, 1}, {3, 2}, {8, 1}, {7, 1}, {0, 0}},   /* x+ S */
    {{1, 2}, {1, 1}, {0, 1}, {2, 1}, {0, 0}},   /* y- V */
    {{4, 1}, {8, 2}, {8, 1}, {6, 2}, {6, 1}},   /* y- S */
    {{1, 1}, {0, 1}, {0, 2}, {2, 1}, {2, 2}},   /* y+ V */
    {{4, 1}, {4, 2}, {8, 1}, {6, 1}, {0, 0}},   /* y+ S */
    {{2, 1}, {0, 1}, {0, 2}, {1, 1}, {1, 2}},   /* z+ V */
    {{5, 1}, {5, 2}, {7, 1}, {6, 1}, {0, 0}},   /* z+ S */
};
static const int lab_nout_tab[10] = {4, 5, 5, 4, 4, 5, 5, 4, 5, 4};

int oracle_lab_nparam(void) { return LAB_NP; }
int oracle_lab_nout(int net) { return lab_nout_tab[net]; }
void oracle_lab_target(int net, int q, int *field, int *h) { *field = lab_tgt[net][q][0]; *h = lab_tgt[net][q][1]; }

static REAL *lab_field(o3d_t *s, int f)
{
    REAL *t[9] = {s->Vx, s->Vy, s->Vz, s->Sxx, s->Syy, s->Szz, s->Syz, s->Sxz, s->Sxy};
    return t[f];
}

/* Columns of face f over a domain of nxs x nys x nzs cells: (a, b)
   extents, a fastest in the column number (a-1) + nA*(b-1). */
static void lab_extent(int f, int nxs, int nys, int nzs, int *nA, int *nB)
{
    if (f < 2) { *nA = nys; *nB = nzs; }
    else if (f < 4) { *nA = nxs; *nB = nzs; }
    else { *nA = nxs; *nB = nys; }
}

/* Domain cell (i, j, k) of column (a, b) at depth d: d >= 1 is strip cell
   d (d = 1 outermost), d = 1 - h the ghost cell at depth h. */
static void lab_cell(int f, int a, int b, int d, int nxs, int nys, int nzs, int *i, int *j, int *k)
{
    int n = f < 2 ? nxs : (f < 4 ? nys : nzs);
    int lo = f == 0 || f == 2;   /* x- and y-; x+, y+ and the bottom z+ count inward from n */
    int p = lo ? d : n + 1 - d;
    if (f < 2) { *i = p; *j = a; *k = b; }
    else if (f < 4) { *i = a; *j = p; *k = b; }
    else { *i = a; *j = b; *k = p; }
}

/* Medium of the column (cell d = 1): vs/vp, vp*dt/h, rho*vp. */
static void lab_medium(o3d_t *s, const float *rho, int f, size_t c0, float *m0, float *m1, float *im)
{
    float l2m = s->lam[c0] + 2.0f * s->mu[c0];
    float h = (float)(f < 2 ? s->dx : (f < 4 ? s->dy : s->dz));
    *m0 = 0.0f; *m1 = 0.0f; *im = 1.0f;
    if (l2m > 0.0f) {
        float vp = sqrtf(l2m / rho[c0]);
        float vs = sqrtf(fmaxf(s->mu[c0], 0.0f) / rho[c0]);
        *im = rho[c0] * vp;
        *m0 = vs / vp;
        *m1 = vp * s->dt / h;
    }
}

/* Features of column (a, b) of face f of the domain at offset (oi, oj, 0)
   with previous ghost values prev[5]; returns N. */
static float lab_features(o3d_t *s, const float *rho, int f, int a, int b, int oi, int oj, int nxs, int nys, int nzs,
                          const float *prev, float *x, float *imp)
{
    int i, j, k;
    lab_cell(f, a, b, 1, nxs, nys, nzs, &i, &j, &k);
    float m0, m1, im;
    lab_medium(s, rho, f, I3(s, k, i + oi, j + oj), &m0, &m1, &im);
    float amax = 0.0f;
    for (int fi = 0; fi < 9; fi++) {
        REAL *F = lab_field(s, fi);
        for (int d = 1; d <= LAB_NS; d++) {
            lab_cell(f, a, b, d, nxs, nys, nzs, &i, &j, &k);
            float v = (float)F[I3(s, k, i + oi, j + oj)];
            if (fi >= 3) v = v / im;
            x[fi * LAB_NS + d - 1] = v;
            amax = fmaxf(amax, fabsf(v));
        }
    }
    for (int q = 0; q < LAB_NG; q++) {
        x[LAB_NF + q] = prev[q];
        amax = fmaxf(amax, fabsf(prev[q]));
    }
    /* N = sqrt(mean square) on values scaled by the largest magnitude;
       columns whose largest magnitude is below 1e-30 count as empty */
    float n = 0.0f;
    if (amax > 1e-30f) {
        float ia = 1.0f / amax, s2 = 0.0f;
        for (int q = 0; q < LAB_NSC; q++) {
            float u = x[q] * ia;
            s2 += u * u;
        }
        n = amax * sqrtf(s2 / (float)LAB_NSC);
        for (int q = 0; q < LAB_NSC; q++) x[q] = x[q] / n;
    }
    x[LAB_NSC] = m0;
    x[LAB_NSC + 1] = m1;
    *imp = im;
    return n;
}

/* BF16 of an FP32 value, nearest even (finite inputs). */
static float lab_bf16(float f)
{
    unsigned u;
    memcpy(&u, &f, 4);
    u += 0x7fffu + ((u >> 16) & 1u);
    u &= 0xffff0000u;
    memcpy(&f, &u, 4);
    return f;
}

/* One net, FP32, sequential sums in index order.  The two hidden layers
   take their inputs rounded to BF16 (the features for layer 1, tanh of
   layer 1 for layer 2), as the matrix cores of the time step do; the
   output layer uses the FP32 features and FP32 layer-2 activations. */
void oracle_lab_forward(const float *p, const float *x, float *y)
{
    float h1[LAB_NH], h2[LAB_NH], xb[LAB_NIN];
    for (int q = 0; q < LAB_NIN; q++) xb[q] = lab_bf16(x[q]);
    for (int h = 0; h < LAB_NH; h++) {
        float z = 0.0f;
        for (int q = 0; q < LAB_NIN; q++) z += p[LAB_W1 + h * LAB_NIN + q] * xb[q];
        h1[h] = lab_bf16(tanhf(z + p[LAB_B1 + h]));
    }
    for (int h = 0; h < LAB_NH; h++) {
        float z = 0.0f;
        for (int g = 0; g < LAB_NH; g++) z += p[LAB_W2 + h * LAB_NH + g] * h1[g];
        h2[h] = tanhf(z + p[LAB_B2 + h]);
    }
    for (int o = 0; o < LAB_NOUT; o++) {
        float z = p[LAB_B3 + o];
        for (int h = 0; h < LAB_NH; h++) z += p[LAB_W3 + o * LAB_NH + h] * h2[h];
        for (int q = 0; q < LAB_NIN; q++) z += p[LAB_WS + o * LAB_NIN + q] * x[q];
        y[o] = z;
    }
}

/* Ghost fill of one phase over the whole grid: params holds the ten nets,
   LAB_NP floats each (net 2*face + phase); the previous ghost values are
   read back from the ghost cells before they are overwritten. */
void o3d_lab_fill(o3d_t *s, int phase, const float *params, const float *rho, float sponge)
{
    /* strip sponge factor of cell d = 1..6, formed in double, rounded to float */
    float sp[LAB_NS];
    for (int d = 1; d <= LAB_NS; d++) {
        double r = (double)(LAB_NS + 1 - d) / LAB_NS;
        sp[d - 1] = (float)(1.0 - (double)sponge * r * r);
    }
    for (int f = 0; f < 5; f++) {
        int net = 2 * f + phase, nA, nB, nout = lab_nout_tab[net];
        lab_extent(f, s->nx, s->ny, s->nz, &nA, &nB);
#pragma omp parallel for schedule(static)
        for (int b = 1; b <= nB; b++)
            for (int a = 1; a <= nA; a++) {
                float x[LAB_NIN], y[LAB_NOUT], prev[LAB_NG], im, m0, m1;
                size_t gc[LAB_NOUT];
                int i, j, k;
                lab_cell(f, a, b, 1, s->nx, s->ny, s->nz, &i, &j, &k);
                lab_medium(s, rho, f, I3(s, k, i, j), &m0, &m1, &im);
                for (int q = 0; q < LAB_NG; q++) {
                    prev[q] = 0.0f;
                    if (q >= nout) continue;
                    int fi = lab_tgt[net][q][0], h = lab_tgt[net][q][1];
                    lab_cell(f, a, b, 1 - h, s->nx, s->ny, s->nz, &i, &j, &k);
                    gc[q] = I3(s, k, i, j);
                    float v = (float)lab_field(s, fi)[gc[q]];
                    prev[q] = fi >= 3 ? v / im : v;
                }
                float N = lab_features(s, rho, f, a, b, 0, 0, s->nx, s->ny, s->nz, prev, x, &im);
                int live = N > 0.0f && m1 > 0.0f;   /* air columns keep zero ghosts */
                if (live) oracle_lab_forward(params + (size_t)net * LAB_NP, x, y);
                for (int q = 0; q < nout; q++) {
                    int fi = lab_tgt[net][q][0];
                    float v = live ? y[q] * N : 0.0f;
                    if (fi >= 3) v = v * im;
                    lab_field(s, fi)[gc[q]] = (REAL)v;
                }
                /* the sponge on this phase's own fields of the column's strip */
                if (sponge > 0.0f)
                    for (int fi = phase == 0 ? 0 : 3; fi < (phase == 0 ? 3 : 9); fi++)
                        for (int d = 1; d <= LAB_NS; d++) {
                            lab_cell(f, a, b, d, s->nx, s->ny, s->nz, &i, &j, &k);
                            lab_field(s, fi)[I3(s, k, i, j)] *= (REAL)sp[d - 1];
                        }
            }
    }
}

/* Training samples of face f, phase `phase`, for the domain of nxs x nys x
   nzs cells at offset (oi, oj, 0) of this (wider) grid.  prev[nA*nB][5]
   holds each column's ghost values of the previous call (in: features;
   out: this step's values).  Every column with N > 0 gives one LAB_REC
   record: x[61], the ghost values / N (stresses / rho*vp) in [61, 66),
   then the column number, N and rho*vp; air columns give none.  Returns
   the number written. */
int o3d_lab_gather(o3d_t *s, int phase, int f, const float *rho, int oi, int oj, int nxs, int nys, int nzs,
                   float *prev, float *out)
{
    int net = 2 * f + phase, nA, nB, n = 0, nout = lab_nout_tab[net];
    lab_extent(f, nxs, nys, nzs, &nA, &nB);
    for (int b = 1; b <= nB; b++)
        for (int a = 1; a <= nA; a++) {
            float x[LAB_NIN], t[LAB_NOUT], im, m0, m1;
            float *pv = prev + (size_t)((a - 1) + nA * (b - 1)) * LAB_NG;
            int i, j, k;
            lab_cell(f, a, b, 1, nxs, nys, nzs, &i, &j, &k);
            lab_medium(s, rho, f, I3(s, k, i + oi, j + oj), &m0, &m1, &im);
            for (int q = 0; q < LAB_NOUT; q++) {
                t[q] = 0.0f;
                if (q >= nout) continue;
                int fi = lab_tgt[net][q][0], h = lab_tgt[net][q][1];
                lab_cell(f, a, b, 1 - h, nxs, nys, nzs, &i, &j, &k);
                float v = (float)lab_field(s, fi)[I3(s, k, i + oi, j + oj)];
                t[q] = fi >= 3 ? v / im : v;
            }
            float N = lab_features(s, rho, f, a, b, oi, oj, nxs, nys, nzs, pv, x, &im);
            for (int q = 0; q < LAB_NG; q++) pv[q] = t[q];
            if (!(N > 0.0f) || !(m1 > 0.0f)) continue;
            float *o = out + (size_t)n * LAB_REC;
            for (int q = 0; q < LAB_REC; q++) o[q] = 0.0f;
            for (int q = 0; q < LAB_NIN; q++) o[q] = x[q];
            for (int q = 0; q < LAB_NOUT; q++) o[LAB_NIN + q] = t[q] / N;
            o[LAB_NIN + 5] = (float)((a - 1) + nA * (b - 1));
            o[LAB_NIN + 6] = N;
            o[LAB_NIN + 7] = im;
            n++;
        }
    return n;
}
