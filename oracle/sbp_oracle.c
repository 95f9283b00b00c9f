/*
 * sbp_oracle.c -- CPU restatement of the reference SBP-SAT hot path.
 *
 * TEST INFRASTRUCTURE ONLY (see sbp_oracle.h).  Built twice from this one
 * source by oracle/Makefile:
 *   liborc.so      strict IEEE (-O2 -ffp-contract=off, no fast-math): the
 *                  parity oracle for the HIP path;
 *   libcpubase.so  the reference's opt flags (-O3 -ffast-math, code/Makefile:17)
 *                  with pthreads: the timed CPU baseline of the benchmark.
 *
 * Every function cites the reference lines it restates.  The region split of
 * partitioned_rhs/rhs.h:363-398 (ll, li, lr, il, ii, ir, rl, ri, rr) is kept
 * as a per-direction choice left / interior / right of the closure, which
 * selects exactly the same stencil for every grid point.
 */
#define _POSIX_C_SOURCE 200809L
#include "sbp_oracle.h"

#include <float.h>
#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

/* PETSC_PI as used by wave_eq_rhs.h:54,63 and wave_eq_sim.cpp:226-228. */
#define ORC_PI 3.14159265358979323846264338327950288419716939937510582097494459

/* ------------------------------------------------------------------------ */
/* Operator tables                                                            */
/* ------------------------------------------------------------------------ */
/* Coefficients of the central SBP first-derivative operators, their norms
 * and inverse norms.  The rational numbers are the operators' definitions and
 * must be identical to code/include/sbpops/D1_central.h:243-273 (stencils),
 * H_central.h:193-212 (closure_quad) and HI_central.h:128-147
 * (closure_invquad).  They are written as the same decimal quotients so the
 * compiled doubles are bit-identical. */
int orc_get_ops(int order, orc_ops *op)
{
  memset(op, 0, sizeof(*op));
  op->order = order;
  if (order == 2) {
    op->r = 1; op->cls = 1; op->w = 2;
    op->inner[0] = -1./2; op->inner[1] = 0; op->inner[2] = 1./2;
    op->Q[0][0] = -1.; op->Q[0][1] = 1;
    op->hq[0] = 1./2;
    op->hqinv[0] = 2.;
    return 0;
  }
  if (order == 4) {
    static const double q4[4][6] = {
      {-24./17, 59./34, -4./17, -3./34, 0, 0},
      {-1./2, 0, 1./2, 0, 0, 0},
      {4./43, -59./86, 0, 59./86, -4./43, 0},
      {3./98, 0, -59./98, 0, 32./49, -4./49}};
    static const double in4[5] = {1./12, -2./3, 0., 2./3, -1./12};
    static const double h4[4] = {17./48, 59./48, 43./48, 49./48};
    static const double hi4[4] = {48./17, 48./59, 48./43, 48./49};
    op->r = 2; op->cls = 4; op->w = 6;
    memcpy(op->inner, in4, sizeof(in4));
    for (int a = 0; a < 4; a++)
      for (int b = 0; b < 6; b++) op->Q[a][b] = q4[a][b];
    memcpy(op->hq, h4, sizeof(h4));
    memcpy(op->hqinv, hi4, sizeof(hi4));
    return 0;
  }
  if (order == 6) {
    static const double q6[6][9] = {
      {-21600./13649, 104009./54596, 30443./81894, -33311./27298, 16863./27298, -15025./163788, 0, 0, 0},
      {-104009./240260, 0, -311./72078, 20229./24026, -24337./48052, 36661./360390, 0, 0, 0},
      {-30443./162660, 311./32532, 0, -11155./16266, 41287./32532, -21999./54220, 0, 0, 0},
      {33311./107180, -20229./21436, 485./1398, 0, 4147./21436, 25427./321540, 72./5359, 0, 0},
      {-16863./78770, 24337./31508, -41287./47262, -4147./15754, 0, 342523./472620, -1296./7877, 144./7877, 0},
      {15025./525612, -36661./262806, 21999./87602, -25427./262806, -342523./525612, 0, 32400./43801, -6480./43801, 720./43801}};
    static const double in6[7] = {-1./60, 3./20, -3./4, 0, 3./4, -3./20, 1./60};
    static const double h6[6] = {13649./43200, 12013./8640, 2711./4320, 5359./4320, 7877./8640, 43801./43200};
    static const double hi6[6] = {43200./13649, 8640./12013, 4320./2711, 4320./5359, 8640./7877, 43200./43801};
    op->r = 3; op->cls = 6; op->w = 9;
    memcpy(op->inner, in6, sizeof(in6));
    for (int a = 0; a < 6; a++)
      for (int b = 0; b < 9; b++) op->Q[a][b] = q6[a][b];
    memcpy(op->hq, h6, sizeof(h6));
    memcpy(op->hqinv, hi6, sizeof(hi6));
    return 0;
  }
  return -1;
}

int orc_components(const sbp_problem *p)
{
  switch (p->physics) {
    case SBP_ADVECTION_1D: case SBP_ADVECTION_2D: return 1;
    case SBP_ACOUSTIC_2D: return 3;   /* [u, v, p]: wave_eq_rhs.h:12-16 */
    case SBP_REFLECTION_1D: return 2; /* [u, v]: reflection_rhs.h:8   */
  }
  return 0;
}

static int is_1d(const sbp_problem *p)
{
  return p->physics == SBP_ADVECTION_1D || p->physics == SBP_REFLECTION_1D;
}

int orc_validate(const sbp_problem *p)
{
  orc_ops op;
  if (!p || orc_get_ops(p->order, &op)) return -1;
  if (orc_components(p) == 0) return -1;
  int taps = op.w > 2 * op.r + 1 ? op.w : 2 * op.r + 1;
  int nmin = 2 * op.cls > taps ? 2 * op.cls : taps;
  if (p->nx < nmin) return -1;
  if (is_1d(p)) { if (p->ny != 1) return -1; }
  else if (p->ny < nmin) return -1;
  if (p->physics == SBP_ACOUSTIC_2D && p->material == SBP_MAT_ARRAYS &&
      (!p->rho_inv || !p->bulk)) return -1;
  /* velocity fields: advection only, and no y-velocity in 1D */
  if ((p->vel_x || p->vel_y) && p->physics != SBP_ADVECTION_1D && p->physics != SBP_ADVECTION_2D) return -1;
  if (p->vel_y && p->physics == SBP_ADVECTION_1D) return -1;
  return 0;
}

double orc_hi(int n, double lo, double hi_end) { return (n - 1) / (hi_end - lo); }

/* ------------------------------------------------------------------------ */
/* D1 application: D1_central.h:57-104 (1D) and :119-231 (2D).               */
/* kind 0 = left closure, 1 = interior, 2 = right closure.                   */
/* ------------------------------------------------------------------------ */
static inline double d1_kind(const orc_ops *op, const double *v, long s, int n, int i,
                             double hi, int kind)
{
  double u = 0;
  if (kind == 0) {               /* apply_left: sum_is Q[i][is] v[is]            */
    for (int k = 0; k < op->w; k++) u += op->Q[i][k] * v[k * s];
  } else if (kind == 1) {        /* apply_interior: sum stencil[is] v[i-r+is]   */
    for (int k = 0; k < 2 * op->r + 1; k++) u += op->inner[k] * v[(long)(i - op->r + k) * s];
  } else {                       /* apply_right: u -= Q[N-i-1][w-is-1] v[N-w+is] */
    for (int k = 0; k < op->w; k++)
      u -= op->Q[n - i - 1][op->w - k - 1] * v[(long)(n - op->w + k) * s];
  }
  return hi * u;
}

static inline int region(const orc_ops *op, int n, int i)
{
  return i < op->cls ? 0 : (i >= n - op->cls ? 2 : 1);
}

double orc_d1(const orc_ops *op, const double *v, long stride, int n, int i, double hi)
{
  return d1_kind(op, v, stride, n, i, hi, region(op, n, i));
}

/* ------------------------------------------------------------------------ */
/* Problem helpers                                                            */
/* ------------------------------------------------------------------------ */
typedef struct {
  const sbp_problem *p;
  orc_ops op;
  int C, nx, ny;
  double hix, hiy;
} orc_setup;

static void setup(const sbp_problem *p, orc_setup *s)
{
  s->p = p;
  orc_get_ops(p->order, &s->op);
  s->C = orc_components(p);
  s->nx = p->nx;
  s->ny = p->ny;
  s->hix = orc_hi(p->nx, p->xl, p->xr);
  s->hiy = is_1d(p) ? 0.0 : orc_hi(p->ny, p->yl, p->yr);
}

/* rho_inv of wave_eq_rhs.h:42-46 (x = xl + i/hi, not i*h). */
static inline double ref_rho_inv(const orc_setup *s, int i, int j)
{
  double x = s->p->xl + i / s->hix;
  double y = s->p->yl + j / s->hiy;
  return 1. / (2 + x * y);
}

/* forcing_u / forcing_v of wave_eq_rhs.h:51-64. */
static inline double ref_forcing_u(const orc_setup *s, int i, int j, double t)
{
  double x = s->p->xl + i / s->hix;
  double y = s->p->yl + j / s->hiy;
  return -(3 * ORC_PI * cos(5 * ORC_PI * t) * cos(3 * ORC_PI * x) * sin(4 * ORC_PI * y) * (x * y + 1)) / (x * y + 2);
}

static inline double ref_forcing_v(const orc_setup *s, int i, int j, double t)
{
  double x = s->p->xl + i / s->hix;
  double y = s->p->yl + j / s->hiy;
  return -(4 * ORC_PI * cos(5 * ORC_PI * t) * cos(4 * ORC_PI * y) * sin(3 * ORC_PI * x) * (x * y + 1)) / (x * y + 2);
}

static inline double mat_rinv(const orc_setup *s, int i, int j)
{
  if (s->p->material == SBP_MAT_REF_RHO) return ref_rho_inv(s, i, j);
  if (s->p->material == SBP_MAT_ARRAYS) return s->p->rho_inv[(long)j * s->nx + i];
  return 1.0;
}

static inline double mat_bulk(const orc_setup *s, int i, int j)
{
  if (s->p->material == SBP_MAT_ARRAYS) return s->p->bulk[(long)j * s->nx + i];
  return 1.0;
}

/* Advection velocities at grid point (i, j): the AppCtx functions a(i) /
 * a(i,j), b(i,j) (advection_1D_sim.cpp:19,70; advection_2D_sim.cpp:21,74-75)
 * sampled on the grid, or the constants ax, ay. */
static inline double vel_x(const orc_setup *s, int i, int j)
{
  return s->p->vel_x ? s->p->vel_x[(long)j * s->nx + i] : s->p->ax;
}

static inline double vel_y(const orc_setup *s, int i, int j)
{
  return s->p->vel_y ? s->p->vel_y[(long)j * s->nx + i] : s->p->ay;
}

/* ------------------------------------------------------------------------ */
/* Region kernels over a block [i0,i1) x [j0,j1)                             */
/* ------------------------------------------------------------------------ */

/* 1D advection: advection_{l,i,r} (advection_rhs.h:13-51): dst_i = -a(i) (D1 u)_i. */
static void adv1d_block(const orc_setup *s, const double *src, double *dst, int i0, int i1)
{
  for (int i = i0; i < i1; i++)
    dst[i] = -vel_x(s, i, 0) * orc_d1(&s->op, src, 1, s->nx, i, s->hix);
}

/* 1D reflection: reflection_{l,i,r} (reflection_rhs.h:10-47):
 * dst(i,1) = D1 src(:,0), dst(i,0) = D1 src(:,1). */
static void refl_block(const orc_setup *s, const double *src, double *dst, int i0, int i1)
{
  for (int i = i0; i < i1; i++) {
    dst[2 * i + 1] = orc_d1(&s->op, src + 0, 2, s->nx, i, s->hix);
    dst[2 * i + 0] = orc_d1(&s->op, src + 1, 2, s->nx, i, s->hix);
  }
}

/* 2D advection: advection_{ll..rr} (advection_rhs.h:156-393):
 * dst = -(a_x(i,j) D_x u + a_y(i,j) D_y u). */
static void adv2d_block(const orc_setup *s, const double *src, double *dst,
                        int i0, int i1, int j0, int j1)
{
  const int nx = s->nx, ny = s->ny;
  for (int j = j0; j < j1; j++) {
    const int ky = region(&s->op, ny, j);
    const double *row = src + (long)nx * j;
    const double *col0 = src;
    for (int i = i0; i < i1; i++) {
      const int kx = region(&s->op, nx, i);
      double dx = d1_kind(&s->op, row, 1, nx, i, s->hix, kx);
      double dy = d1_kind(&s->op, col0 + i, nx, ny, j, s->hiy, ky);
      dst[(long)nx * j + i] = -(vel_x(s, i, j) * dx + vel_y(s, i, j) * dy);
    }
  }
}

/* 2D acoustics: wave_eq_{ll..rr} (wave_eq_rhs.h:76-322) and
 * wave_eq_hom_{ll..rr} (wave_eq_hom_rhs.h:49-296):
 *   F_u = -rho^-1 D_x p (+ f_u),  F_v = -rho^-1 D_y p (+ f_v),  F_p = -(D_x u) - (D_y v).
 * With caller material arrays the pressure row carries K: F_p = -K (D_x u + D_y v). */
static void acoustic_block(const orc_setup *s, double t, const double *src, double *dst,
                           int i0, int i1, int j0, int j1)
{
  const int nx = s->nx, ny = s->ny;
  const long rs = 3L * nx; /* row stride */
  const int mat = s->p->material;
  for (int j = j0; j < j1; j++) {
    const int ky = region(&s->op, ny, j);
    const double *row = src + rs * j;
    for (int i = i0; i < i1; i++) {
      const int kx = region(&s->op, nx, i);
      const double *col = src + 3L * i;
      double dxp = d1_kind(&s->op, row + 2, 3, nx, i, s->hix, kx);
      double dyp = d1_kind(&s->op, col + 2, rs, ny, j, s->hiy, ky);
      double dxu = d1_kind(&s->op, row + 0, 3, nx, i, s->hix, kx);
      double dyv = d1_kind(&s->op, col + 1, rs, ny, j, s->hiy, ky);
      double *F = dst + rs * j + 3L * i;
      if (mat == SBP_MAT_UNIT) {
        F[0] = -dxp;
        F[1] = -dyp;
        F[2] = -dxu - dyv;
      } else if (mat == SBP_MAT_REF_RHO) {
        double ri = ref_rho_inv(s, i, j);
        F[0] = -ri * dxp + ref_forcing_u(s, i, j, t);
        F[1] = -ri * dyp + ref_forcing_v(s, i, j, t);
        F[2] = -dxu - dyv;
      } else {
        double ri = mat_rinv(s, i, j), K = mat_bulk(s, i, j);
        F[0] = -ri * dxp;
        F[1] = -ri * dyp;
        F[2] = -K * (dxu + dyv);
      }
    }
  }
}

/* ------------------------------------------------------------------------ */
/* Boundary conditions on the part of the boundary inside a block            */
/* ------------------------------------------------------------------------ */

/* SAT_bc_l / SAT_bc_r (advection_rhs.h:95-119), tau from the velocity at
 * the boundary point itself.  The right SAT is written with the correct
 * inverse weight closure_invquad[0]; the reference reads
 * closure_invquad[nx-1] there but multiplies it by tau = 0 for a(nx-1) >= 0. */
static void adv1d_bc(const orc_setup *s, const double *src, double *dst, int i0, int i1)
{
  const int nx = s->nx;
  const double hq = s->hix * s->op.hqinv[0];
  if (i0 == 0) {
    double a_l = vel_x(s, 0, 0);
    double tau = -0.5 * (a_l + fabs(a_l));
    dst[0] += 0.5 * tau * (hq * src[0]);
  }
  if (i1 == nx) {
    double a_r = vel_x(s, nx - 1, 0);
    double tau = 0.5 * (a_r - fabs(a_r));
    dst[nx - 1] += 0.5 * tau * (hq * src[nx - 1]);
  }
}

/* proj_dirichlet_bc_{l,r} (reflection_rhs.h:108-120): u_t = 0 at both ends. */
static void refl_bc(const orc_setup *s, double *dst, int i0, int i1)
{
  if (i0 == 0) dst[0] = 0.0;
  if (i1 == s->nx) dst[2 * (s->nx - 1)] = 0.0;
}

/* SAT_bc_{west,south,east,north} (advection_rhs.h:461-528), applied in the
 * bc_serial order W, S, E, N (boundary_conditions.h:84-87); each boundary
 * point takes tau from the velocity at that point (a_x on W/E, a_y on S/N).
 * The south SAT uses the y-spacing and closure_invquad[0]; the reference
 * uses hi[0]*closure_invquad[i] there, multiplied by tau_s = 0 for b <= 0. */
static void adv2d_bc(const orc_setup *s, const double *src, double *dst,
                     int i0, int i1, int j0, int j1)
{
  const int nx = s->nx, ny = s->ny;
  const double hqx = s->hix * s->op.hqinv[0], hqy = s->hiy * s->op.hqinv[0];
  if (i0 == 0)
    for (int j = j0; j < j1; j++) {
      double a_w = vel_x(s, 0, j);
      double tau = -0.5 * (a_w + fabs(a_w));
      dst[(long)nx * j] += 0.5 * tau * (hqx * src[(long)nx * j]);
    }
  if (j0 == 0)
    for (int i = i0; i < i1; i++) {
      double a_s = vel_y(s, i, 0);
      double tau = -0.5 * (a_s + fabs(a_s));
      dst[i] += 0.5 * tau * (hqy * src[i]);
    }
  if (i1 == nx)
    for (int j = j0; j < j1; j++) {
      double a_e = vel_x(s, nx - 1, j);
      double tau = 0.5 * (a_e - fabs(a_e));
      long k = (long)nx * j + nx - 1;
      dst[k] += 0.5 * tau * (hqx * src[k]);
    }
  if (j1 == ny)
    for (int i = i0; i < i1; i++) {
      double a_n = vel_y(s, i, ny - 1);
      double tau = 0.5 * (a_n - fabs(a_n));
      long k = (long)nx * (ny - 1) + i;
      dst[k] += 0.5 * tau * (hqy * src[k]);
    }
}

/* Acoustic boundary terms, W, S, E, N order as bc_serial.
 *  FREE_SURFACE_REF: free_surface_bc_* of wave_eq_rhs.h:417-468 -- F_u -= hix qinv0 p (W),
 *                    F_v -= hiy qinv0 p (S), F_u += ... (E), F_v += ... (N).
 *  FREE_SURFACE:     the same terms multiplied by rho^-1 at the boundary point.
 *  WALL:             normal velocity zero, imposed on the pressure equation:
 *                    F_p -= K hix qinv0 u (W), F_p -= K hiy qinv0 v (S),
 *                    F_p += K hix qinv0 u (E), F_p += K hiy qinv0 v (N). */
static void acoustic_bc(const orc_setup *s, const double *src, double *dst,
                        int i0, int i1, int j0, int j1)
{
  const int nx = s->nx, ny = s->ny;
  const long rs = 3L * nx;
  const double hqx = s->hix * s->op.hqinv[0], hqy = s->hiy * s->op.hqinv[0];
  int bc = s->p->bc == SBP_BC_DEFAULT ? SBP_BC_FREE_SURFACE_REF : s->p->bc;
#define Q(j, i, c) src[rs * (j) + 3L * (i) + (c)]
#define F(j, i, c) dst[rs * (j) + 3L * (i) + (c)]
  if (bc == SBP_BC_FREE_SURFACE_REF) {
    if (i0 == 0)      for (int j = j0; j < j1; j++) F(j, 0, 0) -= hqx * Q(j, 0, 2);
    if (j0 == 0)      for (int i = i0; i < i1; i++) F(0, i, 1) -= hqy * Q(0, i, 2);
    if (i1 == nx)     for (int j = j0; j < j1; j++) F(j, nx - 1, 0) += hqx * Q(j, nx - 1, 2);
    if (j1 == ny)     for (int i = i0; i < i1; i++) F(ny - 1, i, 1) += hqy * Q(ny - 1, i, 2);
  } else if (bc == SBP_BC_FREE_SURFACE) {
    if (i0 == 0)  for (int j = j0; j < j1; j++) F(j, 0, 0) -= mat_rinv(s, 0, j) * (hqx * Q(j, 0, 2));
    if (j0 == 0)  for (int i = i0; i < i1; i++) F(0, i, 1) -= mat_rinv(s, i, 0) * (hqy * Q(0, i, 2));
    if (i1 == nx) for (int j = j0; j < j1; j++) F(j, nx - 1, 0) += mat_rinv(s, nx - 1, j) * (hqx * Q(j, nx - 1, 2));
    if (j1 == ny) for (int i = i0; i < i1; i++) F(ny - 1, i, 1) += mat_rinv(s, i, ny - 1) * (hqy * Q(ny - 1, i, 2));
  } else { /* WALL */
    if (i0 == 0)  for (int j = j0; j < j1; j++) F(j, 0, 2) -= mat_bulk(s, 0, j) * (hqx * Q(j, 0, 0));
    if (j0 == 0)  for (int i = i0; i < i1; i++) F(0, i, 2) -= mat_bulk(s, i, 0) * (hqy * Q(0, i, 1));
    if (i1 == nx) for (int j = j0; j < j1; j++) F(j, nx - 1, 2) += mat_bulk(s, nx - 1, j) * (hqx * Q(j, nx - 1, 0));
    if (j1 == ny) for (int i = i0; i < i1; i++) F(ny - 1, i, 2) += mat_bulk(s, i, ny - 1) * (hqy * Q(ny - 1, i, 1));
  }
#undef Q
#undef F
}

/* Right-hand side on one block: region kernels then boundary functors
 * (the *_serial + *_bc_serial pair of every rhs_serial callback). */
static void rhs_block(const orc_setup *s, double t, const double *src, double *dst,
                      int i0, int i1, int j0, int j1)
{
  switch (s->p->physics) {
    case SBP_ADVECTION_1D:
      adv1d_block(s, src, dst, i0, i1);
      adv1d_bc(s, src, dst, i0, i1);
      break;
    case SBP_REFLECTION_1D:
      refl_block(s, src, dst, i0, i1);
      refl_bc(s, dst, i0, i1);
      break;
    case SBP_ADVECTION_2D:
      adv2d_block(s, src, dst, i0, i1, j0, j1);
      adv2d_bc(s, src, dst, i0, i1, j0, j1);
      break;
    case SBP_ACOUSTIC_2D:
      acoustic_block(s, t, src, dst, i0, i1, j0, j1);
      acoustic_bc(s, src, dst, i0, i1, j0, j1);
      break;
  }
}

int orc_rhs(const sbp_problem *p, double t, const double *src, double *dst)
{
  if (orc_validate(p)) return -1;
  orc_setup s;
  setup(p, &s);
  rhs_block(&s, t, src, dst, 0, s.nx, 0, s.ny);
  return 0;
}

/* ------------------------------------------------------------------------ */
/* RK4: RK4_custom (time_stepping.h:98-152), apply_WAXPY (:232-252),        */
/* rk4_util (:254-278).                                                       */
/* ------------------------------------------------------------------------ */
int orc_rk4(const sbp_problem *p, double *v, double t0, double dt, long nsteps)
{
  return orc_rk4_par(p, v, t0, dt, nsteps, 1);
}

typedef struct {
  const orc_setup *s;
  double *v, *k1, *k2, *k3, *k4, *tmp;
  double t0, dt;
  long nsteps, len;
  int i0, i1, j0, j1;
  pthread_barrier_t *bar;
  int nthreads;
} orc_worker;

static inline void waxpy_block(const orc_setup *s, double *w, double alpha, const double *x,
                               const double *y, int i0, int i1, int j0, int j1)
{
  const int C = s->C;
  for (int j = j0; j < j1; j++) {
    long b = (long)C * ((long)s->nx * j + i0), e = (long)C * ((long)s->nx * j + i1);
    for (long k = b; k < e; k++) w[k] = y[k] + alpha * x[k];
  }
}

static inline void rk4_util_block(const orc_setup *s, double *v, const double *k1,
                                  const double *k2, const double *k3, const double *k4,
                                  double alpha, int i0, int i1, int j0, int j1)
{
  const int C = s->C;
  for (int j = j0; j < j1; j++) {
    long b = (long)C * ((long)s->nx * j + i0), e = (long)C * ((long)s->nx * j + i1);
    for (long k = b; k < e; k++) v[k] += alpha * (k1[k] + 2 * (k2[k] + k3[k]) + k4[k]);
  }
}

static inline void sync_workers(orc_worker *w)
{
  if (w->nthreads > 1) pthread_barrier_wait(w->bar);
}

static void *rk4_worker(void *arg)
{
  orc_worker *w = (orc_worker *)arg;
  const orc_setup *s = w->s;
  const double dt = w->dt, dt2 = 0.5 * dt, dt6 = dt / 6;
  double t = w->t0;
  for (long n = 0; n < w->nsteps; n++) {
    rhs_block(s, t, w->v, w->k1, w->i0, w->i1, w->j0, w->j1);
    sync_workers(w); /* k1 complete before its use is irrelevant (pointwise), but v
                        must stay untouched until every block has read its halo */
    waxpy_block(s, w->tmp, dt2, w->k1, w->v, w->i0, w->i1, w->j0, w->j1);
    sync_workers(w);
    rhs_block(s, t + dt2, w->tmp, w->k2, w->i0, w->i1, w->j0, w->j1);
    sync_workers(w);
    waxpy_block(s, w->tmp, dt2, w->k2, w->v, w->i0, w->i1, w->j0, w->j1);
    sync_workers(w);
    rhs_block(s, t + dt2, w->tmp, w->k3, w->i0, w->i1, w->j0, w->j1);
    sync_workers(w);
    waxpy_block(s, w->tmp, dt, w->k3, w->v, w->i0, w->i1, w->j0, w->j1);
    sync_workers(w);
    rhs_block(s, t + dt, w->tmp, w->k4, w->i0, w->i1, w->j0, w->j1);
    sync_workers(w);
    rk4_util_block(s, w->v, w->k1, w->k2, w->k3, w->k4, dt6, w->i0, w->i1, w->j0, w->j1);
    sync_workers(w);
    t = t + dt;
  }
  return NULL;
}

/* Near-square px x py factorisation of n ranks, as PETSC_DECIDE does for
 * DMDACreate2d (wave_eq_sim.cpp:100-101): px/py close to nx/ny. */
static void decompose(int n, int nx, int ny, int is1d, int *px, int *py)
{
  if (is1d) { *px = n; *py = 1; return; }
  int best = 1; double bestscore = 1e300;
  for (int a = 1; a <= n; a++) {
    if (n % a) continue;
    int b = n / a;
    double sx = (double)nx / a, sy = (double)ny / b;
    double score = sx > sy ? sx / sy : sy / sx;
    if (score < bestscore) { bestscore = score; best = a; }
  }
  *px = best; *py = n / best;
}

int orc_rk4_par(const sbp_problem *p, double *v, double t0, double dt, long nsteps, int nthreads)
{
  if (orc_validate(p)) return -1;
  orc_setup s;
  setup(p, &s);
  long len = (long)s.C * s.nx * s.ny;
  double *buf = (double *)malloc(sizeof(double) * 5 * len);
  if (!buf) return -1;
  if (nthreads < 1) nthreads = 1;
  int px, py;
  decompose(nthreads, s.nx, s.ny, is_1d(p), &px, &py);
  int nt = px * py;
  orc_worker *ws = (orc_worker *)calloc(nt, sizeof(orc_worker));
  pthread_t *th = (pthread_t *)calloc(nt, sizeof(pthread_t));
  pthread_barrier_t bar;
  if (nt > 1) pthread_barrier_init(&bar, NULL, nt);
  for (int b = 0; b < py; b++)
    for (int a = 0; a < px; a++) {
      orc_worker *w = &ws[b * px + a];
      w->s = &s; w->v = v;
      w->k1 = buf; w->k2 = buf + len; w->k3 = buf + 2 * len; w->k4 = buf + 3 * len;
      w->tmp = buf + 4 * len;
      w->t0 = t0; w->dt = dt; w->nsteps = nsteps; w->len = len;
      w->i0 = (int)((long)s.nx * a / px); w->i1 = (int)((long)s.nx * (a + 1) / px);
      w->j0 = (int)((long)s.ny * b / py); w->j1 = (int)((long)s.ny * (b + 1) / py);
      w->bar = &bar; w->nthreads = nt;
    }
  if (nt == 1) {
    rk4_worker(&ws[0]);
  } else {
    for (int k = 0; k < nt; k++) pthread_create(&th[k], NULL, rk4_worker, &ws[k]);
    for (int k = 0; k < nt; k++) pthread_join(th[k], NULL);
    pthread_barrier_destroy(&bar);
  }
  free(th);
  free(ws);
  free(buf);
  return 0;
}

/* ------------------------------------------------------------------------ */
/* Norms                                                                      */
/* ------------------------------------------------------------------------ */
/* Diagonal SBP norm weight of index i on n points (H_central.h:39-58: h*q_i
 * in the closures, h in the interior, mirrored on the right). */
static inline double hweight(const orc_ops *op, int n, int i, double h)
{
  if (i < op->cls) return h * op->hq[i];
  if (i >= n - op->cls) return h * op->hq[n - 1 - i];
  return h;
}

double orc_energy(const sbp_problem *p, const double *v)
{
  if (orc_validate(p)) return -1;
  orc_setup s;
  setup(p, &s);
  const double hx = 1.0 / s.hix, hy = is_1d(p) ? 1.0 : 1.0 / s.hiy;
  double E = 0;
  for (int j = 0; j < s.ny; j++) {
    double wy = is_1d(p) ? 1.0 : hweight(&s.op, s.ny, j, hy);
    for (int i = 0; i < s.nx; i++) {
      double wx = hweight(&s.op, s.nx, i, hx);
      const double *q = v + (long)s.C * ((long)s.nx * j + i);
      double e = 0;
      if (p->physics == SBP_ACOUSTIC_2D) {
        double rho = 1.0 / mat_rinv(&s, i, j);
        double kinv = 1.0 / mat_bulk(&s, i, j);
        e = rho * (q[0] * q[0] + q[1] * q[1]) + kinv * q[2] * q[2];
      } else {
        for (int c = 0; c < s.C; c++) e += q[c] * q[c];
      }
      E += wx * wy * e;
    }
  }
  return E;
}

/* get_norm_1D (H_central.h:70-103) and its 2D tensor-product analogue. */
double orc_hnorm(const sbp_problem *p, const double *v)
{
  if (orc_validate(p)) return -1;
  orc_setup s;
  setup(p, &s);
  const double hx = 1.0 / s.hix, hy = is_1d(p) ? 1.0 : 1.0 / s.hiy;
  double u = 0;
  for (int c = 0; c < s.C; c++)
    for (int j = 0; j < s.ny; j++) {
      double wy = is_1d(p) ? 1.0 : hweight(&s.op, s.ny, j, hy);
      for (int i = 0; i < s.nx; i++) {
        double x = v[(long)s.C * ((long)s.nx * j + i) + c];
        u += wy * hweight(&s.op, s.nx, i, hx) * x * x;
      }
    }
  return sqrt(u);
}

/* error_l2 / error_max (vec_util.h:47-61). */
void orc_error_norms(const sbp_problem *p, const double *v, const double *w,
                     double *l2, double *maxerr)
{
  orc_setup s;
  setup(p, &s);
  long len = (long)s.C * s.nx * s.ny;
  double ss = 0, mx = 0;
  for (long k = 0; k < len; k++) {
    double d = fabs(v[k] - w[k]);
    ss += d * d;
    if (d > mx) mx = d;
  }
  double hprod = 1.0 / s.hix;
  if (!is_1d(p)) hprod *= 1.0 / s.hiy;
  if (l2) *l2 = sqrt(hprod) * sqrt(ss);
  if (maxerr) *maxerr = mx;
}

/* ------------------------------------------------------------------------ */
/* PETSc 3.12 TS step-size rules behind ts_rk4 / ts_rk45 (ts_rk.cpp:15-65):   */
/* TS_EXACTFINALTIME_MATCHSTEP with TSADAPTNONE (TSRK4) or TSADAPTBASIC       */
/* (TSRK5F).  PETSc's source is not part of the reference; these restate      */
/* TSSolve's first-step clip, TSAdaptChoose's MATCHSTEP adjustment            */
/* (matchstepfac {0.01, 2}), TSAdaptChoose_Basic (safety 0.9, reject safety   */
/* 0.5 after a rejected attempt of the same step, clip [0.1, 10], dt in       */
/* [1e-20, 1e50]) and TSStep_RK's max_reject = 10.                            */
/* ------------------------------------------------------------------------ */
static double orc_clip(double x, double lo, double hi) { return x < lo ? lo : (x > hi ? hi : x); }

static double ts_first_step(double ptime, double dt, double t_max)
{
  const double maxdt = t_max - ptime;
  if (dt >= maxdt) return maxdt;
  const double d = fabs(dt - maxdt);  /* PetscIsCloseAtTol(dt, maxdt, 10 eps, 0) */
  if (d <= 10 * DBL_EPSILON * fmax(fabs(dt), fabs(maxdt))) return maxdt;
  return dt;
}

/* After an accepted step ending at t, the proposal h becomes: */
static double ts_matchstep(double t, double h, double t_max)
{
  const double tend = t + h, hmax = t_max - t;
  double next = h;
  if (t < t_max && tend > t_max) next = hmax;
  if (t < t_max && tend < t_max && h * 2.0 > hmax) next = hmax / 2;
  if (t < t_max && tend < t_max && h * (1.0 + 0.01) > hmax) next = hmax;
  return next;
}

void orc_adapt_basic(double enorm, double h, int order, int prev_rejected, int *accept, double *next_h)
{
  double safety = 0.9;
  int acc = 1;
  if (enorm > 1) {
    if (prev_rejected) safety *= 0.5;
    acc = h < (1 + sqrt(DBL_EPSILON)) * 1e-20;
  }
  const double hfac = enorm > 0 ? safety * pow(enorm, -1.0 / order) : INFINITY;
  *accept = acc;
  *next_h = orc_clip(h * orc_clip(hfac, 0.1, 10.0), 1e-20, 1e50);
}

int orc_matchstep_steps(double t0, double t_end, double dt, long max_steps, double *steps, double *t_final)
{
  double t = t0, h = ts_first_step(t0, dt, t_end);
  long n = 0;
  while (t < t_end) {
    if (n >= max_steps) return -1;
    if (steps) steps[n] = h;
    n++;
    const double next = ts_matchstep(t + h, h, t_end);
    t = t + h;
    h = next;
  }
  if (t_final) *t_final = t;
  return (int)n;
}

/* ts_rk4: TSRK4 (the RK4_custom combination) over the MATCHSTEP schedule. */
int orc_ts_rk4(const sbp_problem *p, double *v, double t0, double t_end, double dt, long *nsteps,
               double *t_final, int nthreads)
{
  if (orc_validate(p) || !(dt > 0)) return -1;
  double t = t0, h = ts_first_step(t0, dt, t_end);
  long n = 0;
  while (t < t_end) {
    /* a run of equal steps goes to the (threaded) fixed-step integrator */
    long run = 0;
    double tt = t, hh = h;
    while (tt < t_end && hh == h) {
      const double next = ts_matchstep(tt + hh, hh, t_end);
      tt = tt + hh;
      hh = next;
      run++;
    }
    if (orc_rk4_par(p, v, t, h, run, nthreads)) return -1;
    for (long k = 0; k < run; k++) t = t + h;
    n += run;
    h = hh;
  }
  if (nsteps) *nsteps = n;
  if (t_final) *t_final = t;
  return 0;
}

/* ------------------------------------------------------------------------ */
/* Adaptive Runge-Kutta-Fehlberg 4(5): what ts_rk45 (ts_rk.cpp:41-52) asks   */
/* PETSc for -- TSRK5F (fifth-order solution propagated, fourth-order        */
/* embedded estimate) with the rules above; weighted 2-norm with             */
/* tol = atol + rtol*max(|y|, |y_hat|).                                        */
/* ------------------------------------------------------------------------ */
/* Fehlberg's coefficients (E. Fehlberg, NASA TR R-315, 1969). */
static const double RKF_C[6] = {0.0, 1.0 / 4.0, 3.0 / 8.0, 12.0 / 13.0, 1.0, 1.0 / 2.0};
static const double RKF_A[6][5] = {
    {0, 0, 0, 0, 0},
    {1.0 / 4.0, 0, 0, 0, 0},
    {3.0 / 32.0, 9.0 / 32.0, 0, 0, 0},
    {1932.0 / 2197.0, -7200.0 / 2197.0, 7296.0 / 2197.0, 0, 0},
    {439.0 / 216.0, -8.0, 3680.0 / 513.0, -845.0 / 4104.0, 0},
    {-8.0 / 27.0, 2.0, -3544.0 / 2565.0, 1859.0 / 4104.0, -11.0 / 40.0}};
static const double RKF_B5[6] = {16.0 / 135.0, 0.0, 6656.0 / 12825.0, 28561.0 / 56430.0, -9.0 / 50.0, 2.0 / 55.0};
static const double RKF_B4[6] = {25.0 / 216.0, 0.0, 1408.0 / 2565.0, 2197.0 / 4104.0, -1.0 / 5.0, 0.0};

void orc_rkf45_tableau(double c[6], double a[6][5], double b5[6], double b4[6])
{
  for (int s = 0; s < 6; s++) {
    c[s] = RKF_C[s];
    b5[s] = RKF_B5[s];
    b4[s] = RKF_B4[s];
    for (int j = 0; j < 5; j++) a[s][j] = RKF_A[s][j];
  }
}

int orc_rk45(const sbp_problem *p, double *v, double t0, double t_end, double dt0, double atol,
             double rtol, long max_steps, long *accepted, long *rejected, double *dt_next)
{
  if (orc_validate(p) || !(dt0 > 0) || !(t_end >= t0)) return -1;
  const long n = (long)orc_components(p) * p->nx * (p->ny > 0 ? p->ny : 1);
  double *K[6], *Y = malloc(sizeof(double) * n), *U5 = malloc(sizeof(double) * n);
  int ok = Y && U5;
  for (int s = 0; s < 6; s++) ok = ok && (K[s] = malloc(sizeof(double) * n)) != NULL;
  if (!ok) return -1;
  double e[6];
  for (int s = 0; s < 6; s++) e[s] = RKF_B5[s] - RKF_B4[s];
  double t = t0, h = ts_first_step(t0, dt0, t_end);
  long acc = 0, rej = 0;
  int rc = 0;
  while (t < t_end && !rc) {
    int prev_rejected = 0, rejections = 0;
    for (;;) {
      if (acc + rej >= max_steps) {
        rc = -2;
        break;
      }
      orc_rhs(p, t, v, K[0]);
      for (int s = 1; s < 6; s++) {
        for (long k = 0; k < n; k++) {
          double sum = RKF_A[s][0] * K[0][k];
          for (int j = 1; j < s; j++) sum = sum + RKF_A[s][j] * K[j][k];
          Y[k] = v[k] + h * sum;
        }
        orc_rhs(p, t + RKF_C[s] * h, Y, K[s]);
      }
      double ss = 0;
      for (long k = 0; k < n; k++) {
        double sb = RKF_B5[0] * K[0][k], se = e[0] * K[0][k];
        for (int j = 1; j < 6; j++) {
          sb = sb + RKF_B5[j] * K[j][k];
          se = se + e[j] * K[j][k];
        }
        const double u5 = v[k] + h * sb, err = h * se, y4 = u5 - err;
        const double tol = atol + rtol * fmax(fabs(u5), fabs(y4));
        const double w = err / tol;
        ss += w * w;
        U5[k] = u5;
      }
      const double enorm = sqrt(ss / (double)n);
      if (!isfinite(enorm)) { /* TSErrorWeightedNorm: PETSC_ERR_FP */
        rc = -3;
        break;
      }
      int accept;
      double next;
      orc_adapt_basic(enorm, h, 5, prev_rejected, &accept, &next);
      if (accept) {
        memcpy(v, U5, sizeof(double) * n);
        next = ts_matchstep(t + h, next, t_end);
        t = t + h;
        h = next;
        acc++;
        break;
      }
      rej++;
      h = next;
      prev_rejected = 1;
      if (++rejections > 10) { /* TS_DIVERGED_STEP_REJECTED */
        rc = -3;
        break;
      }
    }
  }
  double dt = h;
  for (int s = 0; s < 6; s++) free(K[s]);
  free(Y);
  free(U5);
  if (accepted) *accepted = acc;
  if (rejected) *rejected = rej;
  if (dt_next) *dt_next = dt;
  return rc;
}
