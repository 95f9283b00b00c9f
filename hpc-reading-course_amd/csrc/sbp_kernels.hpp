// sbp_kernels.hpp -- gfx950 kernels of the SBP-SAT RK4 hot path.
//
// One launch per RK4 stage computes the semi-discrete right-hand side F(t, Y)
// of code/demo/*/*_rhs.h (region kernels + SAT) and folds the classical RK4
// stage arithmetic into the same pass.  The 2D kernels keep four registers in
// HBM (U, Y2, Y3, Y4) and rebuild the RK4 combination in the last stage:
//
//   stage 1: K1 = F(U);   Y2 = U + dt/2 K1                         (reads U)
//   stage 2: K2 = F(Y2);  Y3 = U + dt/2 K2                         (reads Y2, U)
//   stage 3: K3 = F(Y3);  Y4 = U + dt K3                           (reads Y3, U)
//   stage 4: K4 = F(Y4);  U += ((Y2-U) + 2(Y3-U) + (Y4-U))/3 + dt/6 K4
//                                                                  (reads Y4, U, Y2, Y3)
//
// i.e. RK4_custom's update U + dt/6 (K1 + 2(K2+K3) + K4)
// (code/include/time_stepping/time_stepping.h:127-142) with dt/2 K1 = Y2-U,
// dt/2 K2 = Y3-U, dt K3 = Y4-U recovered exactly (each difference of two
// nearby floating-point numbers is exact).  9 reads + 4 writes per value and
// step instead of the 16 of a 3-register (U, Y, accumulator) scheme.  The
// Kahan form compensates the final U += increment with one more register.
// The 1D kernels keep the stage vectors and follow RK4_custom's operation
// order exactly.  Large 2D grids run the two-stage schedule further down
// (pair2d_kernel: stages 1-2 and 3-4 in one launch each, the inner stage on
// chip, 5 words per value and step) with the same arithmetic per point.
//
// Layout in HBM: one plane per component, rows of `pitch` elements, column i
// stored at xoff + i with xoff chosen so that the first interior column
// (i = CLS) starts on a 64-element boundary.
//
// Work split inside one launch:
//   * blocks [0, n_edge_blocks): one thread per boundary-region point (the
//     CLS-deep frame holding every closure row and every SAT point).  Stencil
//     rows come from a table indexed by the point, so closure and interior
//     rows run the same instruction stream (no divergence).
//   * blocks [n_edge_blocks, ...): interior tiles [CLS, N-CLS)^2, BX columns
//     wide and ty_rows tall.  Each thread owns one column and marches up the
//     tile keeping a (2R+1)-row register window of the y-differentiated
//     fields; the x-differentiated fields of the current row go through LDS
//     with an R-wide halo (the DMDA ghost exchange of the reference becomes
//     this intra-device halo).  Block order is remapped so that every XCD gets
//     a contiguous band of tiles (its own L2 then serves the vertical halos).
#pragma once
#include <hip/hip_runtime.h>

#include "sbp_ops.hpp"

namespace sbp {

enum Phys { PH_ADV1D = 1, PH_ADV2D = 2, PH_ACOUSTIC = 3, PH_REFL1D = 4 };
enum Mat { M_UNIT = 0, M_REF = 1, M_ARR = 2 };
enum Bc { B_DEFAULT = 0, B_FSREF = 1, B_FS = 2, B_WALL = 3 };
enum Mode { MODE_RHS = 0, MODE_S1 = 1, MODE_S2 = 2, MODE_S3 = 3, MODE_S4 = 4 };

constexpr double kPi = 3.14159265358979323846264338327950288;  // PETSC_PI
constexpr int kNumXcd = 8;

template <typename T> struct StageArgs {
  const T *yin;   // stage input (stencil operand)
  const T *uin;   // U (stages 2, 3)
  const T *y2;    // Y2 (stage 4 of the 2D kernels)
  const T *y3;    // Y3 (stage 4 of the 2D kernels)
  T *aio;         // S register of the 1D stage sequence
  T *yout;        // next Y (stages 1-3) or F (RHS mode)
  T *uout;        // U (stage 4, read and written)
  T *comp;        // Kahan compensation of U
  T *tk;          // k2 register of the 1D stage sequence
  const T *rinv;  // rho^-1 plane (caller material arrays)
  const T *bulk;  // K plane
  const T *velx;  // advection velocity a(i,j) plane (per-point velocity), or null
  const T *vely;  // advection velocity b(i,j) plane (2D)
  const double *xs, *ys;     // x_i = xl + i/hix, y_j (reference rho and forcing)
  const double *c3x, *s3x;   // cos(3 pi x_i), sin(3 pi x_i)
  const double *c4y, *s4y;   // cos(4 pi y_j), sin(4 pi y_j)
  double c5t;                // cos(5 pi t) at the stage time
  long plane;
  int pitch, xoff, nx, ny;
  T hix, hiy, ax, ay;
  T cy, ca;                  // stage coefficients: Y = U + cy K; dt/6 (stage 4)
  T third;                   // 1/3 (stage 4)
  int bc;
  int tiles_x, ty_rows, n_tiles, n_edge_blocks, n_edge;
  int depth;                 // frame depth of the edge blocks (0: CLS)
};

template <typename T, int ORDER> __constant__ EdgeTable<T, ORDER> c_edge = make_edge_table<T, ORDER>();

// Table-driven derivative at index idx of a strided line of n points.
template <typename T, int ORDER>
__device__ __forceinline__ T dline(const T *__restrict__ f, long stride, int n, int idx, T hi)
{
  using G = Geom<ORDER>;
  // table row and first tap by selects, not branches: the loads of every
  // line of an edge point then sit in one basic block and issue together
  const bool lo = idx < G::CLS, hi_side = idx >= n - G::CLS;
  const int row = lo ? idx : (hi_side ? G::CLS + (n - 1 - idx) : 2 * G::CLS);
  const int start = lo ? 0 : (hi_side ? n - G::TAPS : idx - G::R);
  const T *c = c_edge<T, ORDER>.c[row];
  T acc = T(0);
#pragma unroll
  for (int s = 0; s < G::TAPS; s++) acc += c[s] * f[(long)(start + s) * stride];
  return hi * acc;
}

// Interior stencil on 2R+1 values, zero centre tap skipped.
template <typename T, int ORDER, typename Get>
__device__ __forceinline__ T dint(Get get, T hi)
{
  using O = Ops<ORDER>;
  T acc = T(0);
#pragma unroll
  for (int s = 0; s < 2 * O::R + 1; s++)
    if (O::inner[s] != 0.0) acc += T(O::inner[s]) * get(s);
  return hi * acc;
}

// Acoustic right-hand side at one point (wave_eq_rhs.h:76-322,
// wave_eq_hom_rhs.h:49-296; caller arrays carry K on the pressure row).
template <typename T, int MAT>
__device__ __forceinline__ void acoustic_F(T dxp, T dyp, T dxu, T dyv, T ri, T K, T fu, T fv,
                                           T &Fu, T &Fv, T &Fp)
{
  if constexpr (MAT == M_UNIT) {
    Fu = -dxp;
    Fv = -dyp;
    Fp = -dxu - dyv;
  } else if constexpr (MAT == M_REF) {
    Fu = -ri * dxp + fu;
    Fv = -ri * dyp + fv;
    Fp = -dxu - dyv;
  } else {
    Fu = -ri * dxp;
    Fv = -ri * dyp;
    Fp = -K * (dxu + dyv);
  }
}

// rho^-1 and the manufactured forcing of wave_eq_rhs.h:42-64 from 1D tables.
template <typename T>
__device__ __forceinline__ void ref_material(const StageArgs<T> &a, int i, int j, T &ri, T &fu, T &fv)
{
  const double x = a.xs[i], y = a.ys[j];
  const double xy = x * y;
  ri = T(1. / (2 + xy));
  fu = T(-(3 * kPi * a.c5t * a.c3x[i] * a.s4y[j] * (xy + 1)) / (xy + 2));
  fv = T(-(4 * kPi * a.c5t * a.c4y[j] * a.s3x[i] * (xy + 1)) / (xy + 2));
}

// rho^-1 and the manufactured forcing at (x_i, y_j) and time factor c5t =
// cos(5 pi t), from per-column values (x, cos 3 pi x, sin 3 pi x) and the row
// tables: the expressions of ref_material, so every path rounds them alike.
template <typename T>
__device__ __forceinline__ void ref_material_at(double x, double c3x, double s3x, const double *ys, const double *c4y,
                                                const double *s4y, int j, double c5t, T &ri, T &fu, T &fv)
{
  const double y = ys[j];
  const double xy = x * y;
  ri = T(1. / (2 + xy));
  fu = T(-(3 * kPi * c5t * c3x * s4y[j] * (xy + 1)) / (xy + 2));
  fv = T(-(4 * kPi * c5t * c4y[j] * s3x * (xy + 1)) / (xy + 2));
}

// RK4 stage arithmetic at one point (edge path of the 2D kernels).  Ycen are
// the centre values of the stage operand: U in stage 1, Y4 in stage 4.
template <typename T, bool KAHAN>
__device__ __forceinline__ T stage4_increment(const StageArgs<T> &a, T U, T Y2, T Y3, T Y4, T K)
{
  return ((Y2 - U) + 2 * (Y3 - U) + (Y4 - U)) * a.third + a.ca * K;
}

// Pointwise operands of the RK4 update at one edge point: U (stages 2-4), Y2,
// Y3 and the compensation (stage 4).  edge_point loads them before anything
// else and stores after everything else: the compiler cannot move a load past
// a store through these pointers (they may alias as far as it knows), and
// loads issued late, one load-store pair per component, put several
// dependent memory round trips on the critical path of the edge blocks.
// Distinct components live in distinct planes, and each component's loads
// precede its own stores, so the values are unchanged.
template <typename T, int NC> struct RkIn {
  T U[NC], Y2[NC], Y3[NC], C[NC];
};

template <typename T, int NC, int MODE, bool KAHAN>
__device__ __forceinline__ RkIn<T, NC> rk_load(const StageArgs<T> &a, long o)
{
  RkIn<T, NC> in;
#pragma unroll
  for (int c = 0; c < NC; c++) {
    const long q = o + c * a.plane;
    in.U[c] = in.Y2[c] = in.Y3[c] = in.C[c] = T(0);
    if constexpr (MODE == MODE_S2 || MODE == MODE_S3 || MODE == MODE_S4) in.U[c] = a.uin[q];
    if constexpr (MODE == MODE_S4) {
      in.Y2[c] = a.y2[q];
      in.Y3[c] = a.y3[q];
      if constexpr (KAHAN) in.C[c] = a.comp[q];
    }
  }
  return in;
}

template <typename T, int NC, int MODE, bool KAHAN>
__device__ __forceinline__ void rk_store(const StageArgs<T> &a, long o, const T *K, const T *Ycen,
                                         const RkIn<T, NC> &in)
{
#pragma unroll
  for (int c = 0; c < NC; c++) {
    const long q = o + c * a.plane;
    if constexpr (MODE == MODE_RHS) {
      a.yout[q] = K[c];
    } else if constexpr (MODE == MODE_S1) {
      a.yout[q] = Ycen[c] + a.cy * K[c];
    } else if constexpr (MODE == MODE_S2 || MODE == MODE_S3) {
      a.yout[q] = in.U[c] + a.cy * K[c];
    } else {
      const T U = in.U[c];
      const T inc = stage4_increment<T, KAHAN>(a, U, in.Y2[c], in.Y3[c], Ycen[c], K[c]);
      if constexpr (!KAHAN) {
        a.uout[q] = U + inc;
      } else {
        const T y = inc - in.C[c];
        const T t = U + y;
        a.comp[q] = (t - U) - y;
        a.uout[q] = t;
      }
    }
  }
}

// RK4 stage arithmetic of the 1D kernels in exactly RK4_custom's order
// (time_stepping.h:127-142, rk4_util :267-269): the stage vectors are kept,
//   stage 1: S = k1            Y = U + dt/2 k1
//   stage 2: T = k2            Y = U + dt/2 k2
//   stage 3: S = k1 + 2(k2+k3) Y = U + dt k3
//   stage 4: U = U + dt/6 (S + k4)   (Kahan form: the increment is compensated)
// The 1D problems are latency-bound, so the extra register costs nothing.
template <typename T, int NC, int MODE, bool KAHAN>
__device__ __forceinline__ void rk_point_exact(const StageArgs<T> &a, long o, const T *K, const T *Ucen)
{
#pragma unroll
  for (int c = 0; c < NC; c++) {
    const long q = o + c * a.plane;
    if constexpr (MODE == MODE_RHS) {
      a.yout[q] = K[c];
    } else if constexpr (MODE == MODE_S1) {
      a.aio[q] = K[c];
      a.yout[q] = Ucen[c] + a.cy * K[c];
    } else if constexpr (MODE == MODE_S2) {
      a.tk[q] = K[c];
      a.yout[q] = a.uin[q] + a.cy * K[c];
    } else if constexpr (MODE == MODE_S3) {
      a.aio[q] = a.aio[q] + 2 * (a.tk[q] + K[c]);
      a.yout[q] = a.uin[q] + a.cy * K[c];
    } else {
      const T inc = a.ca * (a.aio[q] + K[c]);
      const T U = a.uout[q];
      if constexpr (!KAHAN) {
        a.uout[q] = U + inc;
      } else {
        const T y = inc - a.comp[q];
        const T t = U + y;
        a.comp[q] = (t - U) - y;
        a.uout[q] = t;
      }
    }
  }
}

// --------------------------------------------------------------------------
// Boundary frame: one thread per point of the frame, CLS deep (or a.depth
// deep: the two-stage interior tiles below leave a frame of CLS + R).  Points
// of the frame beyond the closure rows take the table's interior row.
// --------------------------------------------------------------------------
// The right-hand side K at point (i, j) of a 2D problem (interior, closure
// and SAT terms), with the stencil operand read through a view: component c
// at (i', j') is f[c * fplane + j' * fpitch + i'] (the state in HBM, or a
// patch of it in LDS that holds every point the stencils of (i, j) reach).
// Materials, velocities and the SAT data come from a, at HBM offset o.  Uc
// receives the operand at the point.
template <typename T, int ORDER, int PHYS, int MAT>
__device__ __forceinline__ void edge_rhs(const StageArgs<T> &a, const T *f, long fplane, long fpitch, int i, int j,
                                         long o, T *K, T *Uc)
{
  constexpr bool AC = PHYS == PH_ACOUSTIC;
  const int nx = a.nx, ny = a.ny;
  const long rowb = j * fpitch;  // start of row j
  const long colb = i;           // column i at row 0
  const long q = rowb + i;
  const T qinv0 = T(Ops<ORDER>::hqinv[0]);
  const T hqx = a.hix * qinv0, hqy = a.hiy * qinv0;
  if constexpr (AC) {
    const T *U_ = f, *V_ = f + fplane, *P_ = f + 2 * fplane;
    const T u = U_[q], v = V_[q], p = P_[q];
    T ri = T(1), Kb = T(1), fu = T(0), fv = T(0);
    if constexpr (MAT == M_ARR) {
      ri = a.rinv[o];
      Kb = a.bulk[o];
    }
    const T dxp = dline<T, ORDER>(P_ + rowb, 1, nx, i, a.hix);
    const T dyp = dline<T, ORDER>(P_ + colb, fpitch, ny, j, a.hiy);
    const T dxu = dline<T, ORDER>(U_ + rowb, 1, nx, i, a.hix);
    const T dyv = dline<T, ORDER>(V_ + colb, fpitch, ny, j, a.hiy);
    Uc[0] = u;
    Uc[1] = v;
    Uc[2] = p;
    if constexpr (MAT == M_REF) ref_material(a, i, j, ri, fu, fv);
    acoustic_F<T, MAT>(dxp, dyp, dxu, dyv, ri, Kb, fu, fv, K[0], K[1], K[2]);
    // Boundary terms in bc_serial order W, S, E, N (boundary_conditions.h:84-87).
    const int bc = a.bc == B_DEFAULT ? B_FSREF : a.bc;
    if (bc == B_FSREF) {  // free_surface_bc_* (wave_eq_rhs.h:417-468)
      if (i == 0) K[0] -= hqx * p;
      if (j == 0) K[1] -= hqy * p;
      if (i == nx - 1) K[0] += hqx * p;
      if (j == ny - 1) K[1] += hqy * p;
    } else if (bc == B_FS) {  // the same terms carrying rho^-1
      if (i == 0) K[0] -= ri * (hqx * p);
      if (j == 0) K[1] -= ri * (hqy * p);
      if (i == nx - 1) K[0] += ri * (hqx * p);
      if (j == ny - 1) K[1] += ri * (hqy * p);
    } else {  // wall: normal velocity zero, on the pressure row, carrying K
      if (i == 0) K[2] -= Kb * (hqx * u);
      if (j == 0) K[2] -= Kb * (hqy * v);
      if (i == nx - 1) K[2] += Kb * (hqx * u);
      if (j == ny - 1) K[2] += Kb * (hqy * v);
    }
  } else {  // 2D advection: advection_{ll..rr} + SAT_bc_{west..north}
    // velocities at this point: per-point fields (MAT == M_ARR) or constants;
    // every SAT takes tau from its own boundary point (advection_rhs.h:461-528)
    T ax = a.ax, ay = a.ay;
    if constexpr (MAT == M_ARR) {
      ax = a.velx[o];
      ay = a.vely[o];
    }
    const T u = f[q];
    const T dx = dline<T, ORDER>(f + rowb, 1, nx, i, a.hix);
    const T dy = dline<T, ORDER>(f + colb, fpitch, ny, j, a.hiy);
    Uc[0] = u;
    T k = -(ax * dx + ay * dy);
    if (i == 0) k += T(0.5) * (T(-0.5) * (ax + fabs(ax))) * (hqx * u);
    if (j == 0) k += T(0.5) * (T(-0.5) * (ay + fabs(ay))) * (hqy * u);
    if (i == nx - 1) k += T(0.5) * (T(0.5) * (ax - fabs(ax))) * (hqx * u);
    if (j == ny - 1) k += T(0.5) * (T(0.5) * (ay - fabs(ay))) * (hqy * u);
    K[0] = k;
  }
}

template <typename T, int ORDER, int PHYS, int MAT, int MODE, bool KAHAN>
__device__ __forceinline__ void edge_point(const StageArgs<T> &a, int e)
{
  constexpr int CLS = Ops<ORDER>::CLS;
  constexpr int NC = PHYS == PH_ACOUSTIC ? 3 : 1;
  if (e >= a.n_edge) return;
  const int nx = a.nx, ny = a.ny;
  const int dep = a.depth > 0 ? a.depth : CLS;
  const int nstrip = dep * nx;
  const int nside = dep * (ny - 2 * dep);
  int i, j;
  if (e < nstrip) {  // bottom strip, rows 0..dep-1
    j = e / nx;
    i = e - j * nx;
  } else if (e < 2 * nstrip) {  // top strip
    e -= nstrip;
    j = ny - dep + e / nx;
    i = e - (e / nx) * nx;
  } else if (e < 2 * nstrip + nside) {  // left strip between them
    e -= 2 * nstrip;
    j = dep + e / dep;
    i = e - (e / dep) * dep;
  } else {  // right strip
    e -= 2 * nstrip + nside;
    j = dep + e / dep;
    i = nx - dep + (e - (e / dep) * dep);
  }
  const long o = j * (long)a.pitch + a.xoff + i;
  T K[NC], Uc[NC];
  // every load of the point is issued before the first arithmetic on it
  const RkIn<T, NC> rin = rk_load<T, NC, MODE, KAHAN>(a, o);
  edge_rhs<T, ORDER, PHYS, MAT>(a, a.yin + a.xoff, a.plane, a.pitch, i, j, o, K, Uc);
  rk_store<T, NC, MODE, KAHAN>(a, o, K, Uc, rin);
}

// --------------------------------------------------------------------------
// Interior tile: columns [x0, x0+BX) of [CLS, nx-CLS), rows [j0, j1).
// --------------------------------------------------------------------------

// Memory-access policy of the interior tiles (tuned on gfx950):
//   PF   rows of loads kept in flight ahead of the row being computed;
//   NTP  nontemporal loads of the pointwise operands (U, A, compensation,
//        materials), read exactly once;
//   NTS  nontemporal stores of every output;
//   NTY  nontemporal loads of the stencil operand Y.
template <int PF_, bool NTP_, bool NTS_, bool NTY_> struct TilePolicy {
  static constexpr int PF = PF_;
  static constexpr bool NTP = NTP_, NTS = NTS_, NTY = NTY_;
};
using DefaultPolicy = TilePolicy<1, true, true, false>;

template <bool NT, typename T> __device__ __forceinline__ T ldg(const T *p)
{
  if constexpr (NT) return __builtin_nontemporal_load(p);
  else return *p;
}

template <bool NT, typename T> __device__ __forceinline__ void stg(T *p, T v)
{
  if constexpr (NT) __builtin_nontemporal_store(v, p);
  else *p = v;
}

template <typename T, int NF, int NC> struct RowRegs {
  T win[NF];  // y-differentiated fields at row j+R
  T xo;       // x-only field (u) at row j
  T h[NF];    // halo column of the x-differentiated fields at row j
  T U[NC], Y2[NC], Y3[NC], C[NC];
  T ri, K;
};

template <typename T, int ORDER, int PHYS, int MAT, int MODE, bool KAHAN, int BX, class POL>
__device__ __forceinline__ void interior_tile(const StageArgs<T> &a, int tile)
{
  using O = Ops<ORDER>;
  constexpr int R = O::R, CLS = O::CLS, WN = 2 * R + 1;
  constexpr bool AC = PHYS == PH_ACOUSTIC;
  constexpr int NC = AC ? 3 : 1;
  constexpr int NF = AC ? 2 : 1;   // fields in LDS / in the y window
  constexpr int CS = AC ? 2 : 0;   // field differentiated in x and y (p, or u)
  constexpr int LW = BX + 2 * R;
  constexpr int PF = POL::PF;
  // pointwise values per row: U (stages 2-4), Y2 and Y3 (stage 4), the
  // compensation (Kahan stage 4), materials
  constexpr bool LD_U = MODE >= MODE_S2;
  constexpr bool LD_Y = MODE == MODE_S4;
  constexpr bool LD_C = MODE == MODE_S4 && KAHAN;
  // materials (acoustics: rho^-1, K) or velocity fields (advection: a, b)
  constexpr bool LD_M = MAT == M_ARR;

  __shared__ T lds[2][NF][LW];

  const int tid = threadIdx.x;
  const int tx = tile % a.tiles_x, tyb = tile / a.tiles_x;
  const int x0 = CLS + tx * BX;
  const int i = x0 + tid;
  const bool active = i < a.nx - CLS;
  const int ic = min(i, a.nx - 1);
  const int j0 = CLS + tyb * a.ty_rows;
  const int j1 = min(j0 + a.ty_rows, a.ny - CLS);
  const long plane = a.plane;
  const long pitch = a.pitch;
  const long col = a.xoff + ic;
  const bool hl = tid < R, hr = tid >= BX - R, halo = hl || hr;
  const long hcol = a.xoff + (hl ? i - R : min(i + R, a.nx - 1));
  const int hslot = hl ? tid : tid + 2 * R;
  const int jmax = a.ny - 1;
  const T *__restrict__ yin = a.yin;
  const T *__restrict__ usrc = MODE == MODE_S4 ? a.uout : a.uin;

  T win[NF][WN];
#pragma unroll
  for (int k = 0; k < 2 * R; k++) {
    const long o = (long)(j0 - R + k) * pitch + col;
    win[0][k + 1] = ldg<POL::NTY>(yin + CS * plane + o);
    if constexpr (AC) win[1][k + 1] = ldg<POL::NTY>(yin + plane + o);  // v
  }

  RowRegs<T, NF, NC> q[PF];
  auto fetch = [&](RowRegs<T, NF, NC> &r, int jn) {
    const long oy = (long)min(jn + R, jmax) * pitch;
    const long ox = (long)min(jn, jmax) * pitch;
    r.win[0] = ldg<POL::NTY>(yin + CS * plane + oy + col);
    if constexpr (AC) {
      r.win[1] = ldg<POL::NTY>(yin + plane + oy + col);
      r.xo = ldg<POL::NTY>(yin + ox + col);  // u
    }
    if (halo) {
      r.h[0] = yin[CS * plane + ox + hcol];
      if constexpr (AC) r.h[1] = yin[ox + hcol];
    }
#pragma unroll
    for (int c = 0; c < NC; c++) {
      if constexpr (LD_U) r.U[c] = ldg<POL::NTP>(usrc + c * plane + ox + col);
      if constexpr (LD_Y) {
        r.Y2[c] = ldg<POL::NTP>(a.y2 + c * plane + ox + col);
        r.Y3[c] = ldg<POL::NTP>(a.y3 + c * plane + ox + col);
      }
      if constexpr (LD_C) r.C[c] = ldg<POL::NTP>(a.comp + c * plane + ox + col);
    }
    if constexpr (LD_M) {
      r.ri = ldg<POL::NTP>((AC ? a.rinv : a.velx) + ox + col);
      r.K = ldg<POL::NTP>((AC ? a.bulk : a.vely) + ox + col);
    }
  };
#pragma unroll
  for (int k = 0; k < PF; k++) {
    q[k] = RowRegs<T, NF, NC>{};
    fetch(q[k], j0 + k);
  }

  double xs_i = 0, c3x = 0, s3x = 0;
  if constexpr (MAT == M_REF) {
    xs_i = a.xs[ic];
    c3x = a.c3x[ic];
    s3x = a.s3x[ic];
  }

  for (int j = j0; j < j1; j++) {
    const RowRegs<T, NF, NC> cur = q[0];
#pragma unroll
    for (int k = 0; k < PF - 1; k++) q[k] = q[k + 1];
    fetch(q[PF - 1], j + PF);
#pragma unroll
    for (int f = 0; f < NF; f++) {
#pragma unroll
      for (int k = 0; k < WN - 1; k++) win[f][k] = win[f][k + 1];
      win[f][WN - 1] = cur.win[f];
    }

    const int buf = (j - j0) & 1;
    lds[buf][0][R + tid] = win[0][R];
    if constexpr (AC) lds[buf][1][R + tid] = cur.xo;
    if (halo) {
      lds[buf][0][hslot] = cur.h[0];
      if constexpr (AC) lds[buf][1][hslot] = cur.h[1];
    }
    __syncthreads();

    const T *l0 = &lds[buf][0][tid];
    const T d0x = dint<T, ORDER>([&](int s) { return l0[s]; }, a.hix);
    const T d0y = dint<T, ORDER>([&](int s) { return win[0][s]; }, a.hiy);
    T K[NC], Uc[NC];
    if constexpr (AC) {
      const T *l1 = &lds[buf][1][tid];
      const T dxu = dint<T, ORDER>([&](int s) { return l1[s]; }, a.hix);
      const T dyv = dint<T, ORDER>([&](int s) { return win[1][s]; }, a.hiy);
      T ri = cur.ri, Kb = cur.K, fu = T(0), fv = T(0);
      if constexpr (MAT == M_REF) {
        const double y = a.ys[j];
        const double xy = xs_i * y;
        ri = T(1. / (2 + xy));
        fu = T(-(3 * kPi * a.c5t * c3x * a.s4y[j] * (xy + 1)) / (xy + 2));
        fv = T(-(4 * kPi * a.c5t * a.c4y[j] * s3x * (xy + 1)) / (xy + 2));
      }
      acoustic_F<T, MAT>(d0x, d0y, dxu, dyv, ri, Kb, fu, fv, K[0], K[1], K[2]);
      Uc[0] = cur.xo;
      Uc[1] = win[1][R];
      Uc[2] = win[0][R];
    } else {
      const T ax = LD_M ? cur.ri : a.ax, ay = LD_M ? cur.K : a.ay;
      K[0] = -(ax * d0x + ay * d0y);
      Uc[0] = win[0][R];
    }

    if (active) {
      const long o = (long)j * pitch + col;
      constexpr bool NS = POL::NTS;
#pragma unroll
      for (int c = 0; c < NC; c++) {
        const long qo = o + c * plane;
        if constexpr (MODE == MODE_RHS) {
          stg<NS>(a.yout + qo, K[c]);
        } else if constexpr (MODE == MODE_S1) {
          stg<NS>(a.yout + qo, Uc[c] + a.cy * K[c]);
        } else if constexpr (MODE == MODE_S2 || MODE == MODE_S3) {
          stg<NS>(a.yout + qo, cur.U[c] + a.cy * K[c]);
        } else {
          const T inc = stage4_increment<T, KAHAN>(a, cur.U[c], cur.Y2[c], cur.Y3[c], Uc[c], K[c]);
          if constexpr (!KAHAN) {
            stg<NS>(a.uout + qo, cur.U[c] + inc);
          } else {
            const T y = inc - cur.C[c];
            const T t = cur.U[c] + y;
            stg<NS>(a.comp + qo, (t - cur.U[c]) - y);
            stg<NS>(a.uout + qo, t);
          }
        }
      }
    }
  }
}

template <typename T, int ORDER, int PHYS, int MAT, int MODE, bool KAHAN, int BX, class POL = DefaultPolicy>
__global__ __launch_bounds__(BX) void stage2d_kernel(StageArgs<T> a)
{
  const int b = blockIdx.x;
  if (b < a.n_edge_blocks) {
    edge_point<T, ORDER, PHYS, MAT, MODE, KAHAN>(a, b * BX + threadIdx.x);
    return;
  }
  // n_edge_blocks is a multiple of kNumXcd, so bi % 8 is the XCD this block
  // was dispatched to; hand each XCD a contiguous range of tiles.
  const int bi = b - a.n_edge_blocks;
  const int nt = a.n_tiles;
  const int xcd = bi % kNumXcd, local = bi / kNumXcd;
  const int q = nt / kNumXcd, r = nt % kNumXcd;
  const int tile = xcd * q + min(xcd, r) + local;
  interior_tile<T, ORDER, PHYS, MAT, MODE, KAHAN, BX, POL>(a, tile);
}

// --------------------------------------------------------------------------
// Two RK4 stages per launch (temporal blocking; 2D acoustics and advection).
// --------------------------------------------------------------------------
//   pair 1: K1 = F(U),  Y2 = U + dt/2 K1 on the tile grown by R (on chip),
//           K2 = F(Y2), Y3 = U + dt/2 K2 on the tile.        (reads U; writes Y3)
//   pair 2: K3 = F(Y3), Y4 = U + dt K3 on the tile grown by R (on chip),
//           K1 = F(U),  Y2 = U + dt/2 K1 on the tile (recomputed, not read),
//           K4 = F(Y4), U' = U + ((Y2-U) + 2(Y3-U) + (Y4-U))/3 + dt/6 K4
//                                              (reads Y3, U; writes U')
// 2 + 3 = 5 words per value and step instead of the 13 of one launch per
// stage.  Recomputing K1 from U, which pair 2 reads R rows around every
// output point anyway, costs about 40 FLOP per point and saves a whole
// array pass; K1 is evaluated by the same code on the same values as in
// pair 1, so Y2 is the same to the bit.  Tiles cover [D, nx-D) x [D, ny-D)
// with D = CLS + R, so every point of a grown tile is an interior point of
// the operator.  The D-deep frame is cut into patches that run both stages
// of the launch themselves (frame_patch_pair: the first stage on the patch
// grown into the interior, in LDS, then the second), so a step is two
// launches; pair 2 writes U' into the Y2 array (U is read R points around
// each point by neighbouring tiles); the host swaps the roles of the U and
// Y2 arrays after each step.  Per point the arithmetic is that of the
// one-stage kernels, term for term.
//
// A block of NT threads owns NT mid columns (x0-R .. x0-R+NT-1) and the
// NT-2R output columns inside them, and marches down the rows: iteration k
// takes input row r = j0-2R+k, computes mid row r-R, then output row r-2R
// (pair_tile_rw below).  Two barriers per row.
//
// Threads (mid columns) per two-stage tile; the host sizes the grid with it.
// (With the LDS-ring tile, FP64 order 6 needed 128 to fit three blocks per
// CU; the register-window tile uses 8 KB of LDS and runs 256 everywhere.)
constexpr int pair_nt(int esz, int order) { return (void)esz, (void)order, 256; }

template <typename T> struct PairArgs {
  const T *yin;  // stencil operand of the outer stage: U (pair 1), Y3 (pair 2)
  const T *u;    // U (pair 2)
  const T *y2;   // unused (pair 2 recomputes Y2)
  T *o1;         // pair 2: U' (the Y2 array)
  T *o2;         // pair 1: Y3
  T *comp;       // Kahan compensation of U (pair 2)
  const T *rinv, *bulk;  // acoustics: rho^-1, K planes (M_ARR)
  const T *velx, *vely;  // 2D advection: velocity planes a(i,j), b(i,j) (M_ARR)
  long plane;
  int pitch, xoff, nx, ny;
  T hix, hiy;
  T ax, ay;         // 2D advection: constant velocity (M_UNIT)
  T c1, c2, third;  // pair 1: dt/2, dt/2;  pair 2: dt, dt/6
  T c0;             // pair 2: dt/2 (Y2 = U + dt/2 K1)
  // the reference's rho = 2 + xy with its forcing (M_REF): per-column and
  // per-row tables, and cos(5 pi t) at the times of the outer stage, the
  // inner stage and (pair 2) the recomputed stage 1
  const double *xs, *ys, *c3x, *s3x, *c4y, *s4y;
  double c5_mid, c5_out, c5_k1;
  int tiles_x, ty_rows, n_tiles, n_edge_blocks;
  int nts;  // 1: nontemporal output stores; 0: cached (pair_cached_stores); selects the kernel
};

// Output stores of the register-window tiles: cached when the arrays a step
// streams (U, Y2, Y3 and the materials) fit in the 256 MB Infinity Cache, so
// the next launch reads them from there; nontemporal otherwise, where cached
// stores only evict what the launch still reads.  Measured on gfx950 (ms or
// us per step, nontemporal -> cached): 1025^2 FP64 0.0695 -> 0.0657, 16 x
// 513^2 FP32 ensemble 122 -> 111 us; but 2049^2 FP64 0.176 -> 0.186, 4097^2
// FP32 0.344 -> 0.363, 16 x 513^2 FP64 ensemble 211 -> 239 us.
constexpr double kCachedStoreBytes = 256.0 * 1024 * 1024;
inline bool pair_cached_stores(double streamed_bytes) { return streamed_bytes <= kCachedStoreBytes; }

template <typename T> struct PairRow {
  T in[3];        // stencil operand (u, v, p) at input row r, mid column
  T h[2];         // u, p at input row r, halo column
  T U[3];         // pair 2: U at mid row r-R
  T ri, K;        // materials at mid row r-R
  T C[3];         // pair 2 (Kahan): compensation at output row r-2R
};

// Two-stage tile: each thread keeps its own column's last 2R+1 rows of the
// operand and of the mid stage in registers (the y-derivatives and pointwise
// values come from there), and only the rows being x-differentiated pass
// through LDS, once each: 8 KB of LDS per FP64 block (an earlier form kept
// whole rows of all three fields in LDS rings, 68 KB, and ran up to 7%
// slower).
//   phase 0: stage operand row r (u, p with the x halo) in LDS
//   phase 1: dx of row r into a register FIFO; mid row m = r-R from the
//            windows and the FIFO's oldest entry; its u, p into LDS (pair 2:
//            also the u, p of U at row m)
//   phase 2: dx of mid row m (pair 2: and of U at row m) into second FIFOs;
//            output row o = r-2R (pair 2: K1 at row o from the window of U
//            rows o-R..o+R and the FIFO of its x-derivatives)
// LDS rows are single-buffered: each row written in one phase is read in the
// next, and the barrier between them also separates the read from the next
// iteration's write.
template <typename T, int ORDER, int PHYS, int MAT, int PAIR, bool KAHAN, int NT, int PF, bool NTS>
__device__ __forceinline__ void pair_tile_rw(const PairArgs<T> &a, int tile)
{
  using O = Ops<ORDER>;
  constexpr int R = O::R, D = O::CLS + O::R, WN = 2 * R + 1;
  // acoustics: (u, v, p), u and p differentiated in x, v and p in y;
  // 2D advection: u, differentiated in both
  constexpr bool AC = PHYS == PH_ACOUSTIC;
  constexpr int NC = AC ? 3 : 1;
  static_assert(AC || PHYS == PH_ADV2D, "two-stage tiles: 2D acoustics or 2D advection");
  static_assert(AC || MAT != M_REF, "rho = 2 + xy is an acoustic material");
  constexpr int LW = NT + 2 * R;
  constexpr int BXO = NT - 2 * R;
  constexpr bool LD_M = MAT == M_ARR;

  __shared__ T s_u[LW], s_p[AC ? LW : 1];      // operand row r (x halo included)
  __shared__ T s_mu[NT], s_mp[AC ? NT : 1];    // mid row m
  __shared__ T s_Uu[PAIR == 2 ? NT : 1], s_Up[PAIR == 2 && AC ? NT : 1];  // pair 2: U at row m

  const int t = threadIdx.x;
  const int tx = tile % a.tiles_x, tyb = tile / a.tiles_x;
  const int x0 = D + tx * BXO;
  const int mc = x0 - R + t;
  const int mcc = min(mc, a.nx - 1);
  const bool hl = t < R, hr = t >= NT - R, halo = hl || hr;
  const int hc = hl ? mc - R : min(mc + R, a.nx - 1);
  const int hslot = hl ? t : t + 2 * R;
  const bool out_col = t >= R && t < NT - R && mc < a.nx - D;
  const int j0 = D + tyb * a.ty_rows;
  const int j1 = min(j0 + a.ty_rows, a.ny - D);
  const int nk = (j1 - j0) + 4 * R;
  const long plane = a.plane, pitch = a.pitch;
  const long col = a.xoff + mcc, hcol = a.xoff + hc;
  const int jmax = a.ny - 1;
  const T *__restrict__ yin = a.yin;
  // (a template parameter: stores that differ only in the nontemporal hint
  // on the two sides of a runtime branch are merged into one plain store)
  auto put = [&](T *p, T v) { stg<NTS>(p, v); };

  auto fetch = [&](PairRow<T> &q, int k) {
    const int r = min(j0 - 2 * R + k, jmax);
    const long orow = (long)r * pitch;
#pragma unroll
    for (int c = 0; c < NC; c++) q.in[c] = yin[c * plane + orow + col];
    if (halo) {
      q.h[0] = yin[orow + hcol];
      if constexpr (AC) q.h[1] = yin[2 * plane + orow + hcol];
    }
    const long om = (long)max(r - R, 0) * pitch + col;
    if constexpr (PAIR == 2) {
#pragma unroll
      for (int c = 0; c < NC; c++) q.U[c] = a.u[c * plane + om];
    }
    if constexpr (LD_M) {
      q.ri = (AC ? a.rinv : a.velx)[om];
      q.K = (AC ? a.bulk : a.vely)[om];
    }
    if constexpr (PAIR == 2 && KAHAN) {
      const long oo = (long)max(r - 2 * R, 0) * pitch + col;
#pragma unroll
      for (int c = 0; c < NC; c++) q.C[c] = ldg<true>(a.comp + c * plane + oo);
    }
  };

  double xc = 0, c3c = 0, s3c = 0;  // M_REF: x, cos 3 pi x, sin 3 pi x of this thread's column
  if constexpr (MAT == M_REF) {
    xc = a.xs[mcc];
    c3c = a.c3x[mcc];
    s3c = a.s3x[mcc];
  }

  constexpr int S = PF + 1;
  PairRow<T> q[S];
#pragma unroll
  for (int k = 0; k < S; k++) {
    q[k] = PairRow<T>{};
    if (k < PF) fetch(q[k], k);
  }
  // windows of the last 2R+1 rows (oldest first): operand rows r-2R..r, mid rows m-2R..m
  T win[3][WN], wmid[3][WN];
  // FIFOs of the last R+1 rows: dx of the operand (u, p), dx of the mid stage,
  // (pair 2) dx of U (u, p), materials; (pair 2) window of U over the last
  // 2R+1 mid rows (oldest first; fU[R] is U at the output row)
  constexpr int NU = PAIR == 2 ? WN : R + 1;
  T fdx[2][R + 1], gdx[2][R + 1], hdx[2][R + 1], fri[R + 1], fK[R + 1], fU[NU][3];
#pragma unroll
  for (int s = 0; s < WN; s++)
#pragma unroll
    for (int c = 0; c < 3; c++) win[c][s] = wmid[c][s] = T(0);
#pragma unroll
  for (int p = 0; p <= R; p++) {
    fdx[0][p] = fdx[1][p] = gdx[0][p] = gdx[1][p] = hdx[0][p] = hdx[1][p] = T(0);
    fri[p] = fK[p] = T(1);
  }
#pragma unroll
  for (int p = 0; p < NU; p++)
#pragma unroll
    for (int c = 0; c < 3; c++) fU[p][c] = T(0);

  for (int kb = 0; kb < nk; kb += S) {
#pragma unroll
    for (int pq = 0; pq < S; pq++) {
      const int k = kb + pq;
      const PairRow<T> &cur = q[pq];
      fetch(q[(pq + PF) % S], k + PF);

      // phase 0: operand row r into the windows and (u, p) into LDS
#pragma unroll
      for (int c = 0; c < NC; c++) {
#pragma unroll
        for (int s = 0; s < WN - 1; s++) win[c][s] = win[c][s + 1];
        win[c][WN - 1] = cur.in[c];
      }
      s_u[t + R] = cur.in[0];
      if constexpr (AC) s_p[t + R] = cur.in[2];
      if (halo) {
        s_u[hslot] = cur.h[0];
        if constexpr (AC) s_p[hslot] = cur.h[1];
      }
      __syncthreads();

      // phase 1: dx of row r; mid row m = r - R
#pragma unroll
      for (int p = 0; p < R; p++) {
        fdx[0][p] = fdx[0][p + 1];
        fdx[1][p] = fdx[1][p + 1];
      }
      fdx[0][R] = dint<T, ORDER>([&](int s) { return s_u[t + s]; }, a.hix);
      if constexpr (AC) fdx[1][R] = dint<T, ORDER>([&](int s) { return s_p[t + s]; }, a.hix);
      if (k >= 2 * R) {
        T Um[NC];
#pragma unroll
        for (int c = 0; c < NC; c++) Um[c] = PAIR == 1 ? win[c][R] : cur.U[c];
        T ri = LD_M ? cur.ri : T(1), Kb = LD_M ? cur.K : T(1);
        T F[NC];
        if constexpr (AC) {
          const T dyp = dint<T, ORDER>([&](int s) { return win[2][s]; }, a.hiy);
          const T dyv = dint<T, ORDER>([&](int s) { return win[1][s]; }, a.hiy);
          T fu = T(0), fv = T(0);
          if constexpr (MAT == M_REF)
            ref_material_at<T>(xc, c3c, s3c, a.ys, a.c4y, a.s4y, min(j0 - 3 * R + k, a.ny - 1), a.c5_mid, ri, fu,
                               fv);
          acoustic_F<T, MAT>(fdx[1][0], dyp, fdx[0][0], dyv, ri, Kb, fu, fv, F[0], F[1], F[2]);
        } else {
          const T dy = dint<T, ORDER>([&](int s) { return win[0][s]; }, a.hiy);
          const T ax = LD_M ? ri : a.ax, ay = LD_M ? Kb : a.ay;
          F[0] = -(ax * fdx[0][0] + ay * dy);
        }
#pragma unroll
        for (int c = 0; c < NC; c++) {
#pragma unroll
          for (int s = 0; s < WN - 1; s++) wmid[c][s] = wmid[c][s + 1];
          wmid[c][WN - 1] = Um[c] + a.c1 * F[c];
        }
        s_mu[t] = wmid[0][WN - 1];
        if constexpr (AC) s_mp[t] = wmid[2][WN - 1];
        if constexpr (PAIR == 2) {
          s_Uu[t] = Um[0];
          if constexpr (AC) s_Up[t] = Um[2];
        }
#pragma unroll
        for (int p = 0; p < R; p++) {
          fri[p] = fri[p + 1];
          fK[p] = fK[p + 1];
        }
        fri[R] = ri;
        fK[R] = Kb;
#pragma unroll
        for (int p = 0; p < NU - 1; p++)
#pragma unroll
          for (int c = 0; c < NC; c++) fU[p][c] = fU[p + 1][c];
#pragma unroll
        for (int c = 0; c < NC; c++) fU[NU - 1][c] = Um[c];
      }
      __syncthreads();

      // phase 2: dx of mid row m; output row o = r - 2R
      if (k >= 2 * R) {
#pragma unroll
        for (int p = 0; p < R; p++) {
          gdx[0][p] = gdx[0][p + 1];
          gdx[1][p] = gdx[1][p + 1];
        }
        if constexpr (PAIR == 2) {
#pragma unroll
          for (int p = 0; p < R; p++) {
            hdx[0][p] = hdx[0][p + 1];
            hdx[1][p] = hdx[1][p + 1];
          }
        }
        if (out_col) {
          gdx[0][R] = dint<T, ORDER>([&](int s) { return s_mu[t - R + s]; }, a.hix);
          if constexpr (AC) gdx[1][R] = dint<T, ORDER>([&](int s) { return s_mp[t - R + s]; }, a.hix);
          if constexpr (PAIR == 2) {
            hdx[0][R] = dint<T, ORDER>([&](int s) { return s_Uu[t - R + s]; }, a.hix);
            if constexpr (AC) hdx[1][R] = dint<T, ORDER>([&](int s) { return s_Up[t - R + s]; }, a.hix);
          }
        }
      }
      if (k >= 4 * R && k < nk && out_col) {
        const int o = j0 - 4 * R + k;
        T fu = T(0), fv = T(0), fu1 = T(0), fv1 = T(0);
        if constexpr (MAT == M_REF) {
          T ri_o;
          ref_material_at<T>(xc, c3c, s3c, a.ys, a.c4y, a.s4y, o, a.c5_out, ri_o, fu, fv);
          if constexpr (PAIR == 2) ref_material_at<T>(xc, c3c, s3c, a.ys, a.c4y, a.s4y, o, a.c5_k1, ri_o, fu1, fv1);
        }
        // velocity at the output row (2D advection)
        const T axo = LD_M ? fri[0] : a.ax, ayo = LD_M ? fK[0] : a.ay;
        T F[NC];
        if constexpr (AC) {
          const T dyp = dint<T, ORDER>([&](int s) { return wmid[2][s]; }, a.hiy);
          const T dyv = dint<T, ORDER>([&](int s) { return wmid[1][s]; }, a.hiy);
          acoustic_F<T, MAT>(gdx[1][0], dyp, gdx[0][0], dyv, fri[0], fK[0], fu, fv, F[0], F[1], F[2]);
        } else {
          const T dy = dint<T, ORDER>([&](int s) { return wmid[0][s]; }, a.hiy);
          F[0] = -(axo * gdx[0][0] + ayo * dy);
        }
        const long qo = (long)o * pitch + a.xoff + mc;
        if constexpr (PAIR == 1) {
#pragma unroll
          for (int c = 0; c < NC; c++) put(a.o2 + c * plane + qo, win[c][0] + a.c2 * F[c]);
        } else {
          // K1 = F(U) at the output row and Y2 = U + dt/2 K1, as pair 1 formed them
          T F1[NC];
          if constexpr (AC) {
            const T dyp1 = dint<T, ORDER>([&](int s) { return fU[s][2]; }, a.hiy);
            const T dyv1 = dint<T, ORDER>([&](int s) { return fU[s][1]; }, a.hiy);
            acoustic_F<T, MAT>(hdx[1][0], dyp1, hdx[0][0], dyv1, fri[0], fK[0], fu1, fv1, F1[0], F1[1], F1[2]);
          } else {
            const T dy1 = dint<T, ORDER>([&](int s) { return fU[s][0]; }, a.hiy);
            F1[0] = -(axo * hdx[0][0] + ayo * dy1);
          }
#pragma unroll
          for (int c = 0; c < NC; c++) {
            const T U = fU[R][c];
            const T Y2 = U + a.c0 * F1[c];
            const T Y4 = wmid[c][R];
            const T inc = ((Y2 - U) + 2 * (win[c][0] - U) + (Y4 - U)) * a.third + a.c2 * F[c];
            if constexpr (!KAHAN) {
              put(a.o1 + c * plane + qo, U + inc);
            } else {
              const T y = inc - cur.C[c];
              const T tt = U + y;
              put(a.comp + c * plane + qo, (tt - U) - y);
              put(a.o1 + c * plane + qo, tt);
            }
          }
        }
      }
    }
  }
}

// --------------------------------------------------------------------------
// Frame patches of the two-stage step.  The D-deep frame (D = CLS + R) is cut
// into patches of kPatchW points along each strip: the bottom and top strips
// (D rows, every column) in runs of kPatchW columns, the left and right strips
// (D columns, rows D .. ny-D-1) in runs of kPatchW rows.  A patch block
// evaluates the pair's first stage (1 or 3) on the patch grown by
// RW = TAPS - R - 1 points into the interior and sideways -- every point any
// stencil of the second stage reaches from the patch (closure rows read the
// first TAPS points of their line, the padded interior row reads
// i-R .. i-R+TAPS-1) -- keeps it in LDS, and then evaluates the second stage
// (2 or 4) on the patch from there, so the frame needs no launch of its own.
// Grown points beyond the frame are interior points (the table's interior
// row): the same values the interior tiles form there.  Each point's
// arithmetic is edge_point's, term for term.
// --------------------------------------------------------------------------
constexpr int kPatchW = 64;

// Number of frame patches of an nx x ny grid with a D-deep frame.
__host__ __device__ inline int frame_patch_count(int nx, int ny, int D)
{
  return 2 * ((nx + kPatchW - 1) / kPatchW) + 2 * ((ny - 2 * D + kPatchW - 1) / kPatchW);
}

struct FramePatch {
  int i0, i1, j0, j1;      // the patch: columns [i0, i1), rows [j0, j1)
  int ei0, ei1, ej0, ej1;  // grown by RW into the interior and sideways, clipped to the grid
};

template <int ORDER> __device__ inline FramePatch frame_patch(int p, int nx, int ny)
{
  using G = Geom<ORDER>;
  constexpr int D = G::CLS + G::R, RW = G::TAPS - G::R - 1;
  const int nb = (nx + kPatchW - 1) / kPatchW;
  FramePatch f;
  if (p < 2 * nb) {  // bottom (p < nb) or top strip
    const bool top = p >= nb;
    const int k = top ? p - nb : p;
    f.i0 = k * kPatchW;
    f.i1 = min(f.i0 + kPatchW, nx);
    f.j0 = top ? ny - D : 0;
    f.j1 = top ? ny : D;
    // sideways: RW, and at least the TAPS-point lines of the closure rows
    // (a narrow last patch of right-closure points reads from nx - TAPS)
    f.ei0 = max(min(f.i0 - RW, nx - G::TAPS), 0);
    f.ei1 = min(max(f.i1 + RW, G::TAPS), nx);
    f.ej0 = top ? max(ny - D - RW, 0) : 0;
    f.ej1 = top ? ny : min(D + RW, ny);
  } else {  // left or right strip, rows D .. ny-D-1
    const int nl = (ny - 2 * D + kPatchW - 1) / kPatchW;
    const bool right = p - 2 * nb >= nl;
    const int k = p - 2 * nb - (right ? nl : 0);
    f.j0 = D + k * kPatchW;
    f.j1 = min(f.j0 + kPatchW, ny - D);
    f.i0 = right ? nx - D : 0;
    f.i1 = right ? nx : D;
    f.ej0 = max(f.j0 - RW, 0);
    f.ej1 = min(f.j1 + RW, ny);
    f.ei0 = right ? max(nx - D - RW, 0) : 0;
    f.ei1 = right ? nx : min(D + RW, nx);
  }
  return f;
}

// Stages 1-2 (PAIR 1) or 3-4 (PAIR 2) of frame patch p; eA / eB are the
// arguments of the first / second stage (stage 1: U -> Y2, also stored on the
// patch for stage 4; stage 2: -> Y3; stage 3: -> Y4, kept in LDS only;
// stage 4: U, Y2, Y3, Y4 -> U').
template <typename T, int ORDER, int PHYS, int MAT, int PAIR, bool KAHAN, int NT>
__device__ __forceinline__ void frame_patch_pair(const StageArgs<T> &eA, const StageArgs<T> &eB, int p)
{
  using G = Geom<ORDER>;
  constexpr int R = G::R, D = G::CLS + G::R, RW = G::TAPS - G::R - 1;
  constexpr int NC = PHYS == PH_ACOUSTIC ? 3 : 1;
  constexpr int MA = PAIR == 1 ? MODE_S1 : MODE_S3, MB = PAIR == 1 ? MODE_S2 : MODE_S4;
  constexpr int FP_N = (D + RW) * (kPatchW + 2 * RW);  // points of the largest grown patch
  static_assert(RW >= R, "the grown patch covers the interior stencil");
  __shared__ T s_f[NC][FP_N];
  const int nx = eA.nx, ny = eA.ny;
  const FramePatch f = frame_patch<ORDER>(p, nx, ny);
  const int ew = f.ei1 - f.ei0, ne = ew * (f.ej1 - f.ej0);
  const long pitch = eA.pitch, plane = eA.plane;
  // first stage on the grown patch
  for (int q = threadIdx.x; q < ne; q += NT) {
    const int jj = f.ej0 + q / ew, ii = f.ei0 + q % ew;
    const long o = jj * pitch + eA.xoff + ii;
    T K[NC], Uc[NC];
    const RkIn<T, NC> rin = rk_load<T, NC, MA, KAHAN>(eA, o);
    edge_rhs<T, ORDER, PHYS, MAT>(eA, eA.yin + eA.xoff, plane, pitch, ii, jj, o, K, Uc);
    const bool on = ii >= f.i0 && ii < f.i1 && jj >= f.j0 && jj < f.j1;
#pragma unroll
    for (int c = 0; c < NC; c++) {
      const T y = (MA == MODE_S1 ? Uc[c] : rin.U[c]) + eA.cy * K[c];
      s_f[c][q] = y;
      if (PAIR == 1 && on) eA.yout[o + c * plane] = y;  // Y2 on the frame (stage 4 reads it)
    }
  }
  __syncthreads();
  // second stage on the patch, its operand from LDS: element (c, i, j) of the
  // grown patch is s_f[c][(j - ej0) * ew + (i - ei0)]
  const T *view = &s_f[0][0] - ((long)f.ej0 * ew + f.ei0);
  const int pw = f.i1 - f.i0, np = pw * (f.j1 - f.j0);
  for (int q = threadIdx.x; q < np; q += NT) {
    const int jj = f.j0 + q / pw, ii = f.i0 + q % pw;
    const long o = jj * pitch + eB.xoff + ii;
    T K[NC], Uc[NC];
    const RkIn<T, NC> rin = rk_load<T, NC, MB, KAHAN>(eB, o);
    edge_rhs<T, ORDER, PHYS, MAT>(eB, view, FP_N, ew, ii, jj, o, K, Uc);
    rk_store<T, NC, MB, KAHAN>(eB, o, K, Uc, rin);
  }
}

// Rows of loads in flight in the pair-1 and pair-2 tiles (sweep on gfx950):
// Rows of loads in flight per wave of a two-stage tile, from sweeps on
// gfx950.  FP32 rows take half the registers, and their tiles run three rows
// ahead (4% faster than two on ensembles of 513^2 grids, where one round of
// tiles fills the GPU).  FP64 pair 1 runs three too (4097^2 order 4: 0.189
// against 0.197 ms at 40-row tiles, still three waves per SIMD; order-6
// rho/K: 0.267 against 0.290).  FP64 pair 2 runs three with rho/K arrays
// (order 6: 0.419 against 0.446 ms, order 4: 0.335 against 0.342) and two
// otherwise (rho = K = 1: the same time; rho = 2 + xy: 0.610 against 0.563).
// The ensemble's tiles follow the same rule (FP64, 16 x 513^2 with rho/K
// arrays: 0.188 against 0.205 ms per step with two rows in both launches).
constexpr int pair_pf(int esz, int pair, bool acoustic_arrays = false)
{
  return esz == 4 || pair == 1 || acoustic_arrays ? 3 : 2;
}

// Blocks [0, n_edge_blocks): frame patches, stages 1-2 (pair 1) or 3-4
// (pair 2) with eA / eB the two stages' arguments; the rest: two-stage
// interior tiles (a), XCD-remapped.
template <typename T, int ORDER, int PHYS, int MAT, int PAIR, bool KAHAN, int NT, int PF, bool NTS>
__global__ __launch_bounds__(NT) void pair2d_kernel(StageArgs<T> eA, StageArgs<T> eB, PairArgs<T> a)
{
  const int b = blockIdx.x;
  if (b < a.n_edge_blocks) {
    if (b < frame_patch_count(eA.nx, eA.ny, Ops<ORDER>::CLS + Ops<ORDER>::R))
      frame_patch_pair<T, ORDER, PHYS, MAT, PAIR, KAHAN, NT>(eA, eB, b);
    return;
  }
  const int bi = b - a.n_edge_blocks;
  const int nt = a.n_tiles;
  const int xcd = bi % kNumXcd, local = bi / kNumXcd;
  const int q = nt / kNumXcd, r = nt % kNumXcd;
  const int tile = xcd * q + min(xcd, r) + local;
  pair_tile_rw<T, ORDER, PHYS, MAT, PAIR, KAHAN, NT, PF, NTS>(a, tile);
}

// --------------------------------------------------------------------------
// 1D problems: one thread per point, table-driven rows.
// --------------------------------------------------------------------------
template <typename T, int ORDER, int PHYS, int MODE, bool KAHAN>
__global__ __launch_bounds__(256) void stage1d_kernel(StageArgs<T> a)
{
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= a.nx) return;
  const int nx = a.nx;
  const long o = a.xoff + i;
  const T *f0 = a.yin + a.xoff;
  const T hq = a.hix * T(Ops<ORDER>::hqinv[0]);
  if constexpr (PHYS == PH_ADV1D) {
    // advection_{l,i,r} (advection_rhs.h:13-51) + SAT_bc_{l,r} (:95-119), with
    // the velocity a(i) of this point (per-point field, or the constant)
    const T av = a.velx ? a.velx[o] : a.ax;
    const T u = f0[i];
    T k = -av * dline<T, ORDER>(f0, 1, nx, i, a.hix);
    if (i == 0) k += T(0.5) * (T(-0.5) * (av + fabs(av))) * (hq * u);
    if (i == nx - 1) k += T(0.5) * (T(0.5) * (av - fabs(av))) * (hq * u);
    T K[1] = {k}, Uc[1] = {u};
    rk_point_exact<T, 1, MODE, KAHAN>(a, o, K, Uc);
  } else {
    // reflection_{l,i,r} (reflection_rhs.h:10-47) + proj_dirichlet_bc_{l,r} (:108-120)
    const T *f1 = f0 + a.plane;
    T K[2], Uc[2] = {f0[i], f1[i]};
    K[1] = dline<T, ORDER>(f0, 1, nx, i, a.hix);
    K[0] = dline<T, ORDER>(f1, 1, nx, i, a.hix);
    if (i == 0 || i == nx - 1) K[0] = T(0);
    rk_point_exact<T, 2, MODE, KAHAN>(a, o, K, Uc);
  }
}

// --------------------------------------------------------------------------
// Energy / H-norm reduction (FP64 accumulation), deterministic two passes.
// --------------------------------------------------------------------------
template <typename T> struct NormArgs {
  const double *aos; // AoS host-layout vector (when non-null, read instead of u)
  const T *u;       // state planes
  const T *comp;    // Kahan compensation or null
  const T *rinv, *bulk;
  const double *xs, *ys;
  long plane;
  int pitch, xoff, nx, ny, nc;
  int order, is1d, material, weighted;  // weighted: material energy weights
  double hx, hy;
};

__device__ __forceinline__ double hq_of(int order, int k)
{
  if (order == 2) return c_edge<double, 2>.hq[0];
  if (order == 4) return c_edge<double, 4>.hq[k];
  return c_edge<double, 6>.hq[k];
}

__device__ __forceinline__ double hweight(int order, int n, int i, double h)
{
  const int cls = order == 2 ? 1 : (order == 4 ? 4 : 6);
  if (i < cls) return h * hq_of(order, i);
  if (i >= n - cls) return h * hq_of(order, n - 1 - i);
  return h;
}

__device__ __forceinline__ double block_sum_256(double v, double *sh)
{
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  if (lane == 0) sh[wv] = v;
  __syncthreads();
  double s = 0;
  if (threadIdx.x == 0) s = (sh[0] + sh[1]) + (sh[2] + sh[3]);
  return s;
}

template <typename T>
__global__ __launch_bounds__(256) void norm_partial_kernel(NormArgs<T> a, double *partials)
{
  __shared__ double sh[4];
  double acc = 0;
  const long npts = (long)a.nx * a.ny;
  for (long k = blockIdx.x * 256L + threadIdx.x; k < npts; k += (long)gridDim.x * 256) {
    const int j = (int)(k / a.nx), i = (int)(k - (long)j * a.nx);
    const long o = (long)j * a.pitch + a.xoff + i;
    double w = hweight(a.order, a.nx, i, a.hx);
    if (!a.is1d) w *= hweight(a.order, a.ny, j, a.hy);
    double v[3] = {0, 0, 0};
    for (int c = 0; c < a.nc; c++) {
      double x;
      if (a.aos) {
        x = a.aos[k * a.nc + c];
      } else {
        x = (double)a.u[c * a.plane + o];
        if (a.comp) x -= (double)a.comp[c * a.plane + o];
      }
      v[c] = x;
    }
    double e;
    if (a.weighted && a.nc == 3) {
      double rho = 1.0, kinv = 1.0;
      if (a.material == M_ARR) {
        rho = 1.0 / (double)a.rinv[o];
        kinv = 1.0 / (double)a.bulk[o];
      } else if (a.material == M_REF) {
        rho = 2 + a.xs[i] * a.ys[j];
      }
      e = rho * (v[0] * v[0] + v[1] * v[1]) + kinv * v[2] * v[2];
    } else {
      e = v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
    }
    acc += w * e;
  }
  const double s = block_sum_256(acc, sh);
  if (threadIdx.x == 0) partials[blockIdx.x] = s;
}

static __global__ __launch_bounds__(256) void norm_final_kernel(const double *partials, int n, double *out)
{
  __shared__ double sh[4];
  double acc = 0;
  for (int k = threadIdx.x; k < n; k += 256) acc += partials[k];
  const double s = block_sum_256(acc, sh);
  if (threadIdx.x == 0) out[0] = s;
}

// --------------------------------------------------------------------------
// AoS (reference layout, double) <-> planar device state.
// --------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(256) void aos_to_planar_kernel(const double *__restrict__ aos, T *__restrict__ U,
                                                            T *__restrict__ comp, int nx, int ny, int nc,
                                                            int pitch, int xoff, long plane)
{
  const long k = blockIdx.x * 256L + threadIdx.x;
  if (k >= (long)nx * ny) return;
  const int j = (int)(k / nx), i = (int)(k - (long)j * nx);
  const long o = (long)j * pitch + xoff + i;
  for (int c = 0; c < nc; c++) {
    const double v = aos[k * nc + c];
    const T t = (T)v;
    U[c * plane + o] = t;
    if (comp) comp[c * plane + o] = (T)((double)t - v);  // value = U - comp
  }
}

template <typename T>
__global__ __launch_bounds__(256) void planar_to_aos_kernel(const T *__restrict__ U, const T *__restrict__ comp,
                                                            double *__restrict__ aos, int nx, int ny, int nc,
                                                            int pitch, int xoff, long plane)
{
  const long k = blockIdx.x * 256L + threadIdx.x;
  if (k >= (long)nx * ny) return;
  const int j = (int)(k / nx), i = (int)(k - (long)j * nx);
  const long o = (long)j * pitch + xoff + i;
  for (int c = 0; c < nc; c++) {
    double v = (double)U[c * plane + o];
    if (comp) v -= (double)comp[c * plane + o];
    aos[k * nc + c] = v;
  }
}

// Every stride-th point in both directions (injection onto a nested grid).
template <typename T>
__global__ __launch_bounds__(256) void planar_to_aos_strided_kernel(const T *__restrict__ U, const T *__restrict__ comp,
                                                                    double *__restrict__ aos, int nx, int ny, int nc,
                                                                    int pitch, int xoff, long plane, int stride)
{
  const int cx = (nx - 1) / stride + 1, cy = (ny - 1) / stride + 1;
  const long k = blockIdx.x * 256L + threadIdx.x;
  if (k >= (long)cx * cy) return;
  const int cj = (int)(k / cx), ci = (int)(k - (long)cj * cx);
  const long o = (long)cj * stride * pitch + xoff + (long)ci * stride;
  for (int c = 0; c < nc; c++) {
    double v = (double)U[c * plane + o];
    if (comp) v -= (double)comp[c * plane + o];
    aos[k * nc + c] = v;
  }
}

template <typename T>
__global__ __launch_bounds__(256) void matrix_to_plane_kernel(const double *__restrict__ m, T *__restrict__ P,
                                                              int nx, int ny, int pitch, int xoff)
{
  const long k = blockIdx.x * 256L + threadIdx.x;
  if (k >= (long)nx * ny) return;
  const int j = (int)(k / nx), i = (int)(k - (long)j * nx);
  P[(long)j * pitch + xoff + i] = (T)m[k];
}

// --------------------------------------------------------------------------
// Host-side launch table entry.
// --------------------------------------------------------------------------
struct LaunchGeom {
  int grid;   // blocks
  int block;  // threads
};

using StageLaunchFn = hipError_t (*)(int mode, const void *args, LaunchGeom g, hipStream_t s);

template <typename T, int ORDER, int PHYS, int MAT, bool KAHAN, int BX>
struct Launcher2D {
  template <int MODE> static hipError_t go(const StageArgs<T> &a, LaunchGeom g, hipStream_t s)
  {
    hipLaunchKernelGGL((stage2d_kernel<T, ORDER, PHYS, MAT, MODE, KAHAN, BX>), dim3(g.grid), dim3(BX), 0, s, a);
    return hipGetLastError();
  }
  static hipError_t launch(int mode, const void *args, LaunchGeom g, hipStream_t s)
  {
    const StageArgs<T> &a = *static_cast<const StageArgs<T> *>(args);
    switch (mode) {
      case MODE_RHS: return go<MODE_RHS>(a, g, s);
      case MODE_S1: return go<MODE_S1>(a, g, s);
      case MODE_S2: return go<MODE_S2>(a, g, s);
      case MODE_S3: return go<MODE_S3>(a, g, s);
      default: return go<MODE_S4>(a, g, s);
    }
  }
};

template <typename T, int ORDER, int PHYS, bool KAHAN>
struct Launcher1D {
  template <int MODE> static hipError_t go(const StageArgs<T> &a, LaunchGeom g, hipStream_t s)
  {
    hipLaunchKernelGGL((stage1d_kernel<T, ORDER, PHYS, MODE, KAHAN>), dim3(g.grid), dim3(256), 0, s, a);
    return hipGetLastError();
  }
  static hipError_t launch(int mode, const void *args, LaunchGeom g, hipStream_t s)
  {
    const StageArgs<T> &a = *static_cast<const StageArgs<T> *>(args);
    switch (mode) {
      case MODE_RHS: return go<MODE_RHS>(a, g, s);
      case MODE_S1: return go<MODE_S1>(a, g, s);
      case MODE_S2: return go<MODE_S2>(a, g, s);
      case MODE_S3: return go<MODE_S3>(a, g, s);
      default: return go<MODE_S4>(a, g, s);
    }
  }
};

// The two launches of a two-stage step: part 1 = pair2d_kernel pair 1
// (frame patches and interior tiles, stages 1-2), part 3 = pair 2 (stages
// 3-4).  stage_a / stage_b are the StageArgs of the launch's two stages.
using PairLaunchFn = hipError_t (*)(int part, const void *stage_a, const void *stage_b, const void *pair_args,
                                    int grid, hipStream_t s);

// PF1 / PF2: rows of loads in flight ahead of the row being computed in the
// pair-1 (one stencil operand per row) and pair-2 (operand + U) launches.
template <typename T, int ORDER, int PHYS, int MAT, bool KAHAN, int NT, int PF1, int PF2> struct PairLauncher {
  static hipError_t launch(int part, const void *sa, const void *sb, const void *pa, int grid, hipStream_t s)
  {
    const StageArgs<T> &ea = *static_cast<const StageArgs<T> *>(sa);
    const StageArgs<T> &eb = *static_cast<const StageArgs<T> *>(sb);
    const PairArgs<T> &a = *static_cast<const PairArgs<T> *>(pa);
    if (part == 1 && a.nts)
      hipLaunchKernelGGL((pair2d_kernel<T, ORDER, PHYS, MAT, 1, KAHAN, NT, PF1, true>), dim3(grid), dim3(NT), 0, s, ea,
                         eb, a);
    else if (part == 1)
      hipLaunchKernelGGL((pair2d_kernel<T, ORDER, PHYS, MAT, 1, KAHAN, NT, PF1, false>), dim3(grid), dim3(NT), 0, s, ea,
                         eb, a);
    else if (a.nts)
      hipLaunchKernelGGL((pair2d_kernel<T, ORDER, PHYS, MAT, 2, KAHAN, NT, PF2, true>), dim3(grid), dim3(NT), 0, s, ea,
                         eb, a);
    else
      hipLaunchKernelGGL((pair2d_kernel<T, ORDER, PHYS, MAT, 2, KAHAN, NT, PF2, false>), dim3(grid), dim3(NT), 0, s, ea,
                         eb, a);
    return hipGetLastError();
  }
};

// Per-order, per-precision registries (sbp_inst_o{2,4,6}_{f64,f32}.hip).
// The FP64 units are compiled without FMA contraction so that every
// right-hand side evaluates in the same operation order, and with the same
// roundings, as the reference's loops.  kahan selects the compensated FP32 form.
StageLaunchFn launcher_o2_f64(int phys, int mat, int bx);
StageLaunchFn launcher_o4_f64(int phys, int mat, int bx);
StageLaunchFn launcher_o6_f64(int phys, int mat, int bx);
StageLaunchFn launcher_o2_f32(int phys, int mat, bool kahan, int bx);
StageLaunchFn launcher_o4_f32(int phys, int mat, bool kahan, int bx);
StageLaunchFn launcher_o6_f32(int phys, int mat, bool kahan, int bx);
// Two-stage launchers of 2D acoustics (unit, array or reference materials)
// and 2D advection (constant or per-point velocity); nullptr where none is built.
PairLaunchFn pair_launcher_o2_f64(int phys, int mat);
PairLaunchFn pair_launcher_o4_f64(int phys, int mat);
PairLaunchFn pair_launcher_o6_f64(int phys, int mat);
PairLaunchFn pair_launcher_o2_f32(int phys, int mat, bool kahan);
PairLaunchFn pair_launcher_o4_f32(int phys, int mat, bool kahan);
PairLaunchFn pair_launcher_o6_f32(int phys, int mat, bool kahan);

}  // namespace sbp
