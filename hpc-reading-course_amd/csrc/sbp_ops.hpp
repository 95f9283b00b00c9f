// sbp_ops.hpp -- SBP first-derivative operators, norms and inverse norms as
// compile-time tables, plus the "edge table" the closure kernels index at run
// time.
//
// The coefficients are the definition of the central SBP operators of
// orders 2/4/6 and must equal code/include/sbpops/D1_central.h:243-273,
// H_central.h:193-212 and HI_central.h:128-147 bit for bit; they are written
// as the same decimal quotients so the compiler rounds them identically.
#pragma once

namespace sbp {

template <int ORDER> struct Ops;

template <> struct Ops<2> {
  static constexpr int R = 1, CLS = 1, W = 2;
  static constexpr double inner[3] = {-1. / 2, 0, 1. / 2};
  static constexpr double Q[1][2] = {{-1., 1}};
  static constexpr double hq[1] = {1. / 2};
  static constexpr double hqinv[1] = {2.};
};

template <> struct Ops<4> {
  static constexpr int R = 2, CLS = 4, W = 6;
  static constexpr double inner[5] = {1. / 12, -2. / 3, 0., 2. / 3, -1. / 12};
  static constexpr double Q[4][6] = {{-24. / 17, 59. / 34, -4. / 17, -3. / 34, 0, 0},
                                     {-1. / 2, 0, 1. / 2, 0, 0, 0},
                                     {4. / 43, -59. / 86, 0, 59. / 86, -4. / 43, 0},
                                     {3. / 98, 0, -59. / 98, 0, 32. / 49, -4. / 49}};
  static constexpr double hq[4] = {17. / 48, 59. / 48, 43. / 48, 49. / 48};
  static constexpr double hqinv[4] = {48. / 17, 48. / 59, 48. / 43, 48. / 49};
};

template <> struct Ops<6> {
  static constexpr int R = 3, CLS = 6, W = 9;
  static constexpr double inner[7] = {-1. / 60, 3. / 20, -3. / 4, 0, 3. / 4, -3. / 20, 1. / 60};
  static constexpr double Q[6][9] = {
      {-21600. / 13649, 104009. / 54596, 30443. / 81894, -33311. / 27298, 16863. / 27298, -15025. / 163788, 0, 0, 0},
      {-104009. / 240260, 0, -311. / 72078, 20229. / 24026, -24337. / 48052, 36661. / 360390, 0, 0, 0},
      {-30443. / 162660, 311. / 32532, 0, -11155. / 16266, 41287. / 32532, -21999. / 54220, 0, 0, 0},
      {33311. / 107180, -20229. / 21436, 485. / 1398, 0, 4147. / 21436, 25427. / 321540, 72. / 5359, 0, 0},
      {-16863. / 78770, 24337. / 31508, -41287. / 47262, -4147. / 15754, 0, 342523. / 472620, -1296. / 7877, 144. / 7877, 0},
      {15025. / 525612, -36661. / 262806, 21999. / 87602, -25427. / 262806, -342523. / 525612, 0, 32400. / 43801, -6480. / 43801, 720. / 43801}};
  static constexpr double hq[6] = {13649. / 43200, 12013. / 8640, 2711. / 4320, 5359. / 4320, 7877. / 8640, 43801. / 43200};
  static constexpr double hqinv[6] = {43200. / 13649, 8640. / 12013, 4320. / 2711, 4320. / 5359, 8640. / 7877, 43200. / 43801};
};

constexpr int imax(int a, int b) { return a > b ? a : b; }

// Number of taps of a table-driven row: wide enough for the closure rows and
// the interior stencil.
template <int ORDER> struct Geom {
  static constexpr int R = Ops<ORDER>::R, CLS = Ops<ORDER>::CLS, W = Ops<ORDER>::W;
  static constexpr int TAPS = imax(W, 2 * R + 1);
  static constexpr int ROWS = 2 * CLS + 1;
};

// Edge table: rows 0..CLS-1 are the left closure rows (first tap at index 0),
// rows CLS..2CLS-1 are the right closure rows for m = N-1-i = 0..CLS-1 stored
// as the negated mirror -Q[m][W-1-s] (D1_central.h:95-104) right-aligned so the
// first tap sits at N-TAPS, and row 2CLS is the interior stencil (first tap at
// i-R), zero padded.  Every row reads TAPS consecutive points, all in range.
template <typename T, int ORDER> struct EdgeTable {
  T c[Geom<ORDER>::ROWS][Geom<ORDER>::TAPS];
  T hq[Ops<ORDER>::CLS];
};

template <typename T, int ORDER> constexpr EdgeTable<T, ORDER> make_edge_table()
{
  using O = Ops<ORDER>;
  using G = Geom<ORDER>;
  EdgeTable<T, ORDER> t{};
  for (int a = 0; a < G::ROWS; a++)
    for (int s = 0; s < G::TAPS; s++) t.c[a][s] = T(0);
  for (int i = 0; i < G::CLS; i++)
    for (int s = 0; s < G::W; s++) t.c[i][s] = T(O::Q[i][s]);
  for (int m = 0; m < G::CLS; m++)
    for (int s = 0; s < G::W; s++) t.c[G::CLS + m][G::TAPS - G::W + s] = T(-O::Q[m][G::W - 1 - s]);
  for (int s = 0; s < 2 * G::R + 1; s++) t.c[2 * G::CLS][s] = T(O::inner[s]);
  for (int k = 0; k < G::CLS; k++) t.hq[k] = T(O::hq[k]);
  return t;
}

}  // namespace sbp
