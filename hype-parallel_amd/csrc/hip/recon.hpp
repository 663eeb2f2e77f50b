// recon.hpp -- modal basis, quadrature tables and WENO reconstructions on the
// device (HyPE-2D/reconstruction.c, tables of hype.c:194-305).
//
// Tables are compile-time constants produced by the same basis formulas as
// the reference, so every basis value is bit-identical to the one the
// reference computes at start-up.  Sums over basis functions keep the
// reference's order k = 0..nDOF-1 starting from 0.0; terms whose basis value
// is exactly zero are skipped, which cannot change an IEEE result for finite
// coefficients (the running sum starts at +0.0 and can never become -0.0, so
// adding a signed zero is the identity).
#pragma once
#ifndef HYPE_KNS
#define HYPE_KNS hype  // kernel namespace; the contracted build uses its own
#endif
#include <hip/hip_runtime.h>

#include "fdiv.hpp"

namespace HYPE_KNS {

constexpr double GPT = 0.28867513459481287;  // hype.c:214-230

// reconstruction.c:11-47
constexpr double basis(double x, double y, int n) {
  return n == 0   ? 1.0
         : n == 1 ? x
         : n == 2 ? y
         : n == 3 ? x * x - 1. / 12.
         : n == 4 ? y * y - 1. / 12.
         : n == 5 ? x * y
         : n == 6 ? x * (x * x - 3. / 20.)
         : n == 7 ? y * (y * y - 3. / 20.)
         : n == 8 ? y * (x * x - 1. / 12.)
         : n == 9 ? x * (y * y - 1. / 12.)
                  : 0.0;
}
// reconstruction.c:53-100
constexpr double basis_gx(double x, double y, int n) {
  return n == 1   ? 1.0
         : n == 3 ? 2.0 * x
         : n == 5 ? y
         : n == 6 ? 3.0 * x * x - 3. / 20.
         : n == 8 ? 2.0 * x * y
         : n == 9 ? (y * y - 1. / 12.)
                  : 0.0;
}
constexpr double basis_gy(double x, double y, int n) {
  return n == 2   ? 1.0
         : n == 4 ? 2.0 * y
         : n == 5 ? x
         : n == 7 ? 3.0 * y * y - 3. / 20.
         : n == 8 ? (x * x - 1. / 12.)
         : n == 9 ? 2.0 * x * y
                  : 0.0;
}

// Quadrature for order N (hype.c:56-68, 203-230).  Face points are stored in
// the reference's node order (hype.c:207-229): right face points, left face
// points, top face points, bottom face points; the admissibility nodes are
// exactly these points.
template <int N>
struct Quad {
  static constexpr int NDOF = (N + 1) * (N + 2) / 2;
  static constexpr int NGF = N < 2 ? 1 : 2;
  static constexpr int NGV = N < 2 ? 1 : 4;
  static constexpr int NNODE = 4 * NGF;
  static constexpr double WF = N < 2 ? 1.0 : 0.5;   // wGP_Face
  static constexpr double WV = N < 2 ? 1.0 : 0.25;  // wGP_Vol
  static constexpr double xgf(int q) { return N < 2 ? 0.0 : (q == 0 ? -GPT : GPT); }
  static constexpr int R(int q) { return q; }
  static constexpr int L(int q) { return NGF + q; }
  static constexpr int T(int q) { return 2 * NGF + q; }
  static constexpr int B(int q) { return 3 * NGF + q; }
  static constexpr double nodex(int p) {
    return p < NGF ? 0.5 : p < 2 * NGF ? -0.5 : xgf(p % NGF);
  }
  static constexpr double nodey(int p) {
    return p < 2 * NGF ? xgf(p % NGF) : p < 3 * NGF ? 0.5 : -0.5;
  }
  // volume points (hype.c:217-220): (-g,-g), (-g,g), (g,-g), (g,g)
  static constexpr double volx(int q) { return N < 2 ? 0.0 : (q < 2 ? -GPT : GPT); }
  static constexpr double voly(int q) { return N < 2 ? 0.0 : ((q & 1) ? GPT : -GPT); }
  static constexpr double phin(int p, int k) { return basis(nodex(p), nodey(p), k); }
  static constexpr double phiv(int q, int k) { return basis(volx(q), voly(q), k); }
  static constexpr double gphx(int q, int k) { return basis_gx(volx(q), voly(q), k); }
  static constexpr double gphy(int q, int k) { return basis_gy(volx(q), voly(q), k); }
};

// sum_k s[k]*phi(k) in the reference's order, skipping exact-zero phi
#define HYPE_BASIS_SUM(out, s, NDOF, PHI)               \
  do {                                                  \
    double acc_ = 0.0;                                  \
    _Pragma("unroll") for (int k_ = 0; k_ < NDOF; ++k_) \
      if (PHI(k_) != 0.0) acc_ = acc_ + s[k_] * PHI(k_); \
    out = acc_;                                         \
  } while (0)

// reconstruction.c:106-117
__device__ __forceinline__ double minmod(double a, double b) {
  if (a * b < 0.0) return 0.0;
  if (fabs(a) < fabs(b)) return a;
  return b;
}

__device__ __forceinline__ double pow4(double a) {  // reconstruction.c:123-127
  double a2 = a * a;
  return a2 * a2;
}

// weno3 with every division done by fdiv.hpp's shared-reciprocal sequence.
// Same expressions and operation order as the reference weno3
// (reconstruction.c:144-221), so the coefficients are bitwise those of the
// IEEE-division version whenever all divisors are in range; returns false
// otherwise (the caller then recomputes with plain division).  The divisors
// are pow4(IS + 1e-12) >= 1e-48 and the weight sums; a group is in range when
// the sum of its pow4 terms is at most 2^380 (then every weight and sum lies
// in [2^-380, 2^380]).
__device__ __forceinline__ bool weno3_fast(const double* U_x, const double* U_y, const double* U_xy,
                                           double* coeffs) {
  const double central_cell_wt = 100.0;
  const double r13_3 = 13. / 3.;
  double u_0 = U_xy[0];
  double u_ip1 = U_x[3], u_jp1 = U_y[3], u_ip1jp1 = U_xy[1];
  double u_im1 = U_x[1], u_jm1 = U_y[1], u_ip1jm1 = U_xy[2];
  double u_ip2 = U_x[4], u_jp2 = U_y[4], u_im1jp1 = U_xy[3];
  double u_im2 = U_x[0], u_jm2 = U_y[0], u_im1jm1 = U_xy[4];
  double ux0, ux1, ux2, uxx0, uxx1, uxx2, d0, d1, d2, d3, w0, w1, w2, w3, total, yt, temp;
  bool ok;

  ux0 = -2.0 * u_im1 + 0.5 * u_im2 + 1.5 * u_0;
  uxx0 = 0.5 * (u_im2 - 2.0 * u_im1 + u_0);
  ux1 = 0.5 * (u_ip1 - u_im1);
  uxx1 = 0.5 * (u_im1 - 2.0 * u_0 + u_ip1);
  ux2 = -1.5 * u_0 + 2.0 * u_ip1 - 0.5 * u_ip2;
  uxx2 = 0.5 * (u_0 - 2.0 * u_ip1 + u_ip2);
  d0 = pow4(ux0 * ux0 + r13_3 * uxx0 * uxx0 + SMALL_NUM);
  d1 = pow4(ux1 * ux1 + r13_3 * uxx1 * uxx1 + SMALL_NUM);
  d2 = pow4(ux2 * ux2 + r13_3 * uxx2 * uxx2 + SMALL_NUM);
  ok = (d0 + d1 + d2) <= FDIV_HI;
  w0 = inv_y(d0, recip_nr(d0));
  w1 = div_y(central_cell_wt, d1, recip_nr(d1));
  w2 = inv_y(d2, recip_nr(d2));
  total = w0 + w1 + w2;
  yt = recip_nr(total);
  w0 = div_y(w0, total, yt);
  w1 = div_y(w1, total, yt);
  w2 = div_y(w2, total, yt);
  coeffs[0] = u_0;
  coeffs[1] = w0 * ux0 + w1 * ux1 + w2 * ux2;
  coeffs[3] = w0 * uxx0 + w1 * uxx1 + w2 * uxx2;

  ux0 = -2.0 * u_jm1 + 0.5 * u_jm2 + 1.5 * u_0;
  uxx0 = 0.5 * (u_jm2 - 2.0 * u_jm1 + u_0);
  ux1 = 0.5 * (u_jp1 - u_jm1);
  uxx1 = 0.5 * (u_jm1 - 2.0 * u_0 + u_jp1);
  ux2 = -1.5 * u_0 + 2.0 * u_jp1 - 0.5 * u_jp2;
  uxx2 = 0.5 * (u_0 - 2.0 * u_jp1 + u_jp2);
  d0 = pow4(ux0 * ux0 + r13_3 * uxx0 * uxx0 + SMALL_NUM);
  d1 = pow4(ux1 * ux1 + r13_3 * uxx1 * uxx1 + SMALL_NUM);
  d2 = pow4(ux2 * ux2 + r13_3 * uxx2 * uxx2 + SMALL_NUM);
  ok = ok && (d0 + d1 + d2) <= FDIV_HI;
  w0 = inv_y(d0, recip_nr(d0));
  w1 = div_y(central_cell_wt, d1, recip_nr(d1));
  w2 = inv_y(d2, recip_nr(d2));
  total = w0 + w1 + w2;
  yt = recip_nr(total);
  w0 = div_y(w0, total, yt);
  w1 = div_y(w1, total, yt);
  w2 = div_y(w2, total, yt);
  coeffs[2] = w0 * ux0 + w1 * ux1 + w2 * ux2;
  coeffs[4] = w0 * uxx0 + w1 * uxx1 + w2 * uxx2;

  double c1 = coeffs[1], c2 = coeffs[2], c3 = coeffs[3], c4 = coeffs[4];
  double uxy0 = u_ip1jp1 - u_0 - c1 - c2 - c3 - c4;
  double uxy1 = -u_ip1jm1 + u_0 + c1 - c2 + c3 + c4;
  double uxy2 = -u_im1jp1 + u_0 - c1 + c2 + c3 + c4;
  double uxy3 = u_im1jm1 - u_0 + c1 + c2 - c3 - c4;
  temp = 4.0 * (c3 * c3 + c4 * c4);
  d0 = pow4(temp + uxy0 * uxy0 + SMALL_NUM);
  d1 = pow4(temp + uxy1 * uxy1 + SMALL_NUM);
  d2 = pow4(temp + uxy2 * uxy2 + SMALL_NUM);
  d3 = pow4(temp + uxy3 * uxy3 + SMALL_NUM);
  ok = ok && (d0 + d1 + d2 + d3) <= FDIV_HI;
  w0 = inv_y(d0, recip_nr(d0));
  w1 = inv_y(d1, recip_nr(d1));
  w2 = inv_y(d2, recip_nr(d2));
  w3 = inv_y(d3, recip_nr(d3));
  total = w0 + w1 + w2 + w3;
  yt = recip_nr(total);
  w0 = div_y(w0, total, yt);
  w1 = div_y(w1, total, yt);
  w2 = div_y(w2, total, yt);
  w3 = div_y(w3, total, yt);
  coeffs[5] = w0 * uxy0 + w1 * uxy1 + w2 * uxy2 + w3 * uxy3;
  return ok;
}

// reconstruction.c:223-357: CWENO with one fourth-degree central candidate
// and five third-degree candidates (gamma_hi = gamma_lo = 0.85)
__device__ __forceinline__ void weno4(const double* U_x, const double* U_y, const double* U_xy, double* coeffs) {
  constexpr double r1_6 = 1. / 6., r13_3 = 13. / 3., r7_6 = 7. / 6., r11_120 = 11. / 120., r1_12 = 1. / 12.,
                   r41_60 = 41. / 60.;
  const double gammaHi = 0.85, gammaLo = 0.85;
  double u_0 = U_xy[0];
  double u_ip1 = U_x[3], u_jp1 = U_y[3], u_ip1jp1 = U_xy[1];
  double u_im1 = U_x[1], u_jm1 = U_y[1], u_ip1jm1 = U_xy[2];
  double u_ip2 = U_x[4], u_jp2 = U_y[4], u_im1jp1 = U_xy[3];
  double u_im2 = U_x[0], u_jm2 = U_y[0], u_im1jm1 = U_xy[4];
  double gR3[5], xR3[5], yR3[5], xxR3[5], yyR3[5], xyR3[5], wR3[5];
  const double gR4 = gammaHi;
  gR3[0] = (1.0 - gammaHi) * gammaLo;
#pragma unroll
  for (int i = 1; i < 5; ++i) gR3[i] = 0.25 * (1 - gammaHi) * (1 - gammaLo);

  const double xR4 = r41_60 * (-u_im1 + u_ip1) + r11_120 * (u_im2 - u_ip2);
  const double yR4 = r41_60 * (-u_jm1 + u_jp1) + r11_120 * (u_jm2 - u_jp2);
  const double xxR4 = -u_0 + 0.5 * (u_im1 + u_ip1);
  const double yyR4 = -u_0 + 0.5 * (u_jm1 + u_jp1);
  const double xyR4 = 0.25 * (u_im1jm1 - u_im1jp1 - u_ip1jm1 + u_ip1jp1);
  const double xxxR4 = r1_6 * (u_im1 - u_ip1) + r1_12 * (u_ip2 - u_im2);
  const double yyyR4 = r1_6 * (u_jm1 - u_jp1) + r1_12 * (u_jp2 - u_jm2);
  const double xxyR4 = 0.25 * (-u_im1jm1 + u_im1jp1 - u_ip1jm1 + u_ip1jp1) + 0.5 * (u_jm1 - u_jp1);
  const double xyyR4 = 0.5 * (u_im1 - u_ip1) + 0.25 * (u_ip1jp1 - u_im1jm1 - u_im1jp1 + u_ip1jm1);
  const double IS_R4 = (xR4 + 0.1 * xxxR4) * (xR4 + 0.1 * xxxR4) + r13_3 * (xxR4 * xxR4 + yyR4 * yyR4) +
                       (yR4 + 0.1 * yyyR4) * (yR4 + 0.1 * yyyR4) + 39.05 * (xxxR4 * xxxR4 + yyyR4 * yyyR4) +
                       r7_6 * xyR4 * xyR4 + 4.7 * (xxyR4 * xxyR4 + xyyR4 * xyyR4);
  double wR4 = gR4 / (pow4(IS_R4 + SMALL_NUM));
  double sum = wR4;

  xR3[0] = 0.5 * (-u_im1 + u_ip1);
  yR3[0] = 0.5 * (-u_jm1 + u_jp1);
  xxR3[0] = -u_0 + 0.5 * (u_im1 + u_ip1);
  yyR3[0] = -u_0 + 0.5 * (u_jm1 + u_jp1);
  xyR3[0] = 0.25 * (u_im1jm1 - u_im1jp1 - u_ip1jm1 + u_ip1jp1);

  xR3[1] = 1.5 * u_0 - 2.0 * u_im1 + 0.5 * u_im2;
  yR3[1] = 0.5 * (-u_jm1 + u_jp1);
  xxR3[1] = 0.5 * u_0 - u_im1 + 0.5 * u_im2;
  yyR3[1] = -u_0 + 0.5 * u_jm1 + 0.5 * u_jp1;
  xyR3[1] = 0.5 * (u_im1jm1 - u_im1jp1 - u_jm1 + u_jp1);

  xR3[2] = -1.5 * u_0 + 2.0 * u_ip1 - 0.5 * u_ip2;
  yR3[2] = 0.5 * (-u_jm1 + u_jp1);
  xxR3[2] = 0.5 * u_0 - u_ip1 + 0.5 * u_ip2;
  yyR3[2] = -u_0 + 0.5 * u_jm1 + 0.5 * u_jp1;
  xyR3[2] = 0.5 * (-u_ip1jm1 + u_ip1jp1 + u_jm1 - u_jp1);

  xR3[3] = 0.5 * (-u_im1 + u_ip1);
  yR3[3] = 1.5 * u_0 - 2.0 * u_jm1 + 0.5 * u_jm2;
  xxR3[3] = -u_0 + 0.5 * u_im1 + 0.5 * u_ip1;
  yyR3[3] = 0.5 * u_0 - u_jm1 + 0.5 * u_jm2;
  xyR3[3] = 0.5 * (-u_im1 + u_im1jm1 + u_ip1 - u_ip1jm1);

  xR3[4] = 0.5 * (-u_im1 + u_ip1);
  yR3[4] = -1.5 * u_0 + 2.0 * u_jp1 - 0.5 * u_jp2;
  xxR3[4] = -u_0 + 0.5 * u_im1 + 0.5 * u_ip1;
  yyR3[4] = 0.5 * u_0 - u_jp1 + 0.5 * u_jp2;
  xyR3[4] = 0.5 * (u_im1 - u_im1jp1 - u_ip1 + u_ip1jp1);

#pragma unroll
  for (int i = 0; i < 5; ++i) {
    const double IS = xR3[i] * xR3[i] + yR3[i] * yR3[i] + r13_3 * (xxR3[i] * xxR3[i] + yyR3[i] * yyR3[i]) +
                      r7_6 * xyR3[i] * xyR3[i];
    wR3[i] = gR3[i] / (pow4(IS + SMALL_NUM));
    sum = sum + wR3[i];
  }
  wR4 = wR4 / sum;
#pragma unroll
  for (int i = 0; i < 5; ++i) wR3[i] = wR3[i] / sum;
  const double wt_ratio = wR4 / gR4;

  auto coef = [&](double R4, const double* R3) {
    return wt_ratio * (R4 - gR3[0] * R3[0] - gR3[1] * R3[1] - gR3[2] * R3[2] - gR3[3] * R3[3] - gR3[4] * R3[4]) +
           wR3[0] * R3[0] + wR3[1] * R3[1] + wR3[2] * R3[2] + wR3[3] * R3[3] + wR3[4] * R3[4];
  };
  coeffs[0] = u_0;
  coeffs[1] = coef(xR4, xR3);
  coeffs[2] = coef(yR4, yR3);
  coeffs[3] = coef(xxR4, xxR3);
  coeffs[4] = coef(yyR4, yyR3);
  coeffs[5] = coef(xyR4, xyR3);
  coeffs[6] = wt_ratio * xxxR4;
  coeffs[7] = wt_ratio * yyyR4;
  coeffs[8] = wt_ratio * xxyR4;
  coeffs[9] = wt_ratio * xyyR4;
}

// Stencil of one component: x[0..4] = u(i-2..i+2, j), y[0..4] = u(i, j-2..j+2),
// d[0..4] = u(i,j), u(i+1,j+1), u(i+1,j-1), u(i-1,j+1), u(i-1,j-1)
// (rhs_function_prim.c:332-348).
template <int N>
__device__ __forceinline__ void reconstruct(const double* U_x, const double* U_y, const double* U_xy,
                                            double* coeffs) {
  if constexpr (N == 0) {
    coeffs[0] = U_xy[0];
  } else if constexpr (N == 1) {
    // reconstruction.c:129-142
    double u_0 = U_xy[0];
    coeffs[0] = u_0;
    coeffs[1] = minmod(U_x[3] - u_0, u_0 - U_x[1]);
    coeffs[2] = minmod(U_y[3] - u_0, u_0 - U_y[1]);
  } else if constexpr (N == 3) {
    weno4(U_x, U_y, U_xy, coeffs);
  } else {
    if (__builtin_expect(weno3_fast(U_x, U_y, U_xy, coeffs), 1)) return;
    // reconstruction.c:144-221
    const double central_cell_wt = 100.0;
    const double r13_3 = 13. / 3.;
    double u_0 = U_xy[0];
    double u_ip1 = U_x[3], u_jp1 = U_y[3], u_ip1jp1 = U_xy[1];
    double u_im1 = U_x[1], u_jm1 = U_y[1], u_ip1jm1 = U_xy[2];
    double u_ip2 = U_x[4], u_jp2 = U_y[4], u_im1jp1 = U_xy[3];
    double u_im2 = U_x[0], u_jm2 = U_y[0], u_im1jm1 = U_xy[4];
    double ux0, ux1, ux2, uxx0, uxx1, uxx2, w0, w1, w2, w3, total, temp;

    ux0 = -2.0 * u_im1 + 0.5 * u_im2 + 1.5 * u_0;
    uxx0 = 0.5 * (u_im2 - 2.0 * u_im1 + u_0);
    ux1 = 0.5 * (u_ip1 - u_im1);
    uxx1 = 0.5 * (u_im1 - 2.0 * u_0 + u_ip1);
    ux2 = -1.5 * u_0 + 2.0 * u_ip1 - 0.5 * u_ip2;
    uxx2 = 0.5 * (u_0 - 2.0 * u_ip1 + u_ip2);
    w0 = 1.0 / (pow4(ux0 * ux0 + r13_3 * uxx0 * uxx0 + SMALL_NUM));
    w1 = central_cell_wt / (pow4(ux1 * ux1 + r13_3 * uxx1 * uxx1 + SMALL_NUM));
    w2 = 1.0 / (pow4(ux2 * ux2 + r13_3 * uxx2 * uxx2 + SMALL_NUM));
    total = w0 + w1 + w2;
    w0 = w0 / total;
    w1 = w1 / total;
    w2 = w2 / total;
    coeffs[0] = u_0;
    coeffs[1] = w0 * ux0 + w1 * ux1 + w2 * ux2;
    coeffs[3] = w0 * uxx0 + w1 * uxx1 + w2 * uxx2;

    ux0 = -2.0 * u_jm1 + 0.5 * u_jm2 + 1.5 * u_0;
    uxx0 = 0.5 * (u_jm2 - 2.0 * u_jm1 + u_0);
    ux1 = 0.5 * (u_jp1 - u_jm1);
    uxx1 = 0.5 * (u_jm1 - 2.0 * u_0 + u_jp1);
    ux2 = -1.5 * u_0 + 2.0 * u_jp1 - 0.5 * u_jp2;
    uxx2 = 0.5 * (u_0 - 2.0 * u_jp1 + u_jp2);
    w0 = 1.0 / (pow4(ux0 * ux0 + r13_3 * uxx0 * uxx0 + SMALL_NUM));
    w1 = central_cell_wt / (pow4(ux1 * ux1 + r13_3 * uxx1 * uxx1 + SMALL_NUM));
    w2 = 1.0 / (pow4(ux2 * ux2 + r13_3 * uxx2 * uxx2 + SMALL_NUM));
    total = w0 + w1 + w2;
    w0 = w0 / total;
    w1 = w1 / total;
    w2 = w2 / total;
    coeffs[2] = w0 * ux0 + w1 * ux1 + w2 * ux2;
    coeffs[4] = w0 * uxx0 + w1 * uxx1 + w2 * uxx2;

    double c1 = coeffs[1], c2 = coeffs[2], c3 = coeffs[3], c4 = coeffs[4];
    double uxy0 = u_ip1jp1 - u_0 - c1 - c2 - c3 - c4;
    double uxy1 = -u_ip1jm1 + u_0 + c1 - c2 + c3 + c4;
    double uxy2 = -u_im1jp1 + u_0 - c1 + c2 + c3 + c4;
    double uxy3 = u_im1jm1 - u_0 + c1 + c2 - c3 - c4;
    temp = 4.0 * (c3 * c3 + c4 * c4);
    w0 = 1.0 / (pow4(temp + uxy0 * uxy0 + SMALL_NUM));
    w1 = 1.0 / (pow4(temp + uxy1 * uxy1 + SMALL_NUM));
    w2 = 1.0 / (pow4(temp + uxy2 * uxy2 + SMALL_NUM));
    w3 = 1.0 / (pow4(temp + uxy3 * uxy3 + SMALL_NUM));
    total = w0 + w1 + w2 + w3;
    w0 = w0 / total;
    w1 = w1 / total;
    w2 = w2 / total;
    w3 = w3 / total;
    coeffs[5] = w0 * uxy0 + w1 * uxy1 + w2 * uxy2 + w3 * uxy3;
  }
}

}  // namespace HYPE_KNS
