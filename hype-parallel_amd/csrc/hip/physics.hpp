// physics.hpp -- device point functions of the Euler, Kapila 5-equation and
// Baer-Nunziato systems.
//
// Every expression keeps the operand order of the reference function cited
// beside it, and the library is compiled without FMA contraction, so each
// value is the IEEE-FP64 result of the same operation sequence as the
// reference.  Face directions are template parameters (DIR 0: x-face, normal
// (1,0); DIR 1: y-face, normal (0,1)); products with the zero normal
// component are kept so that even the sign of a zero result matches.
#pragma once
#ifndef HYPE_KNS
#define HYPE_KNS hype  // kernel namespace; the contracted build uses its own
#endif
#include <hip/hip_runtime.h>

namespace HYPE_KNS {

// hype.h:50-52
constexpr double PRS_FLOOR = 1.0e-12;
constexpr double RHO_FLOOR = 1.0e-14;
constexpr double SMALL_NUM = 1.0e-12;

struct Eos {
  double g1, g2, p1, p2;  // GAMMA_1, GAMMA_2, PI_1, PI_2 (hype.h:42-45)
  bool bn_p2_quirk;       // baer_nunziato.c:271
};

// first failing admissibility check of a flux evaluation; kinds follow the
// reference's MPI_Abort sites (see flux functions)
struct Bad {
  int kind;
  double val;
  __device__ void set(int k, double v) {
    if (kind == 0) {
      kind = k;
      val = v;
    }
  }
};

__device__ __forceinline__ double pmax(double a, double b) { return (a < b) ? b : a; }

template <int DIR>
struct Nrm {
  static constexpr double nx = DIR == 0 ? 1.0 : 0.0;
  static constexpr double ny = DIR == 0 ? 0.0 : 1.0;
};

// ------------------------------------------------------------------ Euler
struct Euler {
  static constexpr int NV = 4;
  static constexpr int WAVES_PER_EU = 2;  // register budget of the stage kernel: 256 per lane
  // tile layouts the stage kernel is built for (bit 0: one wave per tile,
  // bit 1: split tiles on two-wave workgroups); the host picks per problem.
  // Split tiles run one wave per SIMD, one-wave tiles two: small grids gain
  // from the shorter sweeps (256^2: 540 -> 601 Mcell-updates/s), large ones
  // from the second wave's latency hiding
  static constexpr int PAIR_MODES = 3;
  // a split tile sends its CFL maximum to memory once, from the lower wave,
  // instead of once per wave: the atomics of the last stage contend for one
  // address (Euler quadrants 256^2: 636 -> 670 Mcell-updates/s)
  static constexpr bool PAIR_CFL_COMBINE = true;
  static constexpr bool HAS_INLET = false;  // no InletBC in euler.c
  // no sound speed is shared between the admissibility check and the flux
  static constexpr bool HAS_ACACHE = false;
  static constexpr int REFX[2] = {1, -1};  // reflective x-wall negates comp 1
  static constexpr int REFY[2] = {2, -1};
  static constexpr bool HAS_NCP = false;
  // components of the face jump term D that can be nonzero
  static constexpr bool DNZ[NV] = {false, false, false, false};

  // euler.c:14-23
  __device__ static void cons2prim(const Eos& e, const double* Q, double* V) {
    double irho = 1.0 / Q[0];
    V[0] = Q[0];
    V[1] = irho * Q[1];
    V[2] = irho * Q[2];
    V[3] = (e.g1 - 1.0) * (Q[3] - 0.5 * irho * (Q[1] * Q[1] + Q[2] * Q[2]));
  }
  // euler.c:29-38
  __device__ static void prim2cons(const Eos& e, const double* V, double* Q) {
    double en = (V[3]) / (e.g1 - 1.0);
    double k = 0.5 * V[0] * (V[1] * V[1] + V[2] * V[2]);
    Q[0] = V[0];
    Q[1] = V[0] * V[1];
    Q[2] = V[0] * V[2];
    Q[3] = k + en;
  }
  // euler.c:44-84
  template <int DIR>
  __device__ static double flux(const Eos& e, const double* Q, double* F, Bad& b) {
    constexpr double nx = Nrm<DIR>::nx, ny = Nrm<DIR>::ny;
    double rho = Q[0];
    double irho = 1.0 / rho;
    double u = irho * Q[1];
    double v = irho * Q[2];
    double un = nx * u + ny * v;
    double rhoun = rho * un;
    double E = Q[3];
    double p = (e.g1 - 1.0) * (E - 0.5 * rho * (u * u + v * v));
    if (rho < RHO_FLOOR) b.set(1, rho);
    if (p < PRS_FLOOR) b.set(2, p);
    F[0] = rhoun;
    F[1] = rhoun * u + nx * p;
    F[2] = rhoun * v + ny * p;
    F[3] = un * (E + p);
    return fabs(un) + sqrt(e.g1 * p * irho);
  }
  // euler.c:197-236 (E shares nothing with prim2cons: different tree)
  template <int DIR>
  __device__ static double flux_prim(const Eos& e, const double* V, double* F, Bad& b) {
    constexpr double nx = Nrm<DIR>::nx, ny = Nrm<DIR>::ny;
    double rho = V[0];
    double u = V[1];
    double v = V[2];
    double un = nx * u + ny * v;
    double rhoun = rho * un;
    double p = V[3];
    double E = p / (e.g1 - 1.0) + 0.5 * rho * (u * u + v * v);
    if (rho < RHO_FLOOR) b.set(1, rho);
    if (p < PRS_FLOOR) b.set(2, p);
    F[0] = rhoun;
    F[1] = rhoun * u + nx * p;
    F[2] = rhoun * v + ny * p;
    F[3] = un * (E + p);
    return fabs(un) + sqrt(e.g1 * p / rho);
  }
  // euler.c:169-191
  __device__ static bool pad(const Eos& e, const double* Q) {
    bool ok = true;
    double rho = Q[0];
    double irho = 1.0 / rho;
    double u = irho * Q[1];
    double v = irho * Q[2];
    double E = Q[3];
    double p = (e.g1 - 1.0) * (E - 0.5 * rho * (u * u + v * v));
    if (rho < RHO_FLOOR) ok = false;
    if (p < PRS_FLOOR) ok = false;
    return ok;
  }
  // euler.c:292-310
  __device__ static bool pad_prim(const Eos&, const double* V) {
    bool ok = true;
    if (V[0] < RHO_FLOOR) ok = false;
    if (V[3] < PRS_FLOOR) ok = false;
    return ok;
  }
  template <bool PRIM>
  __device__ static void vol_ncp(const Eos&, const double*, const double*, const double*, double* B) {
    for (int c = 0; c < NV; ++c) B[c] = 0.0;
  }
  // face jump term: identically zero (euler.c:142-148, 278-286)
  template <bool PRIM, int DIR>
  __device__ static void ncp_face(const Eos&, const double*, const double*, double* D) {
    for (int c = 0; c < NV; ++c) D[c] = 0.0;
  }
};

// ----------------------------------------------------------------- Kapila
struct Kapila {
  static constexpr int NV = 6;
  static constexpr int WAVES_PER_EU = 1;
  // split tiles only: one reconstructed ring row per wave saved at the same
  // one wave per SIMD (1024 x 512 shock-bubble: 952 -> 985 Mcell-updates/s)
  static constexpr int PAIR_MODES = 2;
  // one atomic per split tile (Kapila 1024x512: 1087 -> 1108
  // Mcell-updates/s; the batched set, whose 64 instances already spread the
  // atomics over 64 addresses, 1320 -> 1317)
  static constexpr bool PAIR_CFL_COMBINE = true;
  // left inlet of kapila_5eqn.c:289-320, 485-513 (tabulated per row by the host)
  static constexpr bool HAS_INLET = true;
  // the admissibility check at a node computes the a_sq the primitive flux
  // at that point needs (kapila_5eqn.c:347-349 and 468-470 are the same
  // expression): the stage kernel hands it over instead of recomputing
  static constexpr bool HAS_ACACHE = true;
  static constexpr int REFX[2] = {2, -1};
  static constexpr int REFY[2] = {3, -1};
  static constexpr bool HAS_NCP = true;
  static constexpr bool DNZ[NV] = {false, false, false, false, false, true};

  // kapila_5eqn.c:14-35
  __device__ static void cons2prim(const Eos& e, const double* Q, double* V) {
    double alpha_1 = Q[5];
    double alpha_2 = 1.0 - alpha_1;
    double rho = Q[0] + Q[1];
    double u = Q[2] / rho;
    double v = Q[3] / rho;
    double E = Q[4];
    double en = E - 0.5 * rho * (u * u + v * v);
    double G1m1 = e.g1 - 1.0;
    double G2m1 = e.g2 - 1.0;
    double p = (en * G1m1 * G2m1 - alpha_1 * e.g1 * e.p1 * G2m1 - alpha_2 * e.g2 * e.p2 * G1m1) /
               (alpha_1 * G2m1 + alpha_2 * G1m1);
    V[0] = Q[0] / alpha_1;
    V[1] = Q[1] / alpha_2;
    V[2] = u;
    V[3] = v;
    V[4] = p;
    V[5] = alpha_1;
  }
  // kapila_5eqn.c:41-61
  __device__ static void prim2cons(const Eos& e, const double* V, double* Q) {
    double alpha_1 = V[5];
    double alpha_2 = 1.0 - alpha_1;
    double rho = alpha_1 * V[0] + alpha_2 * V[1];
    double u = V[2];
    double v = V[3];
    double p = V[4];
    double G1m1 = e.g1 - 1.0;
    double G2m1 = e.g2 - 1.0;
    double en = (p * (alpha_1 * G2m1 + alpha_2 * G1m1) + alpha_1 * e.g1 * e.p1 * G2m1 +
                 alpha_2 * e.g2 * e.p2 * G1m1) /
                (G1m1 * G2m1);
    double E = en + 0.5 * rho * (u * u + v * v);
    Q[0] = alpha_1 * V[0];
    Q[1] = alpha_2 * V[1];
    Q[2] = rho * V[2];
    Q[3] = rho * V[3];
    Q[4] = E;
    Q[5] = alpha_1;
  }
  // kapila_5eqn.c:67-113
  template <int DIR>
  __device__ static double flux(const Eos& e, const double* Q, double* F, Bad& b) {
    constexpr double nx = Nrm<DIR>::nx, ny = Nrm<DIR>::ny;
    double alpha_1 = Q[5];
    double alpha_2 = 1.0 - alpha_1;
    double rho_1 = Q[0] / alpha_1;
    double rho_2 = Q[1] / alpha_2;
    double rho = Q[0] + Q[1];
    double u = Q[2] / rho;
    double v = Q[3] / rho;
    double E = Q[4];
    double en = E - 0.5 * rho * (u * u + v * v);
    double G1m1 = e.g1 - 1.0;
    double G2m1 = e.g2 - 1.0;
    double p = (en * G1m1 * G2m1 - alpha_1 * e.g1 * e.p1 * G2m1 - alpha_2 * e.g2 * e.p2 * G1m1) /
               (alpha_1 * G2m1 + alpha_2 * G1m1);
    double un = u * nx + v * ny;
    double a1_sq = e.g1 * (p + e.p1) / rho_1;
    double a2_sq = e.g2 * (p + e.p2) / rho_2;
    double a_sq = (rho_1 * a1_sq * rho_2 * a2_sq) / (rho * (alpha_2 * rho_1 * a1_sq + alpha_1 * rho_2 * a2_sq));
    if (a_sq < PRS_FLOOR) b.set(3, a_sq);
    F[0] = alpha_1 * rho_1 * un;
    F[1] = alpha_2 * rho_2 * un;
    F[2] = rho * u * un + p * nx;
    F[3] = rho * v * un + p * ny;
    F[4] = (E + p) * un;
    F[5] = 0.0;
    return fabs(un) + sqrt(a_sq);
  }
  // kapila_5eqn.c:326-371
  template <int DIR>
  __device__ static double flux_prim(const Eos& e, const double* V, double* F, Bad& b) {
    constexpr double nx = Nrm<DIR>::nx, ny = Nrm<DIR>::ny;
    double alpha_1 = V[5];
    double alpha_2 = 1.0 - alpha_1;
    double rho_1 = V[0];
    double rho_2 = V[1];
    double rho = alpha_1 * rho_1 + alpha_2 * rho_2;
    double u = V[2];
    double v = V[3];
    double p = V[4];
    double G1m1 = e.g1 - 1.0;
    double G2m1 = e.g2 - 1.0;
    double en = (p * (alpha_1 * G2m1 + alpha_2 * G1m1) + alpha_1 * e.g1 * e.p1 * G2m1 +
                 alpha_2 * e.g2 * e.p2 * G1m1) /
                (G1m1 * G2m1);
    double E = en + 0.5 * rho * (u * u + v * v);
    double un = u * nx + v * ny;
    double a1_sq = e.g1 * (p + e.p1) / rho_1;
    double a2_sq = e.g2 * (p + e.p2) / rho_2;
    double a_sq = (rho_1 * a1_sq * rho_2 * a2_sq) / (rho * (alpha_2 * rho_1 * a1_sq + alpha_1 * rho_2 * a2_sq));
    if (a_sq < PRS_FLOOR) b.set(3, a_sq);
    F[0] = alpha_1 * rho_1 * un;
    F[1] = alpha_2 * rho_2 * un;
    F[2] = rho * u * un + p * nx;
    F[3] = rho * v * un + p * ny;
    F[4] = (E + p) * un;
    F[5] = 0.0;
    return fabs(un) + sqrt(a_sq);
  }
  // flux_prim with a_sq supplied (a_in >= 0: the value pad_prim_a computed
  // at the same point, which also guarantees a_sq >= PRS_FLOOR)
  template <int DIR>
  __device__ static double flux_prim_a(const Eos& e, const double* V, double a_in, double* F, Bad& b) {
    if (a_in < 0.0) return flux_prim<DIR>(e, V, F, b);
    constexpr double nx = Nrm<DIR>::nx, ny = Nrm<DIR>::ny;
    double alpha_1 = V[5];
    double alpha_2 = 1.0 - alpha_1;
    double rho_1 = V[0];
    double rho_2 = V[1];
    double rho = alpha_1 * rho_1 + alpha_2 * rho_2;
    double u = V[2];
    double v = V[3];
    double p = V[4];
    double G1m1 = e.g1 - 1.0;
    double G2m1 = e.g2 - 1.0;
    double en = (p * (alpha_1 * G2m1 + alpha_2 * G1m1) + alpha_1 * e.g1 * e.p1 * G2m1 +
                 alpha_2 * e.g2 * e.p2 * G1m1) /
                (G1m1 * G2m1);
    double E = en + 0.5 * rho * (u * u + v * v);
    double un = u * nx + v * ny;
    F[0] = alpha_1 * rho_1 * un;
    F[1] = alpha_2 * rho_2 * un;
    F[2] = rho * u * un + p * nx;
    F[3] = rho * v * un + p * ny;
    F[4] = (E + p) * un;
    F[5] = 0.0;
    return fabs(un) + sqrt(a_in);
  }
  // kapila_5eqn.c:247-283 (alpha_1 test of :278 kept)
  __device__ static bool pad(const Eos& e, const double* Q) {
    double alpha_1 = Q[5];
    double alpha_2 = 1.0 - alpha_1;
    double rho_1 = Q[0] / alpha_1;
    double rho_2 = Q[1] / alpha_2;
    double rho = Q[0] + Q[1];
    double u = Q[2] / rho;
    double v = Q[3] / rho;
    double E = Q[4];
    double en = E - 0.5 * rho * (u * u + v * v);
    double G1m1 = e.g1 - 1.0;
    double G2m1 = e.g2 - 1.0;
    double p = (en * G1m1 * G2m1 - alpha_1 * e.g1 * e.p1 * G2m1 - alpha_2 * e.g2 * e.p2 * G1m1) /
               (alpha_1 * G2m1 + alpha_2 * G1m1);
    double a1_sq = e.g1 * (p + e.p1) / rho_1;
    double a2_sq = e.g2 * (p + e.p2) / rho_2;
    double a_sq = (rho_1 * a1_sq * rho_2 * a2_sq) / (rho * (alpha_2 * rho_1 * a1_sq + alpha_1 * rho_2 * a2_sq));
    if (rho < RHO_FLOOR) return false;
    if (a_sq < PRS_FLOOR) return false;
    if (alpha_1 < 1.0 || alpha_1 > 1.0) return false;
    return true;
  }
  // pad_prim that also returns its a_sq (for flux_prim_a)
  __device__ static bool pad_prim_a(const Eos& e, const double* V, double& a_out) {
    double alpha_1 = V[5];
    double alpha_2 = 1.0 - alpha_1;
    double rho_1 = V[0];
    double rho_2 = V[1];
    double rho = alpha_1 * rho_1 + alpha_2 * rho_2;
    double p = V[4];
    double a1_sq = e.g1 * (p + e.p1) / rho_1;
    double a2_sq = e.g2 * (p + e.p2) / rho_2;
    double a_sq = (rho_1 * a1_sq * rho_2 * a2_sq) / (rho * (alpha_2 * rho_1 * a1_sq + alpha_1 * rho_2 * a2_sq));
    a_out = a_sq;
    if (rho < RHO_FLOOR) return false;
    if (a_sq < PRS_FLOOR) return false;
    if (alpha_1 < 0.0 || alpha_1 > 1.0) return false;
    return true;
  }
  // kapila_5eqn.c:450-479
  __device__ static bool pad_prim(const Eos& e, const double* V) {
    double alpha_1 = V[5];
    double alpha_2 = 1.0 - alpha_1;
    double rho_1 = V[0];
    double rho_2 = V[1];
    double rho = alpha_1 * rho_1 + alpha_2 * rho_2;
    double p = V[4];
    double a1_sq = e.g1 * (p + e.p1) / rho_1;
    double a2_sq = e.g2 * (p + e.p2) / rho_2;
    double a_sq = (rho_1 * a1_sq * rho_2 * a2_sq) / (rho * (alpha_2 * rho_1 * a1_sq + alpha_1 * rho_2 * a2_sq));
    if (rho < RHO_FLOOR) return false;
    if (a_sq < PRS_FLOOR) return false;
    if (alpha_1 < 0.0 || alpha_1 > 1.0) return false;
    return true;
  }
  // B*grad for the primitive path, comp 5 only (kapila_5eqn.c:417-444);
  // arguments: V (6 values), alpha_x, alpha_y, u_x, v_y
  __device__ static double ncp5_prim(const Eos& e, const double* V, double phi_x, double phi_y,
                                     double u_x, double v_y) {
    double alpha_1 = V[5];
    double alpha_2 = 1.0 - alpha_1;
    double rho_1 = V[0];
    double rho_2 = V[1];
    double u = V[2];
    double v = V[3];
    double p = V[4];
    double a_1_sq = e.g1 * (p + e.p1) / rho_1;
    double a_2_sq = e.g2 * (p + e.p2) / rho_2;
    double K = (alpha_1 * alpha_2 * (rho_2 * a_2_sq - rho_1 * a_1_sq)) /
               (alpha_1 * rho_2 * a_2_sq + alpha_2 * rho_1 * a_1_sq);
    return u * phi_x + v * phi_y - K * (u_x + v_y);
  }
  // B*grad for the conservative path, comp 5 only (kapila_5eqn.c:164-198);
  // gx, gy: gradients of all components
  __device__ static double ncp5_cons(const Eos& e, const double* Q, const double* gx, const double* gy) {
    double alpha_1 = Q[5];
    double alpha_2 = 1.0 - alpha_1;
    double rho_1 = Q[0] / alpha_1;
    double rho_2 = Q[1] / alpha_2;
    double rho = Q[0] + Q[1];
    double irho = 1.0 / rho;
    double u = Q[2] / rho;
    double v = Q[3] / rho;
    double E = Q[4];
    double en = E - 0.5 * rho * (u * u + v * v);
    double G1m1 = e.g1 - 1.0;
    double G2m1 = e.g2 - 1.0;
    double p = (en * G1m1 * G2m1 - alpha_1 * e.g1 * e.p1 * G2m1 - alpha_2 * e.g2 * e.p2 * G1m1) /
               (alpha_1 * G2m1 + alpha_2 * G1m1);
    double a1_sq = e.g1 * (p + e.p1) / rho_1;
    double a2_sq = e.g2 * (p + e.p2) / rho_2;
    double K = alpha_1 * alpha_2 * (rho_2 * a2_sq - rho_1 * a1_sq) / (alpha_1 * rho_2 * a2_sq + alpha_2 * rho_1 * a1_sq);
    double alpha1_x = gx[5];
    double alpha1_y = gy[5];
    double u_x = irho * irho * (gx[2] * rho - Q[2] * (gx[0] + gx[1]));
    double v_y = irho * irho * (gy[3] * rho - Q[3] * (gy[0] + gy[1]));
    return u * alpha1_x + v * alpha1_y - K * (u_x + v_y);
  }
  // volume term B(V) grad V at a quadrature point (rhs_function_prim.c:474)
  template <bool PRIM>
  __device__ static void vol_ncp(const Eos& e, const double* V, const double* gx, const double* gy, double* B) {
    for (int c = 0; c < 5; ++c) B[c] = 0.0;
    if (PRIM)
      B[5] = ncp5_prim(e, V, gx[5], gy[5], gx[2], gy[3]);
    else
      B[5] = ncp5_cons(e, V, gx, gy);
  }
  // face jump D = B(mean) (R - L) (riemann.c:95-102 / 130-137)
  template <bool PRIM, int DIR>
  __device__ static void ncp_face(const Eos& e, const double* L, const double* R, double* D) {
    constexpr double nx = Nrm<DIR>::nx, ny = Nrm<DIR>::ny;
    double av[NV], gx[NV], gy[NV];
    for (int c = 0; c < NV; ++c) {
      av[c] = 0.5 * (R[c] + L[c]);
      gx[c] = nx * (R[c] - L[c]);
      gy[c] = ny * (R[c] - L[c]);
    }
    for (int c = 0; c < 5; ++c) D[c] = 0.0;
    if (PRIM)
      D[5] = ncp5_prim(e, av, gx[5], gy[5], gx[2], gy[3]);
    else
      D[5] = ncp5_cons(e, av, gx, gy);
  }
};

// ---------------------------------------------------------- Baer-Nunziato
struct BN {
  static constexpr int NV = 9;
  static constexpr int WAVES_PER_EU = 1;
  // one wave per tile only: two windows of nine components would halve the
  // workgroups an LDS of 160 KB holds
  static constexpr int PAIR_MODES = 1;
  static constexpr bool PAIR_CFL_COMBINE = false;  // one wave per tile
  static constexpr bool HAS_INLET = false;  // no InletBC in baer_nunziato.c
  // no sound speed is shared between the admissibility check and the flux
  static constexpr bool HAS_ACACHE = false;
  static constexpr int REFX[2] = {1, 5};
  static constexpr int REFY[2] = {2, 6};
  static constexpr bool HAS_NCP = true;
  static constexpr bool DNZ[NV] = {false, true, true, true, false, true, true, true, true};

  // baer_nunziato.c:14-41
  __device__ static void cons2prim(const Eos& e, const double* Q, double* V) {
    double phi_1 = Q[8];
    double phi_2 = 1.0 - Q[8];
    double rho_1 = Q[0] / phi_1;
    double rho_2 = Q[4] / phi_2;
    double u_1 = Q[1] / Q[0];
    double v_1 = Q[2] / Q[0];
    double u_2 = Q[5] / Q[4];
    double v_2 = Q[6] / Q[4];
    double E_1 = Q[3] / phi_1;
    double E_2 = Q[7] / phi_2;
    double p_1 = (e.g1 - 1.0) * (E_1 - 0.5 * rho_1 * (u_1 * u_1 + v_1 * v_1)) - e.g1 * e.p1;
    double p_2 = (e.g2 - 1.0) * (E_2 - 0.5 * rho_2 * (u_2 * u_2 + v_2 * v_2)) - e.g2 * e.p2;
    V[0] = rho_1;
    V[1] = u_1;
    V[2] = v_1;
    V[3] = p_1;
    V[4] = rho_2;
    V[5] = u_2;
    V[6] = v_2;
    V[7] = p_2;
    V[8] = phi_1;
  }
  // baer_nunziato.c:47-60
  __device__ static void prim2cons(const Eos& e, const double* V, double* Q) {
    Q[0] = V[8] * V[0];
    Q[1] = Q[0] * V[1];
    Q[2] = Q[0] * V[2];
    Q[3] = Q[0] * (0.5 * (V[1] * V[1] + V[2] * V[2]) + (V[3] + e.g1 * e.p1) / (V[0] * (e.g1 - 1.0)));
    Q[4] = (1.0 - V[8]) * V[4];
    Q[5] = Q[4] * V[5];
    Q[6] = Q[4] * V[6];
    Q[7] = Q[4] * (0.5 * (V[5] * V[5] + V[6] * V[6]) + (V[7] + e.g2 * e.p2) / (V[4] * (e.g2 - 1.0)));
    Q[8] = V[8];
  }
  // baer_nunziato.c:66-131
  template <int DIR>
  __device__ static double flux(const Eos& e, const double* Q, double* F, Bad& b) {
    constexpr double nx = Nrm<DIR>::nx, ny = Nrm<DIR>::ny;
    double phi_1 = Q[8];
    double phi_2 = 1.0 - Q[8];
    double rho_1 = Q[0] / phi_1;
    double rho_2 = Q[4] / phi_2;
    double u_1 = Q[1] / Q[0];
    double v_1 = Q[2] / Q[0];
    double u_2 = Q[5] / Q[4];
    double v_2 = Q[6] / Q[4];
    double E_1 = Q[3] / phi_1;
    double E_2 = Q[7] / phi_2;
    double p_1 = (e.g1 - 1.0) * (E_1 - 0.5 * rho_1 * (u_1 * u_1 + v_1 * v_1)) - e.g1 * e.p1;
    double p_2 = (e.g2 - 1.0) * (E_2 - 0.5 * rho_2 * (u_2 * u_2 + v_2 * v_2)) - e.g2 * e.p2;
    double un_1 = nx * u_1 + ny * v_1;
    double un_2 = nx * u_2 + ny * v_2;
    if (rho_1 < RHO_FLOOR) b.set(4, rho_1);
    if (rho_2 < RHO_FLOOR) b.set(5, rho_2);
    if ((p_1 + e.p1) < PRS_FLOOR) b.set(6, p_1 + e.p1);
    if ((p_2 + e.p2) < PRS_FLOOR) b.set(7, p_2 + e.p2);
    F[0] = phi_1 * rho_1 * un_1;
    F[1] = phi_1 * (rho_1 * u_1 * un_1 + nx * p_1);
    F[2] = phi_1 * (rho_1 * v_1 * un_1 + ny * p_1);
    F[3] = phi_1 * un_1 * (E_1 + p_1);
    F[4] = phi_2 * rho_2 * un_2;
    F[5] = phi_2 * (rho_2 * u_2 * un_2 + nx * p_2);
    F[6] = phi_2 * (rho_2 * v_2 * un_2 + ny * p_2);
    F[7] = phi_2 * un_2 * (E_2 + p_2);
    F[8] = 0.0;
    double s_1 = fabs(un_1) + sqrt(e.g1 * (p_1 + e.p1) / rho_1);
    double s_2 = fabs(un_2) + sqrt(e.g2 * (p_2 + e.p2) / rho_2);
    return pmax(s_1, s_2);
  }
  // baer_nunziato.c:256-321
  template <int DIR>
  __device__ static double flux_prim(const Eos& e, const double* V, double* F, Bad& b) {
    constexpr double nx = Nrm<DIR>::nx, ny = Nrm<DIR>::ny;
    double phi_1 = V[8];
    double phi_2 = 1.0 - V[8];
    double rho_1 = V[0];
    double rho_2 = V[4];
    double u_1 = V[1];
    double v_1 = V[2];
    double u_2 = V[5];
    double v_2 = V[6];
    double p_1 = V[3];
    double p_2 = e.bn_p2_quirk ? V[8] : V[7];
    double E_1 = (p_1 + e.g1 * e.p1) / (e.g1 - 1.0) + 0.5 * rho_1 * (u_1 * u_1 + v_1 * v_1);
    double E_2 = (p_2 + e.g2 * e.p2) / (e.g2 - 1.0) + 0.5 * rho_2 * (u_2 * u_2 + v_2 * v_2);
    double un_1 = nx * u_1 + ny * v_1;
    double un_2 = nx * u_2 + ny * v_2;
    if (rho_1 < RHO_FLOOR) b.set(4, rho_1);
    if (rho_2 < RHO_FLOOR) b.set(5, rho_2);
    if ((p_1 + e.p1) < PRS_FLOOR) b.set(6, p_1 + e.p1);
    if ((p_2 + e.p2) < PRS_FLOOR) b.set(7, p_2 + e.p2);
    F[0] = phi_1 * rho_1 * un_1;
    F[1] = phi_1 * (rho_1 * u_1 * un_1 + nx * p_1);
    F[2] = phi_1 * (rho_1 * v_1 * un_1 + ny * p_1);
    F[3] = phi_1 * un_1 * (E_1 + p_1);
    F[4] = phi_2 * rho_2 * un_2;
    F[5] = phi_2 * (rho_2 * u_2 * un_2 + nx * p_2);
    F[6] = phi_2 * (rho_2 * v_2 * un_2 + ny * p_2);
    F[7] = phi_2 * un_2 * (E_2 + p_2);
    F[8] = 0.0;
    double s_1 = fabs(un_1) + sqrt(e.g1 * (p_1 + e.p1) / rho_1);
    double s_2 = fabs(un_2) + sqrt(e.g2 * (p_2 + e.p2) / rho_2);
    return pmax(s_1, s_2);
  }
  // baer_nunziato.c:207-250
  __device__ static bool pad(const Eos& e, const double* Q) {
    bool ok = true;
    double phi_1 = Q[8];
    double phi_2 = 1.0 - Q[8];
    double rho_1 = Q[0] / phi_1;
    double rho_2 = Q[4] / phi_2;
    double u_1 = Q[1] / Q[0];
    double v_1 = Q[2] / Q[0];
    double u_2 = Q[5] / Q[4];
    double v_2 = Q[6] / Q[4];
    double E_1 = Q[3] / phi_1;
    double E_2 = Q[7] / phi_2;
    double p_1 = (e.g1 - 1.0) * (E_1 - 0.5 * rho_1 * (u_1 * u_1 + v_1 * v_1)) - e.g1 * e.p1;
    double p_2 = (e.g2 - 1.0) * (E_2 - 0.5 * rho_2 * (u_2 * u_2 + v_2 * v_2)) - e.g2 * e.p2;
    if (rho_1 < RHO_FLOOR) ok = false;
    if (rho_2 < RHO_FLOOR) ok = false;
    if ((p_1 + e.p1) < PRS_FLOOR) ok = false;
    if ((p_2 + e.p2) < PRS_FLOOR) ok = false;
    if (phi_1 < 0.0 || phi_1 > 1.0) ok = false;
    return ok;
  }
  // baer_nunziato.c:350-385
  __device__ static bool pad_prim(const Eos& e, const double* V) {
    bool ok = true;
    if (V[0] < RHO_FLOOR) ok = false;
    if (V[4] < RHO_FLOOR) ok = false;
    if ((V[3] + e.p1) < PRS_FLOOR) ok = false;
    if ((V[7] + e.p2) < PRS_FLOOR) ok = false;
    if (V[8] < 0.0 || V[8] > 1.0) ok = false;
    return ok;
  }
  // B*grad, primitive path (baer_nunziato.c:327-344); u_I = u_1, p_I = p_2
  __device__ static void ncp_prim(const double* V, double phi_x, double phi_y, double* B) {
    double u_I = V[1];
    double v_I = V[2];
    double p_I = V[7];
    B[0] = 0.0;
    B[1] = -p_I * phi_x;
    B[2] = -p_I * phi_y;
    B[3] = -p_I * (u_I * phi_x + v_I * phi_y);
    B[4] = 0.0;
    B[5] = p_I * phi_x;
    B[6] = p_I * phi_y;
    B[7] = p_I * (u_I * phi_x + v_I * phi_y);
    B[8] = u_I * phi_x + v_I * phi_y;
  }
  // B*grad, conservative path (baer_nunziato.c:137-165)
  __device__ static void ncp_cons(const Eos& e, const double* Q, double phi_x, double phi_y, double* B) {
    double phi_2 = 1.0 - Q[8];
    double rho_2 = Q[4] / phi_2;
    double u_1 = Q[1] / Q[0];
    double v_1 = Q[2] / Q[0];
    double u_2 = Q[5] / Q[4];
    double v_2 = Q[6] / Q[4];
    double E_2 = Q[7] / phi_2;
    double p_I = (e.g2 - 1.0) * (E_2 - 0.5 * rho_2 * (u_2 * u_2 + v_2 * v_2)) - e.g2 * e.p2;
    B[0] = 0.0;
    B[1] = -p_I * phi_x;
    B[2] = -p_I * phi_y;
    B[3] = -p_I * (u_1 * phi_x + v_1 * phi_y);
    B[4] = 0.0;
    B[5] = p_I * phi_x;
    B[6] = p_I * phi_y;
    B[7] = p_I * (u_1 * phi_x + v_1 * phi_y);
    B[8] = u_1 * phi_x + v_1 * phi_y;
  }
  template <bool PRIM>
  __device__ static void vol_ncp(const Eos& e, const double* V, const double* gx, const double* gy, double* B) {
    if (PRIM)
      ncp_prim(V, gx[8], gy[8], B);
    else
      ncp_cons(e, V, gx[8], gy[8], B);
  }
  template <bool PRIM, int DIR>
  __device__ static void ncp_face(const Eos& e, const double* L, const double* R, double* D) {
    constexpr double nx = Nrm<DIR>::nx, ny = Nrm<DIR>::ny;
    double av[NV];
    for (int c = 0; c < NV; ++c) av[c] = 0.5 * (R[c] + L[c]);
    double phi_x = nx * (R[8] - L[8]);
    double phi_y = ny * (R[8] - L[8]);
    if (PRIM)
      ncp_prim(av, phi_x, phi_y, D);
    else
      ncp_cons(e, av, phi_x, phi_y, D);
  }
};

// Rusanov flux + path-conservative jump (riemann.c:52-105, 112-140).
// Returns max(s_L, s_R) (PetscMax).
// aL, aR: a_sq of the two states when the system shares it with the
// admissibility check (S::HAS_ACACHE), negative otherwise.
template <class S, bool PRIM, int DIR>
__device__ __forceinline__ double riemann(const Eos& e, const double* L, const double* R, double* F,
                                          double* D, Bad& b, double aL = -1.0, double aR = -1.0) {
  constexpr int NV = S::NV;
  double FL[NV], FR[NV];
  double s;
  if (PRIM) {
    double QL[NV], QR[NV];
    S::prim2cons(e, L, QL);
    S::prim2cons(e, R, QR);
    double sl, sr;
    if constexpr (S::HAS_ACACHE) {
      sl = S::template flux_prim_a<DIR>(e, L, aL, FL, b);
      sr = S::template flux_prim_a<DIR>(e, R, aR, FR, b);
    } else {
      sl = S::template flux_prim<DIR>(e, L, FL, b);
      sr = S::template flux_prim<DIR>(e, R, FR, b);
    }
    s = pmax(sl, sr);
#pragma unroll
    for (int c = 0; c < NV; ++c) F[c] = 0.5 * (FR[c] + FL[c] - s * (QR[c] - QL[c]));
  } else {
    double sl = S::template flux<DIR>(e, L, FL, b);
    double sr = S::template flux<DIR>(e, R, FR, b);
    s = pmax(sl, sr);
#pragma unroll
    for (int c = 0; c < NV; ++c) F[c] = 0.5 * (FR[c] + FL[c] - s * (R[c] - L[c]));
  }
  S::template ncp_face<PRIM, DIR>(e, L, R, D);
  return s;
}

}  // namespace HYPE_KNS
