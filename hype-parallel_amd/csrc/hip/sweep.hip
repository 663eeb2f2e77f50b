// sweep.hip -- the fused finite-volume stage kernel for CDNA4 (gfx950).
//
// One wavefront owns a tile: a strip of up to HYPE_TILE_W (62) owned columns
// plus one ring column on each side, and h owned rows.  Lane l works on
// column x0-1+l and the wave marches up the strip one row per iteration:
//
//   * a rolling window of 5 rows x 68 columns of the stage state lives in
//     LDS (component-major, so neighbouring lanes read neighbouring doubles).
//     Each iteration loads one new row, synthesises ghost cells from the
//     boundary index maps (the device replacement for DMGlobalToLocal and the
//     boundary loops, rhs_function_prim.c:44-315) and converts it to
//     primitives (ComputePrimitiveVariables, post_process.c:165-200);
//   * every lane reconstructs its cell (Reconstruct, admissibility check and
//     minmod fallback, rhs_function_prim.c:321-456) into the values at the
//     face quadrature points;
//   * an x-face pairs a lane with its left neighbour through a DPP lane shift
//     and a y-face pairs a row with the previous row kept in registers, so
//     each Rusanov flux + non-conservative jump (riemann.c:112-140) is
//     computed once per face of the tile with no face scratch in memory;
//   * assembly (rhs_function_prim.c:589-602) and the SSP-RK stage update are
//     fused; the stage output goes straight to HBM, and the wave's largest
//     wave speed feeds the device-wide CFL maximum (:617-623).
//
// The build with split tiles (template parameter PR) runs two waves per
// workgroup.  A tile of two or more rows is split at its middle row ym; both
// waves march up.  The lower wave starts on the ring row below the tile as
// above and stops at row ym-1.  The upper wave starts at ym with no ring row
// below it.  It sets aside row ym's bottom-face values and the assembly of
// row ym until both halves are done; then the two waves swap the values
// either side of face ym-1/2 through LDS (one workgroup barrier), each
// computes that face's flux and assembles its boundary row.  One
// reconstructed ring row per wave is saved.  Each system declares the
// layouts it is built for (S::PAIR_MODES); the host picks one per problem.
//
// The build with the left inlet (template parameter IN, Kapila) writes the
// host-tabulated inlet state into the window columns left of the grid.
//
// Arithmetic follows the reference's expression trees and is compiled
// without FMA contraction, so results equal the CPU restatement bit for bit.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../../include/hype2d.h"
#include "../hype_internal.h"
#include "physics.hpp"
#include "recon.hpp"

// Systems this translation unit builds stage kernels for.  Each build (exact
// and contracted) is split in two so that each part gets the scheduler
// options measured best for it: this file builds Euler and Kapila,
// sweep_bn.hip (this source with HYPE_SWEEP_BN) builds Baer-Nunziato, and
// sweep_fc.hip / sweep_fc_bn.hip do the same for the contracted build.
#if defined(HYPE_SWEEP_BN)
#define HYPE_BUILD_EK 0
#define HYPE_BUILD_BN 1
#else
#define HYPE_BUILD_EK 1
#define HYPE_BUILD_BN 0
#endif

namespace HYPE_KNS {

// --------------------------------------------------------------- schemes --
// Stage table of the SSP scheme PETSc uses for order N (hype.c:362-382).
// Buffers: 0 = U (solution), 1 = A, 2 = B.
// op 0 (AXPY):      out = in + (dt/bden) * L           (VecAXPY / VecAYPX)
// op 1 (AXPBYPCZ):  out = (a*X + (dt/bden)*L) + g*Z     (VecAXPBYPCZ)
struct StageDesc {
  int in, out, xb, zb, op;
  double a, g, bden;
};

__host__ __device__ __forceinline__ StageDesc stage_desc(int order, int s) {
  if (order == 0)  // TSEULER: update = U + dt*F, copied back to U afterwards
    return StageDesc{0, 1, 0, 0, 0, 0.0, 0.0, 1.0};
  if (order == 1) {  // TSSSPRKS2, s = 2
    if (s == 0) return StageDesc{0, 1, 0, 0, 0, 0.0, 0.0, 2 - 1.};
    return StageDesc{1, 0, 1, 0, 1, (2 - 1.) / 2, 1. / 2, 2.0};
  }
  if (order == 3) {  // TSSSPRK104: w0 in A/B (ping-pong), w1 built in U at stage 4
    // stages 0-3: U -> A -> B -> A -> B; stage 4 reads B and writes U and A;
    // stages 5-8: A -> B -> A -> B -> A; stage 9 reads A and writes U
    if (s < 4) return StageDesc{s == 0 ? 0 : ((s & 1) ? 1 : 2), (s & 1) ? 2 : 1, 0, 0, 0, 0.0, 0.0, 6.0};
    if (s == 4) return StageDesc{2, 0, 0, 0, 2, 1. / 25, 9. / 25, 6.0};
    if (s < 9) return StageDesc{(s & 1) ? 1 : 2, (s & 1) ? 2 : 1, 0, 0, 0, 0.0, 0.0, 6.0};
    return StageDesc{1, 0, 1, 0, 1, 3. / 5, 1.0, 10.0};
  }
  // TSSSPRKS3, s = 4 (n = 2, r = 2)
  if (s == 0) return StageDesc{0, 1, 0, 0, 0, 0.0, 0.0, 2.0};
  if (s == 1) return StageDesc{1, 2, 0, 0, 0, 0.0, 0.0, 2.0};
  if (s == 2) return StageDesc{2, 1, 0, 2, 1, 1. * 2 / (2 * 2 - 1), (2 - 1.) / (2 * 2 - 1), 2 * (2 * 2 - 1)};
  return StageDesc{1, 0, 0, 0, 0, 0.0, 0.0, 2.0};
}

__host__ __device__ __forceinline__ int num_stages_dev(int order) {
  return order == 0 ? 1 : (order == 1 ? 2 : (order == 2 ? 4 : 10));
}

// packed error word (layout in hype_internal.h): K = face-key bits
__device__ __forceinline__ unsigned long long pack_err(int seq, long key, int kind, double val, int K) {
  const float f = (float)val;
  return ((unsigned long long)(seq & 0xF) << HYPE_ERR_SEQ_SHIFT) |
         (((unsigned long long)key & ((1ull << K) - 1ull)) << (HYPE_ERR_SEQ_SHIFT - K)) |
         ((unsigned long long)(kind & 0xF) << (HYPE_ERR_SEQ_SHIFT - K - 4)) |
         (unsigned long long)(__float_as_uint(f) >> (K - 24));
}

// map a (possibly ghost) index onto the owned range; toggles flip for a
// reflective side (rhs_function_prim.c:84-99,152-167 / 213-228,266-281)
__device__ __forceinline__ int bc_map(int i, int n, int bclo, int bchi, bool& flip) {
  if (i < 0) {
    if (bclo == HYPE_BC_TRANSMISSIVE) return 0;
    if (bclo == HYPE_BC_REFLECTIVE) {
      flip = true;
      return -1 - i;
    }
    return i + n;  // periodic wrap (DMDA_BOUNDARY_PERIODIC)
  }
  if (i >= n) {
    if (bchi == HYPE_BC_TRANSMISSIVE) return n - 1;
    if (bchi == HYPE_BC_REFLECTIVE) {
      flip = true;
      return 2 * (n - 1) - i + 1;
    }
    return i - n;
  }
  return i;
}

__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) {
    double w = __shfl_xor(v, o, HYPE_WAVE);
    v = (v < w) ? w : v;
  }
  return v;
}

__device__ __forceinline__ unsigned long long wave_min_u64(unsigned long long v) {
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) {
    unsigned long long w = __shfl_xor(v, o, HYPE_WAVE);
    v = (w < v) ? w : v;
  }
  return v;
}

// orders a wave's own LDS accesses across lanes: LDS operations of one wave
// complete in issue order, so only the compiler's code motion is fenced
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// global-memory view of a pointer.  The state, output and inlet pointers are
// read from the instance table, so the compiler cannot infer their address
// space and would use flat loads and stores.  Those count on the LDS wait
// counter too, so every LDS read or lane shuffle of a row iteration would
// wait for the next row's prefetch; global instructions count on the vector
// memory counter alone.
template <class T>
using gmem_t = __attribute__((address_space(1))) T;
template <class T>
__device__ __forceinline__ gmem_t<T>* gmem(T* p) {
  return (gmem_t<T>*)p;
}

// whole-wave lane shifts by one lane, as DPP moves (wave_shr:1 = 0x138,
// wave_shl:1 = 0x130): lane i takes lane i-1 (from_left) or lane i+1
// (from_right), and the end lane keeps its own value, as __shfl_up and
// __shfl_down do.  Two VALU moves per double, against two LDS permutes and
// a wait on the LDS counter.
__device__ __forceinline__ double lane_from_left(double v) {
  const int lo = __double2loint(v), hi = __double2hiint(v);
  return __hiloint2double(__builtin_amdgcn_update_dpp(hi, hi, 0x138, 0xf, 0xf, false),
                          __builtin_amdgcn_update_dpp(lo, lo, 0x138, 0xf, 0xf, false));
}
__device__ __forceinline__ double lane_from_right(double v) {
  const int lo = __double2loint(v), hi = __double2hiint(v);
  return __hiloint2double(__builtin_amdgcn_update_dpp(hi, hi, 0x130, 0xf, 0xf, false),
                          __builtin_amdgcn_update_dpp(lo, lo, 0x130, 0xf, 0xf, false));
}

constexpr int NSLOT = 5;
constexpr int COLS = HYPE_WAVE + 4;

template <class S, bool PRIM, int N, bool PR, bool IN>
struct Tile {
  static constexpr int NV = S::NV;
  using Qd = Quad<N>;
  static constexpr int NDOF = Qd::NDOF, NGF = Qd::NGF, NGV = Qd::NGV, NNODE = Qd::NNODE;
  static constexpr bool PAIR = PR;
  static constexpr int WAVES = PAIR ? 2 : 1;
  // the admissibility check's a_sq at each node goes on to the flux there
  static constexpr bool ACACHE = PRIM && S::HAS_ACACHE && N == 2;
  // LDS per workgroup, in doubles: one window per wave, then (paired) the
  // two exchange areas (NGF x (NV + 1) values per lane: face values and the
  // a_sq handed over with them) and the upper wave's deferred row state
  // (G, E, x part and volume term of row ym: 4 x NV per lane)
  static constexpr int WIN = NSLOT * NV * COLS;
  static constexpr int XCH = NGF * (NV + 1) * HYPE_WAVE;
  static constexpr int DEFER = 4 * NV * HYPE_WAVE;
  // then, for long records, per wave, the staging of one row of output
  // records.  Each lane storing its own record writes NV doubles at a stride
  // of NV doubles per instruction; with non-temporal stores, which are not
  // merged in L2, 9-double records wrote 1.8x their bytes to HBM, 6-double
  // records 1.2x.  Staging a row through LDS makes every store a whole line,
  // at the price of two wave barriers per row: even for 9 doubles, and 2-3 %
  // slower for 6
  static constexpr bool STAGE_ROWS = NV > 6;
  static constexpr int OST = STAGE_ROWS ? 2 * NV * HYPE_TILE_W : 0;  // Out and A
  // and, where the pair combines its CFL maximum, one double: the upper
  // wave's maximum
  static constexpr bool COMBINE = PAIR && S::PAIR_CFL_COMBINE;
  // and, on the primitive path of split tiles, per wave, the staging of the
  // four window columns right of the lanes' 64 for 16 rows (they are
  // converted 16 rows at a time).  Not in Euler's one-wave kernel: at two
  // waves per SIMD it has no accumulation registers to spare and the
  // staging spills it to scratch (served stream -2.5 %).
  //
  // The lower wave stages in its exchange area, which it writes only after
  // its row loop; the upper wave needs its own (the upper half's exchange
  // area and deferred state are filled in its first iterations)
  static constexpr bool XBATCH = PRIM && PAIR;
  static constexpr int XST = XBATCH ? 16 * 4 * NV : 0;
  static_assert(XST <= XCH, "the lower wave's staging must fit its exchange area");
  static constexpr int LDS = WAVES * WIN + (PAIR ? 2 * XCH + DEFER : 0) + WAVES * OST + (COMBINE ? 1 : 0) + XST;

  // value of component c at node/face point p from its modal coefficients
  // (rhs_function_prim.c:382-388, 425-440)
  __device__ static __forceinline__ double at_node(const double (&co)[NV][NDOF], int c, int p) {
    double v;
#define PHIN(k) Qd::phin(p, k)
    HYPE_BASIS_SUM(v, co[c], NDOF, PHIN);
#undef PHIN
    return v;
  }
  // value and scaled gradient at volume point q (rhs_function_prim.c:442-455)
  __device__ static __forceinline__ void at_vol(const double (&co)[NV][NDOF], int q, double r1_h, double* V,
                                                double* gx, double* gy) {
#pragma unroll
    for (int c = 0; c < NV; ++c) {
      double a, b, d;
#define PHIV(k) Qd::phiv(q, k)
#define GPX(k) Qd::gphx(q, k)
#define GPY(k) Qd::gphy(q, k)
      HYPE_BASIS_SUM(a, co[c], NDOF, PHIV);
      HYPE_BASIS_SUM(b, co[c], NDOF, GPX);
      HYPE_BASIS_SUM(d, co[c], NDOF, GPY);
#undef PHIV
#undef GPX
#undef GPY
      V[c] = a;
      gx[c] = r1_h * b;
      gy[c] = r1_h * d;
    }
  }

  __device__ static void run(const hype_launch& L, int tile_id, double* lds, int lane, int wave) {
    const hype_tile T = L.tiles[tile_id];
    const hype_inst* I = &L.insts[T.inst];
    hype_dstate* st = I->st;
    const int nx = I->nx, ny = I->ny;
    const int bcL = I->bc[0], bcR = I->bc[1], bcB = I->bc[2], bcT = I->bc[3];
    const double r1_h = I->r1_h;
    const Eos eos{I->g1, I->g2, I->p1, I->p2, (I->quirks & HYPE_QUIRK_BN_P2_IS_PHI) != 0};
    // the first tile of an instance writes its time state (lower wave only)
    const bool leader = (T.flags & 1) != 0 && wave == 0;
    const int mode = L.mode;

    double dt = 0.0;
    int cur = 0;
    StageDesc sd = stage_desc(N, L.stage);
    const bool last_stage = (mode == HYPE_MODE_STAGE) && (L.stage == num_stages_dev(N) - 1);
    const auto Uin = gmem<const double>((mode == HYPE_MODE_STAGE) ? I->buf[sd.in] : (L.rhs_in ? L.rhs_in : I->buf[0]));
    // the stage update's buffers, read from the descriptor once per tile
    // rather than in every row iteration
    const auto Out = gmem((mode == HYPE_MODE_STAGE) ? I->buf[sd.out] : (L.rhs_out ? L.rhs_out : I->buf[1]));
    const auto Xb = gmem<const double>(I->buf[sd.xb]);
    const auto Zb = gmem<const double>(I->buf[sd.zb]);
    const auto Ab = gmem(I->buf[1]);

    // ---- geometry ------------------------------------------------------
    const int x0 = T.x0, y0 = T.y0;
    const int W = T.w, H = T.h;
    const int col = x0 - 1 + lane;
    const bool owned = lane >= 1 && lane <= W;
    const bool xlane = lane >= 1 && lane <= W + 1;
    const int cw = lane + 2;  // window column of this lane's cell
    // split tile (split-tile build, two or more rows): the lower wave owns
    // rows [y0, ym), the upper wave [ym, y0 + H); a one-row tile is the
    // lower wave's alone
    const bool split = PAIR && H >= 2;
    if (PAIR && !split && wave != 0) return;
    const int ym = y0 + (H + 1) / 2;
    const bool upper = split && wave == 1;
    const int own_lo = upper ? ym : y0;
    const int own_hi = (split && !upper) ? ym : y0 + H;
    const int r_first = upper ? ym : y0 - 1;                // lower ring row unless upper half
    const int r_last = (split && !upper) ? ym - 1 : y0 + H;  // upper ring row unless lower half
    const int y_first = upper ? ym + 1 : y0;                // first row whose face r-1/2 is computed
    const int a_first = upper ? ym + 2 : y0 + 1;            // first iteration that assembles row r-1
    double* win = lds + (PAIR ? wave * WIN : 0);
    double* xch = lds + WAVES * WIN;  // [0]: row ym-1 top values, [1]: row ym bottom values
    double* defer = xch + 2 * XCH;
    auto X_ = [&](int side, int q, int c) -> double& { return xch[side * XCH + (q * (NV + 1) + c) * HYPE_WAVE + lane]; };
    auto D_ = [&](int k, int c) -> double& { return defer[(k * NV + c) * HYPE_WAVE + lane]; };
    double* ost = lds + WAVES * WIN + (PAIR ? 2 * XCH + DEFER : 0) + (PAIR ? wave * OST : 0);
    double* pair_smax = lds + WAVES * WIN + (PAIR ? 2 * XCH + DEFER : 0) + WAVES * OST;
    const int rbase = r_first - 2;  // window row of slot 0
    auto slot = [&](int r) { return (r - rbase) % NSLOT; };
    auto W_ = [&](int s, int c, int k) -> double& { return win[(s * NV + c) * COLS + k]; };
    auto sync_window = [&]() {
      if constexpr (PAIR)
        wave_sync();
      else
        __syncthreads();
    };

    // ---- row loader ----------------------------------------------------
    // fetch issues the loads of one cell (source cell from the boundary
    // index maps) and returns its reflection flags; put applies them, so a
    // row's loads need not complete before the next row's are issued
    auto fetch = [&](int rr, int k, double* q) -> int {
      const int cc = min(x0 - 3 + k, nx + 2);
      bool fx = false, fy = false;
      const int sc = bc_map(cc, nx, bcL, bcR, fx);  // an inlet column maps as periodic; put_inlet replaces it
      const int sr = bc_map(rr, ny, bcB, bcT, fy);
      const auto p = Uin + ((size_t)sr * nx + sc) * NV;
#pragma unroll
      for (int c = 0; c < NV; ++c) q[c] = p[c];
      return (fx ? 1 : 0) | (fy ? 2 : 0);
    };
    auto conv = [&](double* q, int refl, double* v) {
      if (refl & 1) {
        q[S::REFX[0]] = -q[S::REFX[0]];
        if (S::REFX[1] >= 0) q[S::REFX[1]] = -q[S::REFX[1]];
      }
      if (refl & 2) {
        q[S::REFY[0]] = -q[S::REFY[0]];
        if (S::REFY[1] >= 0) q[S::REFY[1]] = -q[S::REFY[1]];
      }
      if (PRIM) {
        S::cons2prim(eos, q, v);
      } else {
#pragma unroll
        for (int c = 0; c < NV; ++c) v[c] = q[c];
      }
    };
    auto put = [&](int s, int k, double* q, int refl) {
      double v[NV];
      conv(q, refl, v);
#pragma unroll
      for (int c = 0; c < NV; ++c) W_(s, c, k) = v[c];
    };
    // the four window columns right of the lanes' 64 (grid columns x0+61 ..
    // x0+64) on the primitive path: loaded and converted 16 rows at a time,
    // one cell per lane, into the wave's staging xs ([c][row % 16][column]),
    // and copied into the window a row at a time by lanes 0-3.  The
    // conversion then runs once per 16 rows on every lane instead of every
    // row on 4 lanes.  Window row j (counted from rbase) is staged by the
    // batch of rows 16*(j/16) .. 16*(j/16)+15; rows past r_last+2 are not
    // needed and not loaded
    double* xs = wave == 0 ? xch : lds + WAVES * WIN + (PAIR ? 2 * XCH + DEFER : 0) + WAVES * OST + (COMBINE ? 1 : 0);
    auto xrow = [&](int j0) { return rbase + j0 + (lane >> 2); };
    auto xput = [&](double* q, int refl) {
      double v[NV];
      conv(q, refl, v);
#pragma unroll
      for (int c = 0; c < NV; ++c) xs[c * HYPE_WAVE + lane] = v[c];
    };
    auto xcopy = [&](int s, int j) {
      if (lane < 4) {
#pragma unroll
        for (int c = 0; c < NV; ++c) W_(s, c, HYPE_WAVE + lane) = xs[c * HYPE_WAVE + (j & 15) * 4 + lane];
      }
    };
    // left inlet (Kapila, rhs_function_prim.c:135-144): in the strip at the
    // left edge, the three window columns left of the grid take the state the
    // host tabulated for the ghost row, once the row is stored.  The y-pass
    // that follows in the reference (:204-315) maps transmissive and
    // reflective ghost rows; periodic ones keep their own height.
    const bool inlet_strip = IN && bcL == HYPE_BC_INLET && x0 == 0;
    auto put_inlet = [&](int s, int rr) {
      if constexpr (IN) {
        if (inlet_strip && lane < 3) {
          bool fy = false;
          const int sr = bc_map(rr, ny, bcB, bcT, fy);
          const bool own = (rr < 0 && bcB == HYPE_BC_PERIODIC) || (rr >= ny && bcT == HYPE_BC_PERIODIC);
          const auto p = gmem(I->inlet) + (size_t)((own ? rr : sr) + 3) * NV;
          double v[NV];
#pragma unroll
          for (int c = 0; c < NV; ++c) v[c] = p[c];
          if (fy) {
            v[S::REFY[0]] = -v[S::REFY[0]];
            if (S::REFY[1] >= 0) v[S::REFY[1]] = -v[S::REFY[1]];
          }
#pragma unroll
          for (int c = 0; c < NV; ++c) W_(s, c, lane) = v[c];
        }
      }
    };

    // prologue: rows r_first-2 .. r_first+2, all loads in flight before the
    // first store, and before the time state is read: it decides whether the
    // tile runs at all, and its round trip now overlaps theirs
    double pqa[5][NV], pqb[5][NV];
    int pfa[5], pfb[5];
    const bool pxl = XBATCH && xrow(0) <= r_last + 2;
#pragma unroll
    for (int i = 0; i < 5; ++i) {
      pfa[i] = fetch(rbase + i, lane, pqa[i]);
      pfb[i] = 0;
    }
    if constexpr (XBATCH) {
      if (pxl) pfb[0] = fetch(xrow(0), HYPE_WAVE + (lane & 3), pqb[0]);
    } else {
#pragma unroll
      for (int i = 0; i < 5; ++i)
        if (lane < 4) pfb[i] = fetch(rbase + i, lane + HYPE_WAVE, pqb[i]);
    }

    // ---- time state (single writer per field per launch) -------------
    if (mode == HYPE_MODE_STAGE) {
      if (L.stage == 0) {
        const int step = st->step;
        const double t = st->t;
        // stop after max_steps, at t_end, or after a step (or the priming
        // call) that met an inadmissible state
        if (step >= st->max_steps || !(t < st->t_end) || st->err[(step + 1) & 1] != ~0ull) {
          if (leader && lane == 0) st->done = 1;
          return;
        }
        const double sm = __longlong_as_double((long long)st->smax[step & 1]);
        dt = (I->cfl * I->h) / (2.0 * sm);
        if (st->clip && t + dt > st->t_end) dt = st->t_end - t;
        cur = step;
        if (leader && lane == 0) {
          st->dt = dt;
          st->cur = step;
          st->smax[(step + 1) & 1] = 0ull;
          if (st->dt_hist && step < st->hist_len) st->dt_hist[step] = dt;
        }
      } else {
        if (st->done) return;
        dt = st->dt;
        cur = st->cur;
      }
      if (last_stage && leader && lane == 0) {
        st->t = st->t + dt;
        st->step = cur + 1;
        st->err[(cur + 1) & 1] = ~0ull;  // the next step records there
      }
    }
    const double cdt = dt / sd.bden;

    if constexpr (XBATCH) {
#pragma unroll
      for (int i = 0; i < 5; ++i) {
        put(slot(rbase + i), lane, pqa[i], pfa[i]);
        put_inlet(slot(rbase + i), rbase + i);
      }
      if (pxl) xput(pqb[0], pfb[0]);
      wave_sync();
#pragma unroll
      for (int i = 0; i < 5; ++i) xcopy(slot(rbase + i), i);
    } else {
#pragma unroll
      for (int i = 0; i < 5; ++i) {
        put(slot(rbase + i), lane, pqa[i], pfa[i]);
        if (lane < 4) put(slot(rbase + i), lane + HYPE_WAVE, pqb[i], pfb[i]);
        put_inlet(slot(rbase + i), rbase + i);
      }
    }
    sync_window();

    // ---- carried state -------------------------------------------------
    double topPrev[NGF][NV];
    double topA[NGF];  // a_sq handed over with topPrev (negative: none)
    double Gprev[NV], Eprev[NV], xpartPrev[NV], ncpPrev[NV];
#pragma unroll
    for (int c = 0; c < NV; ++c) {
      Gprev[c] = 0.0;
      Eprev[c] = 0.0;
      xpartPrev[c] = 0.0;
      ncpPrev[c] = 0.0;
#pragma unroll
      for (int q = 0; q < NGF; ++q) topPrev[q][c] = 0.0;
    }
#pragma unroll
    for (int q = 0; q < NGF; ++q) topA[q] = -1.0;
    double smax = 0.0;
    unsigned long long err = ~0ull;
    const int seq = (mode == HYPE_MODE_STAGE) ? L.stage + 1 : 0;
    const double WF = Qd::WF, WV = Qd::WV;

    // flux + jump at the y-face below row fr from the values either side
    // (lower side: top values of row fr-1; upper side: bottom values of fr)
    auto yface = [&](const double (&Lv)[NGF][NV], const double (&Rv)[NGF][NV], const double* aL, const double* aR,
                     int fr, double* G, double* E) {
#pragma unroll
      for (int c = 0; c < NV; ++c) {
        G[c] = 0.0;
        E[c] = 0.0;
      }
#pragma unroll
      for (int q = 0; q < NGF; ++q) {
        double Fq[NV], Dq[NV];
        Bad b{0, 0.0};
        const double sp = riemann<S, PRIM, 1>(eos, Lv[q], Rv[q], Fq, Dq, b, aL[q], aR[q]);
        if (owned) {
          if (sp > smax) smax = sp;
          if (b.kind) {
            const long key = (long)ny * (nx + 1) * NGF + ((long)fr * nx + col) * NGF + q;
            const unsigned long long pk = pack_err(seq, key, b.kind, b.val, I->keybits);
            if (pk < err) err = pk;
          }
        }
#pragma unroll
        for (int c = 0; c < NV; ++c) {
          G[c] = G[c] + WF * Fq[c];
          if (S::DNZ[c]) E[c] = E[c] + WF * Dq[c];
        }
      }
    };

    // the output of the stage.  The stores are non-temporal: the output is
    // read again only by the next launch, long after it has left L2, and
    // streaming it past L2 keeps the window rows there for their re-read by
    // the stage update and the neighbouring tiles.  A short record is stored
    // by its lane (put_out); a long one is put into the LDS staging (slot k:
    // 0 for Out, 1 for A) and the whole wave then stores the row's
    // contiguous span (flush), so every store instruction writes whole
    // cache lines
    auto put_out = [&](double val, gmem_t<double>* dst, size_t i, int k, int c) {
      if constexpr (STAGE_ROWS) ost[(k * HYPE_TILE_W + lane - 1) * NV + c] = val;
      else __builtin_nontemporal_store(val, &dst[i]);
    };
    auto flush = [&](gmem_t<double>* dst, int row, int k) {
      wave_sync();
      const size_t first = ((size_t)row * nx + x0) * NV;
      const double* src = ost + k * HYPE_TILE_W * NV;
      for (int e = lane; e < W * NV; e += HYPE_WAVE) __builtin_nontemporal_store(src[e], &dst[first + e]);
      wave_sync();  // the staging is rewritten by the next row
    };

    // an operand of the stage update read from buffer B.  On the
    // conservative path the window holds the stage input of the owned cells
    // unchanged, so while the row is still in the window (wr) the input is
    // taken from there instead of being read from memory a second time
    auto operand = [&](gmem_t<const double>* B, size_t i, int row, int c, bool wr) -> double {
      if constexpr (!PRIM) {
        if (wr && B == Uin) return W_(slot(row), c, lane + 2);
      }
      return B[i];
    };

    // assembly + stage update of cell (col, row) from the terms of its top
    // and bottom faces, its x-face part and its volume term
    // (rhs_function_prim.c:589-602)
    auto record = [&](int row, const double* Gt, const double* Gb, const double* Et, const double* Eb,
                      const double* xp, const double* nc, bool wr) {
      const size_t idx = ((size_t)row * nx + col) * NV;
      double rhs[NV];
#pragma unroll
      for (int c = 0; c < NV; ++c) {
        const double Y = Gt[c] - Gb[c] + 0.5 * (Et[c] + Eb[c]);
        const double t = xp[c] - r1_h * Y;
        rhs[c] = ((S::HAS_NCP && S::DNZ[c]) ? nc[c] : 0.0) + t;
      }
      if (mode == HYPE_MODE_RHS) {
#pragma unroll
        for (int c = 0; c < NV; ++c) put_out(rhs[c], Out, idx + c, 0, c);
      } else {
        // every operand of the record is loaded before the first store.  The
        // compiler cannot prove the output distinct from the inputs (it is
        // not, for the RK stages that update U in place), so loads
        // interleaved with the stores would each wait for memory: six round
        // trips per row instead of one
        const auto out = Out;
        if (sd.op == 0) {
          double u[NV];
#pragma unroll
          for (int c = 0; c < NV; ++c) u[c] = operand(Uin, idx + c, row, c, wr);
#pragma unroll
          for (int c = 0; c < NV; ++c) put_out(u[c] + cdt * rhs[c], out, idx + c, 0, c);
        } else if (sd.op == 2) {
          // RK104 after its fifth stage: w0 += dt/6 F (VecAXPY); w1 =
          // 1/25 U + 9/25 w0 (VecAXPBYPCZ, c = 0) into U; w0 = 15 w1 - 5 w0
          // (VecAXPBY) into A
          const auto A = Ab;
          double u[NV], o[NV];
#pragma unroll
          for (int c = 0; c < NV; ++c) {
            u[c] = operand(Uin, idx + c, row, c, wr);
            o[c] = out[idx + c];
          }
#pragma unroll
          for (int c = 0; c < NV; ++c) {
            const double w0 = u[c] + cdt * rhs[c];
            const double w1 = sd.a * o[c] + sd.g * w0;
            put_out(w1, out, idx + c, 0, c);
            put_out(15.0 * w1 + -5.0 * w0, A, idx + c, 1, c);
          }
        } else {
          const auto X = Xb;
          const auto Z = Zb;
          double x[NV], z[NV];
#pragma unroll
          for (int c = 0; c < NV; ++c) {
            x[c] = operand(X, idx + c, row, c, wr);
            z[c] = operand(Z, idx + c, row, c, wr);
          }
#pragma unroll
          for (int c = 0; c < NV; ++c) put_out(sd.a * x[c] + cdt * rhs[c] + sd.g * z[c], out, idx + c, 0, c);
        }
      }
    };
    // called by the whole wave (the flush of staged rows needs every lane)
    auto assemble = [&](int row, const double* Gt, const double* Gb, const double* Et, const double* Eb,
                        const double* xp, const double* nc, bool wr) {
      if constexpr (STAGE_ROWS) {
        if (mode == HYPE_MODE_PRIME) return;
        if (owned) record(row, Gt, Gb, Et, Eb, xp, nc, wr);
        flush(Out, row, 0);
        if (mode == HYPE_MODE_STAGE && sd.op == 2) flush(Ab, row, 1);
      } else {
        if (!owned || mode == HYPE_MODE_PRIME) return;
        record(row, Gt, Gb, Et, Eb, xp, nc, wr);
      }
    };

    for (int r = r_first; r <= r_last; ++r) {
      // prefetch the row that enters the window after this iteration
      const bool more = (r + 3 <= r_last + 2);
      const int jn = r - r_first + 5;  // window row of row r+3, from rbase
      const bool xb = XBATCH && more && (jn & 15) == 0;
      const bool xl = xb && xrow(jn) <= r_last + 2;
      double qa[NV], qb[NV];
      int fa = 0, fb = 0;
      if (more) {
        fa = fetch(r + 3, lane, qa);
        if constexpr (XBATCH) {
          if (xl) fb = fetch(xrow(jn), HYPE_WAVE + (lane & 3), qb);
        } else {
          if (lane < 4) fb = fetch(r + 3, lane + HYPE_WAVE, qb);
        }
      }

      const bool ownedRow = (r >= own_lo) && (r < own_hi);
      const bool doY = (r >= y_first);
      const bool doAsm = (r >= a_first);
      const int sm2 = slot(r - 2), sm1 = slot(r - 1), s0 = slot(r), sp1 = slot(r + 1), sp2 = slot(r + 2);

      // ---- reconstruction: modal coefficients of every component --------
      double co[NV][NDOF];
#pragma unroll
      for (int c = 0; c < NV; ++c) {
        double ux[5], uy[5], ud[5];
        ux[0] = W_(s0, c, cw - 2);
        ux[1] = W_(s0, c, cw - 1);
        ux[2] = W_(s0, c, cw);
        ux[3] = W_(s0, c, cw + 1);
        ux[4] = W_(s0, c, cw + 2);
        uy[0] = W_(sm2, c, cw);
        uy[1] = W_(sm1, c, cw);
        uy[2] = ux[2];
        uy[3] = W_(sp1, c, cw);
        uy[4] = W_(sp2, c, cw);
        ud[0] = ux[2];
        ud[1] = W_(sp1, c, cw + 1);
        ud[2] = W_(sm1, c, cw + 1);
        ud[3] = W_(sp1, c, cw - 1);
        ud[4] = W_(sm1, c, cw - 1);
        reconstruct<N>(ux, uy, ud, co[c]);
      }
      // values at the nodes, which are also the face quadrature points
      // (rhs_function_prim.c:382-388, 425-440; the same points, hype.c:207-229)
      double fv[NNODE][NV];
#pragma unroll
      for (int p = 0; p < NNODE; ++p)
#pragma unroll
        for (int c = 0; c < NV; ++c) fv[p][c] = at_node(co, c, p);
      // admissibility at the nodes (rhs_function_prim.c:393-404); with
      // ACACHE the check's a_sq at each node goes on to the flux there
      double an[ACACHE ? NNODE : 1];
      bool pad = true;
#pragma unroll
      for (int p = 0; p < NNODE; ++p) {
        if constexpr (ACACHE)
          pad = S::pad_prim_a(eos, fv[p], an[p]) && pad;
        else
          pad = pad && (PRIM ? S::pad_prim(eos, fv[p]) : S::pad(eos, fv[p]));
      }
      if (!pad) {  // rhs_function_prim.c:407-421 / rhs_function.c:409-421
#pragma unroll
        for (int c = 0; c < NV; ++c) {
          const double w0 = W_(s0, c, cw);
          co[c][0] = w0;
          if constexpr (NDOF > 1) {
            co[c][1] = PRIM ? minmod(W_(s0, c, cw + 1) - w0, w0 - W_(s0, c, cw - 1)) : 0.0;
            co[c][2] = PRIM ? minmod(W_(sp1, c, cw) - w0, w0 - W_(sm1, c, cw)) : 0.0;
          }
#pragma unroll
          for (int k = 3; k < NDOF; ++k) co[c][k] = 0.0;
        }
#pragma unroll
        for (int p = 0; p < NNODE; ++p)
#pragma unroll
          for (int c = 0; c < NV; ++c) fv[p][c] = at_node(co, c, p);
        if constexpr (ACACHE) {
#pragma unroll
          for (int p = 0; p < NNODE; ++p) an[p] = -1.0;  // new values: the flux computes a_sq itself
        }
      }
      if (L.fallback && ownedRow && owned) gmem(L.fallback)[(size_t)r * nx + col] = (unsigned char)(!pad);

      // ---- slide the window: row r-2 is no longer read, row r+3 takes its
      // slot (the wave's own reads above precede this store in program order)
      if (more) {
        const int sn = slot(r + 3);
        put(sn, lane, qa, fa);
        if constexpr (XBATCH) {
          if (xb) {
            if (xl) xput(qb, fb);
            wave_sync();
          }
          xcopy(sn, jn);
        } else {
          if (lane < 4) put(sn, lane + HYPE_WAVE, qb, fb);
        }
        put_inlet(sn, r + 3);
      }

      // ---- volume non-conservative product (rhs_function_prim.c:461-479)
      double ncp[NV];
#pragma unroll
      for (int c = 0; c < NV; ++c) ncp[c] = 0.0;
      if constexpr (S::HAS_NCP) {
        if (mode != HYPE_MODE_PRIME && ownedRow) {
#pragma unroll
          for (int q = 0; q < NGV; ++q) {
            double V[NV], gx[NV], gy[NV], B[NV];
            at_vol(co, q, r1_h, V, gx, gy);
            S::template vol_ncp<PRIM>(eos, V, gx, gy, B);
#pragma unroll
            for (int c = 0; c < NV; ++c)
              if (S::DNZ[c]) ncp[c] = ncp[c] + -WV * B[c];
          }
        }
      }


      // ---- x-faces i-1/2 (rhs_function_prim.c:485-533) -------------------
      double xpart[NV];
      {
        double Fx[NV], Dx[NV];
#pragma unroll
        for (int c = 0; c < NV; ++c) {
          Fx[c] = 0.0;
          Dx[c] = 0.0;
        }
#pragma unroll
        for (int q = 0; q < NGF; ++q) {
          double vl[NV], vr[NV], Fq[NV], Dq[NV];
#pragma unroll
          for (int c = 0; c < NV; ++c) {
            vl[c] = lane_from_left(fv[Qd::R(q)][c]);
            vr[c] = fv[Qd::L(q)][c];
          }
          double al = -1.0, ar = -1.0;
          if constexpr (ACACHE) {
            al = lane_from_left(an[Qd::R(q)]);
            ar = an[Qd::L(q)];
          }
          if (ownedRow) {
            Bad b{0, 0.0};
            const double s = riemann<S, PRIM, 0>(eos, vl, vr, Fq, Dq, b, al, ar);
            if (xlane) {
              if (s > smax) smax = s;
              if (b.kind) {
                const long key = ((long)r * (nx + 1) + col) * NGF + q;
                const unsigned long long pk = pack_err(seq, key, b.kind, b.val, I->keybits);
                if (pk < err) err = pk;
              }
            }
#pragma unroll
            for (int c = 0; c < NV; ++c) {
              Fx[c] = Fx[c] + WF * Fq[c];
              if (S::DNZ[c]) Dx[c] = Dx[c] + WF * Dq[c];
            }
          }
        }
#pragma unroll
        for (int c = 0; c < NV; ++c) {
          const double Fr = lane_from_right(Fx[c]);
          const double Dr = S::DNZ[c] ? lane_from_right(Dx[c]) : 0.0;
          xpart[c] = -r1_h * (Fr - Fx[c] + 0.5 * (Dr + Dx[c]));
        }
      }

      // ---- y-face r-1/2 (rhs_function_prim.c:537-585) --------------------
      double Gy[NV], Ey[NV];
      if (doY) {
        double an_b[NGF];
#pragma unroll
        for (int q = 0; q < NGF; ++q) an_b[q] = ACACHE ? an[Qd::B(q)] : -1.0;
        double fvb[NGF][NV];
#pragma unroll
        for (int q = 0; q < NGF; ++q)
#pragma unroll
          for (int c = 0; c < NV; ++c) fvb[q][c] = fv[Qd::B(q)][c];
        yface(topPrev, fvb, topA, an_b, r, Gy, Ey);
      } else {
#pragma unroll
        for (int c = 0; c < NV; ++c) {
          Gy[c] = 0.0;
          Ey[c] = 0.0;
        }
      }
      if constexpr (PAIR) {
        if (upper && r == ym) {  // row ym's bottom values wait for the lower half
#pragma unroll
          for (int q = 0; q < NGF; ++q) {
#pragma unroll
            for (int c = 0; c < NV; ++c) X_(1, q, c) = fv[Qd::B(q)][c];
            X_(1, q, NV) = ACACHE ? an[Qd::B(q)] : -1.0;
          }
        }
        if (upper && r == ym + 1) {  // so does the assembly of row ym
#pragma unroll
          for (int c = 0; c < NV; ++c) {
            D_(0, c) = Gy[c];
            D_(1, c) = Ey[c];
            D_(2, c) = xpartPrev[c];
            D_(3, c) = ncpPrev[c];
          }
        }
      }

      // ---- assembly + stage update of cell (col, r-1) --------------------
      if (doAsm) assemble(r - 1, Gy, Gprev, Ey, Eprev, xpartPrev, ncpPrev, true);  // rows r-2..r+2 are in the window

      // ---- carry -----------------------------------------------------------
#pragma unroll
      for (int c = 0; c < NV; ++c) {
        Gprev[c] = Gy[c];
        Eprev[c] = Ey[c];
        xpartPrev[c] = xpart[c];
        ncpPrev[c] = ncp[c];
#pragma unroll
        for (int q = 0; q < NGF; ++q) topPrev[q][c] = fv[Qd::T(q)][c];
      }
      if constexpr (ACACHE) {
#pragma unroll
        for (int q = 0; q < NGF; ++q) topA[q] = an[Qd::T(q)];
      }

      sync_window();  // the new row is visible to the next iteration
    }

    const bool want_smax = (mode != HYPE_MODE_STAGE) || last_stage;

    // ---- the two halves of a split tile meet at face ym-1/2 -----------------
    if constexpr (PAIR) {
      if (split) {
        if (!upper) {
#pragma unroll
          for (int q = 0; q < NGF; ++q) {
#pragma unroll
            for (int c = 0; c < NV; ++c) X_(0, q, c) = topPrev[q][c];
            X_(0, q, NV) = topA[q];
          }
        }
        // the CFL maximum goes to memory once per workgroup: the upper wave
        // hands its rows' maximum to the lower wave here.  After the barrier
        // both waves solve the same middle face, and the upper wave adds no
        // other face, so the lower wave's maximum then covers the tile
        if constexpr (COMBINE) {
          const double um = (want_smax && upper) ? wave_max(smax) : 0.0;
          if (want_smax && upper && lane == 0) *pair_smax = um;
        }
        __syncthreads();
        if constexpr (COMBINE) {
          if (want_smax && !upper) {
            const double pm = *pair_smax;
            if (pm > smax) smax = pm;
          }
        }
        double lo[NGF][NV], hi[NGF][NV], alo[NGF], ahi[NGF];
#pragma unroll
        for (int q = 0; q < NGF; ++q) {
#pragma unroll
          for (int c = 0; c < NV; ++c) {
            lo[q][c] = X_(0, q, c);
            hi[q][c] = X_(1, q, c);
          }
          alo[q] = X_(0, q, NV);
          ahi[q] = X_(1, q, NV);
        }
        double Gm[NV], Em[NV];
        yface(lo, hi, alo, ahi, ym, Gm, Em);
        if (!upper) {
          assemble(ym - 1, Gm, Gprev, Em, Eprev, xpartPrev, ncpPrev, false);
        } else {
          double Gt[NV], Et[NV], xp[NV], nc[NV];
#pragma unroll
          for (int c = 0; c < NV; ++c) {
            Gt[c] = D_(0, c);
            Et[c] = D_(1, c);
            xp[c] = D_(2, c);
            nc[c] = D_(3, c);
          }
          assemble(ym, Gt, Gm, Et, Em, xp, nc, false);  // row ym has left the window
        }
      }
    }

    // ---- CFL maximum and error word ----------------------------------------
    if (want_smax && !(COMBINE && split && upper)) {
      const double m = wave_max(smax);
      if (lane == 0) {
        unsigned long long* slotp =
            (mode == HYPE_MODE_RHS) ? &st->smax_rhs : (mode == HYPE_MODE_PRIME ? &st->smax[st->step & 1] : &st->smax[(cur + 1) & 1]);
        atomicMax(slotp, (unsigned long long)__double_as_longlong(m));
      }
    }
    const unsigned long long e = wave_min_u64(err);
    if (lane == 0 && e != ~0ull) {
      unsigned long long* ep =
          (mode == HYPE_MODE_RHS) ? &st->err_rhs : (mode == HYPE_MODE_PRIME ? &st->err[1] : &st->err[cur & 1]);
      atomicMin(ep, e);
    }
  }
};

// Split tiles run one wave per SIMD: the two waves of a workgroup must not
// share a SIMD's issue slots while other SIMDs idle.  A system whose kernel
// would fit two waves per SIMD gets its accumulation registers reserved (an
// empty asm statement that clobbers a255), which leaves room for one wave.
template <class S, bool PR>
struct Occ {
  static constexpr bool ONE_PER_SIMD = PR && S::WAVES_PER_EU > 1;
  static constexpr int WAVES_PER_EU = PR ? 1 : S::WAVES_PER_EU;
};

// IN: the build with the left-inlet ghost columns (Kapila problems with an
// inlet only, so the loop of every other problem is compiled without it)
template <class S, bool PRIM, int N, bool PR, bool IN>
__global__ __launch_bounds__(PR ? 2 * HYPE_WAVE : HYPE_WAVE)
    __attribute__((amdgpu_waves_per_eu(Occ<S, PR>::WAVES_PER_EU))) void stage_kernel(hype_launch L) {
  using TT = Tile<S, PRIM, N, PR, IN>;
  if constexpr (Occ<S, PR>::ONE_PER_SIMD) asm volatile("" ::: "a255");
  __shared__ double lds[TT::LDS];
  __shared__ int next_tile;
  const int lane = (int)threadIdx.x % HYPE_WAVE;
  const int wave = __builtin_amdgcn_readfirstlane((int)threadIdx.x / HYPE_WAVE);
  const int ntiles = L.ntiles_dev ? *L.ntiles_dev : L.ntiles;
  if (L.grid <= 0) {
    if ((int)blockIdx.x < ntiles) TT::run(L, blockIdx.x, lds, lane, wave);
    return;
  }
  for (;;) {
    // one queue pull per workgroup; the second barrier keeps the next pull
    // from overwriting next_tile before every wave has read it
    if (threadIdx.x == 0) next_tile = (int)atomicAdd(&L.queue[0], 1u);
    __syncthreads();
    const int tile = __builtin_amdgcn_readfirstlane(next_tile);
    __syncthreads();
    if (tile >= ntiles) break;
    TT::run(L, tile, lds, lane, wave);
  }
  if (threadIdx.x == 0) {
    __threadfence();
    const unsigned fin = atomicAdd(&L.queue[1], 1u);
    if (fin == gridDim.x - 1) {
      atomicExch(&L.queue[0], 0u);
      atomicExch(&L.queue[1], 0u);
    }
  }
}

template <class S, bool PRIM, int N, bool PR, bool IN>
static hipError_t launch_variant(const hype_launch& L, hipStream_t s) {
  // one workgroup per table entry (grid 0; with a device-resident count the
  // workgroups past it return at once), or a persistent grid pulling tiles
  const int grid = L.grid > 0 ? L.grid : L.ntiles;
  hipLaunchKernelGGL((stage_kernel<S, PRIM, N, PR, IN>), dim3(grid), dim3(HYPE_WAVE * Tile<S, PRIM, N, PR, IN>::WAVES), 0,
                     s, L);
  return hipGetLastError();
}

template <class S, bool PRIM, int N, bool PR>
static hipError_t launch_inlet(const hype_launch& L, hipStream_t s) {
  if (L.v.inlet) {
    if constexpr (S::HAS_INLET) return launch_variant<S, PRIM, N, PR, true>(L, s);
    return hipErrorInvalidValue;
  }
  return launch_variant<S, PRIM, N, PR, false>(L, s);
}

// the tile layout the variant asks for, if the system is built for it
template <class S, bool PRIM, int N>
static hipError_t launch_layout(const hype_launch& L, hipStream_t s) {
  if (L.v.pair) {
    if constexpr ((S::PAIR_MODES & 2) != 0) return launch_inlet<S, PRIM, N, true>(L, s);
  } else {
    if constexpr ((S::PAIR_MODES & 1) != 0) return launch_inlet<S, PRIM, N, false>(L, s);
  }
  return hipErrorInvalidValue;
}

template <class S>
static hipError_t launch_sys(const hype_launch& L, hipStream_t s) {
  const bool p = L.v.prim != 0;
  switch (L.v.order) {
    case 0: return p ? launch_layout<S, true, 0>(L, s) : launch_layout<S, false, 0>(L, s);
    case 1: return p ? launch_layout<S, true, 1>(L, s) : launch_layout<S, false, 1>(L, s);
    case 3: return p ? launch_layout<S, true, 3>(L, s) : launch_layout<S, false, 3>(L, s);
    default: return p ? launch_layout<S, true, 2>(L, s) : launch_layout<S, false, 2>(L, s);
  }
}

template <class S, bool PRIM, int N, bool PR, bool IN>
static int occupancy_variant() {
  int blocks = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, stage_kernel<S, PRIM, N, PR, IN>,
                                                   HYPE_WAVE * Tile<S, PRIM, N, PR, IN>::WAVES, 0) != hipSuccess)
    return 0;
  return blocks;
}

template <class S, bool PRIM, int N, bool PR>
static int occupancy_inlet(const hype_variant& v) {
  if (v.inlet) {
    if constexpr (S::HAS_INLET) return occupancy_variant<S, PRIM, N, PR, true>();
    return 0;
  }
  return occupancy_variant<S, PRIM, N, PR, false>();
}

template <class S, bool PRIM, int N>
static int occupancy_layout(const hype_variant& v) {
  if (v.pair) {
    if constexpr ((S::PAIR_MODES & 2) != 0) return occupancy_inlet<S, PRIM, N, true>(v);
  } else {
    if constexpr ((S::PAIR_MODES & 1) != 0) return occupancy_inlet<S, PRIM, N, false>(v);
  }
  return 0;
}

template <class S>
static int occupancy_sys(const hype_variant& v) {
  const bool p = v.prim != 0;
  switch (v.order) {
    case 0: return p ? occupancy_layout<S, true, 0>(v) : occupancy_layout<S, false, 0>(v);
    case 1: return p ? occupancy_layout<S, true, 1>(v) : occupancy_layout<S, false, 1>(v);
    case 3: return p ? occupancy_layout<S, true, 3>(v) : occupancy_layout<S, false, 3>(v);
    default: return p ? occupancy_layout<S, true, 2>(v) : occupancy_layout<S, false, 2>(v);
  }
}

#if !HYPE_BUILD_EK
// the Baer-Nunziato part of the exact build (sweep_bn.hip)
hipError_t launch_stage_bn(const hype_launch& L, hipStream_t s) { return launch_sys<BN>(L, s); }
int blocks_per_cu_bn(const hype_variant& v) { return occupancy_sys<BN>(v); }
#else
#if !HYPE_BUILD_BN
hipError_t launch_stage_bn(const hype_launch& L, hipStream_t s);
int blocks_per_cu_bn(const hype_variant& v);
#endif

// tile layouts a system's stage kernel is built for (bit 0: one wave per
// tile, bit 1: split tiles)
int pair_modes(int system) {
  switch (system) {
    case HYPE_EULER: return Euler::PAIR_MODES;
    case HYPE_KAPILA: return Kapila::PAIR_MODES;
    default: return BN::PAIR_MODES;
  }
}

int blocks_per_cu(const hype_variant& v) {
  switch (v.system) {
    case HYPE_EULER: return occupancy_sys<Euler>(v);
    case HYPE_KAPILA: return occupancy_sys<Kapila>(v);
#if HYPE_BUILD_BN
    default: return occupancy_sys<BN>(v);
#else
    default: return blocks_per_cu_bn(v);
#endif
  }
}

hipError_t launch_stage(const hype_launch& L, hipStream_t s) {
  if (L.ntiles <= 0 && !L.ntiles_dev) return hipSuccess;
  switch (L.v.system) {
    case HYPE_EULER: return launch_sys<Euler>(L, s);
    case HYPE_KAPILA: return launch_sys<Kapila>(L, s);
#if HYPE_BUILD_BN
    default: return launch_sys<BN>(L, s);
#else
    default: return launch_stage_bn(L, s);
#endif
  }
}

// ---------------------------------------------------- primitive pass --------
template <class S>
__global__ void cons2prim_kernel(const double* U, double* W, long n, Eos e) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double q[S::NV], v[S::NV];
#pragma unroll
  for (int c = 0; c < S::NV; ++c) q[c] = U[i * S::NV + c];
  S::cons2prim(e, q, v);
#pragma unroll
  for (int c = 0; c < S::NV; ++c) W[i * S::NV + c] = v[c];
}

hipError_t launch_cons2prim(int system, const double* U, double* W, long n, Eos e, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  const int bs = 256;
  const unsigned grid = (unsigned)((n + bs - 1) / bs);
  if (system == HYPE_EULER)
    hipLaunchKernelGGL(cons2prim_kernel<Euler>, dim3(grid), dim3(bs), 0, s, U, W, n, e);
  else if (system == HYPE_KAPILA)
    hipLaunchKernelGGL(cons2prim_kernel<Kapila>, dim3(grid), dim3(bs), 0, s, U, W, n, e);
  else
    hipLaunchKernelGGL(cons2prim_kernel<BN>, dim3(grid), dim3(bs), 0, s, U, W, n, e);
  return hipGetLastError();
}
#endif  // HYPE_BUILD_EK

}  // namespace HYPE_KNS

#if !defined(HYPE_CONTRACTED_BUILD) && HYPE_BUILD_EK
// The stage table as the host sees it (tests replay the buffer algebra on
// the CPU against the oracle's TSSSP steps).
extern "C" int hyped_stage_table(int order, int s, int* ibuf, double* coef) {
  if (order < 0 || order > 3 || s < 0 || s >= HYPE_KNS::num_stages_dev(order)) return -1;
  const HYPE_KNS::StageDesc d = HYPE_KNS::stage_desc(order, s);
  ibuf[0] = d.in;
  ibuf[1] = d.out;
  ibuf[2] = d.xb;
  ibuf[3] = d.zb;
  ibuf[4] = d.op;
  coef[0] = d.a;
  coef[1] = d.g;
  coef[2] = d.bden;
  return HYPE_KNS::num_stages_dev(order);
}
#endif
