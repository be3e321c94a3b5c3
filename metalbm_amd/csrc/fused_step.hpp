// fused_step.hpp -- BGK and entropic (ELBM) fused step kernels.
//
// Template parameters:
//   L   Lattice<...> descriptor              (Lattice.h)
//   S   storage type of populations          double | float | _Float16
//   C   compute type                         double | float
//   FS  collision source                     SRC_ZERO | SRC_GUO | SRC_EDM
//   EQ  equilibrium                          TruncationMa3 | Exact
//   SH  populations stored as f_i - w_i
// When C is double and SH is false every operator runs in the reference's
// operation order (parity build); otherwise the shifted/FMA form is used.
#pragma once

#include "mlp.hpp"
#include "step_kernels.hpp"

namespace lbm {

template <class L, class C>
__device__ __forceinline__ void store_fields(const StepArgs& a, int cell, int FS,
                                             C rho, const C* u, const C* F,
                                             double alpha) {
  a.rho_out[cell] = double(rho);
  if (FS == (int)ForcingSchemeType::None) {
    // ForcingScheme<None>::calculateHydrodynamicVelocity (ForcingScheme.h:50-57)
    for (int d = 0; d < L::dimD; ++d) a.uh_out[(size_t)d * a.N + cell] = double(u[d]);
  } else {
    // Generic: u + 0.5 / rho * F (ForcingScheme.h:26-33)
    const C h = C(0.5) / rho;
    for (int d = 0; d < L::dimD; ++d)
      a.uh_out[(size_t)d * a.N + cell] = double(u[d] + F[d] * h);
  }
  if (a.alpha_out) a.alpha_out[cell] = alpha;
  if (a.F_out)
    for (int d = 0; d < L::dimD; ++d) a.F_out[(size_t)d * a.N + cell] = double(F[d]);
}

// FSX: the forcing scheme as configured (None vs Shan-Chen matters only for
// the stored hydrodynamic velocity), passed at run time.
// NT: non-temporal population loads (chosen per configuration by the host).
template <class L, class S, class C, int FS, int EQ, bool SH, int BLOCK, bool NT>
__global__ void __launch_bounds__(BLOCK) k_bgk(const StepArgs a, int FSX) {
  constexpr int Q = L::dimQ;
  constexpr bool REF = std::is_same_v<C, double> && !SH;
  const CellPos c = cell_of_thread<L, BLOCK>(a);
  if (!c.active) return;
  const int cell = self_index<L>(a, c);
  S* __restrict__ dst = static_cast<S*>(a.dst);

  C f[Q];
  pull<L, S, C, NT>(a, c, f);
  C F[3];
  force_of_cell<L, C>(a, c, F);
  const C of = C(a.omega_f), oe = C(a.omega_eq), gc = C(a.guo_coef);
  C rho, u[3] = {C(0), C(0), C(0)}, u2;

  if constexpr (REF) {
    // calculateMoments (Collision.h:72-79)
    moments_ref<L>(f, rho, u, u2);
    const RefEquilibrium<L, EQ, C> E(rho, u, u2);
    C v[3] = {C(0), C(0), C(0)}, v2 = C(0);
    if constexpr (FS == SRC_EDM) {
      // ExactDifferenceMethod source: feq(rho, u + F / rho) - feq (ForcingScheme.h:184-197)
      const C k = C(1.0) / rho;
      for (int d = 0; d < L::dimD; ++d) v[d] = u[d] + F[d] * k;
      v2 = v[0] * v[0];
      for (int d = 1; d < L::dimD; ++d) v2 += v[d] * v[d];
    }
    const RefEquilibrium<L, EQ, C> E2(rho, v, v2);
    // BGK collideAndStream (Collision.h:134-151)
    static_for<Q>([&](auto ic) {
      constexpr int i = decltype(ic)::value;
      const C fe = E.template feq<i>();
      C s = C(0);
      if constexpr (FS == SRC_GUO) s = guo_ref<L, i>(gc, u, F);
      if constexpr (FS == SRC_EDM) s = E2.template feq<i>() - fe;
      const C out = of * f[i] + oe * fe + s;
      store_pop<L, i, S, C>(dst + (size_t)i * a.N + cell, out);
    });
  } else {
    C drho;
    if constexpr (SH) {
      moments_shifted<L>(f, drho, rho, u, u2);
    } else {
      moments_ref<L>(f, rho, u, u2);
      drho = rho - C(1);
    }
    const FastEquilibrium<L, EQ, C> E(u, u2);
    C v[3] = {C(0), C(0), C(0)}, v2 = C(0);
    if constexpr (FS == SRC_EDM) {
      const C k = C(1) / rho;
      for (int d = 0; d < L::dimD; ++d) v[d] = u[d] + F[d] * k;
      v2 = v[0] * v[0];
      for (int d = 1; d < L::dimD; ++d) v2 += v[d] * v[d];
    }
    const FastEquilibrium<L, EQ, C> E2(v, v2);
    C uF = u[0] * F[0];
    for (int d = 1; d < L::dimD; ++d) uF += u[d] * F[d];
    static_for<Q>([&](auto ic) {
      constexpr int i = decltype(ic)::value;
      const C w = C(L::W[i]);
      const C p = E.template p<i>();
      // feq_i in the storage representation: w (drho + rho p) when shifted.
      const C eq = SH ? w * (drho + rho * p) : w * rho * (C(1) + p);
      C s = C(0);
      if constexpr (FS == SRC_GUO) {
        const C cF = cdot<L, i>(F);
        const C cu = cdot<L, i>(u);
        s = (C(3) * gc * w) * (cF - uF + C(3) * cu * cF);
      }
      if constexpr (FS == SRC_EDM) s = rho * w * (E2.template p<i>() - p);
      const C out = of * f[i] + oe * eq + s;
      store_pop<L, i, S, C>(dst + (size_t)i * a.N + cell, out);
    });
  }
  if (a.store) store_fields<L, C>(a, cell, FSX, rho, u, F, 2.0);
}

// ---- entropic collision ------------------------------------------------------

// FP64 view of one cell's pulled populations f and non-equilibrium parts n,
// held in the compute type C (f stored as f - w when SH).  Values are
// converted at each use, so the alpha solve keeps no FP64 copies of them
// (exact for C = double: the view then returns the stored values).
template <class L, class C, bool SH>
struct Fp64View {
  const C* fs;
  const C* ns;
  template <int i>
  __device__ __forceinline__ double f() const {
    return SH ? double(fs[i]) + L::W[i] : double(fs[i]);
  }
  template <int i>
  __device__ __forceinline__ double n() const {
    return double(ns[i]);
  }
  // The same values, converted afresh at each use: the empty asm statement
  // on the stored value keeps the compiler from hoisting the conversions out
  // of an iteration loop, which would hold 2Q doubles live across it.
  template <int i>
  __device__ __forceinline__ double f_fresh() const {
    C v = fs[i];
    asm volatile("" : "+v"(v));
    return SH ? double(v) + L::W[i] : double(v);
  }
  template <int i>
  __device__ __forceinline__ double n_fresh() const {
    C v = ns[i];
    asm volatile("" : "+v"(v));
    return double(v);
  }
};

// ELBM alpha (Collision.h:351-375), first part: the deviation test
// (Collision.h:284-303) and alphaMax (Collision.h:305-326).  Returns true when
// the cell needs the root solve; otherwise alpha and branch are final
// (0: deviation small -> 2; 1: alphaMax < 2 -> 0.95 alphaMax).  p views
// the pulled populations and their non-equilibrium parts in FP64.  REFDIV
// keeps the reference's divisions in the deviation test.
template <class L, bool REFDIV, class V>
__device__ __forceinline__ bool elbm_prelude(const V& p, double& alpha,
                                             double& amax, int& branch) {
  constexpr int Q = L::dimQ;
  bool small = true;
  static_for<Q>([&](auto ic) {
    constexpr int i = decltype(ic)::value;
    const double fi = p.template f<i>(), ni = p.template n<i>();
    if constexpr (REFDIV) {
      const double dev = fabs(ni) / fi;
      if (dev > 1.0e-3) small = false;
    } else {
      if (!(fabs(ni) <= 1.0e-3 * fi) && fi > 0.0) small = false;
    }
  });
  if (small) {
    branch = 0;
    alpha = 2.0;
    return false;
  }
  amax = 2.5;
  if constexpr (REFDIV) {
    static_for<Q>([&](auto ic) {
      constexpr int i = decltype(ic)::value;
      const double fi = p.template f<i>(), ni = p.template n<i>();
      if (ni > 0) {
        const double t = fabs(fi) / ni;
        if (t < amax) amax = t;
      }
    });
  } else {
    // The smallest ratio by cross-multiplied comparisons (num / den < 2.5
    // tracked as a pair) and one division at the end instead of Q: the
    // products round, so two ratios within an ulp may be ordered either
    // way, which moves amax by at most an ulp.
    double num = 2.5, den = 1.0;
    static_for<Q>([&](auto ic) {
      constexpr int i = decltype(ic)::value;
      const double fi = fabs(p.template f<i>()), ni = p.template n<i>();
      if (ni > 0 && fi * den < num * ni) {
        num = fi;
        den = ni;
      }
    });
    if (den != 1.0 || num != 2.5) amax = num / den;
  }
  if (amax < 2.) {
    branch = 1;
    alpha = 0.95 * amax;
    return false;
  }
  return true;
}

// The reference Newton-Raphson (EntropicStep.h:111-140 via solveAlpha,
// Collision.h:328-349): warm start, |step| <= 1e-8, at most 50 iterations,
// the iterate before the last step is returned, alpha = 2 when it fails or
// leaves (1, alphaMax).  G and G' share one log per term.  REFORDER keeps the
// reference's term order g += f A - fm lg (EntropicStep.h:31-62) with
// A_i = ln(f_i / w_i) held per term, for FP64 parity; otherwise the constant
// part is summed once, G = sum f A - sum fm lg, which frees the Q FP64 values
// of A (the solve's register footprint bounds its occupancy).
template <class L, bool REFORDER, class V>
__device__ __forceinline__ double newton_reference(const V& p, double guess,
                                                   double amax, int& branch,
                                                   int& iterations) {
  constexpr int Q = L::dimQ;
  double A[REFORDER ? Q : 1];
  double SA = 0.0;
  static_for<Q>([&](auto ic) {
    constexpr int i = decltype(ic)::value;
    const double a = log(p.template f<i>() / L::W[i]);
    if constexpr (REFORDER)
      A[i] = a;
    else
      SA += p.template f<i>() * a;
  });
  double x = guess, step = 0.0;
  bool ok = false;
  iterations = 0;
  for (int it = 1; it <= 50; ++it) {
    x = x - step;
    iterations = it;
    double g = REFORDER ? 0.0 : SA, dg = 0.0;
    static_for<Q>([&](auto ic) {
      constexpr int i = decltype(ic)::value;
      // outside the parity build, converted afresh per iteration (no 2Q
      // doubles held across the loop)
      double fi, ni;
      if constexpr (REFORDER) {
        fi = p.template f<i>();
        ni = p.template n<i>();
      } else {
        fi = p.template f_fresh<i>();
        ni = p.template n_fresh<i>();
      }
      const double fm = fi - x * ni;
      const double lg = log(fm / L::W[i]);
      if constexpr (REFORDER)
        g += fi * A[i] - fm * lg;
      else
        g -= fm * lg;
      dg += ni * (1 + lg);
    });
    step = g / dg;
    if (fabs(step) <= 1e-8) {
      ok = (x > 1. && x < amax);
      break;
    }
    // A NaN step (an iterate outside the logarithms' domain) makes every
    // later iterate NaN, so the reference's loop runs out its 50 iterations
    // and falls back: end now with the same outcome and iteration count.
    // (One such lane would otherwise hold its wavefront for all 50.)
    if (step != step) {
      iterations = 50;
      break;
    }
  }
  if (!ok) {
    branch = 3;
    return 2.0;
  }
  branch = 2;
  return x;
}

// Residual solvers (modes 1, 2, 4) evaluate G in the reference's log form
// for D3Q27 (register budget) and in the log1p-series form for D2Q9 (cheaper
// where every |alpha x_i| is small).
template <class L>
constexpr bool kLogResidual = L::dimQ > 9;

__device__ __forceinline__ unsigned sample_hash(unsigned cell, unsigned salt) {
  unsigned h = cell * 0x9E3779B1u ^ salt;
  h ^= h >> 16;
  h *= 0x85EBCA6Bu;
  h ^= h >> 13;
  h *= 0xC2B2AE35u;
  return h ^ (h >> 16);
}

// Entropic step, one instantiation per alpha solver MODE (0: reference
// Newton; 1: network start, 2: stored-alpha start, each with a certified
// correction and plain Newton as fallback; 3: that fallback over the
// deferred cells; 4: plain Newton from the stored alpha to the residual
// tolerance for every cell) so that no kernel carries the registers of
// another's solver.  Plain Newton here is the reference's iteration stopped
// at |G| <= residual_tol: it falls back to alpha = 2 exactly where the
// reference Newton does and otherwise reaches the same root.
// Three phases so that the learned mode's wavefront-wide matrix-core
// inference runs with every lane present: (1) pull, moments, fneq and the
// alpha prelude per cell; (2) MODE 1 only: the network predicts alpha for
// every lane's cell at once on the matrix cores; (3) the alpha solve, the
// collision and the stores per cell.  The workgroup stages the network in
// LDS once, before phase 1, and one barrier before phase 2 publishes it (a
// copy per wavefront, as first built, needed four times the LDS and held a
// D3Q27 workgroup pair at the 160 KiB limit).
// Minimum wavefronts per SIMD the register allocator must leave room for
// (the FP64 solve is latency-bound, so occupancy beats a few spills; measured
// on one MI355X at 2048^2 / 256^3): D2Q9 learned, stored-alpha and plain
// Newton four (128 VGPRs with 2 and 4 spilled, 109 without: +10%, +12%,
// +5% over three); D3Q27 learned three (168 VGPRs, none spilled; its launch
// 3.81 -> 3.47 ms against two); the other D3Q27 solvers two (reference
// Newton 2.72 k vs 2.03 k MLUPS at one wavefront with 289 registers; at
// three it spills 42 VGPRs).
// Workgroups of the MODE 3 launch: eight per CU, enough to keep every SIMD
// busy at the entropic kernels' occupancy (two to four wavefronts per SIMD)
// while the list is long.
constexpr int kDeferBlocks = 2048;

// Resets the deferred-solve counter before the correction launch (a kernel,
// not a memset, so that the step can be captured in a graph).
template <int UNUSED = 0>
__global__ void k_defer_reset(unsigned* count) {
  if (threadIdx.x == 0) *count = 0u;
}

template <class L, int MODE>
constexpr int elbm_min_waves() {
  if (L::Type == LatticeType::D2Q9 && (MODE == 1 || MODE == 2 || MODE == 4)) return 4;
  if (L::Type == LatticeType::D3Q27 && MODE == 1) return 3;
  if (L::Type == LatticeType::D3Q27) return 2;
  return 1;
}

template <class L, class S, class C, int FS, int EQ, bool SH, int BLOCK, int MODE>
__global__ void __launch_bounds__(BLOCK)
    __attribute__((amdgpu_waves_per_eu(elbm_min_waves<L, MODE>())))
    k_elbm(const StepArgs a, int FSX) {
  constexpr int Q = L::dimQ;
  constexpr bool REF = std::is_same_v<C, double> && !SH;
  using Shape = MlpShape<L>;
  __shared__ unsigned int s_cnt[6];
  extern __shared__ __align__(16) char s_lds[];
  if (a.stats && threadIdx.x < 6) s_cnt[threadIdx.x] = 0u;
  static_assert(MODE != 1 || Shape::NIN > 0, "no alpha network for this lattice");
  // MODE 1: the workgroup copies the packed network into its LDS now; the
  // copy completes while phase 1 runs, and one barrier before phase 2
  // publishes it.
  using Lay = MfmaLayout<Shape::NIN, Shape::H1, Shape::H2>;
  char* wave_tile = s_lds + Lay::NET + (threadIdx.x >> 6) * Lay::TILE;
  if constexpr (MODE == 1)
    copy_packed_network<Shape::NIN, Shape::H1, Shape::H2>(a.mlp_packed, s_lds);
  // MODE 3: plain Newton for the cells the correction of MODE 1 or 2
  // deferred, one thread per listed cell, the workgroups striding over the
  // list (a launch of at most kDeferBlocks workgroups: when nothing was
  // deferred it costs little more than an empty launch).  Every other mode
  // makes one pass.
  // With the log-form plain Newton (Q > 9) the striding loop's state cost
  // 23 spilled registers, so there MODE 3 makes one pass over a full grid
  // (one thread per possible list entry) instead.
  constexpr bool STRIDED = MODE == 3 && !kLogResidual<L>;
  const int n_listed = MODE == 3 ? (int)*a.defer_count : 0;
  // The striding MODE 3 tallies its statistics per thread over the passes
  // (branches 0-3, Newton iterations, accepted corrections) and adds them up
  // once at the end, so that the passes need no barrier.
  unsigned tally[6] = {0u, 0u, 0u, 0u, 0u, 0u};
  for (int list_base = blockIdx.x * BLOCK;; list_base += gridDim.x * BLOCK) {
  if constexpr (MODE == 3) {
    if (list_base >= n_listed) break;  // uniform over the workgroup
  }
  CellPos c;
  if constexpr (MODE == 3) {
    const int t = list_base + threadIdx.x;
    if (t < n_listed) {
      c = cell_from_index<L>(a, a.defer_list[t]);
    } else {
      c = CellPos{};
      c.active = false;
    }
  } else {
    c = cell_of_thread<L, BLOCK>(a);
  }
  int branch = -1, its = 0;
  bool first = false, solve = false, deferred = false;
  int cell = 0;
  C f[Q], fneq[Q], eqv[Q], F[3];
  C rho = C(1), u[3] = {C(0), C(0), C(0)}, u2 = C(0), drho = C(0);
  double alpha = 2.0, amax = 2.5;
  // Every mode keeps the alpha state (Algorithm.h:103-106): it is the
  // reference Newton's warm start, also when that Newton is the fallback of
  // the residual modes.
  // ---- phase 1 ----
  if (c.active) {
    cell = self_index<L>(a, c);
    pull<L, S, C>(a, c, f);
    force_of_cell<L, C>(a, c, F);
    if constexpr (REF) {
      moments_ref<L>(f, rho, u, u2);
      const RefEquilibrium<L, EQ, C> E(rho, u, u2);
      // fneq parked per cell (Collision.h:232-236), here in registers.
      static_for<Q>([&](auto ic) {
        constexpr int i = decltype(ic)::value;
        fneq[i] = f[i] - E.template feq<i>();
      });
    } else {
      if constexpr (SH) {
        moments_shifted<L>(f, drho, rho, u, u2);
      } else {
        moments_ref<L>(f, rho, u, u2);
        drho = rho - C(1);
      }
      const FastEquilibrium<L, EQ, C> E(u, u2);
      static_for<Q>([&](auto ic) {
        constexpr int i = decltype(ic)::value;
        const C w = C(L::W[i]);
        const C p = E.template p<i>();
        eqv[i] = SH ? w * (drho + rho * p) : w * rho * (C(1) + p);
        fneq[i] = f[i] - eqv[i];
      });
    }
    // The alpha solve runs in FP64 whatever the compute type.
    solve = elbm_prelude<L, REF>(Fp64View<L, C, SH>{f, fneq}, alpha, amax, branch);
  }
  // ---- phase 2 ----
  double start = 2.0;
  if constexpr (MODE == 1) {
    float m[Shape::NIN];
    if (solve) {
      nonequilibrium_moments<L>(fneq, C(1) / rho, m);
    } else {
#pragma unroll
      for (int k = 0; k < Shape::NIN; ++k) m[k] = 0.0f;
    }
    __syncthreads();  // the staged network (every thread of the block reaches here)
    start = mlp_alpha_wave<Shape::NIN, Shape::H1, Shape::H2, typename Shape::Op>(s_lds,
                                                                                wave_tile, m);
  }
  // ---- phase 3 ----
  if (c.active) {
    S* __restrict__ dst = static_cast<S*>(a.dst);
    C* __restrict__ astate = static_cast<C*>(a.alpha_state);
    if (solve) {
      const Fp64View<L, C, SH> p{f, fneq};
      if constexpr (MODE == 0) {
        alpha = newton_reference<L, REF>(p, double(astate[cell]), amax, branch, its);
      } else if constexpr (MODE == 4 || MODE == 3) {
        // plain Newton from the stored alpha to the residual tolerance: the
        // whole solve in MODE 4, the fallback of the cells MODE 1 or 2
        // deferred in MODE 3
        if constexpr (kLogResidual<L>) {
          EvalCurv cv;
          branch = residual_newton_log<L, false, false>(p, amax, double(astate[cell]),
                                                        a.residual_tol, 50, alpha, its, cv)
                       ? 2
                       : 3;
          if (branch == 3) alpha = 2.0;
        } else {
          alpha = newton_residual<L>(p, amax, double(astate[cell]), a.residual_tol, branch,
                                     its);
        }
      } else {
        // MODE 1 starts from the network's alpha, MODE 2 from the stored one;
        // if the solve below is not accepted plain Newton decides -- in the
        // MODE 3 launch that follows, from the untouched stored alpha, with
        // the reference's outcome.
        // MODE 1 makes one certified correction.  MODE 2 (D3Q27) may take a
        // second: from the stored alpha its iterates are the reference
        // Newton's own, so whatever it accepts is the root that Newton
        // converges to.  From the network's alpha a second step would also
        // reach roots from which the reference Newton, started at the stored
        // alpha, diverges to its alpha = 2 fallback.
        if constexpr (MODE == 2) start = double(astate[cell]);
        int ev = 0;
        its = ev;
        bool accepted;
        EvalCurv cv;
        if constexpr (kLogResidual<L>)
          accepted = residual_newton_log<L, true, MODE == 2>(p, amax, start, a.residual_tol,
                                                             MODE == 1 ? 1 : 2, alpha, ev, cv);
        else
          accepted = residual_correction<L>(p, amax, start, a.residual_tol, alpha, ev, cv);
        // The network's root stands only where the reference Newton from the
        // stored alpha provably reaches it; else the cell is deferred and
        // plain Newton from the stored alpha decides (with the reference's
        // alpha = 2 fallbacks).
        if constexpr (MODE == 1)
          accepted = accepted &&
                     reference_reaches(cv, alpha, double(astate[cell]), a.residual_tol);
        if (accepted) {
          branch = 2;
          first = true;
        } else {
          deferred = true;  // listed below, once the wavefront has reconverged
        }
        its = ev;
      }
      // The sample log records the Newton solves: every solve in MODE 0 and
      // MODE 4, the solves the correction deferred in MODE 3 (the cells the
      // network or the stored alpha missed -- data to retrain on).
      if constexpr (Shape::NIN > 0 && (MODE == 0 || MODE == 3 || MODE == 4)) {
        if (a.log_cap > 0 && branch == 2 &&
            sample_hash((unsigned)cell, a.log_salt) < a.log_threshold) {
          const unsigned long long k = atomicAdd(a.log_count, 1ull);
          if ((long long)k < a.log_cap) {
            // recomputed here rather than kept live from phase 2
            float m[Shape::NIN];
            nonequilibrium_moments<L>(fneq, C(1) / rho, m);
            for (int q = 0; q < Shape::NIN; ++q) a.log_x[k * Shape::NIN + q] = m[q];
            a.log_alpha[k] = alpha;
          }
        }
      }
    }
    if (!deferred) {
    astate[cell] = C(alpha);  // Algorithm.h:106
    const C gc = C(a.guo_coef);
    // tau = 1 / (alpha beta); collide with 1 / tau (Collision.h:240, 254).
    const double tau = 1.0 / (alpha * a.beta);
    const C rt = C(1.0 / tau);
    C v[3] = {C(0), C(0), C(0)}, v2 = C(0);
    if constexpr (FS == SRC_EDM) {
      const C k = C(1.0) / rho;
      for (int d = 0; d < L::dimD; ++d) v[d] = u[d] + F[d] * k;
      v2 = v[0] * v[0];
      for (int d = 1; d < L::dimD; ++d) v2 += v[d] * v[d];
    }
    C uF = u[0] * F[0];
    for (int d = 1; d < L::dimD; ++d) uF += u[d] * F[d];
    const RefEquilibrium<L, EQ, C> R2(rho, v, v2);
    const FastEquilibrium<L, EQ, C> E2(v, v2);
    // ELBM collideAndStream (Collision.h:243-258): feq := f - fneq.
    static_for<Q>([&](auto ic) {
      constexpr int i = decltype(ic)::value;
      C s = C(0);
      if constexpr (REF) {
        const C fe = f[i] - fneq[i];
        if constexpr (FS == SRC_GUO) s = guo_ref<L, i>(gc, u, F);
        if constexpr (FS == SRC_EDM) s = R2.template feq<i>() - fe;
      } else {
        const C w = C(L::W[i]);
        if constexpr (FS == SRC_GUO) {
          const C cF = cdot<L, i>(F);
          const C cu = cdot<L, i>(u);
          s = (C(3) * gc * w) * (cF - uF + C(3) * cu * cF);
        }
        if constexpr (FS == SRC_EDM) {
          const C e2 = SH ? w * (drho + rho * E2.template p<i>())
                          : w * rho * (C(1) + E2.template p<i>());
          s = e2 - eqv[i];
        }
      }
      const C out = f[i] - rt * fneq[i] + s;
      store_pop<L, i, S, C>(dst + (size_t)i * a.N + cell, out);
    });
    if (a.store) store_fields<L, C>(a, cell, FSX, rho, u, F, alpha);
    }
  }
  if constexpr (MODE == 1 || MODE == 2) {
    // One counter atomic per wavefront: the lanes' list slots follow from the
    // ballot (about a quarter of the cells defer in a turbulent D3Q27 run).
    const unsigned long long mask = __ballot(deferred);
    if (mask) {
      const int lane = threadIdx.x & 63;
      const int leader = __ffsll((long long)mask) - 1;
      unsigned base = 0u;
      if (lane == leader) base = atomicAdd(a.defer_count, (unsigned)__popcll(mask));
      base = __shfl(base, leader);
      if (deferred)
        a.defer_list[base + (unsigned)__popcll(mask & ((1ull << lane) - 1ull))] = cell;
    }
  }
  if (a.stats) {
    // a deferred cell counts its correction's evaluations here and its
    // branch and Newton iterations in the MODE 3 launch
    if constexpr (STRIDED) {
      // static indices only: a runtime index would put the tally in scratch
      static_for<4>([&](auto bc) {
        constexpr int b = decltype(bc)::value;
        tally[b] += branch == b ? 1u : 0u;
      });
      tally[4] += (unsigned)its;
      tally[5] += first ? 1u : 0u;
    } else {
      __syncthreads();
      if (branch >= 0) atomicAdd(&s_cnt[branch], 1u);
      if (its) atomicAdd(&s_cnt[4], (unsigned)its);
      if (first) atomicAdd(&s_cnt[5], 1u);
    }
  }
  if constexpr (!STRIDED) break;
  }  // list passes
  if (a.stats) {
    if constexpr (STRIDED) {
      __syncthreads();
      static_for<6>([&](auto kc) {
        constexpr int k = decltype(kc)::value;
        if (tally[k]) atomicAdd(&s_cnt[k], tally[k]);
      });
    }
    __syncthreads();
    if (threadIdx.x < 6 && s_cnt[threadIdx.x])
      atomicAdd(&a.stats[(blockIdx.x % kStatSlots) * kStatStride + threadIdx.x],
                (unsigned long long)s_cnt[threadIdx.x]);
  }
}

}  // namespace lbm
