// Cooperative Rosenbrock integrator for large mechanisms: one wavefront
// integrates one cell (KPP_VARIANT_COOP).
//
// The one-lane-per-cell kernels keep a cell's working set in a 64-lane tile
// in HBM; for a mechanism with thousands of LU entries (MCM: 7,123) every
// operand is a memory round trip and a wavefront keeps about one load in
// flight.  Here the 64 lanes of a wavefront share one cell instead:
//  * the iteration matrix G = I/(h gamma) - J and its LU factors live in LDS
//    (LU_NONZERO doubles), next to a scratch region T for the products of
//    Fun/Jac_SP, the triangular solves' vectors and the error-norm terms;
//  * the model routines run as the mechanism compiler's cooperative programs
//    (kppgen --coop): product tables evaluated one product per lane, and
//    list-scheduled chains -- each a value's left-associated sum or
//    elimination sequence, the operands and their order those of the
//    emitted code -- one descriptor per (step, lane);
//  * the vectors of the integrator (Y, Fcn0, Fcn, dF/dT, Ynew, the stage
//    vectors K) are rows of a per-wavefront block in global memory, lanes
//    taking species i = lane, lane+64, ...; the error norm's terms are summed
//    by one lane in species order, as ros_ErrorNorm does.
// Control (accept/reject, step size, counters) is the same code path as
// Cell::attempt, executed identically by every lane, so the wavefront never
// diverges on it.  With FP contraction off the results are the one-lane
// kernels' (and the CPU oracle's) bit for bit.
//
// Program execution: a step reads its operands from LDS, updates the lane's
// running value and stores it to the chain's target (every step stores: a
// partial value is never read, since a target's original value is read only
// by its own chain's first step and other chains read it after the chain's
// last step; an idle step leaves the value unchanged and writes it back to
// the lane's previous target).  The loads of step t+1 are issued before the
// store of step t (the schedule puts every read of a chain's result at least
// two steps after its last step), and LDS operations of one wavefront
// execute in order, so no barrier is needed inside a program.
//
// Descriptor fields are LDS byte offsets from the base of the region they
// address (G at LDS address 0; T at C::GZ, a compile-time constant that the
// LDS instructions take as their immediate offset), so a step spends one
// integer operation per operand; the dynamic LDS of the kernel starts at 0
// because the kernel declares no static LDS.
#pragma once
#include "ros_device.h"

namespace kpphip {

// Rows of a wavefront's block in global memory (one cell at a time).
template <class M, int S>
struct CoopLayout {
  static constexpr int Y = 0;
  static constexpr int FIX = Y + M::NVAR;
  static constexpr int RCT = FIX + (M::NFIX > 0 ? M::NFIX : 1);
  static constexpr int FCN0 = RCT + M::NREACT;
  static constexpr int FCN = FCN0 + M::NVAR;
  static constexpr int DFDT = FCN + M::NVAR;
  static constexpr int YNEW = DFDT + M::NVAR;
  static constexpr int K = YNEW + M::NVAR;
  static constexpr int ROWS = K + S * M::NVAR;
};

constexpr int cmax(int a, int b) { return a > b ? a : b; }

// LDS of one cell: G[LU_NONZERO] at 0, then the T region at C::GZ: a junk
// slot (C::JSLOT, written by idle steps), +0.0 (C::ZSLOT, read by idle
// steps), then the products, Fun's results, the solve vectors X|Z and the
// error terms.
template <class M, class C>
struct CoopLds {
  static_assert(C::GZ == 8u * M::LU_NONZERO && C::JSLOT == 0u && C::ZSLOT == 8u,
                "LDS layout of the cooperative programs");
  static constexpr size_t BYTES = sizeof(double) * (size_t)(M::LU_NONZERO + C::TSIZE);
};

// Global-memory accessor for the emitted rate routines (stride 1).
struct Acc1 {
  using value_type = double;
  gdouble* b;
  __device__ __forceinline__ gdouble& operator[](int i) const { return b[i]; }
};

constexpr unsigned kFacEnd = 0xFFFFFFFFu;

// Typed address spaces matter here: through generic pointers every access
// to LDS is a flat access, which the hardware serialises with global memory
// (each wait for an LDS value becomes a wait for every outstanding load).
using gcdouble = const __attribute__((address_space(1))) double;
using gcu64 = const __attribute__((address_space(1))) unsigned long long;
using gcu32 = const __attribute__((address_space(1))) unsigned;

// ---- the program interpreters, out of line: each is a small loop with few
// live values, so the register allocator keeps its prefetch ring and
// operands in registers whatever the integrator around it holds.

// Descriptors of the next RING steps are requested ahead (a rotating ring:
// slot k holds step base+k until step base+k runs, then step base+k+RING);
// the tables carry idle steps past their end, so requests need no bound.
#ifndef KPPHIP_COOP_RING
#define KPPHIP_COOP_RING 8
#endif
constexpr int kRing = KPPHIP_COOP_RING;

// Product factor value: RCT[j], V[i], F[i] or a constant, without a branch
// (one load from the base the factor's type selects); a slot past the
// product's last factor is 1.0, and x*1.0 == x exactly.
__device__ __forceinline__ double coop_factor(unsigned code, gcdouble* V, gcdouble* R,
                                              gcdouble* F, gcdouble* K) {
  const unsigned idx = (code == kFacEnd) ? 0u : (code & 0x3FFFFFFFu), type = code >> 30;
  gcdouble* base = (type == 0) ? R : (type == 1) ? V : (type == 2) ? F : K;
  const double v = base[idx];
  return code == kFacEnd ? 1.0 : v;
}
// T[q] = product q (A of Fun or B of Jac_SP) at state V, left to right; four
// products per lane at a time (the tables are padded to whole rounds)
template <int MAXF, int NPAD>
__device__ __noinline__ void coop_products(const unsigned* Pg, int N, gcdouble* V, gcdouble* R,
                                           gcdouble* F, gcdouble* K, ldouble* T, int lane) {
  gcu32* P = (gcu32*)Pg;
  constexpr int PU = 4;
  static_assert(NPAD % (TILE * PU) == 0, "product tables are padded to 256");
  for (int q0 = lane; q0 < N; q0 += TILE * PU) {
    unsigned code[PU][MAXF];
#pragma unroll
    for (int u = 0; u < PU; u++)
#pragma unroll
      for (int f = 0; f < MAXF; f++) code[u][f] = P[f * NPAD + q0 + u * TILE];
    double v[PU][MAXF];
#pragma unroll
    for (int u = 0; u < PU; u++)
#pragma unroll
      for (int f = 0; f < MAXF; f++) v[u][f] = coop_factor(code[u][f], V, R, F, K);
#pragma unroll
    for (int u = 0; u < PU; u++) {
      double pr = v[u][0];
#pragma unroll
      for (int f = 1; f < MAXF; f++) pr = pr * v[u][f];
      const int q = q0 + u * TILE;
      if (q < N) T[q] = pr;
    }
  }
}

// LDS at a byte address (the kernel's dynamic LDS starts at address 0)
__device__ __forceinline__ double lds_ld(unsigned addr) { return *(ldouble*)(size_t)addr; }
__device__ __forceinline__ void lds_st(unsigned addr, double v) { *(ldouble*)(size_t)addr = v; }

// A sum program (Vdot or JVS chains over the products in T), two terms of a
// chain per step.  Descriptor: product offset (T) | target offset << 16
// (Fun: T, Jac_SP: G) | second product offset << 32; bit 63 = the chain's
// first step; per-step coefficient tables for both terms.  Step: v = first ?
// c1*T[p1] : acc + c1*T[p1]; v = v + c2*T[p2] (an absent term is -0.0 *
// +0.0); the target gets v (Jac_SP: -v, the Ghimj build; the diagonal's
// +1/(h gamma) is added afterwards).  Products are not written here, so their
// reads run two steps ahead of use.
template <bool JVS, unsigned GZ>
__device__ __noinline__ void coop_sums(const unsigned long long* progg, const double* coefg,
                                       const double* coef2g, int steps, int lane) {
  gcu64* prog = (gcu64*)progg;
  gcdouble* cf = (gcdouble*)coefg;
  gcdouble* cf2 = (gcdouble*)coef2g;
  constexpr unsigned SB = JVS ? 0u : GZ;  // base of the target field
  unsigned long long r[kRing];
  double c[kRing], c2[kRing];
  gcu64* q = prog + lane;
  gcdouble* qc = cf + lane;
  gcdouble* qc2 = cf2 + lane;
#pragma unroll
  for (int k = 0; k < kRing; k++) {
    r[k] = q[k * TILE];
    c[k] = qc[k * TILE];
    c2[k] = qc2[k * TILE];
  }
  q += kRing * TILE;
  qc += kRing * TILE;
  qc2 += kRing * TILE;
  auto p1 = [](unsigned long long d) { return lds_ld(GZ + ((unsigned)d & 0xFFFFu)); };
  auto p2 = [](unsigned long long d) { return lds_ld(GZ + ((unsigned)(d >> 32) & 0xFFFFu)); };
  double x0 = p1(r[0]), x1 = p1(r[1]), y0 = p2(r[0]), y1 = p2(r[1]);
  double acc = 0.0;
  for (int base = 0; base < steps; base += kRing) {
#pragma unroll
    for (int k = 0; k < kRing; k++) {
      const unsigned long long d = r[k];
      // products of step base+k+2 (slot k+2; slots 0 and 1 hold the next
      // round's first two steps, refilled at k = 0 and 1)
      const double x2 = p1(r[(k + 2) % kRing]), y2 = p2(r[(k + 2) % kRing]);
      const double term = c[k] * x0;
      const double sum = acc + term;
      const double v1 = ((long long)d < 0) ? term : sum;
      const double term2 = c2[k] * y0;
      acc = v1 + term2;
      lds_st(SB + ((unsigned)d >> 16), JVS ? -acc : acc);
      x0 = x1;
      x1 = x2;
      y0 = y1;
      y1 = y2;
      // refilled after the last use (see coop_program)
      __builtin_amdgcn_sched_barrier(0);
      r[k] = q[k * TILE];
      c[k] = qc[k * TILE];
      c2[k] = qc2[k * TILE];
    }
    q += kRing * TILE;
    qc += kRing * TILE;
    qc2 += kRing * TILE;
  }
}

// An elimination program over LDS: LU (every field in G) or the triangular
// solves (a in G; b, target and init in T, forward results in X = T[0:NVAR),
// backward in Z = T[NVAR:2NVAR)), two terms of a chain per step.
// Descriptor pair: lo = a1 | b1 << 16 | a2 << 32 | b2 << 48; hi = target |
// init << 16; target bit 2: the first term divides, bit 0: the second
// does, bit 1: some lane of the step divides (the same in every lane); init
// = JUNK (the first address of T) unless the step is a chain's first: a
// solve's inits lie above JUNK, the LU's (in G) below it, so one comparison
// of the whole word tells.  Step: v = first ? X[init] : acc; v = div1 ?
// v / G[a1] : v - G[a1]*X[b1]; v = div2 ? v / G[a2] : v - G[a2]*X[b2] (an
// absent second term reads +0.0 twice); X[target] = v.  The LDS operands of
// step t+1 are read before step t stores.
template <bool SOLVE, unsigned GZ>
__device__ __noinline__ void coop_program(const unsigned long long* prog_lo,
                                          const unsigned long long* prog_hi, int steps,
                                          int lane) {
  constexpr unsigned XB = SOLVE ? GZ : 0u;  // base of b, target, init
  unsigned long long r[kRing], h[kRing];
  gcu64* q = (gcu64*)prog_lo + lane;
  gcu64* qh = (gcu64*)prog_hi + lane;
#pragma unroll
  for (int k = 0; k < kRing; k++) {
    r[k] = q[k * TILE];
    h[k] = qh[k * TILE];
  }
  q += kRing * TILE;
  qh += kRing * TILE;
  struct Ops {
    double ga, xb, ga2, xb2, xi;
  };
  auto operands = [](unsigned long long d, unsigned long long e) {
    const unsigned lo = (unsigned)d, lo2 = (unsigned)(d >> 32), hi = (unsigned)e;
    return Ops{lds_ld(lo & 0xFFFFu), lds_ld(XB + (lo >> 16)), lds_ld(lo2 & 0xFFFFu),
               lds_ld(XB + (lo2 >> 16)), lds_ld(XB + (hi >> 16))};
  };
  double acc = 0.0;
  Ops cur = operands(r[0], h[0]);
  for (int base = 0; base < steps; base += kRing) {
#pragma unroll
    for (int k = 0; k < kRing; k++) {
      const unsigned hi = (unsigned)h[k];
      // step base+k+1: slot k+1, or (k = RING-1) slot 0, refilled at k = 0
      const Ops nxt = operands(r[(k + 1) % kRing], h[(k + 1) % kRing]);
      const bool first = SOLVE ? (hi >= 0x10000u) : (hi < (GZ << 16));
      const double v0 = first ? cur.xi : acc;
      double v1 = v0 - cur.ga * cur.xb;
      double v;
      if (__builtin_amdgcn_readfirstlane(hi) & 2u) {
        // a step with divisions (the barrier keeps the divisions out of the
        // other steps)
        __builtin_amdgcn_sched_barrier(0);
        if (hi & 4u) v1 = v0 / cur.ga;
        v = v1 - cur.ga2 * cur.xb2;
        if (hi & 1u) v = v1 / cur.ga2;
      } else {
        v = v1 - cur.ga2 * cur.xb2;
      }
      lds_st(XB + (hi & 0xFFF8u), v);
      acc = v;
      cur = nxt;
      // the slot is refilled after its last use, so the load can land in
      // the same registers (no copies, and no wait for the ring, at the
      // round's end)
      __builtin_amdgcn_sched_barrier(0);
      r[k] = q[k * TILE];
      h[k] = qh[k * TILE];
    }
    q += kRing * TILE;
    qh += kRing * TILE;
  }
}

// The emitted rate routines, out of line (lane 0 runs them).
template <class M>
__device__ __noinline__ void coop_rates_fixed(double TEMP, double CF, gdouble* R) {
  M::rates_fixed(TEMP, CF, Acc1{R});
}
template <class M>
__device__ __noinline__ void coop_rates_sun(double TEMP, double CF, double SUN, double TIME,
                                            gdouble* R) {
  M::rates_sun(TEMP, CF, SUN, TIME, Acc1{R});
}
template <class M>
__device__ __noinline__ void coop_rates_state(double TEMP, double CF, double SUN, double TIME,
                                              gdouble* V, gdouble* F, gdouble* R) {
  M::rates_state(TEMP, CF, SUN, TIME, Acc1{V}, Acc1{F}, Acc1{R});
}

template <class M, class C, int S>
struct CoopCell {
  using L = CoopLayout<M, S>;
  gdouble* w;  // this wavefront's block in global memory
  ldouble* G;  // LDS: LU entries
  ldouble* T;  // LDS: scratch
  int lane;

  __device__ __forceinline__ gdouble* row(int r) const { return w + r; }
  __device__ __forceinline__ static gcdouble* cst() { return (gcdouble*)C::CONSTS(); }

  // all lanes: f(i) for the species this lane holds
  template <class F>
  __device__ __forceinline__ void vec(F&& f) const {
    for (int i = lane; i < M::NVAR; i += TILE) f(i);
  }
  // every lane sees the others' stores (global memory and LDS)
  __device__ __forceinline__ static void sync() { __syncthreads(); }

  // Fun(V) -> out (a global row): products, then the Vdot chains into the
  // results region of T, copied out
  __device__ __forceinline__ void fun(gcdouble* V, gdouble* out) const {
    coop_products<C::MAXF, C::NA_PAD>(C::APROD(), C::NA, V, (gcdouble*)row(L::RCT),
                                      (gcdouble*)row(L::FIX), cst(), T, lane);
    sync();
    coop_sums<false, C::GZ>(C::FUN(), C::FUN_COEF(), C::FUN_COEF2(), C::FUN_STEPS, lane);
    sync();
    vec([&](int i) { out[i] = lds_ld(C::GZ + C::FO + 8u * (unsigned)i); });
    sync();
  }

  // Jac_SP at (Y, rates) fused with the Ghimj build into G (LDS): G = -JVS,
  // then G[diag] += ghinv.  Returns the first row (1-based) whose diagonal is
  // exactly zero (the KppDecomp test of util/sutil.c:8), or 0.
  __device__ __forceinline__ int jac_ghimj(double ghinv) const {
    coop_products<C::MAXF, C::NB_PAD>(C::BPROD(), C::NB, (gcdouble*)row(L::Y),
                                      (gcdouble*)row(L::RCT), (gcdouble*)row(L::FIX), cst(), T,
                                      lane);
    sync();
    coop_sums<true, C::GZ>(C::JAC(), C::JAC_COEF(), C::JAC_COEF2(), C::JAC_STEPS, lane);
    sync();
    int ising = 0x7FFFFFFF;
    gcu32* dg = (gcu32*)C::DIAG();
    for (int i = lane; i < M::NVAR; i += TILE) {
      const double gd = G[dg[i]] + ghinv;
      G[dg[i]] = gd;
      if (gd == 0.0 && i + 1 < ising) ising = i + 1;
    }
    sync();
    for (int off = 32; off > 0; off >>= 1) {
      const int o = __shfl_xor(ising, off);
      ising = o < ising ? o : ising;
    }
    return ising == 0x7FFFFFFF ? 0 : ising;
  }

  __device__ __forceinline__ void decomp() const {
    coop_program<false, C::GZ>(C::LU(), C::LUH(), C::LU_STEPS, lane);
    sync();
  }
  // solve G x = T[0:NVAR) in place; the result is left in T[NVAR:2NVAR)
  __device__ __forceinline__ void solve() const {
    coop_program<true, C::GZ>(C::SOL(), C::SOLH(), C::SOL_STEPS, lane);
    sync();
  }

  // rate constants at time T (see Cell::rates_at); the serial rate code runs
  // on lane 0, which writes the global RCT row
  __device__ __forceinline__ void rates_at(const RosParams& p, const CellState& s, double Tm,
                                           bool accepted) const {
    if (p.rates_given) return;
    const bool tod = p.sun_mode == KPP_SUN_TOD;
    const bool sun_rates = tod || M::RATES_USE_TIME;
    const bool state_rates = M::RATES_USE_STATE && (accepted || tod || M::RATES_USE_TIME);
    if (!sun_rates && !state_rates) return;
    if (lane == 0) {
      const double sun = tod ? update_sun(Tm) : s.SUN;
      if (sun_rates) coop_rates_sun<M>(s.TEMP, s.CFACTOR, sun, Tm, row(L::RCT));
      if (state_rates)
        coop_rates_state<M>(s.TEMP, s.CFACTOR, sun, Tm, row(L::Y), row(L::FIX), row(L::RCT));
    }
    sync();
  }

  // ---- the integrator; the lane's own species (i = lane + 64 r) are kept
  // in registers (Regs), the rows of the block are what other lanes gather
  static constexpr int NR = (M::NVAR + TILE - 1) / TILE;
  struct Regs {
    double y[NR], f0[NR], fc[NR], dfdt[NR], k[S][NR];
  };
  template <class F>
  __device__ __forceinline__ void own(F&& f) const {
#pragma unroll
    for (int r = 0; r < NR; r++) {
      const int i = lane + r * TILE;
      if (i < M::NVAR) f(r, i);
    }
  }
  __device__ __forceinline__ void load_own(gdouble* rowp, double* v) const {
    own([&](int r, int i) { v[r] = rowp[i]; });
  }

  __device__ __forceinline__ void start(const RosParams& p, int cell, CellState& s,
                                        Regs& g) const {
    const size_t ld = (size_t)p.ldc;
    s.cell = cell;
    s.TEMP = p.temp[cell];
    s.CFACTOR = p.cfactor[cell];
    s.SUN = p.sun ? p.sun[cell] : 0.0;
    gdouble *Y = row(L::Y), *F = row(L::FIX), *R = row(L::RCT);
    own([&](int r, int i) {
      g.y[r] = p.C[i * ld + cell];
      Y[i] = g.y[r];
    });
    for (int i = lane; i < M::NFIX; i += TILE) F[i] = p.C[(M::NVAR + i) * ld + cell];
    if (p.rates_given) {
      for (int i = lane; i < M::NREACT; i += TILE) R[i] = p.rconst[i * ld + cell];
    } else if (p.rpre) {
      const size_t n = (size_t)p.ncells;
      for (int i = lane; i < M::NREACT; i += TILE) R[i] = p.rpre[i * n + cell];
    } else {
      for (int i = lane; i < M::NREACT; i += TILE) R[i] = 0.0;  // NUMBER-rate slots
      sync();
      if (lane == 0) {
        coop_rates_fixed<M>(s.TEMP, s.CFACTOR, R);
        if (p.sun_mode == KPP_SUN_FIXED && !M::RATES_USE_TIME)
          coop_rates_sun<M>(s.TEMP, s.CFACTOR, s.SUN, 0.0, R);
      }
    }
    s.Nfun = p.stat0[0]; s.Njac = p.stat0[1]; s.Nstp = p.stat0[2]; s.Nacc = p.stat0[3];
    s.Nrej = p.stat0[4]; s.Ndec = p.stat0[5]; s.Nsol = p.stat0[6]; s.Nsng = p.stat0[7];
    s.Nfb = 0;
    s.Nrun = 0;
    s.Tstart = p.tstart ? p.tstart[cell] : p.tstart0;
    s.Tend = p.tend ? p.tend[cell] : p.tend0;
    const double span = abs_diff_macro(s.Tend, s.Tstart);
    s.Hmax = (p.Hmax_user == 0.0) ? span : dmin(dabs(p.Hmax_user), span);
    const double hs_user = p.hstart ? p.hstart[cell] : p.Hstart_user;
    double Hstart = (hs_user == 0.0) ? dmax(p.Hmin, KPP_DELTAMIN) : dmin(dabs(hs_user), span);
    s.T = s.Tstart;
    s.Direction = (s.Tend >= s.Tstart) ? +1 : -1;
    sync();
    if (hs_user == 0.0 && p.hstart_auto)
      Hstart = dmax(p.Hmin, dmin(initial_step(p, s, g), span));
    s.Hexit = 0.0;
    s.H = dmin(Hstart, s.Hmax);
    if (dabs(s.H) <= 10.0 * p.Roundoff) s.H = KPP_DELTAMIN;
    s.RejectLastH = false;
    s.RejectMoreH = false;
    s.newstep = true;
  }

  // sum of T[0:n) in index order, computed by lane 0 and seen by all lanes
  __device__ __forceinline__ double ordered_sum(int n) const {
    double acc = 0.0;
    if (lane == 0)
      for (int i = 0; i < n; i++) acc = acc + T[i];
    return __shfl(acc, 0);
  }

  // Fun of the row V into the registers v (the chains write the row OUT)
  __device__ __forceinline__ void fun_to(gcdouble* V, int out_row, double* v) const {
    fun(V, row(out_row));
    load_own(row(out_row), v);
  }

  // the starting step estimate of Cell::initial_step (IPAR[4] = 1)
  __device__ __forceinline__ double initial_step(const RosParams& p, CellState& s,
                                                 Regs& g) const {
    auto sc = [&](int i, double y) {
      return p.vector_tol ? (p.atol[i] + p.rtol[i] * dabs(y)) : (p.atol0 + p.rtol0 * dabs(y));
    };
    rates_at(p, s, s.T, true);
    fun_to((gcdouble*)row(L::Y), L::FCN0, g.f0);
    own([&](int r, int i) {
      const double w = sc(i, g.y[r]), a = g.y[r] / w;
      T[i] = a * a;
    });
    sync();
    double d0 = ordered_sum(M::NVAR);
    sync();
    own([&](int r, int i) {
      const double b = g.f0[r] / sc(i, g.y[r]);
      T[i] = b * b;
    });
    sync();
    double d1 = ordered_sum(M::NVAR);
    d0 = sqrt(d0 / (double)M::NVAR);
    d1 = sqrt(d1 / (double)M::NVAR);
    const double h0 = (d0 < 1.0e-5 || d1 < 1.0e-5) ? 1.0e-6 : 0.01 * (d0 / d1);
    const double dh = (double)s.Direction * h0;
    gdouble* Y1 = row(L::YNEW);
    own([&](int r, int i) { Y1[i] = g.y[r] + dh * g.f0[r]; });
    sync();
    rates_at(p, s, s.T + dh, false);
    double f1[NR];
    fun_to((gcdouble*)Y1, L::FCN, f1);
    own([&](int r, int i) {
      const double c = (f1[r] - g.f0[r]) / sc(i, g.y[r]);
      T[i] = c * c;
    });
    sync();
    double d2 = ordered_sum(M::NVAR);
    sync();
    d2 = sqrt(d2 / (double)M::NVAR) / h0;
    s.Nfun += 2;
    const double dm = dmax(d1, d2);
    const double h1 =
        (dm <= 1.0e-15) ? dmax(1.0e-6, 1.0e-3 * h0) : pow(0.01 / dm, 1.0 / p.coef.ELO);
    return dmin(100.0 * h0, h1);
  }

  __device__ __forceinline__ void finish(const RosParams& p, const CellState& s, const Regs& g,
                                         int ierr, double Texit) const {
    const size_t ld = (size_t)p.ldc;
    const int cell = s.cell;
    own([&](int r, int i) { p.C[i * ld + cell] = g.y[r]; });
    if (lane == 0) {
      p.ierr[cell] = ierr;
      if (p.istat) {
        int* q = p.istat;
        q[0 * ld + cell] = s.Nfun; q[1 * ld + cell] = s.Njac; q[2 * ld + cell] = s.Nstp;
        q[3 * ld + cell] = s.Nacc; q[4 * ld + cell] = s.Nrej; q[5 * ld + cell] = s.Ndec;
        q[6 * ld + cell] = s.Nsol; q[7 * ld + cell] = s.Nsng;
      }
      if (p.rstat) {
        p.rstat[cell] = Texit;
        p.rstat[ld + cell] = s.Hexit;
      }
      if (p.fallbacks) p.fallbacks[cell] = 0;
    }
    sync();
  }

  // One step attempt (Cell::attempt's control flow and arithmetic).
  __device__ __forceinline__ int attempt(const RosParams& p, CellState& s, Regs& g) const {
    constexpr int NVAR = M::NVAR;
    const RosCoef& m = p.coef;
    const double Roundoff = p.Roundoff;
    const bool skip_dfdt = p.dfdt_zero;
    const double Dir = (double)s.Direction;

    if (s.newstep) {
      const bool more = ((s.Direction > 0) && ((s.T - s.Tend) + Roundoff <= 0.0)) ||
                        ((s.Direction < 0) && ((s.Tend - s.T) + Roundoff <= 0.0));
      if (!more) { finish(p, s, g, 1, s.T); return 1; }
      if (s.Nstp > p.max_steps) { finish(p, s, g, -6, s.T); return -6; }
      if (((s.T + 0.1 * s.H) == s.T) || (s.H <= Roundoff)) { finish(p, s, g, -7, s.T); return -7; }
      s.Hexit = s.H;
      s.H = dmin(s.H, abs_diff_macro(s.Tend, s.T));
      rates_at(p, s, s.T, true);
      fun_to((gcdouble*)row(L::Y), L::FCN0, g.f0);
      s.Nfun++;
      if (!p.autonomous) {
        s.Nfun++;
        if (!skip_dfdt) {
          const double Delta = sqrt(Roundoff) * dmax(KPP_DELTAMIN, dabs(s.T));
          rates_at(p, s, s.T + Delta, false);
          fun_to((gcdouble*)row(L::Y), L::DFDT, g.dfdt);
          const double inv = 1.0 / Delta;
          own([&](int r, int i) {
            const double d = g.dfdt[r] + (-1.0) * g.f0[r];
            g.dfdt[r] = (inv == 1.0) ? d : inv * d;
          });
        }
      }
      s.Njac++;
      s.newstep = false;
    }

    rates_at(p, s, s.T, false);
    int Nconsecutive = 0;
    for (;;) {
      const double ghinv = 1.0 / (Dir * s.H * m.Gamma[0]);
      const int ising = jac_ghimj(ghinv);
      s.Ndec++;
      if (ising == 0) break;
      s.Nsng++;
      Nconsecutive++;
      if (Nconsecutive <= 5) {
        s.H = s.H * 0.5;
      } else {
        finish(p, s, g, -8, s.T);
        return -8;
      }
    }
    decomp();

    const double H = s.H;
    gdouble* Ynew = row(L::YNEW);
    auto stage = [&](auto stage_no) __attribute__((always_inline)) {
      constexpr int istage = decltype(stage_no)::value;
      bool from_fcn0 = true;
#pragma unroll
      for (int q = 2; q <= istage; q++)
        if (m.NewF[q - 1]) from_fcn0 = false;
      if (istage > 1 && m.NewF[istage - 1]) {
        own([&](int r, int i) {
          double y = g.y[r];
#pragma unroll
          for (int j = 1; j <= istage - 1; j++) {
            const double a = m.A[(istage - 1) * (istage - 2) / 2 + j - 1];
            if (a != 0.0) y = y + a * g.k[j - 1][r];
          }
          Ynew[i] = y;
        });
        sync();
        const double Tau = s.T + m.Alpha[istage - 1] * Dir * H;
        rates_at(p, s, Tau, false);
        fun_to((gcdouble*)Ynew, L::FCN, g.fc);
        s.Nfun++;
      }
      double HC[S > 1 ? S - 1 : 1];
#pragma unroll
      for (int j = 1; j <= istage - 1; j++)
        HC[j - 1] = m.C[(istage - 1) * (istage - 2) / 2 + j - 1] / (Dir * H);
      const double HG = Dir * H * m.Gamma[istage - 1];
      const bool use_g = (!p.autonomous) && (m.Gamma[istage - 1] != 0.0) && !skip_dfdt &&
                         (HG != 0.0);
      // the stage's right-hand side into T[0:NVAR), solved in LDS
      own([&](int r, int i) {
        double k = from_fcn0 ? g.f0[r] : g.fc[r];
#pragma unroll
        for (int j = 1; j <= istage - 1; j++)
          if (HC[j - 1] != 0.0) k = k + HC[j - 1] * g.k[j - 1][r];
        if (use_g) k = k + HG * g.dfdt[r];
        T[i] = k;
      });
      sync();
      solve();
      own([&](int r, int i) { g.k[istage - 1][r] = T[NVAR + i]; });
      s.Nsol++;
    };
    stages(stage, std::make_integer_sequence<int, S>{});
    // new solution, error estimate and ros_ErrorNorm (int/rosenbrock.c:582-591, :632-659)
    double yn[NR];
    sync();
    own([&](int r, int i) {
      const double y = g.y[r];
      double ynr = y, ye = 0.0;
#pragma unroll
      for (int j = 1; j <= S; j++) {
        const double kj = g.k[j - 1][r];
        if (m.M[j - 1] != 0.0) ynr = ynr + m.M[j - 1] * kj;
        if (m.E[j - 1] != 0.0) ye = ye + m.E[j - 1] * kj;
      }
      yn[r] = ynr;
      const double Ymax = dmax(dabs(y), dabs(ynr));
      const double Scale =
          p.vector_tol ? (p.atol[i] + p.rtol[i] * Ymax) : (p.atol0 + p.rtol0 * Ymax);
      T[i] = (ye * ye) / (Scale * Scale);
    });
    sync();
    double Err = ordered_sum(NVAR);
    sync();
    Err = sqrt(Err / (double)NVAR);
    const double Fac = dmin(p.FacMax, dmax(p.FacMin, p.FacSafe / pow(Err, 1.0 / m.ELO)));
    double Hnew = H * Fac;
    s.Nstp++;
    if ((Err <= 1.0) || (H <= p.Hmin)) {
      if (p.hfirst && s.Nacc == p.stat0[3] && lane == 0) p.hfirst[s.cell] = H;
      s.Nacc++;
      gdouble* Y = row(L::Y);
      own([&](int r, int i) {
        g.y[r] = yn[r];
        Y[i] = yn[r];
      });
      sync();
      s.T += Dir * H;
      Hnew = dmax(p.Hmin, dmin(Hnew, s.Hmax));
      if (s.RejectLastH) Hnew = dmin(Hnew, H);
      s.RejectLastH = false;
      s.RejectMoreH = false;
      s.H = Hnew;
      s.newstep = true;
      s.Nrun = 0;
    } else {
      if (s.Nacc >= 1) s.Nrej++;
      if (s.RejectMoreH) Hnew = H * p.FacRej;
      s.RejectMoreH = s.RejectLastH;
      s.RejectLastH = true;
      s.H = Hnew;
      if (Err != Err || !(dabs(Hnew) <= DBL_MAX)) {
        finish(p, s, g, KPP_IERR_NONFINITE, s.T);
        return KPP_IERR_NONFINITE;
      }
      if (++s.Nrun > KPP_MAX_REJECT) {
        finish(p, s, g, KPP_IERR_REJECTS, s.T);
        return KPP_IERR_REJECTS;
      }
    }
    return 0;
  }
};

// Persistent grid of one-wavefront workgroups; each takes the next cell of
// the queue until the queue is exhausted (every wavefront reaches the exit).
// Register budget: as many wavefronts per SIMD as the LDS lets a compute
// unit hold (160 KiB over the cell's LDS bytes, four SIMDs), at most 8.
#ifndef KPPHIP_COOP_MAX_WAVES
#define KPPHIP_COOP_MAX_WAVES 2
#endif
template <class M, class C>
constexpr int coop_waves_per_eu() {
  const int wgs = (int)((160u * 1024u) / CoopLds<M, C>::BYTES);
  const int w = wgs / 4;
  return w < 1 ? 1 : (w > KPPHIP_COOP_MAX_WAVES ? KPPHIP_COOP_MAX_WAVES : w);
}

template <class M, class C, int S>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(coop_waves_per_eu<M, C>())))
void ros_coop_kernel(RosParams p) {
  using L = CoopLayout<M, S>;
  extern __shared__ double coop_lds[];
  const int lane = threadIdx.x;
  // the programs address LDS absolutely: G at 0, the T region at C::GZ (the
  // LDS-space address of the dynamic array, not its generic one)
  if ((unsigned)(size_t)(ldouble*)coop_lds != 0u) return;
  CoopCell<M, C, S> c{(gdouble*)(p.scratch + (size_t)blockIdx.x * L::ROWS), (ldouble*)coop_lds,
                      (ldouble*)coop_lds + M::LU_NONZERO + 2, lane};
  // the T region's ZERO slot holds +0.0 for the whole launch (idle steps read it)
  if (lane == 0) lds_st(C::GZ + C::ZSLOT, 0.0);
  __syncthreads();
  for (;;) {
    int k = 0;
    if (lane == 0) k = atomicAdd(p.queue, 1);
    k = __shfl(k, 0);
    if (k >= p.ncells) break;
    CellState s;
    typename CoopCell<M, C, S>::Regs g;
    c.start(p, p.order ? p.order[k] : k, s, g);
    while (c.attempt(p, s, g) == 0) {
    }
  }
}

}  // namespace kpphip
