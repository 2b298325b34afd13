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
// running value and, at the chain's end, stores it.  The loads of step t+1
// are issued before the store of step t (the schedule puts every read of a
// chain's result at least two steps after its last step), and LDS
// operations of one wavefront execute in order, so no barrier is needed
// inside a program.
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

// LDS of one cell: G[LU_NONZERO], then T (products, solve vectors X|Z,
// error terms).
template <class M, class C>
struct CoopLds {
  static constexpr int TSIZE = cmax(cmax(C::NA, C::NB), 2 * M::NVAR);
  static constexpr size_t BYTES = sizeof(double) * (size_t)(M::LU_NONZERO + TSIZE);
};

// Plain pointer accessor for the emitted rate routines.
struct PAcc {
  using value_type = double;
  double* b;
  __device__ __forceinline__ double& operator[](int i) const { return b[i]; }
};

constexpr unsigned long long kIdle = ~0ull;
constexpr unsigned kFacEnd = 0xFFFFFFFFu;

template <class M, class C, int S>
struct CoopCell {
  using L = CoopLayout<M, S>;
  double* w;   // this wavefront's block in global memory
  double* G;   // LDS: LU entries
  double* T;   // LDS: scratch
  int lane;

  __device__ __forceinline__ double* row(int r) const { return w + r; }

  // all lanes: f(i) for the species this lane holds
  template <class F>
  __device__ __forceinline__ void vec(F&& f) const {
    for (int i = lane; i < M::NVAR; i += TILE) f(i);
  }
  // every lane sees the others' stores (global memory and LDS)
  __device__ __forceinline__ static void sync() { __syncthreads(); }

  // product factor value: RCT[j], V[i], F[i] or a constant
  __device__ __forceinline__ double factor(unsigned code, const double* V) const {
    const unsigned idx = code & 0x3FFFFFFFu;
    switch (code >> 30) {
      case 0: return row(L::RCT)[idx];
      case 1: return V[idx];
      case 2: return row(L::FIX)[idx];
      default: return C::CONSTS()[idx];
    }
  }
  // T[q] = product q (A of Fun or B of Jac_SP) at state V, left to right
  __device__ __forceinline__ void products(const unsigned* P, int N, const double* V) const {
    for (int q = lane; q < N; q += TILE) {
      double p = factor(P[q], V);
      for (int f = 1; f < C::MAXF; f++) {
        const unsigned code = P[f * N + q];
        if (code == kFacEnd) break;
        p = p * factor(code, V);
      }
      T[q] = p;
    }
  }

  // Fun(V) -> out (a global row): products, then the Vdot chains
  __device__ void fun(const double* V, double* out) const {
    products(C::APROD(), C::NA, V);
    sync();
    const unsigned long long* prog = C::FUN();
    const double* cst = C::CONSTS();
    double acc = 0.0;
    for (int t = 0; t < C::FUN_STEPS; t++) {
      const unsigned long long d = prog[t * TILE + lane];
      if (d == kIdle) continue;
      const double term = ((d >> 50) & 1) ? 0.0 : cst[d & 0xFFFF] * T[(d >> 16) & 0xFFFF];
      acc = ((d >> 48) & 1) ? term : acc + term;
      if ((d >> 49) & 1) out[(d >> 32) & 0xFFFF] = acc;
    }
    sync();
  }

  // Jac_SP at (Y, rates) fused with the Ghimj build into G (LDS): G = -JVS,
  // G[diag] += ghinv.  Returns the first row (1-based) whose diagonal is
  // exactly zero (the KppDecomp test of util/sutil.c:8), or 0.
  __device__ int jac_ghimj(double ghinv) const {
    products(C::BPROD(), C::NB, row(L::Y));
    sync();
    const unsigned long long* prog = C::JAC();
    const double* cst = C::CONSTS();
    double acc = 0.0;
    for (int t = 0; t < C::JAC_STEPS; t++) {
      const unsigned long long d = prog[t * TILE + lane];
      if (d == kIdle) continue;
      const double term = ((d >> 50) & 1) ? 0.0 : cst[d & 0xFFFF] * T[(d >> 16) & 0xFFFF];
      acc = ((d >> 48) & 1) ? term : acc + term;
      if ((d >> 49) & 1) {
        double g = -(acc);
        if ((d >> 51) & 1) g = g + ghinv;
        G[(d >> 32) & 0xFFFF] = g;
      }
    }
    sync();
    int ising = 0x7FFFFFFF;
    const unsigned* dg = C::DIAG();
    for (int i = lane; i < M::NVAR; i += TILE)
      if (G[dg[i]] == 0.0 && i + 1 < ising) ising = i + 1;
    for (int off = 32; off > 0; off >>= 1) {
      const int o = __shfl_xor(ising, off);
      ising = o < ising ? o : ising;
    }
    return ising == 0x7FFFFFFF ? 0 : ising;
  }

  // An elimination program over LDS: LU (X = G) or the solves (X = T).
  // Step: FIRST: acc = X[init(t)]; DIV: acc = acc / G[a], else
  // acc = acc - G[a]*X[b]; LAST: X[t] = acc.
  template <bool SOLVE>
  __device__ __forceinline__ void program(const unsigned long long* prog, int steps) const {
    double* X = SOLVE ? T : G;
    auto init_of = [](unsigned t) {
      // backward solve chains start from the forward result X[i] (t = NVAR+i)
      return (SOLVE && t >= (unsigned)M::NVAR) ? t - (unsigned)M::NVAR : t;
    };
    unsigned long long d = steps > 0 ? prog[lane] : kIdle;
    double ga = 0.0, xb = 0.0, xt = 0.0;
    if (d != kIdle) {
      ga = G[d & 0xFFFF];
      if (!((d >> 50) & 1)) xb = X[(d >> 16) & 0xFFFF];
      if ((d >> 48) & 1) xt = X[init_of((d >> 32) & 0xFFFF)];
    }
    double acc = 0.0;
    for (int t = 0; t < steps; t++) {
      // operands of the next step, read before this step's store
      const unsigned long long dn = (t + 1 < steps) ? prog[(t + 1) * TILE + lane] : kIdle;
      double gan = 0.0, xbn = 0.0, xtn = 0.0;
      if (dn != kIdle) {
        gan = G[dn & 0xFFFF];
        if (!((dn >> 50) & 1)) xbn = X[(dn >> 16) & 0xFFFF];
        if ((dn >> 48) & 1) xtn = X[init_of((dn >> 32) & 0xFFFF)];
      }
      if (d != kIdle) {
        if ((d >> 48) & 1) acc = xt;
        acc = ((d >> 50) & 1) ? acc / ga : acc - ga * xb;
        if ((d >> 49) & 1) X[(d >> 32) & 0xFFFF] = acc;
      }
      d = dn;
      ga = gan;
      xb = xbn;
      xt = xtn;
    }
    sync();
  }
  __device__ void decomp() const { program<false>(C::LU(), C::LU_STEPS); }
  // solve G x = T[0:NVAR) in place; the result is left in T[NVAR:2NVAR)
  __device__ void solve() const { program<true>(C::SOL(), C::SOL_STEPS); }

  // rate constants at time T (see Cell::rates_at); the serial rate code runs
  // on lane 0, which writes the global RCT row
  __device__ void rates_at(const RosParams& p, const CellState& s, double Tm, bool accepted) const {
    if (p.rates_given) return;
    const bool tod = p.sun_mode == KPP_SUN_TOD;
    const bool sun_rates = tod || M::RATES_USE_TIME;
    const bool state_rates = M::RATES_USE_STATE && (accepted || tod || M::RATES_USE_TIME);
    if (!sun_rates && !state_rates) return;
    if (lane == 0) {
      const double sun = tod ? update_sun(Tm) : s.SUN;
      if (sun_rates) M::rates_sun(s.TEMP, s.CFACTOR, sun, Tm, PAcc{row(L::RCT)});
      if (state_rates)
        M::rates_state(s.TEMP, s.CFACTOR, sun, Tm, PAcc{row(L::Y)}, PAcc{row(L::FIX)},
                       PAcc{row(L::RCT)});
    }
    sync();
  }

  __device__ void start(const RosParams& p, int cell, CellState& s) const {
    const size_t ld = (size_t)p.ldc;
    s.cell = cell;
    s.TEMP = p.temp[cell];
    s.CFACTOR = p.cfactor[cell];
    s.SUN = p.sun ? p.sun[cell] : 0.0;
    double *Y = row(L::Y), *F = row(L::FIX), *R = row(L::RCT);
    for (int i = lane; i < M::NVAR; i += TILE) Y[i] = p.C[i * ld + cell];
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
        M::rates_fixed(s.TEMP, s.CFACTOR, PAcc{R});
        if (p.sun_mode == KPP_SUN_FIXED && !M::RATES_USE_TIME)
          M::rates_sun(s.TEMP, s.CFACTOR, s.SUN, 0.0, PAcc{R});
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
    if (hs_user == 0.0 && p.hstart_auto) Hstart = dmax(p.Hmin, dmin(initial_step(p, s), span));
    s.Hexit = 0.0;
    s.H = dmin(Hstart, s.Hmax);
    if (dabs(s.H) <= 10.0 * p.Roundoff) s.H = KPP_DELTAMIN;
    s.RejectLastH = false;
    s.RejectMoreH = false;
    s.newstep = true;
  }

  // sum of T[0:n) in index order, computed by lane 0 and seen by all lanes
  __device__ double ordered_sum(int n) const {
    double acc = 0.0;
    if (lane == 0)
      for (int i = 0; i < n; i++) acc = acc + T[i];
    return __shfl(acc, 0);
  }

  // the starting step estimate of Cell::initial_step (IPAR[4] = 1)
  __device__ double initial_step(const RosParams& p, CellState& s) const {
    double *Y = row(L::Y), *F0 = row(L::FCN0), *Y1 = row(L::YNEW), *F1 = row(L::FCN);
    auto sc = [&](int i, double y) {
      return p.vector_tol ? (p.atol[i] + p.rtol[i] * dabs(y)) : (p.atol0 + p.rtol0 * dabs(y));
    };
    rates_at(p, s, s.T, true);
    fun(Y, F0);
    vec([&](int i) {
      const double w = sc(i, Y[i]), a = Y[i] / w;
      T[i] = a * a;
    });
    sync();
    double d0 = ordered_sum(M::NVAR);
    sync();
    vec([&](int i) {
      const double b = F0[i] / sc(i, Y[i]);
      T[i] = b * b;
    });
    sync();
    double d1 = ordered_sum(M::NVAR);
    d0 = sqrt(d0 / (double)M::NVAR);
    d1 = sqrt(d1 / (double)M::NVAR);
    const double h0 = (d0 < 1.0e-5 || d1 < 1.0e-5) ? 1.0e-6 : 0.01 * (d0 / d1);
    const double dh = (double)s.Direction * h0;
    vec([&](int i) { Y1[i] = Y[i] + dh * F0[i]; });
    sync();
    rates_at(p, s, s.T + dh, false);
    fun(Y1, F1);
    vec([&](int i) {
      const double c = (F1[i] - F0[i]) / sc(i, Y[i]);
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

  __device__ void finish(const RosParams& p, const CellState& s, int ierr, double Texit) const {
    const size_t ld = (size_t)p.ldc;
    const int cell = s.cell;
    const double* Y = row(L::Y);
    for (int i = lane; i < M::NVAR; i += TILE) p.C[i * ld + cell] = Y[i];
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
  __device__ int attempt(const RosParams& p, CellState& s) const {
    constexpr int NVAR = M::NVAR;
    const RosCoef& m = p.coef;
    const double Roundoff = p.Roundoff;
    double *Y = row(L::Y), *Fcn0 = row(L::FCN0), *Fcn = row(L::FCN), *dFdT = row(L::DFDT),
           *Ynew = row(L::YNEW), *K0 = row(L::K);
    const bool skip_dfdt = p.dfdt_zero;
    const double Dir = (double)s.Direction;

    if (s.newstep) {
      const bool more = ((s.Direction > 0) && ((s.T - s.Tend) + Roundoff <= 0.0)) ||
                        ((s.Direction < 0) && ((s.Tend - s.T) + Roundoff <= 0.0));
      if (!more) { finish(p, s, 1, s.T); return 1; }
      if (s.Nstp > p.max_steps) { finish(p, s, -6, s.T); return -6; }
      if (((s.T + 0.1 * s.H) == s.T) || (s.H <= Roundoff)) { finish(p, s, -7, s.T); return -7; }
      s.Hexit = s.H;
      s.H = dmin(s.H, abs_diff_macro(s.Tend, s.T));
      rates_at(p, s, s.T, true);
      fun(Y, Fcn0);
      s.Nfun++;
      if (!p.autonomous) {
        s.Nfun++;
        if (!skip_dfdt) {
          const double Delta = sqrt(Roundoff) * dmax(KPP_DELTAMIN, dabs(s.T));
          rates_at(p, s, s.T + Delta, false);
          fun(Y, dFdT);
          const double inv = 1.0 / Delta;
          vec([&](int i) {
            const double d = dFdT[i] + (-1.0) * Fcn0[i];
            dFdT[i] = (inv == 1.0) ? d : inv * d;
          });
          sync();
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
        finish(p, s, -8, s.T);
        return -8;
      }
    }
    decomp();

    const double H = s.H;
    auto stage = [&](auto stage_no) __attribute__((always_inline)) {
      constexpr int istage = decltype(stage_no)::value;
      bool from_fcn0 = true;
#pragma unroll
      for (int q = 2; q <= istage; q++)
        if (m.NewF[q - 1]) from_fcn0 = false;
      if (istage > 1 && m.NewF[istage - 1]) {
        vec([&](int i) {
          double y = Y[i];
#pragma unroll
          for (int j = 1; j <= istage - 1; j++) {
            const double a = m.A[(istage - 1) * (istage - 2) / 2 + j - 1];
            if (a != 0.0) y = y + a * K0[(j - 1) * NVAR + i];
          }
          Ynew[i] = y;
        });
        sync();
        const double Tau = s.T + m.Alpha[istage - 1] * Dir * H;
        rates_at(p, s, Tau, false);
        fun(Ynew, Fcn);
        s.Nfun++;
      }
      const double* Fc = from_fcn0 ? Fcn0 : Fcn;
      double HC[S > 1 ? S - 1 : 1];
#pragma unroll
      for (int j = 1; j <= istage - 1; j++)
        HC[j - 1] = m.C[(istage - 1) * (istage - 2) / 2 + j - 1] / (Dir * H);
      const double HG = Dir * H * m.Gamma[istage - 1];
      const bool use_g = (!p.autonomous) && (m.Gamma[istage - 1] != 0.0) && !skip_dfdt &&
                         (HG != 0.0);
      // the stage's right-hand side into T[0:NVAR), solved in LDS
      vec([&](int i) {
        double k = Fc[i];
#pragma unroll
        for (int j = 1; j <= istage - 1; j++)
          if (HC[j - 1] != 0.0) k = k + HC[j - 1] * K0[(j - 1) * NVAR + i];
        if (use_g) k = k + HG * dFdT[i];
        T[i] = k;
      });
      sync();
      solve();
      double* Ki = K0 + (istage - 1) * NVAR;
      vec([&](int i) { Ki[i] = T[NVAR + i]; });
      sync();
      s.Nsol++;
    };
    stages(stage, std::make_integer_sequence<int, S>{});
    // new solution, error estimate and ros_ErrorNorm (int/rosenbrock.c:582-591, :632-659)
    vec([&](int i) {
      const double y = Y[i];
      double yn = y, ye = 0.0;
#pragma unroll
      for (int j = 1; j <= S; j++) {
        const double kj = K0[(j - 1) * NVAR + i];
        if (m.M[j - 1] != 0.0) yn = yn + m.M[j - 1] * kj;
        if (m.E[j - 1] != 0.0) ye = ye + m.E[j - 1] * kj;
      }
      Ynew[i] = yn;
      const double Ymax = dmax(dabs(y), dabs(yn));
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
      vec([&](int i) { Y[i] = Ynew[i]; });
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
        finish(p, s, KPP_IERR_NONFINITE, s.T);
        return KPP_IERR_NONFINITE;
      }
      if (++s.Nrun > KPP_MAX_REJECT) {
        finish(p, s, KPP_IERR_REJECTS, s.T);
        return KPP_IERR_REJECTS;
      }
    }
    return 0;
  }
};

// Persistent grid of one-wavefront workgroups; each takes the next cell of
// the queue until the queue is exhausted (every wavefront reaches the exit).
template <class M, class C, int S>
__global__ __launch_bounds__(64) void ros_coop_kernel(RosParams p) {
  using L = CoopLayout<M, S>;
  extern __shared__ double coop_lds[];
  const int lane = threadIdx.x;
  CoopCell<M, C, S> c{p.scratch + (size_t)blockIdx.x * L::ROWS, coop_lds,
                      coop_lds + M::LU_NONZERO, lane};
  for (;;) {
    int k = 0;
    if (lane == 0) k = atomicAdd(p.queue, 1);
    k = __shfl(k, 0);
    if (k >= p.ncells) break;
    CellState s;
    c.start(p, p.order ? p.order[k] : k, s);
    while (c.attempt(p, s) == 0) {
    }
  }
}

}  // namespace kpphip
