// Batched SDIRK integrator for gfx950: KPP's C SDIRK (int/sdirk.c) on the
// mechanism compiler's device functions (Fun, Jac_SP, KppDecomp, KppSolve).
//
// One lane integrates one cell, one pass of SDIRK_Integrator's time loop
// (int/sdirk.c:586-809) per iteration of the lane's loop: the iteration
// matrix E = I/(h gamma) - J when the LU is not reused, the stages with
// their simplified Newton iterations, the error estimate and the step
// control, or a cut step after a Newton failure.  The working vectors are
// rows of a 64-lane tile in HBM (element i of lane l at row*64 + l), as for
// the Rosenbrock kernels.  Operation by operation as the reference (WAXPY's
// and WSCAL's special cases included), so with FP contraction off and the
// same rate constants the results are the CPU oracle's up to the device's
// pow and sqrt (SQRT is pow(x, 0.5) in the reference).
//
// Kept from the reference: SDIRK_PrepareMatrix (:905-960) takes H, SkipJac
// and its singularity flag by value -- a singular matrix halves only the
// local step (the LU is of the halved step, the solves scale with the
// caller's), and the caller never sees -8.  Deviations, as in the oracle:
// parameter errors return their codes (the reference prints them and goes
// on with the setting uninitialised), failures inside the loop return -6/-7
// (the reference returns 0), and a NaN error estimate or non-finite step
// ends the cell with -9, more than KPP_MAX_REJECT failed passes in a row
// with -10.
#pragma once
#include "refill.h"
#include "ros_device.h"

namespace kpphip {

template <class M, int S>
struct SdLayout {
  static constexpr int Y = 0;
  static constexpr int FIX = Y + M::NVAR;
  static constexpr int RCT = FIX + (M::NFIX > 0 ? M::NFIX : 1);
  static constexpr int Z = RCT + M::NREACT;
  static constexpr int G = Z + S * M::NVAR;
  static constexpr int TMP = G + M::NVAR;
  static constexpr int RHS = TMP + M::NVAR;
  static constexpr int SCAL = RHS + M::NVAR;
  static constexpr int FJAC = SCAL + M::NVAR;
  static constexpr int E = FJAC + M::LU_NONZERO;
  static constexpr int ROWS = E + M::LU_NONZERO;
};

struct SdState {
  int cell;
  double T, Tfinal, H, Hmax, Tdirection, Theta, NewtonIncrementOld;
  double TEMP, CFACTOR, SUN;
  double Texit, Hexit, Hnew;  // RSTATUS[1..3]
  bool SkipLU, SkipJac, Reject, FirstStep;
  int Nfun, Njac, Nstp, Nacc, Nrej, Ndec, Nsol, Nsng;
  int Nrun;
};

// KppDecomp that stops at the first exactly zero pivot (util/sutil.c:8),
// leaving the rows before it factored
template <class M>
KPPHIP_MODEL_CALL int call_decomp_checked(dptr G) {
  return M::KppDecompChecked(Acc{G});
}

template <class M, int S>
struct SdCell {
  using L = SdLayout<M, S>;
  using A = Acc;
  double* base;  // lane base inside its tile
  __device__ __forceinline__ A at(int row) const { return Acc{(gdouble*)(base + row * TILE)}; }

  // WSCAL (util/blas.c) with its special cases
  __device__ __forceinline__ static void wscal(double a, A X) {
    if (a == 1.0) return;
    for (int i = 0; i < M::NVAR; i++) {
      if (a == -1.0) X[i] = -X[i];
      else if (a == 0.0) X[i] = 0.0;
      else X[i] = a * X[i];
    }
  }
  // WAXPY: Y += a*X, nothing when a == 0
  __device__ __forceinline__ static void waxpy(double a, A X, A Yv) {
    if (a == 0.0) return;
    for (int i = 0; i < M::NVAR; i++) Yv[i] = Yv[i] + a * X[i];
  }

  // FUN_CHEM/JAC_CHEM (:1309-1361) call Update_SUN and Update_RCONST; rates
  // free of SUN and TIME were computed once per cell (same values), rates
  // reading the concentrations see the accepted state (KPP's global C, which
  // the integrator's Y is)
  __device__ __forceinline__ void rates_at(const RosParams& p, const SdState& s, double T,
                                           bool accepted) const {
    if (p.rates_given) return;
    const bool tod = p.sun_mode == KPP_SUN_TOD;
    const double sun = tod ? update_sun(T) : s.SUN;
    if (tod || M::RATES_USE_TIME) call_rates_sun<M>(s.TEMP, s.CFACTOR, sun, T, at(L::RCT).b);
    if (M::RATES_USE_STATE && (accepted || tod || M::RATES_USE_TIME))
      call_rates_state<M>(s.TEMP, s.CFACTOR, sun, T, at(L::Y).b, at(L::FIX).b, at(L::RCT).b);
  }

  __device__ void scale(const RosParams& p, const SdCoef& m) const {
    A Y = at(L::Y), SC = at(L::SCAL);
    for (int i = 0; i < M::NVAR; i++) {
      const double y = dabs(Y[i]);
      SC[i] = (m.ITOL == 0) ? 1.0 / (p.atol0 + p.rtol0 * y) : 1.0 / (p.atol[i] + p.rtol[i] * y);
    }
  }
  __device__ double norm(A X) const {
    A SC = at(L::SCAL);
    double err = 0.0;
    for (int i = 0; i < M::NVAR; i++) {
      const double v = X[i] * SC[i];
      err = err + v * v;  // pow(v, 2): the correctly rounded square
    }
    return dmax(sqrt(err / (double)M::NVAR), 1.0e-10);
  }

  __device__ void start(const RosParams& p, int cell, SdState& s) const {
    const SdCoef& m = p.sd;
    const size_t ld = (size_t)p.ldc;
    s.cell = cell;
    s.TEMP = p.temp[cell];
    s.CFACTOR = p.cfactor[cell];
    s.SUN = p.sun ? p.sun[cell] : 0.0;
    A Y = at(L::Y), F = at(L::FIX), R = at(L::RCT);
    for (int i = 0; i < M::NVAR; i++) Y[i] = p.C[i * ld + cell];
    for (int i = 0; i < M::NFIX; i++) F[i] = p.C[(M::NVAR + i) * ld + cell];
    if (p.rates_given) {
      for (int i = 0; i < M::NREACT; i++) R[i] = p.rconst[i * ld + cell];
    } else {
      call_rates_fixed<M>(s.TEMP, s.CFACTOR, R.b);
      if (p.sun_mode == KPP_SUN_FIXED && !M::RATES_USE_TIME)
        call_rates_sun<M>(s.TEMP, s.CFACTOR, s.SUN, 0.0, R.b);
    }
    s.Nfun = s.Njac = s.Nstp = s.Nacc = s.Nrej = s.Ndec = s.Nsol = s.Nsng = 0;
    s.Nrun = 0;
    s.Texit = s.Hexit = s.Hnew = 0.0;
    const double T0 = p.tstart ? p.tstart[cell] : p.tstart0;
    s.Tfinal = p.tend ? p.tend[cell] : p.tend0;
    // SDIRK (:388-410) with the ABS macro, then the integrator's first step
    // (:568-577); SIGN(x,y) = (x*y >= 0) ? x : -x
    const double span = abs_diff_macro(s.Tfinal, T0);
    s.Hmax = (m.Hmax_user == 0.0) ? span : dmin(dabs(m.Hmax_user), span);
    const double hs_user = p.hstart ? p.hstart[cell] : m.Hstart_user;
    const double Hstart = (hs_user == 0.0) ? dmax(m.Hmin, p.Roundoff) : dmin(dabs(hs_user), span);
    s.T = T0;
    s.Tdirection = ((1.0 * (s.Tfinal - T0)) >= 0) ? 1.0 : -1.0;
    double H = dmax(dabs(m.Hmin), dabs(Hstart));
    if (dabs(H) <= 10.0 * p.Roundoff) H = 1.0e-06;
    H = dmin(dabs(H), s.Hmax);
    s.H = ((H * s.Tdirection) >= 0) ? H : -H;
    s.SkipLU = s.SkipJac = s.Reject = false;
    s.FirstStep = true;
    s.Theta = 0.0;
    s.NewtonIncrementOld = 0.0;
    scale(p, m);
  }

  __device__ void finish(const RosParams& p, const SdState& s, int ierr) const {
    const size_t ld = (size_t)p.ldc;
    const int cell = s.cell;
    A Y = at(L::Y);
    for (int i = 0; i < M::NVAR; i++) p.C[i * ld + cell] = Y[i];
    p.ierr[cell] = ierr;
    if (p.istat) {
      int* q = p.istat;
      q[0 * ld + cell] = s.Nfun; q[1 * ld + cell] = s.Njac; q[2 * ld + cell] = s.Nstp;
      q[3 * ld + cell] = s.Nacc; q[4 * ld + cell] = s.Nrej; q[5 * ld + cell] = s.Ndec;
      q[6 * ld + cell] = s.Nsol; q[7 * ld + cell] = s.Nsng;
    }
    if (p.rstat) {
      p.rstat[cell] = s.Texit;
      p.rstat[ld + cell] = s.Hexit;
      p.rstat[2 * ld + cell] = s.Hnew;
    }
  }

  // SDIRK_Solve (:963-985): RHS *= 1/(H gamma), KppSolve
  __device__ __forceinline__ void solve(const SdCoef& m, double H, A X, SdState& s) const {
    wscal(1.0 / (H * m.Gamma), X);
    call_solve<M>(at(L::E).b, X.b);
    s.Nsol++;
  }

  // One pass of the time loop; 0 while the cell continues, else its code.
  __device__ int attempt(const RosParams& p, SdState& s) const {
    constexpr int N = M::NVAR;
    const SdCoef& m = p.sd;
    A Y = at(L::Y), F = at(L::FIX), R = at(L::RCT), G = at(L::G), TMP = at(L::TMP),
      RHS = at(L::RHS), E = at(L::E);
    if (!((s.Tfinal - s.T) * s.Tdirection - p.Roundoff > 0.0)) {
      finish(p, s, 1);
      return 1;
    }
    if (!s.SkipLU) {
      // SDIRK_PrepareMatrix with its local H and SkipJac.  The Jacobian is
      // kept between matrices: after a step that reused the LU, a rejection
      // builds the next matrix from the Jacobian of an earlier point, as
      // the reference's FJAC.
      double Hl = s.H;
      bool skipjac = s.SkipJac;
      int cons = 0;
      rates_at(p, s, s.T, true);
      for (;;) {
        const double HGammaInv = 1.0 / (Hl * m.Gamma);
        if (!skipjac) {
          s.Njac++;
          call_jac_sp<M>(Y.b, F.b, R.b, at(L::FJAC).b);
        }
        call_jac_to_ghimj<M>(at(L::FJAC).b, HGammaInv, E.b);
        const int ising = call_decomp_checked<M>(E.b);
        s.Ndec++;
        if (ising == 0) break;
        s.Nsng++;
        if (++cons >= 6) break;
        Hl = 0.5 * Hl;
        skipjac = true;
      }
    }
    if (s.Nstp > m.Max_no_steps) {
      finish(p, s, -6);
      return -6;
    }
    if ((s.T + 0.1 * s.H == s.T) || (dabs(s.H) <= p.Roundoff)) {
      finish(p, s, -7);
      return -7;
    }
    const double H = s.H;
    bool cycle = false;
    double Fac = 0.5;
    for (int is = 0; is < m.S && !cycle; is++) {
      A Zi = at(L::Z + is * N);
      for (int i = 0; i < N; i++) Zi[i] = 0.0;
      for (int i = 0; i < N; i++) G[i] = 0.0;
      for (int j = 0; j < is; j++) {
        A Zj = at(L::Z + j * N);
        waxpy(m.Theta[j][is], Zj, G);
        if (m.StartNewton == 1) waxpy(m.Alpha[j][is], Zj, Zi);
      }
      bool done = false;
      Fac = 0.5;
      double rate = 2.0;
      for (int it = 0; it < m.NewtonMaxit; it++) {
        for (int i = 0; i < N; i++) TMP[i] = Y[i] + Zi[i];
        const double Tf = s.T + m.C[is] * H;
        rates_at(p, s, Tf, false);
        call_fun<M>(TMP.b, F.b, R.b, RHS.b);
        s.Nfun++;
        wscal(H * m.Gamma, RHS);
        waxpy(-1.0, Zi, RHS);
        waxpy(1.0, G, RHS);
        solve(m, H, RHS, s);
        const double inc = norm(RHS);
        if (it == 0) {
          s.Theta = dabs(m.ThetaMin);
          rate = 2.0;
        } else {
          s.Theta = inc / s.NewtonIncrementOld;
          if (s.Theta < 0.99) {
            rate = s.Theta / (1.0 - s.Theta);
            const double pred =
                inc * pow(s.Theta, (double)(m.NewtonMaxit - (it + 1)) / (1.0 - s.Theta));
            if (pred >= m.NewtonTol) {
              const double q = dmin(10.0, pred / m.NewtonTol);
              Fac = 0.8 * pow(q, (-1.0 / (double)(1 + m.NewtonMaxit - it + 1)));
              break;
            }
          } else {
            break;
          }
        }
        s.NewtonIncrementOld = inc;
        waxpy(1.0, RHS, Zi);
        done = (rate * inc <= m.NewtonTol);
        if (done) break;
      }
      if (!done) {
        s.H = Fac * s.H;
        s.Reject = true;
        s.SkipJac = true;
        s.SkipLU = false;
        cycle = true;
      }
    }
    if (cycle) {
      if (!(dabs(s.H) <= DBL_MAX)) {
        finish(p, s, KPP_IERR_NONFINITE);
        return KPP_IERR_NONFINITE;
      }
      if (++s.Nrun > KPP_MAX_REJECT) {
        finish(p, s, KPP_IERR_REJECTS);
        return KPP_IERR_REJECTS;
      }
      return 0;
    }
    // error estimation (:727-737)
    s.Nstp++;
    for (int i = 0; i < N; i++) TMP[i] = 0.0;
    for (int i = 0; i < m.S; i++)
      if (m.E[i] != 0.0) waxpy(m.E[i], at(L::Z + i * N), TMP);
    solve(m, H, TMP, s);
    const double Err = norm(TMP);
    double fac = m.FacSafe * pow(Err, (-1.0 / m.ELO));
    fac = dmax(m.FacMin, dmin(m.FacMax, fac));
    double Hnew = H * fac;
    if (Err < 1.0) {
      s.FirstStep = false;
      s.Nacc++;
      if (p.hfirst && s.Nacc == 1) p.hfirst[s.cell] = H;
      s.T = s.T + H;
      for (int i = 0; i < m.S; i++)
        if (m.D[i] != 0.0) waxpy(m.D[i], at(L::Z + i * N), Y);
      scale(p, m);
      Hnew = s.Tdirection * dmin(dabs(Hnew), s.Hmax);
      s.Texit = s.T;
      s.Hexit = H;
      s.Hnew = Hnew;
      if (s.Reject) Hnew = s.Tdirection * dmin(dabs(Hnew), dabs(H));
      s.Reject = false;
      if ((s.T + Hnew / m.Qmin - s.Tfinal) * s.Tdirection > 0.0) {
        s.H = s.Tfinal - s.T;
      } else {
        const double Hratio = Hnew / H;
        s.SkipLU = (s.Theta <= m.ThetaMin) && (Hratio >= m.Qmin) && (Hratio <= m.Qmax);
        if (!s.SkipLU) s.H = Hnew;
      }
      s.SkipJac = false;
      s.Nrun = 0;
    } else {
      if (s.FirstStep || s.Reject) s.H = m.FacRej * H;
      else s.H = Hnew;
      s.Reject = true;
      s.SkipJac = true;
      s.SkipLU = false;
      if (s.Nacc >= 1) s.Nrej++;
      if (Err != Err || !(dabs(s.H) <= DBL_MAX)) {
        finish(p, s, KPP_IERR_NONFINITE);
        return KPP_IERR_NONFINITE;
      }
      if (++s.Nrun > KPP_MAX_REJECT) {
        finish(p, s, KPP_IERR_REJECTS);
        return KPP_IERR_REJECTS;
      }
    }
    return 0;
  }
};

// One lane per cell for its whole integration (p.lanes cells per wavefront).
template <class M, int S>
__global__ __launch_bounds__(64) KPPHIP_KERNEL_ATTR void sdirk_plain_kernel(RosParams p) {
  using L = SdLayout<M, S>;
  if ((int)threadIdx.x >= p.lanes) return;
  const int k = blockIdx.x * p.lanes + threadIdx.x;
  if (k >= p.ncells) return;
  SdCell<M, S> c{p.scratch + (size_t)blockIdx.x * L::ROWS * TILE + threadIdx.x};
  SdState s;
  c.start(p, p.order ? p.order[k] : k, s);
  while (c.attempt(p, s) == 0) {
  }
}

// A resident grid whose lanes take cells from a work queue (refill_loop).
template <class M, int S>
__global__ __launch_bounds__(64) KPPHIP_KERNEL_ATTR void sdirk_persistent_kernel(RosParams p) {
  using L = SdLayout<M, S>;
  const SdCell<M, S> c{p.scratch + (size_t)blockIdx.x * L::ROWS * TILE + threadIdx.x};
  refill_loop<SdCell<M, S>, SdState>(p, c);
}

}  // namespace kpphip
