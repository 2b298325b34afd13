// Batched fully implicit Runge-Kutta integrator for gfx950: KPP's C
// RungeKutta (int/runge_kutta.c; Radau-2A, Radau-1A, Lobatto-3C, Gauss) on
// the mechanism compiler's device functions, with the complex sparse LU of
// util/sutil.c:28-87 emitted by kppgen (KppDecompCmplxR, KppSolveCmplxR).
//
// One lane integrates one cell, one pass of RK_Integrator's time loop
// (:701-995) per iteration of the lane's loop: F(T, Y), the Jacobian and the
// real and complex factorisations when the LU is not reused, the simplified
// Newton iterations on the transformed stage system, the SDIRK stage of the
// SDIRK error estimator or the classical estimator, and the step control.
// The working vectors are rows of a 64-lane tile in HBM (element i of lane l
// at row*64 + l).  Operation by operation as the reference (WAXPY's zero
// test included), so with FP contraction off and the same rate constants
// the results are the CPU oracle's up to the device's pow (SQRT is
// pow(x, 0.5) in the reference).
//
// Kept from the reference: the Jacobian is kept between factorisations and
// reused after a rejection (so a matrix after a step that reused the LU is
// built from the Jacobian of an earlier point, as the reference's FJAC); the
// three function evaluations of RK_PrepareRHS are not counted in Nfun, the
// solves of RK_ErrorEstimate and of the SDIRK stage not in Nsol; SkipLU is
// left as it was when the step is cut to Tend; the -10 (step size) and -12
// (five singular matrices in a row) messages only printed, the integration
// going on.  Deviations, as in the oracle: -9 (more than Max_no_steps steps)
// ends the cell (the reference prints it and goes on), a NaN error estimate
// or non-finite step ends it with -14, and more than KPP_MAX_REJECT passes
// in a row in which T does not advance with -15.
#pragma once
#include "refill.h"
#include "ros_device.h"

#define KPP_RK_IERR_NONFINITE (-14)
#define KPP_RK_IERR_REJECTS (-15)

namespace kpphip {

template <class M>
struct RkLayout {
  static constexpr int N = M::NVAR;
  static constexpr int Y = 0;
  static constexpr int FIX = Y + N;
  static constexpr int RCT = FIX + (M::NFIX > 0 ? M::NFIX : 1);
  static constexpr int Z1 = RCT + M::NREACT;
  static constexpr int Z2 = Z1 + N, Z3 = Z2 + N, Z4 = Z3 + N;
  static constexpr int DZ1 = Z4 + N, DZ2 = DZ1 + N, DZ3 = DZ2 + N, DZ4 = DZ3 + N;
  static constexpr int G = DZ4 + N, TMP = G + N, FO = TMP + N, FS = FO + N, SCAL = FS + N;
  static constexpr int CONT = SCAL + N;  // CONT[k] at CONT + k*N, k = 0..2
  static constexpr int FJAC = CONT + 3 * N;
  static constexpr int E1 = FJAC + M::LU_NONZERO;
  static constexpr int E2R = E1 + M::LU_NONZERO;
  static constexpr int E2I = E2R + M::LU_NONZERO;
  static constexpr int ROWS = E2I + M::LU_NONZERO;
};

struct RkState {
  int cell;
  double T, Tend, H, Hold, Hacc, ErrOld, Theta, Tdirection, Hmin, Hmax;
  double TEMP, CFACTOR, SUN;
  double Texit, Hexit, Hnew;  // RSTATUS[0..2]
  bool SkipLU, SkipJac, Reject, FirstStep;
  int Nfun, Njac, Nstp, Nacc, Nrej, Ndec, Nsol, Nsng;
  int Nrun;
};


template <class M>
struct RkCell {
  using L = RkLayout<M>;
  using A = Acc;
  static constexpr int N = M::NVAR;
  double* base;  // lane base inside its tile
  __device__ __forceinline__ A at(int row) const { return Acc{(gdouble*)(base + row * TILE)}; }

  // WAXPY (util/blas.c): Y += a*X, nothing when a == 0
  __device__ __forceinline__ static void waxpy(double a, A X, A Yv) {
    if (a == 0.0) return;
    for (int i = 0; i < N; i++) Yv[i] = Yv[i] + a * X[i];
  }

  // FUN_CHEM/JAC_CHEM (:1997-2024) call Update_SUN and Update_RCONST; rates
  // free of SUN and TIME were computed once per cell (same values), rates
  // reading the concentrations see the accepted state (KPP's global C)
  __device__ __forceinline__ void rates_at(const RosParams& p, const RkState& s, double T,
                                           bool accepted) const {
    if (p.rates_given) return;
    const bool tod = p.sun_mode == KPP_SUN_TOD;
    const double sun = tod ? update_sun(T) : s.SUN;
    if (tod || M::RATES_USE_TIME) call_rates_sun<M>(s.TEMP, s.CFACTOR, sun, T, at(L::RCT).b);
    if (M::RATES_USE_STATE && (accepted || tod || M::RATES_USE_TIME))
      call_rates_state<M>(s.TEMP, s.CFACTOR, sun, T, at(L::Y).b, at(L::FIX).b, at(L::RCT).b);
  }

  // RK_ErrorScale (:1063-1083)
  __device__ void scale(const RosParams& p) const {
    A Y = at(L::Y), SC = at(L::SCAL);
    for (int i = 0; i < N; i++) {
      const double y = dabs(Y[i]);
      SC[i] = (p.rk.ITOL == 0) ? 1.0 / (p.atol0 + p.rtol0 * y) : 1.0 / (p.atol[i] + p.rtol[i] * y);
    }
  }
  // RK_ErrorNorm (:1354-1364)
  __device__ double norm(A X) const {
    A SC = at(L::SCAL);
    double s = 0.0;
    for (int i = 0; i < N; i++) {
      const double v = X[i] * SC[i];
      s = s + v * v;
    }
    return dmax(sqrt(s / (double)N), 1.0e-10);
  }

  __device__ void start(const RosParams& p, int cell, RkState& s) const {
    const RkCoef& m = p.rk;
    const size_t ld = (size_t)p.ldc;
    s.cell = cell;
    s.TEMP = p.temp[cell];
    s.CFACTOR = p.cfactor[cell];
    s.SUN = p.sun ? p.sun[cell] : 0.0;
    A Y = at(L::Y), F = at(L::FIX), R = at(L::RCT);
    for (int i = 0; i < N; i++) Y[i] = p.C[i * ld + cell];
    for (int i = 0; i < M::NFIX; i++) F[i] = p.C[(N + i) * ld + cell];
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
    s.Tend = p.tend ? p.tend[cell] : p.tend0;
    // RungeKutta (:520-545) with the unparenthesised ABS, then the
    // integrator's first step (:684-697); SIGN(x,y) = y >= 0 ? ABS(x) : -ABS(x)
    const double span = abs_diff_macro(s.Tend, T0);
    s.Hmin = (m.Hmin_user == 0.0) ? 0.0 : dmin(dabs(m.Hmin_user), span);
    s.Hmax = (m.Hmax_user == 0.0) ? span : dmin(dabs(m.Hmax_user), span);
    const double hs_user = p.hstart ? p.hstart[cell] : m.Hstart_user;
    const double Hstart = (hs_user == 0.0) ? 0.0 : dmin(dabs(hs_user), span);
    s.T = T0;
    s.Tdirection = ((s.Tend - T0) >= 0) ? 1.0 : -1.0;
    double H = dmin(dmax(dabs(s.Hmin), dabs(Hstart)), s.Hmax);
    if (dabs(H) <= 10.0 * p.Roundoff) H = 1.0e-06;
    H = (s.Tdirection >= 0) ? dabs(H) : -dabs(H);
    s.Hold = H;
    s.Reject = s.SkipJac = s.SkipLU = false;
    s.FirstStep = true;
    if ((s.T + H * 1.0001 - s.Tend) * s.Tdirection >= 0.0) H = s.Tend - s.T;
    s.H = H;
    s.Hacc = 0.0;
    s.ErrOld = 0.0;
    s.Theta = 0.0;
    scale(p);
  }

  __device__ void finish(const RosParams& p, const RkState& s, int ierr) const {
    const size_t ld = (size_t)p.ldc;
    const int cell = s.cell;
    A Y = at(L::Y);
    for (int i = 0; i < N; i++) p.C[i * ld + cell] = Y[i];
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

  // RK_PrepareRHS (:1152-1187): R = Z - h A F(Y + Z), uncounted evaluations
  __device__ void prepare_rhs(const RosParams& p, const RkState& s, double H) const {
    const RkCoef& m = p.rk;
    A Y = at(L::Y), TMP = at(L::TMP), FS = at(L::FS);
    A R1 = at(L::DZ1), R2 = at(L::DZ2), R3 = at(L::DZ3);
    A Z1 = at(L::Z1), Z2 = at(L::Z2), Z3 = at(L::Z3);
    for (int i = 0; i < N; i++) {
      R1[i] = Z1[i];
      R2[i] = Z2[i];
      R3[i] = Z3[i];
    }
    for (int j = 0; j < 3; j++) {
      A Zj = at(L::Z1 + j * N);
      for (int i = 0; i < N; i++) TMP[i] = Y[i] + Zj[i];
      rates_at(p, s, s.T + m.C[j] * H, false);
      call_fun<M>(TMP.b, at(L::FIX).b, at(L::RCT).b, FS.b);
      waxpy(-H * m.A[0][j], FS, R1);
      waxpy(-H * m.A[1][j], FS, R2);
      waxpy(-H * m.A[2][j], FS, R3);
    }
  }

  // RK_Solve (:1249-1295): transform, real solve, complex solve, back
  __device__ void solve3(const RosParams& p, double H, RkState& s) const {
    const RkCoef& m = p.rk;
    A R1 = at(L::DZ1), R2 = at(L::DZ2), R3 = at(L::DZ3);
    for (int i = 0; i < N; i++) {
      const double x1 = R1[i] / H, x2 = R2[i] / H, x3 = R3[i] / H;
      R1[i] = m.TinvAinv[0][0] * x1 + m.TinvAinv[0][1] * x2 + m.TinvAinv[0][2] * x3;
      R2[i] = m.TinvAinv[1][0] * x1 + m.TinvAinv[1][1] * x2 + m.TinvAinv[1][2] * x3;
      R3[i] = m.TinvAinv[2][0] * x1 + m.TinvAinv[2][1] * x2 + m.TinvAinv[2][2] * x3;
    }
    call_solve<M>(at(L::E1).b, R1.b);
    call_solve_cmplx<M>(at(L::E2R).b, at(L::E2I).b, R2.b, R3.b);
    for (int i = 0; i < N; i++) {
      const double x1 = R1[i], x2 = R2[i], x3 = R3[i];
      R1[i] = m.T[0][0] * x1 + m.T[0][1] * x2 + m.T[0][2] * x3;
      R2[i] = m.T[1][0] * x1 + m.T[1][1] * x2 + m.T[1][2] * x3;
      R3[i] = m.T[2][0] * x1 + m.T[2][1] * x2 + m.T[2][2] * x3;
    }
    s.Nsol++;
  }

  // RK_ErrorEstimate (:1298-1351); F2 in DZ1, F1 in DZ2 (free after Newton)
  __device__ double error_estimate(const RosParams& p, RkState& s, double H) const {
    const RkCoef& m = p.rk;
    A Y = at(L::Y), FO = at(L::FO), TMP = at(L::TMP), F2 = at(L::DZ1), F1 = at(L::DZ2);
    A Z1 = at(L::Z1), Z2 = at(L::Z2), Z3 = at(L::Z3), E1 = at(L::E1);
    const double HrkE1 = m.E[1] / H, HrkE2 = m.E[2] / H, HrkE3 = m.E[3] / H;
    for (int i = 0; i < N; i++) {
      const double f2 = HrkE1 * Z1[i] + HrkE2 * Z2[i] + HrkE3 * Z3[i];
      F2[i] = f2;
      TMP[i] = m.E[0] * FO[i] + f2;
    }
    call_solve<M>(E1.b, TMP.b);
    if (m.kind == 2 || m.kind == 4) call_solve<M>(E1.b, TMP.b);
    if (m.kind == 4) call_solve<M>(E1.b, TMP.b);
    double Err = norm(TMP);
    if (Err < 1.0) return Err;
    if (s.FirstStep || s.Reject) {
      for (int i = 0; i < N; i++) TMP[i] = Y[i] + TMP[i];
      rates_at(p, s, s.T, false);
      call_fun<M>(TMP.b, at(L::FIX).b, at(L::RCT).b, F1.b);
      s.Nfun++;
      for (int i = 0; i < N; i++) TMP[i] = F1[i] + F2[i];
      call_solve<M>(E1.b, TMP.b);
      Err = norm(TMP);
    }
    return Err;
  }

  __device__ __forceinline__ int end(const RosParams& p, RkState& s, int code) const {
    finish(p, s, code);
    return code;
  }
  // a failed Newton iteration (:855-861, :932-938)
  __device__ __forceinline__ int newton_failed(const RosParams& p, RkState& s, double Fac) const {
    s.H = Fac * s.H;
    s.Reject = true;
    s.SkipJac = true;
    s.SkipLU = false;
    if (!(dabs(s.H) <= DBL_MAX)) return end(p, s, KPP_RK_IERR_NONFINITE);
    if (++s.Nrun > KPP_MAX_REJECT) return end(p, s, KPP_RK_IERR_REJECTS);
    return 0;
  }

  // One pass of the time loop; 0 while the cell continues, else its code.
  __device__ int attempt(const RosParams& p, RkState& s) const {
    const RkCoef& m = p.rk;
    A Y = at(L::Y), F = at(L::FIX), R = at(L::RCT), FO = at(L::FO), TMP = at(L::TMP),
      G = at(L::G);
    A Z1 = at(L::Z1), Z2 = at(L::Z2), Z3 = at(L::Z3), Z4 = at(L::Z4), DZ4 = at(L::DZ4);
    if (!((s.Tend - s.T) * s.Tdirection - p.Roundoff > 0.0)) return end(p, s, 1);
    rates_at(p, s, s.T, true);
    call_fun<M>(Y.b, F.b, R.b, FO.b);
    s.Nfun++;
    if (!s.SkipLU) {
      if (!s.SkipJac) {
        call_jac_sp<M>(Y.b, F.b, R.b, at(L::FJAC).b);
        s.Njac++;
      }
      // RK_Decomp (:1190-1246)
      const double H = s.H;
      call_jac_to_ghimj<M>(at(L::FJAC).b, m.Gamma / H, at(L::E1).b);
      int ising = call_decomp_checked_g<M>(at(L::E1).b);
      if (ising == 0) {
        call_jac_to_ghimj<M>(at(L::FJAC).b, m.Alpha / H, at(L::E2R).b);
        call_diag_fill<M>(at(L::E2I).b, m.Beta / H);
        ising = call_decomp_cmplx<M>(at(L::E2R).b, at(L::E2I).b);
      }
      s.Ndec++;
      if (ising != 0) {
        // the reference prints -12 from the fifth singular matrix in a row
        // on and goes on halving
        s.Nsng++;
        s.H = s.H * 0.5;
        s.Reject = true;
        s.SkipJac = true;
        s.SkipLU = false;
        if (!(dabs(s.H) <= DBL_MAX) || s.H == 0.0) return end(p, s, KPP_RK_IERR_NONFINITE);
        if (++s.Nrun > KPP_MAX_REJECT) return end(p, s, KPP_RK_IERR_REJECTS);
        return 0;
      }
    }
    s.Nstp++;
    if (s.Nstp > m.Max_no_steps) return end(p, s, -9);
    // 0.1*ABS(H) <= ABS(T)*Roundoff: the reference prints -10 and goes on
    const double H = s.H;

    // starting values (:761-769)
    if (s.FirstStep || m.StartNewton == 0) {
      for (int i = 0; i < N; i++) {
        Z1[i] = 0.0;
        Z2[i] = 0.0;
        Z3[i] = 0.0;
      }
    } else {
      A C0 = at(L::CONT), C1 = at(L::CONT + N), C2 = at(L::CONT + 2 * N);
      const double r = H / s.Hold;
      const double x1 = 1.0 + m.C[0] * r, x2 = 1.0 + m.C[1] * r, x3 = 1.0 + m.C[2] * r;
      for (int i = 0; i < N; i++) {
        const double c0 = C0[i], c1 = C1[i], c2 = C2[i];
        Z1[i] = c0 + x1 * (c1 + x1 * c2);
        Z2[i] = c0 + x2 * (c1 + x2 * c2);
        Z3[i] = c0 + x3 * (c1 + x3 * c2);
      }
    }
    // simplified Newton iterations (:771-848)
    bool done = false;
    double Fac = 0.5, rate = 0.0, incOld = 0.0;
    int iter;
    for (iter = 1; iter <= m.NewtonMaxit; iter++) {
      prepare_rhs(p, s, H);
      solve3(p, H, s);
      const double n1 = norm(at(L::DZ1)), n2 = norm(at(L::DZ2)), n3 = norm(at(L::DZ3));
      const double inc = sqrt((n1 * n1 + n2 * n2 + n3 * n3) / 3.0);
      if (iter == 1) {
        s.Theta = dabs(m.ThetaMin);
        rate = 2.0;
      } else {
        s.Theta = inc / incOld;
        if (s.Theta < 0.99) {
          rate = s.Theta / (1.0 - s.Theta);
        } else {
          break;
        }
        if (iter < m.NewtonMaxit) {
          const double pred = inc * pow(s.Theta, (double)(m.NewtonMaxit - iter)) / (1.0 - s.Theta);
          if (pred >= m.NewtonTol) {
            const double q = dmin(10.0, pred / m.NewtonTol);
            Fac = 0.8 * pow(q, (-1.0 / (double)(1 + m.NewtonMaxit - iter)));
            break;
          }
        }
      }
      incOld = dmax(inc, p.Roundoff);
      waxpy(-1.0, at(L::DZ1), Z1);
      waxpy(-1.0, at(L::DZ2), Z2);
      waxpy(-1.0, at(L::DZ3), Z3);
      done = (rate * inc <= m.NewtonTol);
      if (done) break;
    }
    const int iterMain = iter;
    if (!done) return newton_failed(p, s, Fac);

    double Err;
    if (m.SdirkError == 1) {
      // SDIRK stage (:877-939)
      for (int i = 0; i < N; i++) Z4[i] = Z3[i];
      for (int i = 0; i < N; i++) G[i] = 0.0;
      waxpy(m.Bgam[0] * H, FO, G);
      waxpy(m.Theta[0], Z1, G);
      waxpy(m.Theta[1], Z2, G);
      waxpy(m.Theta[2], Z3, G);
      done = false;
      Fac = 0.5;
      for (iter = 1; iter <= m.NewtonMaxit; iter++) {
        for (int i = 0; i < N; i++) TMP[i] = Y[i] + Z4[i];
        rates_at(p, s, s.T + H, false);
        call_fun<M>(TMP.b, F.b, R.b, DZ4.b);
        s.Nfun++;
        waxpy(-1.0 * m.Gamma / H, Z4, DZ4);
        waxpy(m.Gamma / H, G, DZ4);
        call_solve<M>(at(L::E1).b, DZ4.b);
        const double inc = norm(DZ4);
        double thsd;
        if (iter == 1) {
          thsd = dabs(m.ThetaMin);
          rate = 2.0;
        } else {
          thsd = inc / incOld;
          if (thsd < 0.99) {
            rate = thsd / (1.0 - thsd);
            const double pred = inc * pow(thsd, (double)(m.NewtonMaxit - iter)) / (1.0 - thsd);
            if (pred >= m.NewtonTol) {
              const double q = dmin(10.0, pred / m.NewtonTol);
              Fac = 0.8 * pow(q, (-1.0 / (double)(1 + m.NewtonMaxit - iter)));
              break;
            }
          } else {
            break;
          }
        }
        incOld = inc;
        waxpy(1.0, DZ4, Z4);
        done = (rate * inc <= m.NewtonTol);
        if (done) break;
      }
      if (!done) return newton_failed(p, s, Fac);
      // error estimate (:965-971)
      for (int i = 0; i < N; i++) DZ4[i] = 0.0;
      if (m.D[0] != 0.0) waxpy(m.D[0], Z1, DZ4);
      if (m.D[1] != 0.0) waxpy(m.D[1], Z2, DZ4);
      if (m.D[2] != 0.0) waxpy(m.D[2], Z3, DZ4);
      waxpy(-1.0, Z4, DZ4);
      Err = norm(DZ4);
    } else {
      Err = error_estimate(p, s, H);
    }

    // new step size (:982-987, the main loop's iteration count)
    Fac = pow(Err, (-1.0 / m.ELO)) *
          dmin(m.FacSafe, (1.0 + 2 * m.NewtonMaxit) / (double)(iterMain + 2 * m.NewtonMaxit));
    Fac = dmin(m.FacMax, dmax(m.FacMin, Fac));
    double Hnew = Fac * H;

    if (Err < 1.0) {
      s.FirstStep = false;
      s.Nacc++;
      if (p.hfirst && s.Nacc == 1) p.hfirst[s.cell] = H;
      if (m.Gustafsson == 1) {
        // predictive controller of Gustafsson (:996-1005)
        if (s.Nacc > 1) {
          double fg = m.FacSafe * (H / s.Hacc) * pow(Err * Err / s.ErrOld, -0.25);
          fg = dmin(m.FacMax, dmax(m.FacMin, fg));
          Fac = dmin(Fac, fg);
          Hnew = Fac * H;
        }
        s.Hacc = H;
        s.ErrOld = dmax(1.0e-02, Err);
      }
      s.Hold = H;
      const double Tprev = s.T;
      s.T = s.T + H;
      if (m.D[0] != 0.0) waxpy(m.D[0], Z1, Y);
      if (m.D[1] != 0.0) waxpy(m.D[1], Z2, Y);
      if (m.D[2] != 0.0) waxpy(m.D[2], Z3, Y);
      if (m.StartNewton == 1) {
        // RK_Interpolate "make" (:1117-1134)
        A C0 = at(L::CONT), C1 = at(L::CONT + N), C2 = at(L::CONT + 2 * N);
        const double c0 = m.C[0], c1 = m.C[1], c2 = m.C[2];
        const double den = (c2 - c1) * (c1 - c0) * (c0 - c2);
        for (int i = 0; i < N; i++) {
          const double z1 = Z1[i], z2 = Z2[i], z3 = Z3[i];
          C0[i] = (-c2 * c2 * c1 * z1 + z3 * c1 * c0 * c0 + c1 * c1 * c2 * z1 -
                   c1 * c1 * c0 * z3 + c2 * c2 * c0 * z2 - z2 * c2 * c0 * c0) / den - z3;
          C1[i] = -(c0 * c0 * (z3 - z2) + c1 * c1 * (z1 - z3) + c2 * c2 * (z2 - z1)) / den;
          C2[i] = (c0 * (z3 - z2) + c1 * (z1 - z3) + c2 * (z2 - z1)) / den;
        }
      }
      scale(p);
      s.Texit = s.T;
      s.Hnew = Hnew;
      s.Hexit = H;
      Hnew = s.Tdirection * dmin(dmax(dabs(Hnew), s.Hmin), s.Hmax);
      if (s.Reject) Hnew = s.Tdirection * dmin(dabs(Hnew), dabs(H));
      s.Reject = false;
      if ((s.T + Hnew / m.Qmin - s.Tend) * s.Tdirection >= 0.0) {
        s.H = s.Tend - s.T;
      } else {
        const double Hratio = Hnew / H;
        s.SkipLU = (s.Theta <= m.ThetaMin) && (Hratio >= m.Qmin) && (Hratio <= m.Qmax);
        if (!s.SkipLU) s.H = Hnew;
      }
      s.SkipJac = false;
      // an accepted step of T + H == T does not count as progress
      if (s.T != Tprev) s.Nrun = 0;
      else if (++s.Nrun > KPP_MAX_REJECT) return end(p, s, KPP_RK_IERR_REJECTS);
    } else {
      if (s.FirstStep || s.Reject) s.H = m.FacRej * H;
      else s.H = Hnew;
      s.Reject = true;
      s.SkipJac = true;
      s.SkipLU = false;
      if (s.Nacc >= 1) s.Nrej++;
      if (Err != Err || !(dabs(s.H) <= DBL_MAX)) return end(p, s, KPP_RK_IERR_NONFINITE);
      if (++s.Nrun > KPP_MAX_REJECT) return end(p, s, KPP_RK_IERR_REJECTS);
    }
    return 0;
  }
};

// One lane per cell for its whole integration (p.lanes cells per wavefront).
template <class M>
__global__ __launch_bounds__(64) KPPHIP_KERNEL_ATTR void rk_plain_kernel(RosParams p) {
  using L = RkLayout<M>;
  if ((int)threadIdx.x >= p.lanes) return;
  const int k = blockIdx.x * p.lanes + threadIdx.x;
  if (k >= p.ncells) return;
  RkCell<M> c{p.scratch + (size_t)blockIdx.x * L::ROWS * TILE + threadIdx.x};
  RkState s;
  c.start(p, p.order ? p.order[k] : k, s);
  while (c.attempt(p, s) == 0) {
  }
}

// A resident grid whose lanes take cells from a work queue (refill_loop).
template <class M>
__global__ __launch_bounds__(64) KPPHIP_KERNEL_ATTR void rk_persistent_kernel(RosParams p) {
  using L = RkLayout<M>;
  const RkCell<M> c{p.scratch + (size_t)blockIdx.x * L::ROWS * TILE + threadIdx.x};
  refill_loop<RkCell<M>, RkState>(p, c);
}

}  // namespace kpphip
