// Batched discrete adjoint Rosenbrock integrator for gfx950: KPP's C
// RosenbrockADJ (int/rosenbrock_adj.c) with AdjointType 1 (none) and 2
// (discrete) on the mechanism compiler's device functions, the adjoint's
// model functions (Hessian, HessTR_Vec, JacTR_SP_Vec, KppSolveTR) included.
//
// One lane integrates one cell: ros_FwdInt (:949-1186) forward, pushing a
// checkpoint per accepted step (T, H, the stage states, the stage vectors and,
// with SaveLU, the factored matrix) into the lane's column of a checkpoint
// tile in HBM, then ros_DadjInt (:1189-1324) backward over the checkpoints,
// all the cell's adjoint vectors together.  The working vectors are rows of a
// 64-lane tile (element i of lane l at row*64 + l), the checkpoints rows of
// a second tile, one slot of 2 + 2*S*NVAR (+ LU_NONZERO) rows per step.
// Operation by operation as the reference (WAXPY's zero test, WSCAL's special
// cases), so with FP contraction off and the same rate constants the results
// are the CPU oracle's up to the device's pow and sqrt (SQRT is pow(x, 0.5)
// in the reference).
//
// Kept from the reference: ros_PrepareMatrix halves only its own copy of H
// (the stages use the caller's H with that LU) and gives up after the sixth
// singular matrix (-8); Hexit (RSTATUS[2]) is set to H before the step is
// cut to Tend; the forward ends with -6 after Max_no_steps (199999) steps;
// the backward loop leaves out the first accepted step (stack_ptr > 0) unless
// the caller asks for every step.  Deviations, as in the oracle: a NaN error
// estimate or a non-finite step ends the cell with -10, more than
// KPP_MAX_REJECT rejections in a row with -11, a full checkpoint buffer with
// -12 (the reference exits the process).
#pragma once
#include "ros_device.h"

#define KPP_ADJ_IERR_NONFINITE (-10)
#define KPP_ADJ_IERR_REJECTS (-11)
#define KPP_ADJ_IERR_CHECKPOINTS (-12)

namespace kpphip {

template <class M>
KPPHIP_MODEL_CALL void call_hessian(dptr V, dptr F, dptr R, dptr H) {
  M::Hessian(Acc{V}, Acc{F}, Acc{R}, Acc{H});
}
template <class M>
KPPHIP_MODEL_CALL void call_hesstr_vec(dptr H, dptr U1, dptr U2, dptr O) {
  M::HessTR_Vec(Acc{H}, Acc{U1}, Acc{U2}, Acc{O});
}
template <class M>
KPPHIP_MODEL_CALL void call_jactr_vec(dptr J, dptr U, dptr O) {
  M::JacTR_SP_Vec(Acc{J}, Acc{U}, Acc{O});
}
// (LU)^T x = b from row X into the distinct row XX
template <class M>
KPPHIP_MODEL_CALL void call_solve_tr(dptr J, dptr X, dptr XX) {
  M::KppSolveTR(Acc{J}, Acc{X}, Acc{XX});
}

// the tile rows of one lane: fixed part, then per adjoint vector a
// Lambda_a (N), U_a (S*N) and V_a (S*N)
template <class M>
struct AdjLayout {
  static constexpr int N = M::NVAR;
  static constexpr int Y = 0;
  static constexpr int FIX = Y + N;
  static constexpr int RCT = FIX + (M::NFIX > 0 ? M::NFIX : 1);
  static constexpr int YNEW = RCT + M::NREACT;
  static constexpr int FCN0 = YNEW + N, FCN = FCN0 + N, DFDT = FCN + N, YERR = DFDT + N;
  static constexpr int TMP = YERR + N, TMP2 = TMP + N;
  static constexpr int K = TMP2 + N;             // stage vectors, KPP_SMAX * N
  static constexpr int YST = K + KPP_SMAX * N;   // stage states, KPP_SMAX * N
  static constexpr int JAC0 = YST + KPP_SMAX * N;
  static constexpr int GHIMJ = JAC0 + M::LU_NONZERO;
  static constexpr int DJDT = GHIMJ + M::LU_NONZERO;
  static constexpr int HES = DJDT + M::LU_NONZERO;
  static constexpr int FIXED = HES + M::NHESS;
};

struct AdjState {
  int cell;
  double T, Tend, H, Hmin, Hmax, Hstart, Direction;
  double TEMP, CFACTOR, SUN;
  double Texit, Hexit, Hnew;  // RSTATUS[1..3]
  int Nfun, Njac, Nstp, Nacc, Nrej, Ndec, Nsol, Nsng;
  int ptr;  // last checkpoint pushed (-1: none)
};

template <class M>
struct AdjCell {
  using L = AdjLayout<M>;
  using A = Acc;
  static constexpr int N = M::NVAR;
  double* base;  // lane base inside its work tile
  double* chk;   // lane base inside its checkpoint tile
  int S, CR;     // stages; rows per checkpoint slot

  __device__ __forceinline__ A at(int row) const { return Acc{(gdouble*)(base + row * TILE)}; }
  __device__ __forceinline__ A ck(int slot, int row) const {
    return Acc{(gdouble*)(chk + ((size_t)slot * CR + row) * TILE)};
  }
  __device__ __forceinline__ int lam(const RosParams& p, int a) const {
    return L::FIXED + a * N * (1 + 2 * S);
  }
  __device__ __forceinline__ int urow(const RosParams& p, int a, int i) const {
    return lam(p, a) + N + i * N;
  }
  __device__ __forceinline__ int vrow(const RosParams& p, int a, int i) const {
    return lam(p, a) + N + S * N + i * N;
  }

  // WAXPY / WSCAL / WCOPY (util/blas.c)
  __device__ __forceinline__ static void waxpy(int n, double a, A X, A Yv) {
    if (a == 0.0) return;
    for (int i = 0; i < n; i++) Yv[i] = Yv[i] + a * X[i];
  }
  __device__ __forceinline__ static void wscal(int n, double a, A X) {
    if (a == 1.0) return;
    if (a == -1.0) {
      for (int i = 0; i < n; i++) X[i] = -X[i];
    } else if (a == 0.0) {
      for (int i = 0; i < n; i++) X[i] = 0.0;
    } else {
      for (int i = 0; i < n; i++) X[i] = a * X[i];
    }
  }
  __device__ __forceinline__ static void wcopy(int n, A X, A Yv) {
    for (int i = 0; i < n; i++) Yv[i] = X[i];
  }

  // FunTemplate/JacTemplate/HessTemplate (:2474-2544) call Update_SUN and
  // Update_RCONST; rates free of SUN and TIME were computed once per cell,
  // rates reading the concentrations are recomputed at every call from the
  // integrator's state Y (KPP's global C)
  __device__ __forceinline__ void rates_at(const RosParams& p, const AdjState& s,
                                           double T) const {
    if (p.rates_given) return;
    const bool tod = p.sun_mode == KPP_SUN_TOD;
    const double sun = tod ? update_sun(T) : s.SUN;
    if (tod || M::RATES_USE_TIME) call_rates_sun<M>(s.TEMP, s.CFACTOR, sun, T, at(L::RCT).b);
    if (M::RATES_USE_STATE)
      call_rates_state<M>(s.TEMP, s.CFACTOR, sun, T, at(L::Y).b, at(L::FIX).b, at(L::RCT).b);
  }
  __device__ void fun(const RosParams& p, AdjState& s, double T, A Yv, A out) const {
    rates_at(p, s, T);
    call_fun<M>(Yv.b, at(L::FIX).b, at(L::RCT).b, out.b);
    s.Nfun++;
  }
  __device__ void jac(const RosParams& p, AdjState& s, double T, A Yv, A out) const {
    rates_at(p, s, T);
    call_jac_sp<M>(Yv.b, at(L::FIX).b, at(L::RCT).b, out.b);
    s.Njac++;
  }

  __device__ void start(const RosParams& p, int cell, AdjState& s) const {
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
      if (M::RATES_USE_STATE && p.sun_mode == KPP_SUN_FIXED && !M::RATES_USE_TIME)
        call_rates_state<M>(s.TEMP, s.CFACTOR, s.SUN, 0.0, Y.b, F.b, R.b);
    }
    for (int a = 0; a < p.adj.nadj; a++) {
      A La = at(lam(p, a));
      for (int i = 0; i < N; i++) La[i] = p.adj.lambda[(size_t)(a * N + i) * ld + cell];
    }
    s.Nfun = s.Njac = s.Nstp = s.Nacc = s.Nrej = s.Ndec = s.Nsol = s.Nsng = 0;
    s.Texit = s.Hexit = s.Hnew = 0.0;
    s.ptr = -1;
    s.T = p.tstart ? p.tstart[cell] : p.tstart0;
    s.Tend = p.tend ? p.tend[cell] : p.tend0;
    // RosenbrockADJ (:385-430) with the unparenthesised ABS of the reference
    const double span = abs_diff_macro(s.Tend, s.T);
    s.Hmin = p.Hmin;
    s.Hmax = (p.Hmax_user == 0.0) ? span : dmin(dabs(p.Hmax_user), span);
    s.Hstart = (p.Hstart_user == 0.0) ? dmax(p.Hmin, 1.0e-5) : dmin(dabs(p.Hstart_user), span);
    s.Direction = (s.Tend >= s.T) ? 1.0 : -1.0;
  }

  __device__ void finish(const RosParams& p, const AdjState& s, int ierr) const {
    const size_t ld = (size_t)p.ldc;
    const int cell = s.cell;
    A Y = at(L::Y);
    for (int i = 0; i < N; i++) p.C[i * ld + cell] = Y[i];
    for (int a = 0; a < p.adj.nadj; a++) {
      A La = at(lam(p, a));
      for (int i = 0; i < N; i++) p.adj.lambda[(size_t)(a * N + i) * ld + cell] = La[i];
    }
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

  // ros_PrepareMatrix (:1813-1872), H by value: true if still singular
  // after five halvings
  __device__ bool prepare(const RosParams& p, AdjState& s, double H) const {
    A J0 = at(L::JAC0), G = at(L::GHIMJ);
    int nconsecutive = 0;
    for (;;) {
      // Ghimj = -Jac0, then ghinv added on the diagonal
      const double ghinv = 1.0 / (s.Direction * H * p.coef.Gamma[0]);
      call_jac_to_ghimj<M>(J0.b, ghinv, G.b);
      const int ising = call_decomp_checked_g<M>(G.b);
      s.Ndec++;
      if (ising == 0) return false;
      s.Nsng++;
      nconsecutive++;
      if (nconsecutive <= 5) H = H * 0.5;
      else return true;
    }
  }

  // ros_ErrorNorm (:1747-1770) before its floor
  __device__ double error_norm(const RosParams& p) const {
    A Y = at(L::Y), Yn = at(L::YNEW), E = at(L::YERR);
    double err = 0.0;
    for (int i = 0; i < N; i++) {
      const double ymax = dmax(dabs(Y[i]), dabs(Yn[i]));
      const double sc = p.vector_tol ? p.atol[i] + p.rtol[i] * ymax : p.atol0 + p.rtol0 * ymax;
      const double q = E[i] / sc;
      err = err + q * q;
    }
    return sqrt(err / (double)N);
  }

  // ros_FwdInt (:949-1186): 1 on success, or the error code
  __device__ int forward(const RosParams& p, AdjState& s) const {
    const RosCoef& m = p.coef;
    const bool discrete = p.adj.adjoint != 0;
    A Y = at(L::Y), Yn = at(L::YNEW), F0 = at(L::FCN0), Fc = at(L::FCN);
    A dFdT = at(L::DFDT), Ye = at(L::YERR);
    double H = dmin(dmax(dabs(s.Hmin), dabs(s.Hstart)), dabs(s.Hmax));
    if (dabs(H) <= 10.0 * p.Roundoff) H = 1.0e-5;
    H = s.Direction * H;
    bool RejectLastH = false, RejectMoreH = false;
    int Nrun = 0;
    for (int i = 0; i < N; i++) Yn[i] = 0.0;
    while (((s.Direction > 0) && ((s.T - s.Tend) + p.Roundoff <= 0.0)) ||
           ((s.Direction < 0) && ((s.Tend - s.T) + p.Roundoff <= 0.0))) {
      if (s.Nstp > p.max_steps) return -6;
      if (((s.T + 0.1 * H) == s.T) || (H <= p.Roundoff)) return -7;
      s.Hexit = H;
      H = dmin(H, abs_diff_macro(s.Tend, s.T));
      fun(p, s, s.T, Y, F0);
      if (!p.autonomous) {
        const double Delta = sqrt(p.Roundoff) * dmax(1.0e-5, dabs(s.T));
        fun(p, s, s.T + Delta, Y, dFdT);
        waxpy(N, -1.0, F0, dFdT);
        wscal(N, 1.0 / Delta, dFdT);
      }
      jac(p, s, s.T, Y, at(L::JAC0));
      for (;;) {
        if (prepare(p, s, H)) return -8;
        for (int istage = 0; istage < m.S; istage++) {
          A Ki = at(L::K + istage * N);
          if (istage == 0) {
            wcopy(N, F0, Fc);
            if (discrete) {
              wcopy(N, Y, at(L::YST));
              wcopy(N, Y, Yn);
            }
          } else if (m.NewF[istage]) {
            wcopy(N, Y, Yn);
            for (int j = 0; j < istage; j++)
              waxpy(N, m.A[istage * (istage - 1) / 2 + j], at(L::K + j * N), Yn);
            fun(p, s, s.T + m.Alpha[istage] * s.Direction * H, Yn, Fc);
          }
          if (istage > 0 && discrete) wcopy(N, Yn, at(L::YST + istage * N));
          wcopy(N, Fc, Ki);
          for (int j = 0; j < istage; j++) {
            const double HC = m.C[istage * (istage - 1) / 2 + j] / (s.Direction * H);
            waxpy(N, HC, at(L::K + j * N), Ki);
          }
          if (!p.autonomous && m.Gamma[istage] != 0.0)
            waxpy(N, s.Direction * H * m.Gamma[istage], dFdT, Ki);
          call_solve<M>(at(L::GHIMJ).b, Ki.b);
          s.Nsol++;
        }
        wcopy(N, Y, Yn);
        for (int j = 0; j < m.S; j++) waxpy(N, m.M[j], at(L::K + j * N), Yn);
        wscal(N, 0.0, Ye);
        for (int j = 0; j < m.S; j++) waxpy(N, m.E[j], at(L::K + j * N), Ye);
        double Err = error_norm(p);
        if (Err != Err) return KPP_ADJ_IERR_NONFINITE;
        Err = dmax(Err, 1.0e-10);
        const double Fac = dmin(p.FacMax, dmax(p.FacMin, p.FacSafe / pow(Err, 1.0 / m.ELO)));
        double Hnew = H * Fac;
        s.Nstp++;
        if ((Err <= 1.0) || (H <= s.Hmin)) {
          s.Nacc++;
          if (discrete) {  // ros_DPush (:779-815)
            if (s.ptr + 1 >= p.adj.nchk) return KPP_ADJ_IERR_CHECKPOINTS;
            const int k = ++s.ptr;
            ck(k, 0)[0] = H;
            ck(k, 1)[0] = s.T;
            for (int r = 0; r < m.S * N; r++) {
              ck(k, 2)[r] = at(L::YST)[r];
              ck(k, 2 + m.S * N)[r] = at(L::K)[r];
            }
            if (p.adj.save_lu) {
              A G = at(L::GHIMJ), D = ck(k, 2 + 2 * m.S * N);
              for (int r = 0; r < M::LU_NONZERO; r++) D[r] = G[r];
            }
          }
          wcopy(N, Yn, Y);
          s.T = s.T + s.Direction * H;
          Hnew = dmax(s.Hmin, dmin(Hnew, s.Hmax));
          if (RejectLastH) Hnew = dmin(Hnew, H);
          s.Hexit = H;
          s.Hnew = Hnew;
          s.Texit = s.T;
          RejectLastH = RejectMoreH = false;
          H = Hnew;
          Nrun = 0;
          break;
        }
        if (RejectMoreH) Hnew = H * p.FacRej;
        RejectMoreH = RejectLastH;
        RejectLastH = true;
        H = Hnew;
        if (s.Nacc >= 1) s.Nrej++;
        if (!(dabs(H) <= DBL_MAX)) return KPP_ADJ_IERR_NONFINITE;
        if (++Nrun > KPP_MAX_REJECT) return KPP_ADJ_IERR_REJECTS;
      }
    }
    return 1;
  }

  // ros_DadjInt (:1189-1324)
  __device__ void backward(const RosParams& p, AdjState& s) const {
    const RosCoef& m = p.coef;
    const int nadj = p.adj.nadj;
    const int last = p.adj.first_step ? 0 : 1;
    A Jac = at(L::JAC0), dJdT = at(L::DJDT), Hes = at(L::HES);
    A Tmp = at(L::TMP), Tmp2 = at(L::TMP2);
    while (s.ptr >= last) {
      const int k = s.ptr--;
      const double H = ck(k, 0)[0];
      const double T = ck(k, 1)[0];
      A Yst0 = ck(k, 2);
      A G = p.adj.save_lu ? ck(k, 2 + 2 * m.S * N) : at(L::GHIMJ);
      if (!p.adj.save_lu) {
        jac(p, s, T, Yst0, Jac);
        const double Tau = 1.0 / (s.Direction * H * m.Gamma[0]);
        call_jac_to_ghimj<M>(Jac.b, Tau, G.b);
        (void)call_decomp_checked_g<M>(G.b);
        s.Ndec++;
      }
      rates_at(p, s, T);
      call_hessian<M>(Yst0.b, at(L::FIX).b, at(L::RCT).b, Hes.b);
      for (int istage = m.S - 1; istage >= 0; istage--) {
        for (int a = 0; a < nadj; a++) {
          A Ui = at(urow(p, a, istage));
          wcopy(N, at(lam(p, a)), Ui);
          wscal(N, m.M[istage], Ui);
        }
        for (int j = istage + 1; j < m.S; j++) {
          const double HA = m.A[j * (j - 1) / 2 + istage];
          const double HC = m.C[j * (j - 1) / 2 + istage] / (s.Direction * H);
          for (int a = 0; a < nadj; a++) {
            waxpy(N, HA, at(vrow(p, a, j)), at(urow(p, a, istage)));
            waxpy(N, HC, at(urow(p, a, j)), at(urow(p, a, istage)));
          }
        }
        for (int a = 0; a < nadj; a++) {
          A Ui = at(urow(p, a, istage));
          call_solve_tr<M>(G.b, Ui.b, Tmp.b);
          wcopy(N, Tmp, Ui);
          s.Nsol++;
        }
        jac(p, s, T + m.Alpha[istage] * s.Direction * H, ck(k, 2 + istage * N), Jac);
        for (int a = 0; a < nadj; a++)
          call_jactr_vec<M>(Jac.b, at(urow(p, a, istage)).b, at(vrow(p, a, istage)).b);
      }
      if (!p.autonomous) {  // ros_JacTimeDerivative with the stage-1 Jacobian
        const double Delta = sqrt(p.Roundoff) * dmax(1.0e-5, dabs(T));
        jac(p, s, T + Delta, Yst0, dJdT);
        waxpy(M::LU_NONZERO, -1.0, Jac, dJdT);
        wscal(M::LU_NONZERO, 1.0 / Delta, dJdT);
      }
      for (int istage = 0; istage < m.S; istage++) {
        for (int a = 0; a < nadj; a++) {
          A La = at(lam(p, a));
          waxpy(N, 1.0, at(vrow(p, a, istage)), La);
          call_hesstr_vec<M>(Hes.b, at(urow(p, a, istage)).b, ck(k, 2 + m.S * N + istage * N).b,
                             Tmp.b);
          waxpy(N, 1.0, Tmp, La);
        }
      }
      if (!p.autonomous) {
        for (int a = 0; a < nadj; a++) {
          for (int i = 0; i < N; i++) Tmp[i] = 0.0;
          for (int istage = 0; istage < m.S; istage++)
            waxpy(N, m.Gamma[istage], at(urow(p, a, istage)), Tmp);
          call_jactr_vec<M>(dJdT.b, Tmp.b, Tmp2.b);
          waxpy(N, H, Tmp2, at(lam(p, a)));
        }
      }
    }
  }
};

// One lane per cell for its whole integration, forward then backward
// (p.lanes cells per wavefront).
template <class M>
__global__ __launch_bounds__(64) KPPHIP_KERNEL_ATTR void adj_plain_kernel(RosParams p) {
  using L = AdjLayout<M>;
  if ((int)threadIdx.x >= p.lanes) return;
  const int k = blockIdx.x * p.lanes + threadIdx.x;
  if (k >= p.ncells) return;
  const int S = p.coef.S;
  const int rows = L::FIXED + p.adj.nadj * M::NVAR * (1 + 2 * S);
  const int CR = 2 + 2 * S * M::NVAR + (p.adj.save_lu ? M::LU_NONZERO : 0);
  AdjCell<M> c{p.scratch + (size_t)blockIdx.x * rows * TILE + threadIdx.x,
               p.adj.chk + (size_t)blockIdx.x * p.adj.nchk * CR * TILE + threadIdx.x, S, CR};
  AdjState s;
  c.start(p, p.order ? p.order[k] : k, s);
  int ierr = c.forward(p, s);
  if (ierr >= 0 && p.adj.adjoint) {
    c.backward(p, s);
    ierr = 1;
  }
  c.finish(p, s, ierr);
}

}  // namespace kpphip
