// Batched Rosenbrock integrator for gfx950: device side.
//
// One lane integrates one grid cell with the algorithm of KPP's
// int/rosenbrock.c (RosenbrockIntegrator :450-627, ros_PrepareMatrix
// :686-737, ros_ErrorNorm :632-659, ros_FunTimeDerivative :663-682); the
// model functions are the mechanism compiler's straight-line HIP output.
//
// Control flow.  The integrator is written as a state machine that performs
// ONE step attempt per iteration of the lane's loop.  All lanes of a
// wavefront execute the same attempt body whatever step their cell is at,
// so a wavefront only diverges at accept/reject and when a lane's cell ends.
// The plain kernel gives each lane one cell; the persistent kernel refills
// finished lanes from a global queue (batched with a wavefront ballot), so
// its wavefronts stay full while uneven cells drain.
//
// Memory layout.  Per-cell working vectors live in HBM as 64-lane tiles:
// element i of lane l of tile t is at scratch[(t*ROWS + row + i)*64 + l], so
// a wavefront-wide access of one element is one contiguous 512-byte segment
// and element offsets are compile-time constants from the lane's base.
//
// Arithmetic follows the reference operation by operation (comments at each
// step), so with FP contraction off and the same rate constants the results
// are the CPU oracle's bit for bit; exp/pow/cos from the device math library
// are the only source of difference.
#pragma once
#include <hip/hip_runtime.h>
#include <float.h>
#include <stdint.h>

#include <utility>

#include "kpphip_params.h"

// Build knobs (set per mechanism by kpp_amd/build.py):
//  KPPHIP_WAVES_PER_EU  minimum resident wavefronts per SIMD the register
//                       allocator must allow (0: compiler default)
//  KPPHIP_MODEL_INLINE  1: inline the model routines into the kernels;
//                       0: call them out of line (one copy per kernel)
#ifndef KPPHIP_WAVES_PER_EU
#define KPPHIP_WAVES_PER_EU 0
#endif
#ifndef KPPHIP_MODEL_INLINE
#define KPPHIP_MODEL_INLINE 0
#endif
//  KPPHIP_PIPELINED     must be 0.  The memory-pipelined KppDecompPF /
//                       KppSolvePF (kppgen --pipelined) are bitwise the plain
//                       routines on the host, but with them a persistent-kernel
//                       parity run stopped making progress on the GPU for a
//                       reason not found; they stay out of the integrator.
#ifndef KPPHIP_PIPELINED
#define KPPHIP_PIPELINED 0
#endif
#if KPPHIP_PIPELINED
#error "KPPHIP_PIPELINED=1 is disabled: the pipelined LU stalled a persistent-kernel run"
#endif
#if KPPHIP_WAVES_PER_EU > 0
#define KPPHIP_KERNEL_ATTR __attribute__((amdgpu_waves_per_eu(KPPHIP_WAVES_PER_EU)))
#else
#define KPPHIP_KERNEL_ATTR
#endif
// The out-of-line routines must stay internal (no `used`): the compiler
// then derives their register budget from their callers' launch bounds and
// waves-per-EU; an externally visible routine is compiled for the default
// 1024-thread workgroup, i.e. capped at 128 VGPRs whatever the kernel allows.
#if KPPHIP_MODEL_INLINE
#define KPPHIP_MODEL_CALL __device__ __forceinline__
#else
#define KPPHIP_MODEL_CALL __device__ __noinline__
#endif

namespace kpphip {

constexpr int TILE = 64;

// Tiles live in global memory.  Saying so in the pointer type matters for
// the out-of-line model routines: through a generic pointer every access is
// a flat load/store, whose results may return out of order, so the compiler
// has to drain all outstanding loads before using any of them; global loads
// return in order and can be kept in flight back to back.
using gdouble = __attribute__((address_space(1))) double;

// Element accessor for a lane inside a 64-lane tile.  value_type is the
// arithmetic type of the emitted routines' locals.
struct Acc {
  using value_type = double;
  gdouble* __restrict__ b;
  __device__ __forceinline__ gdouble& operator[](int i) const { return b[i * TILE]; }
};
// One lane's LDS mirror of LU entries (KppDecompL/KppSolveL): slot r at
// b[r*TILE], so the 64 lanes of a wavefront read 512 consecutive bytes.
using ldouble = __attribute__((address_space(3))) double;
struct AccL {
  using value_type = double;
  ldouble* b;
  __device__ __forceinline__ ldouble& operator[](int i) const { return b[i * TILE]; }
};
using gfloat = __attribute__((address_space(1))) float;
struct AccF {
  using value_type = float;
  gfloat* __restrict__ b;
  __device__ __forceinline__ gfloat& operator[](int i) const { return b[i * TILE]; }
};

// Row layout of a tile for mechanism M and S stages (a row holds one double
// per lane).  The mixed-precision integrator (MIX) adds an FP32 copy of the
// iteration matrix (two FP32 elements per row), an FP32 solve vector, and the
// FP64 first solution and residual of the refinement sweep.
template <class M, int S, bool MIX = false>
struct Layout {
  static constexpr int Y = 0;
  static constexpr int FIX = Y + M::NVAR;
  static constexpr int RCT = FIX + (M::NFIX > 0 ? M::NFIX : 1);
  static constexpr int FCN0 = RCT + M::NREACT;
  static constexpr int FCN = FCN0 + M::NVAR;
  static constexpr int DFDT = FCN + M::NVAR;
  static constexpr int YNEW = DFDT + M::NVAR;
  static constexpr int K = YNEW + M::NVAR;
  static constexpr int G = K + S * M::NVAR;
  static constexpr int G32 = G + M::LU_NONZERO;                 // FP32 rows
  static constexpr int X32 = G32 + (M::LU_NONZERO + 1) / 2;     // FP32 rows
  static constexpr int XR = X32 + (M::NVAR + 1) / 2;
  static constexpr int RR = XR + M::NVAR;
  static constexpr int ROWS = MIX ? RR + M::NVAR : G + M::LU_NONZERO;
};

// Refinement of the FP32 solves: the correction d of the one FP64 sweep must
// satisfy ||d||_inf <= kRefineTol * ||x||_inf, else the cell's attempt goes
// on with the FP64 factorisation (the sweep's contraction factor is about
// ||d||/||x||, so the refined solution is then accurate to ~1e-10).
constexpr double kRefineTol = 1e-5;

// util/UpdateSun.c:1-20
__device__ __forceinline__ double update_sun(double TIME) {
  const double SunRise = 4.5, SunSet = 19.5;
  const double PI = 3.14159265358979;
  double Thour = TIME / 3600.0;
  double Tlocal = Thour - (double)(((int)Thour / 24) * 24);
  if ((Tlocal >= SunRise) && (Tlocal <= SunSet)) {
    double Ttmp = (2.0 * Tlocal - SunRise - SunSet) / (SunSet - SunRise);
    if (Ttmp > 0) Ttmp = Ttmp * Ttmp;
    else Ttmp = -Ttmp * Ttmp;
    return (1.0 + cos(PI * Ttmp)) / 2.0;
  }
  return 0.0;
}

// MAX/MIN/ABS macros of int/rosenbrock.c:1-3 (not fmin/fmax: NaN order)
__device__ __forceinline__ double dmax(double a, double b) { return (a >= b) ? a : b; }
__device__ __forceinline__ double dmin(double b, double c) { return (b < c) ? b : c; }
__device__ __forceinline__ double dabs(double x) { return (x >= 0) ? x : -x; }
// ABS(a-b) as the unparenthesised macro of int/rosenbrock.c:3 expands it:
// ((a-b) >= 0) ? (a-b) : (-a-b).  For a < b (backward integration) this is
// -a-b, which makes Rosenbrock() stop with -7 at once; reproduced on purpose.
__device__ __forceinline__ double abs_diff_macro(double a, double b) {
  return ((a - b) >= 0) ? (a - b) : (-a - b);
}

// Model calls kept out of line: one copy of each large straight-line routine
// per kernel bounds instruction-cache footprint and register pressure.  The
// arrays a routine touches are distinct rows of the lane's tile; declaring
// them __restrict__ lets the compiler keep values in registers across stores
// and issue loads early (otherwise every store to X in KppSolve would order
// all later loads of G behind it).
using dptr = gdouble* __restrict__;
template <class M>
KPPHIP_MODEL_CALL void call_fun(dptr V, dptr F, dptr R, dptr out) {
  M::Fun(Acc{V}, Acc{F}, Acc{R}, Acc{out});
}
// Fun_SPLIT of a mechanism compiled with #FUNCTION SPLIT (the F90 flavour)
template <class M>
KPPHIP_MODEL_CALL void call_fun_split(dptr V, dptr F, dptr R, dptr out) {
  M::FunSplit(Acc{V}, Acc{F}, Acc{R}, Acc{out});
}
template <class M>
KPPHIP_MODEL_CALL int call_jac_ghimj(dptr V, dptr F, dptr R, double ghinv, dptr G) {
  return M::JacGhimj(Acc{V}, Acc{F}, Acc{R}, ghinv, Acc{G});
}
template <class M>
KPPHIP_MODEL_CALL void call_decomp(dptr G) {
  M::KppDecomp(Acc{G});
}
template <class M>
KPPHIP_MODEL_CALL void call_solve(dptr G, dptr X) {
  M::KppSolve(Acc{G}, Acc{X});
}
// LU with the first NL mirror slots in LDS (NL <= M::LDS_MAX)
template <class M, int NL>
KPPHIP_MODEL_CALL void call_decomp_l(dptr G, ldouble* Gs) {
  M::template KppDecompL<NL>(Acc{G}, AccL{Gs});
}
template <class M, int NL>
KPPHIP_MODEL_CALL void call_solve_l(dptr G, dptr X, ldouble* Gs) {
  M::template KppSolveL<NL>(Acc{G}, Acc{X}, AccL{Gs});
}
using fptr = gfloat* __restrict__;
template <class M>
KPPHIP_MODEL_CALL void call_decomp32(fptr G) {
  M::KppDecomp(AccF{G});
}
template <class M>
KPPHIP_MODEL_CALL void call_solve32(fptr G, fptr X) {
  M::KppSolve(AccF{G}, AccF{X});
}
template <class M>
KPPHIP_MODEL_CALL void call_residual(dptr G, dptr X, dptr B, dptr R) {
  M::Residual(Acc{G}, Acc{X}, Acc{B}, Acc{R});
}
template <class M>
KPPHIP_MODEL_CALL void call_rates_sun(double TEMP, double CFACTOR, double SUN, double TIME,
                                      dptr R) {
  M::rates_sun(TEMP, CFACTOR, SUN, TIME, Acc{R});
}
template <class M>
KPPHIP_MODEL_CALL void call_rates_state(double TEMP, double CFACTOR, double SUN, double TIME,
                                        dptr V, dptr F, dptr R) {
  M::rates_state(TEMP, CFACTOR, SUN, TIME, Acc{V}, Acc{F}, Acc{R});
}
template <class M>
KPPHIP_MODEL_CALL void call_rates_fixed(double TEMP, double CFACTOR, dptr R) {
  M::rates_fixed(TEMP, CFACTOR, Acc{R});
}
// the integrators that keep the Jacobian between factorisations (SDIRK,
// Runge-Kutta) and the complex LU of the Runge-Kutta integrator
template <class M>
KPPHIP_MODEL_CALL void call_jac_sp(dptr V, dptr F, dptr R, dptr J) {
  M::Jac_SP(Acc{V}, Acc{F}, Acc{R}, Acc{J});
}
template <class M>
KPPHIP_MODEL_CALL void call_jac_to_ghimj(dptr J, double ghinv, dptr G) {
  M::JacToGhimj(Acc{J}, ghinv, Acc{G});
}
template <class M>
KPPHIP_MODEL_CALL void call_diag_fill(dptr G, double v) {
  M::DiagFill(Acc{G}, v);
}
template <class M>
KPPHIP_MODEL_CALL int call_decomp_cmplx(dptr GR, dptr GI) {
  return M::KppDecompCmplxR(Acc{GR}, Acc{GI});
}
template <class M>
KPPHIP_MODEL_CALL void call_solve_cmplx(dptr GR, dptr GI, dptr XR, dptr XI) {
  M::KppSolveCmplxR(Acc{GR}, Acc{GI}, Acc{XR}, Acc{XI});
}
template <class M>
KPPHIP_MODEL_CALL int call_decomp_checked_g(dptr G) {
  return M::KppDecompChecked(Acc{G});
}

// Storage of a lane's working set.
//
// TileMem: one column of a 64-lane tile in global memory (every mechanism);
// the model routines are called out of line on row pointers.  With NL > 0
// the factored values of the NL most re-read LU entries live in the lane's
// LDS mirror `lds` instead (persistent kernel).
template <class M, int NL = 0>
struct TileMem {
  using A = Acc;
  static constexpr bool kRegisters = false;
  double* base;  // lane base pointer inside its tile
  ldouble* lds;  // NL > 0: the lane's mirror slots (slot r at lds[r*TILE])
  __device__ __forceinline__ A at(int row) const { return Acc{(gdouble*)(base + row * TILE)}; }
  __device__ __forceinline__ static void fun(A V, A F, A R, A out) {
    call_fun<M>(V.b, F.b, R.b, out.b);
  }
  __device__ __forceinline__ static void fun_split(A V, A F, A R, A out) {
    call_fun_split<M>(V.b, F.b, R.b, out.b);
  }
  __device__ __forceinline__ static int jac_ghimj(A V, A F, A R, double ghinv, A G) {
    return call_jac_ghimj<M>(V.b, F.b, R.b, ghinv, G.b);
  }
  __device__ __forceinline__ void decomp(A G) const {
    if constexpr (NL > 0)
      call_decomp_l<M, NL>(G.b, lds);
    else
      call_decomp<M>(G.b);
  }
  __device__ __forceinline__ void solve(A G, A X) const {
    if constexpr (NL > 0)
      call_solve_l<M, NL>(G.b, X.b, lds);
    else
      call_solve<M>(G.b, X.b);
  }
  __device__ __forceinline__ static void rates_sun(double T, double CF, double SUN, double TIME,
                                                   A R) {
    call_rates_sun<M>(T, CF, SUN, TIME, R.b);
  }
  __device__ __forceinline__ static void rates_state(double T, double CF, double SUN,
                                                     double TIME, A V, A F, A R) {
    call_rates_state<M>(T, CF, SUN, TIME, V.b, F.b, R.b);
  }
  __device__ __forceinline__ static void rates_fixed(double T, double CF, A R) {
    call_rates_fixed<M>(T, CF, R.b);
  }
};

// RegMem: a private array of the lane (small mechanisms).  Every element
// offset is a compile-time constant -- the model routines are inlined and the
// integrator's loops over species and reactions unrolled -- so the array
// lives in VGPRs and an attempt touches no memory but the cell's inputs and
// outputs.  Same operations in the same order as with TileMem.
struct RegAcc {
  using value_type = double;
  double* b;
  __device__ __forceinline__ double& operator[](int i) const { return b[i]; }
};
template <class M, int ROWS>
struct RegMem {
  using A = RegAcc;
  static constexpr bool kRegisters = true;
  mutable double a[ROWS];
  __device__ __forceinline__ A at(int row) const { return RegAcc{a + row}; }
  __device__ __forceinline__ static void fun(A V, A F, A R, A out) { M::Fun(V, F, R, out); }
  __device__ __forceinline__ static void fun_split(A V, A F, A R, A out) {
    M::FunSplit(V, F, R, out);
  }
  __device__ __forceinline__ static int jac_ghimj(A V, A F, A R, double ghinv, A G) {
    return M::JacGhimj(V, F, R, ghinv, G);
  }
  __device__ __forceinline__ static void decomp(A G) { M::KppDecomp(G); }
  __device__ __forceinline__ static void solve(A G, A X) { M::KppSolve(G, X); }
  __device__ __forceinline__ static void rates_sun(double T, double CF, double SUN, double TIME,
                                                   A R) {
    M::rates_sun(T, CF, SUN, TIME, R);
  }
  __device__ __forceinline__ static void rates_state(double T, double CF, double SUN,
                                                     double TIME, A V, A F, A R) {
    M::rates_state(T, CF, SUN, TIME, V, F, R);
  }
  __device__ __forceinline__ static void rates_fixed(double T, double CF, A R) {
    M::rates_fixed(T, CF, R);
  }
};

// Mechanisms whose working set is small enough for RegMem (at most this many
// LU entries: small_strato has 19; saprc99 920).
template <class M>
constexpr bool fits_registers() {
  return M::LU_NONZERO <= 64 && M::NREACT <= 64;
}

// for (i = 0; i < N; i++) f(i): fully unrolled for RegMem (compile-time
// indices), an ordinary loop otherwise; the calls run in increasing i either way.
template <class F, int... I>
__device__ __forceinline__ void vloop_seq(F& f, std::integer_sequence<int, I...>) {
  (f(I), ...);
}
template <bool Unroll, int N, class F>
__device__ __forceinline__ void vloop(F&& f) {
  if constexpr (Unroll) {
    vloop_seq(f, std::make_integer_sequence<int, N>{});
  } else {
    for (int i = 0; i < N; i++) f(i);
  }
}

// stage(1), ..., stage(S) as S separate calls with constant arguments
template <class F, int... I>
__device__ __forceinline__ void stages(F& f, std::integer_sequence<int, I...>) {
  (f(std::integral_constant<int, I + 1>{}), ...);
}

// Per-lane integration state between attempts.
struct CellState {
  int cell;
  double T, Tstart, Tend, H, Hmax, Hexit;
  double TEMP, CFACTOR, SUN;
  int Direction;
  bool RejectLastH, RejectMoreH, newstep;
  int Nfun, Njac, Nstp, Nacc, Nrej, Ndec, Nsol, Nsng;
  int Nfb;   // attempts that fell back from the FP32 to the FP64 factorisation
  int Nrun;  // rejected attempts since the last accepted one
  double Tacc, Hnew;  // f90: T and predicted step after the last accepted step
  double FacOld, ErrOld;  // the H211b filter's memory of the previous attempt
};

template <class M, int S, bool MIX = false, class Mem = TileMem<M>>
struct Cell {
  using L = Layout<M, S, MIX>;
  using A = typename Mem::A;
  static constexpr bool U = Mem::kRegisters;  // unroll the species/reaction loops
  static_assert(!(MIX && U), "the mixed-precision integrator works on tiles");
  Mem mem;
  __device__ __forceinline__ A at(int row) const { return mem.at(row); }
  // FP32 rows (tiles only): element i of this lane at float index 2*row*TILE + i*TILE
  __device__ __forceinline__ AccF at32(int row) const {
    return AccF{(gfloat*)((float*)(mem.base - lane_of()) + 2 * row * TILE + lane_of())};
  }
  __device__ __forceinline__ int lane_of() const { return (int)(threadIdx.x & (TILE - 1)); }

  // FunTemplate's model call: Fun, or for the F90 flavour of a mechanism
  // compiled with #FUNCTION SPLIT, Fun_SPLIT (KPP_FUN_OR_FUN_SPLIT,
  // src/code.c:279-280, 361-362; the C integrators always call Fun)
  __device__ __forceinline__ static void fcn(const RosParams& p, A V, A F, A R, A out) {
    if constexpr (M::FUN_SPLIT) {
      if (p.f90) {
        Mem::fun_split(V, F, R, out);
        return;
      }
    }
    Mem::fun(V, F, R, out);
  }

  // Rate constants at time T (int/rosenbrock.c:1257-1296 call Update_SUN and
  // Update_RCONST before every Fun/Jac).  Rates free of SUN and TIME were
  // computed once per cell with the same inputs, so not recomputing them
  // changes no value.  Rates that read the concentrations see the last
  // accepted state (KPP's global C, which Rosenbrock updates in place on
  // acceptance), so between acceptances they are recomputed only when SUN or
  // TIME can have changed; `accepted` marks the first call after a new state.
  __device__ __forceinline__ void rates_at(const RosParams& p, const CellState& s, double T,
                                           bool accepted) const {
    if (p.rates_given || p.rates_frozen) return;
    const bool tod = p.sun_mode == KPP_SUN_TOD;
    const double sun = tod ? update_sun(T) : s.SUN;
    if (tod || M::RATES_USE_TIME) Mem::rates_sun(s.TEMP, s.CFACTOR, sun, T, at(L::RCT));
    if (M::RATES_USE_STATE && (accepted || tod || M::RATES_USE_TIME))
      Mem::rates_state(s.TEMP, s.CFACTOR, sun, T, at(L::Y), at(L::FIX), at(L::RCT));
  }

  // Load a cell into this lane's slot; the set-up part of Rosenbrock()
  // (int/rosenbrock.c:286-336) and RosenbrockIntegrator (:495-507).
  __device__ void start(const RosParams& p, int cell, CellState& s) const {
    const size_t ld = (size_t)p.ldc;
    s.cell = cell;
    s.TEMP = p.temp[cell];
    s.CFACTOR = p.cfactor[cell];
    s.SUN = p.sun ? p.sun[cell] : 0.0;
    A Y = at(L::Y), F = at(L::FIX), R = at(L::RCT);
    vloop<U, M::NVAR>([&](int i) { Y[i] = p.C[i * ld + cell]; });
    vloop<U, M::NFIX>([&](int i) { F[i] = p.C[(M::NVAR + i) * ld + cell]; });
    if (p.rates_given) {
      vloop<U, M::NREACT>([&](int i) { R[i] = p.rconst[i * ld + cell]; });
    } else if (p.rpre) {
      // computed by ros_rates_prepass with the same inputs and formulae
      const size_t n = (size_t)p.ncells;
      vloop<U, M::NREACT>([&](int i) { R[i] = p.rpre[i * n + cell]; });
    } else {
      if constexpr (U) vloop<U, M::NREACT>([&](int i) { R[i] = 0.0; });  // NUMBER-rate slots
      Mem::rates_fixed(s.TEMP, s.CFACTOR, R);
      if (p.sun_mode == KPP_SUN_FIXED && !M::RATES_USE_TIME)
        Mem::rates_sun(s.TEMP, s.CFACTOR, s.SUN, 0.0, R);
    }
    s.Tstart = p.tstart ? p.tstart[cell] : p.tstart0;
    if (p.rates_frozen && !p.rates_given) {
      // every rate class once, at the cell's entry time and state
      if (M::RATES_USE_TIME) Mem::rates_sun(s.TEMP, s.CFACTOR, s.SUN, s.Tstart, R);
      if (M::RATES_USE_STATE) Mem::rates_state(s.TEMP, s.CFACTOR, s.SUN, s.Tstart, Y, F, R);
    }
    s.Nfun = p.stat0[0]; s.Njac = p.stat0[1]; s.Nstp = p.stat0[2]; s.Nacc = p.stat0[3];
    s.Nrej = p.stat0[4]; s.Ndec = p.stat0[5]; s.Nsol = p.stat0[6]; s.Nsng = p.stat0[7];
    s.Nfb = 0;
    s.Nrun = 0;
    s.Tend = p.tend ? p.tend[cell] : p.tend0;
    s.Tacc = 0.0;
    s.Hnew = 0.0;
    s.FacOld = 1.0;
    s.ErrOld = 1.0;
    if (p.f90) {
      // int/rosenbrock.f90: Hmax, Hstart (:364-385) and the first step of
      // ros_Integrator (:540-552), with a true ABS and the sign of direction
      const double span = dabs(s.Tend - s.Tstart);
      s.Hmax = (p.Hmax_user == 0.0) ? span : dmin(dabs(p.Hmax_user), span);
      const double hs_user = p.hstart ? p.hstart[cell] : p.Hstart_user;
      const double Hstart =
          (hs_user == 0.0) ? dmax(p.Hmin, KPP_DELTAMIN_F90) : dmin(dabs(hs_user), span);
      s.T = s.Tstart;
      s.Direction = (s.Tend >= s.Tstart) ? +1 : -1;
      double H = dmin(dmax(dabs(p.Hmin), dabs(Hstart)), dabs(s.Hmax));
      if (dabs(H) <= 10.0 * p.Roundoff) H = KPP_DELTAMIN_F90;
      s.H = (double)s.Direction * H;
      s.Hexit = 0.0;
      s.RejectLastH = false;
      s.RejectMoreH = false;
      s.newstep = true;
      return;
    }
    const double span = abs_diff_macro(s.Tend, s.Tstart);
    s.Hmax = (p.Hmax_user == 0.0) ? span : dmin(dabs(p.Hmax_user), span);
    const double hs_user = p.hstart ? p.hstart[cell] : p.Hstart_user;
    double Hstart = (hs_user == 0.0) ? dmax(p.Hmin, KPP_DELTAMIN) : dmin(dabs(hs_user), span);
    s.T = s.Tstart;
    s.Direction = (s.Tend >= s.Tstart) ? +1 : -1;
    if (hs_user == 0.0 && p.hstart_auto) Hstart = dmax(p.Hmin, dmin(initial_step(p, s), span));
    s.Hexit = 0.0;
    s.H = dmin(Hstart, s.Hmax);
    if (dabs(s.H) <= 10.0 * p.Roundoff) s.H = KPP_DELTAMIN;
    s.RejectLastH = false;
    s.RejectMoreH = false;
    s.newstep = true;
  }

  // Starting step estimate (no counterpart in int/rosenbrock.c, whose
  // default start is DeltaMin = 1e-6 s; selected with IPAR[4] = 1 when no
  // starting step is given): the algorithm of Hairer, Norsett & Wanner,
  // Solving ODEs I, II.4 ("Starting step size"), with the error norm's
  // weights sc_i = ATOL_i + RTOL_i |y0_i|:
  //   d0 = ||y0||, d1 = ||f(t0, y0)||, h0 = 0.01 d0/d1 (1e-6 if either is
  //   below 1e-5), y1 = y0 + h0 f(t0, y0), d2 = ||f(t0+h0, y1) - f(t0, y0)||/h0,
  //   h1 = (0.01/max(d1, d2))^(1/ELO) (max(1e-6, 1e-3 h0) if that max is
  //   below 1e-15), h = min(100 h0, h1).
  // Two function evaluations per cell, counted in Nfun.  Uses the tile rows
  // Fcn0, Ynew and Fcn, which the first attempt overwrites.
  __device__ double initial_step(const RosParams& p, CellState& s) const {
    constexpr int NVAR = M::NVAR;
    A Y = at(L::Y), F = at(L::FIX), R = at(L::RCT), F0 = at(L::FCN0), Y1 = at(L::YNEW),
      F1 = at(L::FCN);
    auto sc = [&](int i, double y) {
      return p.vector_tol ? (p.atol[i] + p.rtol[i] * dabs(y)) : (p.atol0 + p.rtol0 * dabs(y));
    };
    rates_at(p, s, s.T, true);
    Mem::fun(Y, F, R, F0);
    double d0 = 0.0, d1 = 0.0;
    vloop<U, NVAR>([&](int i) {
      const double w = sc(i, Y[i]), a = Y[i] / w, b = F0[i] / w;
      d0 = d0 + a * a;
      d1 = d1 + b * b;
    });
    d0 = sqrt(d0 / (double)NVAR);
    d1 = sqrt(d1 / (double)NVAR);
    const double h0 = (d0 < 1.0e-5 || d1 < 1.0e-5) ? 1.0e-6 : 0.01 * (d0 / d1);
    const double dh = (double)s.Direction * h0;
    vloop<U, NVAR>([&](int i) { Y1[i] = Y[i] + dh * F0[i]; });
    rates_at(p, s, s.T + dh, false);
    Mem::fun(Y1, F, R, F1);
    double d2 = 0.0;
    vloop<U, NVAR>([&](int i) {
      const double c = (F1[i] - F0[i]) / sc(i, Y[i]);
      d2 = d2 + c * c;
    });
    d2 = sqrt(d2 / (double)NVAR) / h0;
    s.Nfun += 2;
    const double dm = dmax(d1, d2);
    const double h1 = (dm <= 1.0e-15) ? dmax(1.0e-6, 1.0e-3 * h0) : pow(0.01 / dm, 1.0 / p.coef.ELO);
    return dmin(100.0 * h0, h1);
  }

  __device__ void finish(const RosParams& p, const CellState& s, int ierr, double Texit) const {
    const size_t ld = (size_t)p.ldc;
    const int cell = s.cell;
    A Y = at(L::Y);
    vloop<U, M::NVAR>([&](int i) { p.C[i * ld + cell] = Y[i]; });
    p.ierr[cell] = ierr;
    if (p.istat) {
      int* q = p.istat;
      q[0 * ld + cell] = s.Nfun; q[1 * ld + cell] = s.Njac; q[2 * ld + cell] = s.Nstp;
      q[3 * ld + cell] = s.Nacc; q[4 * ld + cell] = s.Nrej; q[5 * ld + cell] = s.Ndec;
      q[6 * ld + cell] = s.Nsol; q[7 * ld + cell] = s.Nsng;
    }
    if (p.rstat) {
      if (p.f90) {
        // RSTATUS(1:3) of int/rosenbrock.f90: T, H and Hnew of the last
        // accepted step (zero if none)
        p.rstat[cell] = s.Tacc;
        p.rstat[ld + cell] = s.Hexit;
        p.rstat[2 * ld + cell] = s.Hnew;
      } else {
        p.rstat[cell] = Texit;
        p.rstat[ld + cell] = s.Hexit;
      }
    }
    if (p.fallbacks) p.fallbacks[cell] = s.Nfb;
  }

  // One step attempt.  Returns 0 while the cell continues, else its IERR
  // (and the lane's cell is finished).
  __device__ int attempt(const RosParams& p, CellState& s) const {
    constexpr int NVAR = M::NVAR;
    const RosCoef& m = p.coef;
    const double Roundoff = p.Roundoff;
    A Y = at(L::Y), F = at(L::FIX), R = at(L::RCT), Fcn0 = at(L::FCN0), Fcn = at(L::FCN),
      dFdT = at(L::DFDT), Ynew = at(L::YNEW), K0 = at(L::K), G = at(L::G);
    const bool skip_dfdt = p.dfdt_zero;
    const bool f90 = p.f90 != 0;
    const double Dir = (double)s.Direction;

    if (s.newstep) {
      // time loop test and step guards (int/rosenbrock.c:510-524)
      const bool more = ((s.Direction > 0) && ((s.T - s.Tend) + Roundoff <= 0.0)) ||
                        ((s.Direction < 0) && ((s.Tend - s.T) + Roundoff <= 0.0));
      if (!more) { finish(p, s, 1, s.T); return 1; }
      if (s.Nstp > p.max_steps) { finish(p, s, -6, s.T); return -6; }
      if (((s.T + 0.1 * s.H) == s.T) || (s.H <= Roundoff)) { finish(p, s, -7, s.T); return -7; }
      if (p.f90) {
        s.H = dmin(s.H, dabs(s.Tend - s.T));
      } else {
        s.Hexit = s.H;
        s.H = dmin(s.H, abs_diff_macro(s.Tend, s.T));
      }
      // Fcn0 = F(T, Y)
      rates_at(p, s, s.T, true);
      fcn(p, Y, F, R, Fcn0);
      s.Nfun++;
      // dF/dT by forward difference (int/rosenbrock.c:663-682)
      if (!p.autonomous) {
        s.Nfun++;
        if (!skip_dfdt) {
          const double Delta = sqrt(Roundoff) * dmax(KPP_DELTAMIN, dabs(s.T));
          rates_at(p, s, s.T + Delta, false);
          fcn(p, Y, F, R, dFdT);
          const double inv = 1.0 / Delta;
          vloop<U, NVAR>([&](int i) {
            const double d = dFdT[i] + (-1.0) * Fcn0[i];
            dFdT[i] = (inv == 1.0) ? d : inv * d;
          });
        }
      }
      // Jac0 at (T, Y): counted here, evaluated inside every Ghimj build
      // below from the same (T, Y, rates), hence with the same values.
      s.Njac++;
      s.newstep = false;
    }

    // ros_PrepareMatrix: Ghimj = -Jac0 + I/(H*gamma1), KppDecomp, halve H on
    // an exactly zero diagonal, fail after 5 consecutive singular matrices.
    rates_at(p, s, s.T, false);
    int Nconsecutive = 0;
    for (;;) {
      const double ghinv = 1.0 / (Dir * s.H * m.Gamma[0]);
      const int ising = Mem::jac_ghimj(Y, F, R, ghinv, G);
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
    // FP64: KppDecomp in place.  MIX: factor an FP32 copy and keep the FP64
    // matrix for the residuals of the refinement sweeps.
    bool lu64 = !MIX;
    if constexpr (MIX) {
      AccF G32 = at32(L::G32);
      for (int i = 0; i < M::LU_NONZERO; i++) G32[i] = (float)G[i];
      call_decomp32<M>(G32.b);
    } else {
      mem.decomp(G);
    }

    const double H = s.H;
    auto stage = [&](auto stage_no) __attribute__((always_inline)) {
      constexpr int istage = decltype(stage_no)::value;
      A Ki = at(L::K + (istage - 1) * NVAR);
      // Fcn of this stage: output of the latest stage q<=istage with NewF,
      // or Fcn0 (stage 1 copies Fcn0; int/rosenbrock.c:547-560)
      bool from_fcn0 = true;
#pragma unroll
      for (int q = 2; q <= istage; q++)
        if (m.NewF[q - 1]) from_fcn0 = false;
      if (istage > 1 && m.NewF[istage - 1]) {
        vloop<U, NVAR>([&](int i) {
          double y = Y[i];
#pragma unroll
          for (int j = 1; j <= istage - 1; j++) {
            const double a = m.A[(istage - 1) * (istage - 2) / 2 + j - 1];
            if (a != 0.0 || f90) y = y + a * K0[(j - 1) * NVAR + i];
          }
          Ynew[i] = y;
        });
        const double Tau = s.T + m.Alpha[istage - 1] * Dir * H;
        rates_at(p, s, Tau, false);
        fcn(p, Ynew, F, R, Fcn);
        s.Nfun++;
      }
      // the row holding this stage's function values (tiles; in registers the
      // choice is made per value below, so every index stays static)
      A Fc = from_fcn0 ? Fcn0 : Fcn;
      double HC[S > 1 ? S - 1 : 1];
#pragma unroll
      for (int j = 1; j <= istage - 1; j++)
        HC[j - 1] = m.C[(istage - 1) * (istage - 2) / 2 + j - 1] / (Dir * H);
      const double HG = Dir * H * m.Gamma[istage - 1];
      const bool use_g = (!p.autonomous) && (m.Gamma[istage - 1] != 0.0) && !skip_dfdt &&
                         (HG != 0.0 || f90);
      // F90 adds HG*dFdT whenever Gamma_i /= 0; where dF/dT is exactly +0.0
      // (not evaluated) that is HG*(+0.0)
      const bool zero_g = f90 && (!p.autonomous) && (m.Gamma[istage - 1] != 0.0) && skip_dfdt;
      vloop<U, NVAR>([&](int i) {
        double k;
        if constexpr (U) {
          // both rows read, one chosen bit for bit (a select of the values,
          // which the compiler cannot turn back into a select of rows)
          const unsigned long long a0 = __double_as_longlong(Fcn0[i]);
          const unsigned long long a1 = __double_as_longlong(Fcn[i]);
          const unsigned long long sel = from_fcn0 ? ~0ull : 0ull;
          k = __longlong_as_double((long long)((a0 & sel) | (a1 & ~sel)));
        } else {
          k = Fc[i];
        }
#pragma unroll
        for (int j = 1; j <= istage - 1; j++)
          if (HC[j - 1] != 0.0 || f90) k = k + HC[j - 1] * K0[(j - 1) * NVAR + i];
        if (use_g) k = k + HG * dFdT[i];
        if (zero_g) k = k + HG * 0.0;
        Ki[i] = k;
      });
      if constexpr (MIX) {
       if (!lu64) {
        // x1 = FP32 solve of b; r = b - A x1 in FP64; x = x1 + FP32 solve of r
        AccF G32 = at32(L::G32), X32 = at32(L::X32);
        Acc XR = at(L::XR), RR = at(L::RR);
        for (int i = 0; i < NVAR; i++) X32[i] = (float)Ki[i];
        call_solve32<M>(G32.b, X32.b);
        for (int i = 0; i < NVAR; i++) XR[i] = (double)X32[i];
        call_residual<M>(G.b, XR.b, Ki.b, RR.b);
        for (int i = 0; i < NVAR; i++) X32[i] = (float)RR[i];
        call_solve32<M>(G32.b, X32.b);
        // the max of |d| and |x| forgets a NaN (dmax(NaN, y) = y), and an
        // infinite correction passes inf <= 1e-5*inf: non-finite values are
        // flagged element by element instead
        double dmx = 0.0, xmx = 0.0;
        bool bad = false;
        for (int i = 0; i < NVAR; i++) {
          const double d = (double)X32[i], x = XR[i] + d;
          XR[i] = x;
          bad |= !isfinite(d) || !isfinite(x);
          dmx = dmax(dmx, dabs(d));
          xmx = dmax(xmx, dabs(x));
        }
        if (!bad && dmx <= kRefineTol * xmx) {
          for (int i = 0; i < NVAR; i++) Ki[i] = XR[i];
        } else {
          // the sweep did not converge (or produced a non-finite value):
          // this and the remaining stages use the FP64 factorisation
          lu64 = true;
          s.Nfb++;
          mem.decomp(G);
          mem.solve(G, Ki);
        }
       } else {
        mem.solve(G, Ki);
       }
      } else {
        mem.solve(G, Ki);
      }
      s.Nsol++;
    };
    // one copy of the stage per istage (a compile-time constant in each)
    stages(stage, std::make_integer_sequence<int, S>{});
    // new solution, error estimate and ros_ErrorNorm (:582-591, :632-659)
    double Err = 0.0;
    vloop<U, NVAR>([&](int i) {
      const double y = Y[i];
      double yn = y, ye = 0.0;
#pragma unroll
      for (int j = 1; j <= S; j++) {
        const double kj = K0[(j - 1) * NVAR + i];
        if (m.M[j - 1] != 0.0 || f90) yn = yn + m.M[j - 1] * kj;
        if (m.E[j - 1] != 0.0 || f90) ye = ye + m.E[j - 1] * kj;
      }
      Ynew[i] = yn;
      const double Ymax = dmax(dabs(y), dabs(yn));
      const double Scale =
          p.vector_tol ? (p.atol[i] + p.rtol[i] * Ymax) : (p.atol0 + p.rtol0 * Ymax);
      if (!f90) {
        Err = Err + (ye * ye) / (Scale * Scale);
      } else if (!p.np_w || p.np_w[i] != 0.0) {
        // int/rosenbrock.f90:718-728: (Yerr/Scale)**2 over non-passive species
        const double t = ye / Scale;
        Err = Err + t * t;
      }
    });
    Err = sqrt(Err / (double)(f90 ? p.np_count : NVAR));
    if (f90) {
      // ros_ErrorNorm's floor MAX(Err, 1e-10); a NaN estimate ends the cell
      // (the floor would turn it into an accepted step)
      if (Err != Err) {
        finish(p, s, KPP_IERR_NONFINITE, s.T);
        return KPP_IERR_NONFINITE;
      }
      Err = dmax(Err, 1.0e-10);
    }
    // step size control (:594-617), or the H211b filter of
    // int/rosenbrock_h211b_qssa.f90:882-892 (no FacMin/FacMax bounds)
    double Fac;
    if (p.h211b) {
      const double e = 1.0 / (p.bb * p.kk);
      Fac = pow(1.0 / Err, e) * pow(1.0 / s.ErrOld, e) * pow(s.FacOld, -1.0 / p.bb);
      s.FacOld = Fac;
      s.ErrOld = Err;
    } else {
      Fac = dmin(p.FacMax, dmax(p.FacMin, p.FacSafe / pow(Err, 1.0 / m.ELO)));
    }
    double Hnew = H * Fac;
    s.Nstp++;
    if ((Err <= 1.0) || (H <= p.Hmin)) {
      // step log: size of the first accepted step of this call
      if (p.hfirst && s.Nacc == p.stat0[3]) p.hfirst[s.cell] = H;
      s.Nacc++;
      if (p.nonneg)
        vloop<U, NVAR>([&](int i) { Y[i] = dmax(Ynew[i], 0.0); });  // MAX(Ynew, ZERO)
      else
        vloop<U, NVAR>([&](int i) { Y[i] = Ynew[i]; });
      s.T += Dir * H;
      Hnew = dmax(p.Hmin, dmin(Hnew, s.Hmax));
      if (s.RejectLastH) Hnew = dmin(Hnew, H);
      if (f90) {
        s.Hexit = H;
        s.Hnew = Hnew;
        s.Tacc = s.T;
      }
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
      // Bounds of the reject loop, which int/rosenbrock.c:537-619 does not
      // have: a NaN error estimate (a NaN in the cell's inputs, rates or
      // Jacobian) or a step size that is no longer finite would reject
      // forever, and so would step factors of 1 (FacMin = FacRej = 1).
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

// Rate constants a cell starts with (what Cell::start computes), for all
// cells at once, so that a persistent-kernel lane refilled with a new cell
// copies them instead of evaluating exp/pow while its wavefront waits.
template <class M>
__global__ __launch_bounds__(256) void ros_rates_prepass(RosParams p) {
  const int cell = blockIdx.x * blockDim.x + threadIdx.x;
  if (cell >= p.ncells) return;
  double* base = p.rpre + cell;
  const size_t n = (size_t)p.ncells;
  struct S {
    double* b;
    size_t n;
    __device__ __forceinline__ double& operator[](int i) const { return b[(size_t)i * n]; }
  } r{base, n};
  const double TEMP = p.temp[cell], CF = p.cfactor[cell];
  M::rates_fixed(TEMP, CF, r);
  if (p.sun_mode == KPP_SUN_FIXED && !M::RATES_USE_TIME)
    M::rates_sun(TEMP, CF, p.sun ? p.sun[cell] : 0.0, 0.0, r);
}

// Plain kernel: one lane per cell for the whole integration.  The first
// p.lanes lanes of each wavefront carry cells (p.lanes < 64 puts more
// wavefronts in flight for small batches).
template <class M, int S, bool MIX = false>
__global__ __launch_bounds__(64) KPPHIP_KERNEL_ATTR void ros_plain_kernel(RosParams p) {
  using L = Layout<M, S, MIX>;
  if ((int)threadIdx.x >= p.lanes) return;
  const int k = blockIdx.x * p.lanes + threadIdx.x;
  if (k >= p.ncells) return;
  Cell<M, S, MIX> c{p.scratch + (size_t)blockIdx.x * L::ROWS * TILE + threadIdx.x};
  CellState s;
  c.start(p, p.order ? p.order[k] : k, s);
  while (c.attempt(p, s) == 0) {
  }
}

// Register kernel (small mechanisms, fits_registers<M>()): one lane per cell
// for its whole integration like the plain kernel, with the working set in
// the lane's VGPRs (RegMem) instead of a tile in global memory.
template <class M, int S>
__global__ __launch_bounds__(64) KPPHIP_KERNEL_ATTR void ros_reg_kernel(RosParams p) {
  if ((int)threadIdx.x >= p.lanes) return;
  const int k = blockIdx.x * p.lanes + threadIdx.x;
  if (k >= p.ncells) return;
  Cell<M, S, false, RegMem<M, Layout<M, S, false>::ROWS>> c;
  CellState s;
  c.start(p, p.order ? p.order[k] : k, s);
  while (c.attempt(p, s) == 0) {
  }
}

// Persistent work-stealing kernel: a fixed grid of wavefronts; a lane whose
// cell has ended takes the next cell of the queue.  Refills are batched per
// wavefront: when at least refill_min lanes are idle (or none is active),
// one lane claims as many queue entries as there are idle lanes with a
// single atomic and the idle lanes take consecutive entries.  A wavefront
// leaves once the queue is exhausted and none of its lanes is active, so
// every wavefront reaches the exit.  Lanes at or above p.lanes never work.
//
// NL > 0 (FP64 only): the factored values of the NL most re-read LU entries
// of each lane are kept in LDS (dynamic shared memory of NL*TILE doubles per
// one-wavefront workgroup) instead of the tile; see KppDecompL.
template <class M, int S, bool MIX = false, int NL = 0>
__global__ __launch_bounds__(64) KPPHIP_KERNEL_ATTR void ros_persistent_kernel(RosParams p) {
  static_assert(NL == 0 || (!MIX && NL <= M::LDS_MAX), "LDS mirror: FP64, NL <= LDS_MAX");
  using L = Layout<M, S, MIX>;
  const int lane = threadIdx.x;
  extern __shared__ double kpp_lds_mirror[];
  Cell<M, S, MIX, TileMem<M, NL>> c{{p.scratch + (size_t)blockIdx.x * L::ROWS * TILE + lane,
                                     (ldouble*)(kpp_lds_mirror + lane)}};
  CellState s;
  const bool worker = lane < p.lanes;
  bool active = false;
  bool exhausted = false;  // wavefront-uniform
  for (;;) {
    const uint64_t idle = __ballot(worker && !active);
    const int n_idle = __popcll(idle);
    const int n_active = __popcll(__ballot(active));
    if (!exhausted && n_idle > 0 && (n_idle >= p.refill_min || n_active == 0)) {
      const int first = __ffsll((long long)idle) - 1;
      int base = 0;
      if (lane == first) base = atomicAdd(p.queue, n_idle);
      base = __shfl(base, first);
      if (base + n_idle >= p.ncells) exhausted = true;
      if (worker && !active) {
        const int rank = __popcll(idle & ((1ull << lane) - 1ull));
        const int k = base + rank;
        if (k < p.ncells) {
          c.start(p, p.order ? p.order[k] : k, s);
          active = true;
        }
      }
    }
    if (__ballot(active) == 0) break;
    if (active && c.attempt(p, s) != 0) active = false;
  }
}

}  // namespace kpphip
