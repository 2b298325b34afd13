"""The mechanism compiler's cooperative programs (kppgen --coop), run on the
host by a lockstep emulation of the device interpreter (kpp_amd/csrc/
ros_coop.h): 64 lanes, and at every step all lanes first read the operands
of the next step, then compute the current step and store.  A schedule that
read a value in the step after it is stored, or a table entry naming a
wrong operand, changes the results.

Bar: bitwise equality with the generated straight-line routines (Fun,
JacGhimj, KppDecomp, KppSolve) of the same header, which the oracle's
own model functions match bit for bit (tests/test_kppgen_lds.py, the GPU
parity tests), on random states of saprc99, saprcnov, small_strato and MCM.
"""
import ctypes
import subprocess
from pathlib import Path

import numpy as np
import pytest

from oracle import ref

GEN = Path(__file__).resolve().parents[1] / "kpp_amd" / "_gen"

SHIM = """#pragma once
#include <cmath>
#define __device__
#define __forceinline__ inline
#define __noinline__
#define __builtin_amdgcn_sched_barrier(x) ((void)0)
using std::exp; using std::pow; using std::log; using std::log10; using std::sqrt; using std::cos;
using std::acos; using std::fabs;
"""

PROG = r"""#include "%(mech)s_hip.h"
#include <cstdint>
#include <cstdlib>
#include <cstring>
struct HA {
  using value_type = double;
  double* b;
  double& operator[](int i) const { return b[i]; }
};
using M = kppm_%(mech)s;
using C = kppm_%(mech)s_coop;
static const unsigned long long IDLE = 0;  // idle (lane, step); real steps have bit 52
extern "C" int enabled() { return C::ENABLED ? 1 : 0; }

// ---- reference: the straight-line routines ----
extern "C" void ref_fun(double* V, double* F, double* R, double* out) {
  M::Fun(HA{V}, HA{F}, HA{R}, HA{out});
}
extern "C" int ref_jac(double* V, double* F, double* R, double ghinv, double* G) {
  return M::JacGhimj(HA{V}, HA{F}, HA{R}, ghinv, HA{G});
}
extern "C" void ref_lu(double* G, double* X) {
  M::KppDecomp(HA{G});
  M::KppSolve(HA{G}, HA{X});
}

#if %(enabled)d
static double factor(unsigned code, const double* V, const double* F, const double* R) {
  const unsigned idx = code & 0x3FFFFFFFu;
  switch (code >> 30) {
    case 0: return R[idx];
    case 1: return V[idx];
    case 2: return F[idx];
    default: return C::CONSTS()[idx];
  }
}
static void products(const unsigned* P, int N, int NPAD, const double* V, const double* F,
                     const double* R, double* T) {
  for (int q = 0; q < N; q++) {
    double p = factor(P[q], V, F, R);
    for (int f = 1; f < C::MAXF; f++) {
      const unsigned code = P[f * NPAD + q];
      if (code == 0xFFFFFFFFu) break;
      p = p * factor(code, V, F, R);
    }
    T[q] = p;
  }
}
// sum chains in lockstep; the coefficient comes from the per-step table
// (which must agree with the constant pool index of the descriptor)
template <class Store>
static void sums(const unsigned long long* prog, const double* coef, int steps, const double* T,
                 Store store) {
  double acc[64] = {0};
  for (int t = 0; t < steps; t++)
    for (int l = 0; l < 64; l++) {
      const unsigned long long d = prog[t * 64 + l];
      if (d == IDLE) continue;
      if (!((d >> 52) & 1)) abort();
      const double c = coef[t * 64 + l];
      if (!(c == C::CONSTS()[d & 0xFFFF]) && !((d >> 50) & 1)) abort();
      const double term = ((d >> 50) & 1) ? 0.0 : c * T[(d >> 16) & 0xFFFF];
      acc[l] = ((d >> 48) & 1) ? term : acc[l] + term;
      if ((d >> 49) & 1) store(d, acc[l]);
    }
  // the padding past the last step is idle
  for (int t = steps; t < steps + C::PAD; t++)
    for (int l = 0; l < 64; l++)
      if (prog[t * 64 + l] != IDLE) abort();
}
extern "C" void coop_fun(double* V, double* F, double* R, double* out) {
  static double T[C::NA > 0 ? C::NA : 1];
  products(C::APROD(), C::NA, C::NA_PAD, V, F, R, T);
  sums(C::FUN(), C::FUN_COEF(), C::FUN_STEPS, T,
       [&](unsigned long long d, double a) { out[(d >> 32) & 0xFFFF] = a; });
}
extern "C" void coop_jac(double* V, double* F, double* R, double ghinv, double* G) {
  static double T[C::NB > 0 ? C::NB : 1];
  products(C::BPROD(), C::NB, C::NB_PAD, V, F, R, T);
  sums(C::JAC(), C::JAC_COEF(), C::JAC_STEPS, T, [&](unsigned long long d, double a) {
    double g = -(a);
    if ((d >> 51) & 1) g = g + ghinv;
    G[(d >> 32) & 0xFFFF] = g;
  });
}
// elimination program in lockstep: at step t every lane reads the operands
// of step t+1, then computes step t and stores
static void program(const unsigned long long* prog, int steps, double* G, double* X, bool solve,
                    int nvar) {
  auto init_of = [&](unsigned t) { return (solve && t >= (unsigned)nvar) ? t - nvar : t; };
  double acc[64] = {0}, ga[64] = {0}, xb[64] = {0}, xt[64] = {0};
  unsigned long long d[64];
  // every lane reads its three operands each step, idle ones entry 0
  auto load = [&](int t, int l, double& a, double& b, double& c) {
    const unsigned long long e = prog[t * 64 + l];
    a = G[e & 0xFFFF];
    b = X[(e >> 16) & 0xFFFF];
    c = X[init_of((e >> 32) & 0xFFFF)];
    return e;
  };
  for (int l = 0; l < 64; l++) d[l] = load(0, l, ga[l], xb[l], xt[l]);
  for (int t = 0; t < steps; t++) {
    unsigned long long dn[64];
    double gan[64] = {0}, xbn[64] = {0}, xtn[64] = {0};
    for (int l = 0; l < 64; l++) dn[l] = load(t + 1, l, gan[l], xbn[l], xtn[l]);
    for (int l = 0; l < 64; l++) {
      if (d[l] != IDLE && !((d[l] >> 52) & 1)) abort();
      if ((d[l] >> 48) & 1) acc[l] = xt[l];
      acc[l] = ((d[l] >> 50) & 1) ? acc[l] / ga[l] : acc[l] - ga[l] * xb[l];
      if ((d[l] >> 49) & 1) X[(d[l] >> 32) & 0xFFFF] = acc[l];
    }
    for (int l = 0; l < 64; l++) {
      d[l] = dn[l]; ga[l] = gan[l]; xb[l] = xbn[l]; xt[l] = xtn[l];
    }
  }
}
extern "C" void coop_lu(double* G, double* X) {
  static double T[2 * M::NVAR];
  program(C::LU(), C::LU_STEPS, G, G, false, M::NVAR);
  for (int i = 0; i < M::NVAR; i++) T[i] = X[i];
  program(C::SOL(), C::SOL_STEPS, G, T, true, M::NVAR);
  for (int i = 0; i < M::NVAR; i++) X[i] = T[M::NVAR + i];
}
#else
extern "C" void coop_fun(double*, double*, double*, double*) {}
extern "C" void coop_jac(double*, double*, double*, double, double*) {}
extern "C" void coop_lu(double*, double*) {}
#endif
"""

MECHS = ["small_strato", "saprc99", "saprcnov", "mcm"]


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


@pytest.fixture(scope="module", params=MECHS)
def coop(request, tmp_path_factory):
    mech = request.param
    hdr = GEN / f"{mech}_hip.h"
    if not hdr.exists():
        pytest.skip(f"{hdr} not generated (run the build first)")
    enabled = int(f"struct kppm_{mech}_coop {{\nstatic constexpr bool ENABLED = true;"
                  in hdr.read_text())
    d = tmp_path_factory.mktemp(f"coop_{mech}")
    (d / "hip").mkdir()
    (d / "hip" / "hip_runtime.h").write_text(SHIM)
    (d / "prog.cpp").write_text(PROG % {"mech": mech, "enabled": enabled})
    so = d / "libcoop.so"
    subprocess.run(["g++", "-std=c++17", "-O1", "-ffp-contract=off", "-shared", "-fPIC", "-w",
                    "-I", str(d), "-I", str(GEN), str(d / "prog.cpp"), "-o", str(so)],
                   check=True)
    L = ctypes.CDLL(str(so))
    L.ref_jac.restype = ctypes.c_int
    L.ref_jac.argtypes = [ctypes.c_void_p] * 3 + [ctypes.c_double, ctypes.c_void_p]
    L.coop_jac.argtypes = L.ref_jac.argtypes
    return mech, L


def test_coop_programs_present(coop):
    mech, L = coop
    assert L.enabled() == 1, f"no cooperative programs for {mech}"


def test_coop_fun_and_jac_bitwise(coop):
    mech, L = coop
    s = ref.sizes(mech)
    nv, nf, nr, nnz = s["NVAR"], s["NFIX"], s["NREACT"], s["LU_NONZERO"]
    rng = np.random.default_rng(21)
    for _ in range(3):
        V = 10.0 ** rng.uniform(-8, 12, nv)
        F = 10.0 ** rng.uniform(8, 19, max(nf, 1))
        R = 10.0 ** rng.uniform(-20, 2, nr)
        a, b = np.zeros(nv), np.full(nv, np.nan)
        L.ref_fun(_p(V), _p(F), _p(R), _p(a))
        L.coop_fun(_p(V), _p(F), _p(R), _p(b))
        assert np.array_equal(a, b)
        ghinv = float(rng.uniform(1.0, 1e3))
        g1, g2 = np.zeros(nnz), np.full(nnz, np.nan)
        L.ref_jac(_p(V), _p(F), _p(R), ghinv, _p(g1))
        L.coop_jac(_p(V), _p(F), _p(R), ghinv, _p(g2))
        assert np.array_equal(g1, g2) and np.array_equal(np.signbit(g1), np.signbit(g2))


def test_coop_lu_and_solve_bitwise(coop):
    mech, L = coop
    s = ref.sizes(mech)
    nv = s["NVAR"]
    icol = np.zeros(s["LU_NONZERO"], np.int32)
    crow = np.zeros(nv + 1, np.int32)
    diag = np.zeros(nv + 1, np.int32)
    ref.lib(mech).ros_ref_lu_pattern(_p(icol), _p(crow), _p(diag))
    rng = np.random.default_rng(5)
    for _ in range(3):
        G = rng.uniform(-1.0, 1.0, s["LU_NONZERO"])
        for i in range(nv):
            G[diag[i]] = 2.0 + np.abs(G[crow[i]:crow[i + 1]]).sum()
        X = rng.uniform(-1.0, 1.0, nv)
        Gr, Xr = G.copy(), X.copy()
        L.ref_lu(_p(Gr), _p(Xr))
        Gc, Xc = G.copy(), X.copy()
        L.coop_lu(_p(Gc), _p(Xc))
        assert np.array_equal(Gr, Gc)
        assert np.array_equal(Xr, Xc)
