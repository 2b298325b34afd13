"""The mechanism compiler's cooperative programs (kppgen --coop), run on the
host by a lockstep emulation of the device interpreters (kpp_amd/csrc/
ros_coop.h) over the cell's LDS image (byte addresses, G at 0, the T region
at GZ): 64 lanes; the sum programs read each step's product two steps
ahead, the elimination programs (two terms of a chain per step) read the
operands of step t+1 before step t stores; every step stores, idle steps included, and the device's whole
rounds of 8 steps run into the idle padding.  The LDS starts as NaN except
the +0.0 slot, which must stay +0.0.  A schedule that read a value before it
is final, an idle step that changed a value, or a table entry naming a wrong
operand changes the results.

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
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <vector>
struct HA {
  using value_type = double;
  double* b;
  double& operator[](int i) const { return b[i]; }
};
using M = kppm_%(mech)s;
using C = kppm_%(mech)s_coop;
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
// the cell's LDS as the device sees it: byte addresses, G at 0, T at GZ
static std::vector<double> lds;
static double& at(unsigned addr) {
  if (addr %% 8 || addr / 8 >= lds.size()) abort();
  return lds[addr / 8];
}
static constexpr int RING = 8;
static int rounds(int steps) { return (steps + RING - 1) / RING * RING; }
static void fresh() {
  // everything NaN (a read of memory no step should read shows in the
  // results) except the ZERO slot
  lds.assign(M::LU_NONZERO + C::TSIZE, std::nan(""));
  at(C::GZ + C::ZSLOT) = 0.0;
}
static void zero_slot_intact() {
  if (!(at(C::GZ + C::ZSLOT) == 0.0) || std::signbit(at(C::GZ + C::ZSLOT))) abort();
}
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
                     const double* R) {
  for (int q = 0; q < N; q++) {
    double p = factor(P[q], V, F, R);
    for (int f = 1; f < C::MAXF; f++) {
      const unsigned code = P[f * NPAD + q];
      if (code == 0xFFFFFFFFu) break;
      p = p * factor(code, V, F, R);
    }
    at(C::GZ + 16 + 8 * q) = p;
  }
}
// a sum program in lockstep (two terms per step), as the device runs it:
// whole rounds of RING steps, each step's products read two steps ahead
// (before the stores of the two steps in between), every step storing
static void sums(const unsigned long long* prog, const double* coef, const double* coef2,
                 int steps, bool jvs) {
  const unsigned sb = jvs ? 0u : C::GZ;
  const int n = rounds(steps);
  if (n + 2 > steps + C::PAD) abort();
  double acc[64] = {0}, x0[64], x1[64], x2[64], y0[64], y1[64], y2[64];
  auto prod = [&](int t, int l) { return at(C::GZ + ((unsigned)prog[t * 64 + l] & 0xFFFFu)); };
  auto prod2 = [&](int t, int l) {
    return at(C::GZ + ((unsigned)(prog[t * 64 + l] >> 32) & 0xFFFFu));
  };
  for (int l = 0; l < 64; l++) {
    x0[l] = prod(0, l); x1[l] = prod(1, l); y0[l] = prod2(0, l); y1[l] = prod2(1, l);
  }
  for (int t = 0; t < n; t++) {
    for (int l = 0; l < 64; l++) { x2[l] = prod(t + 2, l); y2[l] = prod2(t + 2, l); }
    for (int l = 0; l < 64; l++) {
      const unsigned long long d = prog[t * 64 + l];
      const double term = coef[t * 64 + l] * x0[l];
      const double sum = acc[l] + term;
      const double v1 = ((long long)d < 0) ? term : sum;
      acc[l] = v1 + coef2[t * 64 + l] * y0[l];
      at(sb + ((unsigned)d >> 16)) = jvs ? -acc[l] : acc[l];
      x0[l] = x1[l]; x1[l] = x2[l]; y0[l] = y1[l]; y1[l] = y2[l];
    }
  }
  zero_slot_intact();
}
extern "C" void coop_fun(double* V, double* F, double* R, double* out) {
  fresh();
  products(C::APROD(), C::NA, C::NA_PAD, V, F, R);
  sums(C::FUN(), C::FUN_COEF(), C::FUN_COEF2(), C::FUN_STEPS, false);
  for (int i = 0; i < M::NVAR; i++) out[i] = at(C::GZ + C::FO + 8 * i);
}
extern "C" void coop_jac(double* V, double* F, double* R, double ghinv, double* G) {
  fresh();
  products(C::BPROD(), C::NB, C::NB_PAD, V, F, R);
  sums(C::JAC(), C::JAC_COEF(), C::JAC_COEF2(), C::JAC_STEPS, true);
  for (int i = 0; i < M::NVAR; i++) at(8 * C::DIAG()[i]) = at(8 * C::DIAG()[i]) + ghinv;
  for (int n = 0; n < M::LU_NONZERO; n++) G[n] = at(8 * C::GSLOT()[n]);
}
// an elimination program in lockstep (two terms per step): at step t every
// lane reads the operands of step t+1, then computes step t and stores
static void program(const unsigned long long* prog, const unsigned long long* progh, int steps,
                    bool solve) {
  const unsigned xb = solve ? C::GZ : 0u, noinit = solve ? C::JSLOT : C::GZ + C::JSLOT;
  const int n = rounds(steps);
  if (n + 1 > steps + C::PAD) abort();
  struct Ops { double ga, xb, ga2, xb2, xi; };
  double acc[64] = {0};
  Ops cur[64], nxt[64];
  auto load = [&](int t, int l) {
    const unsigned long long e = prog[t * 64 + l];
    const unsigned lo = (unsigned)e, lo2 = (unsigned)(e >> 32), hi = (unsigned)progh[t * 64 + l];
    return Ops{at(lo & 0xFFFFu), at(xb + (lo >> 16)), at(lo2 & 0xFFFFu), at(xb + (lo2 >> 16)),
               at(xb + (hi >> 16))};
  };
  for (int l = 0; l < 64; l++) cur[l] = load(0, l);
  for (int t = 0; t < n; t++) {
    for (int l = 0; l < 64; l++) nxt[l] = load(t + 1, l);
    const unsigned step_div = (unsigned)progh[t * 64] & 2u;
    for (int l = 0; l < 64; l++) {
      const unsigned hi = (unsigned)progh[t * 64 + l];
      if (((hi & 2u) != step_div) || ((hi & 5u) && !step_div)) abort();
      // the device's test of a chain's first step, and what it stands for
      const bool first = solve ? (hi >= 0x10000u) : (hi < (C::GZ << 16));
      if (first != ((hi >> 16) != noinit)) abort();
      const Ops& o = cur[l];
      const double v0 = first ? o.xi : acc[l];
      const double v1 = (hi & 4u) ? v0 / o.ga : v0 - o.ga * o.xb;
      const double v = (hi & 1u) ? v1 / o.ga2 : v1 - o.ga2 * o.xb2;
      at(xb + (hi & 0xFFF8u)) = v;
      acc[l] = v;
    }
    for (int l = 0; l < 64; l++) cur[l] = nxt[l];
  }
  zero_slot_intact();
}
extern "C" void coop_lu(double* G, double* X) {
  fresh();
  for (int n = 0; n < M::LU_NONZERO; n++) at(8 * C::GSLOT()[n]) = G[n];
  program(C::LU(), C::LUH(), C::LU_STEPS, false);
  for (int n = 0; n < M::LU_NONZERO; n++) G[n] = at(8 * C::GSLOT()[n]);
  for (int i = 0; i < M::NVAR; i++) at(C::GZ + 16 + 8 * i) = X[i];
  program(C::SOL(), C::SOLH(), C::SOL_STEPS, true);
  for (int i = 0; i < M::NVAR; i++) X[i] = at(C::GZ + 16 + 8 * (M::NVAR + i));
}
#else
extern "C" void coop_fun(double*, double*, double*, double*) {}
extern "C" void coop_jac(double*, double*, double*, double, double*) {}
extern "C" void coop_lu(double*, double*) {}
#endif
"""

MECHS = ["small_strato", "saprc99", "saprcnov", "linear6", "mcm"]


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
