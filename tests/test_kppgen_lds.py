"""Routines of the generated HIP model header compiled for the host: the
LDS-mirror variants of the LU routines (KppDecompL<NL>/KppSolveL<NL>) and the
reaction-major Fun.

They are KppDecomp/KppSolve with the factored values of the NL most re-read
LU entries kept in a second array (the lane's LDS slots on the GPU); the bar
is bitwise equality with KppDecomp/KppSolve for every mirror size up to
LDS_MAX, on random diagonally dominant matrices of the saprc99 and saprcnov
patterns: the triangular solves read every LU entry, so any entry stored
to or read from the wrong place shows in X.

Fun is emitted reaction by reaction (each product term added to the running
sums of the species it feeds right after it is formed); the bar is bitwise
equality with the oracle's Fun, which evaluates each species' sum as one
left-associated chain in reaction order (KPP's form).
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
using std::exp; using std::pow; using std::log; using std::sqrt;
"""

PROG = """#include "%(mech)s_hip.h"
struct HA {
  using value_type = double;
  double* b;
  double& operator[](int i) const { return b[i]; }
};
using M = kppm_%(mech)s;
template <int NL>
static void run(double* G, double* Gs, double* X) {
  M::KppDecompL<NL>(HA{G}, HA{Gs});
  M::KppSolveL<NL>(HA{G}, HA{X}, HA{Gs});
}
extern "C" int lds_max() { return M::LDS_MAX; }
extern "C" void hip_fun(double* V, double* F, double* R, double* out) {
  M::Fun(HA{V}, HA{F}, HA{R}, HA{out});
}
extern "C" void ref_lu(double* G, double* X) {
  M::KppDecomp(HA{G});
  M::KppSolve(HA{G}, HA{X});
}
extern "C" int split_lu(int nl, double* G, double* Gs, double* X) {
  switch (nl) {
    case 0: run<0>(G, Gs, X); return 0;
    case 1: run<1>(G, Gs, X); return 0;
    case 7: run<7>(G, Gs, X); return 0;
    case 20: run<20>(G, Gs, X); return 0;
    case 24: run<24>(G, Gs, X); return 0;
  }
  return -1;
}
"""


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


@pytest.fixture(scope="module", params=["saprc99", "saprcnov"])
def host_lu(request, tmp_path_factory):
    mech = request.param
    hdr = GEN / f"{mech}_hip.h"
    if not hdr.exists():
        pytest.skip(f"{hdr} not generated (run the build first)")
    d = tmp_path_factory.mktemp(f"lds_{mech}")
    (d / "hip").mkdir()
    (d / "hip" / "hip_runtime.h").write_text(SHIM)
    (d / "prog.cpp").write_text(PROG % {"mech": mech})
    so = d / "liblds.so"
    subprocess.run(["g++", "-std=c++17", "-O1", "-ffp-contract=off", "-shared", "-fPIC",
                    "-I", str(d), "-I", str(GEN), str(d / "prog.cpp"), "-o", str(so)],
                   check=True)
    L = ctypes.CDLL(str(so))
    for fn in (L.ref_lu, L.split_lu):
        fn.restype = ctypes.c_int if fn is L.split_lu else None
    return mech, L


def _matrix(mech, rng):
    s = ref.sizes(mech)
    nv = s["NVAR"]
    icol = np.zeros(s["LU_NONZERO"], np.int32)
    crow = np.zeros(nv + 1, np.int32)
    diag = np.zeros(nv + 1, np.int32)
    ref.lib(mech).ros_ref_lu_pattern(_p(icol), _p(crow), _p(diag))
    G = rng.uniform(-1.0, 1.0, s["LU_NONZERO"])
    for i in range(nv):
        G[diag[i]] = 2.0 + np.abs(G[crow[i]:crow[i + 1]]).sum()
    return G, rng.uniform(-1.0, 1.0, nv)


def test_split_lu_bitwise(host_lu):
    mech, L = host_lu
    assert L.lds_max() == 24
    rng = np.random.default_rng(5)
    for _ in range(4):
        G0, X0 = _matrix(mech, rng)
        Gr, Xr = G0.copy(), X0.copy()
        L.ref_lu(_p(Gr), _p(Xr))
        for nl in (0, 1, 7, 20, 24):
            G, X = G0.copy(), X0.copy()
            Gs = np.full(64, np.nan)
            assert L.split_lu(nl, _p(G), _p(Gs), _p(X)) == 0
            assert np.array_equal(X, Xr), nl
            # entries not mirrored hold their factored values in G; the
            # mirrored ones keep the unfactored value there
            moved = np.nonzero(G != Gr)[0]
            assert len(moved) <= nl
            assert set(Gr[moved]) <= set(Gs[:nl])
            assert np.isnan(Gs[nl:]).all()


def test_reaction_major_fun_bitwise(host_lu):
    mech, L = host_lu
    s = ref.sizes(mech)
    nv, nf, nr = s["NVAR"], s["NFIX"], s["NREACT"]
    R = ref.lib(mech)
    rng = np.random.default_rng(13)
    for _ in range(5):
        # magnitudes spread over many decades, as in a chemistry state
        V = 10.0 ** rng.uniform(-8, 12, nv)
        F = 10.0 ** rng.uniform(8, 19, max(nf, 1))
        K = 10.0 ** rng.uniform(-20, 2, nr)
        a, b = np.zeros(nv), np.zeros(nv)
        L.hip_fun(_p(V), _p(F), _p(K), _p(a))
        R.ros_ref_fun(_p(V), _p(F), _p(K), _p(b))
        assert np.array_equal(a, b)



# ---- the memory-pipelined LU (kppgen --pipelined; not used by the integrator)
PF_PROG = """#include "%(mech)s_hip.h"
struct HA {
  using value_type = double;
  double* b;
  double& operator[](int i) const { return b[i]; }
};
using M = kppm_%(mech)s;
extern "C" void ref_lu(double* G, double* X) { M::KppDecomp(HA{G}); M::KppSolve(HA{G}, HA{X}); }
extern "C" void pf_lu(double* G, double* X) { M::KppDecompPF(HA{G}); M::KppSolvePF(HA{G}, HA{X}); }
"""


@pytest.mark.parametrize("mech", ["saprc99", "saprcnov"])
def test_pipelined_lu_bitwise_on_reused_buffer(tmp_path, mech):
    """KppDecompPF/KppSolvePF against KppDecomp/KppSolve, bitwise, on a
    sequence of matrices written into one reused buffer, every entry
    overwritten before each factorisation as JacGhimj does in a persistent
    lane's tile.  (They pass; the stall seen with them on the GPU is not an
    arithmetic or ordering error of the routines that a host build shows.)"""
    from kpp_amd import build
    exe = build.OBJ / "kppgen"
    if not exe.exists() or not (build.kpp_home() / "models" / f"{mech}.def").exists():
        pytest.skip("kppgen or the KPP models not available")
    gen = tmp_path / "gen"
    gen.mkdir()
    subprocess.run([str(exe), "--kpp-home", str(build.kpp_home()), "--root", mech, "-o", str(gen),
                    "--pipelined", str(build.kpp_home() / "models" / f"{mech}.def")],
                   check=True, capture_output=True)
    (tmp_path / "hip").mkdir()
    (tmp_path / "hip" / "hip_runtime.h").write_text(SHIM)
    (tmp_path / "prog.cpp").write_text(PF_PROG % {"mech": mech})
    so = tmp_path / "libpf.so"
    subprocess.run(["g++", "-std=c++17", "-O1", "-ffp-contract=off", "-shared", "-fPIC", "-w",
                    "-I", str(tmp_path), "-I", str(gen), str(tmp_path / "prog.cpp"), "-o", str(so)],
                   check=True)
    L = ctypes.CDLL(str(so))
    rng = np.random.default_rng(17)
    Gbuf = np.zeros(ref.sizes(mech)["LU_NONZERO"])
    Xbuf = np.zeros(ref.sizes(mech)["NVAR"])
    for _ in range(6):
        G0, X0 = _matrix(mech, rng)
        Gr, Xr = G0.copy(), X0.copy()
        L.ref_lu(_p(Gr), _p(Xr))
        Gbuf[:] = G0   # the same buffer every time, holding the previous factors before
        Xbuf[:] = X0
        L.pf_lu(_p(Gbuf), _p(Xbuf))
        assert np.array_equal(Gbuf, Gr)
        assert np.array_equal(Xbuf, Xr)
