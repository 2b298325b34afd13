"""ctypes wrapper of the CPU oracle (oracle/ros_ref.c).

TEST INFRASTRUCTURE ONLY: used by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py as the checker or the reported CPU baseline.
The libraries are built by kpp_amd.build.build_oracle() into oracle/build/.
"""
from __future__ import annotations

import ctypes
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
BUILD = HERE / "build"
_libs = {}


def build_native(mech: str, out_dir) -> Path:
    """The timing build of the oracle for this host's CPU (gcc -O2
    -march=native -fopenmp) from oracle/ros_ref.c and the mechanism's
    generated C model, into out_dir; used by the CPU baseline so the timing
    build targets the machine it runs on."""
    import subprocess
    gen = HERE.parent / "kpp_amd" / "_gen"
    out = Path(out_dir) / f"libros_ref_{mech}_native.so"
    subprocess.run(["gcc", "-O2", "-march=native", "-fPIC", "-shared", "-fopenmp", "-w",
                    f'-DKPP_MODEL_HEADER="{mech}_model.h"', "-I", str(gen), str(HERE / "ros_ref.c"),
                    "-o", str(out), "-lm"], check=True, capture_output=True)
    return out


def lib(mech: str, fast: bool = False, path=None) -> ctypes.CDLL:
    """The oracle library of a mechanism: the parity build, the fast build
    (fast=True), or the library at `path` (e.g. build_native's)."""
    key = (mech, fast, str(path) if path else None)
    if key in _libs:
        return _libs[key]
    p = Path(path) if path else BUILD / (f"libros_ref_{mech}_fast.so" if fast else
                                         f"libros_ref_{mech}.so")
    if not p.exists():
        raise FileNotFoundError(f"{p} missing: run kpp_amd.build.build_oracle()")
    L = ctypes.CDLL(str(p))
    vp, ci, cd = ctypes.c_void_p, ctypes.c_int, ctypes.c_double
    L.ros_ref_batch.argtypes = [ci, vp, vp, vp, vp, ci, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp,
                                vp, vp, ci]
    L.ros_ref_update_sun.restype = cd
    L.ros_ref_update_sun.argtypes = [cd]
    L.ros_ref_wlamch.restype = cd
    L.ros_ref_rconst.argtypes = [cd, cd, cd, cd, vp]
    L.ros_ref_rconst_c.argtypes = [cd, cd, cd, cd, vp, vp]
    L.ros_ref_fun.argtypes = [vp, vp, vp, vp]
    L.ros_ref_jac_sp.argtypes = [vp, vp, vp, vp]
    L.ros_ref_decomp.argtypes = [vp]
    L.ros_ref_solve.argtypes = [vp, vp]
    L.ros_ref_prepare.argtypes = [vp, cd, cd, vp]
    L.ros_ref_error_norm.restype = cd
    L.ros_ref_error_norm.argtypes = [vp, vp, vp, vp, vp, ci]
    L.ros_ref_method.argtypes = [ci, vp, vp, vp, vp, vp, vp, vp, vp, vp]
    L.ros_ref_species_name.restype = ctypes.c_char_p
    L.ros_ref_species_name.argtypes = [ci]
    L.ros_ref_lu_pattern.argtypes = [vp, vp, vp]
    L.ros_ref_initialize.argtypes = [vp]
    L.ros_ref_sizes.argtypes = [vp]
    _libs[key] = L
    return L


def _p(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def sizes(mech: str) -> dict:
    out = np.zeros(6, np.int32)
    lib(mech).ros_ref_sizes(_p(out))
    return dict(zip(["NVAR", "NFIX", "NSPEC", "NREACT", "NONZERO", "LU_NONZERO"], map(int, out)))


def integrate(batch, nthreads: int = 0, fast: bool = False, hstart=None, rconst=None,
              IPAR=None, RPAR=None, sun_mode=None, C=None, path=None):
    """Integrate a workloads.Batch cell by cell; returns a dict of results."""
    L = lib(batch.mech, fast, path)
    sz = sizes(batch.mech)
    n = batch.ncells
    Cw = np.ascontiguousarray(batch.C if C is None else C, np.float64).copy()
    ierr = np.zeros(n, np.int32)
    istat = np.zeros((8, n), np.int32)
    rstat = np.zeros((2, n), np.float64)
    hacc1 = np.zeros(n, np.float64)
    f = lambda a: np.ascontiguousarray(a, np.float64)
    ip = np.ascontiguousarray(batch.IPAR if IPAR is None else IPAR, np.int32)
    rp = f(batch.RPAR if RPAR is None else RPAR)
    hs = None if hstart is None else f(hstart)
    rc = None if rconst is None else f(rconst)
    temp, cf, sun, ts, te = f(batch.temp), f(batch.cfactor), f(batch.sun), f(batch.tstart), f(batch.tend)
    atol, rtol = f(batch.atol), f(batch.rtol)
    L.ros_ref_batch(n, _p(Cw), _p(temp), _p(cf), _p(sun),
                    batch.sun_mode if sun_mode is None else sun_mode, _p(ts), _p(te), _p(atol),
                    _p(rtol), _p(ip), _p(rp), _p(hs), _p(rc), _p(ierr), _p(istat), _p(rstat),
                    _p(hacc1), nthreads)
    return {"C": Cw, "ierr": ierr, "istat": istat, "rstat": rstat, "hacc1": hacc1,
            "nvar": sz["NVAR"]}


def method(code: int):
    L = lib_any()
    S = ctypes.c_int()
    A = np.zeros(15); Cc = np.zeros(15); M = np.zeros(6); E = np.zeros(6)
    Al = np.zeros(6); G = np.zeros(6); ELO = ctypes.c_double(); NewF = np.zeros(6, np.int32)
    ok = L.ros_ref_method(code, ctypes.byref(S), _p(A), _p(Cc), _p(M), _p(E), _p(Al), _p(G),
                          ctypes.byref(ELO), _p(NewF))
    if not ok:
        return None
    s = S.value
    return {"S": s, "A": A, "C": Cc, "M": M[:s], "E": E[:s], "Alpha": Al[:s], "Gamma": G[:s],
            "ELO": ELO.value, "NewF": NewF[:s]}


def lib_any():
    for p in sorted(BUILD.glob("libros_ref_*.so")):
        if not p.name.endswith("_fast.so"):
            return lib(p.name[len("libros_ref_"):-3])
    raise FileNotFoundError("no oracle library built")
