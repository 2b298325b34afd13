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
    L.ros_ref_method_f90.argtypes = [ci, vp, vp, vp, vp, vp, vp, vp, vp, vp]
    L.ros_ref_f90_batch.argtypes = [ci, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp,
                                    vp, ci, ci]
    L.ros_ref_f90_merge.argtypes = [vp, vp, vp, vp]
    L.sdirk_ref_batch.argtypes = [ci, vp, vp, vp, vp, ci, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp,
                                  ci]
    L.sdirk_ref_method.argtypes = [ci, vp, vp, vp, vp, vp, vp, vp, vp]
    L.rk_ref_batch.argtypes = [ci, vp, vp, vp, vp, ci, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp,
                               ci]
    L.rk_ref_method.argtypes = [ci, ci, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp]
    L.rk_ref_decomp_cmplx.argtypes = [vp, vp]
    if hasattr(L, "ros_ref_fun_split"):
        L.ros_ref_fun_split.argtypes = [vp, vp, vp, vp]
    L.rk_ref_solve_cmplx.argtypes = [vp, vp, vp, vp]
    if hasattr(L, "adj_ref_batch"):  # mechanisms with the adjoint's model functions
        L.adj_ref_batch.argtypes = [ci, vp, vp, vp, vp, ci, vp, vp, vp, vp, vp, vp, vp, ci, vp,
                                    ci, ci, vp, vp, vp, ci]
        L.adj_ref_hessian.argtypes = [vp, vp, vp, vp, vp, vp, vp]
        L.adj_ref_hesstr_vec.argtypes = [vp, vp, vp, vp]
        L.adj_ref_jactr_vec.argtypes = [vp, vp, vp]
        L.adj_ref_jac_vec.argtypes = [vp, vp, vp]
        L.adj_ref_solve_tr.argtypes = [vp, vp, vp]
    L.ros_ref_species_name.restype = ctypes.c_char_p
    L.ros_ref_species_name.argtypes = [ci]
    L.ros_ref_lu_pattern.argtypes = [vp, vp, vp]
    L.ros_ref_initialize.argtypes = [vp]
    L.ros_ref_sizes.argtypes = [vp]
    _libs[key] = L
    return L


def _p(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def sizes(mech: str, path=None) -> dict:
    out = np.zeros(6, np.int32)
    lib(mech, path=path).ros_ref_sizes(_p(out))
    return dict(zip(["NVAR", "NFIX", "NSPEC", "NREACT", "NONZERO", "LU_NONZERO"], map(int, out)))


def species_name(mech: str, k: int) -> str:
    """the name of species k (KPP order) of the mechanism"""
    return lib(mech).ros_ref_species_name(k).decode()


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


def integrate_f90(batch, ICNTRL_U=None, RCNTRL_U=None, passive=None, rconst=None,
                  nthreads: int = 0, fast: bool = False, C=None, tin=None, tout=None,
                  h211b: bool = False):
    """The F90 INTEGRATE(TIN, TOUT, ICNTRL_U, RCNTRL_U, ISTATUS_U, RSTATUS_U,
    IERR_U) of int/rosenbrock.f90 cell by cell on a workloads.Batch (its
    tstart/tend unless tin/tout are given; its atol/rtol, whose length picks
    vector or scalar tolerances only through ICNTRL(2)).  Returns C, ierr,
    istatus [8, n] and rstatus [3, n] (Texit, Hexit, Hnew).  h211b=True: the
    INTEGRATE of int/rosenbrock_h211b_qssa.f90 (H211b control with
    RCNTRL(15), RCNTRL(16) = bb, kk; no passive species)."""
    if h211b and passive is not None:
        raise ValueError("the H211b integrator has no passive species")
    L = lib(batch.mech, fast)
    sz = sizes(batch.mech)
    n = batch.ncells
    f = lambda a: np.ascontiguousarray(a, np.float64)
    Cw = f(batch.C if C is None else C).copy()
    ierr = np.zeros(n, np.int32)
    ist = np.zeros((8, n), np.int32)
    rst = np.zeros((3, n), np.float64)
    ic = None if ICNTRL_U is None else np.ascontiguousarray(ICNTRL_U, np.int32)
    rc_ = None if RCNTRL_U is None else f(RCNTRL_U)
    pas = None if passive is None else np.ascontiguousarray(passive, np.int32)
    rcon = None if rconst is None else f(rconst)
    t0 = f(batch.tstart if tin is None else np.broadcast_to(tin, (n,)))
    t1 = f(batch.tend if tout is None else np.broadcast_to(tout, (n,)))
    temp, cf, sun = f(batch.temp), f(batch.cfactor), f(batch.sun)
    atol, rtol = f(batch.atol), f(batch.rtol)
    L.ros_ref_f90_batch(n, _p(Cw), _p(temp), _p(cf), _p(sun), _p(t0), _p(t1), _p(atol),
                        _p(rtol), _p(ic), _p(rc_), _p(pas), _p(rcon), _p(ierr), _p(ist), _p(rst),
                        1 if h211b else 0, nthreads)
    return {"C": Cw, "ierr": ierr, "istat": ist, "rstat": rst, "nvar": sz["NVAR"]}


def integrate_sdirk(batch, ICNTRL=None, RCNTRL=None, rconst=None, nthreads: int = 0,
                    fast: bool = False, C=None, sun_mode=None):
    """KPP's C SDIRK (int/sdirk.c) cell by cell on a workloads.Batch over its
    [tstart, tend]; ICNTRL/RCNTRL as SDIRK takes them (0-based, no merge;
    None = all zero).  Returns C, ierr, istat [8, n] (ISTATUS[1..8]) and
    rstat [3, n] (Texit, Hexit, Hnew)."""
    L = lib(batch.mech, fast)
    sz = sizes(batch.mech)
    n = batch.ncells
    f = lambda a: np.ascontiguousarray(a, np.float64)
    Cw = f(batch.C if C is None else C).copy()
    ierr = np.zeros(n, np.int32)
    ist = np.zeros((8, n), np.int32)
    rst = np.zeros((3, n), np.float64)
    ic = np.zeros(20, np.int32) if ICNTRL is None else np.ascontiguousarray(ICNTRL, np.int32)
    rc = np.zeros(20) if RCNTRL is None else f(RCNTRL)
    rcon = None if rconst is None else f(rconst)
    temp, cf, sun, ts, te = (f(batch.temp), f(batch.cfactor), f(batch.sun), f(batch.tstart),
                             f(batch.tend))
    atol, rtol = f(batch.atol), f(batch.rtol)
    L.sdirk_ref_batch(n, _p(Cw), _p(temp), _p(cf), _p(sun),
                      batch.sun_mode if sun_mode is None else sun_mode, _p(ts), _p(te),
                      _p(atol), _p(rtol), _p(ic), _p(rc), _p(rcon), _p(ierr), _p(ist), _p(rst),
                      nthreads)
    return {"C": Cw, "ierr": ierr, "istat": ist, "rstat": rst, "nvar": sz["NVAR"]}


def sdirk_method(code: int):
    """SDIRK table of ICNTRL[2] = code as the integrator uses it: S, Gamma,
    C, D, E, ELO, Theta[j, i], Alpha[j, i] (int/sdirk.c's rkTheta[j][i]);
    known=False for a code the switch sends to its default (Sdirk2a)."""
    L = lib_any()
    S = ctypes.c_int()
    G = ctypes.c_double()
    ELO = ctypes.c_double()
    C = np.zeros(5); D = np.zeros(5); E = np.zeros(5)
    Th = np.zeros(25); Al = np.zeros(25)
    r = L.sdirk_ref_method(code, ctypes.byref(S), ctypes.byref(G), _p(C), _p(D), _p(E),
                           ctypes.byref(ELO), _p(Th), _p(Al))
    s = S.value
    return {"S": s, "Gamma": G.value, "C": C[:s], "D": D[:s], "E": E[:s], "ELO": ELO.value,
            "Theta": Th.reshape(5, 5)[:s, :s], "Alpha": Al.reshape(5, 5)[:s, :s],
            "known": r == 1}


def integrate_rk(batch, ICNTRL=None, RCNTRL=None, rconst=None, nthreads: int = 0,
                 fast: bool = False, C=None, sun_mode=None):
    """KPP's C fully implicit Runge-Kutta (int/runge_kutta.c) cell by cell on a
    workloads.Batch over its [tstart, tend]; ICNTRL/RCNTRL as RungeKutta
    takes them (0-based, no merge; None = all zero).  Returns C, ierr, istat
    [8, n] (ISTATUS[0..7]) and rstat [3, n] (Texit, Hacc, Hnew)."""
    L = lib(batch.mech, fast)
    sz = sizes(batch.mech)
    n = batch.ncells
    f = lambda a: np.ascontiguousarray(a, np.float64)
    Cw = f(batch.C if C is None else C).copy()
    ierr = np.zeros(n, np.int32)
    ist = np.zeros((8, n), np.int32)
    rst = np.zeros((3, n), np.float64)
    ic = np.zeros(20, np.int32) if ICNTRL is None else np.ascontiguousarray(ICNTRL, np.int32)
    rc = np.zeros(20) if RCNTRL is None else f(RCNTRL)
    rcon = None if rconst is None else f(rconst)
    temp, cf, sun, ts, te = (f(batch.temp), f(batch.cfactor), f(batch.sun), f(batch.tstart),
                             f(batch.tend))
    atol, rtol = f(batch.atol), f(batch.rtol)
    L.rk_ref_batch(n, _p(Cw), _p(temp), _p(cf), _p(sun),
                   batch.sun_mode if sun_mode is None else sun_mode, _p(ts), _p(te),
                   _p(atol), _p(rtol), _p(ic), _p(rc), _p(rcon), _p(ierr), _p(ist), _p(rst),
                   nthreads)
    return {"C": Cw, "ierr": ierr, "istat": ist, "rstat": rst, "nvar": sz["NVAR"]}


def integrate_adj(batch, Lambda, ICNTRL=None, RCNTRL=None, first_step: bool = False,
                  cap: int = 200000, rconst=None, nthreads: int = 0, fast: bool = False,
                  C=None, path=None):
    """KPP's C discrete adjoint Rosenbrock (int/rosenbrock_adj.c) cell by cell
    on a workloads.Batch over its [tstart, tend].  Lambda [NADJ*NVAR, n] (row
    a*NVAR + i is Lambda[a][i]) holds dg/dy(tend) on entry and dg/dy(tstart)
    on return.  first_step=False keeps the reference's backward loop, which
    leaves out the first accepted step; True differentiates every step.  cap
    checkpoints per cell (the reference's bufsize is 200000).  Returns C,
    Lambda, ierr, istat [8, n] (ISTATUS[1..8]) and rstat [3, n] (Texit, Hexit,
    Hnew)."""
    L = lib(batch.mech, fast, path)
    n = batch.ncells
    f = lambda a: np.ascontiguousarray(a, np.float64)
    Cw = f(batch.C if C is None else C).copy()
    Lw = f(Lambda).copy()
    nadj = Lw.shape[0] // sizes(batch.mech, path)["NVAR"]
    ierr = np.zeros(n, np.int32)
    ist = np.zeros((8, n), np.int32)
    rst = np.zeros((3, n), np.float64)
    ic = np.zeros(20, np.int32) if ICNTRL is None else np.ascontiguousarray(ICNTRL, np.int32)
    rc = np.zeros(20) if RCNTRL is None else f(RCNTRL)
    rcon = None if rconst is None else f(rconst)
    temp, cf, sun, ts, te = (f(batch.temp), f(batch.cfactor), f(batch.sun), f(batch.tstart),
                             f(batch.tend))
    atol, rtol = f(batch.atol), f(batch.rtol)
    L.adj_ref_batch(n, _p(Cw), _p(temp), _p(cf), _p(sun), batch.sun_mode, _p(ts), _p(te),
                    _p(atol), _p(rtol), _p(ic), _p(rc), _p(rcon), nadj, _p(Lw),
                    1 if first_step else 0, cap, _p(ierr), _p(ist), _p(rst), nthreads)
    return {"C": Cw, "Lambda": Lw, "ierr": ierr, "istat": ist, "rstat": rst}


def rk_method(code: int, sdirk_error: int = 0):
    """The Runge-Kutta tables of ICNTRL[2] = code with the error estimator of
    ICNTRL[9] = sdirk_error (b0 depends on it), or None for a method that is
    not built.  kind: 1 Radau-2A, 2 Radau-1A, 3 Lobatto-3C, 4 Gauss."""
    L = lib_any()
    kind = ctypes.c_int()
    A = np.zeros(9); B = np.zeros(3); C = np.zeros(3); D = np.zeros(3); E = np.zeros(4)
    Bh = np.zeros(4); Bg = np.zeros(5); Th = np.zeros(3); gabe = np.zeros(4)
    T = np.zeros(9); TA = np.zeros(9)
    if not L.rk_ref_method(code, sdirk_error, ctypes.byref(kind), _p(A), _p(B), _p(C), _p(D),
                           _p(E), _p(Bh), _p(Bg), _p(Th), _p(gabe), _p(T), _p(TA)):
        return None
    return {"kind": kind.value, "A": A.reshape(3, 3), "B": B, "C": C, "D": D, "E": E,
            "Bhat": Bh, "Bgam": Bg, "Theta": Th, "Gamma": gabe[0], "Alpha": gabe[1],
            "Beta": gabe[2], "ELO": gabe[3], "T": T.reshape(3, 3), "TinvAinv": TA.reshape(3, 3)}


def f90_merge(ICNTRL_U=None, RCNTRL_U=None):
    """INTEGRATE's merge of the caller's ICNTRL_U/RCNTRL_U with its defaults."""
    L = lib_any()
    ic = np.zeros(20, np.int32)
    rc = np.zeros(20)
    iu = None if ICNTRL_U is None else np.ascontiguousarray(ICNTRL_U, np.int32)
    ru = None if RCNTRL_U is None else np.ascontiguousarray(RCNTRL_U, np.float64)
    L.ros_ref_f90_merge(_p(iu), _p(ru), _p(ic), _p(rc))
    return ic, rc


def method(code: int, f90: bool = False):
    """Method table of IPAR[3] = code (int/rosenbrock.c), or with f90=True of
    ICNTRL(3) = code (int/rosenbrock.f90's map, incl. Rang3)."""
    L = lib_any()
    S = ctypes.c_int()
    A = np.zeros(15); Cc = np.zeros(15); M = np.zeros(6); E = np.zeros(6)
    Al = np.zeros(6); G = np.zeros(6); ELO = ctypes.c_double(); NewF = np.zeros(6, np.int32)
    fn = L.ros_ref_method_f90 if f90 else L.ros_ref_method
    ok = fn(code, ctypes.byref(S), _p(A), _p(Cc), _p(M), _p(E), _p(Al), _p(G),
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
