"""Host-side binding of the batched Rosenbrock C ABI (include/kpphip.h).

Mirrors KPP's integrator interface for many cells at once:
``rosenbrock(...)`` takes the arguments of ``Rosenbrock(Y, Tstart, Tend,
AbsTol, RelTol, ode_Fun, ode_Jac, RPAR, IPAR)`` (int/rosenbrock.c:137-446)
with Y/Tstart/Tend per cell, keeps the IPAR/RPAR index meanings, and
returns Rosenbrock's per-cell IERR codes plus the IPAR[10..17] and
RPAR[10..11] outputs per cell.  Device memory is handled with torch tensors;
the arithmetic runs in the HIP library, which must be present: there is no
CPU fallback.
"""
from __future__ import annotations

import ctypes
import json
import os
from dataclasses import dataclass
from pathlib import Path

import numpy as np

PKG = Path(__file__).resolve().parent
LIB_PATH = Path(os.environ.get("KPPHIP_LIB", PKG / "lib" / "libkpphip.so"))
GEN = PKG / "_gen"

SUN_FIXED = 0
SUN_TOD = 1
PLAIN = 0
PERSISTENT = 1
REGISTER = 2  # one lane per cell, working set in VGPRs (small mechanisms only)
COOP = 3      # one wavefront per cell, LU in LDS (large mechanisms)

# IPAR[3] codes of int/rosenbrock.c:218-225
ROS2, ROS3, ROS4, RODAS3, RODAS4 = 1, 2, 3, 4, 5
METHOD_STAGES = {ROS2: 2, ROS3: 3, ROS4: 4, RODAS3: 4, RODAS4: 6}

IERR_TEXT = {
    1: "success",
    -1: "Improper value for maximal no of steps",
    -2: "Selected Rosenbrock method not implemented",
    -3: "Hmin/Hmax/Hstart must be positive",
    -4: "FacMin/FacMax/FacRej must be positive",
    -5: "Improper tolerance values",
    -6: "No of steps exceeds maximum bound",
    -7: "Step size too small (T + H/10 = T) or H < Roundoff",
    -8: "Matrix is repeatedly singular",
    # this library's bounds of the reject loop (none in int/rosenbrock.c)
    -9: "Error estimate is NaN or the step size is not finite",
    -10: "More than 100000 rejected steps in a row",
}


class KppError(RuntimeError):
    pass


class _Batch(ctypes.Structure):
    _fields_ = [
        ("ncells", ctypes.c_int), ("ld", ctypes.c_int),
        ("C", ctypes.c_void_p), ("temp", ctypes.c_void_p), ("cfactor", ctypes.c_void_p),
        ("sun", ctypes.c_void_p), ("sun_mode", ctypes.c_int),
        ("tstart", ctypes.c_void_p), ("tend", ctypes.c_void_p),
        ("tstart0", ctypes.c_double), ("tend0", ctypes.c_double),
        ("hstart", ctypes.c_void_p), ("rconst", ctypes.c_void_p),
        ("ierr", ctypes.c_void_p), ("istat", ctypes.c_void_p), ("rstat", ctypes.c_void_p),
        ("order", ctypes.c_void_p), ("hfirst", ctypes.c_void_p),
        ("fallbacks", ctypes.c_void_p),
    ]


class _Launch(ctypes.Structure):
    _fields_ = [("variant", ctypes.c_int), ("nblocks", ctypes.c_int), ("refill_min", ctypes.c_int),
                ("lanes", ctypes.c_int), ("lu_fp32", ctypes.c_int),
                ("lds_entries", ctypes.c_int)]


_lib = None


def library() -> ctypes.CDLL:
    """Load libkpphip.so; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise KppError(f"{LIB_PATH} is missing: build it with python -m kpp_amd.build "
                       "(there is no CPU fallback for the solver)")
    L = ctypes.CDLL(str(LIB_PATH))
    vp, ci, cd, cs = ctypes.c_void_p, ctypes.c_int, ctypes.c_double, ctypes.c_char_p
    L.kpp_mech_count.restype = ci
    L.kpp_mech_name.restype = cs
    L.kpp_mech_name.argtypes = [ci]
    L.kpp_mech_info.argtypes = [cs, vp]
    L.kpp_mech_species.restype = cs
    L.kpp_mech_initial.argtypes = [cs, vp]
    L.kpp_mech_species.argtypes = [cs, ci]
    L.kpp_mech_lu_pattern.argtypes = [cs, vp, vp, vp]
    L.kpp_check_params.argtypes = [vp, vp, vp, vp, ci]
    L.kpp_workspace_bytes.restype = ctypes.c_size_t
    L.kpp_workspace_bytes.argtypes = [cs, ci, ctypes.POINTER(_Launch), ci]
    L.kpp_rosenbrock_dev.argtypes = [cs, ctypes.POINTER(_Batch), vp, vp, vp, vp,
                                     ctypes.POINTER(_Launch), vp, ctypes.c_size_t, vp]
    for f in ("kpp_fun_dev", "kpp_jac_sp_dev"):
        getattr(L, f).argtypes = [cs, ci, ci, vp, vp, vp, vp, vp]
    L.kpp_decomp_dev.argtypes = [cs, ci, ci, vp, vp, vp]
    L.kpp_solve_dev.argtypes = [cs, ci, ci, vp, vp, vp]
    L.kpp_rates_dev.argtypes = [cs, ci, ci, vp, vp, vp, vp, vp, vp]
    L.kpp_rates_c_dev.argtypes = [cs, ci, ci, vp, vp, vp, vp, vp, vp, vp]
    L.kpp_ctx_create.argtypes = [ctypes.POINTER(vp), cs, ci, ci, ci]
    L.kpp_ctx_destroy.argtypes = [vp]
    L.kpp_set_state.argtypes = [vp, vp, vp, vp, vp]
    L.kpp_rosenbrock_batch.argtypes = [vp, vp, vp, vp, vp, vp, vp, vp]
    L.kpp_get_state.argtypes = [vp, vp, vp, vp, vp]
    L.kpp_ctx_set_tol.argtypes = [vp, vp, vp, ci]
    L.INTEGRATE_BATCH.argtypes = [vp, cd, cd]
    L.kpp_integrate_f90_dev.argtypes = [cs, ctypes.POINTER(_Batch), vp, vp, vp, vp, vp,
                                        ctypes.POINTER(_Launch), vp, ctypes.c_size_t, vp]
    L.kpp_workspace_bytes_f90.restype = ctypes.c_size_t
    L.kpp_workspace_bytes_f90.argtypes = [cs, ci, ctypes.POINTER(_Launch), ci]
    L.kpp_check_params_f90.argtypes = [vp, vp, vp, vp, ci]
    L.INTEGRATE_F90_BATCH.argtypes = [vp, cd, cd, vp, vp, vp, vp, vp]
    L.kpp_ctx_set_passive.argtypes = [vp, vp]
    L.kpp_integrate_h211b_dev.argtypes = [cs, ctypes.POINTER(_Batch), vp, vp, vp, vp,
                                          ctypes.POINTER(_Launch), vp, ctypes.c_size_t, vp]
    L.kpp_workspace_bytes_sdirk.restype = ctypes.c_size_t
    L.kpp_workspace_bytes_sdirk.argtypes = [cs, ci, ci]
    L.kpp_sdirk_dev.argtypes = [cs, ctypes.POINTER(_Batch), vp, vp, vp, vp, vp, ctypes.c_size_t,
                                vp]
    L.kpp_workspace_bytes_rk.restype = ctypes.c_size_t
    L.kpp_workspace_bytes_rk.argtypes = [cs, ci]
    L.kpp_rk_dev.argtypes = [cs, ctypes.POINTER(_Batch), vp, vp, vp, vp, vp, ctypes.c_size_t,
                             vp]
    L.kpp_workspace_bytes_adj.restype = ctypes.c_size_t
    L.kpp_workspace_bytes_adj.argtypes = [cs, ci, vp, ci, ci]
    L.kpp_rosenbrock_adj_dev.argtypes = [cs, ctypes.POINTER(_Batch), vp, vp, vp, vp, ci, vp, ci,
                                         ci, vp, ctypes.c_size_t, vp]
    _lib = L
    return L


@dataclass(frozen=True)
class Mechanism:
    name: str
    nvar: int
    nfix: int
    nspec: int
    nreact: int
    nonzero: int
    lu_nonzero: int
    rates_use_time: bool
    n_sun_rates: int
    rates_use_state: bool
    n_state_rates: int
    species: tuple
    stats: dict

    @property
    def var_species(self):
        return self.species[: self.nvar]


def mechanisms() -> list[str]:
    L = library()
    return [L.kpp_mech_name(i).decode() for i in range(L.kpp_mech_count())]


def mechanism(name: str) -> Mechanism:
    L = library()
    out = np.zeros(10, np.int32)
    r = L.kpp_mech_info(name.encode(), out.ctypes.data)
    if r != 0:
        raise KppError(f"unknown mechanism {name!r} (compiled: {mechanisms()})")
    sp = tuple(L.kpp_mech_species(name.encode(), k).decode() for k in range(out[2]))
    stats_path = GEN / f"{name}_stats.json"
    stats = json.loads(stats_path.read_text()) if stats_path.exists() else {}
    return Mechanism(name, *map(int, out[:6]), bool(out[6]), int(out[7]), bool(out[8]),
                     int(out[9]), sp, stats)


def lu_pattern(name: str):
    m = mechanism(name)
    icol = np.zeros(m.lu_nonzero, np.int32)
    crow = np.zeros(m.nvar + 1, np.int32)
    diag = np.zeros(m.nvar + 1, np.int32)
    library().kpp_mech_lu_pattern(name.encode(), icol.ctypes.data, crow.ctypes.data,
                                  diag.ctypes.data)
    return icol, crow, diag


def check_params(IPAR, RPAR, atol, rtol, nvar) -> int:
    IPAR = np.ascontiguousarray(IPAR, np.int32)
    RPAR = np.ascontiguousarray(RPAR, np.float64)
    atol = np.ascontiguousarray(atol, np.float64)
    rtol = np.ascontiguousarray(rtol, np.float64)
    return library().kpp_check_params(IPAR.ctypes.data, RPAR.ctypes.data, atol.ctypes.data,
                                      rtol.ctypes.data, nvar)


def _ptr(t):
    return None if t is None else ctypes.c_void_p(t.data_ptr())


def _stream_ptr(stream):
    if stream is None:
        import torch
        stream = torch.cuda.current_stream()
    return ctypes.c_void_p(stream.cuda_stream)


def default_ipar(method: int, scalar_tol: bool = True, autonomous: bool = False,
                 max_steps: int = 0):
    IPAR = np.zeros(20, np.int32)
    IPAR[0] = 1 if autonomous else 0
    IPAR[1] = 1 if scalar_tol else 0
    IPAR[2] = max_steps
    IPAR[3] = method
    return IPAR


class Workspace:
    """Device workspace for kpp_rosenbrock_dev, grown on demand."""

    def __init__(self, device="cuda"):
        self.device = device
        self.buf = None

    def get(self, nbytes: int):
        import torch
        if self.buf is None or self.buf.numel() < nbytes:
            self.buf = torch.empty(int(nbytes), dtype=torch.uint8, device=self.device)
        return self.buf


_default_ws = {}


def workspace_bytes(mech: str, method: int, ncells: int, variant=PLAIN, nblocks=0,
                    refill_min=0, lanes=0, lu_fp32=0, lds_entries=0) -> int:
    l = _Launch(variant, nblocks, refill_min, lanes, lu_fp32, lds_entries)
    return int(library().kpp_workspace_bytes(mech.encode(), method, ctypes.byref(l), ncells))


def rosenbrock(mech: str, C, temp, cfactor, *, tstart, tend, atol, rtol, IPAR, RPAR=None,
               sun=None, sun_mode=SUN_FIXED, hstart=None, rconst=None, order=None,
               hfirst=None, ierr=None, istat=None, rstat=None, variant=PLAIN, nblocks=0, refill_min=0,
               lanes=0, lu_fp32=0, lds_entries=0, fallbacks=None, workspace=None, stream=None):
    """Batched Rosenbrock on device tensors (float64 SoA, C is [NSPEC, ncells]).

    tstart/tend: per-cell tensors or floats.  atol/rtol: host sequences of
    NVAR entries (IPAR[1]==0) or 1.  Returns (code, ierr, istat, rstat) with
    code 0 or Rosenbrock's parameter error; launches asynchronously on
    ``stream`` (torch's current stream by default).  ``hfirst`` (float64
    [ncells], optional) receives each cell's first accepted step size.
    ``lu_fp32=1`` selects the mixed-precision integrator (FP32 LU, one FP64
    refinement sweep per solve, per-cell FP64 fallback counted in the int32
    [ncells] ``fallbacks``).  ``lds_entries`` (persistent kernel, 0/20/24)
    keeps the factored values of that many LU entries per cell in LDS.
    """
    import torch
    L = library()
    n = int(C.shape[1])
    dev = C.device
    if C.dtype != torch.float64 or not C.is_contiguous():
        raise KppError("C must be a contiguous float64 tensor [NSPEC, ncells]")
    if ierr is None:
        ierr = torch.empty(n, dtype=torch.int32, device=dev)
    if istat is None:
        istat = torch.empty((8, n), dtype=torch.int32, device=dev)
    if rstat is None:
        rstat = torch.empty((2, n), dtype=torch.float64, device=dev)
    b = _Batch()
    b.ncells = n
    b.ld = n
    b.C = C.data_ptr()
    b.temp = temp.data_ptr()
    b.cfactor = cfactor.data_ptr()
    b.sun = sun.data_ptr() if sun is not None else None
    b.sun_mode = sun_mode
    if torch.is_tensor(tstart):
        b.tstart = tstart.data_ptr()
    else:
        b.tstart0 = float(tstart)
    if torch.is_tensor(tend):
        b.tend = tend.data_ptr()
    else:
        b.tend0 = float(tend)
    b.hstart = hstart.data_ptr() if hstart is not None else None
    b.rconst = rconst.data_ptr() if rconst is not None else None
    b.ierr = ierr.data_ptr()
    b.istat = istat.data_ptr()
    b.rstat = rstat.data_ptr()
    b.order = order.data_ptr() if order is not None else None
    b.hfirst = hfirst.data_ptr() if hfirst is not None else None
    b.fallbacks = fallbacks.data_ptr() if fallbacks is not None else None
    IPAR = np.ascontiguousarray(IPAR, np.int32)
    RPAR = np.zeros(20) if RPAR is None else np.ascontiguousarray(RPAR, np.float64)
    atol = np.ascontiguousarray(np.atleast_1d(atol), np.float64)
    rtol = np.ascontiguousarray(np.atleast_1d(rtol), np.float64)
    l = _Launch(variant, nblocks, refill_min, lanes, lu_fp32, lds_entries)
    nbytes = L.kpp_workspace_bytes(mech.encode(), int(IPAR[3]), ctypes.byref(l), n)
    if workspace is None:
        workspace = _default_ws.setdefault(str(dev), Workspace(dev))
    ws = workspace.get(max(nbytes, 256))
    code = L.kpp_rosenbrock_dev(mech.encode(), ctypes.byref(b), IPAR.ctypes.data,
                                RPAR.ctypes.data, atol.ctypes.data, rtol.ctypes.data,
                                ctypes.byref(l), ctypes.c_void_p(ws.data_ptr()),
                                ctypes.c_size_t(ws.numel()), _stream_ptr(stream))
    if code <= -100:
        raise KppError(f"kpp_rosenbrock_dev failed with {code}")
    return code, ierr, istat, rstat


def integrate_f90(mech: str, C, temp, cfactor, *, tin, tout, atol, rtol, ICNTRL=None,
                  RCNTRL=None, sun=None, passive=None, hstart=None, rconst=None, ierr=None,
                  istat=None, rstat=None, variant=PLAIN, nblocks=0, lanes=0, workspace=None,
                  stream=None, h211b=False, order=None):
    """The F90 INTEGRATE (int/rosenbrock.f90:43-125) on device tensors, one
    call per cell over [tin, tout] (per-cell tensors or floats).

    ICNTRL/RCNTRL: the caller's ICNTRL_U/RCNTRL_U (20 entries; None = absent),
    merged as INTEGRATE merges them.  atol/rtol: NVAR entries (ICNTRL(2) = 0)
    or 1.  passive: NVAR flags (1 = left out of the error norm) or None.
    Returns (code, ierr, istat [8, n] = ISTATUS(1:8), rstat [3, n] = RSTATUS(1:3)).
    h211b=True: the INTEGRATE of int/rosenbrock_h211b_qssa.f90 (kpp_integrate_h211b_dev;
    H211b control with RCNTRL(15), RCNTRL(16); no passive species).
    """
    import torch
    L = library()
    n = int(C.shape[1])
    dev = C.device
    if C.dtype != torch.float64 or not C.is_contiguous():
        raise KppError("C must be a contiguous float64 tensor [NSPEC, ncells]")
    if ierr is None:
        ierr = torch.empty(n, dtype=torch.int32, device=dev)
    if istat is None:
        istat = torch.empty((8, n), dtype=torch.int32, device=dev)
    if rstat is None:
        rstat = torch.empty((3, n), dtype=torch.float64, device=dev)
    b = _Batch()
    b.ncells = n
    b.ld = n
    b.C = C.data_ptr()
    b.temp = temp.data_ptr()
    b.cfactor = cfactor.data_ptr()
    b.sun = sun.data_ptr() if sun is not None else None
    if torch.is_tensor(tin):
        b.tstart = tin.data_ptr()
    else:
        b.tstart0 = float(tin)
    if torch.is_tensor(tout):
        b.tend = tout.data_ptr()
    else:
        b.tend0 = float(tout)
    b.hstart = hstart.data_ptr() if hstart is not None else None
    b.rconst = rconst.data_ptr() if rconst is not None else None
    b.ierr = ierr.data_ptr()
    b.istat = istat.data_ptr()
    b.rstat = rstat.data_ptr()
    if order is not None:  # launch order (cell permutation); results per cell unchanged
        if order.dtype != torch.int32 or order.numel() != n:
            raise KppError("order must be an int32 tensor of ncells cell indices")
        b.order = order.data_ptr()
    ic = None if ICNTRL is None else np.ascontiguousarray(ICNTRL, np.int32)
    rc = None if RCNTRL is None else np.ascontiguousarray(RCNTRL, np.float64)
    pas = None if passive is None else np.ascontiguousarray(passive, np.int32)
    atol = np.ascontiguousarray(np.atleast_1d(atol), np.float64)
    rtol = np.ascontiguousarray(np.atleast_1d(rtol), np.float64)
    l = _Launch(variant, nblocks, 0, lanes, 0, 0)
    icm = 0 if ic is None or ic[2] == 0 else int(ic[2])
    nbytes = L.kpp_workspace_bytes_f90(mech.encode(), icm, ctypes.byref(l), n)
    if workspace is None:
        workspace = _default_ws.setdefault(str(dev), Workspace(dev))
    ws = workspace.get(max(nbytes, 256))
    ptr = lambda a: None if a is None else a.ctypes.data
    if h211b:
        if pas is not None:
            raise KppError("the H211b integrator has no passive species")
        code = L.kpp_integrate_h211b_dev(mech.encode(), ctypes.byref(b), ptr(ic), ptr(rc),
                                         atol.ctypes.data, rtol.ctypes.data, ctypes.byref(l),
                                         ctypes.c_void_p(ws.data_ptr()),
                                         ctypes.c_size_t(ws.numel()), _stream_ptr(stream))
    else:
        code = L.kpp_integrate_f90_dev(mech.encode(), ctypes.byref(b), ptr(ic), ptr(rc),
                                       atol.ctypes.data, rtol.ctypes.data, ptr(pas),
                                       ctypes.byref(l), ctypes.c_void_p(ws.data_ptr()),
                                       ctypes.c_size_t(ws.numel()), _stream_ptr(stream))
    if code <= -100:
        raise KppError(f"kpp_integrate_f90_dev failed with {code}")
    return code, ierr, istat, rstat


def _implicit(which, mech, C, temp, cfactor, tstart, tend, atol, rtol, ICNTRL, RCNTRL, sun,
              sun_mode, hstart, rconst, ierr, istat, rstat, workspace, stream, adj=None,
              order=None):
    """kpp_sdirk_dev / kpp_rk_dev / kpp_rosenbrock_adj_dev on device tensors
    (the shared batch set-up); adj = (nadj, Lambda, nchk, first_step)"""
    import torch
    L = library()
    n = int(C.shape[1])
    dev = C.device
    if C.dtype != torch.float64 or not C.is_contiguous():
        raise KppError("C must be a contiguous float64 tensor [NSPEC, ncells]")
    ierr = torch.empty(n, dtype=torch.int32, device=dev) if ierr is None else ierr
    istat = torch.empty((8, n), dtype=torch.int32, device=dev) if istat is None else istat
    rstat = torch.empty((3, n), dtype=torch.float64, device=dev) if rstat is None else rstat
    b = _Batch()
    b.ncells = n
    b.ld = n
    b.C = C.data_ptr()
    b.temp = temp.data_ptr()
    b.cfactor = cfactor.data_ptr()
    b.sun = sun.data_ptr() if sun is not None else None
    b.sun_mode = sun_mode
    if torch.is_tensor(tstart):
        b.tstart = tstart.data_ptr()
    else:
        b.tstart0 = float(tstart)
    if torch.is_tensor(tend):
        b.tend = tend.data_ptr()
    else:
        b.tend0 = float(tend)
    b.hstart = hstart.data_ptr() if hstart is not None else None
    b.rconst = rconst.data_ptr() if rconst is not None else None
    b.ierr = ierr.data_ptr()
    b.istat = istat.data_ptr()
    b.rstat = rstat.data_ptr()
    if order is not None:
        if order.dtype != torch.int32 or order.numel() != n:
            raise KppError("order must be an int32 tensor of ncells cell indices")
        b.order = order.data_ptr()
    ic = np.zeros(20, np.int32) if ICNTRL is None else np.ascontiguousarray(ICNTRL, np.int32)
    rc = np.zeros(20) if RCNTRL is None else np.ascontiguousarray(RCNTRL, np.float64)
    atol = np.ascontiguousarray(np.atleast_1d(atol), np.float64)
    rtol = np.ascontiguousarray(np.atleast_1d(rtol), np.float64)
    if which == "sdirk":
        nbytes = L.kpp_workspace_bytes_sdirk(mech.encode(), int(ic[2]), n)
        fn = L.kpp_sdirk_dev
    elif which == "rk":
        nbytes = L.kpp_workspace_bytes_rk(mech.encode(), n)
        fn = L.kpp_rk_dev
    else:
        nadj, lam, nchk, first = adj
        nbytes = L.kpp_workspace_bytes_adj(mech.encode(), n, ic.ctypes.data, nadj, nchk)
        fn = L.kpp_rosenbrock_adj_dev
        if nbytes == 0 and int(ic[2]) in (0, 1, 2, 3, 4, 5, 7):
            raise KppError(f"the adjoint integrator is not built for {mech}")
    if workspace is None:
        workspace = _default_ws.setdefault(str(dev), Workspace(dev))
    ws = workspace.get(max(nbytes, 256))
    head = (mech.encode(), ctypes.byref(b), ic.ctypes.data, rc.ctypes.data, atol.ctypes.data,
            rtol.ctypes.data)
    tail = (ctypes.c_void_p(ws.data_ptr()), ctypes.c_size_t(ws.numel()), _stream_ptr(stream))
    if which == "adj":
        code = fn(*head, nadj, ctypes.c_void_p(lam.data_ptr() if lam is not None else 0), nchk,
                  1 if first else 0, *tail)
    else:
        code = fn(*head, *tail)
    if code <= -100:
        raise KppError(f"{fn.__name__} failed with {code}")
    return code, ierr, istat, rstat


def sdirk(mech: str, C, temp, cfactor, *, tstart, tend, atol, rtol, ICNTRL=None, RCNTRL=None,
          sun=None, sun_mode=SUN_FIXED, hstart=None, rconst=None, ierr=None, istat=None,
          rstat=None, workspace=None, stream=None, order=None):
    """KPP's C SDIRK (int/sdirk.c) on device tensors, one lane per cell.
    ICNTRL/RCNTRL: 20 entries as SDIRK takes them (None = zeros).  Returns
    (code, ierr, istat [8, n] = ISTATUS[1..8], rstat [3, n] = RSTATUS[1..3])."""
    return _implicit("sdirk", mech, C, temp, cfactor, tstart, tend, atol, rtol, ICNTRL, RCNTRL,
                     sun, sun_mode, hstart, rconst, ierr, istat, rstat, workspace, stream,
                     order=order)


def runge_kutta(mech: str, C, temp, cfactor, *, tstart, tend, atol, rtol, ICNTRL=None,
                RCNTRL=None, sun=None, sun_mode=SUN_FIXED, hstart=None, rconst=None, ierr=None,
                istat=None, rstat=None, workspace=None, stream=None, order=None):
    """KPP's C fully implicit Runge-Kutta (int/runge_kutta.c) on device
    tensors, one lane per cell.  ICNTRL/RCNTRL: 20 entries as RungeKutta
    takes them (None = zeros: Radau-2A, classical error estimate,
    Gustafsson's controller).  Returns (code, ierr, istat [8, n] =
    ISTATUS[0..7], rstat [3, n] = RSTATUS[0..2])."""
    return _implicit("rk", mech, C, temp, cfactor, tstart, tend, atol, rtol, ICNTRL, RCNTRL,
                     sun, sun_mode, hstart, rconst, ierr, istat, rstat, workspace, stream,
                     order=order)


def rosenbrock_adj(mech: str, C, temp, cfactor, *, tstart, tend, atol, rtol, Lambda,
                   ICNTRL=None, RCNTRL=None, nchk=4096, first_step=False, sun=None,
                   sun_mode=SUN_FIXED, rconst=None, ierr=None, istat=None, rstat=None,
                   workspace=None, stream=None, order=None):
    """KPP's C discrete adjoint Rosenbrock (int/rosenbrock_adj.c) on device
    tensors, one lane per cell.  Lambda: float64 [NADJ*NVAR, ncells] (row
    a*NVAR + i = Lambda[a][i]), dg/dy at tend on entry, at tstart on return;
    nchk checkpoint slots per cell; first_step False keeps the reference's
    backward loop (the first accepted step left out).  Returns (code, ierr,
    istat [8, n] = ISTATUS[1..8], rstat [3, n] = RSTATUS[1..3])."""
    import torch
    nv = mechanism(mech).nvar
    nadj = 0 if Lambda is None else int(Lambda.shape[0]) // nv
    if Lambda is not None:
        if (Lambda.dtype != torch.float64 or not Lambda.is_contiguous() or
                Lambda.shape[0] != nadj * nv or Lambda.shape[1] != C.shape[1]):
            raise KppError("Lambda must be a contiguous float64 tensor [NADJ*NVAR, ncells]")
    return _implicit("adj", mech, C, temp, cfactor, tstart, tend, atol, rtol, ICNTRL, RCNTRL,
                     sun, sun_mode, None, rconst, ierr, istat, rstat, workspace, stream,
                     adj=(nadj, Lambda, int(nchk), bool(first_step)), order=order)


def _comp(fn, *args):
    r = fn(*args)
    if r != 0:
        raise KppError(f"{fn.__name__} failed with {r}")


def fun(mech, V, F, RCT, out, stream=None):
    n = V.shape[1]
    _comp(library().kpp_fun_dev, mech.encode(), n, n, _ptr(V), _ptr(F), _ptr(RCT), _ptr(out),
          _stream_ptr(stream))


def jac_sp(mech, V, F, RCT, out, stream=None):
    n = V.shape[1]
    _comp(library().kpp_jac_sp_dev, mech.encode(), n, n, _ptr(V), _ptr(F), _ptr(RCT), _ptr(out),
          _stream_ptr(stream))


def decomp(mech, JVS, ising, stream=None):
    n = JVS.shape[1]
    _comp(library().kpp_decomp_dev, mech.encode(), n, n, _ptr(JVS), _ptr(ising),
          _stream_ptr(stream))


def solve(mech, JVS, X, stream=None):
    n = X.shape[1]
    _comp(library().kpp_solve_dev, mech.encode(), n, n, _ptr(JVS), _ptr(X), _stream_ptr(stream))


def rates(mech, temp, cfactor, sun, time, out, stream=None, C=None):
    """Rate constants per cell; C ([NSPEC, n]) is the state that rate code
    reading concentrations sees (required for such mechanisms)."""
    n = temp.shape[0]
    if C is not None:
        _comp(library().kpp_rates_c_dev, mech.encode(), n, n, _ptr(temp), _ptr(cfactor),
              _ptr(sun), _ptr(time), _ptr(C), _ptr(out), _stream_ptr(stream))
        return
    _comp(library().kpp_rates_dev, mech.encode(), n, n, _ptr(temp), _ptr(cfactor), _ptr(sun),
          _ptr(time), _ptr(out), _stream_ptr(stream))


# ---- autotuner ----
def autotune(mech: str, C, C0, temp, cfactor, *, tstart, tend, atol, rtol, IPAR, RPAR=None,
             sun=None, sun_mode=SUN_FIXED, hstart=None, order=None, stream=None):
    """Fastest persistent-kernel launch shape on this batch: (variant, nblocks,
    refill_min, lanes, lds_entries), milliseconds, candidates tried.  C is
    reset from C0."""
    import torch
    L = library()
    _bind_extras(L)
    n = int(C.shape[1])
    b = _Batch()
    b.ncells = n
    b.ld = n
    b.C = C.data_ptr()
    b.temp = temp.data_ptr()
    b.cfactor = cfactor.data_ptr()
    b.sun = sun.data_ptr() if sun is not None else None
    b.sun_mode = sun_mode
    b.tstart = tstart.data_ptr()
    b.tend = tend.data_ptr()
    b.hstart = hstart.data_ptr() if hstart is not None else None
    b.order = order.data_ptr() if order is not None else None
    ierr = torch.empty(n, dtype=torch.int32, device=C.device)
    b.ierr = ierr.data_ptr()
    IPAR = np.ascontiguousarray(IPAR, np.int32)
    RPAR = np.zeros(20) if RPAR is None else np.ascontiguousarray(RPAR, np.float64)
    atol = np.ascontiguousarray(np.atleast_1d(atol), np.float64)
    rtol = np.ascontiguousarray(np.atleast_1d(rtol), np.float64)
    nbytes = int(L.kpp_autotune_workspace_bytes(mech.encode(), int(IPAR[3]), n))
    ws = torch.empty(max(nbytes, 256), dtype=torch.uint8, device=C.device)
    best = _Launch()
    ms = ctypes.c_float()
    tried = ctypes.c_int()
    r = L.kpp_autotune(mech.encode(), ctypes.byref(b), ctypes.c_void_p(C0.data_ptr()),
                       IPAR.ctypes.data, RPAR.ctypes.data, atol.ctypes.data, rtol.ctypes.data,
                       ctypes.c_void_p(ws.data_ptr()), ctypes.c_size_t(ws.numel()),
                       _stream_ptr(stream), ctypes.byref(best), ctypes.byref(ms),
                       ctypes.byref(tried))
    if r != 0:
        raise KppError(f"kpp_autotune failed with {r}")
    return ((best.variant, best.nblocks, best.refill_min, best.lanes, best.lds_entries),
            float(ms.value),
            int(tried.value))


# ---- graphs ----
class Graph:
    """A recorded sequence of library calls (kpp_capture_begin/end); replay with launch()."""

    def __init__(self, handle):
        self._h = handle

    def launch(self, stream=None):
        r = library().kpp_graph_launch(self._h, _stream_ptr(stream))
        if r != 0:
            raise KppError(f"kpp_graph_launch failed with {r}")

    def close(self):
        if self._h:
            library().kpp_graph_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def capture(fn, stream):
    """Record the library calls fn() makes on ``stream`` (a non-default torch
    stream, made current during the call) into a Graph.  Everything fn needs
    must be allocated beforehand: nothing may allocate while recording."""
    import torch
    L = library()
    _bind_extras(L)
    sp = ctypes.c_void_p(stream.cuda_stream)
    r = L.kpp_capture_begin(sp)
    if r != 0:
        raise KppError(f"kpp_capture_begin failed with {r}")
    h = ctypes.c_void_p()
    try:
        with torch.cuda.stream(stream):
            fn()
    finally:
        r = L.kpp_capture_end(sp, ctypes.byref(h))
    if r != 0:
        raise KppError(f"kpp_capture_end failed with {r}")
    return Graph(h)


def _bind_extras(L):
    if getattr(L, "_extras_bound", False):
        return
    vp = ctypes.c_void_p
    L.kpp_autotune_workspace_bytes.restype = ctypes.c_size_t
    L.kpp_autotune_workspace_bytes.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.c_int]
    L.kpp_autotune.argtypes = [ctypes.c_char_p, vp, vp, vp, vp, vp, vp, vp, ctypes.c_size_t, vp,
                               vp, vp, vp]
    L.kpp_capture_begin.argtypes = [vp]
    L.kpp_capture_end.argtypes = [vp, vp]
    L.kpp_graph_launch.argtypes = [vp, vp]
    L.kpp_graph_destroy.argtypes = [vp]
    L.kpp_graph_destroy.restype = None
    L._extras_bound = True

