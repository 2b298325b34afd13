"""The C-ABI library loads and exports what include/kpphip.h declares;
host-side logic (parameter checks, sizes) without touching a GPU."""
import ctypes
import re
from pathlib import Path

import numpy as np
import pytest

from kpp_amd import solver
from oracle import ref

HEADER = Path(__file__).resolve().parents[1] / "include" / "kpphip.h"


def declared_functions():
    text = re.sub(r"/\*.*?\*/", "", HEADER.read_text(), flags=re.S)
    names = re.findall(r"^\s*(?:const\s+)?[\w\s\*]+?\b(\w+)\s*\(", text, re.M)
    return sorted(set(n for n in names if n not in {"if", "return", "sizeof"}))


def test_header_declares_the_entry_points():
    names = declared_functions()
    for n in ("kpp_rosenbrock_dev", "kpp_fun_dev", "kpp_jac_sp_dev", "kpp_decomp_dev",
              "kpp_solve_dev", "kpp_rates_dev", "kpp_ctx_create", "INTEGRATE_BATCH"):
        assert n in names


def test_library_exports_every_declared_symbol():
    L = solver.library()
    for n in declared_functions():
        assert hasattr(L, n), n


def test_mechanism_registry():
    assert set(solver.mechanisms()) >= {"small_strato", "saprc99"}
    for m in ("small_strato", "saprc99"):
        info = solver.mechanism(m)
        s = ref.sizes(m)
        assert (info.nvar, info.nfix, info.nspec, info.nreact, info.nonzero, info.lu_nonzero) == (
            s["NVAR"], s["NFIX"], s["NSPEC"], s["NREACT"], s["NONZERO"], s["LU_NONZERO"])
        L = ref.lib(m)
        assert list(info.species) == [L.ros_ref_species_name(k).decode() for k in range(s["NSPEC"])]
        icol, crow, diag = solver.lu_pattern(m)
        ic = np.zeros_like(icol); cr = np.zeros_like(crow); dg = np.zeros_like(diag)
        p = lambda a: a.ctypes.data_as(ctypes.c_void_p)
        L.ros_ref_lu_pattern(p(ic), p(cr), p(dg))
        assert (ic == icol).all() and (cr == crow).all() and (dg == diag).all()
    with pytest.raises(solver.KppError):
        solver.mechanism("nope")


def test_check_params_codes():
    nv = 5
    ok = solver.default_ipar(solver.ROS3)
    at, rt = np.array([1.0]), np.array([1e-3])
    assert solver.check_params(ok, np.zeros(20), at, rt, nv) == 0
    ip = ok.copy(); ip[2] = -5
    assert solver.check_params(ip, np.zeros(20), at, rt, nv) == -1
    for m in (0, 6, 7, -1):
        ip = ok.copy(); ip[3] = m
        assert solver.check_params(ip, np.zeros(20), at, rt, nv) == -2
    for k, code in [(0, -3), (1, -3), (2, -3), (3, -4), (4, -4), (5, -4), (6, -4)]:
        rp = np.zeros(20); rp[k] = -1e-3
        assert solver.check_params(ok, rp, at, rt, nv) == code
    assert solver.check_params(ok, np.zeros(20), np.array([0.0]), rt, nv) == -5
    assert solver.check_params(ok, np.zeros(20), at, np.array([1.0]), nv) == -5
    assert solver.check_params(ok, np.zeros(20), at, np.array([2e-15]), nv) == -5
    vec = ok.copy(); vec[1] = 0
    atv = np.ones(nv); rtv = np.full(nv, 1e-3); rtv[4] = 0.0
    assert solver.check_params(vec, np.zeros(20), atv, rtv, nv) == -5
    rtv[4] = 1e-3
    assert solver.check_params(vec, np.zeros(20), atv, rtv, nv) == 0
    # order of the checks: a bad method is reported before a bad Hmin
    ip = ok.copy(); ip[3] = 9
    rp = np.zeros(20); rp[0] = -1
    assert solver.check_params(ip, rp, at, rt, nv) == -2


def test_workspace_sizes():
    for m, meth in (("small_strato", solver.ROS3), ("saprc99", solver.RODAS3)):
        a = solver.workspace_bytes(m, meth, 1000)
        b = solver.workspace_bytes(m, meth, 2000)
        assert 0 < a < b
        p = solver.workspace_bytes(m, meth, 100000, variant=solver.PERSISTENT, nblocks=64)
        assert 0 < p < solver.workspace_bytes(m, meth, 100000)
    assert solver.workspace_bytes("saprc99", 6, 10) == 0


def test_register_variant_geometry():
    # the register kernel needs no tiles: only the queue head and tolerances
    r = solver.workspace_bytes("small_strato", solver.ROS3, 4096, variant=solver.REGISTER)
    assert 0 < r <= 1024 < solver.workspace_bytes("small_strato", solver.ROS3, 4096)
    # compiled for small mechanisms only: saprc99 (920 LU entries) is refused
    assert solver.workspace_bytes("saprc99", solver.RODAS3, 4096, variant=solver.REGISTER) == 0
    # and not with the FP32 factorisation
    assert solver.workspace_bytes("small_strato", solver.ROS3, 4096, variant=solver.REGISTER,
                                  lu_fp32=1) == 0


def test_lds_mirror_geometry():
    # saprc99 and saprcnov are built with 20- and 24-entry LDS mirrors for the
    # persistent FP64 kernel; the tiles are unchanged, so is the workspace
    for mech in ("saprc99", "saprcnov"):
        w0 = solver.workspace_bytes(mech, solver.RODAS3, 4096, variant=solver.PERSISTENT)
        for lds in (20, 24):
            assert solver.workspace_bytes(mech, solver.RODAS3, 4096, variant=solver.PERSISTENT,
                                          lds_entries=lds) == w0 > 0
    # refused: other sizes, other kernels, the FP32 factorisation, a mechanism
    # built without a mirror
    for mech, kw in (("saprc99", {"variant": solver.PERSISTENT, "lds_entries": 32}),
                     ("saprc99", {"variant": solver.PERSISTENT, "lds_entries": -1}),
                     ("saprc99", {"variant": solver.PLAIN, "lds_entries": 24}),
                     ("saprc99", {"variant": solver.PERSISTENT, "lds_entries": 24, "lu_fp32": 1}),
                     ("small_strato", {"variant": solver.PERSISTENT, "lds_entries": 20})):
        assert solver.workspace_bytes(mech, solver.RODAS3, 4096, **kw) == 0, (mech, kw)
