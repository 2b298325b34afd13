"""GPU tests of the MCM mechanism (F90 rate code, state-dependent rates) and of
the mixed-precision integrator (FP32 LU, one FP64 refinement sweep per solve,
per-cell FP64 fallback).

Bars:
 * MCM rate constants on the device (kpp_rates_c_dev) against the oracle's
   C translation at the same state: |rel| <= 1e-12 (device exp/pow/log10/cos
   vs glibc; parity at the libm boundary is not pinned by the reference);
 * MCM integrations with the oracle's rate constants supplied: all eight
   counters identical in every cell and |rel| <= 1e-10 for species above
   ATOL/RTOL in every cell;
 * MCM integrations, device rates, against the oracle: same IERR everywhere,
   identical step counts for >= 90 % of cells, on those cells |rel| <= 1e-6,
   and on every cell |rel| <= 1e-2 (a few RTOL: a cell whose step sequence
   differs through the device's exp/pow differs by the integration error),
   species above ATOL/RTOL;
 * the mixed-precision path against the oracle (rate constants supplied) on
   every cell of mcm_mixed (96 cells) and saprc99_uneven (1,024 cells): the
   same IERR and |rel| <= 1e-6 for species above ATOL/RTOL;
 * mixed precision against FP64 on the same cells: every cell succeeds; for
   species above ATOL/RTOL the two runs agree to a few RTOL (max |rel| <=
   1e-2), and against an FP64 run at RTOL 1e-8 the mixed run's error is at
   most twice the FP64 run's (plus 1e-6).
"""
import dataclasses

import numpy as np
import pytest

from kpp_amd import solver, workloads
from oracle import ref

pytestmark = pytest.mark.gpu


def _run(b, gpu, rconst=None, **kw):
    import torch
    d = b.to_device(gpu)
    fb = torch.zeros(b.ncells, dtype=torch.int32, device=gpu)
    rc = None if rconst is None else torch.from_numpy(np.ascontiguousarray(rconst)).to(gpu)
    code, ierr, istat, rstat = solver.rosenbrock(
        b.mech, d["C"], d["temp"], d["cfactor"], tstart=d["tstart"], tend=d["tend"],
        atol=b.atol, rtol=b.rtol, IPAR=b.IPAR, RPAR=b.RPAR, sun=d["sun"], sun_mode=b.sun_mode,
        fallbacks=fb, rconst=rc, **kw)
    torch.cuda.synchronize()
    assert code == 0
    return {"C": d["C"].cpu().numpy(), "ierr": ierr.cpu().numpy(), "istat": istat.cpu().numpy(),
            "fallbacks": fb.cpu().numpy()}


def _oracle_rates(b):
    """rate constants of each cell at its initial state (the oracle's)"""
    nr = ref.sizes(b.mech)["NREACT"]
    L = ref.lib(b.mech)
    R = np.zeros((nr, b.ncells))
    tmp = np.zeros(nr)
    for c in range(b.ncells):
        Cc = np.ascontiguousarray(b.C[:, c])
        L.ros_ref_rconst_c(b.temp[c], b.sun[c], b.cfactor[c], b.tstart[c], Cc.ctypes.data,
                           tmp.ctypes.data)
        R[:, c] = tmp
    return R


def _rel(a, b, nvar, thr):
    m = np.abs(b[:nvar]) > thr
    return np.abs(a[:nvar] - b[:nvar])[m] / np.abs(b[:nvar])[m]


def test_mcm_sizes():
    m = solver.mechanism("mcm")
    assert (m.nvar, m.nfix, m.nreact) == (610, 0, 1944)
    assert m.rates_use_state and m.n_state_rates > 0 and m.n_sun_rates > 0


def test_mcm_rates_match_oracle(gpu):
    import torch
    b = workloads.mcm_mixed(ncells=128, seed=61)
    n = b.ncells
    nr = ref.sizes("mcm")["NREACT"]
    L = ref.lib("mcm")
    want = np.zeros((nr, n))
    tmp = np.zeros(nr)
    for c in range(n):
        Cc = np.ascontiguousarray(b.C[:, c])
        L.ros_ref_rconst_c(b.temp[c], b.sun[c], b.cfactor[c], 0.0, Cc.ctypes.data, tmp.ctypes.data)
        want[:, c] = tmp
    d = b.to_device(gpu)
    out = torch.zeros((nr, n), dtype=torch.float64, device=gpu)
    tm = torch.zeros(n, dtype=torch.float64, device=gpu)
    solver.rates("mcm", d["temp"], d["cfactor"], d["sun"], tm, out, C=d["C"])
    torch.cuda.synchronize()
    got = out.cpu().numpy()
    # NUMBER rates are folded into Fun/Jac_SP and left as zero by the device
    types = np.array(solver.mechanism("mcm").stats["rate_types"])
    rows = types != 1
    g, w = got[rows], want[rows]
    assert np.all((g == 0) == (w == 0))
    nz = w != 0
    assert np.max(np.abs(g[nz] - w[nz]) / np.abs(w[nz])) <= 1e-12


@pytest.mark.parametrize("variant", [solver.PLAIN, solver.PERSISTENT])
def test_mcm_rates_given_matches_oracle(gpu, variant):
    b = workloads.mcm_mixed(ncells=96, seed=62)
    R = _oracle_rates(b)
    o = ref.integrate(b, rconst=R)
    g = _run(b, gpu, rconst=R, variant=variant)
    assert np.array_equal(g["ierr"], o["ierr"]) and np.all(o["ierr"] == 1)
    assert np.array_equal(g["istat"], o["istat"])  # all 8 counters, every cell
    nvar = solver.mechanism("mcm").nvar
    thr = float(b.atol[0] / b.rtol[0])
    r = _rel(g["C"], o["C"], nvar, thr)
    print("mcm rates given: max rel", r.max())
    assert r.max() <= 1e-10


@pytest.mark.parametrize("variant", [solver.PLAIN, solver.PERSISTENT])
def test_mcm_integration_matches_oracle(gpu, variant):
    b = workloads.mcm_mixed(ncells=96, seed=62)
    o = ref.integrate(b)
    g = _run(b, gpu, variant=variant)
    assert np.array_equal(g["ierr"], o["ierr"]) and np.all(o["ierr"] == 1)
    same = np.all(g["istat"] == o["istat"], axis=0)
    print("identical step counts:", same.mean())
    assert same.mean() >= 0.9
    nvar = solver.mechanism("mcm").nvar
    thr = float(b.atol[0] / b.rtol[0])
    r = _rel(g["C"][:, same], o["C"][:, same], nvar, thr)
    assert r.max() <= 1e-6
    r_all = _rel(g["C"], o["C"], nvar, thr)
    print("all cells: max rel", r_all.max())
    assert r_all.max() <= 1e-2


@pytest.mark.parametrize("wl,n", [("mcm_mixed", 96), ("saprc99_uneven", 1024)])
@pytest.mark.parametrize("variant", [solver.PLAIN, solver.PERSISTENT])
def test_mixed_precision_matches_oracle(gpu, wl, n, variant):
    b = workloads.WORKLOADS[wl](ncells=n, seed=64)
    R = _oracle_rates(b)
    o = ref.integrate(b, rconst=R)
    g = _run(b, gpu, rconst=R, variant=variant, lu_fp32=1)
    assert np.array_equal(g["ierr"], o["ierr"]) and np.all(o["ierr"] == 1)
    nvar = solver.mechanism(b.mech).nvar
    thr = float(b.atol[0] / b.rtol[0])
    r = _rel(g["C"], o["C"], nvar, thr)
    same = np.all(g["istat"] == o["istat"], axis=0)
    print(wl, "mixed vs oracle: max rel", r.max(), "identical counters", same.mean(),
          "fallback attempts", int(g["fallbacks"].sum()))
    assert r.max() <= 1e-6


@pytest.mark.parametrize("wl,n", [("saprc99_uneven", 4096), ("mcm_mixed", 256)])
def test_mixed_precision_against_fp64(gpu, wl, n):
    b = workloads.WORKLOADS[wl](ncells=n, seed=63)
    nvar = solver.mechanism(b.mech).nvar
    thr = float(b.atol[0] / b.rtol[0])
    r64 = _run(b, gpu, variant=solver.PERSISTENT)
    r32 = _run(b, gpu, variant=solver.PERSISTENT, lu_fp32=1)
    assert np.all(r64["ierr"] == 1) and np.all(r32["ierr"] == 1)
    assert np.all(r64["fallbacks"] == 0)
    d = _rel(r32["C"], r64["C"], nvar, thr)
    print(wl, "mixed vs fp64 max rel", d.max(), "fallback attempts",
          int(r32["fallbacks"].sum()), "cells", int((r32["fallbacks"] > 0).sum()))
    assert d.max() <= 1e-2
    tight = dataclasses.replace(b, rtol=np.full_like(b.rtol, 1e-8))
    rt = _run(tight, gpu, variant=solver.PERSISTENT)
    e64 = _rel(r64["C"], rt["C"], nvar, thr).max()
    e32 = _rel(r32["C"], rt["C"], nvar, thr).max()
    print(wl, "error vs RTOL 1e-8: fp64", e64, "mixed", e32)
    assert e32 <= 2 * e64 + 1e-6
