"""Mechanisms built with the front end's newer commands on the GPU: the
strato_fam test mechanism (kpp_amd/mechs/strato_fam.kpp; #FAMILIES of odd
oxygen and #FUNCTION SPLIT) through the C Rosenbrock path (Fun) and the F90
flavour (Fun_SPLIT), against the oracle (pinned in tests/test_kpp_families.py).

Bars: with the oracle's rate constants, every counter identical and the
states to 1e-10 (both paths, the same arithmetic as the oracle's); with
device rates (time-of-day SUN, device cos), IERR everywhere, counters in at
least 99 % of cells and 1e-5 above ATOL/RTOL; and on the device the family
balance POx - LOx - (O + O1D + O3) kept to 1e-9 of the magnitudes
(the rounding of ~100 stiff solves; a missed family term drifts by O(1)).
"""
import ctypes

import numpy as np
import pytest

from kpp_amd import solver, workloads
from oracle import ref

pytestmark = pytest.mark.gpu
MECH = "strato_fam"


def _rates(b):
    p = lambda a: a.ctypes.data_as(ctypes.c_void_p)
    L = ref.lib(b.mech)
    nr = ref.sizes(b.mech)["NREACT"]
    R = np.zeros((nr, b.ncells))
    tmp = np.zeros(nr)
    for c in range(b.ncells):
        L.ros_ref_rconst(b.temp[c], b.sun[c], b.cfactor[c], b.tstart[c], p(tmp))
        R[:, c] = tmp
    return R


def _t(a, dev):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).to(dev)


def _rel(a, b, nv, floor):
    big = np.abs(b[:nv]) > floor
    r = np.abs(a[:nv] - b[:nv]) / np.maximum(np.abs(b[:nv]), 1e-300)
    return float(r[big].max()) if big.any() else 0.0


def _c_path(b, dev, rconst=None):
    import torch
    d = b.to_device(dev)
    code, ierr, istat, rstat = solver.rosenbrock(
        MECH, d["C"], d["temp"], d["cfactor"], tstart=d["tstart"], tend=d["tend"],
        atol=b.atol, rtol=b.rtol, IPAR=b.IPAR, RPAR=b.RPAR, sun=d["sun"], sun_mode=b.sun_mode,
        rconst=None if rconst is None else _t(rconst, dev))
    torch.cuda.synchronize()
    return code, {"C": d["C"].cpu().numpy(), "ierr": ierr.cpu().numpy(),
                  "istat": istat.cpu().numpy()}


def _f90_path(b, dev, ic, rconst=None):
    import torch
    d = b.to_device(dev)
    code, ierr, istat, rstat = solver.integrate_f90(
        MECH, d["C"], d["temp"], d["cfactor"], tin=d["tstart"], tout=d["tend"], atol=b.atol,
        rtol=b.rtol, ICNTRL=ic, sun=d["sun"], rconst=None if rconst is None else _t(rconst, dev))
    torch.cuda.synchronize()
    return code, {"C": d["C"].cpu().numpy(), "ierr": ierr.cpu().numpy(),
                  "istat": istat.cpu().numpy(), "rstat": rstat.cpu().numpy()}


def test_c_path_rates_given(gpu):
    b = workloads.strato_fam_tod(ncells=256)
    nv = ref.sizes(MECH)["NVAR"]
    R = _rates(b)
    o = ref.integrate(b, rconst=R)
    c, g = _c_path(b, gpu, R)
    assert c == 0 and (o["ierr"] == 1).all()
    assert np.array_equal(g["ierr"], o["ierr"]) and np.array_equal(g["istat"], o["istat"])
    assert _rel(g["C"], o["C"], nv, 1.0 / 1e-3) <= 1e-10


def test_f90_fun_split_rates_given(gpu):
    b = workloads.strato_fam_tod(ncells=256)
    nv = ref.sizes(MECH)["NVAR"]
    R = _rates(b)
    ic = np.zeros(20, np.int32)
    ic[1], ic[2], ic[14] = 1, 2, 1  # scalar tolerances, Ros3, rates at every call
    o = ref.integrate_f90(b, ic, rconst=R)
    c, g = _f90_path(b, gpu, ic, R)
    assert c == 0 and (o["ierr"] == 1).all()
    assert np.array_equal(g["ierr"], o["ierr"]) and np.array_equal(g["istat"], o["istat"])
    assert _rel(g["C"], o["C"], nv, 1.0 / 1e-3) <= 1e-10


def test_device_rates_and_family_balance(gpu):
    b = workloads.strato_fam_tod(ncells=512)
    sz = ref.sizes(MECH)
    nv = sz["NVAR"]
    o = ref.integrate(b)
    c, g = _c_path(b, gpu)
    assert c == 0 and np.array_equal(g["ierr"], o["ierr"]) and (g["ierr"] == 1).all()
    same = (g["istat"] == o["istat"]).all(axis=0)
    assert same.mean() >= 0.99
    assert _rel(g["C"][:, same], o["C"][:, same], nv, 1.0 / 1e-3) <= 1e-5
    i = {ref.species_name(MECH, k): k for k in range(sz["NSPEC"])}
    C0, C1 = b.C, g["C"]
    inv = lambda C: C[i["POx"]] - C[i["LOx"]] - (C[i["O"]] + C[i["O1D"]] + C[i["O3"]])
    scale = C1[i["O3"]] + C1[i["POx"]] + C1[i["LOx"]]
    assert np.all(np.abs(inv(C1) - inv(C0)) <= 1e-9 * scale)
