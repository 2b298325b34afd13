"""KPP's C fully implicit Runge-Kutta integrator on the GPU (kpp_rk_dev)
against the oracle's restatement of int/runge_kutta.c (pinned in
tests/test_rk_oracle.py).

As for SDIRK, the libm boundary matters: the reference's norms take SQRT =
pow(x, 0.5) (glibc's differs from the device's correctly rounded sqrt by one
ULP for about 0.08 % of arguments) and the Newton iterations compare norms
and pow(Theta, ...) against NewtonTol, so a one-ULP difference can move a
Newton iteration, and the results then differ at the Newton tolerance's
level.  Bars:
 * with the oracle's rate constants: every cell has the oracle's IERR; at
   least 95 % of cells have all eight counters; species above ATOL/RTOL agree
   to 1e-6 in those cells and to 2e-3 in all -- every method, both error
   estimators, scalar and vector tolerances, the classical controller;
 * with device-computed rates (fixed SUN, time-of-day SUN): the same IERR in
   at least 98 % of cells, the oracle's counters in at least 95 %, 1e-6
   (saprc99) or 1e-5 (small_strato, device cos in Update_SUN) above
   ATOL/RTOL in cells with the oracle's counters;
 * the linear6 closed-form scenarios of the CPU test on the device (counters
   exactly the model's, RSTATUS to 1e-9), a singular first matrix and five
   in a row handled as the oracle does, parameter errors in every cell, MCM
   (not built) refused.
"""
import ctypes

import numpy as np
import pytest

from kpp_amd import solver, workloads
from oracle import ref

pytestmark = pytest.mark.gpu


def _t(a, dev):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).to(dev)


def _oracle_rates(b):
    p = lambda a: a.ctypes.data_as(ctypes.c_void_p)
    L = ref.lib(b.mech)
    nr = ref.sizes(b.mech)["NREACT"]
    R = np.zeros((nr, b.ncells))
    tmp = np.zeros(nr)
    for c in range(b.ncells):
        L.ros_ref_rconst(b.temp[c], b.sun[c], b.cfactor[c], b.tstart[c], p(tmp))
        R[:, c] = tmp
    return R


def _gpu(b, dev, ic, rc=None, rconst=None, atol=None, rtol=None):
    import torch
    d = b.to_device(dev)
    code, ierr, istat, rstat = solver.runge_kutta(
        b.mech, d["C"], d["temp"], d["cfactor"], tstart=d["tstart"], tend=d["tend"],
        atol=b.atol if atol is None else atol, rtol=b.rtol if rtol is None else rtol,
        ICNTRL=ic, RCNTRL=rc, sun=d["sun"], sun_mode=b.sun_mode,
        rconst=None if rconst is None else _t(rconst, dev))
    torch.cuda.synchronize()
    return code, {"C": d["C"].cpu().numpy(), "ierr": ierr.cpu().numpy(),
                  "istat": istat.cpu().numpy(), "rstat": rstat.cpu().numpy()}


def _ic(vector=False, method=1, sdirk_error=0, classical=False):
    ic = np.zeros(20, np.int32)
    ic[1] = 0 if vector else 1
    ic[2] = method
    ic[9] = sdirk_error
    ic[10] = 1 if classical else 0
    return ic


def _rel(a, b, nv, floor):
    big = np.abs(b[:nv]) > floor
    r = np.abs(a[:nv] - b[:nv]) / np.maximum(np.abs(b[:nv]), 1e-300)
    return float(r[big].max()) if big.any() else 0.0


def _bar(g, o, nv):
    assert np.array_equal(g["ierr"], o["ierr"])
    same = (g["istat"] == o["istat"]).all(axis=0)
    assert same.mean() >= 0.95, same.mean()
    assert _rel(g["C"][:, same], o["C"][:, same], nv, 1.0 / 1e-3) <= 1e-6
    assert _rel(g["C"], o["C"], nv, 1.0 / 1e-3) <= 2e-3
    # step sizes carry the error estimate's last-ULP differences through the
    # controller's pow: a few 1e-4 at most where the counters agree
    for k, tol in ((0, 1e-12), (1, 2e-3), (2, 2e-3)):
        assert np.allclose(g["rstat"][k][same], o["rstat"][k][same], rtol=tol, atol=0.0), k


@pytest.mark.parametrize("method", [1, 2, 3, 4])
@pytest.mark.parametrize("sdirk_error", [0, 1])
def test_rates_given_every_method(gpu, method, sdirk_error):
    b = workloads.saprc99_uneven(ncells=256)
    nv = ref.sizes("saprc99")["NVAR"]
    R = _oracle_rates(b)
    ic = _ic(method=method, sdirk_error=sdirk_error)
    o = ref.integrate_rk(b, ic, rconst=R)
    c, g = _gpu(b, gpu, ic, rconst=R)
    assert c == 0 and (o["ierr"] == 1).all()
    _bar(g, o, nv)


def test_vector_tolerances_classical_controller(gpu):
    b = workloads.saprc99_uneven(ncells=256, seed=13)
    nv = ref.sizes("saprc99")["NVAR"]
    rng = np.random.default_rng(2)
    b.atol = 10.0 ** rng.uniform(-1, 1, nv)
    b.rtol = np.full(nv, 1e-3)
    R = _oracle_rates(b)
    ic = _ic(vector=True, method=3, classical=True)
    ic[5] = 1  # zero Newton start
    rc = np.zeros(20)
    rc[[3, 4, 9, 10]] = [0.3, 5.0, 0.9, 1.3]  # FacMin, FacMax, Qmin, Qmax
    o = ref.integrate_rk(b, ic, rc, rconst=R)
    c, g = _gpu(b, gpu, ic, rc, rconst=R)
    assert c == 0 and (o["ierr"] == 1).all()
    _bar(g, o, nv)


@pytest.mark.parametrize("wl,floor_tol", [("saprc99_uneven", 1e-6), ("small_strato_tod", 1e-5)])
def test_device_rates(gpu, wl, floor_tol):
    b = workloads.WORKLOADS[wl](ncells=512)
    nv = ref.sizes(b.mech)["NVAR"]
    ic = _ic(method=1, sdirk_error=1)
    o = ref.integrate_rk(b, ic)
    c, g = _gpu(b, gpu, ic)
    assert c == 0
    assert (g["ierr"] == o["ierr"]).mean() >= 0.98
    ok = (g["ierr"] == 1) & (o["ierr"] == 1)
    assert ok.mean() >= 0.98
    same = (g["istat"] == o["istat"]).all(axis=0) & ok
    assert same.mean() >= 0.95
    assert _rel(g["C"][:, same], o["C"][:, same], nv, 1.0 / 1e-3) <= floor_tol


def test_linear6_model_scenarios(gpu):
    from tests.test_rk_oracle import SCEN_RK, _model_rk
    from tests.test_step_control import _batch, _x1_rates
    for code, se, lam, h0 in SCEN_RK:
        b = _batch(np.array([1.0e6]), 0.0, 3600.0, rtol=1e-4, atol=1.0)
        ic = _ic(method=code, sdirk_error=se)
        rc = np.zeros(20)
        rc[2] = h0
        R = _x1_rates(lam)
        c, g = _gpu(b, gpu, ic, rc, rconst=R)
        o = ref.integrate_rk(b, ic, rc, rconst=R)
        y, st, rs = _model_rk(code, se, lam, 1.0e6, 0.0, 3600.0, h0, 1.0, 1e-4)
        want = [st[k] for k in ("Nfun", "Njac", "Nstp", "Nacc", "Nrej", "Ndec", "Nsol")]
        assert c == 0 and int(g["ierr"][0]) == 1
        assert list(g["istat"][:7, 0]) == want, (code, se, lam)
        assert np.array_equal(g["istat"], o["istat"])
        for k in range(3):
            assert g["rstat"][k, 0] == pytest.approx(rs[k], rel=1e-9), k
        assert g["C"][0, 0] == pytest.approx(y, rel=1e-9)


def test_singular_matrices(gpu):
    from tests.test_step_control import _batch, _rates
    m = ref.rk_method(1)
    h0 = 2.0 ** -8
    rc = np.zeros(20)
    rc[2] = h0
    ic = _ic(method=1)
    # one singular first matrix: halved, counted, the integration goes on
    b = _batch(np.array([1.0]), 0.0, 2.0 ** -5)
    R = _rates(1, g=0.0)
    R[0, 0] = m["Gamma"] / h0
    o = ref.integrate_rk(b, ic, rc, rconst=R)
    c, g = _gpu(b, gpu, ic, rc, rconst=R)
    assert c == 0 and int(g["ierr"][0]) == 1 and g["istat"][7, 0] == 1
    assert np.array_equal(g["istat"], o["istat"])
    assert g["C"][0, 0] == pytest.approx(o["C"][0, 0], rel=1e-12)
    # five in a row: reported by the reference (-12), the integration goes on
    b = _batch(np.array([1.0]), 0.0, h0)
    b.C[:5] = 1.0
    R = _rates(1, g=0.0)
    for k in range(5):
        R[k, 0] = m["Gamma"] / (h0 * 2.0 ** -k)
    o = ref.integrate_rk(b, ic, rc, rconst=R)
    c, g = _gpu(b, gpu, ic, rc, rconst=R)
    assert int(o["ierr"][0]) == 1 and o["istat"][7, 0] == 5
    assert np.array_equal(g["ierr"], o["ierr"]) and np.array_equal(g["istat"], o["istat"])
    assert np.allclose(g["C"][:6, 0], o["C"][:6, 0], rtol=1e-12, atol=0.0)


def test_errors_and_unsupported(gpu):
    b = workloads.saprc99_uneven(ncells=64)
    for idx, val, want in ((3, -1, -1), (2, 5, -13), (4, -3, -2)):
        ic = _ic()
        ic[idx] = val
        c, g = _gpu(b, gpu, ic)
        assert c == want and (g["ierr"] == want).all()
    rc = np.zeros(20)
    rc[9] = 2.0  # Qmin > 1
    c, g = _gpu(b, gpu, _ic(), rc)
    assert c == -7 and (g["ierr"] == -7).all()
    c, g = _gpu(b, gpu, _ic(), rtol=np.array([0.0]))
    assert c == -8 and (g["ierr"] == -8).all()
    bm = workloads.mcm_mixed(ncells=4)
    with pytest.raises(solver.KppError):
        _gpu(bm, gpu, _ic())


def test_launch_order_leaves_results_unchanged(gpu):
    """The work-queue kernels of SDIRK and Runge-Kutta take cells in the
    given launch order; each cell's integration is its own, so a permuted
    order gives bit-identical states and counters."""
    import torch
    b = workloads.saprc99_uneven(ncells=300, seed=5)
    perm = torch.from_numpy(np.random.default_rng(1).permutation(300).astype(np.int32)).to(gpu)
    for fn in (solver.sdirk, solver.runge_kutta):
        outs = []
        for order in (None, perm):
            d = b.to_device(gpu)
            code, ierr, istat, rstat = fn(
                b.mech, d["C"], d["temp"], d["cfactor"], tstart=d["tstart"], tend=d["tend"],
                atol=b.atol, rtol=b.rtol, ICNTRL=_ic(), sun=d["sun"], sun_mode=b.sun_mode,
                order=order)
            torch.cuda.synchronize()
            assert code == 0 and bool((ierr == 1).all())
            outs.append((d["C"].cpu().numpy(), istat.cpu().numpy(), rstat.cpu().numpy()))
        for x, y in zip(*outs):
            assert np.array_equal(x, y)
