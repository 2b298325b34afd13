"""KPP's C discrete adjoint Rosenbrock on the GPU (kpp_rosenbrock_adj_dev)
against the oracle's restatement of int/rosenbrock_adj.c (pinned in
tests/test_adj_oracle.py by finite differences, a restart identity and a
closed form).

The device's pow and sqrt stand where the reference calls glibc's pow
(Fac = FacSafe/pow(Err, 1/ELO), SQRT = pow(x, 0.5)), so a step-size decision
can differ by one ULP and move the step sequence of a cell.  Bars:
 * fixed steps (Hmin = Hmax = Hstart: no step control), every method, all
   NVAR adjoint vectors: counters identical, states and Lambda to 1e-10
   relative with the oracle's rate constants, 1e-6 with device rates (whose
   last-ULP differences the finite-difference dJ/dT divides by its delta;
   non-autonomous, dF/dT and dJ/dT included), with
   SaveLU and without, the reference's loop and every step;
 * adaptive steps with the oracle's rate constants (saprc99, Rodas3 and
   Ros3, vector tolerances): IERR everywhere, counters in at least 95 % of
   cells, states above ATOL/RTOL to 1e-8 and Lambda to 1e-4 of its largest
   entry in those cells (where the counters agree the step sizes still
   differ at 1e-5, the pow of the controller);
 * with device rates (time-of-day SUN, device cos in Update_SUN): IERR
   everywhere, counters in at least 90 %, 1e-5 in those cells;
 * the linear decay closed form lambda(t0) = Y(tend)/Y(t0) lambda(tend);
 * parameter errors, the checkpoint limit (-12), forward only (AdjointType
   1), MCM (not built) refused.
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


def _gpu(b, dev, lam, ic, rc=None, rconst=None, nchk=4096, first_step=False, C=None):
    import torch
    d = b.to_device(dev)
    if C is not None:
        d["C"] = _t(C, dev)
    L = _t(lam, dev) if lam is not None else None
    code, ierr, istat, rstat = solver.rosenbrock_adj(
        b.mech, d["C"], d["temp"], d["cfactor"], tstart=d["tstart"], tend=d["tend"],
        atol=b.atol, rtol=b.rtol, Lambda=L, ICNTRL=ic, RCNTRL=rc, nchk=nchk,
        first_step=first_step, sun=d["sun"], sun_mode=b.sun_mode,
        rconst=None if rconst is None else _t(rconst, dev))
    torch.cuda.synchronize()
    return code, {"C": d["C"].cpu().numpy(), "Lambda": None if L is None else L.cpu().numpy(),
                  "ierr": ierr.cpu().numpy(), "istat": istat.cpu().numpy(),
                  "rstat": rstat.cpu().numpy()}


def _ic(method, vector=False, autonomous=False, save_lu=False, adj_type=0):
    ic = np.zeros(20, np.int32)
    ic[0] = 1 if autonomous else 0
    ic[1] = 0 if vector else 1
    ic[2] = method
    ic[6] = adj_type
    ic[7] = 1 if save_lu else 0
    return ic


def _rel(a, b):
    return float(np.max(np.abs(a - b) / np.maximum(np.abs(b), 1e-300)))


@pytest.mark.parametrize("method", [1, 2, 3, 4, 5, 7])
@pytest.mark.parametrize("save_lu,first_step", [(False, False), (True, True)])
def test_fixed_steps_every_method(gpu, method, save_lu, first_step):
    b = workloads.small_strato_tod(ncells=128)
    b.tend = b.tstart + 1800.0
    nv = ref.sizes(b.mech)["NVAR"]
    rc = np.zeros(20)
    rc[0] = rc[1] = rc[2] = 60.0
    lam = np.zeros((nv * nv, b.ncells))
    for a in range(nv):
        lam[a * nv + a] = 1.0
    ic = _ic(method, save_lu=save_lu)
    R = _oracle_rates(b) if method == 3 else None  # one method with given rates
    o = ref.integrate_adj(b, lam, ic, rc, first_step=first_step, rconst=R)
    c, g = _gpu(b, gpu, lam, ic, rc, rconst=R, first_step=first_step)
    assert c == 0 and (o["ierr"] == 1).all()
    assert np.array_equal(g["ierr"], o["ierr"]) and np.array_equal(g["istat"], o["istat"])
    # device rates: the device cos in Update_SUN moves the rates by an ULP, and
    # the finite-difference dJ/dT of the backward pass divides that by delta
    tol = 1e-10 if R is not None else 1e-6
    assert _rel(g["C"][:nv], o["C"][:nv]) <= tol
    scale = np.abs(o["Lambda"]).max(axis=0)
    assert np.max(np.abs(g["Lambda"] - o["Lambda"]) / scale) <= tol
    assert np.allclose(g["rstat"], o["rstat"], rtol=1e-12, atol=0.0)


@pytest.mark.parametrize("method", [2, 4])
def test_adaptive_rates_given(gpu, method):
    b = workloads.saprc99_uneven(ncells=256)
    nv = ref.sizes("saprc99")["NVAR"]
    rng = np.random.default_rng(3)
    b.atol = np.full(nv, 1.0)
    b.rtol = np.full(nv, 1e-3)
    lam = rng.standard_normal((2 * nv, b.ncells))
    R = _oracle_rates(b)
    ic = _ic(method, vector=True)
    o = ref.integrate_adj(b, lam, ic, rconst=R)
    c, g = _gpu(b, gpu, lam, ic, rconst=R)
    assert c == 0 and (o["ierr"] == 1).all()
    assert np.array_equal(g["ierr"], o["ierr"])
    same = (g["istat"] == o["istat"]).all(axis=0)
    assert same.mean() >= 0.95, same.mean()
    big = np.abs(o["C"][:nv]) > 1.0 / 1e-3
    r = np.abs(g["C"][:nv] - o["C"][:nv]) / np.maximum(np.abs(o["C"][:nv]), 1e-300)
    assert r[big & same[None, :]].max() <= 1e-8
    scale = np.abs(o["Lambda"]).max(axis=0)
    d = np.abs(g["Lambda"] - o["Lambda"]) / scale
    assert d[:, same].max() <= 1e-4  # the step sizes differ at 1e-5


def test_device_rates(gpu):
    b = workloads.small_strato_tod(ncells=512)
    nv = ref.sizes(b.mech)["NVAR"]
    lam = np.ones((nv, b.ncells))
    ic = _ic(4)
    o = ref.integrate_adj(b, lam, ic)
    c, g = _gpu(b, gpu, lam, ic)
    assert c == 0 and np.array_equal(g["ierr"], o["ierr"]) and (g["ierr"] == 1).all()
    same = (g["istat"] == o["istat"]).all(axis=0)
    assert same.mean() >= 0.90, same.mean()
    big = np.abs(o["C"][:nv]) > 1.0 / 1e-3
    r = np.abs(g["C"][:nv] - o["C"][:nv]) / np.maximum(np.abs(o["C"][:nv]), 1e-300)
    assert r[big & same[None, :]].max() <= 1e-5
    scale = np.abs(o["Lambda"]).max(axis=0)
    assert (np.abs(g["Lambda"] - o["Lambda"]) / scale)[:, same].max() <= 1e-5


@pytest.mark.parametrize("method", [1, 2, 3, 4, 5, 7])
def test_linear_decay_closed_form(gpu, method):
    from tests.test_step_control import _batch, _x1_rates
    nv = ref.sizes("linear6")["NVAR"]
    for lam_rate in (-2e-3, -5e-2, 5e-3):
        b = _batch(np.array([1.0e6]), 0.0, 3600.0, rtol=1e-4, atol=1.0)
        L0 = np.zeros((nv, 1))
        L0[0] = 2.5
        ic = _ic(method)
        R = _x1_rates(lam_rate)
        c, g = _gpu(b, gpu, L0, ic, rconst=R, first_step=True)
        o = ref.integrate_adj(b, L0, ic, rconst=R, first_step=True)
        assert c == 0 and int(g["ierr"][0]) == 1
        assert np.array_equal(g["istat"], o["istat"])
        assert g["Lambda"][0, 0] == pytest.approx(2.5 * g["C"][0, 0] / 1.0e6, rel=1e-12)
        assert g["Lambda"][0, 0] == pytest.approx(o["Lambda"][0, 0], rel=1e-12)


def test_errors_limits_and_unsupported(gpu):
    b = workloads.small_strato_tod(ncells=64)
    nv = ref.sizes(b.mech)["NVAR"]
    lam = np.ones((nv, b.ncells))
    for idx, val, want in ((3, 100, -1), (2, 6, -2), (6, 5, -9), (6, 3, -13)):
        ic = _ic(4)
        ic[idx] = val
        c, g = _gpu(b, gpu, lam, ic)
        assert c == want and (g["ierr"] == want).all()
    rc = np.zeros(20)
    rc[3] = -0.5
    c, g = _gpu(b, gpu, lam, _ic(4), rc)
    assert c == -4 and (g["ierr"] == -4).all()
    # too few checkpoint slots: -12 in every cell that needs more
    c, g = _gpu(b, gpu, lam, _ic(4), nchk=3)
    o = ref.integrate_adj(b, lam, _ic(4), cap=3)
    assert c == 0 and np.array_equal(g["ierr"], o["ierr"]) and (g["ierr"] == -12).any()
    # forward only: the states of the adjoint run, Lambda untouched
    c, g1 = _gpu(b, gpu, lam, _ic(4, adj_type=1))
    c2, g2 = _gpu(b, gpu, lam, _ic(4))
    assert c == 0 and (g1["ierr"] == 1).all() and np.array_equal(g1["Lambda"], lam)
    assert np.array_equal(g1["C"], g2["C"])
    bm = workloads.mcm_mixed(ncells=4)
    with pytest.raises(solver.KppError):
        _gpu(bm, gpu, None, _ic(4))


def test_launch_order_leaves_results_unchanged(gpu):
    """Cells launched in a permuted order (lane k integrates cell order[k])
    give bit-identical states, Lambda and counters."""
    import torch
    b = workloads.small_strato_tod(ncells=200)
    nv = ref.sizes(b.mech)["NVAR"]
    perm = torch.from_numpy(np.random.default_rng(4).permutation(200).astype(np.int32)).to(gpu)
    outs = []
    for order in (None, perm):
        d = b.to_device(gpu)
        lam = torch.ones((nv, 200), dtype=torch.float64, device=gpu)
        code, ierr, istat, rstat = solver.rosenbrock_adj(
            b.mech, d["C"], d["temp"], d["cfactor"], tstart=d["tstart"], tend=d["tend"],
            atol=b.atol, rtol=b.rtol, Lambda=lam, ICNTRL=_ic(4), sun=d["sun"],
            sun_mode=b.sun_mode, order=order)
        torch.cuda.synchronize()
        assert code == 0 and bool((ierr == 1).all())
        outs.append((d["C"].cpu().numpy(), lam.cpu().numpy(), istat.cpu().numpy()))
    for x, y in zip(*outs):
        assert np.array_equal(x, y)
