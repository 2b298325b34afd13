"""The F90 integrator on the GPU (kpp_integrate_f90_dev, INTEGRATE_F90_BATCH)
against the oracle's restatement of int/rosenbrock.f90 (tests/test_f90_oracle.py
pins that restatement).

Bars:
 * with the oracle's rate constants supplied: every cell has the oracle's
   IERR and its eight ISTATUS counters, RSTATUS(1:3) = Texit, Hexit, Hnew
   agree to 1e-12 / 1e-7 / 1e-7 (relative), and species above ATOL/RTOL to
   1e-10 -- all ICNTRL(3) methods (Rang3 included), vector and scalar
   tolerances, passive species, ICNTRL(16) clipping, plain, persistent and
   register kernels;
 * with device-computed rates (fixed SUN, time-of-day SUN, rates held at
   their TIN values): same IERR everywhere, the oracle's step counts for at
   least 99 % of cells, 1e-8 (saprc99) or 1e-5 (small_strato, device cos in
   Update_SUN) for species above ATOL/RTOL;
 * the linear6 closed-form scenarios of the CPU test on the device;
 * parameter errors -1/-2/-5 in every cell; the cooperative kernel and mixed
   precision are argument errors;
 * INTEGRATE_F90_BATCH (host buffers) equals kpp_integrate_f90_dev bit for bit;
 * the H211b flavour (kpp_integrate_h211b_dev, int/rosenbrock_h211b_qssa.f90):
   the linear6 closed-form scenarios of the CPU test (counters exact, step
   sizes to 1e-7 as there), and on saprc99 with the oracle's rate constants
   the oracle's IERR everywhere, its counters in at least 99 % of cells
   (the unclamped filter carries a one-ULP difference of the device's pow
   into every later step size, so a cell can meet a decision boundary),
   RSTATUS to 1e-7 and species above ATOL/RTOL to 1e-9 in those cells, 1e-6
   in all; the QSSA options and passive species are argument errors.
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


def _ic(**kw):
    ic = np.zeros(20, np.int32)
    for k, v in kw.items():
        ic[int(k[1:]) - 1] = v
    return ic


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


def _gpu(b, dev, ICNTRL=None, RCNTRL=None, rconst=None, passive=None, atol=None, rtol=None,
         **kw):
    import torch
    d = b.to_device(dev)
    code, ierr, istat, rstat = solver.integrate_f90(
        b.mech, d["C"], d["temp"], d["cfactor"], tin=d["tstart"], tout=d["tend"],
        atol=b.atol if atol is None else atol, rtol=b.rtol if rtol is None else rtol,
        ICNTRL=ICNTRL, RCNTRL=RCNTRL, sun=d["sun"], passive=passive,
        rconst=None if rconst is None else _t(rconst, dev), **kw)
    torch.cuda.synchronize()
    return code, {"C": d["C"].cpu().numpy(), "ierr": ierr.cpu().numpy(),
                  "istat": istat.cpu().numpy(), "rstat": rstat.cpu().numpy()}


def _rel(a, b, nv, floor):
    big = np.abs(b[:nv]) > floor
    r = np.abs(a[:nv] - b[:nv]) / np.maximum(np.abs(b[:nv]), 1e-300)
    return float(r[big].max()) if big.any() else 0.0


def _exact_bar(g, o, nv, floor):
    assert np.array_equal(g["ierr"], o["ierr"])
    assert np.array_equal(g["istat"], o["istat"])
    for k, tol in ((0, 1e-12), (1, 1e-7), (2, 1e-7)):
        assert np.allclose(g["rstat"][k], o["rstat"][k], rtol=tol, atol=0.0), k
    assert _rel(g["C"], o["C"], nv, floor) <= 1e-10


def _vector_tol(b, nv):
    b.atol = np.full(nv, float(np.atleast_1d(b.atol)[0]))
    b.rtol = np.full(nv, float(np.atleast_1d(b.rtol)[0]))
    return b


@pytest.mark.parametrize("code", [0, 1, 2, 3, 5, 6, 7])
def test_rates_given_every_method(gpu, code):
    b = workloads.saprc99_uneven(ncells=256)
    nv = ref.sizes("saprc99")["NVAR"]
    R = _oracle_rates(b)
    ic = _ic(i2=1, i3=code, i15=1)
    o = ref.integrate_f90(b, ic, rconst=R)
    variant = solver.PERSISTENT if code in (0, 6) else solver.PLAIN
    c, g = _gpu(b, gpu, ic, rconst=R, variant=variant)
    assert c == 0 and (o["ierr"] == 1).all()
    _exact_bar(g, o, nv, 1.0 / 1e-3)


def test_vector_tolerances_passive_and_clipping(gpu):
    b = _vector_tol(workloads.saprc99_uneven(ncells=256), ref.sizes("saprc99")["NVAR"])
    nv = len(b.atol)
    rng = np.random.default_rng(5)
    b.atol = 10.0 ** rng.uniform(-1, 1, nv)
    pas = (rng.uniform(size=nv) < 0.2).astype(np.int32)
    R = _oracle_rates(b)
    ic = _ic(i3=0, i15=1, i16=1)
    rc = np.zeros(20)
    rc[[0, 3, 4]] = [1e-9, 0.3, 4.0]  # Hmin, FacMin, FacMax
    o = ref.integrate_f90(b, ic, rc, passive=pas, rconst=R)
    c, g = _gpu(b, gpu, ic, rc, rconst=R, passive=pas, variant=solver.PERSISTENT)
    assert c == 0
    _exact_bar(g, o, nv, 1.0 / 1e-3)
    assert (g["C"][:nv] >= 0.0).all()
    # the passive set changes the integration
    o2 = ref.integrate_f90(b, ic, rc, rconst=R)
    assert not np.array_equal(o2["istat"], o["istat"])


def test_register_kernel_small_strato(gpu):
    b = workloads.small_strato_tod(ncells=300)
    nv = ref.sizes("small_strato")["NVAR"]
    for code in (2, 6):
        ic = _ic(i2=1, i3=code)  # ICNTRL(15) = 5: Update_SUN and Update_RCONST
        o = ref.integrate_f90(b, ic)
        c, g = _gpu(b, gpu, ic, variant=solver.REGISTER)
        _, gp = _gpu(b, gpu, ic, variant=solver.PLAIN)
        assert c == 0 and (g["ierr"] == 1).all() and (o["ierr"] == 1).all()
        for k in ("C", "ierr", "istat", "rstat"):
            assert np.array_equal(g[k], gp[k]), k
        same = (g["istat"] == o["istat"]).all(axis=0)
        assert same.mean() >= 0.99
        assert _rel(g["C"], o["C"], nv, 1.0 / 1e-3) <= 1e-5


@pytest.mark.parametrize("opt", [1, -1, 4])
def test_device_rates(gpu, opt):
    """ICNTRL(15) = 1 (rates at every call, fixed SUN), -1 and 4 (rates held
    at their TIN values)"""
    b = workloads.saprc99_uneven(ncells=512)
    nv = ref.sizes("saprc99")["NVAR"]
    ic = _ic(i2=1, i15=opt)
    o = ref.integrate_f90(b, ic)
    c, g = _gpu(b, gpu, ic, variant=solver.PERSISTENT)
    assert c == 0
    assert np.array_equal(g["ierr"], o["ierr"]) and (g["ierr"] == 1).all()
    same = (g["istat"] == o["istat"]).all(axis=0)
    assert same.mean() >= 0.99
    assert _rel(g["C"], o["C"], nv, 1.0 / 1e-3) <= 1e-8


def test_linear6_closed_form_on_device(gpu):
    from tests.test_f90_oracle import SCEN, _model_f90
    from tests.test_step_control import _batch, _x1_rates
    for code, lam, h0 in SCEN:
        b = _batch(np.array([1.0e6]), 0.0, 3600.0)
        rc = np.zeros(20)
        rc[2] = h0
        ic = _ic(i2=1, i3=code)
        R = _x1_rates(lam)
        c, g = _gpu(b, gpu, ic, rc, rconst=R)
        o = ref.integrate_f90(b, ic, rc, rconst=R)
        w = _model_f90(code, lam, 1.0e6, 0.0, 3600.0, h0, 1.0, 1e-3)
        assert c == 0 and np.array_equal(g["istat"], o["istat"])
        st = g["istat"][:, 0]
        assert (st[2], st[3], st[4]) == (w["Nstp"], w["Nacc"], w["Nrej"])
        assert g["rstat"][1, 0] == pytest.approx(w["Hexit"], rel=1e-9)
        assert g["rstat"][2, 0] == pytest.approx(w["Hnew"], rel=1e-9)


def test_errors(gpu):
    b = workloads.saprc99_uneven(ncells=64)
    for ic, want in ((_ic(i3=8), -2), (_ic(i4=-1), -1)):
        c, g = _gpu(b, gpu, ic)
        assert c == want and (g["ierr"] == want).all()
    c, g = _gpu(b, gpu, _ic(i2=1), rtol=np.array([1.0]))
    assert c == -5 and (g["ierr"] == -5).all()
    with pytest.raises(solver.KppError):
        _gpu(b, gpu, _ic(i2=1), variant=solver.COOP)
    from tests.test_step_control import _batch
    bn = _batch(np.array([1.0e6, np.nan]), 0.0, 3600.0)
    R = np.zeros((12, 2))
    R[6] = 1e-3
    c, g = _gpu(bn, gpu, _ic(i2=1), rconst=R)
    assert list(g["ierr"]) == [1, -9]


def test_integrate_f90_batch_host_api(gpu):
    """INTEGRATE_F90_BATCH on host buffers equals the device entry point"""
    b = _vector_tol(workloads.saprc99_uneven(ncells=200, seed=11), ref.sizes("saprc99")["NVAR"])
    L = solver.library()
    mech = b.mech
    n, nsp = b.ncells, b.C.shape[0]
    ctx = ctypes.c_void_p()
    assert L.kpp_ctx_create(ctypes.byref(ctx), mech.encode(), n, 4, solver_flag_persistent()) == 0
    try:
        f = lambda a: np.ascontiguousarray(a, np.float64)
        C, temp, cf, sun = f(b.C), f(b.temp), f(b.cfactor), f(b.sun)
        assert L.kpp_set_state(ctx, C.ctypes.data, temp.ctypes.data, cf.ctypes.data,
                               sun.ctypes.data) == 0
        assert L.kpp_ctx_set_tol(ctx, f(b.atol).ctypes.data, f(b.rtol).ctypes.data, 1) == 0
        pas = np.zeros(len(b.atol), np.int32)
        pas[::7] = 1
        assert L.kpp_ctx_set_passive(ctx, pas.ctypes.data) == 0
        ic = _ic(i3=6, i15=1)
        ist = np.zeros((8, n), np.int32)
        rst = np.zeros((3, n))
        ier = np.zeros(n, np.int32)
        t0, t1 = float(b.tstart[0]), float(b.tend[0])
        assert L.INTEGRATE_F90_BATCH(ctx, t0, t1, ic.ctypes.data, None, ist.ctypes.data,
                                     rst.ctypes.data, ier.ctypes.data) == 0
        Cout = np.zeros_like(C)
        assert L.kpp_get_state(ctx, Cout.ctypes.data, None, None, None) == 0
    finally:
        L.kpp_ctx_destroy(ctx)
    b.tstart[:] = t0
    b.tend[:] = t1
    c, g = _gpu(b, gpu, ic, passive=pas, variant=solver.PERSISTENT)
    assert c == 0
    assert np.array_equal(Cout, g["C"]) and np.array_equal(ist, g["istat"])
    assert np.array_equal(rst, g["rstat"]) and np.array_equal(ier, g["ierr"])


def test_h211b_linear6_closed_form(gpu):
    from tests.test_f90_oracle import SCEN_H211B, _model_f90
    from tests.test_step_control import _batch, _x1_rates
    for code, lam, h0, bb, kk in SCEN_H211B:
        b = _batch(np.array([1.0e6]), 0.0, 3600.0)
        rc = np.zeros(20)
        rc[[2, 14, 15]] = [h0, bb, kk]
        ic = _ic(i2=1, i3=code)
        R = _x1_rates(lam)
        c, g = _gpu(b, gpu, ic, rc, rconst=R, h211b=True)
        o = ref.integrate_f90(b, ic, rc, rconst=R, h211b=True)
        w = _model_f90(code, lam, 1.0e6, 0.0, 3600.0, h0, 1.0, 1e-3, h211b=(bb, kk))
        assert c == 0 and np.array_equal(g["istat"], o["istat"]), code
        st = g["istat"][:, 0]
        assert (st[2], st[3], st[4]) == (w["Nstp"], w["Nacc"], w["Nrej"])
        assert g["rstat"][1, 0] == pytest.approx(w["Hexit"], rel=1e-7)
        assert g["rstat"][2, 0] == pytest.approx(w["Hnew"], rel=1e-7)
        assert g["C"][0, 0] == pytest.approx(o["C"][0, 0], rel=1e-8)


@pytest.mark.parametrize("code,bb,kk", [(0, 2.0, 3.0), (5, 2.0, 4.0), (6, 4.0, 3.0)])
def test_h211b_rates_given(gpu, code, bb, kk):
    b = workloads.saprc99_uneven(ncells=256)
    nv = ref.sizes("saprc99")["NVAR"]
    R = _oracle_rates(b)
    ic = _ic(i2=1, i3=code, i15=1)
    rc = np.zeros(20)
    rc[[14, 15]] = [bb, kk]
    o = ref.integrate_f90(b, ic, rc, rconst=R, h211b=True)
    c, g = _gpu(b, gpu, ic, rc, rconst=R, h211b=True, variant=solver.PERSISTENT)
    assert c == 0 and (o["ierr"] == 1).all()
    assert np.array_equal(g["ierr"], o["ierr"])
    same = (g["istat"] == o["istat"]).all(axis=0)
    assert same.mean() >= 0.99
    # the filter compounds the pow of two error estimates: step sizes to 1e-4
    for k, tol in ((0, 1e-12), (1, 1e-4), (2, 1e-4)):
        assert np.allclose(g["rstat"][k][same], o["rstat"][k][same], rtol=tol, atol=0.0), k
    assert _rel(g["C"][:, same], o["C"][:, same], nv, 1.0 / 1e-3) <= 1e-9
    assert _rel(g["C"], o["C"], nv, 1.0 / 1e-3) <= 1e-6
    # the filter changes the integration
    p = ref.integrate_f90(b, ic, rconst=R)
    assert not np.array_equal(p["istat"], o["istat"])


def test_h211b_errors(gpu):
    b = workloads.saprc99_uneven(ncells=64)
    with pytest.raises(solver.KppError):
        _gpu(b, gpu, _ic(i2=1, i18=1), h211b=True)
    with pytest.raises(solver.KppError):
        _gpu(b, gpu, _ic(i2=1), passive=np.zeros(ref.sizes("saprc99")["NVAR"], np.int32),
             h211b=True)
    c, g = _gpu(b, gpu, _ic(i3=8), h211b=True)
    assert c == -2 and (g["ierr"] == -2).all()


def solver_flag_persistent():
    return 0x10  # KPP_CTX_PERSISTENT
