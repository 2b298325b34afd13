"""The oracle's F90 flavour (int/rosenbrock.f90) pinned on the CPU.

KPP's Fortran 90 integrator is a separate implementation from the C one;
gfortran is absent here, so nothing runs it.  What pins the restatement
(oracle/ros_ref.c, ros_ref_f90_batch):

* the ICNTRL(3) method map (int/rosenbrock.f90:321-340: 0 and 4 Rodas3,
  1 Ros2, 2 Ros3, 3 Ros4, 5 Rodas4, 6 Rang3, 7 Rodas3.1, anything else -2),
  with the Rosenbrock order conditions for Rang3 (a W-method of order 3 by
  Rang & Angermann, BIT 45 (2005), which with the exact Jacobian satisfies
  the order-3 conditions) and the two tables that the F90 file computes
  instead of printing (Ros2's g = 1 + 1/sqrt(2), Rodas3's C(6) = -8/3);
* INTEGRATE's merge of the caller's arrays (:57-80: ICNTRL(15) = 5 unless
  the caller gives a nonzero value; RCNTRL_U only where > 0) and the
  parameter checks' codes (:306-433);
* the step sequence of ros_Integrator (:492-686) on linear6 (one species
  with y' = lam*y, the others zero), predicted from the stability functions
  of the method and its embedded formula (test_step_control.py) and the F90
  controller as written in int/rosenbrock.f90: first step MIN(MAX(|Hmin|,
  |Hstart|), |Hmax|) with Hstart = 1e-5 by default (:378, :540-546), error
  norm sqrt(sum (Yerr/Scale)^2 / NonPassiveSpc_Count) floored at 1e-10
  (:716-728), and RSTATUS(1:3) = T, H and Hnew of the last accepted step
  (:651-662), not the C integrator's Hexit taken before the step is cut to
  Tend;
* the passive-species error norm (src/gen.c:3355-3384) by the same closed
  form with one species counted;
* ICNTRL(16) = 1 (:651-656): an accepted step whose closed-form value is
  negative is clipped to zero, and the rest of the integration follows the
  clipped value;
* the H211b flavour (int/rosenbrock_h211b_qssa.f90): the same step sequence
  with the H211b filter of :882-892 in place of the controller (Fac =
  (1/Err)^(1/(bb kk)) (1/Err_prev)^(1/(bb kk)) Fac_prev^(-1/bb), memory
  starting at 1 (:663-664) and updated on every attempt, no FacMin/FacMax
  bounds), selected only when RCNTRL(15) and RCNTRL(16) are both set
  (:490-514); a clamped factor or a filter without memory planted in the
  model disagrees with the oracle.
"""
import numpy as np
import pytest

from oracle import ref
from tests.test_oracle import _order_conditions, _standard_form
from tests.test_step_control import NV, _batch, _stability, _x1_rates

MECH = "linear6"


@pytest.fixture(scope="module", autouse=True)
def _have_linear6():
    try:
        ref.sizes(MECH)
    except (FileNotFoundError, OSError):
        pytest.skip("oracle for linear6 not built")


def _icntrl(**kw):
    """ICNTRL_U with 1-based Fortran indices: _icntrl(i3=6, i16=1)"""
    ic = np.zeros(20, np.int32)
    for k, v in kw.items():
        ic[int(k[1:]) - 1] = v
    return ic


def test_method_map():
    c = {k: ref.method(k) for k in range(1, 6)}
    c[7] = ref.method(7)
    f = {k: ref.method(k, f90=True) for k in range(9)}
    assert f[8] is None and ref.method(-1, f90=True) is None
    for k in (2, 3, 5, 7):  # same tables as the C integrator
        for key in ("A", "C", "M", "E", "Alpha", "Gamma"):
            assert np.array_equal(f[k][key], c[k][key]), (k, key)
    for k in (0, 4):
        assert f[k]["S"] == 4 and list(f[k]["NewF"]) == [1, 0, 1, 1]
        assert f[k]["C"][5] == -(8.0 / 3.0) and c[4]["C"][5] == -2.66666666666667
    g = 1.0 + 1.0 / np.sqrt(2.0)
    assert f[1]["Gamma"][0] == g and f[1]["A"][0] == 1.0 / g
    assert c[1]["Gamma"][0] == 1.70710678118655
    r = f[6]
    assert r["S"] == 4 and r["ELO"] == 3.0 and list(r["NewF"]) == [1, 1, 1, 1]


def test_rang3_order_conditions():
    m = ref.method(6, f90=True)
    alpha, Gam, b, bhat, g = _standard_form(m)
    # the tabled alpha_i and gamma_i are the row sums (15 printed digits)
    assert np.allclose(alpha.sum(1), m["Alpha"], atol=1e-12)
    assert np.allclose(Gam.sum(1), m["Gamma"], atol=1e-12)
    for q, r in _order_conditions(alpha, Gam, b, g).items():
        if q <= 3:
            assert np.abs(r).max() <= 1e-12, (q, r)
    # the embedded formula: order 2, not 3
    for q, r in _order_conditions(alpha, Gam, bhat, g).items():
        if q <= 2:
            assert np.abs(r).max() <= 1e-12, (q, r)
    assert np.abs(_order_conditions(alpha, Gam, bhat, g)[3]).max() > 1e-6


def test_integrate_merge():
    ic, rc = ref.f90_merge()
    assert ic[14] == 5 and np.count_nonzero(ic) == 1 and not rc.any()
    iu = _icntrl(i1=1, i3=6, i15=-1, i16=1)
    ru = np.zeros(20)
    ru[[0, 2, 4]] = [-1.0, 2.5, 3.0]  # a negative RCNTRL_U entry is not copied
    ic, rc = ref.f90_merge(iu, ru)
    assert list(ic[[0, 2, 14, 15]]) == [1, 6, -1, 1]
    assert rc[0] == 0.0 and rc[2] == 2.5 and rc[4] == 3.0


def test_parameter_errors():
    b = _batch(np.array([1.0e6, 2.0e6]), 0.0, 3600.0)
    R = np.repeat(_x1_rates(-2e-3), 2, axis=1)
    assert list(ref.integrate_f90(b, _icntrl(i3=8), rconst=R)["ierr"]) == [-2, -2]
    assert list(ref.integrate_f90(b, _icntrl(i4=-3), rconst=R)["ierr"]) == [-1, -1]
    b.rtol[:] = 1.0
    out = ref.integrate_f90(b, _icntrl(i2=1), rconst=R)
    assert list(out["ierr"]) == [-5, -5]
    assert not out["istat"].any() and not out["rstat"].any()


def _model_f90(code, lam, y0, t0, t1, hstart, atol, rtol, npc=NV, nonneg=False, hmin=0.0,
               fac=(0.2, 6.0, 0.1, 0.9), mistake=None, h211b=None):
    """ros_Integrator's step sequence on y' = lam*y (species X1, the others
    zero): npc species count in the error norm; h211b = (bb, kk) for the
    H211b filter."""
    R, E, m = _stability_f90(code)
    facmin, facmax, facrej, facsafe = fac
    roundoff = 2.0 ** -52
    hmax = abs(t1 - t0)
    dmin = 1e-6 if mistake == "c_start" else 1e-5
    hs = max(hmin, dmin) if hstart == 0.0 else min(abs(hstart), hmax)
    h = min(max(abs(hmin), abs(hs)), abs(hmax))
    y, t = y0, t0
    nstp = nacc = nrej = 0
    rej_last = rej_more = False
    texit = hexit = hnew_out = 0.0
    clipped = 0
    facold = errold = 1.0
    while t - t1 + roundoff <= 0.0:
        h_before_cut = h
        h = min(h, abs(t1 - t))
        while True:
            z = lam * h
            ynew = R(z) * y
            yerr = E(z) * y
            sc = atol + rtol * max(abs(y), abs(ynew))
            err = max(np.sqrt((yerr / sc) ** 2 / npc), 1e-10)
            if h211b:
                bb, kk = h211b
                f = (1.0 / err) ** (1.0 / (bb * kk)) * (1.0 / errold) ** (1.0 / (bb * kk)) \
                    * facold ** (-1.0 / bb)
                if mistake == "h211b_clamp":
                    f = min(facmax, max(facmin, f))
                if mistake != "h211b_nomem":
                    facold, errold = f, err
                hnew = h * f
            else:
                hnew = h * min(facmax, max(facmin, facsafe / err ** (1.0 / m["ELO"])))
            nstp += 1
            if err <= 1.0 or h <= hmin:
                nacc += 1
                if nonneg and ynew < 0.0:
                    ynew = 0.0
                    clipped += 1
                y = ynew
                t += h
                hnew = max(hmin, min(hnew, hmax))
                if rej_last:
                    hnew = min(hnew, h)
                hexit, hnew_out, texit = h, hnew, t
                if mistake == "c_hexit":
                    hexit = h_before_cut
                rej_last = rej_more = False
                h = hnew
                break
            if rej_more:
                hnew = h * facrej
            rej_more = rej_last
            rej_last = True
            h = hnew
            if nacc >= 1:
                nrej += 1
    return {"y": y, "Texit": texit, "Hexit": hexit, "Hnew": hnew_out, "Nstp": nstp,
            "Nacc": nacc, "Nrej": nrej, "clipped": clipped}


def _stability_f90(code):
    m = ref.method(code, f90=True)
    alpha, Gam, b, bhat, _ = _standard_form(m)
    B = alpha + Gam
    e = b - bhat
    one = np.ones(len(b))

    def R(z):
        return 1.0 + z * (b @ np.linalg.solve(np.eye(len(b)) - z * B, one))

    def E(z):
        return z * (e @ np.linalg.solve(np.eye(len(b)) - z * B, one))

    return R, E, m


# (ICNTRL(3), lambda, RCNTRL(3)): rejections in a row at the start, the
# default 1e-5 start, a growth rejected after acceptances, Rang3
SCEN = [(0, -2e-3, 3600.0), (2, -2e-3, 3600.0), (6, -2e-3, 3600.0), (1, -1e-3, 1800.0),
        (5, -2e-3, 0.0), (3, 5e-3, 3600.0), (7, -5e-2, 3600.0), (6, 5e-3, 0.0)]


def _check(out, want, code, tol_h=1e-9, tol_y=1e-10):
    m = ref.method(code, f90=True)
    st = out["istat"][:, 0]
    assert int(out["ierr"][0]) == 1
    assert (st[2], st[3], st[4]) == (want["Nstp"], want["Nacc"], want["Nrej"])
    newf = int(np.sum(m["NewF"][1:]))
    assert st[0] == want["Nacc"] * 2 + want["Nstp"] * newf  # Fcn0 + dF/dT per step
    assert st[1] == want["Nacc"] and st[5] == want["Nstp"] and st[6] == m["S"] * want["Nstp"]
    assert out["rstat"][0, 0] == pytest.approx(want["Texit"], rel=1e-13)
    assert out["rstat"][1, 0] == pytest.approx(want["Hexit"], rel=tol_h)
    assert out["rstat"][2, 0] == pytest.approx(want["Hnew"], rel=tol_h)
    assert out["C"][0, 0] == pytest.approx(want["y"], rel=tol_y, abs=1e-300)


@pytest.mark.parametrize("code,lam,h0", SCEN)
def test_step_sequence_matches_f90_controller(code, lam, h0):
    y0 = 1.0e6
    b = _batch(np.array([y0]), 0.0, 3600.0)
    rc = np.zeros(20)
    rc[2] = h0
    out = ref.integrate_f90(b, _icntrl(i2=1, i3=code), rc, rconst=_x1_rates(lam))
    want = _model_f90(code, lam, y0, 0.0, 3600.0, h0, 1.0, 1e-3)
    _check(out, want, code)
    assert np.all(out["C"][1:NV, 0] == 0.0)


def test_scenarios_exercise_rejections():
    """most scenarios reject (FacRej after two in a row, no growth after a
    rejection), and one rejects after its first acceptance"""
    w = [_model_f90(c, lam, 1.0e6, 0.0, 3600.0, h0, 1.0, 1e-3) for c, lam, h0 in SCEN]
    assert sum(x["Nstp"] - x["Nacc"] >= 2 for x in w) >= 4
    assert any(x["Nrej"] >= 1 for x in w)


@pytest.mark.parametrize("mistake", ["c_hexit", "c_start"])
def test_planted_c_rules_are_caught(mistake):
    """The C integrator's rules planted in the model -- Hexit as the step
    before the cut to Tend (int/rosenbrock.c:523), the 1e-6 default start --
    disagree with the oracle on at least one scenario."""
    caught = 0
    for code, lam, h0 in SCEN:
        b = _batch(np.array([1.0e6]), 0.0, 3600.0)
        rc = np.zeros(20)
        rc[2] = h0
        out = ref.integrate_f90(b, _icntrl(i2=1, i3=code), rc, rconst=_x1_rates(lam))
        w = _model_f90(code, lam, 1.0e6, 0.0, 3600.0, h0, 1.0, 1e-3, mistake=mistake)
        st = out["istat"][:, 0]
        if ((st[2], st[3], st[4]) != (w["Nstp"], w["Nacc"], w["Nrej"]) or
                abs(out["rstat"][1, 0] - w["Hexit"]) > 1e-9 * abs(out["rstat"][1, 0])):
            caught += 1
    assert caught >= 1


def test_passive_species_norm():
    """With X2..X6 passive the norm counts one species: err = |Yerr/Scale|,
    a factor sqrt(6) above the all-species norm, so the steps differ."""
    code, lam, h0 = 0, -2e-3, 3600.0
    b = _batch(np.array([1.0e6]), 0.0, 3600.0)
    rc = np.zeros(20)
    rc[2] = h0
    pas = np.array([0, 1, 1, 1, 1, 1])
    out = ref.integrate_f90(b, _icntrl(i2=1, i3=code), rc, passive=pas, rconst=_x1_rates(lam))
    want = _model_f90(code, lam, 1.0e6, 0.0, 3600.0, h0, 1.0, 1e-3, npc=1)
    _check(out, want, code)
    assert want["Nstp"] != _model_f90(code, lam, 1.0e6, 0.0, 3600.0, h0, 1.0, 1e-3)["Nstp"]


def _negative_step_case():
    """(method, lambda, h0) whose first accepted closed-form step is negative:
    R(h0*lam) < 0 with loose tolerances so that it is accepted"""
    for code in (2, 3, 1, 0):
        R, E, m = _stability_f90(code)
        for lam in (-1.0, -0.3, -0.1, -0.03):
            for h0 in (3600.0, 1000.0, 300.0, 100.0, 30.0, 10.0):
                if R(lam * h0) < -1e-3:
                    return code, lam, h0
    raise AssertionError("no negative stability value found")


@pytest.mark.parametrize("nonneg", [0, 1])
def test_nonnegative_clipping(nonneg):
    code, lam, h0 = _negative_step_case()
    y0, atol, rtol = 1.0, 1.0e6, 0.5
    b = _batch(np.array([y0]), 0.0, 3600.0, atol=atol, rtol=rtol)
    rc = np.zeros(20)
    rc[2] = h0
    out = ref.integrate_f90(b, _icntrl(i2=1, i3=code, i16=nonneg), rc, rconst=_x1_rates(lam))
    want = _model_f90(code, lam, y0, 0.0, 3600.0, h0, atol, rtol, nonneg=bool(nonneg))
    _check(out, want, code)
    if nonneg:
        assert want["clipped"] >= 1 and out["C"][0, 0] == 0.0
    else:
        assert out["C"][0, 0] != 0.0


# (ICNTRL(3), lambda, RCNTRL(3), bb, kk)
SCEN_H211B = [(0, -2e-3, 3600.0, 2.0, 3.0), (5, -2e-3, 0.0, 2.0, 4.0), (6, 5e-3, 0.0, 4.0, 3.0),
              (2, -2e-3, 3600.0, 2.0, 3.0), (1, -1e-3, 1800.0, 2.0, 2.0),
              (7, -5e-2, 3600.0, 3.0, 3.0)]


def _h211b_run(code, lam, h0, bb, kk):
    b = _batch(np.array([1.0e6]), 0.0, 3600.0)
    rc = np.zeros(20)
    rc[[2, 14, 15]] = [h0, bb, kk]
    return ref.integrate_f90(b, _icntrl(i2=1, i3=code), rc, rconst=_x1_rates(lam), h211b=True)


@pytest.mark.parametrize("code,lam,h0,bb,kk", SCEN_H211B)
def test_h211b_step_sequence(code, lam, h0, bb, kk):
    """Counters exact.  The filter has no clamp, so the rounding noise of the
    embedded error estimate (the cancellation in sum E_j K_j, large relative
    to errors near the 1e-10 floor) reaches every later step size at weight
    1/(bb kk): step sizes to 1e-7, the state to 1e-8 (a few 1e-9 seen)."""
    out = _h211b_run(code, lam, h0, bb, kk)
    want = _model_f90(code, lam, 1.0e6, 0.0, 3600.0, h0, 1.0, 1e-3, h211b=(bb, kk))
    _check(out, want, code, tol_h=1e-7, tol_y=1e-8)
    plain = _model_f90(code, lam, 1.0e6, 0.0, 3600.0, h0, 1.0, 1e-3)
    assert (want["Nstp"], want["Hnew"]) != (plain["Nstp"], plain["Hnew"])


@pytest.mark.parametrize("mistake", ["h211b_clamp", "h211b_nomem"])
def test_h211b_planted_mistakes_are_caught(mistake):
    caught = 0
    for code, lam, h0, bb, kk in SCEN_H211B:
        out = _h211b_run(code, lam, h0, bb, kk)
        w = _model_f90(code, lam, 1.0e6, 0.0, 3600.0, h0, 1.0, 1e-3, mistake=mistake,
                       h211b=(bb, kk))
        st = out["istat"][:, 0]
        if ((st[2], st[3], st[4]) != (w["Nstp"], w["Nacc"], w["Nrej"]) or
                abs(out["rstat"][2, 0] - w["Hnew"]) > 1e-7 * abs(out["rstat"][2, 0])):
            caught += 1
    assert caught >= 1


def test_h211b_needs_both_parameters():
    """bb or kk unset: the F90 controller (:490-514); RCNTRL(15:16) mean
    nothing to int/rosenbrock.f90"""
    b = _batch(np.array([1.0e6]), 0.0, 3600.0)
    R = _x1_rates(-2e-3)
    ic = _icntrl(i2=1, i3=0)
    rc = np.zeros(20)
    rc[2] = 3600.0
    base = ref.integrate_f90(b, ic, rc, rconst=R)
    for bb, kk, flag in ((2.0, 0.0, True), (0.0, 3.0, True), (2.0, 3.0, False)):
        r = rc.copy()
        r[[14, 15]] = [bb, kk]
        o = ref.integrate_f90(b, ic, r, rconst=R, h211b=flag)
        assert np.array_equal(o["istat"], base["istat"])
        assert np.array_equal(o["rstat"], base["rstat"]) and np.array_equal(o["C"], base["C"])
    with pytest.raises(ValueError):
        ref.integrate_f90(b, ic, rc, passive=np.zeros(NV, np.int32), rconst=R, h211b=True)


def test_f90_nan_cell_ends_with_minus9():
    b = _batch(np.array([1.0e6, np.nan]), 0.0, 3600.0)
    R = np.repeat(_x1_rates(-1e-3), 2, axis=1)
    out = ref.integrate_f90(b, _icntrl(i2=1), rconst=R)
    assert list(out["ierr"]) == [1, -9]
    assert out["istat"][2, 1] == 0  # ended before its first attempt counted


def test_saprc99_f90_runs_and_counts(kpp_home):
    """Rodas3 and Rang3 on saprc99 cells with fixed SUN (ICNTRL(15) = 1),
    vector tolerances: every cell ends at TOUT; the counters add up."""
    from kpp_amd import workloads
    b = workloads.saprc99_uneven(ncells=64)
    nv = ref.sizes("saprc99")["NVAR"]
    b.atol = np.full(nv, 1.0)
    b.rtol = np.full(nv, 1e-3)
    for code in (0, 6):
        m = ref.method(code, f90=True)
        newf = int(np.sum(m["NewF"][1:]))
        out = ref.integrate_f90(b, _icntrl(i3=code, i15=1))
        assert (out["ierr"] == 1).all()
        st = out["istat"].astype(np.int64)
        # per step: Fcn0, dF/dT and one Jacobian; per attempt: one LU, S
        # solves and the stages' new functions; no singular matrices
        assert (st[1] == st[3]).all() and (st[5] == st[2]).all() and (st[7] == 0).all()
        assert (st[6] == m["S"] * st[2]).all() and (st[2] >= st[3]).all()
        assert (st[0] == 2 * st[3] + newf * st[2]).all()
        assert np.allclose(out["rstat"][0], b.tend, rtol=1e-14)
        assert (out["rstat"][1] > 0).all() and (out["rstat"][2] > 0).all()
