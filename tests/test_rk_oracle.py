"""The oracle's restatement of KPP's C fully implicit Runge-Kutta integrator
(int/runge_kutta.c) pinned on the CPU.

* The four methods' tables (int/runge_kutta.c:1369-1863) against the theory
  of their families (Hairer & Wanner, Solving ODEs II, IV.5): c_i the row
  sums of A; the quadrature order B(p) (Radau-2A and Radau-1A 5, Lobatto-3C
  4, Gauss 6) and the simplifying assumptions C(q), D(r) each family
  satisfies (Radau-2A C(3) D(2), Radau-1A C(2) D(3), Lobatto-3C C(2) D(2),
  Gauss C(3) D(3)); and what the integrator derives from them: D = b A^-1
  (the new solution from Z), T^-1 A^-1 T = diag(gamma, [[alpha, -beta],
  [beta, alpha]]) with TinvAinv = T^-1 A^-1 (the transformed stage system of
  RK_Solve), the classical error estimator E = (bhat - (0, b)) / gamma in
  terms of (f0, Z) where the reference prints bhat, and for Radau-2A
  (no bhat printed) Hairer-Wanner's RADAU5 estimator E[1:]/E[0] =
  (-(13 + 7 sqrt 6)/3, (-13 + 7 sqrt 6)/3, -1/3) with E[0] = b0; the SDIRK
  estimator's Bgam a quadrature of order 3 on (0, c, 1) with Bgam[4] =
  1/gamma and Theta = Bgam[1:4] A^-1; b0 chosen by ICNTRL[9].
* KppDecompCmplxR/KppSolveCmplxR (util/sutil.c:28-87) against numpy's dense
  complex solve on saprc99's LU pattern, and the reference's singularity
  test (a zero real or imaginary diagonal).
* Whole integrations: y' = lam*y on linear6 against exp(lam t) for every
  method and both error estimators (Radau-1A with the SDIRK estimator
  excepted: its printed SDIRK stage reproduces the method's own solution on
  linear problems, the estimate is zero and the steps grow unchecked -- the
  reference's behaviour, shown from the tables); small_strato with
  time-of-day SUN against scipy's Radau at rtol 1e-12; the counters'
  structure.
* The step sequence on y' = lam*y against an independent model of the
  algorithm (simplified Newton on the 3-stage system written with dense
  complex algebra, the estimators and Gustafsson's controller from their
  definitions): counters exact, RSTATUS to 1e-9.
* The reference's own behaviour kept: the parameter checks run in order and
  the last failing one's code is returned; a singular first matrix halves
  the step and is counted; five in a row are only reported (-12 printed by
  the reference), the integration goes on.
"""
import numpy as np
import pytest

from kpp_amd import workloads
from oracle import ref
from tests.test_step_control import NV, _batch, _rates, _x1_rates

# ICNTRL[2] -> (name, quadrature order p, C(q), D(r))
FAMILY = {1: ("Radau-2A", 5, 3, 2), 2: ("Lobatto-3C", 4, 2, 2), 3: ("Gauss", 6, 3, 3),
          4: ("Radau-1A", 5, 2, 3)}


@pytest.fixture(scope="module", autouse=True)
def _have_linear6():
    try:
        ref.sizes("linear6")
    except (FileNotFoundError, OSError):
        pytest.skip("oracle for linear6 not built")


@pytest.mark.parametrize("code", [1, 2, 3, 4])
@pytest.mark.parametrize("sdirk_error", [0, 1])
def test_tables_against_theory(code, sdirk_error):
    name, p, q, r = FAMILY[code]
    m = ref.rk_method(code, sdirk_error)
    A, b, c = m["A"], m["B"], m["C"]
    assert np.allclose(A.sum(1), c, atol=1e-15)
    for k in range(1, p + 1):
        assert abs(b @ c ** (k - 1) - 1 / k) < 1e-15, (name, "B", k)
    assert abs(b @ c ** p - 1 / (p + 1)) > 1e-4  # exactly order p
    for k in range(1, q + 1):
        assert np.abs(A @ c ** (k - 1) - c ** k / k).max() < 1e-15, (name, "C", k)
    assert np.abs(A @ c ** q - c ** (q + 1) / (q + 1)).max() > 1e-4
    for k in range(1, r + 1):
        assert np.abs((b * c ** (k - 1)) @ A - b * (1 - c ** k) / k).max() < 1e-15, (name, "D", k)
    Ai = np.linalg.inv(A)
    assert np.allclose(m["D"], b @ Ai, atol=1e-14)
    T = m["T"]
    blk = np.array([[m["Gamma"], 0, 0], [0, m["Alpha"], -m["Beta"]], [0, m["Beta"], m["Alpha"]]])
    assert np.allclose(np.linalg.inv(T) @ Ai @ T, blk, atol=1e-13)
    assert np.allclose(m["TinvAinv"], np.linalg.inv(T) @ Ai, atol=2e-14)
    ev = np.sort_complex(np.linalg.eigvals(Ai))
    assert np.allclose(ev, np.sort_complex([m["Gamma"], m["Alpha"] + 1j * m["Beta"],
                                            m["Alpha"] - 1j * m["Beta"]]), atol=1e-13)
    # estimators
    E = m["E"]
    if code == 1:
        s6 = np.sqrt(6.0)
        assert E[0] == (0.02 if sdirk_error else 0.05)
        assert np.allclose(E[1:] / E[0], [-(13 + 7 * s6) / 3, (-13 + 7 * s6) / 3, -1 / 3],
                           rtol=1e-15)
        assert m["ELO"] == 4.0
    else:
        bh = m["Bhat"]
        assert bh[0] == {2: (0.2 if sdirk_error else 0.5), 3: 0.1, 4: 0.1}[code]
        w = np.concatenate([[E[0]], E[1:] @ A])
        assert np.allclose(w, (bh - np.concatenate([[0.0], b])) / m["Gamma"], atol=1e-15)
    Bg = m["Bgam"]
    assert Bg[4] == pytest.approx(1 / m["Gamma"], rel=1e-15)
    assert np.allclose(m["Theta"], Bg[1:4] @ Ai, atol=1e-14)
    cg = np.concatenate([[0.0], c, [1.0]])
    for k in (1, 2, 3):
        assert abs(Bg @ cg ** (k - 1) - 1 / k) < 1e-15, (name, "Bgam", k)


def test_unbuilt_methods():
    assert ref.rk_method(5) is None  # Lobatto-3A, "not yet implemented"
    assert ref.rk_method(6) is None and ref.rk_method(-1) is None
    assert ref.rk_method(0)["kind"] == ref.rk_method(1)["kind"] == 1


def _pattern(mech):
    icol = np.zeros(ref.sizes(mech)["LU_NONZERO"], np.int32)
    nv = ref.sizes(mech)["NVAR"]
    crow = np.zeros(nv + 1, np.int32)
    diag = np.zeros(nv + 1, np.int32)
    ref.lib(mech).ros_ref_lu_pattern(ref._p(icol), ref._p(crow), ref._p(diag))
    return icol, crow, diag


def test_complex_lu_against_dense_solve(kpp_home):
    """values on saprc99's pattern (fill-in entries zero), a dominant complex
    diagonal: the factorisation and solve agree with numpy to 1e-12"""
    mech = "saprc99"
    L = ref.lib(mech)
    icol, crow, diag = _pattern(mech)
    nv = len(crow) - 1
    rng = np.random.default_rng(4)
    JR = np.zeros(len(icol))
    JI = np.zeros(len(icol))
    M = np.zeros((nv, nv), complex)
    # the Jacobian's own entries: the pattern of Jac_SP (non-fill) is not
    # known here, so every pattern entry gets a value; the elimination is
    # closed on the pattern, so the result is the dense LU's
    for i in range(nv):
        for k in range(crow[i], crow[i + 1]):
            j = icol[k]
            v = rng.normal() + 1j * rng.normal() * 0.3
            if k == diag[i]:
                v = 40.0 + rng.uniform(1, 2) + 1j * rng.uniform(1, 3)
            JR[k], JI[k] = v.real, v.imag
            M[i, j] = v
    x = rng.normal(size=nv) + 1j * rng.normal(size=nv)
    rhs = M @ x
    gr, gi = JR.copy(), JI.copy()
    assert L.rk_ref_decomp_cmplx(ref._p(gr), ref._p(gi)) == 0
    xr, xi = rhs.real.copy(), rhs.imag.copy()
    L.rk_ref_solve_cmplx(ref._p(gr), ref._p(gi), ref._p(xr), ref._p(xi))
    assert np.allclose(xr + 1j * xi, x, rtol=1e-12, atol=1e-12)
    # the reference's test: row k+1 reported when its real (or imaginary)
    # diagonal is zero as its elimination starts
    for part in ("re", "im"):
        gr, gi = JR.copy(), JI.copy()
        (gr if part == "re" else gi)[diag[7]] = 0.0
        assert L.rk_ref_decomp_cmplx(ref._p(gr), ref._p(gi)) == 8, part


@pytest.mark.parametrize("code", [1, 2, 3, 4])
@pytest.mark.parametrize("sdirk_error", [0, 1])
def test_linear_decay(code, sdirk_error):
    lam, y0 = -2e-3, 1.0e6
    b = _batch(np.array([y0]), 0.0, 3600.0, rtol=1e-6, atol=1e-3)
    ic = np.zeros(20, np.int32)
    ic[1], ic[2], ic[9] = 1, code, sdirk_error
    out = ref.integrate_rk(b, ic, rconst=_x1_rates(lam))
    assert int(out["ierr"][0]) == 1
    if (code, sdirk_error) == (4, 1):
        # the reference's tables: Radau-1A's SDIRK stage reproduces its own
        # solution on linear problems (estimate identically zero), so the
        # steps grow by FacMax and the result is far from exp(lam t)
        m = ref.rk_method(4, 1)
        A, D, Bg, g = m["A"], m["D"], m["Bgam"], m["Gamma"]
        for z in (-0.1, -1.0, -10.0):
            Z = np.linalg.solve(np.eye(3) - z * A, z * A @ np.ones(3))
            Z4 = (Bg[0] * z + m["Theta"] @ Z + z / g) / (1 - z / g)
            assert abs(D @ Z - Z4) < 1e-14
        assert out["istat"][4, 0] == 0 and out["rstat"][1, 0] > 1000.0
    else:
        assert out["C"][0, 0] == pytest.approx(y0 * np.exp(lam * 3600.0), rel=1e-5)
    assert np.all(out["C"][1:NV, 0] == 0.0)
    nfun, njac, nstp, nacc, nrej, ndec, nsol, nsng = out["istat"][:, 0]
    assert nsng == 0 and nacc <= nstp and njac <= nstp and ndec <= nstp
    assert nsol >= nstp  # one RK_Solve per Newton iteration, at least one per step
    assert out["rstat"][0, 0] == pytest.approx(3600.0, rel=1e-14)


def test_small_strato_against_scipy_radau():
    from scipy.integrate import solve_ivp
    from tests.test_oracle import _rhs
    b = workloads.small_strato_tod(ncells=6)
    b.atol = np.array([1e-3])
    b.rtol = np.array([1e-8])
    ys = []
    for c in range(b.ncells):
        f = _rhs("small_strato", b, c)
        ys.append(solve_ivp(f, (b.tstart[c], b.tend[c]), b.C[:5, c], method="Radau",
                            rtol=1e-12, atol=1e-6).y[:, -1])
    for code in (1, 2, 3, 4):
        for se in (0, 1):
            if (code, se) == (4, 1):
                continue  # Radau-1A's SDIRK estimator: see test_linear_decay
            ic = np.zeros(20, np.int32)
            ic[1], ic[2], ic[9] = 1, code, se
            o = ref.integrate_rk(b, ic)
            assert (o["ierr"] == 1).all(), (code, se)
            # Gauss is not L-stable (|R(-inf)| = 1) and its classical
            # estimate (three solves with E1) damps the stiff error: O1D,
            # the stiff species, is off by up to 80 % there; the others hold
            floor = 1e3 if (code, se) == (3, 0) else 1.0
            for c in range(b.ncells):
                y = ys[c]
                big = np.abs(y) > floor
                assert np.allclose(o["C"][:5, c][big], y[big], rtol=2e-5), (code, se, c)


def test_parameter_checks_last_code_wins():
    b = _batch(np.array([1.0e6]), 0.0, 3600.0)
    R = _x1_rates(-2e-3)

    def run(ic_set=(), rc_set=(), rtol=None):
        ic = np.zeros(20, np.int32)
        ic[1] = 1
        rc = np.zeros(20)
        for k, v in ic_set:
            ic[k] = v
        for k, v in rc_set:
            rc[k] = v
        bb = _batch(np.array([1.0e6]), 0.0, 3600.0)
        if rtol is not None:
            bb.rtol = np.array([rtol])
        return int(ref.integrate_rk(bb, ic, rc, rconst=R)["ierr"][0])

    assert run([(2, 5)]) == -13 and run([(2, 7)]) == -13
    assert run([(3, -1)]) == -1 and run([(4, -1)]) == -2
    assert run(rc_set=[(4, 0.5)]) == -4 and run(rc_set=[(3, 2.0)]) == -4
    assert run(rc_set=[(6, 1.0)]) == -4 and run(rc_set=[(7, 1.5)]) == -5
    assert run(rc_set=[(8, 1e-20)]) == -6 and run(rc_set=[(9, 2.0)]) == -7
    assert run(rtol=0.0) == -8
    # several failing: the last check's code
    assert run([(2, 9), (3, -1)]) == -1
    assert run([(3, -1)], [(9, 2.0)], rtol=0.0) == -8
    # Hmin, Hmax, Hstart are taken with ABS: no -3
    assert run(rc_set=[(0, -1.0), (1, -100.0), (2, -10.0)]) == 1
    del b


def test_singular_first_matrix_halves_and_counts():
    """gamma/h - lam = 0 for the first matrix E1 (lam = gamma/h0 on X1): the
    step is halved, Nsng = 1, and the integration goes on"""
    m = ref.rk_method(1)
    h0 = 2.0 ** -8
    b = _batch(np.array([1.0]), 0.0, 2.0 ** -5)
    R = _rates(1, g=0.0)
    R[0, 0] = m["Gamma"] / h0
    ic = np.zeros(20, np.int32)
    ic[1] = 1
    rc = np.zeros(20)
    rc[2] = h0
    out = ref.integrate_rk(b, ic, rc, rconst=R)
    st = out["istat"][:, 0]
    assert int(out["ierr"][0]) == 1 and st[7] == 1
    assert np.isfinite(out["C"][0, 0])


def test_five_singular_matrices_are_reported_and_passed():
    """X1..X5 growth rates gamma/(h0 2^-k): every halving meets the next zero
    diagonal; from the fifth in a row the reference prints -12 and goes on
    halving, so the sixth matrix is regular and the integration ends with 1"""
    m = ref.rk_method(1)
    h0 = 2.0 ** -8
    b = _batch(np.array([1.0]), 0.0, h0)
    b.C[:5] = 1.0
    R = _rates(1, g=0.0)
    for k in range(5):
        R[k, 0] = m["Gamma"] / (h0 * 2.0 ** -k)
    ic = np.zeros(20, np.int32)
    ic[1] = 1
    rc = np.zeros(20)
    rc[2] = h0
    out = ref.integrate_rk(b, ic, rc, rconst=R)
    st = out["istat"][:, 0]
    assert int(out["ierr"][0]) == 1 and st[7] == 5
    assert np.all(np.isfinite(out["C"][:, 0]))


def _model_rk(code, se, lam, y0, t0, t1, hstart, atol, rtol, n=NV, mistake=None):
    """RK_Integrator on y' = lam*y (X1; the other species zero, counted in
    the norm's N), written from the method's definition with dense algebra:
    the stage system (I - h lam A) Z = h lam A 1 y solved by simplified
    Newton whose matrix is that of the step H_E of the last factorisation
    (A^-1/H_E - lam), starting from zero or the previous step's collocation
    polynomial; the SDIRK stage Z4 = G + (h/gamma) f(y + Z4) or the classical
    estimate (gamma/H_E - lam)^-k (E0 f0 + E.Z/h); Gustafsson's controller."""
    m = ref.rk_method(code, se)
    A, C, D, E, Bg, Th, g = m["A"], m["C"], m["D"], m["E"], m["Bgam"], m["Theta"], m["Gamma"]
    Ai = np.linalg.inv(A)
    facmin, facmax, facrej, facsafe, thetamin, ntol, qmin, qmax, maxit = (
        0.2, 6.0 if mistake == "facmax6" else 8.0, 0.1, 0.9, 1e-3, 3e-2, 1.0, 1.2, 8)
    rnd = 2.0 ** -52
    nsolves = {1: 1, 2: 1, 3: 3, 4: 2}[code]  # classical estimate: solves with E1
    y, t = y0, t0
    hmax = abs(t1 - t0)
    h = min(max(0.0, abs(hstart)), hmax)
    if h <= 10 * rnd:
        h = 1e-6
    hold = h
    if t + h * 1.0001 - t1 >= 0:
        h = t1 - t
    st = dict(Nfun=0, Njac=0, Nstp=0, Nacc=0, Nrej=0, Ndec=0, Nsol=0)
    reject = skipjac = skiplu = False
    first = True
    hacc = errold = 0.0
    theta = 0.0
    HE = None
    Zold = None
    rs = (0.0, 0.0, 0.0)

    def nrm(v, yy):
        return max(np.sqrt((v / (atol + rtol * abs(yy))) ** 2 / n), 1e-10)

    while (t1 - t) - rnd > 0:
        f0 = lam * y
        st["Nfun"] += 1
        if not skiplu:
            if not skipjac:
                st["Njac"] += 1
            st["Ndec"] += 1
            HE = h
        st["Nstp"] += 1
        Mi = np.linalg.inv(Ai / HE - lam * np.eye(3))
        if first:
            Z = np.zeros(3)
        else:
            p = np.polyfit(C, Zold, 2)
            Z = np.polyval(p, 1 + C * h / hold) - Zold[2]
        done, fac, inc_old = False, 0.5, 0.0
        it = 0
        for it in range(1, maxit + 1):
            R = Z - h * A @ (lam * (y + Z))
            dZ = Mi @ (Ai @ R / h)
            st["Nsol"] += 1
            inc = np.sqrt(sum(nrm(v, y) ** 2 for v in dZ) / 3)
            if it == 1:
                theta, rate = thetamin, 2.0
            else:
                theta = inc / inc_old
                if theta >= 0.99:
                    break
                rate = theta / (1 - theta)
                if it < maxit:
                    pred = inc * theta ** (maxit - it) / (1 - theta)
                    if pred >= ntol:
                        fac = 0.8 * min(10.0, pred / ntol) ** (-1.0 / (1 + maxit - it))
                        break
            inc_old = max(inc, rnd)
            Z = Z - dZ
            done = rate * inc <= ntol
            if done:
                break
        it_main = it if done else maxit + 1
        if done and se:
            Z4 = Z[2]
            G = Bg[0] * h * f0 + Th @ Z
            done, fac = False, 0.5
            for it4 in range(1, maxit + 1):
                st["Nfun"] += 1
                d4 = (lam * (y + Z4) - g / h * Z4 + g / h * G) / (g / HE - lam)
                inc = nrm(d4, y)
                if it4 == 1:
                    rate = 2.0
                else:
                    th = inc / inc_old
                    if th >= 0.99:
                        break
                    rate = th / (1 - th)
                    pred = inc * th ** (maxit - it4) / (1 - th)
                    if pred >= ntol:
                        fac = 0.8 * min(10.0, pred / ntol) ** (-1.0 / (1 + maxit - it4))
                        break
                inc_old = inc
                Z4 = Z4 + d4
                done = rate * inc <= ntol
                if done:
                    break
        if not done:
            h = fac * h
            reject = skipjac = True
            skiplu = False
            continue
        if se:
            err = nrm(D @ Z - Z4, y)
        else:
            F2 = E[1:] @ Z / h
            x = E[0] * f0 + F2
            for _ in range(nsolves):
                x = x / (g / HE - lam)
            err = nrm(x, y)
            if err >= 1 and (first or reject):
                st["Nfun"] += 1
                err = nrm((lam * (y + x) + F2) / (g / HE - lam), y)
        fac = err ** (-1.0 / m["ELO"]) * min(facsafe, (1 + 2 * maxit) / (it_main + 2 * maxit))
        fac = min(facmax, max(facmin, fac))
        hnew = fac * h
        if err < 1:
            first = False
            st["Nacc"] += 1
            if st["Nacc"] > 1 and mistake != "no_gustafsson":
                fg = facsafe * (h / hacc) * (err * err / errold) ** -0.25
                fac = min(fac, min(facmax, max(facmin, fg)))
                hnew = fac * h
            hacc, errold = h, max(1e-2, err)
            hold = h
            t = t + h
            y = y + D @ Z
            Zold = Z.copy()
            rs = (t, h, hnew)
            hnew = min(max(abs(hnew), 0.0), hmax)
            if reject and mistake != "no_reject_cap":
                hnew = min(hnew, abs(h))
            reject = False
            if t + hnew / qmin - t1 >= 0:
                h = t1 - t
            else:
                skiplu = (theta <= thetamin and qmin <= hnew / h <= qmax
                          and mistake != "no_lu_reuse")
                if not skiplu:
                    h = hnew
            skipjac = False
        else:
            h = facrej * h if (first or reject) else hnew
            reject = skipjac = True
            skiplu = False
            if st["Nacc"] >= 1:
                st["Nrej"] += 1
    return y, st, rs


# (ICNTRL[2], ICNTRL[9], lambda, RCNTRL[2]): rejected large first steps,
# the default 1e-6 start, growth, both estimators
SCEN_RK = [(1, 0, -2e-3, 3600.0), (1, 1, -2e-3, 3600.0), (2, 0, -5e-3, 0.0), (2, 1, -2e-3, 900.0),
           (3, 0, -2e-3, 3600.0), (3, 1, 3e-3, 0.0), (4, 0, -1e-3, 1800.0), (1, 1, 4e-3, 60.0)]


@pytest.mark.parametrize("code,se,lam,h0", SCEN_RK)
def test_step_sequence_matches_model(code, se, lam, h0):
    y0, atol, rtol = 1.0e6, 1.0, 1e-4
    b = _batch(np.array([y0]), 0.0, 3600.0, rtol=rtol, atol=atol)
    ic = np.zeros(20, np.int32)
    ic[1], ic[2], ic[9] = 1, code, se
    rc = np.zeros(20)
    rc[2] = h0
    out = ref.integrate_rk(b, ic, rc, rconst=_x1_rates(lam))
    y, st, rs = _model_rk(code, se, lam, y0, 0.0, 3600.0, h0, atol, rtol)
    got = out["istat"][:, 0]
    want = [st[k] for k in ("Nfun", "Njac", "Nstp", "Nacc", "Nrej", "Ndec", "Nsol")]
    assert int(out["ierr"][0]) == 1
    assert list(got[:7]) == want, (list(got[:7]), want)
    assert got[7] == 0
    for k in range(3):
        assert out["rstat"][k, 0] == pytest.approx(rs[k], rel=1e-9), k
    assert out["C"][0, 0] == pytest.approx(y, rel=1e-9)


def test_scenarios_exercise_the_controller():
    """the scenarios reject (FacRej, Gustafsson), reuse the LU (Qmin..Qmax)
    and run more than one Newton iteration"""
    rej = reuse = multi = 0
    for code, se, lam, h0 in SCEN_RK:
        _, st, _ = _model_rk(code, se, lam, 1.0e6, 0.0, 3600.0, h0, 1.0, 1e-4)
        rej += st["Nstp"] > st["Nacc"]
        reuse += st["Ndec"] < st["Nstp"]
        multi += st["Nsol"] > st["Nstp"]
    assert rej >= 3 and reuse >= 2 and multi >= 3


@pytest.mark.parametrize("mistake", ["facmax6", "no_reject_cap", "no_gustafsson", "no_lu_reuse"])
def test_planted_controller_mistakes_are_caught(mistake):
    """each plausible slip in the model (Rosenbrock's FacMax 6, growth after
    a rejection, no predictive controller, no LU reuse) makes at least one
    scenario disagree with the oracle"""
    caught = 0
    for code, se, lam, h0 in SCEN_RK:
        b = _batch(np.array([1.0e6]), 0.0, 3600.0, rtol=1e-4, atol=1.0)
        ic = np.zeros(20, np.int32)
        ic[1], ic[2], ic[9] = 1, code, se
        rc = np.zeros(20)
        rc[2] = h0
        out = ref.integrate_rk(b, ic, rc, rconst=_x1_rates(lam))
        _, st, rs = _model_rk(code, se, lam, 1.0e6, 0.0, 3600.0, h0, 1.0, 1e-4, mistake=mistake)
        want = [st[k] for k in ("Nfun", "Njac", "Nstp", "Nacc", "Nrej", "Ndec", "Nsol")]
        if (list(out["istat"][:7, 0]) != want or
                abs(out["rstat"][2, 0] - rs[2]) > 1e-9 * abs(rs[2])):
            caught += 1
    assert caught >= 1
