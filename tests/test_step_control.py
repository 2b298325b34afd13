"""The oracle's step-size control pinned by closed forms (CPU).

Mechanism: linear6 (kpp_amd/mechs/linear6.kpp), six independent species
X_k with dX_k/dt = (g_k - d_k) X_k when the fixed species P is 1 and the
caller supplies the rate constants g_k (reactions G1..G6) and d_k (D1..D6).

For y' = lambda*y a Rosenbrock method gives, with z = h*lambda,
    y_new = R(z) y,     y_new - y_hat = Ehat(z) y,
    R(z)    = 1 + z b^T (I - z B)^{-1} 1,
    Ehat(z) = z (b - b_hat)^T (I - z B)^{-1} 1,    B = alpha + Gamma
(Hairer & Wanner, Solving ODEs II, IV.7: the stability function of a
Rosenbrock method, and the same for the embedded formula), with (alpha,
Gamma, b, b_hat) recovered from KPP's (A, C, M, E) tables as in
test_oracle.py.  The step sequence then follows from the controller as the
KPP manual states it (docs/source/num_methods/controlling_the_integrator.rst:
601-622: FacMin/FacMax bound the step change, FacRej is the decrease after
multiple rejections, FacSafe the safety factor) and Hairer & Wanner IV.8
(h_new = h * min(FacMax, max(FacMin, FacSafe * err^(-1/ELO))), no increase
right after a rejection), with KPP's documented outputs: Nrej counts
rejections except at the very beginning
(docs/source/num_methods/output_from_integrators.rst:95), and Hexit is the
step size the last step started from (int/rosenbrock.c:523, used as the
next call's Hstart, :250-253).

The scenarios are chosen so that each rule is exercised (at least two
rejections in a row, so FacRej acts; an acceptance right after a rejection
whose predicted step would have grown), which the tests assert about the
model before comparing, so that a wrong FacRej rule, step growth after a
rejection or a wrong Hexit fails.

Also here: the two bounds of the reject loop that the reference lacks (-9
for a NaN error estimate or a non-finite step, -10 for more than 100000
rejections in a row), and the singular-matrix path of ros_PrepareMatrix
(int/rosenbrock.c:686-737: halve H on an exactly zero diagonal, -8 after
more than five in a row), whose exact zeros linear6 can produce.
"""
import numpy as np
import pytest

from kpp_amd import workloads
from oracle import ref

MECH = "linear6"
NV = 6


@pytest.fixture(scope="module", autouse=True)
def _have_linear6():
    try:
        ref.sizes(MECH)
    except (FileNotFoundError, OSError):
        pytest.skip("oracle for linear6 not built")


def _standard_form(m):
    s = m["S"]
    A = np.zeros((s, s))
    C = np.zeros((s, s))
    for i in range(2, s + 1):
        for j in range(1, i):
            A[i - 1, j - 1] = m["A"][(i - 1) * (i - 2) // 2 + j - 1]
            C[i - 1, j - 1] = m["C"][(i - 1) * (i - 2) // 2 + j - 1]
    g = m["Gamma"][0]
    Gam = np.linalg.inv(np.diag(np.full(s, 1.0 / g)) - C)
    return A @ Gam, Gam, m["M"] @ Gam, m["E"] @ Gam


def _stability(code):
    m = ref.method(code)
    alpha, Gam, b, e = _standard_form(m)
    B = alpha + Gam
    one = np.ones(len(b))

    def R(z):
        return 1.0 + z * (b @ np.linalg.solve(np.eye(len(b)) - z * B, one))

    def E(z):
        return z * (e @ np.linalg.solve(np.eye(len(b)) - z * B, one))

    return R, E, m


def _model(code, lam, y0, t0, t1, hstart, atol, rtol, fac=(0.2, 6.0, 0.1, 0.9), hmin=0.0,
           mistake=None):
    """Step sequence of the documented controller on y' = lam*y (species X1,
    the other five species zero).  `mistake` plants one error in the model
    (for the test that the comparison catches it)."""
    R, E, m = _stability(code)
    facmin, facmax, facrej, facsafe = fac
    roundoff = 2.0 ** -52
    hmax = t1 - t0
    h = min(hstart, hmax)
    y, t = y0, t0
    nstp = nacc = nrej = 0
    rej_last = rej_more = False
    hexit = 0.0
    used_facrej = no_growth_binding = 0
    while t - t1 + roundoff <= 0.0:
        hexit = h
        h = min(h, t1 - t)
        if mistake == "hexit_truncated":
            hexit = h
        while True:
            z = lam * h
            ynew = R(z) * y
            yerr = E(z) * y
            sc = atol + rtol * max(abs(y), abs(ynew))
            err = np.sqrt((yerr / sc) ** 2 / NV)
            hnew = h * min(facmax, max(facmin, facsafe / err ** (1.0 / m["ELO"])))
            nstp += 1
            if err <= 1.0 or h <= hmin:
                nacc += 1
                y = ynew
                t += h
                hnew = max(hmin, min(hnew, hmax))
                if rej_last and mistake != "growth_after_reject":
                    if hnew > h:
                        no_growth_binding += 1
                    hnew = min(hnew, h)
                rej_last = rej_more = False
                h = hnew
                break
            if nacc >= 1:
                nrej += 1
            if rej_more or (mistake == "facrej_first" and rej_last is False):
                hnew = h * facrej
                used_facrej += 1
            rej_more = rej_last
            rej_last = True
            h = hnew
    return {"y": y, "T": t, "Hexit": hexit, "Nstp": nstp, "Nacc": nacc, "Nrej": nrej,
            "facrej": used_facrej, "no_growth": no_growth_binding}


def _batch(y0s, ts, te, rtol=1e-3, atol=1.0, method=4):
    """linear6 cells: X1 = y0s[c], X2..X6 = 0, P = 1."""
    n = len(y0s)
    b = workloads.linear6(ncells=n, method=method, rtol=rtol, atol=atol)
    b.C[:] = 0.0
    b.C[0] = y0s
    b.C[NV] = 1.0
    b.tstart[:] = ts
    b.tend[:] = te
    return b


def _rates(n, g=None, d=None):
    R = np.zeros((12, n))
    if g is not None:
        R[:6] = g
    if d is not None:
        R[6:] = d
    return R


def _x1_rates(lam):
    """rate constants making dX1/dt = lam*X1 (growth G1 or decay D1)"""
    R = _rates(1)
    R[0 if lam > 0 else 6, 0] = abs(lam)
    return R


# scenarios (method, lambda, Hstart): decays whose large first steps the error
# control rejects at least twice in a row, and a growth that is also
# rejected after its first acceptance (Nrej > 0)
SCEN = [(2, -2e-3, 3600.0), (4, -2e-3, 3600.0), (4, -5e-2, 3600.0), (1, -1e-3, 1800.0),
        (3, -2e-3, 3600.0), (5, -2e-3, 3600.0), (3, 5e-3, 3600.0)]


@pytest.mark.parametrize("code,lam,h0", SCEN)
def test_step_sequence_matches_closed_form_controller(code, lam, h0):
    y0, t0, t1 = 1.0e6, 0.0, 3600.0
    b = _batch(np.array([y0]), t0, t1, method=code)
    b.IPAR[3] = code
    out = ref.integrate(b, hstart=np.array([h0]), rconst=_x1_rates(lam))
    want = _model(code, lam, y0, t0, t1, h0, 1.0, 1e-3)
    # the scenario exercises the rules under test
    assert want["facrej"] >= 1 and want["no_growth"] >= 1, want
    if lam > 0:
        assert want["Nrej"] >= 1
    assert int(out["ierr"][0]) == 1
    st = out["istat"][:, 0]
    assert (st[2], st[3], st[4]) == (want["Nstp"], want["Nacc"], want["Nrej"])
    m = ref.method(code)
    newf = int(np.sum(m["NewF"][1:]))
    assert st[0] == want["Nacc"] * 2 + want["Nstp"] * newf  # Fcn0 + dF/dT per step
    assert st[1] == want["Nacc"] and st[5] == want["Nstp"] and st[6] == m["S"] * want["Nstp"]
    assert st[7] == 0
    assert out["rstat"][0, 0] == pytest.approx(want["T"], rel=1e-13)
    assert out["rstat"][1, 0] == pytest.approx(want["Hexit"], rel=1e-9)
    assert out["C"][0, 0] == pytest.approx(want["y"], rel=1e-10)
    assert np.all(out["C"][1:NV, 0] == 0.0)


@pytest.mark.parametrize("fac", [(0.5, 3.0, 0.3, 0.8), (0.1, 10.0, 0.05, 0.95)])
def test_user_step_factors(fac):
    code, d, h0 = 4, 2e-3, 3600.0
    b = _batch(np.array([1.0e6]), 0.0, 3600.0, method=code)
    b.RPAR[3:7] = fac
    out = ref.integrate(b, hstart=np.array([h0]), rconst=_x1_rates(-d))
    want = _model(code, -d, 1.0e6, 0.0, 3600.0, h0, 1.0, 1e-3, fac=fac)
    assert want["facrej"] >= 1
    st = out["istat"][:, 0]
    assert (st[2], st[3], st[4]) == (want["Nstp"], want["Nacc"], want["Nrej"])
    assert out["rstat"][1, 0] == pytest.approx(want["Hexit"], rel=1e-9)


@pytest.mark.parametrize("mistake", ["facrej_first", "growth_after_reject", "hexit_truncated"])
def test_planted_controller_mistakes_are_caught(mistake):
    """A plausible controller mistake planted in the model disagrees with the
    oracle on at least one of the scenarios (the comparison discriminates)."""
    caught = 0
    for code, lam, h0 in SCEN:
        b = _batch(np.array([1.0e6]), 0.0, 3600.0, method=code)
        out = ref.integrate(b, hstart=np.array([h0]), rconst=_x1_rates(lam))
        st, hexit = out["istat"][:, 0], float(out["rstat"][1, 0])
        w = _model(code, lam, 1.0e6, 0.0, 3600.0, h0, 1.0, 1e-3, mistake=mistake)
        if ((st[2], st[3], st[4]) != (w["Nstp"], w["Nacc"], w["Nrej"]) or
                abs(hexit - w["Hexit"]) > 1e-9 * abs(hexit)):
            caught += 1
    assert caught >= 1


def test_nan_cell_ends_with_minus9():
    b = _batch(np.array([1.0e6, 1.0e6, np.nan]), 0.0, 3600.0)
    R = _rates(3, d=np.full((6, 3), 1e-3))
    R[6, 1] = np.nan  # a NaN rate constant
    out = ref.integrate(b, rconst=R)
    assert list(out["ierr"]) == [1, -9, -9]
    # the NaN ends the cell at the first rejection, within a few attempts
    assert out["istat"][2, 1] <= 3 and out["istat"][2, 2] <= 3


def test_unit_step_factors_end_with_minus10():
    """FacMin = FacRej = 1 keep the step size after a rejection: without the
    bound the reject loop would never end."""
    b = _batch(np.array([1.0e6]), 0.0, 3600.0)
    b.RPAR[3] = 1.0
    b.RPAR[5] = 1.0
    out = ref.integrate(b, hstart=np.array([3600.0]), rconst=_x1_rates(-2e-3))
    assert int(out["ierr"][0]) == -10
    assert int(out["istat"][2, 0]) == 100001  # attempts: the first 100000 rejections + 1
    assert int(out["istat"][3, 0]) == 0


def _singular_case(nzero, method=4, h0=2.0 ** -8):
    """growth rates g_k = 2^(k-1) * fl(1/fl(h0*gamma)) for the first `nzero`
    species: the diagonal 1/(h*gamma) - g_k of I/(h gamma) - J is exactly
    zero at the k-th try (h = h0/2^(k-1)), so the first nzero tries are
    singular."""
    gam = ref.method(method)["Gamma"][0]
    r = 1.0 / (h0 * gam)
    g = np.zeros((6, 1))
    for k in range(nzero):
        g[k, 0] = 2.0 ** k * r
    for k in range(nzero, 6):
        g[k, 0] = 0.25 * r  # never exactly singular for h <= h0
    b = _batch(np.array([1.0]), 0.0, h0, method=method)
    b.C[:NV] = 1.0
    return b, _rates(1, g=g), h0


def test_singular_matrix_halves_step():
    b, R, h0 = _singular_case(5)
    out = ref.integrate(b, hstart=np.array([h0]), rconst=R)
    st = out["istat"][:, 0]
    assert int(out["ierr"][0]) == 1
    assert st[7] == 5                     # Nsng: five singular tries
    assert st[5] == st[2] + 5             # Ndec: one decomposition per attempt + the singular ones
    assert out["rstat"][0, 0] == pytest.approx(h0, rel=1e-14)
    # the first step is taken with h0/32: it needs at least 32 steps
    assert st[3] >= 32


def test_six_singular_matrices_end_with_minus8():
    b, R, h0 = _singular_case(6)
    out = ref.integrate(b, hstart=np.array([h0]), rconst=R)
    st = out["istat"][:, 0]
    assert int(out["ierr"][0]) == -8
    assert st[7] == 6 and st[5] == 6 and st[2] == 0 and st[3] == 0
    assert out["rstat"][0, 0] == 0.0      # Texit: the cell did not move
    assert np.all(out["C"][:NV, 0] == 1.0)
