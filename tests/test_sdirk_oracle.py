"""The oracle's restatement of KPP's C SDIRK integrator (int/sdirk.c) pinned
on the CPU.

* The five methods' Butcher tableaux, as int/sdirk.c:988-1306 prints them
  (rkA, rkB, rkBhat; written out below as the method definitions), satisfy
  the Runge-Kutta order conditions of their order (Sdirk2a/2b: 2; Sdirk3a:
  2 with 3 stages, "adjoint-invariant", int/sdirk.f90:3-5; Sdirk4a/4b: 4)
  and their embedded formulas one order less; c_i are the row sums of A.
* The coefficients the integrator actually uses are derived from the
  tableau, and the oracle's values must equal them: Z = h A k gives
  D = b A^-1 (the new solution from the stage increments Z), E = (b - bhat)
  A^-1 (the error estimate), Theta[m, i] = sum_{j<i} a_ij (A^-1)_jm (the
  explicit part of stage i), gamma = a_ii and C = c; the starting values
  Z_i^0 = sum_{j<i} Alpha[j, i] Z_j extrapolate exactly when Z is a
  polynomial in c of degree <= i without constant term.
* Whole integrations: the linear decay of linear6 against exp(lam t), and
  small_strato (time-of-day SUN) against scipy's Radau at rtol 1e-12; the
  counters obey the integrator's structure (each Newton iteration one
  function evaluation and one solve, each completed step one more solve).
* The reference's own behaviour kept in the restatement: SDIRK_PrepareMatrix
  halving only its local step on a singular matrix (a singular first matrix
  is counted, the integration goes on), an unknown ICNTRL[2] selecting
  Sdirk2a, and the parameter checks' codes (returned, a documented deviation:
  the reference prints them and goes on).
"""
import numpy as np
import pytest

from kpp_amd import workloads
from oracle import ref
from tests.test_step_control import _batch, _x1_rates

S2 = 0.7071067811865475244008443621048490
G2A = 0.2928932188134524755991556378951510
G2B = 1.707106781186547524400844362104849
G3A = 0.2113248654051871177454256097490213
G4A = 0.2666666666666666666666666666666667

# (A lower triangular with the diagonal, b, bhat, order)
TABLEAUX = {
    1: ([[G2A], [S2, G2A]], [S2, G2A],
        [0.6666666666666666666666666666666667, 0.3333333333333333333333333333333333], 2),
    2: ([[G2B], [-0.707106781186547524400844362104849, G2B]],
        [-0.707106781186547524400844362104849, G2B],
        [0.6666666666666666666666666666666667, 0.3333333333333333333333333333333333], 2),
    3: ([[G3A], [G3A, G3A], [G3A, 0.5773502691896257645091487805019573, G3A]],
        [G3A, 0.5773502691896257645091487805019573, G3A],
        [G3A, 0.6477918909913548037576239837516312, 0.1408832436034580784969504064993475], 2),
    4: ([[G4A], [0.5, G4A],
         [0.3541539528432732316227461858529820, -0.5415395284327323162274618585298197e-01, G4A],
         [0.8515494131138652076337791881433756e-01, -0.6484332287891555171683963466229754e-01,
          0.7915325296404206392428857585141242e-01, G4A],
         [2.100115700566932777970612055999074, -0.7677800284445976813343102185062276,
          2.399816361080026398094746205273880, -2.998818699869028161397714709433394, G4A]],
        [2.100115700566932777970612055999074, -0.7677800284445976813343102185062276,
         2.399816361080026398094746205273880, -2.998818699869028161397714709433394, G4A],
        [2.885264204387193942183851612883390, -0.1458793482962771337341223443218041,
         2.390008682465139866479830743628554, -4.129393538556056674929560012190140, 0.0], 4),
    5: ([[0.25], [0.5, 0.25], [0.34, -0.40e-01, 0.25],
         [0.2727941176470588235294117647058824, -0.5036764705882352941176470588235294e-01,
          0.2757352941176470588235294117647059e-01, 0.25],
         [1.041666666666666666666666666666667, -1.020833333333333333333333333333333,
          7.812500000000000000000000000000000, -7.083333333333333333333333333333333, 0.25]],
        [1.041666666666666666666666666666667, -1.020833333333333333333333333333333,
         7.812500000000000000000000000000000, -7.083333333333333333333333333333333, 0.25],
        [1.069791666666666666666666666666667, -0.894270833333333333333333333333333,
         7.695312500000000000000000000000000, -7.083333333333333333333333333333333,
         0.212500000000000000000000000000000], 4),
}


def _tab(code):
    rows, b, bh, p = TABLEAUX[code]
    s = len(b)
    A = np.zeros((s, s))
    for i, r in enumerate(rows):
        A[i, :len(r)] = r
    return A, np.array(b), np.array(bh), p


def _rk_order(A, b):
    c = A.sum(1)
    return {1: [b.sum() - 1], 2: [b @ c - 0.5],
            3: [b @ c**2 - 1 / 3, b @ (A @ c) - 1 / 6],
            4: [b @ c**3 - 0.25, b @ (c * (A @ c)) - 1 / 8, b @ (A @ c**2) - 1 / 12,
                b @ (A @ (A @ c)) - 1 / 24]}


@pytest.fixture(scope="module", autouse=True)
def _have_linear6():
    try:
        ref.sizes("linear6")
    except (FileNotFoundError, OSError):
        pytest.skip("oracle for linear6 not built")


@pytest.mark.parametrize("code", [1, 2, 3, 4, 5])
def test_tableau_order_conditions(code):
    A, b, bh, p = _tab(code)
    for q, r in _rk_order(A, b).items():
        if q <= p:
            assert np.abs(r).max() <= 1e-14, (q, r)
    for q, r in _rk_order(A, bh).items():
        if q <= p - 1:
            assert np.abs(r).max() <= 1e-14, ("embedded", q, r)
    assert np.abs(_rk_order(A, bh)[p]).max() > 1e-6
    assert np.allclose(np.diag(A), A[0, 0])  # singly diagonally implicit


@pytest.mark.parametrize("code", [1, 2, 3, 4, 5])
def test_integrator_coefficients_from_tableau(code):
    A, b, bh, p = _tab(code)
    m = ref.sdirk_method(code)
    s = len(b)
    assert m["known"] and m["S"] == s and m["ELO"] == (2 if p <= 3 else 4)
    Ai = np.linalg.inv(A)
    assert m["Gamma"] == A[0, 0]
    assert np.allclose(m["C"], A.sum(1), rtol=1e-14, atol=1e-15)
    assert np.allclose(m["D"], b @ Ai, rtol=1e-13, atol=1e-13)
    assert np.allclose(m["E"], (b - bh) @ Ai, rtol=1e-12, atol=1e-12)
    Th = np.zeros((s, s))
    for i in range(s):
        for mm in range(i):
            Th[mm, i] = sum(A[i, j] * Ai[j, mm] for j in range(i))
    assert np.allclose(m["Theta"], Th, rtol=1e-12, atol=1e-12)
    c = A.sum(1)
    for i in range(1, s):
        for k in range(1, i + 1):
            got = sum(m["Alpha"][j, i] * c[j] ** k for j in range(i))
            assert got == pytest.approx(c[i] ** k, rel=1e-12, abs=1e-13), (i, k)


def test_unknown_method_is_sdirk2a():
    m9, m1 = ref.sdirk_method(9), ref.sdirk_method(1)
    assert not m9["known"] and m1["known"]
    for k in ("C", "D", "E", "Theta", "Alpha"):
        assert np.array_equal(m9[k], m1[k])
    b = _batch(np.array([1.0e6]), 0.0, 3600.0)
    ic = np.zeros(20, np.int32)
    ic[1] = 1
    outs = []
    for code in (9, 1):
        ic[2] = code
        outs.append(ref.integrate_sdirk(b, ic, rconst=_x1_rates(-2e-3)))
    assert np.array_equal(outs[0]["C"], outs[1]["C"])
    assert np.array_equal(outs[0]["istat"], outs[1]["istat"])


@pytest.mark.parametrize("code", [1, 2, 3, 4, 5])
def test_linear_decay_and_counters(code):
    lam, y0 = -2e-3, 1.0e6
    b = _batch(np.array([y0]), 0.0, 3600.0, rtol=1e-6, atol=1e-3)
    ic = np.zeros(20, np.int32)
    ic[1], ic[2] = 1, code
    out = ref.integrate_sdirk(b, ic, rconst=_x1_rates(lam))
    assert int(out["ierr"][0]) == 1
    exact = y0 * np.exp(lam * 3600.0)
    assert out["C"][0, 0] == pytest.approx(exact, rel=2e-4)
    st = out["istat"][:, 0]
    nfun, njac, nstp, nacc, nrej, ndec, nsol, nsng = st
    assert nsol == nfun + nstp and nacc <= nstp and nsng == 0 and ndec >= 1
    assert out["rstat"][0, 0] == pytest.approx(3600.0, rel=1e-14)
    assert 0 < out["rstat"][1, 0] <= 3600.0


def test_small_strato_against_scipy_radau():
    """the order-4 methods against Radau at rtol 1e-12 (the order-2 methods
    need very small steps near sunrise at such tolerances and can end with
    -7, as the algorithm does)"""
    from scipy.integrate import solve_ivp
    from tests.test_oracle import _rhs
    b = workloads.small_strato_tod(ncells=8)
    b.atol = np.array([1e-3])
    for code, rt, bar in ((5, 1e-7, 1e-4), (4, 1e-5, 1e-3)):
        b.rtol = np.array([rt])
        ic = np.zeros(20, np.int32)
        ic[1], ic[2] = 1, code
        o = ref.integrate_sdirk(b, ic)
        assert (o["ierr"] == 1).all()
        for c in range(b.ncells):
            f = _rhs("small_strato", b, c)
            sol = solve_ivp(f, (b.tstart[c], b.tend[c]), b.C[:5, c], method="Radau",
                            rtol=1e-12, atol=1e-6)
            y = sol.y[:, -1]
            big = np.abs(y) > 1.0
            assert np.allclose(o["C"][:5, c][big], y[big], rtol=bar), (code, c)


def test_singular_first_matrix_is_counted_and_passed():
    """linear6 with growth rate 1/(h0 gamma) on X1: the first iteration matrix
    has an exactly zero diagonal.  SDIRK_PrepareMatrix halves only its local
    step (the matrix of h0/2 is regular), counts one singular matrix, and the
    integration goes on with the caller's step (no -8)."""
    m = ref.sdirk_method(4)
    h0 = 2.0 ** -8
    g = 1.0 / (h0 * m["Gamma"])
    b = _batch(np.array([1.0]), 0.0, 2.0 ** -6)
    R = np.zeros((12, 1))
    R[0, 0] = g
    ic = np.zeros(20, np.int32)
    ic[1], ic[2] = 1, 4
    rc = np.zeros(20)
    rc[2] = h0
    out = ref.integrate_sdirk(b, ic, rc, rconst=R)
    st = out["istat"][:, 0]
    assert int(out["ierr"][0]) == 1 and st[7] >= 1
    assert np.isfinite(out["C"][0, 0])


def test_parameter_errors():
    b = _batch(np.array([1.0e6]), 0.0, 3600.0)
    R = _x1_rates(-2e-3)
    cases = [((3, -1), None, -1), ((4, -2), None, -2), (None, (0, -1.0), -3),
             (None, (1, -1.0), -3), (None, (2, -1.0), -3), (None, (3, -0.5), -4),
             (None, (6, -0.5), -4)]
    for icx, rcx, want in cases:
        ic = np.zeros(20, np.int32)
        ic[1] = 1
        rc = np.zeros(20)
        if icx:
            ic[icx[0]] = icx[1]
        if rcx:
            rc[rcx[0]] = rcx[1]
        out = ref.integrate_sdirk(b, ic, rc, rconst=R)
        assert int(out["ierr"][0]) == want, (icx, rcx)
        assert not out["istat"].any()
    # the scalar tolerance test as written: AbsTol > 0 and RelTol == 0 is -5
    b.rtol = np.array([0.0])
    ic = np.zeros(20, np.int32)
    ic[1] = 1
    assert int(ref.integrate_sdirk(b, ic, rconst=R)["ierr"][0]) == -5
