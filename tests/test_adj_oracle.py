"""The oracle's discrete adjoint Rosenbrock (oracle/rosadj_ref.h, after
int/rosenbrock_adj.c) and the adjoint's model functions the mechanism
compiler emits (Hessian, HessTR_Vec, Jac_SP_Vec, JacTR_SP_Vec, KppSolveTR;
src/gen.c:1327-1600, 1829-1870, 2035-2106), pinned by what the mathematics
fixes rather than by a second copy of the formulas:

 * the Hessian against central differences of Jac_SP (the rate laws are
   polynomial of degree three at most, so central differences with a
   large step are exact up to rounding), HessTR_Vec
   against the dense contraction of that Hessian, Jac_SP_Vec/JacTR_SP_Vec
   against the dense Jacobian, KppSolveTR against A^T x = b;
 * the adjoint of a fixed-step integration (Hmin = Hmax = Hstart, so every
   step is accepted with the same H) against central differences of the
   forward map, every Jacobian entry at once (NADJ = NVAR), every method:
   autonomous (ICNTRL[0] = 1, the map is smooth) to 1e-7 of the normalised
   sensitivity, non-autonomous to 1e-4 (the forward's finite-difference
   dF/dT makes the map rough at about 1e-6, which bounds the differences);
   the planted mistakes (no Hessian term, no dJ/dT term, the solve not
   transposed) each fail that bar;
 * the reference's backward loop, which leaves out the first accepted step,
   equal to the complete adjoint of the integration restarted after that
   step;
 * a linear decay with adaptive steps (linear6): lambda(t0) = Y(tend)/Y(t0)
   lambda(tend), the product of the stability functions of the steps taken;
 * the parameter checks and the checkpoint limit.
"""
import ctypes

import numpy as np
import pytest

from kpp_amd import workloads
from oracle import ref

MECHS = ["small_strato", "saprc99"]


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def _pattern(mech):
    sz = ref.sizes(mech)
    icol = np.zeros(sz["LU_NONZERO"], np.int32)
    crow = np.zeros(sz["NVAR"] + 1, np.int32)
    diag = np.zeros(sz["NVAR"] + 1, np.int32)
    ref.lib(mech).ros_ref_lu_pattern(_p(icol), _p(crow), _p(diag))
    rows = np.zeros(sz["LU_NONZERO"], np.int32)
    for i in range(sz["NVAR"]):
        rows[crow[i]:crow[i + 1]] = i
    return rows, icol


def _state(mech, seed=0):
    sz = ref.sizes(mech)
    rng = np.random.default_rng(seed)
    # magnitudes near one: the differences below then lose nothing to
    # cancellation, and every term still enters with a distinct weight
    V = rng.uniform(0.5, 2.0, sz["NVAR"])
    F = rng.uniform(0.5, 2.0, max(sz["NFIX"], 1))
    R = rng.uniform(0.5, 2.0, sz["NREACT"])
    return sz, V, F, R


def _dense_jac(mech, V, F, R):
    sz = ref.sizes(mech)
    rows, cols = _pattern(mech)
    jvs = np.zeros(sz["LU_NONZERO"])
    ref.lib(mech).ros_ref_jac_sp(_p(V), _p(F), _p(R), _p(jvs))
    J = np.zeros((sz["NVAR"], sz["NVAR"]))
    J[rows, cols] = jvs
    return J, jvs


def _hessian(mech, V, F, R):
    L = ref.lib(mech)
    nh = L.adj_ref_nhess()
    hs = np.zeros(max(nh, 1))
    ii, jj, kk = (np.zeros(max(nh, 1), np.int32) for _ in range(3))
    L.adj_ref_hessian(_p(V), _p(F), _p(R), _p(hs), _p(ii), _p(jj), _p(kk))
    n = ref.sizes(mech)["NVAR"]
    H = np.zeros((n, n, n))
    for m in range(nh):  # 0-based in the C output, j <= k stored once
        i, j, k = ii[m], jj[m], kk[m]
        H[i, j, k] = H[i, k, j] = hs[m]
    return hs, H


@pytest.mark.parametrize("mech", MECHS)
def test_hessian_against_jacobian_differences(mech):
    sz, V, F, R = _state(mech)
    n = sz["NVAR"]
    _, H = _hessian(mech, V, F, R)
    Hfd = np.zeros((n, n, n))
    for k in range(n):
        e = 0.5 * V[k]  # central differences are exact for quadratic Jacobians
        Vp, Vm = V.copy(), V.copy()
        Vp[k] += e
        Vm[k] -= e
        Hfd[:, :, k] = (_dense_jac(mech, Vp, F, R)[0] - _dense_jac(mech, Vm, F, R)[0]) / (2 * e)
    scale = np.abs(Hfd).max(axis=(1, 2), keepdims=True) + 1e-300
    assert np.max(np.abs(H - Hfd) / scale) <= 1e-9
    assert np.count_nonzero(H) > 0


@pytest.mark.parametrize("mech", MECHS)
def test_vector_products_and_transposed_solve(mech):
    sz, V, F, R = _state(mech, 1)
    n = sz["NVAR"]
    L = ref.lib(mech)
    rng = np.random.default_rng(5)
    U1, U2 = rng.standard_normal(n), rng.standard_normal(n)
    hs, H = _hessian(mech, V, F, R)
    htu = np.zeros(n)
    L.adj_ref_hesstr_vec(_p(hs), _p(U1), _p(U2), _p(htu))
    want = np.einsum("ijk,i,j->k", H, U1, U2)
    assert np.allclose(htu, want, rtol=1e-12, atol=1e-12 * np.abs(want).max())
    J, jvs = _dense_jac(mech, V, F, R)
    out = np.zeros(n)
    L.adj_ref_jac_vec(_p(jvs), _p(U1), _p(out))
    assert np.allclose(out, J @ U1, rtol=1e-12, atol=1e-12 * np.abs(J @ U1).max())
    L.adj_ref_jactr_vec(_p(jvs), _p(U1), _p(out))
    assert np.allclose(out, J.T @ U1, rtol=1e-12, atol=1e-12 * np.abs(J.T @ U1).max())
    # (1/(h gamma) - J)^T x = b through the LU factors
    A = -jvs.copy()
    rows, cols = _pattern(mech)
    A[rows == cols] += 1.0 / (60.0 * 0.43)
    Ad = np.zeros((n, n))
    Ad[rows, cols] = A
    assert L.ros_ref_decomp(_p(A)) == 0
    x = np.zeros(n)
    L.adj_ref_solve_tr(_p(A), _p(U2), _p(x))
    assert np.allclose(Ad.T @ x, U2, rtol=1e-9, atol=1e-9 * np.abs(U2).max())


def _fixed_step_batch(ncells=3):
    b = workloads.small_strato_tod(ncells=ncells)
    b.tend = b.tstart + 1800.0
    rc = np.zeros(20)
    rc[0] = rc[1] = rc[2] = 60.0
    return b, rc


def _fd_error(b, rc, ic, eps, first_step=True):
    nv = ref.sizes(b.mech)["NVAR"]
    lam = np.zeros((nv * nv, b.ncells))
    for a in range(nv):
        lam[a * nv + a] = 1.0
    o = ref.integrate_adj(b, lam, ic, rc, first_step=first_step)
    assert (o["ierr"] == 1).all(), o["ierr"]
    worst = 0.0
    for j in range(nv):
        e = eps * np.abs(b.C[j])
        Cp, Cm = b.C.copy(), b.C.copy()
        Cp[j] += e
        Cm[j] -= e
        yp = ref.integrate_adj(b, lam, ic, rc, C=Cp)["C"][:nv]
        ym = ref.integrate_adj(b, lam, ic, rc, C=Cm)["C"][:nv]
        fd = (yp - ym) / (2 * e)
        adj = np.array([o["Lambda"][k * nv + j] for k in range(nv)])
        sc = np.abs(b.C[j]) / np.maximum(np.abs(o["C"][:nv]), 1e-30)
        err = np.abs(adj - fd) * sc
        worst = max(worst, float(np.max(np.where(np.isfinite(err), err, np.inf))))
    return worst, o


@pytest.mark.parametrize("method", [1, 2, 3, 4, 5, 7])
def test_fixed_step_adjoint_matches_differences(method):
    b, rc = _fixed_step_batch()
    ic = np.zeros(20, np.int32)
    ic[0], ic[1], ic[2] = 1, 1, method
    worst, o = _fd_error(b, rc, ic, 1e-6)
    assert worst <= 1e-7, worst
    nv = ref.sizes(b.mech)["NVAR"]
    S = {1: 2, 2: 3, 3: 4, 4: 4, 5: 6, 7: 4}[method]
    # every step accepted with the same H; one Jac per step and per stage back
    assert o["istat"][2, 0] == o["istat"][3, 0] == 30 and o["istat"][4, 0] == 0
    assert o["istat"][1, 0] == 30 + 30 * (S + 1)
    assert o["istat"][6, 0] == 30 * S * (1 + nv)


def test_non_autonomous_and_planted_mistakes():
    b, rc = _fixed_step_batch()
    ic = np.zeros(20, np.int32)
    ic[1], ic[2] = 1, 2
    worst, _ = _fd_error(b, rc, ic, 1e-5)
    assert worst <= 1e-4, worst
    L = ref.lib(b.mech)
    try:
        for plant in (1, 2, 3):
            L.adj_ref_plant(plant)
            bad, _ = _fd_error(b, rc, ic, 1e-5)
            assert bad > 1e-3, (plant, bad)  # ten times the bar at least
    finally:
        L.adj_ref_plant(0)


def test_reference_backward_loop_skips_first_step():
    b, rc = _fixed_step_batch()
    nv = ref.sizes(b.mech)["NVAR"]
    ic = np.zeros(20, np.int32)
    ic[1], ic[2] = 1, 4
    lam = np.ones((nv, b.ncells))
    o_ref = ref.integrate_adj(b, lam, ic, rc, first_step=False)
    # restart after the first step: Y1 at t0 + h, then the complete adjoint
    b1 = workloads.small_strato_tod(ncells=b.ncells)
    b1.tend = b.tstart + 60.0
    y1 = ref.integrate_adj(b1, lam, ic, rc)["C"]
    b2 = workloads.small_strato_tod(ncells=b.ncells)
    b2.tstart = b.tstart + 60.0
    b2.tend = b.tend
    o_rest = ref.integrate_adj(b2, lam, ic, rc, first_step=True, C=y1)
    assert np.allclose(o_ref["C"], o_rest["C"], rtol=1e-13, atol=0.0)
    assert np.allclose(o_ref["Lambda"], o_rest["Lambda"], rtol=1e-11, atol=0.0)
    o_all = ref.integrate_adj(b, lam, ic, rc, first_step=True)
    assert np.max(np.abs(o_all["Lambda"] - o_ref["Lambda"]) / np.abs(o_ref["Lambda"])) > 1e-3


@pytest.mark.parametrize("method", [1, 2, 3, 4, 5, 7])
@pytest.mark.parametrize("lam", [-2e-3, -5e-2, 5e-3])
def test_linear_decay_closed_form(method, lam):
    from tests.test_step_control import _batch, _x1_rates
    b = _batch(np.array([1.0e6]), 0.0, 3600.0, rtol=1e-4, atol=1.0)
    ic = np.zeros(20, np.int32)
    ic[1], ic[2] = 1, method
    nv = ref.sizes("linear6")["NVAR"]
    L0 = np.zeros((nv, 1))
    L0[0] = 2.5
    o = ref.integrate_adj(b, L0, ic, rconst=_x1_rates(lam), first_step=True)
    assert int(o["ierr"][0]) == 1 and o["istat"][3, 0] > 3
    assert o["Lambda"][0, 0] == pytest.approx(2.5 * o["C"][0, 0] / 1.0e6, rel=1e-12)
    assert np.all(o["Lambda"][1:, 0] == 0.0)


def test_parameter_checks_and_checkpoint_limit():
    b, rc = _fixed_step_batch(2)
    nv = ref.sizes(b.mech)["NVAR"]
    lam = np.ones((nv, 2))

    def run(ic_set=(), rc_set=(), **kw):
        ic = np.zeros(20, np.int32)
        ic[1] = 1
        r = np.zeros(20)
        for k, v in ic_set:
            ic[k] = v
        for k, v in rc_set:
            r[k] = v
        return ref.integrate_adj(b, lam, ic, r, **kw)["ierr"]

    assert (run([(3, 100)]) == -1).all()          # any Max_no_steps
    assert (run([(2, 6)]) == -2).all()            # no such method
    assert (run([(5, 7)]) == -2).all()            # CadjMethod
    assert (run([(6, 5)]) == -9).all()            # AdjointType
    assert (run([(6, 3)]) == -13).all()           # continuous: not restated
    assert (run(rc_set=[(0, -1.0)]) == -3).all()
    assert (run(rc_set=[(3, -0.5)]) == -4).all()
    b.rtol = np.array([1.0])
    assert (run() == -5).all()
    b.rtol = np.array([1e-3])
    assert (run([(6, 1)]) == 1).all()             # no adjoint: forward only
    assert (run(cap=10) == -12).all()
