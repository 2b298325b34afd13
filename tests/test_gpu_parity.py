"""GPU parity: the HIP path against the CPU oracle through the C ABI.

Bars (stated per test):
 * model functions (Fun, Jac_SP, KppDecomp, KppSolve): bit-exact;
 * rate constants: |rel| <= 1e-13 (device exp/pow vs glibc; parity of the
   libm boundary itself is not pinned by the reference);
 * whole integrations with the oracle's own rate constants supplied: the
   same step counts for every cell and |rel| <= 1e-10 for species above
   ATOL (the step controller's pow(Err,1/ELO) and the norm's square root
   come from the device math library), Texit to 1e-12 and the last step
   size Hexit to 1e-7;
 * whole integrations with device-computed rates: same IERR everywhere,
   same step count for >= 99 % of cells, |rel| <= 1e-5 (small_strato, whose
   SUN comes from the device cos) or 1e-8 (saprc99, saprcnov) for species
   above ATOL.
"""
import ctypes

import numpy as np
import pytest

from kpp_amd import solver, workloads
from oracle import ref

pytestmark = pytest.mark.gpu

MECHS = ["small_strato", "saprc99", "saprcnov"]


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def _t(a, dev):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).to(dev)


def _oracle_rates(b):
    L = ref.lib(b.mech)
    nr = ref.sizes(b.mech)["NREACT"]
    R = np.zeros((nr, b.ncells))
    tmp = np.zeros(nr)
    for c in range(b.ncells):
        L.ros_ref_rconst(b.temp[c], b.sun[c], b.cfactor[c], b.tstart[c], _p(tmp))
        R[:, c] = tmp
    return R


def _gpu_run(b, dev, variant=solver.PLAIN, rconst=None, IPAR=None, RPAR=None, hstart=None,
             atol=None, rtol=None, **kw):
    import torch
    d = b.to_device(dev)
    code, ierr, istat, rstat = solver.rosenbrock(
        b.mech, d["C"], d["temp"], d["cfactor"], tstart=d["tstart"], tend=d["tend"],
        atol=b.atol if atol is None else atol, rtol=b.rtol if rtol is None else rtol,
        IPAR=b.IPAR if IPAR is None else IPAR, RPAR=b.RPAR if RPAR is None else RPAR,
        sun=d["sun"], sun_mode=b.sun_mode, variant=variant,
        rconst=None if rconst is None else _t(rconst, dev),
        hstart=None if hstart is None else _t(hstart, dev), **kw)
    torch.cuda.synchronize()
    return {"code": code, "C": d["C"].cpu().numpy(), "ierr": ierr.cpu().numpy(),
            "istat": istat.cpu().numpy(), "rstat": rstat.cpu().numpy()}


def _rel(g, o, nvar, atol=1.0):
    a, r = g["C"][:nvar], o["C"][:nvar]
    big = np.abs(r) > atol
    rel = np.abs(a - r) / np.maximum(np.abs(r), 1e-300)
    return rel[big]


# ---------------------------------------------------------------- components
@pytest.mark.parametrize("mech", MECHS)
def test_model_functions_bit_exact(gpu, mech):
    import torch
    s = ref.sizes(mech)
    L = ref.lib(mech)
    nv, nf, nr, nnz = s["NVAR"], s["NFIX"], s["NREACT"], s["LU_NONZERO"]
    n = 512
    rng = np.random.default_rng(3)
    V = rng.uniform(0.1, 10.0, (nv, n))
    F = rng.uniform(0.1, 10.0, (nf, n))
    R = rng.uniform(0.1, 10.0, (nr, n))
    dV, dF, dR = _t(V, gpu), _t(F, gpu), _t(R, gpu)
    out = torch.empty((nv, n), dtype=torch.float64, device=gpu)
    solver.fun(mech, dV, dF, dR, out)
    jv = torch.empty((nnz, n), dtype=torch.float64, device=gpu)
    solver.jac_sp(mech, dV, dF, dR, jv)
    torch.cuda.synchronize()
    g_vd, g_jv = out.cpu().numpy(), jv.cpu().numpy()
    # diagonally shifted matrix so the decomposition is well defined
    icol, crow, diag = solver.lu_pattern(mech)
    G = -jv.cpu().numpy()
    G[diag[:nv]] += 50.0
    gG = _t(G, gpu)
    ising = torch.empty(n, dtype=torch.int32, device=gpu)
    solver.decomp(mech, gG, ising)
    X = rng.uniform(-1.0, 1.0, (nv, n))
    gX = _t(X, gpu)
    solver.solve(mech, gG, gX)
    torch.cuda.synchronize()
    for c in range(0, n, 37):
        v, f, r = V[:, c].copy(), F[:, c].copy(), R[:, c].copy()
        vd = np.zeros(nv)
        L.ros_ref_fun(_p(v), _p(f), _p(r), _p(vd))
        assert (vd == g_vd[:, c]).all()
        j = np.zeros(nnz)
        L.ros_ref_jac_sp(_p(v), _p(f), _p(r), _p(j))
        assert (j == g_jv[:, c]).all()
        g = G[:, c].copy()
        assert L.ros_ref_decomp(_p(g)) == 0 == int(ising[c])
        assert (g == gG[:, c].cpu().numpy()).all()
        x = X[:, c].copy()
        L.ros_ref_solve(_p(g), _p(x))
        assert (x == gX[:, c].cpu().numpy()).all()


def test_decomp_reports_first_zero_diagonal(gpu):
    import torch
    mech = "saprc99"
    s = ref.sizes(mech)
    icol, crow, diag = solver.lu_pattern(mech)
    n = 4
    G = np.random.default_rng(1).uniform(1.0, 2.0, (s["LU_NONZERO"], n))
    G[diag[5], 1] = 0.0
    G[diag[40], 2] = 0.0
    G[diag[7], 3] = 0.0
    G[diag[60], 3] = 0.0
    gG = _t(G, gpu)
    ising = torch.empty(n, dtype=torch.int32, device=gpu)
    solver.decomp(mech, gG, ising)
    torch.cuda.synchronize()
    L = ref.lib(mech)
    want = [L.ros_ref_decomp(_p(G[:, c].copy())) for c in range(n)]
    assert ising.cpu().tolist() == want == [0, 6, 41, 8]


@pytest.mark.parametrize("mech", MECHS)
def test_rate_constants(gpu, mech):
    import torch
    s = ref.sizes(mech)
    n = 256
    rng = np.random.default_rng(7)
    T = rng.uniform(200.0, 320.0, n)
    cf = rng.uniform(1e12, 3e13, n)
    sun = rng.uniform(0.0, 1.0, n)
    tm = rng.uniform(0.0, 86400.0, n)
    out = torch.empty((s["NREACT"], n), dtype=torch.float64, device=gpu)
    solver.rates(mech, _t(T, gpu), _t(cf, gpu), _t(sun, gpu), _t(tm, gpu), out)
    torch.cuda.synchronize()
    got = out.cpu().numpy()
    L = ref.lib(mech)
    st = solver.mechanism(mech).stats
    number = np.array(st["rate_types"]) == 1
    tmp = np.zeros(s["NREACT"])
    for c in range(n):
        L.ros_ref_rconst(T[c], sun[c], cf[c], tm[c], _p(tmp))
        want = np.where(number, 0.0, tmp)
        rel = np.abs(got[:, c] - want) / np.maximum(np.abs(want), 1e-300)
        assert rel.max() <= 1e-13


# ------------------------------------------------------------ integrations
@pytest.mark.parametrize("wl,variant", [("saprc99_uneven", solver.PLAIN),
                                        ("saprc99_uneven", solver.PERSISTENT),
                                        ("saprcnov_uneven", solver.PERSISTENT)])
def test_saprc99_rates_given_matches_oracle(gpu, wl, variant):
    b = workloads.WORKLOADS[wl](ncells=1024)
    R = _oracle_rates(b)
    o = ref.integrate(b, rconst=R)
    g = _gpu_run(b, gpu, variant=variant, rconst=R)
    assert g["code"] == 0
    assert (g["ierr"] == 1).all() and (o["ierr"] == 1).all()
    assert (g["istat"] == o["istat"]).all()          # all 8 counters, every cell
    assert _rel(g, o, o["nvar"]).max() <= 1e-10
    # Texit: the same sum of the same steps; Hexit = H*FacSafe/Err^(1/ELO) of the
    # last step, where Err is the norm of the embedded error estimate, a difference
    # of stage vectors: their 1e-10 agreement can lose digits to cancellation there
    assert np.allclose(g["rstat"][0], o["rstat"][0], rtol=1e-12, atol=0)
    hrel = np.abs(g["rstat"][1] - o["rstat"][1]) / np.abs(o["rstat"][1])
    print(f"{wl}: max relative difference of Hexit {hrel.max():.3e}")
    assert hrel.max() <= 1e-7


def test_small_strato_time_of_day(gpu):
    b = workloads.small_strato_tod(ncells=4096)
    o = ref.integrate(b)
    for v in (solver.PLAIN, solver.PERSISTENT):
        g = _gpu_run(b, gpu, variant=v)
        assert (g["ierr"] == 1).all() and (o["ierr"] == 1).all()
        assert (g["istat"][2] == o["istat"][2]).mean() >= 0.99
        rel = _rel(g, o, o["nvar"])
        assert rel.max() <= 1e-5 and np.median(rel) <= 1e-9


@pytest.mark.parametrize("wl", ["saprc99_day", "saprc99_uneven", "saprcnov_uneven"])
def test_saprc99_device_rates(gpu, wl):
    b = workloads.WORKLOADS[wl](ncells=1024)
    o = ref.integrate(b)
    g = _gpu_run(b, gpu)
    assert (g["ierr"] == 1).all() and (o["ierr"] == 1).all()
    assert (g["istat"][2] == o["istat"][2]).mean() >= 0.99
    assert _rel(g, o, o["nvar"]).max() <= 1e-8


def test_plain_and_persistent_identical(gpu):
    b = workloads.saprc99_uneven(ncells=3000)  # not a multiple of 64
    g0 = _gpu_run(b, gpu, variant=solver.PLAIN)
    for refill in (1, 16, 64):
        g1 = _gpu_run(b, gpu, variant=solver.PERSISTENT, nblocks=7, refill_min=refill)
        assert (g0["C"] == g1["C"]).all()
        assert (g0["istat"] == g1["istat"]).all() and (g0["ierr"] == g1["ierr"]).all()


def test_processing_order_does_not_change_results(gpu):
    import torch
    b = workloads.saprc99_uneven(ncells=640)
    g0 = _gpu_run(b, gpu, variant=solver.PERSISTENT)
    perm = np.random.default_rng(2).permutation(b.ncells).astype(np.int32)
    g1 = _gpu_run(b, gpu, variant=solver.PERSISTENT, order=_t(perm, gpu))
    assert (g0["C"] == g1["C"]).all() and (g0["istat"] == g1["istat"]).all()


@pytest.mark.parametrize("method", [solver.ROS2, solver.ROS3, solver.ROS4, solver.RODAS3,
                                    solver.RODAS4])
@pytest.mark.parametrize("scalar_tol,autonomous", [(True, False), (False, True)])
def test_methods_and_options(gpu, method, scalar_tol, autonomous):
    b = workloads.small_strato_tod(ncells=512, method=method)
    IPAR = solver.default_ipar(method, scalar_tol=scalar_tol, autonomous=autonomous)
    nv = ref.sizes(b.mech)["NVAR"]
    atol = b.atol if scalar_tol else np.linspace(0.5, 2.0, nv)
    rtol = b.rtol if scalar_tol else np.linspace(1e-3, 2e-3, nv)
    bb = workloads.Batch(**{**b.__dict__, "IPAR": IPAR, "atol": atol, "rtol": rtol})
    R = _oracle_rates(bb)
    o = ref.integrate(bb, rconst=R)
    g = _gpu_run(bb, gpu, rconst=R)
    assert (g["ierr"] == o["ierr"]).all()
    assert (g["istat"] == o["istat"]).all()
    assert _rel(g, o, nv).max() <= 1e-10


def test_parameter_errors_match_rosenbrock(gpu):
    b = workloads.small_strato_tod(ncells=64)
    base = solver.default_ipar(solver.ROS3)
    cases = []
    ip = base.copy(); ip[2] = -1; cases.append((ip, np.zeros(20), b.atol, b.rtol, -1))
    ip = base.copy(); ip[3] = 0; cases.append((ip, np.zeros(20), b.atol, b.rtol, -2))
    ip = base.copy(); ip[3] = 7; cases.append((ip, np.zeros(20), b.atol, b.rtol, -2))
    for k in (0, 1, 2):
        rp = np.zeros(20); rp[k] = -1.0; cases.append((base, rp, b.atol, b.rtol, -3))
    for k in (3, 4, 5, 6):
        rp = np.zeros(20); rp[k] = -1.0; cases.append((base, rp, b.atol, b.rtol, -4))
    cases.append((base, np.zeros(20), np.array([0.0]), b.rtol, -5))
    cases.append((base, np.zeros(20), b.atol, np.array([1.0]), -5))
    cases.append((base, np.zeros(20), b.atol, np.array([1e-16]), -5))
    for ip, rp, at, rt, want in cases:
        bb = workloads.Batch(**{**b.__dict__, "IPAR": ip, "RPAR": rp, "atol": at, "rtol": rt})
        o = ref.integrate(bb)
        g = _gpu_run(bb, gpu)
        assert g["code"] == want
        assert (g["ierr"] == want).all() and (o["ierr"] == want).all()
        assert (g["C"] == b.C).all()  # nothing integrated


def test_too_many_steps_and_step_too_small(gpu):
    b = workloads.saprc99_uneven(ncells=256)
    R = _oracle_rates(b)
    o0 = ref.integrate(b, rconst=R)
    ip = b.IPAR.copy()
    ip[2] = int(np.median(o0["istat"][2]))  # -6 for cells needing more steps
    bb = workloads.Batch(**{**b.__dict__, "IPAR": ip})
    o = ref.integrate(bb, rconst=R)
    g = _gpu_run(bb, gpu, rconst=R)
    assert set(np.unique(o["ierr"])) == {1, -6}
    assert (g["ierr"] == o["ierr"]).all() and (g["istat"] == o["istat"]).all()
    assert np.array_equal(g["rstat"], o["rstat"]) or np.allclose(g["rstat"], o["rstat"], rtol=1e-12)
    # T + 0.1*H == T at the first step: -7 (int/rosenbrock.c:517)
    ts = np.full(b.ncells, 1e12)  # ulp(1e12) > 0.1*DeltaMin
    bb = workloads.Batch(**{**b.__dict__, "tstart": ts, "tend": ts + 3600.0})
    o = ref.integrate(bb, rconst=R)
    g = _gpu_run(bb, gpu, rconst=R)
    assert (o["ierr"] == -7).all() and (g["ierr"] == -7).all()
    assert (g["istat"] == o["istat"]).all()


def test_hstart_per_cell_and_restart(gpu):
    # Hexit of one call as Hstart of the next, as INTEGRATE does
    # (int/rosenbrock.c:131)
    b = workloads.saprc99_day(ncells=256)
    R = _oracle_rates(b)
    o1 = ref.integrate(b, rconst=R)
    g1 = _gpu_run(b, gpu, rconst=R)
    b2 = workloads.Batch(**{**b.__dict__, "C": o1["C"], "tstart": b.tend, "tend": b.tend + 3600})
    o2 = ref.integrate(b2, rconst=R, hstart=o1["rstat"][1])
    g2 = _gpu_run(b2, gpu, rconst=R, hstart=g1["rstat"][1])
    assert (g2["istat"] == o2["istat"]).all()
    assert _rel(g2, o2, o2["nvar"]).max() <= 1e-10


def test_backward_integration_quirk(gpu):
    """Tend < Tstart: KPP's ABS macro (int/rosenbrock.c:3) expands
    ABS(Tend-Tstart) to -Tend-Tstart, Hmax becomes negative and Rosenbrock
    stops with -7 before the first step.  The device reproduces it."""
    b = workloads.small_strato_tod(ncells=128)
    bb = workloads.Batch(**{**b.__dict__, "tstart": b.tend, "tend": b.tstart})
    R = _oracle_rates(bb)
    o = ref.integrate(bb, rconst=R)
    g = _gpu_run(bb, gpu, rconst=R)
    assert (o["ierr"] == -7).all() and (g["ierr"] == -7).all()
    assert (g["istat"] == 0).all() and (o["istat"] == 0).all()
    assert (g["rstat"][0] == bb.tstart).all() and (g["C"] == bb.C).all()


def test_context_api_and_integrate_batch(gpu):
    """Host-buffer API (kpp_ctx_*) and INTEGRATE_BATCH over consecutive
    intervals, with STEPMIN carried as INTEGRATE does."""
    L = solver.library()
    b = workloads.saprc99_day(ncells=200)
    ctx = ctypes.c_void_p()
    assert L.kpp_ctx_create(ctypes.byref(ctx), b"saprc99", b.ncells, solver.RODAS3, 0) == 0
    try:
        C = np.ascontiguousarray(b.C)
        assert L.kpp_set_state(ctx, _p(C), _p(b.temp), _p(b.cfactor), _p(b.sun)) == 0
        Cref = C.copy()
        hs = np.zeros(b.ncells)
        for k in range(2):
            t0, t1 = 43200.0 + 1800 * k, 43200.0 + 1800 * (k + 1)
            assert L.INTEGRATE_BATCH(ctx, t0, t1) == 0
            bb = workloads.Batch(**{**b.__dict__, "C": Cref, "tstart": np.full(b.ncells, t0),
                                    "tend": np.full(b.ncells, t1)})
            o = ref.integrate(bb, hstart=hs)
            Cref, hs = o["C"], o["rstat"][1]
        got = np.zeros_like(C)
        ierr = np.zeros(b.ncells, np.int32)
        assert L.kpp_get_state(ctx, _p(got), _p(ierr), None, None) == 0
        assert (ierr == 1).all()
        assert _rel({"C": got}, {"C": Cref}, o["nvar"]).max() <= 1e-8
    finally:
        L.kpp_ctx_destroy(ctx)


# ------------------------------------------------------------ register kernel
@pytest.mark.parametrize("method", [solver.ROS2, solver.ROS3, solver.ROS4, solver.RODAS3,
                                    solver.RODAS4])
def test_register_kernel_bit_identical_to_plain(gpu, method):
    """small_strato with the working set in VGPRs: the same operations in the
    same order as the tile kernels, so every cell's state, counters, Texit and
    Hexit are bit for bit the plain kernel's (time-of-day SUN)."""
    b = workloads.small_strato_tod(ncells=2000, method=method)  # not a multiple of 64
    g0 = _gpu_run(b, gpu, variant=solver.PLAIN)
    g1 = _gpu_run(b, gpu, variant=solver.REGISTER)
    assert g1["code"] == 0 and (g1["ierr"] == g0["ierr"]).all()
    assert np.array_equal(g1["C"], g0["C"])
    assert (g1["istat"] == g0["istat"]).all()
    assert np.array_equal(g1["rstat"], g0["rstat"])


def test_register_kernel_matches_oracle(gpu):
    b = workloads.small_strato_tod(ncells=4096)
    o = ref.integrate(b)
    g = _gpu_run(b, gpu, variant=solver.REGISTER, lanes=16)
    assert (g["ierr"] == 1).all() and (o["ierr"] == 1).all()
    assert (g["istat"][2] == o["istat"][2]).mean() >= 0.99
    rel = _rel(g, o, o["nvar"])
    assert rel.max() <= 1e-5 and np.median(rel) <= 1e-9


def test_register_kernel_refused_for_large_mechanisms(gpu):
    b = workloads.saprc99_uneven(ncells=64)
    with pytest.raises(solver.KppError):
        _gpu_run(b, gpu, variant=solver.REGISTER)


# ------------------------------------------------------------ LDS mirror of LU entries
@pytest.mark.parametrize("lds", [20, 24])
@pytest.mark.parametrize("wl,method", [("saprc99", solver.RODAS3), ("saprc99", solver.ROS2),
                                       ("saprc99", solver.RODAS4), ("saprcnov", solver.RODAS3)])
def test_lds_mirror_bit_identical_to_plain(gpu, lds, wl, method):
    """Persistent kernel with the factored values of the most re-read LU
    entries kept in LDS (KppDecompL/KppSolveL): the same operations in the
    same order, so every cell's state, counters, Texit and Hexit are bit for
    bit the plain kernel's, for 2-, 4- and 6-stage methods."""
    if wl == "saprc99":
        b = workloads.saprc99_uneven(ncells=3000, method=method)  # not a multiple of 64
    else:
        b = workloads.saprcnov_uneven(ncells=1500, method=method)
    g0 = _gpu_run(b, gpu, variant=solver.PLAIN)
    for refill in (1, 16):
        g1 = _gpu_run(b, gpu, variant=solver.PERSISTENT, nblocks=11, refill_min=refill,
                      lds_entries=lds)
        assert g1["code"] == 0 and (g1["ierr"] == g0["ierr"]).all()
        assert np.array_equal(g1["C"], g0["C"])
        assert (g1["istat"] == g0["istat"]).all()
        assert np.array_equal(g1["rstat"], g0["rstat"])


def test_lds_mirror_refused(gpu):
    """A mirror the mechanism is not built with, a size other than 0/20/24,
    or a mirror outside the FP64 persistent kernel is an argument error."""
    big = workloads.saprc99_uneven(ncells=64)
    small = workloads.small_strato_tod(ncells=64)
    for b, kw in ((small, {"variant": solver.PERSISTENT, "lds_entries": 24}),
                  (big, {"variant": solver.PERSISTENT, "lds_entries": 32}),
                  (big, {"variant": solver.PLAIN, "lds_entries": 24}),
                  (big, {"variant": solver.PERSISTENT, "lds_entries": 24, "lu_fp32": 1})):
        with pytest.raises(solver.KppError):
            _gpu_run(b, gpu, **kw)

