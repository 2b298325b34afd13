"""The cooperative integrator (KPP_VARIANT_COOP: one wavefront per cell, the
LU in LDS, the model routines as the mechanism compiler's list-scheduled
programs) against the one-lane-per-cell kernels and the oracle.

Bars:
 * every cell's state, IERR, counters, Texit and Hexit bit for bit those of
   the plain kernel (same operations in the same order), for MCM, saprc99,
   saprcnov and small_strato, 2-, 3-, 4- and 6-stage methods, fixed and
   time-of-day SUN, and the estimated starting step (IPAR[4] = 1);
 * with the oracle's rate constants supplied, the oracle's counters in every
   cell and |rel| <= 1e-10 for species above ATOL/RTOL (MCM);
 * the singular-matrix path and the reject-loop bounds of linear6 as the
   oracle has them;
 * mixed precision and LDS mirrors are argument errors on this variant.
"""
import numpy as np
import pytest

from kpp_amd import solver, workloads
from oracle import ref

pytestmark = pytest.mark.gpu


def _t(a, dev):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).to(dev)


def _run(b, dev, variant, rconst=None, hstart=None, IPAR=None, **kw):
    import torch
    d = b.to_device(dev)
    code, ierr, istat, rstat = solver.rosenbrock(
        b.mech, d["C"], d["temp"], d["cfactor"], tstart=d["tstart"], tend=d["tend"],
        atol=b.atol, rtol=b.rtol, IPAR=b.IPAR if IPAR is None else IPAR, RPAR=b.RPAR,
        sun=d["sun"], sun_mode=b.sun_mode, variant=variant,
        rconst=None if rconst is None else _t(rconst, dev),
        hstart=None if hstart is None else _t(hstart, dev), **kw)
    torch.cuda.synchronize()
    assert code == 0
    return {"C": d["C"].cpu().numpy(), "ierr": ierr.cpu().numpy(), "istat": istat.cpu().numpy(),
            "rstat": rstat.cpu().numpy()}


def _same(a, b):
    for k in ("C", "ierr", "istat", "rstat"):
        assert np.array_equal(a[k], b[k], equal_nan=True), k


@pytest.mark.parametrize("wl,n,method", [
    ("mcm_mixed", 300, solver.RODAS3),
    ("saprc99_uneven", 700, solver.RODAS3),
    ("saprc99_uneven", 300, solver.ROS2),
    ("saprc99_uneven", 300, solver.RODAS4),
    ("saprcnov_uneven", 500, solver.RODAS3),
    ("small_strato_tod", 400, solver.ROS3),
])
def test_coop_bit_identical_to_plain(gpu, wl, n, method):
    b = workloads.WORKLOADS[wl](ncells=n, method=method)
    g0 = _run(b, gpu, solver.PLAIN)
    g1 = _run(b, gpu, solver.COOP)
    assert (g1["ierr"] == 1).all()
    _same(g1, g0)


def test_coop_small_grid_and_estimated_start(gpu):
    """A grid of 7 wavefronts over 500 cells (each wavefront takes many
    cells from the queue), with IPAR[4] = 1."""
    b = workloads.saprc99_uneven(ncells=500)
    ip = b.IPAR.copy()
    ip[4] = 1
    g0 = _run(b, gpu, solver.PLAIN, IPAR=ip)
    g1 = _run(b, gpu, solver.COOP, IPAR=ip, nblocks=7)
    _same(g1, g0)


def test_coop_mcm_rates_given_matches_oracle(gpu):
    b = workloads.mcm_mixed(ncells=96, seed=62)
    nr = ref.sizes("mcm")["NREACT"]
    L = ref.lib("mcm")
    R = np.zeros((nr, b.ncells))
    tmp = np.zeros(nr)
    for c in range(b.ncells):
        Cc = np.ascontiguousarray(b.C[:, c])
        L.ros_ref_rconst_c(b.temp[c], b.sun[c], b.cfactor[c], 0.0, Cc.ctypes.data, tmp.ctypes.data)
        R[:, c] = tmp
    o = ref.integrate(b, rconst=R)
    g = _run(b, gpu, solver.COOP, rconst=R)
    assert np.array_equal(g["ierr"], o["ierr"]) and (o["ierr"] == 1).all()
    assert np.array_equal(g["istat"], o["istat"])
    nv = ref.sizes("mcm")["NVAR"]
    thr = float(b.atol[0] / b.rtol[0])
    m = np.abs(o["C"][:nv]) > thr
    rel = np.abs(g["C"][:nv] - o["C"][:nv])[m] / np.abs(o["C"][:nv])[m]
    assert rel.max() <= 1e-10


def test_coop_linear6_singular_and_bounds(gpu):
    # the singular cases of test_gpu_linear6 (5 and 6 singular tries) and a NaN cell
    method = solver.RODAS3
    h0 = 2.0 ** -8
    gam = ref.method(method)["Gamma"][0]
    r = 1.0 / (h0 * gam)
    R = np.zeros((12, 3))
    for k in range(6):
        R[k, 0] = 2.0 ** k * r if k < 5 else 0.25 * r
        R[k, 1] = 2.0 ** k * r
    R[6, 2] = np.nan
    b = workloads.linear6(ncells=3, method=method)
    b.C[:6] = 1.0
    b.tend[:] = h0
    hs = np.full(3, h0)
    o = ref.integrate(b, rconst=R, hstart=hs)
    assert list(o["ierr"]) == [1, -8, -9]
    g = _run(b, gpu, solver.COOP, rconst=R, hstart=hs)
    assert np.array_equal(g["ierr"], o["ierr"]) and np.array_equal(g["istat"], o["istat"])
    assert np.array_equal(g["rstat"][0], o["rstat"][0])
    ok = np.isfinite(o["C"][:6])
    assert np.allclose(g["C"][:6][ok], o["C"][:6][ok], rtol=1e-12, atol=0)


def test_coop_refused_combinations(gpu):
    b = workloads.mcm_mixed(ncells=64)
    for kw in ({"lu_fp32": 1}, {"lds_entries": 20}, {"lanes": 32}):
        with pytest.raises(solver.KppError):
            _run(b, gpu, solver.COOP, **kw)
