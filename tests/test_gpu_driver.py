"""The C box-model driver (kpp_amd/drv/box_driver.c) on the GPU.

It calls the C ABI exactly as a host model would (kpp_ctx_create,
kpp_set_state, INTEGRATE_BATCH per output interval, kpp_get_state).  Bars:
 * plain context: its final state equals, bit for bit, the same sequence of
   calls made through ctypes on inputs regenerated here with the driver's own
   generator; every cell returns IERR = 1.  The ctypes path itself is checked
   against the oracle in test_gpu_parity.py::test_context_api_and_integrate_batch;
 * the tuned context (KPP_CTX_AUTOTUNE | KPP_CTX_LEARNED: autotuned persistent
   kernel, learned starting steps and launch order), and the persistent one
   with estimated starting steps (KPP_CTX_HSTART_AUTO), on saprc99: every cell
   IERR = 1, and against the oracle run through the same intervals from the
   starting steps the driver reports (then STEPMIN carried): the same step
   count for >= 99 % of cells and |rel| <= 1e-8 for species above ATOL
   (test_gpu_parity.py::test_saprc99_device_rates's bar).
"""
import ctypes
import math
import subprocess
from pathlib import Path

import numpy as np
import pytest

from kpp_amd import solver

pytestmark = pytest.mark.gpu

DRIVER = Path(__file__).resolve().parents[1] / "kpp_amd" / "lib" / "box_driver"


def _inputs(mech, ncells):
    L = solver.library()
    m = solver.mechanism(mech)
    init = np.zeros(m.nspec)
    assert L.kpp_mech_initial(mech.encode(), init.ctypes.data_as(ctypes.c_void_p)) == 0
    st = 12345
    mask = (1 << 64) - 1
    C = np.zeros((m.nspec, ncells))
    for c in range(ncells):
        for s in range(m.nspec):
            f = 1.0
            if s < m.nvar:
                st = (st + 0x9E3779B97F4A7C15) & mask
                z = st
                z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & mask
                z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & mask
                z = z ^ (z >> 31)
                u = (z >> 11) * (1.0 / 9007199254740992.0)
                f = math.pow(10.0, 2.0 * u - 1.0)
            C[s, c] = init[s] * f
    return m, C


def test_box_driver_matches_ctypes(gpu, tmp_path):
    if not DRIVER.exists():
        pytest.skip("box_driver not built")
    mech, n, method, t0, t1, dt = "small_strato", 300, solver.ROS3, 43200.0, 50400.0, 1200.0
    out = tmp_path / "out.bin"
    r = subprocess.run([str(DRIVER), mech, str(n), str(method), str(t0), str(t1), str(dt),
                        str(out)], capture_output=True, text=True, timeout=300)
    print(r.stdout, r.stderr)
    assert r.returncode == 0, r.stderr
    raw = out.read_bytes()
    m, C = _inputs(mech, n)
    got = np.frombuffer(raw[:8 * m.nspec * n], np.float64).reshape(m.nspec, n)
    ierr_d = np.frombuffer(raw[8 * m.nspec * n:8 * m.nspec * n + 4 * n], np.int32)
    assert np.all(ierr_d == 1)

    L = solver.library()
    ctx = ctypes.c_void_p()
    p = lambda a: a.ctypes.data_as(ctypes.c_void_p)
    assert L.kpp_ctx_create(ctypes.byref(ctx), mech.encode(), n, method, solver.SUN_TOD) == 0
    try:
        Cw = np.ascontiguousarray(C)
        temp = np.full(n, 270.0)
        cf = np.ones(n)
        assert L.kpp_set_state(ctx, p(Cw), p(temp), p(cf), None) == 0
        t = t0
        while t < t1:
            assert L.INTEGRATE_BATCH(ctx, t, t + dt) == 0
            t += dt
        want = np.zeros_like(Cw)
        ierr = np.zeros(n, np.int32)
        assert L.kpp_get_state(ctx, p(want), p(ierr), None, None) == 0
    finally:
        L.kpp_ctx_destroy(ctx)
    assert np.all(ierr == 1)
    assert np.array_equal(got, want)


def _oracle_intervals(mech, C, temp, cf, t0, t1, dt, h0, method, ipar4=0):
    """The oracle through INTEGRATE's sequence of calls (STEPMIN carried)."""
    from kpp_amd import workloads
    from oracle import ref
    n = C.shape[1]
    b = workloads.Batch(mech, method, C.copy(), np.full(n, temp), np.full(n, cf), np.zeros(n),
                        solver.SUN_TOD, np.zeros(n), np.zeros(n), np.array([1.0]),
                        np.array([1e-3]), solver.default_ipar(method), np.zeros(20),
                        np.arange(n))
    b.IPAR[4] = ipar4
    hs = h0.copy()
    t = t0
    while t < t1:
        b.tstart[:] = t
        b.tend[:] = t + dt
        o = ref.integrate(b, hstart=hs)
        b.C = o["C"]
        hs = o["rstat"][1]
        t += dt
    return o


@pytest.mark.parametrize("flags", ["tuned", "persistent,hstart_auto"])
def test_box_driver_tuned_matches_oracle(gpu, tmp_path, flags):
    if not DRIVER.exists():
        pytest.skip("box_driver not built")
    mech, n, method, t0, t1, dt = "saprc99", 2048, solver.RODAS3, 43200.0, 50400.0, 3600.0
    temp, cf = 298.0, 2.4476e13
    out = tmp_path / "out.bin"
    r = subprocess.run([str(DRIVER), mech, str(n), str(method), str(t0), str(t1), str(dt),
                        str(out), flags, str(temp), str(cf)], capture_output=True, text=True,
                       timeout=600)
    print(r.stdout, r.stderr)
    assert r.returncode == 0, r.stderr
    m, C = _inputs(mech, n)
    raw = out.read_bytes()
    k = 8 * m.nspec * n
    got = np.frombuffer(raw[:k], np.float64).reshape(m.nspec, n)
    ierr = np.frombuffer(raw[k:k + 4 * n], np.int32)
    h0 = np.frombuffer(raw[k + 4 * n:], np.float64).copy()
    assert np.all(ierr == 1)
    if flags == "tuned":
        assert np.all(h0 > 0)            # learned steps on the first interval
    else:
        assert np.all(h0 == 0)           # estimated inside the integrator (IPAR[4] = 1)
    o = _oracle_intervals(mech, C, temp, cf, t0, t1, dt, h0, method,
                          ipar4=1 if "hstart_auto" in flags else 0)
    assert np.all(o["ierr"] == 1)
    nv = m.nvar
    big = np.abs(o["C"][:nv]) > 1.0
    rel = np.abs(got[:nv] - o["C"][:nv])[big] / np.abs(o["C"][:nv])[big]
    print(flags, "max rel vs oracle", rel.max())
    assert rel.max() <= 1e-8
