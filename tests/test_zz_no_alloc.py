"""The workspace-free entry point kpp_rosenbrock_dev_static (no counterpart in
KPP; for callers that cannot allocate device memory) against
kpp_rosenbrock_dev, through the C ABI.

Bar: bit for bit the same state, IERR, counters, Texit and Hexit in every
cell (the same kernels; the persistent kernel computes each cell's rate
constants at its start instead of in a pre-pass, with the same formulae),
on two calls in a row, the second with a smaller batch, for the persistent
kernel (its grid cut to what the library's array holds), the plain kernel
and the register kernel; a batch too large for the array is refused.
"""
import ctypes

import numpy as np
import pytest

from kpp_amd import solver, workloads

pytestmark = pytest.mark.gpu

KPP_E_WORKSPACE = -104


def _bind():
    L = solver.library()
    vp = ctypes.c_void_p
    L.kpp_rosenbrock_dev_static.argtypes = [ctypes.c_char_p, ctypes.POINTER(solver._Batch), vp, vp,
                                            vp, vp, ctypes.POINTER(solver._Launch), vp]
    L.kpp_rosenbrock_dev_static.restype = ctypes.c_int
    return L


def _static_run(b, dev, variant, lanes=0):
    """kpp_rosenbrock_dev_static on a workloads batch; returns host results."""
    import torch
    L = _bind()
    d = b.to_device(dev)
    n = b.ncells
    ierr = torch.empty(n, dtype=torch.int32, device=dev)
    istat = torch.empty((8, n), dtype=torch.int32, device=dev)
    rstat = torch.empty((2, n), dtype=torch.float64, device=dev)
    bb = solver._Batch()
    bb.ncells, bb.ld = n, n
    bb.C, bb.temp, bb.cfactor = d["C"].data_ptr(), d["temp"].data_ptr(), d["cfactor"].data_ptr()
    bb.sun = d["sun"].data_ptr()
    bb.sun_mode = b.sun_mode
    bb.tstart, bb.tend = d["tstart"].data_ptr(), d["tend"].data_ptr()
    bb.ierr, bb.istat, bb.rstat = ierr.data_ptr(), istat.data_ptr(), rstat.data_ptr()
    IPAR = np.ascontiguousarray(b.IPAR, np.int32)
    RPAR = np.ascontiguousarray(b.RPAR, np.float64)
    atol = np.ascontiguousarray(b.atol, np.float64)
    rtol = np.ascontiguousarray(b.rtol, np.float64)
    l = solver._Launch(variant, 0, 0, lanes, 0, 0)
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    code = L.kpp_rosenbrock_dev_static(b.mech.encode(), ctypes.byref(bb), IPAR.ctypes.data,
                                       RPAR.ctypes.data, atol.ctypes.data, rtol.ctypes.data,
                                       ctypes.byref(l), stream)
    torch.cuda.synchronize()
    return code, {"C": d["C"].cpu().numpy(), "ierr": ierr.cpu().numpy(),
                  "istat": istat.cpu().numpy(), "rstat": rstat.cpu().numpy()}


def _dev_run(b, dev, variant, lanes=0):
    import torch
    d = b.to_device(dev)
    code, ierr, istat, rstat = solver.rosenbrock(
        b.mech, d["C"], d["temp"], d["cfactor"], tstart=d["tstart"], tend=d["tend"],
        atol=b.atol, rtol=b.rtol, IPAR=b.IPAR, RPAR=b.RPAR, sun=d["sun"], sun_mode=b.sun_mode,
        variant=variant, lanes=lanes)
    torch.cuda.synchronize()
    assert code == 0
    return {"C": d["C"].cpu().numpy(), "ierr": ierr.cpu().numpy(), "istat": istat.cpu().numpy(),
            "rstat": rstat.cpu().numpy()}


def _same(a, b):
    for k in ("C", "ierr", "istat", "rstat"):
        assert np.array_equal(a[k], b[k]), k


@pytest.mark.parametrize("wl,variant,sizes", [
    ("saprc99_uneven", solver.PERSISTENT, (20000, 3000)),
    ("saprc99_uneven", solver.PLAIN, (4000, 1000)),
    ("small_strato_tod", solver.REGISTER, (5000, 700)),
    ("mcm_mixed", solver.PERSISTENT, (512, 200)),
])
def test_static_entry_matches_workspace_entry(gpu, wl, variant, sizes):
    for n in sizes:  # two calls in a row, the second smaller
        b = workloads.WORKLOADS[wl](ncells=n)
        code, got = _static_run(b, gpu, variant)
        assert code == 0
        want = _dev_run(b, gpu, variant)
        _same(got, want)
        assert (got["ierr"] == 1).all()


def test_static_entry_refuses_a_batch_beyond_its_array(gpu):
    # plain kernel: one 64-cell tile of 1802 doubles per lane (saprc99,
    # Rodas3) per 64 cells: 1,048,576 cells need 15 GB of tiles
    b = workloads.saprc99_uneven(ncells=1048576)
    code, _ = _static_run(b, gpu, solver.PLAIN, lanes=64)
    assert code == KPP_E_WORKSPACE
