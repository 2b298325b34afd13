"""GPU tests of graph capture, the autotuner and the launch-shape independence of results.

Bars:
 * a recorded tuned step (scheduler features, network, bucket sort, integrator)
   replays bit for bit like the eager calls, on three replays separated by the
   program's own eager work and synchronisations;
 * the autotuner tries its candidates, leaves C equal to C0, and its choice
   integrates every cell bit for bit like the default launch.
"""
import os

import numpy as np
import pytest

from kpp_amd import learned, solver, workloads

pytestmark = pytest.mark.gpu

WEIGHTS = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "kpp_amd",
                       "weights", "saprc99_step_mlp.npz")


def _batch(gpu, n, seed):
    b = workloads.saprc99_uneven(ncells=n, seed=seed)
    return b, b.to_device(gpu)


@pytest.mark.parametrize("lds", [0, 20])
def test_graph_replays_match_eager(gpu, lds):
    """Replays of the recorded tuned step (scheduler + persistent kernel, with
    and without the LDS mirror of the LU) match the eager step bit for bit."""
    import torch
    b, d = _batch(gpu, 8192, 51)
    model = learned.StepModel.load(WEIGHTS, gpu)
    n = b.ncells
    C = d["C"]
    C0 = C.clone()
    ierr = torch.empty(n, dtype=torch.int32, device=gpu)
    istat = torch.empty((8, n), dtype=torch.int32, device=gpu)
    rstat = torch.empty((2, n), dtype=torch.float64, device=gpu)
    ws = solver.Workspace(gpu)
    span = float(np.max(b.tend - b.tstart))
    X = model.features(C, d["temp"], d["sun"])
    sched = model.schedule(X, 1e-6, span)

    def step():
        model.features(C, d["temp"], d["sun"], out=X)
        model.schedule(X, 1e-6, span, out=sched)
        code, _, _, _ = solver.rosenbrock(
            b.mech, C, d["temp"], d["cfactor"], tstart=d["tstart"], tend=d["tend"], atol=b.atol,
            rtol=b.rtol, IPAR=b.IPAR, RPAR=b.RPAR, sun=d["sun"], sun_mode=b.sun_mode,
            hstart=sched[0], order=sched[2], ierr=ierr, istat=istat, rstat=rstat,
            variant=solver.PERSISTENT, lds_entries=lds, workspace=ws)
        assert code == 0

    C.copy_(C0)
    step()
    torch.cuda.synchronize()
    ref = [t.cpu().clone() for t in (C, istat, ierr, rstat, sched[0], sched[2])]
    assert bool((ref[2] == 1).all())

    s = torch.cuda.Stream()
    g = solver.capture(step, s)
    try:
        for k in range(3):
            # the program's own eager work between replays
            C.copy_(C0)
            istat.zero_()
            ierr.fill_(-99)
            rstat.zero_()
            sched[0].zero_()
            torch.cuda.synchronize()
            g.launch(s)
            s.synchronize()
            got = [t.cpu() for t in (C, istat, ierr, rstat, sched[0], sched[2])]
            for a, r, name in zip(got, ref, ("C", "istat", "ierr", "rstat", "hstart", "order")):
                assert torch.equal(a, r), (k, name)
    finally:
        g.close()


def test_capture_rejects_vector_tolerances(gpu):
    import torch
    b, d = _batch(gpu, 256, 52)
    nvar = solver.mechanism(b.mech).nvar
    IPAR = np.array(b.IPAR)
    IPAR[1] = 0
    ws = solver.Workspace(gpu)
    ierr = torch.empty(b.ncells, dtype=torch.int32, device=gpu)
    s = torch.cuda.Stream()
    codes = []

    def step():
        ierr.zero_()  # one recorded node, so the graph is not empty
        try:
            solver.rosenbrock(b.mech, d["C"], d["temp"], d["cfactor"], tstart=d["tstart"],
                              tend=d["tend"], atol=np.ones(nvar), rtol=np.full(nvar, 1e-3),
                              IPAR=IPAR, RPAR=b.RPAR, sun=d["sun"], sun_mode=b.sun_mode,
                              ierr=ierr, workspace=ws)
        except solver.KppError as e:
            codes.append(str(e))

    # size the workspace first (nothing may allocate while recording)
    solver.rosenbrock(b.mech, d["C"].clone(), d["temp"], d["cfactor"], tstart=d["tstart"],
                      tend=d["tend"], atol=b.atol, rtol=b.rtol, IPAR=b.IPAR, RPAR=b.RPAR,
                      sun=d["sun"], sun_mode=b.sun_mode, ierr=ierr, workspace=ws)
    torch.cuda.synchronize()
    g = solver.capture(step, s)
    g.close()
    assert codes and "-102" in codes[0]


def test_autotune_choice_integrates_identically(gpu):
    import torch
    b, d = _batch(gpu, 65536, 53)
    C = d["C"]
    C0 = C.clone()
    best, ms, tried = solver.autotune(b.mech, C, C0, d["temp"], d["cfactor"], tstart=d["tstart"],
                                      tend=d["tend"], atol=b.atol, rtol=b.rtol, IPAR=b.IPAR,
                                      RPAR=b.RPAR, sun=d["sun"], sun_mode=b.sun_mode)
    torch.cuda.synchronize()
    print("autotune:", best, ms, tried)
    # eight launch shapes of the plain tile layout, three with an LDS mirror
    assert tried == 11 and ms > 0
    assert best[4] in (0, 20, 24)
    assert best[0] == solver.PERSISTENT and 1 <= best[3] <= 64 and best[1] >= 1
    assert torch.equal(C, C0)
    outs = []
    for launch in ({"variant": solver.PERSISTENT},
                   {"variant": best[0], "nblocks": best[1], "refill_min": best[2],
                    "lanes": best[3], "lds_entries": best[4]}):
        Cw = C0.clone()
        code, ierr, istat, rstat = solver.rosenbrock(
            b.mech, Cw, d["temp"], d["cfactor"], tstart=d["tstart"], tend=d["tend"], atol=b.atol,
            rtol=b.rtol, IPAR=b.IPAR, RPAR=b.RPAR, sun=d["sun"], sun_mode=b.sun_mode, **launch)
        torch.cuda.synchronize()
        assert code == 0 and bool((ierr == 1).all())
        outs.append((Cw.cpu(), istat.cpu(), rstat.cpu()))
    for a, r in zip(outs[0], outs[1]):
        assert torch.equal(a, r)
