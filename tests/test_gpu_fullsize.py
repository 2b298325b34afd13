"""GPU tests at the benchmark's full size: 1,048,576 uneven saprc99 cells (Rodas3,
RTOL 1e-3, ATOL 1, 1 h; the bench batch, seed 3), through properties that do not
need the oracle to finish at this size.

Bars:
 * every cell ends with IERR = 1 at Tend (RPAR[10] per cell, to rounding), on the plain kernel,
   the persistent kernel and the tuned path (learned schedule + persistent kernel);
 * the plain and the persistent kernel, started the same way, give every cell the
   same final state and the same eight counters bit for bit (one cell's arithmetic
   does not depend on the lane, wavefront or order it runs in);
 * sulfur is conserved: saprc99's only sulfur species are SO2 and H2SO4
   (models/saprc99.spc: SO2 = S + 2O, H2SO4 = 2H + S + 4O) and its only reaction
   that touches them is <44> OH + SO2 = HO2 + H2SO4 (models/saprc99.eqn:46), so
   SO2 + H2SO4 is a linear invariant of the ODE, which Rosenbrock methods keep to
   rounding error: relative drift <= 1e-12 per cell on every path;
 * the tuned path's starting steps and launch order change each cell's step
   sequence, not its accuracy class: for species above ATOL/RTOL the tuned and
   plain final states agree to 1e-2 relative (RTOL 1e-3 solutions of the same ODE).
"""
import os

import numpy as np
import pytest

from kpp_amd import learned, solver, workloads

pytestmark = pytest.mark.gpu

WEIGHTS = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "kpp_amd",
                       "weights", "saprc99_step_mlp.npz")
NCELLS = 1048576


def _run(b, d, C0, variant, hstart=None, order=None):
    import torch
    C = C0.clone()
    code, ierr, istat, rstat = solver.rosenbrock(
        b.mech, C, d["temp"], d["cfactor"], tstart=d["tstart"], tend=d["tend"], atol=b.atol,
        rtol=b.rtol, IPAR=b.IPAR, RPAR=b.RPAR, sun=d["sun"], sun_mode=b.sun_mode,
        hstart=hstart, order=order, variant=variant)
    assert code == 0
    torch.cuda.synchronize()
    return C, ierr, istat, rstat


def test_full_batch_properties(gpu):
    import torch
    b = workloads.saprc99_uneven(ncells=NCELLS, seed=3)
    d = b.to_device(gpu)
    C0 = d["C"].clone()
    mech = solver.mechanism(b.mech)
    print(f"full-size batch: {b.ncells} saprc99 cells", flush=True)

    plain = _run(b, d, C0, solver.PLAIN)
    pers = _run(b, d, C0, solver.PERSISTENT)
    assert torch.equal(plain[0], pers[0])
    assert torch.equal(plain[2], pers[2])
    del pers

    model = learned.StepModel.load(WEIGHTS, gpu)
    X = model.features(C0, d["temp"], d["sun"])
    hs, _, order = model.schedule(X, 1e-6, float(np.max(b.tend - b.tstart)))
    tuned = _run(b, d, C0, solver.PERSISTENT, hstart=hs, order=order)

    tend = torch.as_tensor(b.tend, device=gpu)
    iso2, ih2so4 = mech.species.index("SO2"), mech.species.index("H2SO4")
    s0 = (C0[iso2] + C0[ih2so4]).double()
    for name, (C, ierr, istat, rstat) in (("plain", plain), ("tuned", tuned)):
        assert bool((ierr == 1).all()), name
        assert bool(torch.allclose(rstat[0], tend, rtol=1e-12, atol=0.0)), name
        assert bool(torch.isfinite(C).all()), name
        drift = ((C[iso2] + C[ih2so4]) - s0).abs() / s0
        assert float(drift.max()) <= 1e-12, (name, float(drift.max()))

    nv = mech.nvar
    ref = plain[0][:nv]
    big = ref.abs() > float(b.atol[0]) / float(b.rtol[0])
    rel = ((tuned[0][:nv] - ref).abs() / ref.abs())[big]
    assert float(rel.max()) <= 1e-2, float(rel.max())
    # the learned schedule needs fewer attempts than the default start
    assert float(tuned[2][2].double().mean()) < 0.8 * float(plain[2][2].double().mean())
