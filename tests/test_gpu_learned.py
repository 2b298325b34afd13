"""GPU tests of the learned step scheduler (BF16 MFMA network, step logs, schedule).

Bars:
 * forward pass: |out - oracle| <= 1e-2 * max|oracle| (oracle: float64 over
   the same BF16-rounded weights, inputs and activations; the device sums in
   FP32, so BF16 roundings of hidden activations can differ by one unit);
 * gradients: relative Frobenius error <= 2e-2 per parameter block and
   cosine >= 0.999 against the oracle's gradient (same rounding model);
 * Adam: the updated FP32 weights equal the closed-form update of the device
   gradient to 1e-6 relative;
 * step log: with the oracle's rate constants supplied (identical step
   counts), the device's first accepted step equals the oracle's to 1e-12
   relative (step sizes after a rejection pass through the device pow);
 * gain: a network trained on logged steps cuts the mean step attempts of
   held-out cells by >= 10 % against Rosenbrock's default start, with every
   cell still succeeding.
"""
import ctypes

import numpy as np
import pytest

from kpp_amd import learned, solver, workloads
from oracle import mlp_ref as R
from oracle import ref

pytestmark = pytest.mark.gpu


def _model(gpu, seed=0):
    m = learned.StepModel(mech="saprc99", device=gpu, seed=seed)
    import torch
    # non-zero biases so that a dropped bias term shows
    rng = np.random.default_rng(seed + 7)
    w = m.w.cpu().numpy()
    _, b_off = m.offsets()
    for l in range(4):
        nb = 128 if l < 3 else 16
        w[b_off[l]:b_off[l] + nb] = rng.uniform(-0.2, 0.2, nb)
    m.w = torch.tensor(w, device=gpu)
    m.wb = torch.empty(m.w.numel(), dtype=torch.bfloat16, device=gpu)
    m.sync()
    return m


def _inputs(m, n, seed):
    import torch
    rng = np.random.default_rng(seed)
    X = np.zeros((n, m.in_pad))
    X[:, :m.nin] = R.bf16(rng.standard_normal((n, m.nin)))
    return X, torch.tensor(X, dtype=torch.bfloat16, device=m.device)


def test_forward_matches_oracle(gpu):
    import torch
    m = _model(gpu)
    n = 1000  # not a multiple of 16: exercises the row guards
    X, Xd = _inputs(m, n, 1)
    out = m.forward(Xd)
    torch.cuda.synchronize()
    O, _, _ = R.forward(m.w.cpu().double().numpy(), X, m.in_pad, emulate=True)
    got = out.cpu().double().numpy()
    assert np.max(np.abs(got - O)) <= 1e-2 * np.max(np.abs(O))
    assert np.all(got[:, 9:] == O[:, 9:])  # padded outputs: bias only


def test_gradients_match_oracle(gpu):
    import torch
    m = _model(gpu, seed=2)
    n = 2048
    X, Xd = _inputs(m, n, 3)
    rng = np.random.default_rng(4)
    y = rng.standard_normal(n).astype(np.float32)
    bk = rng.integers(0, 8, n).astype(np.int32)
    wr = (rng.uniform(size=n) > 0.1).astype(np.float32)
    loss, g = m.grad(Xd, torch.tensor(y, device=gpu), torch.tensor(bk, device=gpu),
                     torch.tensor(wr, device=gpu))
    torch.cuda.synchronize()
    L, gref = R.grads(m.w.cpu().double().numpy(), X, y.astype(np.float64), bk,
                      wr.astype(np.float64), m.in_pad, emulate=True)
    assert abs(float(loss.item()) - L) <= 1e-3 * abs(L)
    gd = g.cpu().double().numpy()
    for l, ((Wd, bd), (Wr, br)) in enumerate(zip(R.unpack(gd, m.in_pad), R.unpack(gref, m.in_pad))):
        for a, b, name in ((Wd, Wr, "W"), (bd, br, "b")):
            rel = np.linalg.norm(a - b) / np.linalg.norm(b)
            cos = float((a * b).sum() / (np.linalg.norm(a) * np.linalg.norm(b)))
            print(f"layer {l} {name}: rel {rel:.2e} cos {cos:.6f}")
            assert rel <= 2e-2 and cos >= 0.999, (l, name, rel, cos)


def test_adam_step(gpu):
    import torch
    m = _model(gpu, seed=5)
    n = 512
    _, Xd = _inputs(m, n, 6)
    y = torch.zeros(n, device=gpu)
    bk = torch.zeros(n, dtype=torch.int32, device=gpu)
    wr = torch.ones(n, device=gpu)
    w0 = m.w.cpu().double().numpy()
    g = torch.empty_like(m.w)
    m1 = torch.zeros_like(m.w)
    m2 = torch.zeros_like(m.w)
    loss = torch.zeros(1, device=gpu)
    ws = m.work(int(learned._lib().kpp_mlp_work_bytes(n)))
    net = m.net()
    r = learned._lib().kpp_mlp_train_step(ctypes.byref(net), learned._p(Xd), learned._p(y),
                                          learned._p(bk), learned._p(wr), n, learned._p(g),
                                          learned._p(m1), learned._p(m2), 1, 1e-3,
                                          learned._p(loss), learned._p(ws), ws.numel(),
                                          learned._stream())
    assert r == 0
    torch.cuda.synchronize()
    gd = g.cpu().double().numpy()
    w1, _, _ = R.adam(w0, gd, 0 * w0, 0 * w0, 1, 1e-3)
    assert np.allclose(m.w.cpu().double().numpy(), w1, rtol=1e-6, atol=1e-9)
    assert np.array_equal(m.wb.cpu().float().numpy(), m.w.cpu().to(torch.bfloat16).float().numpy())


@pytest.mark.parametrize("centres", [None, [12.0, 14.0, 15.5, 17.5, 39.0, 46.5, 53.0, 63.5]])
def test_schedule_outputs(gpu, centres):
    import torch
    m = _model(gpu, seed=8)
    n = 3000
    _, Xd = _inputs(m, n, 9)
    m.mu_y, m.sd_y = -1.0, 2.0
    m.centres = None if centres is None else np.array(centres, np.float32)
    hs, bk, order = m.schedule(Xd, 1e-6, 3600.0)
    out = m.forward(Xd)
    torch.cuda.synchronize()
    O = out.cpu().double().numpy()
    expo = (O[:, 0].astype(np.float32) * np.float32(2.0) + np.float32(-1.0)).astype(np.float64)
    h_ref = np.clip(10.0 ** expo, 1e-6, 3600.0)
    assert np.allclose(hs.cpu().numpy(), h_ref, rtol=1e-6)
    b = bk.cpu().numpy()
    assert np.array_equal(b, O[:, 1:9].argmax(1))
    o = order.cpu().numpy()
    assert np.array_equal(np.sort(o), np.arange(n))
    # launch order: expected step count (bucket centres) or expected bucket
    # index under the softmax, most steps first, on 64 levels between the
    # first and last centre (recomputed here in float64: a level can differ by
    # one where the device's FP32 softmax rounds the other way)
    lg = O[:, 1:9]
    p = np.exp(lg - lg.max(1, keepdims=True))
    c = np.arange(8.0) if centres is None else np.array(centres, np.float64)
    eb = (p * c).sum(1) / p.sum(1)
    key = 63 - np.clip(np.floor((eb - c[0]) * 63.0 / (c[-1] - c[0]) + 0.5), 0, 63).astype(int)
    ko = key[o]
    assert np.all(np.diff(ko) >= -1)
    assert np.mean(np.diff(ko) >= 0) > 0.99
    # stable: within one key, cells keep their index order
    same = np.diff(ko) == 0
    assert np.mean(np.diff(o)[same] > 0) > 0.99


def test_schedule_rejects_bad_centres(gpu):
    m = _model(gpu, seed=8)
    _, Xd = _inputs(m, 64, 9)
    m.centres = np.array([1, 2, 3, 4, 4, 5, 6, 7], np.float32)  # not increasing
    with pytest.raises(solver.KppError):
        m.schedule(Xd, 1e-6, 3600.0)


def test_features_match_oracle(gpu):
    import torch
    m = learned.StepModel(mech="saprc99", device=gpu)
    b = workloads.saprc99_uneven(ncells=777, seed=21)
    d = b.to_device(gpu)
    m.fit_inputs(d["C"], d["temp"], d["sun"])
    X = m.features(d["C"], d["temp"], d["sun"])
    torch.cuda.synchronize()
    ref_X = R.features(b.C, b.temp, b.sun, m.mu.cpu().numpy(), m.sd.cpu().numpy(), m.cfloor,
                       m.nspec, m.in_pad)
    got = X.cpu().double().numpy()
    # one BF16 unit where float32 log10 rounding differs
    assert np.max(np.abs(got - ref_X)) <= 2.0 ** -7 * max(1.0, np.max(np.abs(ref_X)))
    assert np.mean(got == ref_X) > 0.99


def test_step_log_matches_oracle(gpu):
    """First accepted step: device == oracle when the step sequences agree."""
    import torch
    b = workloads.saprc99_uneven(ncells=256, seed=31)
    L = ref.lib(b.mech)
    nr = ref.sizes(b.mech)["NREACT"]
    rc = np.zeros((nr, b.ncells))
    tmp = np.zeros(nr)
    for c in range(b.ncells):
        L.ros_ref_rconst(b.temp[c], b.sun[c], b.cfactor[c], b.tstart[c],
                         tmp.ctypes.data_as(ctypes.c_void_p))
        rc[:, c] = tmp
    span = b.tend - b.tstart
    o = ref.integrate(b, hstart=span, rconst=rc)
    d = b.to_device(gpu)
    hf = torch.zeros(b.ncells, dtype=torch.float64, device=gpu)
    code, ierr, istat, _ = solver.rosenbrock(
        b.mech, d["C"], d["temp"], d["cfactor"], tstart=d["tstart"], tend=d["tend"], atol=b.atol,
        rtol=b.rtol, IPAR=b.IPAR, RPAR=b.RPAR, sun=d["sun"], sun_mode=b.sun_mode,
        hstart=torch.tensor(span, device=gpu), rconst=torch.tensor(rc, device=gpu), hfirst=hf,
        variant=solver.PERSISTENT)
    torch.cuda.synchronize()
    assert code == 0
    assert np.array_equal(istat.cpu().numpy(), o["istat"])
    h = hf.cpu().numpy()
    assert np.all(o["hacc1"] > 0)
    assert np.max(np.abs(h - o["hacc1"]) / o["hacc1"]) <= 1e-12
    assert np.mean(h == o["hacc1"]) > 0.5


def test_learned_start_cuts_steps(gpu):
    """Train on step logs of 65,536 cells; held-out cells need >= 10 % fewer attempts."""
    import torch
    wl = workloads.saprc99_uneven
    model, metrics = learned.train_model("saprc99", wl, [41], 65536, gpu, epochs=20, log=print)
    b = wl(ncells=16384, seed=42)
    d = b.to_device(gpu)
    X = model.features(d["C"], d["temp"], d["sun"])
    span = float(np.max(b.tend - b.tstart))
    hs, bk, order = model.schedule(X, 1e-6, span)
    res = {}
    for name, kw in (("default", {}), ("learned", {"hstart": hs, "order": order})):
        C = d["C"].clone()
        code, ierr, istat, _ = solver.rosenbrock(
            b.mech, C, d["temp"], d["cfactor"], tstart=d["tstart"], tend=d["tend"], atol=b.atol,
            rtol=b.rtol, IPAR=b.IPAR, RPAR=b.RPAR, sun=d["sun"], sun_mode=b.sun_mode,
            variant=solver.PERSISTENT, **kw)
        torch.cuda.synchronize()
        assert code == 0 and bool((ierr == 1).all())
        res[name] = (float(istat[2].double().mean()), C.cpu().numpy())
    print(metrics, {k: v[0] for k, v in res.items()})
    assert res["learned"][0] <= 0.9 * res["default"][0]
    # both runs meet the same tolerance: for species above ATOL/RTOL they agree to a few RTOL
    nvar = solver.mechanism("saprc99").nvar
    a, c = res["learned"][1][:nvar], res["default"][1][:nvar]
    big = np.abs(c) > 1e3
    assert np.max(np.abs(a - c)[big] / np.abs(c)[big]) < 1e-2
