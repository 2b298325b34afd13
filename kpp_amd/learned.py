"""Learned step scheduler: training and inference around the C ABI of mlp.hip.

A small network predicts, from a cell's initial state (log10 concentrations
of every species, TEMP, SUN), the size of its first accepted Rosenbrock step
and a step-count bucket.  At launch time the prediction sets each cell's
starting step (Rosenbrock's RPAR[2], per cell) and the launch order (cells
predicted to need the most steps first).  Rosenbrock's error control still
accepts or rejects every step, so a poor prediction costs steps, not
accuracy.

Step logs come from the solver itself (``log_steps``): one integration
started at the largest allowed step (the whole interval) records each cell's
first accepted step -- the step the error control settles on at t0 -- and a
second integration started at that step records the step count the
scheduler's policy leads to.  Training runs on the GPU through
``kpp_mlp_train_step`` (forward, loss, backward on BF16 matrix cores, Adam
on FP32 master weights).

CLI:  python -m kpp_amd.learned --out kpp_amd/weights/saprc99_step_mlp.npz
"""
from __future__ import annotations

import ctypes
import time
from dataclasses import dataclass, field

import numpy as np

from . import solver

HID, NOUT, NBUCKET = 128, 16, 8


class _Net(ctypes.Structure):
    _fields_ = [("nin", ctypes.c_int), ("in_pad", ctypes.c_int), ("w", ctypes.c_void_p),
                ("wb", ctypes.c_void_p)]


def _lib():
    L = solver.library()
    if not getattr(L, "_mlp_bound", False):
        vp, ci, cs, cf, cd = (ctypes.c_void_p, ctypes.c_int, ctypes.c_size_t, ctypes.c_float,
                              ctypes.c_double)
        L.kpp_mlp_param_count.restype = cs
        L.kpp_mlp_param_count.argtypes = [ci]
        L.kpp_mlp_param_offsets.argtypes = [ci, vp, vp]
        L.kpp_mlp_sync_bf16.argtypes = [vp, vp]
        L.kpp_mlp_features.argtypes = [ci, ci, ci, vp, vp, vp, vp, vp, cd, ci, vp, vp]
        L.kpp_mlp_work_bytes.restype = cs
        L.kpp_mlp_work_bytes.argtypes = [ci]
        L.kpp_mlp_forward.argtypes = [vp, vp, ci, vp, vp, cs, vp]
        L.kpp_mlp_grad.argtypes = [vp, vp, vp, vp, vp, ci, vp, vp, vp, cs, vp]
        L.kpp_mlp_train_step.argtypes = [vp, vp, vp, vp, vp, ci, vp, vp, vp, ci, cf, vp, vp, cs,
                                         vp]
        L.kpp_mlp_schedule_work_bytes.restype = cs
        L.kpp_mlp_schedule_work_bytes.argtypes = [ci]
        L.kpp_mlp_schedule.argtypes = [vp, vp, ci, cf, cf, cd, cd, vp, vp, vp, vp, vp, cs, vp]
        L._mlp_bound = True
    return L


def _check(r, what):
    if r != 0:
        raise solver.KppError(f"{what} failed with {r}")


def _p(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else None


def _stream():
    import torch
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


@dataclass
class StepModel:
    """Network parameters (device), input/output standardisation, bucket edges."""
    mech: str
    device: object
    seed: int = 0
    cfloor: float = 1.0e-3
    nin: int = 0
    in_pad: int = 0
    w: object = None        # torch float32 [P]
    wb: object = None       # torch bfloat16 [P]
    mu: object = None       # torch float32 [in_pad]
    sd: object = None
    mu_y: float = 0.0
    sd_y: float = 1.0
    h_scale: float = 1.0    # starting step = h_scale x the predicted first step
    edges: np.ndarray = field(default_factory=lambda: np.zeros(NBUCKET - 1))
    # mean step count of each bucket in the training data: the launch order
    # sorts by the expected count under the bucket softmax (None: by the
    # expected bucket index)
    centres: np.ndarray = None
    _work: object = None

    def __post_init__(self):
        import torch
        info = solver.mechanism(self.mech)
        self.nspec = info.nspec
        self.nin = self.nspec + 2
        self.in_pad = (self.nin + 31) // 32 * 32
        P = int(_lib().kpp_mlp_param_count(self.in_pad))
        if self.w is None:
            self.w = torch.tensor(self.init_params(), dtype=torch.float32, device=self.device)
        self.wb = torch.empty(P, dtype=torch.bfloat16, device=self.device)
        if self.mu is None:
            self.mu = torch.zeros(self.in_pad, dtype=torch.float32, device=self.device)
            self.sd = torch.ones(self.in_pad, dtype=torch.float32, device=self.device)
        self.sync()

    # ---- parameters ----
    def offsets(self):
        w_off = (ctypes.c_size_t * 4)()
        b_off = (ctypes.c_size_t * 4)()
        _check(_lib().kpp_mlp_param_offsets(self.in_pad, w_off, b_off), "kpp_mlp_param_offsets")
        return list(w_off), list(b_off)

    def init_params(self):
        """Uniform(+-sqrt(6/fan_in)) weights on the used inputs, zero biases and padding."""
        rng = np.random.default_rng(self.seed)
        P = int(_lib().kpp_mlp_param_count(self.in_pad))
        w_off, _ = self.offsets()
        flat = np.zeros(P, np.float32)
        shapes = [(HID, self.in_pad, self.nin), (HID, HID, HID), (HID, HID, HID),
                  (NOUT, HID, HID)]
        for l, (no, ki, used) in enumerate(shapes):
            W = np.zeros((no, ki), np.float32)
            lim = np.sqrt(6.0 / used) * (0.5 if l == 3 else 1.0)
            rows = no if l < 3 else 1 + NBUCKET
            W[:rows, :used] = rng.uniform(-lim, lim, (rows, used))
            flat[w_off[l]:w_off[l] + no * ki] = W.ravel()
        return flat

    def net(self) -> _Net:
        return _Net(self.nin, self.in_pad, self.w.data_ptr(), self.wb.data_ptr())

    def sync(self):
        n = self.net()
        _check(_lib().kpp_mlp_sync_bf16(ctypes.byref(n), _stream()), "kpp_mlp_sync_bf16")

    def work(self, nbytes):
        import torch
        if self._work is None or self._work.numel() < nbytes:
            self._work = torch.empty(max(nbytes, 256), dtype=torch.uint8, device=self.device)
        return self._work

    # ---- inputs ----
    def features(self, C, temp, sun, out=None):
        """BF16 inputs [n, in_pad] of the cells of a device batch (C is [NSPEC, ld])."""
        import torch
        n = int(temp.shape[0])
        X = out if out is not None else torch.empty((n, self.in_pad), dtype=torch.bfloat16,
                                                    device=self.device)
        _check(_lib().kpp_mlp_features(n, int(C.shape[1]), self.nspec, _p(C), _p(temp), _p(sun),
                                       _p(self.mu), _p(self.sd), self.cfloor, self.in_pad, _p(X),
                                       _stream()), "kpp_mlp_features")
        return X

    def fit_inputs(self, C, temp, sun):
        """Set the input standardisation from a device batch."""
        import torch
        raw = torch.log10(torch.clamp(C[:self.nspec].T, min=self.cfloor)).float()
        cols = [raw, temp.float()[:, None],
                (sun.float() if sun is not None else torch.zeros_like(temp).float())[:, None]]
        A = torch.cat(cols, 1)
        mu = torch.zeros(self.in_pad, dtype=torch.float32, device=self.device)
        sd = torch.ones(self.in_pad, dtype=torch.float32, device=self.device)
        mu[:self.nin] = A.mean(0)
        s = A.std(0)
        sd[:self.nin] = torch.where(s > 1e-6, s, torch.ones_like(s))
        self.mu, self.sd = mu, sd

    # ---- network ----
    def forward(self, X):
        import torch
        n = int(X.shape[0])
        out = torch.empty((n, NOUT), dtype=torch.float32, device=self.device)
        ws = self.work(int(_lib().kpp_mlp_work_bytes(n)))
        net = self.net()
        _check(_lib().kpp_mlp_forward(ctypes.byref(net), _p(X), n, _p(out), _p(ws), ws.numel(),
                                      _stream()), "kpp_mlp_forward")
        return out

    def grad(self, X, y, bucket, wreg):
        """(summed loss, gradient of the mean loss) for one batch, on the device."""
        import torch
        n = int(X.shape[0])
        g = torch.empty_like(self.w)
        loss = torch.zeros(1, dtype=torch.float32, device=self.device)
        ws = self.work(int(_lib().kpp_mlp_work_bytes(n)))
        net = self.net()
        _check(_lib().kpp_mlp_grad(ctypes.byref(net), _p(X), _p(y), _p(bucket), _p(wreg), n, _p(g),
                                   _p(loss), _p(ws), ws.numel(), _stream()), "kpp_mlp_grad")
        return loss, g

    def train(self, X, y, bucket, wreg, *, epochs=40, batch=8192, lr=3e-3, seed=0, log=None):
        """Adam over shuffled mini-batches; cosine learning-rate decay."""
        import torch
        n = int(X.shape[0])
        g = torch.empty_like(self.w)
        m1 = torch.zeros_like(self.w)
        m2 = torch.zeros_like(self.w)
        loss = torch.zeros(1, dtype=torch.float32, device=self.device)
        ws = self.work(int(_lib().kpp_mlp_work_bytes(batch)))
        net = self.net()
        L = _lib()
        gen = torch.Generator(device="cpu").manual_seed(seed)
        steps_per_epoch = n // batch
        total = epochs * steps_per_epoch
        step = 0
        hist = []
        for ep in range(epochs):
            perm = torch.randperm(n, generator=gen).to(self.device)
            Xs, ys, bs, ws_ = X[perm], y[perm], bucket[perm], wreg[perm]
            loss.zero_()
            for k in range(steps_per_epoch):
                step += 1
                sl = slice(k * batch, (k + 1) * batch)
                lr_t = lr * 0.5 * (1.0 + np.cos(np.pi * (step - 1) / total))
                _check(L.kpp_mlp_train_step(ctypes.byref(net), _p(Xs[sl]), _p(ys[sl]),
                                            _p(bs[sl]), _p(ws_[sl]), batch, _p(g), _p(m1), _p(m2),
                                            step, float(lr_t), _p(loss), _p(ws), ws.numel(),
                                            _stream()), "kpp_mlp_train_step")
            # loss of the last batch of the epoch
            hist.append(float(loss.item()) / batch)
            if log:
                log(f"[learned] epoch {ep + 1}/{epochs}: last-batch loss {hist[-1]:.4f}")
        return hist

    def schedule(self, X, hmin, hmax, want_order=True, out=None):
        """(hstart float64 [n], bucket int32 [n], order int32 [n] or None) on the device.

        ``out``: a previous call's result tuple to write into (no allocation)."""
        import torch
        n = int(X.shape[0])
        if out is not None:
            hs, bk, order = out
        else:
            hs = torch.empty(n, dtype=torch.float64, device=self.device)
            bk = torch.empty(n, dtype=torch.int32, device=self.device)
            order = torch.empty(n, dtype=torch.int32, device=self.device) if want_order else None
        ws = self.work(int(_lib().kpp_mlp_schedule_work_bytes(n)))
        net = self.net()
        mu = float(self.mu_y + np.log10(self.h_scale))  # 10^(o*sd + mu) * h_scale
        cen = None
        if self.centres is not None:
            cen = (ctypes.c_float * NBUCKET)(*[float(c) for c in self.centres])
        _check(_lib().kpp_mlp_schedule(ctypes.byref(net), _p(X), n, mu,
                                       float(self.sd_y), float(hmin), float(hmax), cen, _p(hs),
                                       _p(bk), _p(order), _p(ws), ws.numel(), _stream()),
               "kpp_mlp_schedule")
        return hs, bk, order

    # ---- targets ----
    def targets(self, hfirst, nstp):
        """Standardised log10 first step, bucket index, regression weight (device tensors)."""
        import torch
        ok = hfirst > 0
        lh = torch.log10(torch.where(ok, hfirst, torch.ones_like(hfirst)))
        y = ((lh - self.mu_y) / self.sd_y).float()
        edges = torch.tensor(self.edges, dtype=torch.float64, device=self.device)
        b = torch.bucketize(nstp.double(), edges, right=True).int()
        return y, b, ok.float()

    def fit_targets(self, hfirst, nstp):
        import torch
        ok = hfirst > 0
        lh = torch.log10(hfirst[ok])
        self.mu_y = float(lh.mean())
        self.sd_y = float(lh.std()) or 1.0
        q = torch.quantile(nstp.double().cpu(), torch.linspace(0, 1, NBUCKET + 1, dtype=torch.float64)[1:-1])
        self.edges = q.numpy() + 0.5
        self.centres = bucket_centres(nstp.cpu().numpy(), self.edges)

    # ---- persistence ----
    def save(self, path):
        np.savez(path, mech=np.array(self.mech), w=self.w.cpu().numpy(),
                 mu=self.mu.cpu().numpy(), sd=self.sd.cpu().numpy(),
                 mu_y=np.float64(self.mu_y), sd_y=np.float64(self.sd_y), edges=self.edges,
                 h_scale=np.float64(self.h_scale),
                 **({} if self.centres is None else {"centres": np.asarray(self.centres,
                                                                          np.float32)}),
                 cfloor=np.float64(self.cfloor), in_pad=np.int32(self.in_pad))

    @classmethod
    def load(cls, path, device):
        import torch
        z = np.load(path, allow_pickle=False)
        m = cls(mech=str(z["mech"]), device=device, cfloor=float(z["cfloor"]),
                w=torch.tensor(z["w"], dtype=torch.float32, device=device),
                mu=torch.tensor(z["mu"], dtype=torch.float32, device=device),
                sd=torch.tensor(z["sd"], dtype=torch.float32, device=device),
                mu_y=float(z["mu_y"]), sd_y=float(z["sd_y"]), edges=np.array(z["edges"]),
                h_scale=float(z["h_scale"]) if "h_scale" in z.files else 1.0,
                centres=np.array(z["centres"], np.float32) if "centres" in z.files else None)
        if m.in_pad != int(z["in_pad"]):
            raise ValueError("weights file does not match the mechanism's input width")
        return m


WEIGHTS_MAGIC = b"KPPWGT1\0"


def export_flat(npz_path, out_path, nspec=None):
    """Write a trained .npz as the flat weights image C hosts load with
    kpp_sched_load (layout in kpp_amd/csrc/sched.cpp).  numpy only: runs
    without a GPU (the build exports the shipped weights this way)."""
    import json
    import struct
    z = np.load(npz_path, allow_pickle=False)
    mech = str(z["mech"])
    if nspec is None:
        nspec = json.loads((solver.GEN / f"{mech}_stats.json").read_text())["NSPEC"]
    in_pad = int(z["in_pad"])
    w = np.ascontiguousarray(z["w"], "<f4")
    has_c = "centres" in z.files
    cen = np.ascontiguousarray(z["centres"] if has_c else np.zeros(NBUCKET), "<f4")
    hs = float(z["h_scale"]) if "h_scale" in z.files else 1.0
    name = mech.encode()
    if len(name) > 31:
        raise ValueError("mechanism name too long for the weights image")
    with open(out_path, "wb") as f:
        f.write(WEIGHTS_MAGIC)
        f.write(name.ljust(32, b"\0"))
        f.write(struct.pack("<4i", nspec + 2, in_pad, w.size, 1 if has_c else 0))
        f.write(struct.pack("<4d", float(z["cfloor"]), float(z["mu_y"]), float(z["sd_y"]), hs))
        f.write(cen.tobytes())
        f.write(np.ascontiguousarray(z["mu"], "<f4").tobytes())
        f.write(np.ascontiguousarray(z["sd"], "<f4").tobytes())
        f.write(w.tobytes())


def bucket_centres(nstp, edges):
    """Mean step count of each bucket (the buckets of ``targets``), increasing.

    An empty bucket takes the middle of its neighbours' edges."""
    nstp = np.asarray(nstp, np.float64)
    b = np.searchsorted(np.asarray(edges, np.float64), nstp, side="right")
    e = np.concatenate([[nstp.min()], np.asarray(edges, np.float64), [nstp.max()]])
    c = np.array([nstp[b == k].mean() if (b == k).any() else 0.5 * (e[k] + e[k + 1])
                  for k in range(NBUCKET)])
    for k in range(1, NBUCKET):  # strictly increasing, as the schedule requires
        c[k] = max(c[k], np.nextafter(np.float32(c[k - 1]), np.float32(np.inf)))
    return c.astype(np.float32)


def log_steps(b, d, ws=None):
    """Step logs of a device batch: (hfirst, nstp).

    hfirst: first accepted step of an integration started at the whole
    interval (the largest step Rosenbrock allows); nstp: step attempts of an
    integration started at hfirst.  Inputs are left unchanged.
    """
    import torch
    n = b.ncells
    C = d["C"].clone()
    hfirst = torch.zeros(n, dtype=torch.float64, device=C.device)
    span = (d["tend"] - d["tstart"]).abs()
    kw = dict(tstart=d["tstart"], tend=d["tend"], atol=b.atol, rtol=b.rtol, IPAR=b.IPAR,
              RPAR=b.RPAR, sun=d["sun"], sun_mode=b.sun_mode, variant=solver.PERSISTENT,
              workspace=ws)
    code, ierr, istat, _ = solver.rosenbrock(b.mech, C, d["temp"], d["cfactor"], hstart=span,
                                             hfirst=hfirst, **kw)
    if code != 0:
        raise solver.KppError(f"step-log integration returned {code}")
    C.copy_(d["C"])
    hs = torch.where(hfirst > 0, hfirst, torch.full_like(hfirst, 1e-6))
    code, ierr2, istat2, _ = solver.rosenbrock(b.mech, C, d["temp"], d["cfactor"], hstart=hs, **kw)
    if code != 0:
        raise solver.KppError(f"step-log integration returned {code}")
    return hfirst, istat2[2].clone()


def train_model(mech, wl_fn, seeds, ncells, device, *, epochs=40, batch=8192, lr=3e-3,
                holdout=0.1, log=print):
    """Log steps on fresh batches of a workload, train a StepModel, report holdout metrics."""
    import torch
    t0 = time.time()
    feats, hf, ns = [], [], []
    model = StepModel(mech=mech, device=device)
    for i, s in enumerate(seeds):
        b = wl_fn(ncells=ncells, seed=s)
        d = b.to_device(device)
        if i == 0:
            model.fit_inputs(d["C"], d["temp"], d["sun"])
        h, n = log_steps(b, d)
        feats.append(model.features(d["C"], d["temp"], d["sun"]))
        hf.append(h)
        ns.append(n)
        del d
    X = torch.cat(feats)
    hfirst = torch.cat(hf)
    nstp = torch.cat(ns)
    torch.cuda.synchronize()
    log(f"[learned] logged {X.shape[0]} samples in {time.time() - t0:.1f}s")
    N = X.shape[0]
    perm = torch.randperm(N, generator=torch.Generator().manual_seed(1)).to(device)
    nho = int(N * holdout)
    ho, tr = perm[:nho], perm[nho:]
    model.fit_targets(hfirst[tr], nstp[tr])
    y, bk, wr = model.targets(hfirst, nstp)
    t1 = time.time()
    hist = model.train(X[tr], y[tr], bk[tr], wr[tr], epochs=epochs, batch=batch, lr=lr, log=log)
    torch.cuda.synchronize()
    log(f"[learned] trained {epochs} epochs on {len(tr)} samples in {time.time() - t1:.1f}s")
    out = model.forward(X[ho])
    ok = wr[ho] > 0
    rmse = float(torch.sqrt(((out[:, 0] - y[ho])[ok] ** 2).mean()) * model.sd_y)
    acc = float((out[:, 1:1 + NBUCKET].argmax(1).int() == bk[ho]).float().mean())
    near = float(((out[:, 1:1 + NBUCKET].argmax(1).int() - bk[ho]).abs() <= 1).float().mean())
    # starting-step scale: the one of a few that needs the fewest attempts on
    # (training) cells of the first batch
    b = wl_fn(ncells=min(ncells, 65536), seed=seeds[0])
    d = b.to_device(device)
    span = float(np.max(np.abs(b.tend - b.tstart)))
    Xs = model.features(d["C"], d["temp"], d["sun"])
    tried = {}
    for sc in (1.0, 1.5, 2.0):
        model.h_scale = sc
        hs, _, order = model.schedule(Xs, 1e-6, span)
        code, ierr, istat, _ = solver.rosenbrock(
            b.mech, d["C"].clone(), d["temp"], d["cfactor"], tstart=d["tstart"], tend=d["tend"],
            atol=b.atol, rtol=b.rtol, IPAR=b.IPAR, RPAR=b.RPAR, sun=d["sun"],
            sun_mode=b.sun_mode, hstart=hs, order=order, variant=solver.PERSISTENT)
        tried[sc] = float(istat[2].double().mean())
    model.h_scale = min(tried, key=tried.get)
    log(f"[learned] starting-step scale {model.h_scale} (attempts per cell {tried})")
    metrics = {"samples": int(N), "holdout": int(nho), "log10_h_rmse": rmse,
               "bucket_accuracy": acc, "bucket_within_1": near, "h_scale": model.h_scale,
               "h_scale_attempts": tried, "loss_history": hist}
    log(f"[learned] holdout: log10(h) RMSE {rmse:.3f}, bucket accuracy {acc:.3f} "
        f"(within one bucket {near:.3f})")
    return model, metrics


def main():
    import argparse
    import json
    import torch
    from . import workloads
    ap = argparse.ArgumentParser(description="train the step scheduler on solver step logs")
    ap.add_argument("--workload", default="saprc99_uneven")
    ap.add_argument("--seeds", default="101,102")
    ap.add_argument("--ncells", type=int, default=1048576)
    ap.add_argument("--epochs", type=int, default=40)
    ap.add_argument("--out", required=True)
    ap.add_argument("--centres-for", default=None,
                    help="instead of training: add bucket centres to this trained weights file "
                         "from the step logs of the same seeds and training split, write --out")
    a = ap.parse_args()
    wl = workloads.WORKLOADS[a.workload]
    dev = torch.device("cuda:0")
    mech = wl(ncells=64, seed=0).mech
    if a.centres_for:
        model = StepModel.load(a.centres_for, dev)
        ns = []
        for s in [int(x) for x in a.seeds.split(",")]:
            b = wl(ncells=a.ncells, seed=s)
            ns.append(log_steps(b, b.to_device(dev))[1])
        nstp = torch.cat(ns)
        N = nstp.shape[0]
        perm = torch.randperm(N, generator=torch.Generator().manual_seed(1)).to(dev)
        tr = perm[int(N * 0.1):]  # the training split of train_model
        model.centres = bucket_centres(nstp[tr].cpu().numpy(), model.edges)
        model.save(a.out)
        print(json.dumps({"centres": [float(c) for c in model.centres]}))
        return
    model, metrics = train_model(mech, wl, [int(s) for s in a.seeds.split(",")], a.ncells, dev,
                                 epochs=a.epochs)
    model.save(a.out)
    print(json.dumps(metrics))


if __name__ == "__main__":
    main()
