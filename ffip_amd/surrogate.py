"""Design surrogate on the matrix cores.

No counterpart in the reference.  A small 3-D convolutional network maps the
design density of the inverse-design workload (21 x 101 x 101 density nodes of a
100 x 100 x 20-cell two-medium box) to the Ez DFT intensity on the 21 x 21
objective patch at four frequencies.  It is trained on (density, patch DFT)
pairs that the GPU FDTD solver produces, and it chooses the step lengths of an
adjoint gradient ascent; every design it proposes is re-judged by the full FDTD
solve (and the adjoint solve when accepted), which stays the fallback.

Network: four 3x3x3 convolutions (channels 1 -> 32 -> 32 -> 32 -> 4), each
striding 2 along z (21 -> 11 -> 6 -> 3 -> 2 planes) with x/y dilations 1, 2, 4, 8
(receptive field +-15 cells), ReLU after the first three; the head averages over
z and crops the 21 x 21 patch; the 4 output channels are the intensities
|E_f|^2 at the 4 frequencies (the objective is the patch sum of the 0.1
channel).  Each convolution runs as im2col + a BF16 MFMA GEMM with FP32
accumulation (ffip_sg_*), its backward pass as two more GEMMs + col2im; Adam
updates FP32 master weights and refreshes their bf16 copies.  torch only holds
the device memory and supplies the stream.
"""
from __future__ import annotations

import ctypes
import math
import time

import numpy as np

from . import _lib
from ._lib import FfipError, _check

FREQS = (0.09, 0.095, 0.1, 0.105)
PATCH = 21  # objective patch nodes per side (20 x 20 cells)


def _torch():
    import torch
    if not torch.cuda.is_available():
        raise FfipError("the surrogate kernels need the GPU; there is no CPU path")
    return torch


def _vp(t):
    return ctypes.c_void_p(t.data_ptr())


def _arr(vals):
    a = (ctypes.c_int64 * len(vals))(*[int(v) for v in vals])
    return ctypes.cast(a, ctypes.c_void_p), a  # keep the array alive with the pointer


def _stream():
    return ctypes.c_void_p(_torch().cuda.current_stream().cuda_stream)


# ---------------------------------------------------------------- kernel calls
def gemm(A, B, C, M, N, K, ta=False, tb=False, lda=None, ldb=None, ldc=None, accumulate=False, ksplit=1,
         work=None):
    """C (+)= op(A) op(B) on the matrix cores (A, B bf16 tensors, C FP32).  With
    a FP32 workspace `work` of at least ksplit * M * N elements, the split-K
    partial tiles are added in a fixed order (the same bits on every run)."""
    lda = lda if lda is not None else (M if ta else K)
    ldb = ldb if ldb is not None else (K if tb else N)
    ldc = ldc if ldc is not None else N
    if work is None:
        _check(_lib.lib().ffip_sg_gemm(int(ta), int(tb), M, N, K, _vp(A), lda, _vp(B), ldb, _vp(C), ldc,
                                       int(accumulate), int(ksplit), _stream()))
    else:
        _check(_lib.lib().ffip_sg_gemm_ws(int(ta), int(tb), M, N, K, _vp(A), lda, _vp(B), ldb, _vp(C), ldc,
                                          int(accumulate), int(ksplit), _vp(work), int(work.numel()), _stream()))


def conv_out(shape, stride):
    return tuple((n - 1) // s + 1 for n, s in zip(shape, stride))


class Conv3:
    """One 3x3x3 convolution layer, channels-last, weights [27*cin, cout]."""

    def __init__(self, cin, cout, stride, dil, relu, rng):
        torch = _torch()
        self.cin, self.cout, self.stride, self.dil, self.relu = cin, cout, tuple(stride), tuple(dil), relu
        K = 27 * cin
        w = rng.standard_normal((K, cout)) * math.sqrt(2.0 / K)
        dev = "cuda"
        self.w = torch.tensor(w, dtype=torch.float32, device=dev)
        self.wb = self.w.to(torch.bfloat16)
        self.b = torch.zeros(cout, dtype=torch.float32, device=dev)
        self.gw = torch.zeros_like(self.w)
        self.gb = torch.zeros_like(self.b)
        self.m = [torch.zeros_like(self.w), torch.zeros_like(self.b)]
        self.v = [torch.zeros_like(self.w), torch.zeros_like(self.b)]
        self.cache = None
        self.work = None  # split-K workspace of the weight gradient

    def geom(self, batch, shape):
        return [batch, self.cin] + list(shape) + list(self.stride) + list(self.dil)

    def forward(self, x, batch, shape, keep):
        torch = _torch()
        out = conv_out(shape, self.stride)
        M = batch * out[0] * out[1] * out[2]
        K = 27 * self.cin
        col = torch.empty((M, K), dtype=torch.bfloat16, device="cuda")
        g, keep_g = _arr(self.geom(batch, shape))
        _check(_lib.lib().ffip_sg_im2col(g, _vp(x), _vp(col), _stream()))
        Y = torch.empty((M, self.cout), dtype=torch.float32, device="cuda")
        gemm(col, self.wb, Y, M, self.cout, K)
        a = torch.empty((M, self.cout), dtype=torch.bfloat16, device="cuda")
        Z = torch.empty_like(Y) if keep else None
        _check(_lib.lib().ffip_sg_bias_act(M, self.cout, _vp(Y), _vp(self.b), int(self.relu), _vp(a),
                                           _vp(Z) if Z is not None else None, _stream()))
        self.cache = (batch, shape, col, Z, M) if keep else None
        return a, out

    def backward(self, dA, need_dx=True):
        """dA: FP32 [M, cout] gradient of this layer's output; accumulates gw, gb;
        returns the FP32 gradient of the input [batch, z, y, x, cin]."""
        torch = _torch()
        batch, shape, col, Z, M = self.cache
        K = 27 * self.cin
        dZb = torch.empty((M, self.cout), dtype=torch.bfloat16, device="cuda")
        dZf = torch.empty((M, self.cout), dtype=torch.float32, device="cuda")
        _check(_lib.lib().ffip_sg_act_back(M * self.cout, _vp(dA), _vp(Z), int(self.relu), _vp(dZb), _vp(dZf),
                                           _stream()))
        _check(_lib.lib().ffip_sg_colsum(M, self.cout, _vp(dZf), _vp(self.gb), 0, _stream()))
        # weight gradient col^T dZ: long K (= positions), split over workgroups,
        # the partial tiles added in a fixed order so that training is repeatable
        ks = max(1, M // 8192)
        if self.work is None or self.work.numel() < ks * K * self.cout:
            self.work = torch.empty(ks * K * self.cout, dtype=torch.float32, device="cuda")
        gemm(col, dZb, self.gw, K, self.cout, M, ta=True, lda=K, ldb=self.cout, ksplit=ks, work=self.work)
        if not need_dx:
            return None
        dcol = torch.empty((M, K), dtype=torch.float32, device="cuda")
        gemm(dZb, self.wb, dcol, M, K, self.cout, tb=True, lda=self.cout, ldb=self.cout)
        dX = torch.empty((batch,) + tuple(shape) + (self.cin,), dtype=torch.float32, device="cuda")
        g, keep_g = _arr(self.geom(batch, shape))
        _check(_lib.lib().ffip_sg_col2im(g, _vp(dcol), _vp(dX), _stream()))
        return dX

    def adam(self, t, lr, b1=0.9, b2=0.999, eps=1e-8):
        hp = (ctypes.c_float * 4)(lr, b1, b2, eps)
        hpp = ctypes.cast(hp, ctypes.c_void_p)
        _check(_lib.lib().ffip_sg_adam(self.w.numel(), _vp(self.w), _vp(self.gw), _vp(self.m[0]), _vp(self.v[0]),
                                       hpp, t, _vp(self.wb), _stream()))
        _check(_lib.lib().ffip_sg_adam(self.b.numel(), _vp(self.b), _vp(self.gb), _vp(self.m[1]), _vp(self.v[1]),
                                       hpp, t, None, _stream()))


class Surrogate:
    """density [B, 21, 101, 101] in [0, 1] -> patch intensities |E_f|^2 [B, 21, 21, 4]."""

    def __init__(self, seed=0, shape=(21, 101, 101), crop=(40, 40, PATCH, PATCH), width=32, nout=len(FREQS)):
        rng = np.random.default_rng(seed)
        self.shape, self.crop, self.nout = tuple(shape), crop, nout
        dil = [(1, 1, 1), (1, 2, 2), (1, 4, 4), (1, 8, 8)]
        chans = [1, width, width, width, nout]
        self.layers = [Conv3(chans[q], chans[q + 1], (2, 1, 1), dil[q], q < 3, rng) for q in range(4)]
        self.scale = 1.0  # target normalisation (fields / scale are learned)
        self.t = 0
        self._head = None

    def _input(self, rho):
        torch = _torch()
        r = torch.as_tensor(np.asarray(rho, dtype=np.float32), device="cuda")
        return (2.0 * r - 1.0).to(torch.bfloat16).contiguous()  # [B, z, y, x] == channels-last with C = 1

    def forward(self, rho, keep=False):
        torch = _torch()
        x = self._input(rho)
        batch = x.shape[0]
        shape = self.shape
        for L in self.layers:
            x, shape = L.forward(x, batch, shape, keep)
        y0, x0, ny, nx = self.crop
        h = [batch, shape[0], shape[1], shape[2], self.nout, y0, x0, ny, nx]
        self._head = (h, shape, batch)
        out = torch.empty((batch, ny, nx, self.nout), dtype=torch.float32, device="cuda")
        hp, keep_h = _arr(h)
        _check(_lib.lib().ffip_sg_head(hp, _vp(x), _vp(out), _stream()))
        return out

    def backward(self, dout, need_dx=False):
        torch = _torch()
        h, shape, batch = self._head
        dA = torch.empty((batch,) + tuple(shape) + (self.nout,), dtype=torch.float32, device="cuda")
        hp, keep_h = _arr(h)
        _check(_lib.lib().ffip_sg_head_back(hp, _vp(dout.contiguous()), _vp(dA), _stream()))
        g = dA.reshape(-1, self.nout)
        for q in range(len(self.layers) - 1, -1, -1):
            L = self.layers[q]
            dx = L.backward(g, need_dx=need_dx or q > 0)
            if q > 0:
                g = dx.reshape(-1, L.cin)
            elif need_dx:
                return dx.reshape((batch,) + tuple(self.shape)) * 2.0  # d/d rho of the 2 rho - 1 input map
        return None

    # ----------------------------------------------------------- training
    def _steps(self, rhos, tgt, order, batch, lr):
        tot, cnt = 0.0, 0
        for s in range(0, len(order) - batch + 1, batch):
            idx = order[s:s + batch]
            out = self.forward(rhos[idx], keep=True)
            diff = out - tgt[idx]
            tot += float((diff * diff).mean()) * len(idx)
            cnt += len(idx)
            self.backward(2.0 * diff / diff.numel())
            self.t += 1
            for L in self.layers:
                L.adam(self.t, lr)
        return tot / max(cnt, 1)

    def fit(self, rhos, targets, epochs=60, batch=8, lr=2e-3, seed=0, log=None, augment=False):
        """Mean-squared error on targets [N, 21, 21, nout] divided by self.scale
        (set here from the targets and kept for later fine-tuning); augment adds
        the mirror images (mirror_augment)."""
        torch = _torch()
        if augment:
            rhos, targets = mirror_augment(rhos, targets)
        self.scale = float(np.mean(np.abs(targets))) or 1.0
        tgt = torch.as_tensor(np.asarray(targets, dtype=np.float32) / self.scale, device="cuda")
        rng = np.random.default_rng(seed)
        hist = []
        for ep in range(epochs):
            hist.append(self._steps(rhos, tgt, rng.permutation(len(rhos)), batch, lr))
            if log:
                log("epoch %d: train mse %.4e" % (ep, hist[-1]))
        return hist

    def fine_tune(self, rhos, targets, steps=20, batch=4, lr=1e-3, seed=0, augment=False):
        """A few Adam steps on a small set of (density, target) pairs (online
        updates from the solves of an optimisation run)."""
        torch = _torch()
        if augment:
            rhos, targets = mirror_augment(rhos, targets)
        tgt = torch.as_tensor(np.asarray(targets, dtype=np.float32) / self.scale, device="cuda")
        rng = np.random.default_rng(seed + self.t)
        n = len(rhos)
        b = min(batch, n)
        for _ in range(steps):
            self._steps(rhos, tgt, rng.choice(n, size=b, replace=False), b, lr)

    def predict(self, rho):
        """Patch intensities |E_f|^2 [B, 21, 21, 4] (numpy) for densities [B, 21, 101, 101]."""
        return self.forward(np.asarray(rho)[None] if np.ndim(rho) == 3 else rho).cpu().numpy() * self.scale

    def values(self, rhos, fidx=FREQS.index(0.1)):
        """F = sum over the patch of |E(f)|^2 for a batch of densities (forward only)."""
        out = self.forward(np.asarray(rhos), keep=False)
        return (out[..., fidx].sum(dim=(1, 2)) * self.scale).cpu().numpy().astype(np.float64)

    def objective(self, rho, fidx=FREQS.index(0.1)):
        """F = sum over the patch of |E(f)|^2 (the inverse-design objective) and
        its gradient with respect to the density, from the network."""
        torch = _torch()
        out = self.forward(np.asarray(rho)[None], keep=True)
        F = float(out[..., fidx].sum().item()) * self.scale
        dout = torch.zeros_like(out)
        dout[..., fidx] = self.scale
        g = self.backward(dout, need_dx=True)
        return F, g[0].cpu().numpy().astype(np.float64)


# ---------------------------------------------------------------- data and loop
def design_problem(steps=3000, precision="fp32"):
    """The inverse-design workload (200x200x100, design box 100x100x20, eps 1..12,
    F = sum |DFT Ez(0.1)|^2 on the 20x20 patch at z = +30) with an extra patch
    monitor at the four surrogate frequencies."""
    from .adjoint import c3_problem
    from .geom import Vector3
    p = c3_problem(steps=steps, precision=precision)
    mon = p.fwd.add_dft_fields(center=Vector3(0, 0, 30.0), size=Vector3(20, 20, 0), frequency=list(FREQS),
                               field_component="Ez")
    return p, mon


def patch_fields(mon):
    xs = np.linspace(-10.0, 10.0, PATCH)
    v = mon(np.array(FREQS), [30.0], xs, xs)[:, 0]  # [f, y, x]
    return np.moveaxis(v, 0, -1)  # [y, x, f]


def mirror_augment(rhos, targets):
    """The inverse-design workload is mirror-symmetric in x and in y (grid, CPML,
    source sheet, design nodes and patch are all centred), so the density flipped
    along x and/or y has the patch intensities flipped the same way: four
    training pairs per solve.  rhos [N, z, y, x], targets [N, y, x, f]."""
    rhos = np.asarray(rhos)
    targets = np.asarray(targets)
    r = [rhos, rhos[..., ::-1], rhos[..., ::-1, :], rhos[..., ::-1, ::-1]]
    t = [targets, targets[:, :, ::-1, :], targets[:, ::-1, :, :], targets[:, ::-1, ::-1, :]]
    return np.ascontiguousarray(np.concatenate(r)), np.ascontiguousarray(np.concatenate(t))


def to_intensity(fields):
    """[..., f] complex patch fields -> [..., f] intensities |E|^2 (the network's targets)."""
    return (np.abs(fields) ** 2).astype(np.float32)


def generate(seeds, steps=3000, precision="fp32", log=None):
    """Solve the forward problem for the blurred random densities of `seeds`;
    returns (densities [N, 21, 101, 101], patch fields [N, 21, 21, 4])."""
    from .workloads import blurred_density
    p, mon = design_problem(steps, precision)
    rhos, fields = [], []
    t0 = time.time()
    for q, s in enumerate(seeds):
        rho = blurred_density(p.shape, seed=int(s), radius=3, sigma=1.5)
        p.objective(rho)
        rhos.append(rho.astype(np.float32))
        fields.append(patch_fields(mon))
        if log and (q + 1) % 32 == 0:
            log("%d designs solved (%.1f s)" % (q + 1, time.time() - t0))
    return np.stack(rhos), np.stack(fields)


class FixedChoice:
    """Stands in for the network in `ascent` (an ablation): always picks the
    candidate step length of index `k` (3 = the longest of 0.5, 1, 2, 4 x alpha)."""

    def __init__(self, k=3):
        self.k = int(k)

    def values(self, trials):
        v = np.zeros(len(trials))
        v[self.k] = 1.0
        return v

    def fine_tune(self, *args, **kwargs):
        pass


def ascent(problem, rho0, iters, surrogate=None, monitor=None, alpha0=0.05, inner=1, tune_steps=20, log=None,
           augment=False):
    """Normalised projected-gradient ascent of F on [0, 1] with full-solve
    acceptance.  Without a surrogate, step lengths are found by backtracking
    with forward solves.  With one, `inner` surrogate-guided steps (step length
    chosen on the network, later directions from its gradient) produce the
    candidate, which one forward solve accepts or rejects (rejection falls back
    to backtracking); after every forward solve the network is fine-tuned on
    the run's latest solves (the patch `monitor` supplies their targets).  The run stops early when
    backtracking finds no ascent step longer than 1e-4.  Returns the design, F,
    the history [(full solves so far, F)] of accepted steps and the total number
    of full solves (forward and adjoint), including those of a final failed
    search."""
    count = [0]
    seen_r, seen_t = [], []

    def forward(c):
        F = problem.objective(c)
        count[0] += 1
        if surrogate is not None and monitor is not None:
            seen_r.append(np.asarray(c, dtype=np.float32))
            seen_t.append(to_intensity(patch_fields(monitor)))
            surrogate.fine_tune(np.stack(seen_r[-16:]), np.stack(seen_t[-16:]), steps=tune_steps, augment=augment)
        return F

    F = forward(rho0)
    rho, alpha = np.asarray(rho0, dtype=np.float64), alpha0
    hist = [(count[0], F)]
    for it in range(iters):
        g = problem.gradient()
        count[0] += 1
        # projected gradient: entries pushing a bound-held density outwards are
        # dropped before normalising, so the step moves the free densities
        g = np.where(((rho >= 1) & (g > 0)) | ((rho <= 0) & (g < 0)), 0.0, g)
        d = g / (np.abs(g).max() + 1e-300)
        cand = None
        picked = None
        if surrogate is not None:
            c, step, first = rho, d, None
            for k in range(inner):
                lens = alpha * np.array([0.5, 1.0, 2.0, 4.0])
                trials = np.stack([np.clip(c + a * step, 0, 1) for a in lens])
                best = int(np.argmax(surrogate.values(trials.astype(np.float32))))
                picked = [0.5, 1.0, 2.0, 4.0][best] if picked is None else picked
                c = trials[best]
                first = lens[best] if first is None else first
                if k + 1 < inner:
                    gs = surrogate.objective(c)[1]
                    step = gs / (np.abs(gs).max() + 1e-300)
            Fc = forward(c)
            if Fc > F:
                # the network may lengthen the step, never shorten it; success
                # grows it as in the plain run
                cand, F, alpha = c, Fc, max(alpha, first) * 1.5
        while cand is None and alpha > 1e-4:
            c = np.clip(rho + alpha * d, 0, 1)
            Fc = forward(c)
            if Fc > F:
                cand, F = c, Fc
                alpha *= 1.5
            else:
                alpha *= 0.5
        if cand is None:
            if log:
                log("iter %d: no ascent step longer than 1e-4, stopping (full solves %d)" % (it, count[0]))
            break
        rho = cand
        hist.append((count[0], F))
        if log:
            log("iter %d: F = %.6e, full solves %d%s" % (it, F, count[0],
                                                       "" if picked is None else ", picked %g x alpha" % picked))
    return rho, F, hist, count[0]


# ---------------------------------------------------------------- experiment
def main(argv=None):
    """Surrogate-assisted design experiment on the inverse-design workload.

        python -m ffip_amd.surrogate data  --out FIELDS.npz [--designs 512]
        python -m ffip_amd.surrogate study --data FIELDS.npz --out RESULT.json [--iters 25]

    `data` solves the forward problem for the blurred random densities of seeds
    1000.. (FP32 storage) and saves the patch fields; `study` trains the network
    on all but the last 64 designs (with their mirror images), reports its
    held-out error, then runs the plain, the surrogate-assisted and, as an
    ablation, the same policy with the longest candidate step always chosen (no
    network), from the same start, and counts full solves."""
    import argparse
    import json
    import sys

    ap = argparse.ArgumentParser(prog="python -m ffip_amd.surrogate")
    ap.add_argument("stage", choices=["data", "study"])
    ap.add_argument("--out", required=True)
    ap.add_argument("--data")
    ap.add_argument("--designs", type=int, default=512)
    ap.add_argument("--steps", type=int, default=3000)
    ap.add_argument("--epochs", type=int, default=100)
    ap.add_argument("--iters", type=int, default=25)
    ap.add_argument("--inner", type=int, default=1)
    ap.add_argument("--tune-steps", type=int, default=20)
    ap.add_argument("--heldout", type=int, default=64)
    ap.add_argument("--no-augment", action="store_true", help="train without the mirror images")
    a = ap.parse_args(argv)

    def log(msg):
        print("[surrogate] " + msg, flush=True)

    from .workloads import blurred_density
    if a.stage == "data":
        seeds = np.arange(1000, 1000 + a.designs)
        rhos, fields = generate(seeds, steps=a.steps, log=log)
        np.savez(a.out, seeds=seeds, fields=fields, steps=a.steps)
        log("saved %s" % a.out)
        return 0
    d = np.load(a.data)
    seeds, fields = d["seeds"], d["fields"]
    shape = (21, 101, 101)
    rhos = np.stack([blurred_density(shape, seed=int(s), radius=3, sigma=1.5) for s in seeds]).astype(np.float32)
    ntr = len(seeds) - a.heldout
    net = Surrogate(seed=0)
    t0 = time.time()
    hist = net.fit(rhos[:ntr], to_intensity(fields[:ntr]), epochs=a.epochs, log=log, augment=not a.no_augment)
    t_train = time.time() - t0
    pred = np.concatenate([net.predict(rhos[q:q + 8]) for q in range(ntr, len(seeds), 8)])
    true = to_intensity(fields[ntr:])
    field_err = float(np.linalg.norm(pred - true) / np.linalg.norm(true))
    fi = FREQS.index(0.1)
    F_true = true[..., fi].sum(axis=(1, 2))
    F_pred = pred[..., fi].sum(axis=(1, 2))
    F_err = float(np.median(np.abs(F_pred - F_true) / F_true))
    corr = float(np.corrcoef(F_pred, F_true)[0, 1])
    log("held-out: intensity rel L2 %.3f, objective median rel err %.3f, corr %.3f (train %.1f s)"
        % (field_err, F_err, corr, t_train))

    rho0 = blurred_density(shape, seed=2000, radius=3, sigma=1.5)
    runs = {}
    for name, sg in (("plain", None), ("surrogate", net), ("longest-step", FixedChoice(3))):
        p, mon = design_problem(a.steps, "fp32")
        t0 = time.time()
        rho, F, h, total = ascent(p, rho0, a.iters, surrogate=sg, monitor=mon, inner=a.inner,
                                  tune_steps=a.tune_steps, log=lambda m: log(name + " " + m),
                                  augment=not a.no_augment)
        runs[name] = {"final_F": F, "history": h, "total_solves": total, "seconds": time.time() - t0}
    target = runs["plain"]["final_F"]
    # the highest objective every run reaches, and the full solves each needs for it
    common = min(r["final_F"] for r in runs.values())
    for name, r in runs.items():
        h = r["history"]
        r["solves_to_plain_final"] = next((s for s, F in h if F >= target), None)
        r["solves_to_common_F"] = next((s for s, F in h if F >= common), None)
        r["accepted_steps"] = len(h) - 1
    res = {"train_mse": hist, "train_seconds": t_train, "heldout_intensity_rel_l2": field_err,
           "heldout_objective_median_rel_err": F_err, "heldout_objective_corr": corr,
           "designs": int(len(seeds)), "train_designs": int(ntr), "iters": a.iters, "inner": a.inner,
           "tune_steps": a.tune_steps, "augment": not a.no_augment, "runs": runs}
    res["common_F"] = common
    for name in ("surrogate", "longest-step"):
        if runs[name]["solves_to_plain_final"]:
            res["solve_ratio_at_plain_final_" + name] = (runs["plain"]["solves_to_plain_final"] /
                                                        runs[name]["solves_to_plain_final"])
        res["solve_ratio_at_common_F_" + name] = (runs["plain"]["solves_to_common_F"] /
                                                 runs[name]["solves_to_common_F"])
    with open(a.out, "w") as f:
        json.dump(res, f, indent=1, default=float)
    log(json.dumps({k: v for k, v in res.items() if k not in ("train_mse", "runs")}, default=float))
    for name, r in runs.items():
        log("%s: final F %.6e, %d accepted steps, %d full solves, %s solves to reach the plain run's final F, "
            "%s to reach %.6e (the highest F every run reaches)"
            % (name, r["final_F"], r["accepted_steps"], r["total_solves"], r["solves_to_plain_final"],
               r["solves_to_common_F"], common))
    return 0


if __name__ == "__main__":
    import sys
    sys.exit(main())
