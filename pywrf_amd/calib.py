"""Truncated-backprop calibration: a water-vapour source bias through the
transport adjoint, and a theta-forcing bias through the acoustic loop's.

The model's water vapour is transported by the step's RK3 scalar transport
(rk_scalar_tend's advection and rk_update_scalar, wrf_dynamics.py:3982-4119)
with the mass fluxes of a recorded run, plus an additive source b per RK3
step; b is a (41, ny, nx) FP32 field.  Observations come from a truth run
with a source b_true.  Calibration runs windows of `window` RK3 steps: each
window starts from the model's own state at its first step (the one
checkpoint per RK3 step that truncated backprop keeps: the state, detached),
runs the window through adjoint.TracerStep, and one backward pass through the
hand-written adjoint kernels gives d(loss)/db; Adam updates b after each
window.  The windows of an epoch cover the recorded steps in order.

    python -m pywrf_amd.calib --config config2 --steps 36 --window 6 --epochs 30
"""
import math
import time

import torch

from . import adjoint


def smooth_field(shape, scale, seed, device="cuda"):
    """A smooth random field: a few low-wavenumber sines in each direction."""
    gen = torch.Generator().manual_seed(seed)
    nz, ny, nx = shape
    k = torch.arange(nz, dtype=torch.float64).view(nz, 1, 1) / nz
    j = torch.arange(ny, dtype=torch.float64).view(1, ny, 1) / ny
    i = torch.arange(nx, dtype=torch.float64).view(1, 1, nx) / nx
    out = torch.zeros(shape, dtype=torch.float64)
    for _ in range(4):
        a, p = torch.rand(3, generator=gen) * 3 + 0.5, torch.rand(3, generator=gen) * 6.283
        amp = float(torch.rand(1, generator=gen)) + 0.5
        out += amp * torch.sin(a[0] * 3.1416 * k + p[0]) * torch.sin(a[1] * 6.283 * j + p[1]) * \
            torch.sin(a[2] * 6.283 * i + p[2])
    out *= scale / float(out.abs().max())
    return out.float().to(device)


def interior_mask(g, device="cuda"):
    """1 on the points the transport updates (the mass points of the domain), 0 elsewhere."""
    m = torch.zeros((g.nz, g.ny, g.nx), device=device)
    m[g.kts:g.kde - 1, g.jts:g.jde - 1, g.its:g.ide - 1] = 1.0
    return m


class Calibration:
    """State of a calibration: the recorded transport, the truth trajectory,
    the bias being fitted and its Adam moments."""

    def __init__(self, config="config2", nsteps=36, window=6, bias_scale=2e-6, seed=0, log=None):
        import pywrf_amd
        from .gen import synthetic_state
        from .step import RunConfig
        self.log = log or (lambda *a: None)
        t0 = time.perf_counter()
        s = synthetic_state(config, device="cuda")
        self.q0 = s.moist[1].clone()
        self.grid = s.grid
        self.steps = adjoint.record_transport(s, pywrf_amd.ops(), RunConfig(), nsteps)
        torch.cuda.synchronize()
        self.log("calib: recorded %d steps of %s in %.1f s" % (nsteps, config, time.perf_counter() - t0))
        self.window = window
        self.mask = interior_mask(self.grid)
        self.b_true = smooth_field(self.q0.shape, bias_scale, seed) * self.mask
        with torch.no_grad():
            q = self.q0
            self.truth = []
            for st in self.steps:
                q = adjoint.tracer_step(q, st) + self.b_true
                self.truth.append(q)
        self.b = torch.zeros_like(self.q0)
        self.m = torch.zeros_like(self.q0)
        self.v = torch.zeros_like(self.q0)
        self.t = 0
        self.scale = float(torch.stack([x.square().mean() for x in self.truth]).mean())

    def bias_error(self):
        return float((self.b - self.b_true).norm() / self.b_true.norm())

    def window_loss(self, n0, q_start, b):
        """Mean squared misfit over the window's steps (normalised by the
        truth's mean square), and the model state at the window's end."""
        q = q_start
        loss = 0.0
        for n in range(n0, min(n0 + self.window, len(self.steps))):
            q = adjoint.TracerStep.apply(q, self.steps[n]) + b
            loss = loss + ((q - self.truth[n]) ** 2).mean() / self.scale
        return loss, q

    def epoch(self, lr=1e-7, beta1=0.9, beta2=0.999, eps=1e-20):
        """One pass over the recorded steps, one Adam update per window."""
        q = self.q0
        total = 0.0
        for n0 in range(0, len(self.steps), self.window):
            b = self.b.clone().requires_grad_()
            loss, q_end = self.window_loss(n0, q.detach(), b)
            loss.backward()
            g = b.grad * self.mask
            self.t += 1
            self.m.mul_(beta1).add_(g, alpha=1 - beta1)
            self.v.mul_(beta2).addcmul_(g, g, value=1 - beta2)
            mh = self.m / (1 - beta1 ** self.t)
            vh = self.v / (1 - beta2 ** self.t)
            self.b -= lr * mh / (vh.sqrt() + eps)
            total += float(loss.detach())
            q = q_end
        return total


def calibrate(config="config2", nsteps=36, window=6, epochs=30, lr=2e-7, log=None):
    """Run a calibration; returns its history and timings."""
    log = log or (lambda *a: None)
    torch.cuda.reset_peak_memory_stats()
    cal = Calibration(config, nsteps, window, log=log)
    hist = [{"epoch": 0, "bias_error": cal.bias_error()}]
    # timings: forward transport steps alone, and gradient (forward + backward) steps
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    with torch.no_grad():
        q = cal.q0
        for st in cal.steps:
            q = adjoint.tracer_step(q, st)
    torch.cuda.synchronize()
    fwd = len(cal.steps) / (time.perf_counter() - t0)
    t0 = time.perf_counter()
    for e in range(1, epochs + 1):
        loss = cal.epoch(lr * 0.5 * (1.0 + math.cos(math.pi * (e - 1) / epochs)))     # cosine decay
        hist.append({"epoch": e, "loss": loss, "bias_error": cal.bias_error()})
        log("calib: epoch %d loss %.4e bias error %.4f" % (e, loss, hist[-1]["bias_error"]))
    torch.cuda.synchronize()
    grad = epochs * len(cal.steps) / (time.perf_counter() - t0)
    return {"config": config, "steps": nsteps, "window": window, "epochs": epochs, "history": hist,
            "transport_steps_per_s": fwd, "gradient_steps_per_s": grad,
            "peak_hbm_gb": torch.cuda.max_memory_allocated() / 1e9}


def fit_forcing_bias(config="config4", stage=3, iters=60, lr=None, scale=0.1, seed=3, log=None):
    """A theta-forcing bias fitted through an RK stage's acoustic loop.

    The loop (solver.py:1349-1493: the stage's acoustic substeps, recorded
    from a model run by adjoint.record_acoustic_loop) is run from its
    recorded start with the large-step theta tendency t_tend + b_true: the
    truth.  The model starts from b = 0 and fits b to the truth's theta at
    the loop's end through adjoint.AcousticLoop (one checkpoint; the
    substeps run again in its backward, through the hand-written adjoints of
    advance_uv, advance_mu_t, advance_w and the calls around them), Adam.
    b_true: a smooth field of `scale` times max |t_tend| on the mass points.
    Returns the history of the misfit (relative to the truth's departure from
    the unbiased run) and of the relative error of b."""
    import pywrf_amd
    from .gen import synthetic_state
    from .step import RunConfig
    log = log or (lambda *a: None)
    s = synthetic_state(config, device="cuda")
    x0, subs = adjoint.record_acoustic_loop(s, pywrf_amd.ops(), RunConfig(), stage=stage)
    forcing = adjoint.loop_forcing(subs)
    it = adjoint.FORCING.index("t_tend")
    tt = forcing[it]
    mask = interior_mask(s.grid, device=tt.device)
    b_true = smooth_field(tuple(tt.shape), scale * float(tt.abs().max()), seed, device=tt.device) * mask
    ti = adjoint.ACOUSTIC_STATE.index("t")

    def theta_end(b):
        f = list(forcing)
        f[it] = tt + b
        return adjoint.AcousticLoop.apply(subs, *x0, *f)[ti]

    with torch.no_grad():
        truth, plain = theta_end(b_true), theta_end(torch.zeros_like(tt))
    norm = float(((truth - plain) ** 2).sum())
    b = torch.zeros_like(tt, requires_grad=True)
    lr = lr if lr is not None else 0.05 * float(b_true.abs().max())
    opt = torch.optim.Adam([b], lr=lr)
    sched = torch.optim.lr_scheduler.CosineAnnealingLR(opt, iters)
    hist = []
    t0 = time.perf_counter()
    for n in range(iters + 1):
        opt.zero_grad()
        loss = ((theta_end(b) - truth) ** 2).sum() / norm
        err = float((b.detach() - b_true).norm() / b_true.norm())
        lv = float(loss.detach())
        hist.append({"iter": n, "loss": lv, "bias_error": err})
        if n % 10 == 0 or n == iters:
            log("forcing fit %3d: misfit %.3e  bias error %.3e" % (n, lv, err))
        if n == iters:
            break
        loss.backward()
        opt.step()
        sched.step()
    torch.cuda.synchronize()
    return {"config": config, "stage": stage, "substeps": len(subs), "iters": iters,
            "seconds": time.perf_counter() - t0, "history": hist}


def main(argv=None):
    import argparse
    import json
    import sys
    ap = argparse.ArgumentParser()
    ap.add_argument("--config", default="config2")
    ap.add_argument("--steps", type=int, default=36)
    ap.add_argument("--window", type=int, default=6)
    ap.add_argument("--epochs", type=int, default=30)
    ap.add_argument("--lr", type=float, default=2e-7)
    a = ap.parse_args(argv)
    out = calibrate(a.config, a.steps, a.window, a.epochs, a.lr, log=lambda *x: print(*x, file=sys.stderr, flush=True))
    print(json.dumps(out))


if __name__ == "__main__":
    main()
