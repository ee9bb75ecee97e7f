#!/usr/bin/env python
"""RK3 steps/s of the HIP ARW step on one MI355X at 240 x 240 x 41 FP32.

    python bench.py [--steps K] [--warmup W] [--gpus 1] [--eager] [--no-cpu-baseline]
    python bench.py --drift 60        # RMS drift of 60 steps against the oracle in FP64

Workload: the seeded 240 x 240 x 41 state (3 Gaussian hills, 20 m/s jet at
10 km with N(0, 2 m/s) noise, 6 warm bubbles, 6 moisture species; dt = 54 s,
4 acoustic substeps, 5-wide spec/relax lateral boundaries with constant
boundary values), synthetic data from pywrf_amd.gen.  A "step" is one full
RK3 step (3 stages, 7 acoustic substeps, moisture transport, WSM6, boundary
updates), all on the device; the state is resident in HBM before timing.

The timed steps are replays of HIP graphs of up to 10 steps (pywrf_amd.graph)
unless --eager.  The warmup steps run eagerly (the first loads the code
objects).  One JSON line is printed with the steps/s, the HBM roofline of the
dominant kernel (timed with HIP events on the launch stream) and the CPU
baseline (the oracle restatement of the reference, one step on the host).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_HBM_GBS = 8000.0
COPY_REF_FLOATS = 64 * 1024 * 1024   # 256 MiB copy used to scale the HBM byte counters
# Measured bytes per launch at 240x240x41 (FETCH_SIZE x 2.000 + WRITE_SIZE x 1.000, the scale of a
# 256 MiB copy), from profiles/measured_kernels_240x240x41_hbm_bytes.md.  Only valid at that size.
TRAFFIC_240 = {"advance_w": 213.9e6, "advance_uv": 194.8e6, "advance_mu_t": 131.8e6, "wsm6": 159.1e6,
               "rk_scalar_tend_moist": 588.0e6}
FIELD_PASSES_PER_STEP = 700          # algorithmic 3-D field passes per RK3 step (7 acoustic substeps included)


def _log(*a):
    print(*a, file=sys.stderr, flush=True)


def kernel_candidates(s, cfg):
    """(name, calls per RK3 step, bytes per launch, launch()) for the kernels
    that carry the step.  Bytes are algorithmic: each distinct 3-D array the
    kernel must read or write once, 4 bytes per point (2-D arrays included)."""
    import pywrf_amd
    o = pywrf_amd.ops()
    g = s.grid
    n3 = g.nz * g.ny * g.nx * 4
    n2 = g.ny * g.nx * 4
    I18 = g.ix18
    dts = cfg.dt / cfg.soundsteps
    P18 = g.phys18()

    def adv_w():
        o.advance_w(s.w, s.rw_tend, s.ww, s.w_save, s.u, s.v, s.mu, s.mut, s.muave, s.muts, s.c1h, s.c2h, s.c1f,
                    s.c2f, s.c3h, s.c4h, s.c3f, s.c4f, s.t_2save, s.t, s.t_save, s.ph, s.ph_save, s.phb, s.ph_tend,
                    s.ht, s.c2a, s.cqw, s.alt, s.alb, s.a, s.alpha, s.gamma, cfg.rdx, cfg.rdy, dts, 300.0, 0.1,
                    s.dnw, s.fnm, s.fnp, s.rdnw, s.rdn, s.cf1, s.cf2, s.cf3, s.msftx, s.msfty, False, *I18)

    def adv_uv():
        o.advance_uv(s.u, s.ru_tend, s.v, s.rv_tend, s.p, s.pb, s.ph, s.php, s.alt, s.al, s.mu, s.muu, s.cqu, s.muv,
                     s.cqv, s.mudf, s.c1h, s.c2h, s.c1f, s.c2f, s.c3h, s.c4h, s.c3f, s.c4f, s.msfux, s.msfuy,
                     s.msfvx, s.msfvx_inv, s.msfvy, cfg.rdx, cfg.rdy, dts, s.cf1, s.cf2, s.cf3, s.fnm, s.fnp, 0.01,
                     s.rdnw, 1, True, False, *I18)

    def adv_mu_t():
        o.advance_mu_t(s.ww, s.ww1, s.u, s.u_save, s.v, s.v_save, s.mu, s.mut, s.muave, s.muts, s.muu, s.muv, s.mudf,
                       s.c1h, s.c2h, s.c1f, s.c2f, s.c3h, s.c4h, s.c3f, s.c4f, s.ru_m, s.rv_m, s.ww_m, s.t, s.t_save,
                       s.t_2save, s.t_tend, s.mu_tend, cfg.rdx, cfg.rdy, dts, 0.1, s.dnw, s.fnm, s.fnp, s.rdnw,
                       s.msfux, s.msfuy, s.msfvx, s.msfvx_inv, s.msfvy, s.msftx, s.msfty, 2, *I18)

    def wsm6():
        m = s.th_phy.clone(), [x.clone() for x in s.moist[1:7]]
        o.microphysics_driver(m[0], s.rho, s.pi_phy, s.p_phy, s.ht, s.dz8w, s.p8w, cfg.dt, cfg.dx, cfg.dy, *m[1],
                              s.rainnc, s.rainncv, s.snownc, s.snowncv, s.sr, s.refl_10cm, 1, 1, s.graupelnc,
                              s.graupelncv, 0, 0, 0, 0, 0, 0, *P18)

    mf = (s.msfux, s.msfuy, s.msfvx, s.msfvx_inv, s.msfvy, s.msftx, s.msfty)
    sct = (cfg.rdx, cfg.rdy, s.rdn, s.rdnw, 0.0, 0.0, s.xkhh, 0, 0.12)

    def moist_transport():
        # rk_scalar_tend for the six moisture species on a later RK stage (advection only): one launch
        o.rk_scalar_tend(1, 6, False, 2, 0.5 * cfg.dt, s.ru_m, s.rv_m, s.ww_m, s.muts, s.mub, s.mu_1, s.c1h, s.c2h,
                         s.alt, s.moist_old, s.moist, s.moist_tend, s.advect_tend, s.h_tendency, s.z_tendency,
                         s.RQVFTEN, s.qv_base, True, s.fnm, s.fnp, *mf, *sct, 2, *I18)

    return [
        ("advance_w", 7, 20 * n3 + 6 * n2, adv_w),
        ("advance_uv", 7, 14 * n3 + 6 * n2, adv_uv),
        ("advance_mu_t", 7, 13 * n3 + 12 * n2, adv_mu_t),
        ("wsm6", 1, 18 * n3 + 7 * n2, wsm6),
        # reads 6 species + ru, rv, ww; writes 6 advection tendencies + zeroes h_tendency, z_tendency
        ("rk_scalar_tend_moist", 3, 17 * n3 + 1 * n2, moist_transport),
    ]


def time_kernel(fn, reps):
    import torch
    fn()
    torch.cuda.synchronize()
    start, end = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    start.record()
    for _ in range(reps):
        fn()
    end.record()
    end.synchronize()
    return start.elapsed_time(end) / reps * 1e-3        # seconds per launch (host-side marshalling overlaps)


def cpu_baseline(s_cpu, cfg):
    import torch
    import oracle
    from pywrf_amd.step import rk3_step
    threads = min(16, os.cpu_count() or 1)
    torch.set_num_threads(threads)
    t0 = time.perf_counter()
    rk3_step(s_cpu, oracle.ops(), cfg)
    dt = time.perf_counter() - t0
    return {"value": round(1.0 / dt, 5), "unit": "RK3 steps/s", "cores": threads, "kind": "port",
            "sample": "1 RK3 step of the same 240x240x41 seeded state, FP32, oracle restatement of the reference "
                      "(torch CPU, %d threads), %.1f s" % (threads, dt)}


def drift(config, nsteps, cfg):
    """RMS drift after nsteps: rms(x - x64) / rms(x64) per prognostic field, for
    the HIP run (FP32, device) and for the oracle's FP32 run (the reference's
    arithmetic on the host), both against the oracle's FP64 run of the same
    seeded state."""
    import torch
    import oracle
    import pywrf_amd
    from pywrf_amd.gen import synthetic_state
    from pywrf_amd.state import clone_state
    from pywrf_amd.step import rk3_step
    torch.set_num_threads(min(16, os.cpu_count() or 1))
    names = ("u", "v", "w", "t", "ph", "p", "al", "mu")

    def fields(s):
        out = {n: getattr(s, n).detach().double().cpu() for n in names}
        out["qv"] = s.moist[1].detach().double().cpu()
        out["qc+qr+qi+qs+qg"] = s.moist[2:7].sum(0).detach().double().cpu()
        return out

    s32 = synthetic_state(config)
    gpu = clone_state(s32, device="cuda")
    o = pywrf_amd.ops()
    t0 = time.perf_counter()
    for _ in range(nsteps):
        rk3_step(gpu, o, cfg)
    torch.cuda.synchronize()
    _log("drift: %d GPU steps in %.1f s" % (nsteps, time.perf_counter() - t0))
    hip = fields(gpu)
    ops = oracle.ops()
    ref = {}
    for dt in (torch.float64, torch.float32):
        s = synthetic_state(config, dtype=dt)
        t0 = time.perf_counter()
        for n in range(nsteps):
            rk3_step(s, ops, cfg)
            if (n + 1) % 10 == 0:
                _log("drift: oracle %s step %d (%.1f s)" % (dt, n + 1, time.perf_counter() - t0))
        ref[dt] = fields(s)
    rms = lambda x: float(torch.sqrt(torch.mean(x * x)))
    out = {"drift_steps": nsteps, "config": config, "grid": [s32.grid.nz, s32.grid.ny, s32.grid.nx],
           "metric": "rms(x - x_fp64) / rms(x_fp64) after %d RK3 steps" % nsteps,
           "hip_fp32": {}, "oracle_fp32": {}}
    r64 = ref[torch.float64]
    for n in r64:
        den = rms(r64[n]) or 1.0
        out["hip_fp32"][n] = float("%.3e" % (rms(hip[n] - r64[n]) / den))
        out["oracle_fp32"][n] = float("%.3e" % (rms(ref[torch.float32][n] - r64[n]) / den))
    out["finite"] = all(bool(torch.isfinite(v).all()) for v in hip.values())
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=30)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--config", default="config2")
    ap.add_argument("--eager", action="store_true")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--kernel-reps", type=int, default=20)
    ap.add_argument("--drift", type=int, default=0, metavar="N",
                    help="instead of timing: run N steps on the GPU and the oracle in FP64 and FP32 on the host, "
                         "print the RMS drift of each field from the FP64 run")
    ap.add_argument("--kernels-only", action="store_true",
                    help="run each measured kernel --kernel-reps times and a plain copy of one 3-D field "
                         "(the reference for counter scaling), print nothing else; for rocprofv3 --pmc runs")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: this build runs on one GPU only (the domain is not sharded across GPUs); "
              "run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import torch
    import pywrf_amd
    from pywrf_amd.gen import synthetic_state, CONFIGS
    from pywrf_amd.step import rk3_step, RunConfig
    from pywrf_amd.state import clone_state

    cfg = RunConfig()
    if args.drift > 0:
        drift(args.config, args.drift, cfg)
        return
    _log("bench: generating %s state" % args.config)
    s_cpu = synthetic_state(args.config)
    s = clone_state(s_cpu, device="cuda")
    g = s.grid
    ops = pywrf_amd.ops()
    _log("bench: eager warmup step")
    rk3_step(s, ops, cfg)
    torch.cuda.synchronize()
    if args.kernels_only:
        src = s.u.clone()
        dst = torch.empty_like(src)
        big = torch.ones(COPY_REF_FLOATS, device="cuda")
        big_dst = torch.empty_like(big)
        for name, _, nbytes, fn in kernel_candidates(s, cfg):
            for _ in range(args.kernel_reps):
                fn()
        for _ in range(args.kernel_reps):
            dst.copy_(src)                  # one 3-D field: 9.45 MB read and written at 240^2
        for _ in range(args.kernel_reps):
            big_dst.copy_(big)              # counter scale reference: 256 MiB read and written, beyond every cache
        torch.cuda.synchronize()
        _log("bench: kernels-only run done (copies of %d and %d bytes)" % (src.numel() * 4, big.numel() * 4))
        return
    step = lambda: rk3_step(s, ops, cfg)
    for _ in range(max(args.warmup - 1, 0)):
        rk3_step(s, ops, cfg)
    torch.cuda.synchronize()
    if args.eager:
        runs = [(step, args.steps)]
    else:
        # the timed steps as replays of graphs of up to 10 steps each
        from pywrf_amd.graph import StepGraph
        per = min(10, args.steps)
        full, rest = divmod(args.steps, per)
        _log("bench: capturing a %d-step HIP graph%s" % (per, " (+ a %d-step one)" % rest if rest else ""))
        runs = [(StepGraph(s, cfg, ops, steps=per).replay, full)]
        if rest:
            runs.append((StepGraph(s, cfg, ops, steps=rest).replay, 1))
    torch.cuda.synchronize()
    _log("bench: timing %d steps" % args.steps)
    t0 = time.perf_counter()
    for fn, times in runs:
        for _ in range(times):
            fn()
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    ms = elapsed / args.steps * 1e3
    finite = bool(torch.isfinite(s.p).all()) and bool(torch.isfinite(s.u).all())

    # dominant kernel: largest (launch time x launches per step), timed with HIP events
    n3 = g.nz * g.ny * g.nx
    kern = []
    for name, calls, nbytes, fn in kernel_candidates(s, cfg):
        sec = time_kernel(fn, args.kernel_reps)
        kern.append((sec * calls, name, sec, nbytes))
        _log("bench: %-14s %8.1f us/launch  %7.1f GB/s  x%d per step" % (name, sec * 1e6, nbytes / sec / 1e9, calls))
    kern.sort(reverse=True)
    _, kname, ksec, kbytes = kern[0]
    achieved = kbytes / ksec / 1e9
    traffic = TRAFFIC_240.get(kname) if (g.nz, g.ny, g.nx) == (41, 240, 240) else None
    roofline = {"bound": "hbm", "kernel": kname, "achieved": round(achieved, 1), "peak": PEAK_HBM_GBS,
                "unit": "GB/s", "frac": round(achieved / PEAK_HBM_GBS, 4),
                "traffic": None if traffic is None else round(traffic),
                "bytes_per_launch": kbytes, "us_per_launch": round(ksec * 1e6, 2)}
    step_bytes = FIELD_PASSES_PER_STEP * n3 * 4
    out = {
        "metric": "RK3 steps/s @240x240x41 FP32",
        "value": round(1e3 / ms, 3),
        "unit": "RK3 steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "%s: %dx%dx%d seeded state, dt=54 s, 4 acoustic substeps, WSM6, %s" % (
            args.config, g.nx, g.ny, g.nz, "eager" if args.eager else "HIP graph"),
            "grid": [g.nz, g.ny, g.nx]},
        "hbm_gbs_step": round(step_bytes / (ms * 1e-3) / 1e9, 1),
        "step_roofline_frac": round(step_bytes / (ms * 1e-3) / 1e9 / PEAK_HBM_GBS, 4),
        "finite": finite,
        "roofline": roofline,
    }
    if not args.no_cpu_baseline:
        _log("bench: CPU baseline (oracle, one step)")
        out["cpu_baseline"] = cpu_baseline(s_cpu, cfg)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
