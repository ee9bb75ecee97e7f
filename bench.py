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
objects).  One JSON line is printed with the steps/s; the HBM roofline of
the step (700 field passes per step) and of the dominant kernel, timed inside
a step with HIP events on the launch stream; the WSM6 surrogate against the
plain scheme (steps/s, 60-step error, fallback fraction, the matrix-core
roofline of its layers); and the CPU baseline (the oracle restatement of the
reference on the host's cores).
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
# 256 MiB copy), from profiles/measured_kernels_240x240x41_hbm_bytes.md and
# profiles/scalar_transport_240x240x41_hbm_bytes.md.  Only valid at that size; the species transport
# figure is its later-stage launch (six moisture and two tracer species, advection only).
TRAFFIC_240 = {"advance_w": 213.9e6, "advance_uv": 194.8e6, "advance_mu_t": 131.8e6, "microphysics_driver": 159.1e6,
               "advect_scalar_fresh": 101.0e6, "rk_scalar_tend_pair": 778.2e6}
# the HIP kernel each function launches
KERNEL_OF = {"advance_w": "k_advance_w", "advance_uv": "k_advance_uv", "advance_mu_t": "k_advance_mu_t",
             "microphysics_driver": "k_wsm6", "pgf_buoy_coriolis_curvature": "k_pgf_buoy_cor",
             "advect_scalar_fresh": "k_scalar_adv", "rk_scalar_tend_pair": "k_scalar_adv"}
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


def call_bytes(name, rk_call, g):
    """Algorithmic bytes of one call of a hot-path function at grid g: each
    distinct array it must read or write once, 4 bytes per point (3-D fields
    n3, 2-D fields n2).  None for functions outside the table (and for the
    first-stage scalar transport, which also runs the diffusion)."""
    n3 = g.nz * g.ny * g.nx * 4
    n2 = g.ny * g.nx * 4
    table = {
        "advance_w": 20 * n3 + 6 * n2,
        "advance_uv": 14 * n3 + 6 * n2,
        "advance_mu_t": 13 * n3 + 12 * n2,
        "microphysics_driver": 18 * n3 + 7 * n2,
        # pressure gradient + buoyancy + Coriolis + curvature: 15 fields read, 3 tendencies read and written
        "pgf_buoy_coriolis_curvature": 21 * n3 + 12 * n2,
        # theta advection: t, ru, rv, ww read, t_tend written
        "advect_scalar_fresh": 5 * n3 + n2,
        # 6 moisture + 2 tracer species read, 8 advection tendencies written, ru_m, rv_m, ww_m read,
        # h_tendency and z_tendency zeroed (stages 2 and 3: advection only)
        "rk_scalar_tend_pair": None if rk_call == 0 else 21 * n3 + n2,
    }
    return table.get(name)


def in_step_times(s, ops, cfg, spin_cycles=200_000):
    """One eager RK3 step with every call bracketed by HIP events on the
    launching stream, each call preceded by a GPU spin (torch.cuda._sleep) so
    that the events and the call are enqueued before the GPU reaches them:
    the events then time the call's kernels alone, with the cache contents
    the step itself leaves (no flush, no host gap).  Returns
    {name: [(call index, seconds)]}."""
    import torch
    from pywrf_amd.step import rk3_step
    pending = []
    counts = {}

    class Timed:
        def __getattr__(self, name):
            f = getattr(ops, name)
            if not callable(f):
                return f

            def call(*args):
                torch.cuda._sleep(spin_cycles)
                e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                e0.record()
                r = f(*args)
                e1.record()
                n = counts.get(name, 0)
                counts[name] = n + 1
                pending.append((name, n, e0, e1))
                return r
            return call

    rk3_step(s, Timed(), cfg)
    # the cost of the event pair itself, subtracted from every call
    empty = []
    for _ in range(8):
        torch.cuda._sleep(spin_cycles)
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        e1.record()
        empty.append((e0, e1))
    torch.cuda.synchronize()
    base = sorted(e0.elapsed_time(e1) for e0, e1 in empty)[len(empty) // 2] * 1e-3
    out = {}
    for name, n, e0, e1 in pending:
        out.setdefault(name, []).append((n, max(e0.elapsed_time(e1) * 1e-3 - base, 0.0)))
    return out


def dominant_roofline(times, g):
    """The roofline object for the HIP kernel with the largest in-step time.

    Calls are grouped by the kernel they launch (several functions launch the
    same kernel: the theta advection and the species transport both run
    k_scalar_adv).  A call whose bytes are not in the table (the first-stage
    species transport, which also launches the diffusion kernel) counts, for
    choosing the kernel, as the mean of the same function's counted calls;
    the achieved rate is the counted calls' bytes over their summed time."""
    kernels = {}
    for name, calls in times.items():
        counted = [(n, t) for n, t in calls if call_bytes(name, n, g)]
        if not counted:
            continue
        k = kernels.setdefault(KERNEL_OF.get(name, name), {"t": 0.0, "b": 0.0, "n": 0, "est": 0.0, "fns": []})
        tsum = sum(t for _, t in counted)
        k["t"] += tsum
        k["b"] += sum(call_bytes(name, n, g) for n, _ in counted)
        k["n"] += len(counted)
        k["est"] += tsum / len(counted) * len(calls)
        k["fns"].append(name)
        k.setdefault("traffic", []).extend([TRAFFIC_240.get(name)] * len(counted))
    if not kernels:
        return None
    kernel, k = max(kernels.items(), key=lambda kv: kv[1]["est"])
    achieved = k["b"] / k["t"] / 1e9
    traffic = None
    if (g.nz, g.ny, g.nx) == (41, 240, 240) and None not in k["traffic"]:
        traffic = round(sum(k["traffic"]) / len(k["traffic"]))       # mean over the timed launches
    return {"bound": "hbm", "kernel": kernel, "functions": k["fns"], "achieved": round(achieved, 1),
            "peak": PEAK_HBM_GBS, "unit": "GB/s", "frac": round(achieved / PEAK_HBM_GBS, 4), "traffic": traffic,
            "bytes_per_launch": round(k["b"] / k["n"]), "us_per_launch": round(k["t"] / k["n"] * 1e6, 2),
            "launches_timed": k["n"], "us_per_step_est": round(k["est"] * 1e6, 1),
            "timing": "in-step (eager step, each call spin-separated, HIP events)"}


def _cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(s_cpu, cfg, steps=3):
    """The oracle (the CPU restatement of the reference's step, torch on the
    host) on the same seeded state, with every core this process may run on:
    one warm-up step, then the median of `steps` timed steps."""
    import statistics
    import torch
    import oracle
    from pywrf_amd.step import rk3_step
    threads = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    torch.set_num_threads(threads)
    rk3_step(s_cpu, oracle.ops(), cfg)
    times = []
    for _ in range(steps):
        t0 = time.perf_counter()
        rk3_step(s_cpu, oracle.ops(), cfg)
        times.append(time.perf_counter() - t0)
    dt = statistics.median(times)
    return {"value": round(1.0 / dt, 5), "unit": "RK3 steps/s", "cores": threads, "kind": "port",
            "sample": "median of %d RK3 steps (after 1 warm-up) of the same 240x240x41 seeded state, FP32, oracle "
                      "restatement of the reference (torch CPU, %d threads, %s), %.2f s/step"
                      % (steps, threads, _cpu_model(), dt)}


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


def surrogate_leg(args):
    """The learned WSM6 replacement against the plain scheme: steps/s of the
    whole step with each (HIP graphs, replays interleaved), the microphysics
    call's own time inside a step of each, 60-step qv / theta RMS difference,
    the fraction of columns that fell back, and the matrix-core roofline of
    the network's three layers over all 227^2 columns."""
    import torch
    from pywrf_amd import surrogate as sg
    if not os.path.exists(args.surrogate):
        _log("bench: no trained surrogate at %s, surrogate leg skipped" % args.surrogate)
        return None
    sur = sg.Surrogate.load(args.surrogate)
    _log("bench: surrogate vs WSM6, 60 steps each")
    rec = sg.compare_runs(sur, nsteps=60)
    # the inline network (k_sg_fused: inputs packed on chip, three layers) on the
    # 227^2 tile columns of the seeded state after one step
    import ctypes
    import pywrf_amd
    from pywrf_amd import _lib
    from pywrf_amd.gen import synthetic_state
    from pywrf_amd.step import rk3_step, RunConfig
    # the microphysics call alone inside an eager step of each scheme (HIP
    # events, as in_step_times), after one warm-up step on the same state
    mp = {}
    for name, mops in (("wsm6", pywrf_amd.ops()), ("surrogate", sg.surrogate_ops(sur))):
        s = synthetic_state("config2", device="cuda")
        rk3_step(s, mops, RunConfig())
        t = in_step_times(s, mops, RunConfig())
        mp[name] = sum(x for _, x in t["microphysics_driver"]) * 1e6
    rec["microphysics_us_in_step"] = {k: round(v, 1) for k, v in mp.items()}
    rec["microphysics_speedup"] = mp["wsm6"] / mp["surrogate"]
    s = synthetic_state("config2", device="cuda")
    rk3_step(s, pywrf_amd.ops(), RunConfig())
    g = _lib.grid(s.th_phy.shape, s.grid.phys18())
    gc = g.contents
    M = ((gc.ite - 1) - gc.its) * ((gc.jte - 1) - gc.jts)
    ldy = (M + 127) // 128 * 128
    Y = torch.empty((sg.FOUT + 1, ldy), device="cuda")
    net = sur.struct()
    m = s.moist
    ptrs = [ctypes.c_void_p(t.data_ptr()) for t in (s.th_phy, s.rho, s.pi_phy, s.p_phy, m[1], m[2], m[3], m[4],
                                                     m[5], m[6])]

    def fused():
        _lib.check("pywrf_wsm6_surrogate_outputs", _lib.LIB.pywrf_wsm6_surrogate_outputs(
            g, ctypes.addressof(net), *ptrs, ctypes.c_void_p(Y.data_ptr()), ldy, 1, None, 0,
            ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)))

    X = torch.randn(M, sg.KIN, device="cuda").to(torch.bfloat16)
    ws = sg.Workspace((M + 31) // 32 * 32)
    flops = 2.0 * M * (sg.FIN * sg.HID + sg.HID * sg.HID + sg.HID * sg.FOUT)
    reps = 10
    for name, f in (("mfma_roofline", fused), ("gemm_roofline", lambda: sg.forward(sur.net, X, ws))):
        for _ in range(3):
            f()
        torch.cuda._sleep(200_000)
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            f()
        e1.record()
        e1.synchronize()
        sec = e0.elapsed_time(e1) * 1e-3 / reps
        kernel = ("k_sg_fused (inputs packed on chip, 3 layers)" if name == "mfma_roofline"
                  else "k_mlp_gemm (3 layers, training path; inputs already packed)")
        rec[name] = {"bound": "mfma", "kernel": kernel, "achieved": round(flops / sec / 1e12, 1), "peak": 2500.0,
                     "unit": "TFLOP/s", "frac": round(flops / sec / 1e12 / 2500.0, 4), "flops_per_call": flops,
                     "us_per_call": round(sec * 1e6, 1)}
    return {k: (round(v, 6) if isinstance(v, float) else v) for k, v in rec.items()}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=30)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--config", default="config2")
    ap.add_argument("--eager", action="store_true")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--surrogate", default=os.path.join(ROOT, "pywrf_amd", "data", "wsm6_surrogate.pt"),
                    help="trained WSM6 surrogate to compare with the plain scheme over 60 steps ('' to skip)")
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

    # the dominant kernel, timed inside a step (see in_step_times)
    n3 = g.nz * g.ny * g.nx
    times = in_step_times(s, ops, cfg)
    per_fn = sorted(((sum(t for _, t in c), n, len(c)) for n, c in times.items()), reverse=True)
    for tsum, name, ncalls in per_fn[:8]:
        _log("bench: in-step %-30s %8.1f us/step  %2d calls" % (name, tsum * 1e6, ncalls))
    roofline = dominant_roofline(times, g)
    step_bytes = FIELD_PASSES_PER_STEP * n3 * 4
    out = {
        "metric": "RK3 steps/s @240x240x41 FP32 + HBM GB/s; surrogate vs WSM6 speedup at fixed err",
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
    out["in_step_top"] = {n: round(t * 1e6, 1) for t, n, _ in per_fn[:6]}
    if args.surrogate and (g.nz, g.ny, g.nx) == (41, 240, 240):
        out["surrogate"] = surrogate_leg(args)
    if not args.no_cpu_baseline:
        _log("bench: CPU baseline (oracle, 1 + 3 steps)")
        out["cpu_baseline"] = cpu_baseline(s_cpu, cfg)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
