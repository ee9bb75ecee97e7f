"""One RK3 step captured as a HIP graph.

The step launches several hundred small kernels; replaying them from a graph
removes the per-call host cost (ctypes marshalling, Python driver) from the
critical path.  The driver rebinds many state fields to fresh tensors during
a step (as the reference does); inside the capture every rebound field is
copied back into the tensor it replaced, so each replay reads and writes the
same buffers and replays can follow each other.

Scalars that stay fixed from step to step (dt, the acoustic step, the
coefficients) are baked into the graph.  The one that changes, the boundary
time dtbc = dt * (step - it_bdy * bdy_interval), is formed on the device at
the start of each captured step from a device step counter (DeviceClock),
which is set from the host's counters before every replay; the kernels that
use it (relax_bdytend, relax_bdy_scalar, spec_bdy_final) read it from device
memory, so replays are exact for time-varying lateral boundaries too.
"""
import torch

from .step import rk3_step, DeviceClock


class StepGraph:
    """`steps` consecutive RK3 steps captured as one graph (the rebound fields
    are copied back once per replay, so longer graphs amortise the copies)."""

    def __init__(self, s, cfg, ops=None, steps=1, nan_flag=None):
        """nan_flag: a step.NanFlag whose scans are captured with the steps
        (read it with nan_flag.check after replays)."""
        from . import ops as _ops
        ops = ops or _ops()
        if steps < 1:
            raise ValueError("steps must be >= 1")
        self.s, self.cfg, self.ops, self.steps = s, cfg, ops, steps
        names = [n for n, v in vars(s).items() if isinstance(v, torch.Tensor)]
        orig = {n: getattr(s, n) for n in names}
        step0, it0 = s.step, s.it_bdy
        self.clock = DeviceClock(s)
        torch.cuda.synchronize()
        self.graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(self.graph):
            for _ in range(steps):
                rk3_step(s, ops, cfg, clock=self.clock, nan_flag=nan_flag)
            for n in names:
                t = getattr(s, n)
                if t is not orig[n]:
                    orig[n].copy_(t)
        for n in names:
            setattr(s, n, orig[n])
        s.step, s.it_bdy = step0, it0

    def replay(self, n=1):
        """Advance the state by n replays of the graph, n * steps RK3 steps
        (asynchronous)."""
        s, cfg = self.s, self.cfg
        for _ in range(n):
            self.clock.sync_from(s)
            for _ in range(self.steps):
                s.step += 1
                if s.step % cfg.bdy_interval == 1:
                    s.it_bdy = int(s.step / cfg.bdy_interval)
            self.graph.replay()
