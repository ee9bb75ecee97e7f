"""One RK3 step captured as a HIP graph.

The step launches several hundred small kernels; replaying them from a graph
removes the per-call host cost (ctypes marshalling, Python driver) from the
critical path.  The driver rebinds many state fields to fresh tensors during
a step (as the reference does); inside the capture every rebound field is
copied back into the tensor it replaced, so each replay reads and writes the
same buffers and replays can follow each other.

Scalars the driver computes on the host are baked into the graph: dt, the
acoustic step and the boundary time dtbc = dt * (step - it_bdy *
bdy_interval).  dtbc multiplies only the boundary tendencies, so a captured
step is valid for every step when those tendencies are zero (the case for a
constant lateral boundary); StepGraph checks this at capture time and refuses
otherwise -- run such cases eagerly.
"""
import torch

from .step import rk3_step


class StepGraph:
    """`steps` consecutive RK3 steps captured as one graph (the rebound fields
    are copied back once per replay, so longer graphs amortise the copies)."""

    def __init__(self, s, cfg, ops=None, steps=1):
        from . import ops as _ops
        ops = ops or _ops()
        for var, arrs in s.bdy.items():
            for a in arrs[4:]:
                if bool((a != 0).any()):
                    raise ValueError("boundary tendency of %r is not zero: dtbc changes every step, "
                                     "capture is only valid for constant boundaries" % var)
        if steps < 1:
            raise ValueError("steps must be >= 1")
        self.s, self.cfg, self.ops, self.steps = s, cfg, ops, steps
        names = [n for n, v in vars(s).items() if isinstance(v, torch.Tensor)]
        orig = {n: getattr(s, n) for n in names}
        step0, it0 = s.step, s.it_bdy
        torch.cuda.synchronize()
        self.graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(self.graph):
            for _ in range(steps):
                rk3_step(s, ops, cfg)
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
            for _ in range(self.steps):
                s.step += 1
                if s.step % cfg.bdy_interval == 1:
                    s.it_bdy = int(s.step / cfg.bdy_interval)
            self.graph.replay()
