"""Write the reference's own autograd gradients of advance_w and
advect_scalar as fixtures.

Run in the build container only (needs /root/reference):

    python tests/golden/make_adjoint_golden.py

The seeded 40 x 40 x 41 state with clouds (config1c) is advanced into its
first RK3 step by pywrf_amd.step.rk3_step driving the reference functions
(ref_shim.py), and the arguments of the first advance_w call and of the
first advect_scalar call are kept.  The reference's own functions are then
differentiated by torch autograd (FP32, the reference's arithmetic):

* advance_w (wrf_dynamics.py:3641-3796), the loss
  <lT, t_2ave> + <lW, w> + <lP, ph> with seeded lT, lW, lP, with respect to
  w, ph, t_2 and t_2ave;
* advect_scalar (wrf_dynamics.py:1841-2028), the loss <lam, tendency>, with
  respect to the advected field.

adjoint_golden.npz holds the seeds of the sensitivities and the gradients.
The fixtures are data; no reference source is copied.
"""
import os
import sys

import numpy as np
import torch

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, HERE)

from ref_shim import load_reference          # noqa: E402
from pywrf_amd.gen import synthetic_state   # noqa: E402
from pywrf_amd.step import rk3_step, RunConfig  # noqa: E402

SEED_W, SEED_ADV = 7, 2
W_POS = {"w": 0, "t_2ave": 18, "t_2": 19, "ph": 21}     # positions in advance_w's signature


class FirstCalls:
    def __init__(self, ops, names):
        self._ops, self._names, self.args = ops, names, {}

    def __getattr__(self, n):
        f = getattr(self._ops, n)
        if n not in self._names:
            return f

        def rec(*args):
            if n not in self.args:
                self.args[n] = [a.clone() if isinstance(a, torch.Tensor) else a for a in args]
            return f(*args)
        return rec


def lams(shape, seed, n):
    g = torch.Generator().manual_seed(seed)
    return [torch.randn(shape, generator=g) for _ in range(n)]


def main():
    torch.set_num_threads(8)
    s = synthetic_state("config1c")
    g = s.grid
    ref = load_reference(g.ny, g.nx)
    rec = FirstCalls(ref, ("advance_w", "advect_scalar"))
    rk3_step(s, rec, RunConfig())
    out = {}
    # advance_w
    a = list(rec.args["advance_w"])
    x = {n: a[p].clone().requires_grad_() for n, p in W_POS.items()}
    for n, p in W_POS.items():
        a[p] = x[n]
    res = ref.advance_w(*a)
    lT, lW, lP = lams(tuple(s.w.shape), SEED_W, 3)
    loss = (res[0] * lT).sum() + (res[1] * lW).sum() + (res[2] * lP).sum()
    loss.backward()
    for n in W_POS:
        out["advance_w.d" + n] = x[n].grad.numpy()
    # advect_scalar: the gradient with respect to the field (its first two arguments)
    b = list(rec.args["advect_scalar"])
    q = b[0].clone().requires_grad_()
    b[0] = q
    b[1] = q
    b[2] = torch.zeros_like(b[2])
    tend = ref.advect_scalar(*b)
    (lam,) = lams(tuple(q.shape), SEED_ADV, 1)
    (tend * lam).sum().backward()
    out["advect_scalar.dfield"] = q.grad.numpy()
    np.savez_compressed(os.path.join(HERE, "adjoint_golden.npz"), **out)
    print({k: float(np.abs(v).max()) for k, v in out.items()})


if __name__ == "__main__":
    main()
