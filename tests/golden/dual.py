"""Call-by-call comparison of two implementations of the hot-path functions.

`DualOps(a, b)` exposes every function of namespace `a`; each call runs `a`
and `b` on separate copies of the same arguments (aliasing between arguments
preserved), records per-call differences of all returned tensors and all
tensor arguments (which the functions may update in place), then writes `a`'s
results back into the caller's tensors so that the driver proceeds on `a`'s
trajectory.  Used by make_golden.py (a = the reference) and by the parity
tests (a = the oracle, b = the HIP path).
"""
import torch


def _clone_args(args, device=None):
    memo = {}
    out = []
    for x in args:
        if isinstance(x, torch.Tensor):
            if id(x) not in memo:
                memo[id(x)] = x.detach().clone() if device is None else x.detach().to(device, copy=True)
            out.append(memo[id(x)])
        else:
            out.append(x)
    return out


def _as_tuple(r):
    return r if isinstance(r, tuple) else (r,)


def rel_err(x, y):
    """Normwise max|x-y| / max|y| over finite entries (inf/nan must match)."""
    x = x.double().cpu()
    y = y.double().cpu()
    fin = torch.isfinite(y)
    if not torch.equal(torch.isfinite(x), fin):
        return float("inf")
    if not bool(fin.any()):
        return 0.0
    d = (x[fin] - y[fin]).abs().max().item()
    s = y[fin].abs().max().item()
    return d / s if s > 0 else d


def local_err(x, y, before):
    """rel_err over the points either implementation changed (x or y differs
    from `before`): the error of a function that writes only a strip or a
    ring, measured on that strip against its own values.  0.0 when neither
    changed anything."""
    x = x.double().cpu()
    y = y.double().cpu()
    b = before.double().cpu()
    m = (x != b) | (y != b)
    m &= torch.isfinite(y) & torch.isfinite(x)
    if not bool(m.any()):
        return 0.0
    d = (x[m] - y[m]).abs().max().item()
    s = y[m].abs().max().item()
    return d / s if s > 0 else d


class DualOps:
    """local: names of functions whose updated arguments are also compared on
    the points they changed (local_err, recorded as "local_arg<n>")."""

    def __init__(self, a, b, b_device=None, on_call=None, local=()):
        self._a, self._b, self._bdev = a, b, b_device
        self.records = []
        self._on_call = on_call
        self._count = {}
        self._local = set(local)

    def __getattr__(self, name):
        fa = getattr(self._a, name)
        fb = getattr(self._b, name)

        def call(*args):
            n = self._count.get(name, 0)
            self._count[name] = n + 1
            args_a = _clone_args(args)
            args_b = _clone_args(args, self._bdev)
            pre = _clone_args(args) if name in self._local else None
            ra = _as_tuple(fa(*args_a))
            rb = _as_tuple(fb(*args_b))
            errs = []
            for i, (x, y) in enumerate(zip(rb, ra)):
                if isinstance(y, torch.Tensor):
                    errs.append(("ret%d" % i, rel_err(x, y)))
            for i, (x, y) in enumerate(zip(args_b, args_a)):
                if isinstance(y, torch.Tensor):
                    errs.append(("arg%d" % i, rel_err(x, y)))
                    if pre is not None and pre[i].shape == y.shape:
                        errs.append(("local_arg%d" % i, local_err(x, y, pre[i])))
            rec = dict(name=name, call=n, errs=errs)
            self.records.append(rec)
            if self._on_call is not None:
                self._on_call(rec, args_a, ra)
            # write a's results back into the caller's objects
            back = {}
            for orig, ca in zip(args, args_a):
                if isinstance(orig, torch.Tensor) and id(ca) not in back:
                    orig.copy_(ca)
                    back[id(ca)] = orig
            out = tuple(back.get(id(r), r) if isinstance(r, torch.Tensor) else r for r in ra)
            return out if len(out) > 1 else out[0]
        return call


def stats_of(x):
    """[sum, sum|x|, max|x|, #non-finite] of a tensor over its finite entries."""
    x = x.detach().double().cpu()
    fin = torch.isfinite(x)
    xf = x[fin]
    if xf.numel() == 0:
        return [0.0, 0.0, 0.0, float((~fin).sum())]
    return [float(xf.sum()), float(xf.abs().sum()), float(xf.abs().max()), float((~fin).sum())]


class Recorder:
    """Wrap a namespace of hot-path functions and record stats_of() of every
    returned tensor and every tensor argument after each call."""

    def __init__(self, ops):
        self._ops = ops
        self.rows = []
        self.enabled = True

    def __getattr__(self, name):
        f = getattr(self._ops, name)

        def call(*args):
            r = f(*args)
            if self.enabled:
                rr = r if isinstance(r, tuple) else (r,)
                self.rows.append((name, [stats_of(x) for x in rr if isinstance(x, torch.Tensor)],
                                  [stats_of(x) for x in args if isinstance(x, torch.Tensor)]))
            return r
        return call
