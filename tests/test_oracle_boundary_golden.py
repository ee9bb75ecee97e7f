"""The oracle's lateral-boundary functions against the reference's own calls
(CPU).

tests/golden/bdycalls_bdy24.npz holds, for every call of step 2 of the bdy24
workload (24x24x41, boundary tendencies towards a perturbed state) that reads
the boundary arrays, the arguments the reference received and the array it
returned (written by make_golden.py, which drove the reference functions).
The oracle is run on exactly those arguments and must return the same array
bit for bit.  The fixture is also shown to be able to see the mistakes a
boundary kernel could make: swapped start / end strips, the wrong strip
column, a dropped tendency term -- each planted in the oracle's inputs here
changes the result.
"""
import json
import os

import numpy as np
import pytest
import torch

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
BDY_FUNCS = ("relax_bdytend", "spec_bdytend", "relax_bdy_scalar", "spec_bdy_scalar", "spec_bdy_final")
# position of the first of the eight (or four) boundary arrays in each signature
# (wrf_boundary.py:52, 188, 418, 454, 513) and whether the tendency arrays follow the values
FIRST_BDY = {"relax_bdytend": (2, 8), "spec_bdytend": (1, 8), "relax_bdy_scalar": (5, 8),
             "spec_bdy_scalar": (1, 8), "spec_bdy_final": (5, 8)}
DTBC_POS = {"relax_bdytend": 14, "relax_bdy_scalar": 16, "spec_bdy_final": 16}


def _load():
    z = np.load(os.path.join(GOLD, "bdycalls_bdy24.npz"))
    calls = json.load(open(os.path.join(GOLD, "bdycalls_bdy24.json")))
    out = {}
    for key, scalars in calls.items():
        args = [torch.from_numpy(z["%s.arg%d" % (key, n)].copy()) if a is None else a
                for n, a in enumerate(scalars)]
        out[key] = (args, torch.from_numpy(z["%s.out" % key]))
    return out


CALLS = _load()


def _run(key, args):
    import oracle
    fname = key.split(".")[0]
    args = [a.clone() if isinstance(a, torch.Tensor) else a for a in args]
    return getattr(oracle.ops(), fname)(*args)


def test_fixture_covers_every_boundary_function_with_live_tendencies():
    names = {k.split(".")[0] for k in CALLS}
    assert names == set(BDY_FUNCS)
    assert len(CALLS) == 5 + 5 + 1 + 1 + 5
    for key, (args, _) in CALLS.items():
        fname = key.split(".")[0]
        p0, n = FIRST_BDY[fname]
        tends = args[p0 + 4:p0 + n]
        assert all(float(t.abs().max()) > 0 for t in tends), key
        if fname in DTBC_POS:
            assert args[DTBC_POS[fname]] > 0, key


@pytest.mark.parametrize("key", sorted(CALLS))
def test_oracle_equals_reference_call(key):
    args, ref = CALLS[key]
    mine = _run(key, args)
    assert mine.shape == ref.shape
    assert torch.equal(mine, ref), (key, float((mine - ref).abs().max()))


def _swap(args, a, b):
    args = list(args)
    args[a], args[b] = args[b], args[a]
    return args


def _changed(key, args):
    _, ref = CALLS[key]
    return not torch.equal(_run(key, args), ref)


@pytest.mark.parametrize("key", sorted(CALLS))
def test_fixture_sees_planted_strip_mistakes(key):
    """Swapping the start and end strips (x and y), reading strip column 1
    for column 0, or dropping the tendencies must each change the result."""
    args, _ = CALLS[key]
    fname = key.split(".")[0]
    p0, _n = FIRST_BDY[fname]
    t0 = p0 + 4
    # swapped x strips, swapped y strips of the tendency arrays
    assert _changed(key, _swap(args, t0, t0 + 1)), "xs/xe swap unseen"
    assert _changed(key, _swap(args, t0 + 2, t0 + 3)), "ys/ye swap unseen"
    # strip column 1 read in place of column 0
    shifted = list(args)
    bx = args[t0].clone()
    bx[..., 0] = bx[..., 1]
    shifted[t0] = bx
    assert _changed(key, shifted), "column shift unseen"
    # tendencies dropped
    dropped = list(args)
    for m in range(4):
        dropped[t0 + m] = torch.zeros_like(args[t0 + m])
    assert _changed(key, dropped), "dropped tendency unseen"
    if fname in ("relax_bdytend", "relax_bdy_scalar", "spec_bdy_final") and float(args[p0].abs().max()) > 0:
        # the value strips too (mu's are zero: the initial state has mu = 0)
        assert _changed(key, _swap(args, p0, p0 + 1)), "value xs/xe swap unseen"
        assert _changed(key, _swap(args, p0 + 2, p0 + 3)), "value ys/ye swap unseen"
