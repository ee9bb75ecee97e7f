"""Index-box helper shared by the oracle modules (test infrastructure only).

The reference writes every statement as a slice assignment over a box
[k0,k1) x [j0,j1) x [i0,i1) of (k, j, i) arrays, reading neighbours through
shifted copies of the same box.  `Box` expresses exactly that: `b(a, dk, dj,
di)` is the shifted 3-D view, `b.m(a, dj, di)` a 2-D (j, i) field broadcast
over k, `b.c(a, dk)` a 1-D level array broadcast over (j, i).
"""
import torch


class Box:
    __slots__ = ("k0", "k1", "j0", "j1", "i0", "i1")

    def __init__(self, k0, k1, j0, j1, i0, i1):
        self.k0, self.k1, self.j0, self.j1, self.i0, self.i1 = k0, k1, j0, j1, i0, i1

    def __call__(self, a, dk=0, dj=0, di=0):
        return a[self.k0 + dk:self.k1 + dk, self.j0 + dj:self.j1 + dj, self.i0 + di:self.i1 + di]

    def m(self, a, dj=0, di=0):
        return a[self.j0 + dj:self.j1 + dj, self.i0 + di:self.i1 + di]

    def c(self, a, dk=0):
        return a[self.k0 + dk:self.k1 + dk].reshape(-1, 1, 1)

    def put(self, dst, value, dk=0, dj=0, di=0):
        self(dst, dk, dj, di).copy_(torch.as_tensor(value, dtype=dst.dtype).expand_as(self(dst, dk, dj, di)))


def zeros_like3(a, nz=None):
    """Fresh zero scratch with the dtype of `a` (FP32 or FP64 runs)."""
    shape = a.shape if nz is None else (nz,) + tuple(a.shape[-2:])
    return torch.zeros(shape, dtype=a.dtype)


def sgn(x):
    """Fortran SIGN(1., x): +1 where x >= 0 else -1 (dtype of x)."""
    one = torch.ones((), dtype=x.dtype)
    return torch.where(x >= 0, one, -one)


def flux5(qm3, qm2, qm1, q0, qp1, qp2, ua):
    """5th-order upwind face value, wrf_dynamics.py:1322-1329 (time_step > 0)."""
    f6 = (37. * (q0 + qm1) - 8. * (qp1 + qm2) + (qp2 + qm3)) / 60.0
    d = (qp2 - qm3) - 5. * (qp1 - qm2) + 10. * (q0 - qm1)
    return f6 - sgn(ua) * d / 60.0


def flux3(qm2, qm1, q0, qp1, ua):
    """3rd-order upwind face value, wrf_dynamics.py:1313-1320 (time_step > 0)."""
    f4 = (7 * (q0 + qm1) - (qp1 + qm2)) / 12.0
    d = (qp1 - qm2) - 3. * (q0 - qm1)
    return f4 + sgn(ua) * d / 12.0
