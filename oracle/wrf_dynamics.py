"""Oracle restatement of pywrf/wrf_dynamics.py (test infrastructure only).

Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg import
this package; the product (pywrf_amd) never does.
"""
from .dyn_prep import *    # noqa: F401,F403
from .dyn_adv import *     # noqa: F401,F403
from .dyn_tend import *    # noqa: F401,F403
from .dyn_small import *   # noqa: F401,F403
from .dyn_scalar import *  # noqa: F401,F403
