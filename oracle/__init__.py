"""CPU FP64 oracle for DCell's level-set update (test infrastructure only).

Only tests/, ``__graft_entry__.smoke()`` and bench.py's ``cpu_baseline`` leg
import this package.  It is the checker, never the product path.
"""
from .oracle import *  # noqa: F401,F403
