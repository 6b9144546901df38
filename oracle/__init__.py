"""Test infrastructure: ctypes binding of the CPU oracle (h4_oracle.c).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this package.  The product path never does.
"""
from .h4o import Oracle, load, D_TIME_MIN, D_TIME_MAX, ETA_DEFAULT, softening_e2  # noqa: F401
