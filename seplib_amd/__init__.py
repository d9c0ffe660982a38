"""seplib_amd: the seplib per-timestep hot path on AMD Instinct MI355X.

The product is the C-ABI library seplib_amd/libsep.so (headers in
include/), built by ``python -m seplib_amd.build``.  This package adds the
ctypes mirror of that ABI (sep.py), a driver that reads like the
reference's example programs (sim.py) and seeded workload generators
(workloads.py).
"""
