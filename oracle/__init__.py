"""CPU oracle for the hex8 FEM hot path (TEST INFRASTRUCTURE ONLY).

This package restates, in NumPy/SciPy (and in plain C for the timed CPU
baseline), the algorithm of the reference's CPU path so that the HIP
implementation can be checked against it.  It is never imported by the
product package ``fem_taichi_amd``; only ``tests/``, ``__graft_entry__.smoke()``
and the ``cpu_baseline`` leg of ``bench.py`` may use it, and only as the
checker.

Every function cites the reference file:line it follows.  Parity status of
each function (what pins it) is listed in ``oracle/fem_oracle.py``'s header.
"""
