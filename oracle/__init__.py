"""CPU oracle for the IGA Navier-Stokes hot path (test infrastructure only).

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s cpu_baseline leg
may import this package, and only as the checker.  The product path
(``tubularflowiga_amd``) never imports it.
"""
