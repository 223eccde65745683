"""FP64 CPU oracle of the reference PG-DG step (TEST INFRASTRUCTURE ONLY).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load
this package, and only as the checker - never as a compute path.
"""
