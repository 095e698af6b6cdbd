"""CPU oracle of the explicit finite-volume step -- TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may import this package,
and only as the checker.  The product never imports it.
"""
