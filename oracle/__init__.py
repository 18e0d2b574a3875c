"""FP64 CPU oracle of the Swift-Hohenberg step -- TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load
this package, and only as the checker.  The product library never uses it.
"""
