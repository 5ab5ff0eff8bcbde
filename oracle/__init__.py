"""Test infrastructure: the CPU oracle (C restatement of the reference path).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may use
this package, and only as the checker or the timed CPU baseline.
"""
