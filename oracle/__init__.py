"""The oracle: a CPU restatement of the reference hot path (test infrastructure).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
import this package, and only as the checker or the measured CPU baseline.
"""
