"""CPU oracle of the regular-force path (test infrastructure only).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may use
this package, and only as the checker or the timed CPU baseline; the product
library never links or calls it.
"""
