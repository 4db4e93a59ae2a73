"""CPU oracle of the SA single-layer path -- TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may use
this package, and only as the checker or the timed CPU baseline.  The shipped
library never imports it.
"""
from .pyoracle import Oracle, build  # noqa: F401
