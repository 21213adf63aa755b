"""CPU oracle for the MulSKIPS KMC event engine -- TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
import this package.  The product (mulskips_amd) never does.
"""
from .oracle import Oracle, EVENT_DTYPE, build, numberslist  # noqa: F401
