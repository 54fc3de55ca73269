"""CPU oracle: a NumPy/SciPy restatement of rextlfung/mslr-recon's reconstruction path.

TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg may import this package, and only as the checker or the timed
CPU baseline -- never as part of the product path (mslr-recon_amd/), which has no
CPU fallback.

Each function cites the reference file:line it restates.  Pins (what checks each
function against something other than itself) are listed in mslr_ref.PINS and
exercised by tests/test_oracle.py.
"""
from .mslr_ref import *  # noqa: F401,F403
from .mslr_ref import _gr_restart, _patch_starts, _workers  # noqa: F401,E402
