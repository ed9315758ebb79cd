"""ttc.ttc_util: TTCargs, fuseIndices and the helpers of ttc/ttc_util.py in
the reference, provided by ttc_amd.ttc_util."""
from ttc_amd.ttc_util import *  # noqa: F401,F403
