"""ttc.sql_util: the tuning database queries of ttc/sql_util.py in the
reference, provided by ttc_amd.sql_util."""
from ttc_amd.sql_util import *  # noqa: F401,F403
