"""`ttc`: the import name of HPAC/TTC, for drop-in use.

Code written against the reference runs unchanged: `import ttc` then
`ttc.ttc.generateTransposition(args)` and `from ttc import ttc_util`
(test-framework/ttc_test.py:27-28, 152), and the `ttc` command
(scripts/ttc:2, `from ttc.ttc import main`).  The modules are those of
ttc_amd, the MI355X (gfx950) implementation.
"""
from . import sql_util, ttc, ttc_util  # noqa: F401
from .ttc import main  # noqa: F401

__all__ = ["main"]
