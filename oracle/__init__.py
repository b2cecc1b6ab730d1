"""CPU oracle for the SIMP hot path (test infrastructure only).

This package is the checker, never the product: only ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` may
import it.  The product path (``pytopo3d_amd``) never imports anything from here
and fails loudly when its HIP library is missing.
"""

from .simp import *  # noqa: F401,F403
