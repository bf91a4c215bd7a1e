"""CPU oracle for the finitevolX QG hot path (test infrastructure only).

This package is the checker, never the product: only ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py``
may import it.  The shipped path (``finitevolx_amd``) runs the HIP
library and fails loudly when it is missing; it never falls back here.

``oracle.fvx`` is a NumPy float64 restatement of the reference algorithm;
every function cites the reference file:line it follows (paths are
relative to the finitevolX 0.0.40 source tree).
"""

from oracle import fvx  # noqa: F401
