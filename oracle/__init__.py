"""CPU oracle for the iterative-reconstruction hot path (test infrastructure).

Two libraries live here:

* ``liboracle.so`` -- the C restatement of the reference algorithm
  (``trace_oracle.c``), built by ``make -C oracle oracle``;
* ``_ref/libtrace_ref.so`` -- the reference's own hot-path sources compiled
  by path (``make -C oracle ref``; only where the reference checkout exists),
  driven by the harness in ``oracle/ref/``.

Only tests, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may use this package.  It is the checker, never the product.
"""
from .pyoracle import (  # noqa: F401
    ALG_NAMES, APMLR, MLEM, PML, SIRT, Oracle, Params, RefLib, load_oracle,
    load_ref, ref_available)
