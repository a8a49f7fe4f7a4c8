"""Test infrastructure: CPU restatement (and a torch FP64 twin) of the reference march.

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may use this package, and only as the checker or the timed CPU
baseline.  The product path (``angularspectrum_heterogeneous_amd``) never
imports it.
"""
