"""CPU oracle for the N-body hot path -- TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import this package, and only as the checker (or as the timed
CPU baseline).  The product (``nbodyssey_amd``) never imports it.

Parity status of each oracle function (what pins it) is listed in the header
of ``nbody_oracle.c`` and in ``numpy_ref.py``.
"""
