"""CPU oracle for the DDSCAT Green-tensor MATVEC / Krylov / far-field path.

TEST INFRASTRUCTURE ONLY.  Nothing in the product library or its Python
bindings imports this package; only ``tests/``, ``__graft_entry__.smoke()``
and the ``cpu_baseline`` leg of ``bench.py`` use it, and only as the checker
(or, for the baseline, as the timed CPU restatement).
"""
