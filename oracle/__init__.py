"""Checkers for the (T) triples path.

TEST INFRASTRUCTURE ONLY.  Nothing under ``oracle/`` is imported, linked or run
by the product path (``libcc_amd/``); only ``tests/``, ``__graft_entry__.smoke()``
and ``bench.py``'s cpu_baseline leg use it, and only as the checker or as the
timed CPU baseline.

* ``oracle.dense``  numpy closed forms (W, E[T], E_ST, v1, dE/dVOE) and the
                    independent textbook (T) formula from raw integrals.
* ``oracle.port``   ctypes wrapper of ``port.c``, the C restatement of the
                    reference CPU path (12/6 DGEMMs per tuple, trant3 steps).
* ``oracle.ref``    runs the reference's own per-tuple Fortran routines,
                    compiled under ``oracle/_ref`` by ``oracle/Makefile``.
"""
