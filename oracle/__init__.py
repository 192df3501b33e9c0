"""FP64 CPU oracle for the velocity-space hot path (TEST INFRASTRUCTURE ONLY).

This package is a plain numpy/scipy restatement of the reference algorithm
(ut-padas/boltzmann, BESolver/python).  It exists to check the HIP product
path, never to be the product:  only ``tests/``, ``__graft_entry__.smoke()``
and the ``cpu_baseline`` leg of ``bench.py`` may import it.  Every function
cites the reference file:line it restates.

Pinning status (how each part is checked against something other than
itself) is listed per module header; functions that nothing but a
code-following restatement pins are marked "parity unpinned".
"""
