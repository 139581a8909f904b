"""CPU oracle for the doubly periodic Stokes spectral Ewald path.

Test infrastructure only: the C restatement of the reference MEX kernels
(``se2p_oracle.c``), a ctypes binding to it (``cref``) and closed-form direct
sums (``direct_sums``).  Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` import this package; the product package
``fasttools2d_amd`` never does.
"""
