"""CPU oracle for the implicit Vlasov-Poisson timestep -- TEST INFRASTRUCTURE ONLY.

This package restates, in NumPy/SciPy (``vlasov_oracle``) and in plain C with
OpenMP (``c/vv1d_oracle.c``), the algorithm of the viVlasov1D split-mode,
matrix-free, nonlinear implicit-midpoint timestep.  Every function cites the
reference file:line it follows.

It is the checker, never the product:

* only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
  ``bench.py`` may import, call, link or run anything in this package;
* the product path (``vivlasov1d_amd``) never imports it and fails loudly when
  its HIP library is missing -- there is no CPU fallback.

Parity pins (how each oracle function is checked against something other than
itself) are listed in the header of ``vlasov_oracle``.
"""
