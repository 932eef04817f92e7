"""FP64 CPU oracle for lime's Lindblad/Redfield RK4 density-matrix path.

TEST INFRASTRUCTURE ONLY.  Nothing in ``lime_amd`` imports this package.  Only
``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py``
may import it, and only as the checker (or, for ``cpu_baseline``, as the timed
restatement of lime's NumPy path) -- never as the thing measured or shipped.

Modules
-------
``lindblad``  RK4 + Lindblad right-hand side (both lime's 2+5K matmul operation
              sequence and the H_eff-folded form), observables, QRT correlations.
``superop``   Liouville-space (row-major vec) superoperator algebra, the exact RK4
              amplification matrix, and the eigen-series ("eseries") solver.
``redfield``  Bloch-Redfield tensor (Liouville form) and its Hilbert-space form.
``spectra``   lime's Fourier-transform convention (fft / ifft / fft2).

Pinning (what fixes each function independently of its own formula) is listed in the
module headers; every function is exercised by ``tests/test_oracle.py``.
"""
from . import lindblad, superop, redfield, spectra  # noqa: F401
