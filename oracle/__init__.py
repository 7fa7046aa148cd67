"""CPU oracle for jrystal's direct-minimisation total-energy path.

TEST INFRASTRUCTURE ONLY.  This package is a NumPy/SciPy float64 restatement
of the reference algorithm (sail-sg/jrystal, JAX).  It exists to check the
HIP implementation in ``jrystal_amd`` and to provide the CPU baseline leg of
``bench.py``.  Only ``tests/``, ``__graft_entry__.smoke()`` and
``bench.py``'s ``cpu_baseline`` leg may import it, and only as the checker or
the baseline -- never as the thing measured or shipped.  ``jrystal_amd`` never
imports it.

Modules (each function cites the reference file:line it follows):

* ``oracle.crystal``    -- Crystal container, extxyz reader, unit constants
* ``oracle.grid``       -- G/R/k grids, masks, FFT-size table
* ``oracle.pw``         -- QR orthonormalisation, scatter/gather, wave/density
* ``oracle.energy``     -- kinetic / Hartree / external / LDA-x / Ewald / total
* ``oracle.adjoint``    -- closed-form gradient dE/d(w_re, w_im)
* ``oracle.occupation`` -- fixed (uniform, gamma) and parameterised
                            (simplex projector, idempotent) occupations,
                            Fermi-Dirac entropy and their gradients
* ``oracle.workloads``  -- the seeded synthetic inputs used by tests and bench
* ``oracle.optim``      -- optimiser restatements (optax-style Adam,
                            convergence window, Teter-preconditioned descent)

Pinning: every function is checked in ``tests/test_oracle_*.py`` against the
reference's own golden values (``crystal_test.py``), closed forms, brute-force
sums on tiny boxes, a second autodiff system (torch CPU autograd) and central
finite differences.  Functions pinned only by a closed form that the
reference's third-party dependency would define are marked "parity unpinned"
in their docstrings (``lda_x`` low-density threshold, optax step rounding).
"""
