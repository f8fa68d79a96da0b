"""CPU oracle: a numpy restatement of FiniteMPTools' hot-path algorithms.

TEST INFRASTRUCTURE ONLY.  This package is the checker for the HIP engine.
Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import it; the engine (``finitemptools_amd``) never does and
has no CPU fallback.

Every function cites the reference file:line it restates.  The reference is
Julia + TensorKit + KrylovKit and cannot run here (no Julia runtime), so the
oracle is pinned by closed forms, brute force and exact diagonalisation in
``tests/test_oracle.py``:

- contractions (ac_prime, ac2_prime, c_prime, updateleft/right): brute-force
  Kronecker-product H_eff, special cases (w = d = 1 reduces to L x R^T),
  Hermiticity, environment/expectation consistency.          -> pinned
- simple_lanczos_solver: k = n recovers scipy.linalg.eigh.   -> pinned
- exponentiate: scipy.linalg.expm on the dense H_eff.         -> pinned
- DMRG2 sweeps + MPSTruncation: ED energies (scipy eigsh), Bethe-ansatz
  trend, truncation semantics on hand-made spectra.           -> pinned
- freefermion.tfim_quench_sz (closed form of the TFIM quench used to check
  the L = 100 TDVP run): exact state-vector evolution at L = 8. -> pinned
"""
