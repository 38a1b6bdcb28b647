"""CPU oracle for the maxLq RCG hot path.

TEST INFRASTRUCTURE ONLY.  Nothing in ``maxlq_amd`` imports this package;
only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may use it, and only as the checker / CPU baseline.

``lq_oracle``  -- numpy FP64 restatement of the reference field operators
                  (function_ops.f90), FFT conventions (fftwfunction.f90) and
                  optimizer (optimization.f90), following the reference's
                  operation order (complex-to-complex transforms, the same
                  dealias chain, the same scalar line-search arithmetic).
``seeded``     -- the project-defined seeded divergence-free field generator.

Parity status of every oracle function is listed in ``lq_oracle.PINS``; the
tests under ``tests/test_oracle_*.py`` hold the checks that pin them.
"""
