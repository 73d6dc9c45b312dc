"""CPU FP64 oracle for the batched rigid-body Brownian-dynamics step.

TEST INFRASTRUCTURE ONLY.  Nothing in the product (``pystarc_amd``) imports
this package; only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` use it, and only as the checker (or as
the timed CPU baseline), never as the thing measured or shipped.

Each module restates one part of PySTARC's GPU batch path in plain NumPy,
vectorised over trajectories, and cites the reference file:line it follows
(paths relative to the PySTARC source tree):

* ``forces``     - pystarc/forces/gpu_batch_engine.py:65-657 (grid gather,
                   Yukawa multipole far-field, reverse Born), torque
                   definition from pystarc/forces/engine.py:133-184.
* ``simulator``  - pystarc/simulation/gpu_batch_simulator.py:159-1751
                   (step sequence with the reference's NumPy RNG draw order).
* ``physics``    - Romberg k_b / return probability, Wilson CI, quaternion
                   helpers, Stokes/RPY scalars.
* ``rpy12``      - Zuk et al. 2014 two-body 12x12 RPY mobility + Cholesky
                   (no reference counterpart; pinned by continuity,
                   symmetry, positive definiteness and the reference's
                   scalar relative diffusion).
* ``mlp``        - FP64 forward/backward of the 9-256x4-13 SiLU surrogate.

Pinning: every function is checked in ``tests/test_oracle_*.py`` against
known answers held by the reference's own tests (tests/test_pystarc.py),
closed forms, invariants and brute force.  Functions without such a check
would be marked "parity unpinned" here; at present every function has one.
"""
