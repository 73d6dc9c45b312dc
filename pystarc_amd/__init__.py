"""MI355X-native batched rigid-body Brownian-dynamics step with the call
surface of PySTARC's GPU batch path (GPUBatchForceEngine, GPUBatchSimulator).

The compute path is libpystarc_hip.so (hand-written gfx950 HIP kernels behind
a C ABI, include/pystarc_hip.h); this package is the host-side mirror of the
reference interface plus the synthetic benchmark systems.
"""
__version__ = "0.1.0"
