"""Unit system of the batch BD path: lengths in A, time in ps, energy in kBT
(pystarc/global_defs/constants.py:31-85)."""
import math

EPS0 = 0.000142                       # vacuum permittivity, e^2/(kBT A)
CONV = 6.022e23 * 1e-30 / 1e-12 / 1e-3  # A^3/ps -> M^-1 s^-1 (gpu_batch_simulator.py:87)
KB_SI = 1.380649e-23                  # J/K
T_DEFAULT = 298.15                    # K
ETA_WATER = 1.002e-3                  # Pa s
ANG_TO_M = 1.0e-10
PS_TO_S = 1.0e-12
MU_REF = 0.243                        # water viscosity in kBT ps/A^3 (gpu_batch_simulator.py:742)
ALPHA_DEFAULT = 1.0 / (4.0 * math.pi)  # Born desolvation alpha, 0.07957747
QB_FACTOR = 1.1
OVERLAP_THRESHOLD = 1.5               # A (gpu_batch_simulator.py:774)
