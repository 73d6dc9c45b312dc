from .rotne_prager import (MobilityTensor, rpy_offdiagonal, stokes_rotational_diffusion,  # noqa: F401
                           stokes_translational_diffusion)
