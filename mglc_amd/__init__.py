"""mglc_amd -- MI355X-native D3Q19 lattice Boltzmann step (hand-written HIP
kernels behind the C-ABI of include/mglc_lbm.h)."""
from .lbm import Batch, Cavity, CavityConfig, LbmError, ThermalConfig, load_library  # noqa: F401
