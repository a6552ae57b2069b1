"""lbm -- MI355X-native D2Q9 lattice-Boltzmann engine (gfx950 HIP kernels behind a ctypes C-ABI).

Host-side mirror of the reference's driver surface (lbm_smemaa_main.py, lbm_mine2.py,
lbm_smemaa_endfused_main.py): lattice size, tau_inv / Cs, BC mask, steps per launch k, AA parity.
"""
from . import geometry
from ._lib import BGK, CHANNEL, MRT_LES, PERIODIC, LbmError, LbmParams, load, params

__all__ = ["geometry", "load", "params", "LbmParams", "LbmError", "MRT_LES", "BGK", "CHANNEL", "PERIODIC",
           "Simulation"]


def __getattr__(name):
    # torch is only needed for device-resident simulations; keep `import lbm` light for CPU tools
    if name == "Simulation":
        from .sim import Simulation
        return Simulation
    raise AttributeError(name)
