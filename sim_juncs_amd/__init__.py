"""MI355X-native FDTD time-stepping engine for sim_juncs junction simulations.

The hot path (Yee E/H updates, Drude-Lorentz polarisation, CPML, sources, monitors, DFT) runs
in hand-written HIP kernels in libsjf.so behind the C ABI of include/sjf.h.
"""
from .engine import Engine, COMPONENTS  # noqa: F401
from . import materials  # noqa: F401
