"""octlb_amd -- MI355X-native per-time-step AMR D3Q19 lattice-Boltzmann update.

The product is the C-ABI shared library ``liboctlb_gpu.so`` (HIP, gfx950);
this package holds its sources (``csrc/``), the ctypes binding (``abi``), the
brick-forest and case builders (``forest``, ``cases``) and the solver wrapper
(``solver``) that mirrors the reference's TimeLoop interface.
"""
__all__ = ["abi"]
