"""biblbm_amd -- MI355X (gfx950) D3Q19 Guo-VOF lattice Boltzmann and
immersed-boundary coupling behind a C-ABI (include/biblbm.h).

`import biblbm_amd` does not load the GPU library; `biblbm_amd.capi`,
`biblbm_amd.lattice` and `biblbm_amd.ibm` do, and raise if it is not built.
"""
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libbiblbm.so")

__all__ = ["HERE", "LIB_PATH"]
