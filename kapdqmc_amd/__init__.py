"""MI355X-native projector-DQMC inner loop (t-U model) behind a C-ABI HIP library.

Host-side mirrors of the reference's model constructors live in ``models``; the GPU entry
points (``DQMCBatch``, ``phy_update``, ``Initial_s``) live in ``dqmc``.
"""
from .models import tU_Hubbard_Para, tV_Hubbard_Para, tUV_Hubbard_Para  # noqa: F401
from .geometry import area_index, nnK_Matrix, xy_i, i_xy, nn2idx  # noqa: F401


def __getattr__(name):
    if name in ("DQMCBatch", "phy_update", "Initial_s"):
        from . import dqmc
        return getattr(dqmc, name)
    raise AttributeError(name)
