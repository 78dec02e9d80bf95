"""LBM collision and streaming (reference physics/lbm.py)."""
from .. import kernels

collide_mrt = kernels.collide_mrt
collide_srt = kernels.collide_srt
collide_srt_variable_tau = kernels.collide_srt_variable_tau
stream = kernels.stream

__all__ = ["collide_mrt", "collide_srt", "collide_srt_variable_tau", "stream"]
