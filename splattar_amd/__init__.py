"""MI355X-native differentiable 2D-Gaussian-surfel training step.

The compute path is libgs_hip.so (hand-written gfx950 HIP kernels behind the
C ABI of include/gs_hip.h); this package is its host-side mirror of the
reference op API (see api.py) plus synthetic scene generators (scenes.py).
Importing the API without the built library raises: there is no CPU path.
"""

from ._lib import GsError, LIB_PATH, header_symbols, lib  # noqa: F401
from .api import (Camera, DeviceBuffer, GpuContext, NormalLoss, Optimizer, PrefixSum,  # noqa: F401
                  RadixSorter, Rasterizer, SsimLoss, TileBinner, TrackConfig, Tracker, TrainConfig, Trainer,
                  mcmc_add_noise,
                  mcmc_sample_pairs, ply_pack_rows, write_3dgs_ply)

__all__ = ["GpuContext", "DeviceBuffer", "PrefixSum", "RadixSorter", "TileBinner", "Rasterizer",
           "SsimLoss", "NormalLoss", "Optimizer", "Trainer", "TrainConfig", "Camera", "GsError", "lib",
           "header_symbols", "LIB_PATH", "write_3dgs_ply", "ply_pack_rows", "mcmc_add_noise", "mcmc_sample_pairs", "Tracker", "TrackConfig"]
