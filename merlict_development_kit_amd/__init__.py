"""MI355X-native photon propagator for merlict sceneries.

The hot path -- merlict's per-photon propagation through the scenery frame
tree -- runs in hand-written HIP kernels for CDNA4 (``csrc/propagate.hip``)
behind the C ABI of ``include/mlgpu.h``.  This package holds the host side:
scene builders and flattening (C++, in ``libmlgpu.so``) and thin Python
bindings for tests and benchmarks.
"""
from .scenery import SceneDescription
from .propagation import (GpuScene, PhotonBatch, PropagationConfig, PropagationResult,
                          propagate_photons_in_frame_with_config)
from . import sources
from .arrivals import ArrivalBinning, histogram_arrivals
from .bunches import Bunch, propagate_bunches
from . import corsika

__all__ = [
    "SceneDescription",
    "GpuScene",
    "PhotonBatch",
    "PropagationConfig",
    "PropagationResult",
    "propagate_photons_in_frame_with_config",
    "sources",
    "ArrivalBinning",
    "histogram_arrivals",
    "Bunch",
    "propagate_bunches",
    "corsika",
]
