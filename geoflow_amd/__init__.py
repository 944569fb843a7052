"""geoflow_amd -- MI355X-native GeoFLOW spectral-element operator path.

The compute lives in ``libgeoflow_hip.so`` (hand-written HIP for gfx950 behind
the C ABI declared in ``include/geoflow_hip.h``).  This module binds that C
ABI with ctypes and offers a thin host-side mirror of the reference operator
interface (GGridBox / GGFX / GMass / GHelmholtz / GCG / GExRKStepper) for
Python callers and tests.  Device memory and streams come from PyTorch; there
is no CPU fallback: if the HIP library is missing every entry point raises.
"""
import ctypes
import os

import numpy as np

from . import _capi
from ._capi import (GF_OK, GF_BDY_DIRICHLET, GF_BDY_PERIODIC, GF_BDY_NONE, GF_DSS_SUM,
                    GF_DSS_SMOOTH, GF_DSS_MAX, GF_DSS_MIN, GF_NORM_INF, GF_NORM_EUC, GF_NORM_L2,
                    GF_NORM_L1, GF_NORM_ASSEMBLED_L2, GeoflowError, lib, lib_path)
from .grid import (gll_basis, box_grid, icos_grid, terrain_box_grid, deformed_geometry2d,
                   ggfx_build, BoxGrid, IcosGrid, TerrainBoxGrid, TerrainBoxGrid3D,
                   DeformedGrid2D)

__all__ = ["lib", "lib_path", "gll_basis", "box_grid", "icos_grid", "terrain_box_grid",
           "deformed_geometry2d", "ggfx_build", "BoxGrid", "IcosGrid", "TerrainBoxGrid",
           "TerrainBoxGrid3D", "DeformedGrid2D", "GeoflowError", "Context", "Mesh"]


def __getattr__(name):
    # device-side classes import torch lazily so host-only use stays light
    if name in ("Context", "Mesh"):
        from . import device
        return getattr(device, name)
    raise AttributeError(name)
