"""edxraster_amd: EDXRaster's raster-pipeline hot path on MI355X (gfx950).

The pipeline (vertex transform, homogeneous clipping, triangle setup,
order-preserving 32x32 tile binning, hierarchical edge-function
rasterisation, perspective-correct interpolation, Blinn-Phong shading and the
depth test) runs as HIP kernels in libedxraster.so behind a C ABI
(include/edxraster.h).  This package is the Python mirror of the reference's
Renderer / Mesh API over that ABI.
"""
from . import _lib
from .renderer import Mesh, Renderer, look_at, perspective, raster_matrix, default_material, load_bmp, load_image
from ._lib import (SHADER_LAMBERT_ALBEDO, SHADER_BLINN_PHONG_REF, SHADER_VC_BLINN_PHONG, SHADER_TEXTURED_LAMBERT,
                   FILTER_NEAREST, FILTER_LINEAR, FILTER_TRILINEAR, EdxrError)

__all__ = ["Mesh", "Renderer", "look_at", "perspective", "raster_matrix", "default_material",
           "SHADER_LAMBERT_ALBEDO", "SHADER_BLINN_PHONG_REF", "SHADER_VC_BLINN_PHONG", "EdxrError"]
