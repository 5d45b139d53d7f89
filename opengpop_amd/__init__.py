"""OpenGPOP-MI355X: orbit propagation (RK8 startup + Gauss-Jackson 8) with JGM3 spherical-harmonic
gravity and pairwise conjunction screening, as hand-written HIP kernels for gfx950 behind a
C ABI (include/ogp.h). This package holds the ctypes binding and the Python mirror of the
reference's entry points.
"""
from ._capi import (DATA_DIR, OBJECT_NONFINITE_INPUT, OBJECT_NONFINITE_RESULT, OBJECT_OK,
                    FORCE_CONST, FORCE_DAMPED, FORCE_EXP, FORCE_GRAVITY, FORCE_NONE,
                    FORCE_POLY7, FORCE_POLY8, FORCE_POLY9, LIB_PATH, QUIRKS_REFERENCE,
                    QUIRK_BACKWARD_TIMES, QUIRK_RK8_ROW7, SOURCE_DRAG, SOURCE_MOON, SOURCE_SUN,
                    OgpError, lib)
from .engine import ClosestApproach, ClosestApproachData, Context, GpuPropagator

__all__ = ["Context", "GpuPropagator", "ClosestApproach", "ClosestApproachData", "OgpError",
           "lib", "LIB_PATH", "DATA_DIR", "FORCE_GRAVITY", "FORCE_NONE", "FORCE_CONST",
           "FORCE_POLY7", "FORCE_POLY8", "FORCE_POLY9", "FORCE_EXP", "FORCE_DAMPED",
           "QUIRKS_REFERENCE", "QUIRK_RK8_ROW7", "QUIRK_BACKWARD_TIMES", "SOURCE_SUN",
           "SOURCE_MOON", "SOURCE_DRAG", "OBJECT_OK", "OBJECT_NONFINITE_INPUT",
           "OBJECT_NONFINITE_RESULT"]
