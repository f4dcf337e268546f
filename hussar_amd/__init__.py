"""hussar_amd — an MI355X (gfx950) backend for libhussar's Monte Carlo FMCW radar integrator.

The host API mirrors libhussar's (TriangleMesh, Scene, NFAntenna, RFConfig, FrameConfig,
PathTracer and a Backend constructed from (mesh, integrator)); the per-path work runs as HIP
kernels behind the C ABI of include/hussar_hip.h (hussar_amd/libhussar_hip.so).
"""
from .types import (FrameConfig, Knobs, NFAntenna, RFConfig, Scene, TriangleMesh, SPEED_OF_LIGHT,
                    GHz, MHz, kHz, Hz, ms, us, ns, mm, cm, rotation_y)
from .integrator import Guide, HipBackend, PathTracer, device_info, knobs_default
from . import io, scenes

__all__ = ["FrameConfig", "Knobs", "NFAntenna", "RFConfig", "Scene", "TriangleMesh", "SPEED_OF_LIGHT",
           "Guide", "HipBackend", "PathTracer", "device_info", "knobs_default", "io", "scenes", "rotation_y",
           "GHz", "MHz", "kHz", "Hz", "ms", "us", "ns", "mm", "cm"]
