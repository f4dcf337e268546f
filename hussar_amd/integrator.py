"""The integrator and backend API of libhussar, over the gfx950 C ABI.

* ``PathTracer``  — hussar::PathTracer (libhussar/include/hussar/integrators/path.h:133-476):
                    the public knobs, ``configureFrame``, ``run(backend, scene, samples,
                    interruptFlag)`` and ``fetchFrame``.
* ``HipBackend``  — the third backend beside cpu::Backend / gpu::Backend
                    (libhussar/include/hussar/arch/cpu.h:17-63, gpu.h:17-47): constructed from
                    ``(mesh, integrator)``, ``run(scene, budget, interruptFlag)``.
* ``Guide``       — the guiding distribution PathTracer owns (path.h:135-138).

As in the reference, the backend binds the integrator at construction and the integrator's
run drives the backend; the accumulation frame lives on the device, next to the mesh.
"""
from __future__ import annotations

import ctypes as C
from typing import Optional

import numpy as np

from . import _lib
from ._lib import check, lib
from .types import FrameConfig, Knobs, Scene, TriangleMesh, c_device_info, c_knobs, c_run_stats


def knobs_default() -> Knobs:
    k = c_knobs()
    lib().hsr_knobs_default(C.byref(k))
    return Knobs(bool(k.onlyIndirect), k.maxDepth, bool(k.useGeometricalOptics), bool(k.poDiffraction),
                 bool(k.doGuiding), bool(k.clearBeforeIteration), bool(k.correctPhase), bool(k.filterGuiding),
                 bool(k.doFiltering), bool(k.filteringSphere), k.filteringMin, k.filteringMax,
                 k.filteringRadius, k.primaryOrder, bool(k.produceDebugImage))


def device_info(device: int = 0) -> dict:
    d = c_device_info()
    check(lib().hsr_hip_device_info(device, C.byref(d)))
    return {"name": d.name.decode(), "arch": d.arch.decode(), "compute_units": d.compute_units, "xcds": d.xcds,
            "hbm_bytes": d.hbm_bytes, "lds_bytes_per_cu": d.lds_bytes_per_cu, "clock_mhz": d.clock_mhz,
            "lds_bytes_per_block": d.lds_bytes_per_block}


class Guide:
    """GuidingWrapper<BTree<2, ...>> with PathTracer::setup's settings (path.h:229-232)."""

    def __init__(self, backend: "HipBackend", uniformProb: float = 0.1, splitThreshold: float = 0.005,
                 secondMoment: bool = True):
        h = C.c_void_p()
        check(lib().hsr_guide_create(backend._h, uniformProb, splitThreshold, int(secondMoment), C.byref(h)))
        self._h = h
        self.backend = backend

    def reset(self):
        check(lib().hsr_guide_reset(self._h))

    def step(self):
        check(lib().hsr_guide_step(self._h))

    def sampling(self):
        n = C.c_int32()
        check(lib().hsr_guide_sampling_size(self._h, C.byref(n)))
        child = np.empty(n.value, np.int32)
        p0 = np.empty(n.value, np.float32)
        check(lib().hsr_guide_get_sampling(self._h, child.ctypes.data_as(C.POINTER(C.c_int32)),
                                           p0.ctypes.data_as(C.POINTER(C.c_float)), n.value))
        return child, p0

    def set_sampling(self, child, p0):
        child = np.ascontiguousarray(child, np.int32)
        p0 = np.ascontiguousarray(p0, np.float32)
        check(lib().hsr_guide_set_sampling(self._h, child.ctypes.data_as(C.POINTER(C.c_int32)),
                                           p0.ctypes.data_as(C.POINTER(C.c_float)), child.size))

    def training(self):
        n = C.c_int32()
        check(lib().hsr_guide_training_size(self._h, C.byref(n)))
        child = np.empty(n.value, np.int32)
        stat = np.empty(n.value, np.float64)
        check(lib().hsr_guide_get_training(self._h, child.ctypes.data_as(C.POINTER(C.c_int32)),
                                           stat.ctypes.data_as(C.POINTER(C.c_double)), n.value))
        return child, stat

    def close(self):
        if getattr(self, "_h", None):
            lib().hsr_guide_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class PathTracer(Knobs):
    """hussar::PathTracer: knobs (path.h:153-169) plus the frame and the guiding state."""

    def __init__(self, **knobs):
        super().__init__(**knobs)
        self._frame_config: Optional[FrameConfig] = None
        self._backend: Optional[HipBackend] = None
        self.guiding: Optional[Guide] = None
        # PathTracer state read by Backend::run (path.h:138-141, 458-464)
        self.sampleIndexOffset = 0
        self.currentSampleWeight = 1.0
        self.isFinalIteration = False

    # Integrator::configureFrame (integrator.h:60-62)
    def configureFrame(self, config: FrameConfig) -> None:
        self._frame_config = config
        if self._backend is not None:
            self._backend._configure(config)

    def _bind(self, backend: "HipBackend") -> None:
        self._backend = backend
        if self._frame_config is not None:
            backend._configure(self._frame_config)

    def _guide(self, backend: "HipBackend") -> Optional[Guide]:
        if not self.doGuiding:
            return None
        if self.guiding is None or self.guiding.backend is not backend:
            self.guiding = Guide(backend)
        return self.guiding

    # PathTracer::run (path.h:171-224), the doubling guiding schedule on the device side
    def run(self, backend: "HipBackend", scene: Scene, samples: int, interruptFlag=None) -> None:
        if backend is not self._backend:
            raise ValueError("backend was constructed for another integrator")
        flag = _flag_ptr(interruptFlag)
        check(lib().hsr_path_tracer_run(backend._h, C.byref(scene.to_c()), C.byref(self.to_c()), int(samples),
                                        self._guide(backend)._h if self.doGuiding else None, flag))

    # an explicit schedule: `passes` training passes of `train` paths, then `final` paths
    def run_schedule(self, backend: "HipBackend", scene: Scene, passes: int, train: int, final: int,
                     interruptFlag=None) -> None:
        if backend is not self._backend:
            raise ValueError("backend was constructed for another integrator")
        flag = _flag_ptr(interruptFlag)
        check(lib().hsr_path_tracer_run_schedule(backend._h, C.byref(scene.to_c()), C.byref(self.to_c()),
                                                 int(passes), int(train), int(final),
                                                 self._guide(backend)._h if self.doGuiding else None, flag))

    # PathTracer::fetchFrame (path.h:453-455): frame / totalWeight, complex64 [chirp, sample, channel]
    def fetchFrame(self, precision: str = "complex64") -> np.ndarray:
        return self._backend.fetch(precision)[0]

    def clearFrame(self) -> None:
        check(lib().hsr_hip_clear_frame(self._backend._h))


def _flag_ptr(flag):
    if flag is None:
        return None
    if isinstance(flag, C.c_uint8):
        return C.pointer(flag)
    if isinstance(flag, C.c_bool):
        return C.cast(C.pointer(flag), C.POINTER(C.c_uint8))
    raise TypeError("interruptFlag must be a ctypes.c_bool or c_uint8 (set .value = 1 to interrupt)")


class HipBackend:
    """The gfx950 backend: ``HipBackend(mesh, integrator)`` then ``run(scene, budget)``."""

    def __init__(self, mesh: TriangleMesh, integrator: Optional[PathTracer] = None, device: int = 0):
        V = np.ascontiguousarray(mesh.vertexBuffer, np.float32)
        I = np.ascontiguousarray(mesh.indexBuffer, np.int32)
        h = C.c_void_p()
        check(lib().hsr_hip_create(V.ctypes.data_as(C.POINTER(C.c_float)), V.shape[0],
                                   I.ctypes.data_as(C.POINTER(C.c_int32)), I.shape[0], device, C.byref(h)))
        self._h = h
        self.device = device
        self.pipeline = 1  # the library default: wavefront stages
        self.integrator = integrator
        self.frame_config: Optional[FrameConfig] = None
        if integrator is not None:
            integrator._bind(self)

    def set_velocities(self, velocities: Optional[np.ndarray]) -> None:
        """Per-triangle velocities [m/s] (triangle order of the mesh), or None for a static mesh."""
        if velocities is None:
            check(lib().hsr_hip_set_velocities(self._h, None, 0))
            return
        v = np.ascontiguousarray(velocities, np.float32).reshape(-1, 3)
        check(lib().hsr_hip_set_velocities(self._h, v.ctypes.data_as(C.POINTER(C.c_float)), v.shape[0]))

    def _configure(self, config: FrameConfig) -> None:
        check(lib().hsr_hip_configure_frame(self._h, C.byref(config.to_c())))
        self.frame_config = config

    # Backend::run (cpu.h:44-46): the integrator's current offset / weight / final flag
    def run(self, scene: Scene, budget: int, interruptFlag=None) -> None:
        it = self.integrator
        guide = it._guide(self) if it.doGuiding else None
        check(lib().hsr_hip_run(self._h, C.byref(scene.to_c()), C.byref(it.to_c()), int(it.sampleIndexOffset),
                                int(budget), float(it.currentSampleWeight), int(it.isFinalIteration),
                                guide._h if guide else None, _flag_ptr(interruptFlag)))

    # lower-level entry used by tests and the benchmark
    def run_raw(self, scene: Scene, knobs: Knobs, offset: int, budget: int, sample_weight: float = 1.0,
                is_final: bool = False, guide: Optional[Guide] = None, interruptFlag=None) -> None:
        check(lib().hsr_hip_run(self._h, C.byref(scene.to_c()), C.byref(knobs.to_c()), int(offset), int(budget),
                                float(sample_weight), int(is_final), guide._h if guide else None,
                                _flag_ptr(interruptFlag)))

    def clear(self) -> None:
        check(lib().hsr_hip_clear_frame(self._h))

    def fetch(self, precision: str = "complex64"):
        fc = self.frame_config
        n = fc.sampleCount()
        tw = C.c_double()
        if precision == "complex128":
            buf = np.empty(2 * n, np.float64)
            check(lib().hsr_hip_fetch_frame(self._h, None, buf.ctypes.data_as(C.POINTER(C.c_double)), C.byref(tw)))
            z = buf[0::2] + 1j * buf[1::2]
        else:
            buf = np.empty(2 * n, np.float32)
            check(lib().hsr_hip_fetch_frame(self._h, buf.ctypes.data_as(C.POINTER(C.c_float)), None, C.byref(tw)))
            z = buf.view(np.complex64)
        return z.reshape(fc.chirpCount, fc.samplesPerChirp, fc.channelCount), tw.value

    def fetch_fft(self, precision: str = "complex64") -> np.ndarray:
        """Frame::fft of the frame (frame / totalWeight), computed on the device."""
        fc = self.frame_config
        n = fc.sampleCount()
        if precision == "complex128":
            buf = np.empty(2 * n, np.float64)
            check(lib().hsr_hip_fetch_frame_fft(self._h, None, buf.ctypes.data_as(C.POINTER(C.c_double))))
            z = buf[0::2] + 1j * buf[1::2]
        else:
            buf = np.empty(2 * n, np.float32)
            check(lib().hsr_hip_fetch_frame_fft(self._h, buf.ctypes.data_as(C.POINTER(C.c_float)), None))
            z = buf.view(np.complex64)
        return z.reshape(fc.chirpCount, fc.samplesPerChirp, fc.channelCount)

    def peak(self, fourier: bool = False):
        """Frame::argmax and Frame::frequencyEstimation on the device: ((chirp, sample, channel),
        refined position per axis, magnitude)."""
        idx = (C.c_int32 * 3)()
        ref = (C.c_double * 3)()
        mag = C.c_double()
        check(lib().hsr_hip_frame_peak(self._h, int(fourier), idx, ref, C.byref(mag)))
        return tuple(idx), tuple(ref), mag.value

    def set_option(self, option: int, value: int) -> None:
        check(lib().hsr_hip_set_option(self._h, int(option), int(value)))
        if option == _lib.OPT_PIPELINE:
            self.pipeline = int(value)

    def stats(self) -> dict:
        s = c_run_stats()
        check(lib().hsr_hip_last_stats(self._h, C.byref(s)))
        return {f: (list(v) if isinstance(v, C.Array) else v) for f, v in ((f, getattr(s, f)) for f, _ in s._fields_)}

    def hit_record(self, paths: int, maxDepth: int, rxCount: int):
        hits = np.empty((paths, maxDepth), np.int32)
        nee = np.empty((paths, maxDepth, rxCount), np.int8)
        check(lib().hsr_hip_get_hit_record(self._h, hits.ctypes.data_as(C.POINTER(C.c_int32)),
                                           nee.ctypes.data_as(C.POINTER(C.c_int8)), paths, maxDepth, rxCount))
        return hits, nee

    def close(self):
        if getattr(self, "_h", None):
            lib().hsr_hip_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
