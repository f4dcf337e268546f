"""Python binding of the serving runtime (hsr_server_* of include/hussar_hip.h).

    srv = Server(slots=4)
    sid = srv.open_scene(mesh)
    t = srv.submit(sid, scene, knobs, FrameConfig(1, 256, 4), paths=1 << 20)
    frame, info = srv.wait(t)      # complex64 [chirp, sample, channel], per-request timings

Scenes stay resident on the device, with a guiding tree once trained; each request runs on a
cleared frame in one of the server's slots (own stream, buffers and guide), admitted when its
own device footprint fits. There is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass
from typing import Optional

import numpy as np

from ._lib import check, lib
from .types import (FrameConfig, Knobs, Scene, TriangleMesh, c_request, c_request_info, c_server_config,
                    c_server_stats)


GUIDE_PER_REQUEST, GUIDE_SCENE, GUIDE_TRAIN = 0, 1, 2  # hsr_request.guideMode


def request_mix(count: int, rate: float, seed: int = 3) -> list:
    """The serving mix: Poisson arrivals at `rate`/s; per request one of three scenes, a radar
    pose within 20 m of the origin looking down +z (height 0.75 m), bins U{128..512},
    receivers U{1..12} at λ/2, paths 2^U{18..24}; every draw from SplitMix64 streams of
    `seed`. Returns dicts with t (arrival [s]), scene, pos, bins, channels, paths."""
    from .scenes import uniforms
    u = uniforms(seed, 8 * count).reshape(count, 8)
    t = np.cumsum(-np.log1p(-u[:, 0]) / rate)
    out = []
    for i in range(count):
        r = 20.0 * np.sqrt(u[i, 1])
        phi = 2 * np.pi * u[i, 2]
        out.append({"t": float(t[i]), "scene": int(u[i, 3] * 3), "pos": (r * np.cos(phi), 0.75, r * np.sin(phi)),
                    "bins": 128 + int(u[i, 4] * 385), "channels": 1 + int(u[i, 5] * 12),
                    "paths": 1 << (18 + int(u[i, 6] * 7))})
    return out


@dataclass
class RequestInfo:
    status: int
    slot: int
    queue_ms: float
    run_ms: float
    latency_ms: float
    totalWeight: float


class Server:
    def __init__(self, device: int = 0, slots: int = 4, queueCapacity: int = 0, maxDepth: int = 0,
                 hbmBudget: int = 0, pathsPerBatch: int = 0, frameReserveBytes: int = 0, graphs: bool = False):
        cfg = c_server_config(device, slots, queueCapacity, maxDepth, hbmBudget, pathsPerBatch, frameReserveBytes,
                              1 if graphs else 0, 0)
        h = C.c_void_p()
        check(lib().hsr_server_create(C.byref(cfg), C.byref(h)))
        self._h = h
        self._shapes = {}

    def open_scene(self, mesh: TriangleMesh) -> int:
        V = np.ascontiguousarray(mesh.vertexBuffer, np.float32)
        I = np.ascontiguousarray(mesh.indexBuffer, np.int32)
        sid = C.c_int32()
        check(lib().hsr_server_open_scene(self._h, V.ctypes.data_as(C.POINTER(C.c_float)), V.shape[0],
                                          I.ctypes.data_as(C.POINTER(C.c_int32)), I.shape[0], C.byref(sid)))
        return sid.value

    def set_velocities(self, scene_id: int, velocities: Optional[np.ndarray]) -> None:
        if velocities is None:
            check(lib().hsr_server_set_velocities(self._h, scene_id, None, 0))
            return
        v = np.ascontiguousarray(velocities, np.float32).reshape(-1, 3)
        check(lib().hsr_server_set_velocities(self._h, scene_id, v.ctypes.data_as(C.POINTER(C.c_float)), v.shape[0]))

    def close_scene(self, scene_id: int) -> None:
        check(lib().hsr_server_close_scene(self._h, scene_id))

    @staticmethod
    def _request(scene_id, scene, knobs, frame, paths, guide_mode) -> c_request:
        r = c_request()
        r.sceneId = scene_id
        r.guideMode = int(guide_mode)
        r.paths = int(paths)
        r.scene = scene.to_c()
        r.knobs = knobs.to_c()
        r.frame = frame.to_c()
        return r

    def request_bytes(self, scene_id: int, scene: Scene, knobs: Knobs, frame: FrameConfig, paths: int,
                      guide_mode: int = 0) -> int:
        """The request's admission footprint: device bytes it needs on an empty slot."""
        n = C.c_int64()
        r = self._request(scene_id, scene, knobs, frame, paths, guide_mode)
        check(lib().hsr_server_request_bytes(self._h, C.byref(r), C.byref(n)))
        return n.value

    def submit(self, scene_id: int, scene: Scene, knobs: Knobs, frame: FrameConfig, paths: int,
               guide_mode: int = 0) -> int:
        """guide_mode: GUIDE_PER_REQUEST (the reference's schedule per request), GUIDE_SCENE (one
        final pass with the scene's resident tree; trains it first if the scene has none) or
        GUIDE_TRAIN (the reference's schedule, whose tree becomes the scene's)."""
        r = self._request(scene_id, scene, knobs, frame, paths, guide_mode)
        t = C.c_int64()
        check(lib().hsr_server_submit(self._h, C.byref(r), C.byref(t)))
        self._shapes[t.value] = (max(1, frame.chirpCount), frame.samplesPerChirp, frame.channelCount)
        return t.value

    def poll(self, ticket: int) -> bool:
        d = C.c_int32()
        check(lib().hsr_server_poll(self._h, ticket, C.byref(d)))
        return bool(d.value)

    def wait(self, ticket: int, fetch: bool = True):
        shape = self._shapes.pop(ticket, None)
        info = c_request_info()
        out = None
        if fetch and shape is not None:
            out = np.empty(int(np.prod(shape)), np.complex64)
            ptr = out.ctypes.data_as(C.POINTER(C.c_float))
            check(lib().hsr_server_wait(self._h, ticket, ptr, out.size, C.byref(info)))
            out = out.reshape(shape)
        else:
            check(lib().hsr_server_wait(self._h, ticket, None, 0, C.byref(info)))
        return out, RequestInfo(info.status, info.slot, info.queue_ms, info.run_ms, info.latency_ms, info.totalWeight)

    def stats(self) -> dict:
        s = c_server_stats()
        check(lib().hsr_server_get_stats(self._h, C.byref(s)))
        return {f: getattr(s, f) for f, _ in s._fields_}

    def reset_stats(self) -> None:
        check(lib().hsr_server_reset_stats(self._h))

    def close(self) -> None:
        if getattr(self, "_h", None):
            lib().hsr_server_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
