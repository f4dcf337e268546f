"""Host-side mirrors of libhussar's scene, antenna, chirp and frame types.

The names, fields and units follow the reference so that code written against libhussar
reads the same here:

* ``RFConfig``      — radar::RFConfig   (libradar/include/radar/radar.h:113-136)
* ``FrameConfig``   — radar::FrameConfig (radar.h:141-164)
* ``TriangleMesh``  — hussar::TriangleMesh with addQuad/addBox (libhussar/include/hussar/core/mesh.h:14-55)
* ``NFAntenna``     — hussar::NFAntenna + AWRAngularDistribution (libhussar/include/hussar/core/emitter.h:159-214)
* ``Scene``         — hussar::Scene (libhussar/include/hussar/core/scene.h:10-16), with the
                      multi-receiver extension (``rxs``)

Each class converts itself into the plain-old-data structs of ``include/hussar_hip.h``
(``to_c()``); those structs are defined here with ctypes.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field
from typing import List, Sequence

import numpy as np

MAX_CHANNELS = 16
MAX_TX = 4

# Units as in libradar/include/radar/units.h:8-37 (values in SI).
GHz, MHz, kHz, Hz = 1e9, 1e6, 1e3, 1.0
s_, ms, us, ns = 1.0, 1e-3, 1e-6, 1e-9
m_, cm, mm = 1.0, 1e-2, 1e-3

# radar.h:107 `constexpr float SPEED_OF_LIGHT = 299792458;` (299792448 after float rounding)
SPEED_OF_LIGHT = float(np.float32(299792458.0))


class c_rf_config(C.Structure):
    _fields_ = [("startFreq", C.c_float), ("freqSlope", C.c_float), ("adcRate", C.c_float),
                ("idleTime", C.c_float), ("rampTime", C.c_float), ("antennaDelay", C.c_float)]


class c_frame_config(C.Structure):
    _fields_ = [("chirpCount", C.c_int32), ("samplesPerChirp", C.c_int32), ("channelCount", C.c_int32)]


class c_antenna(C.Structure):
    _fields_ = [("position", C.c_float * 3), ("rotation", C.c_float * 9)]


class c_scene(C.Structure):
    _fields_ = [("rfConfig", c_rf_config), ("tx", c_antenna), ("rx", c_antenna * MAX_CHANNELS),
                ("rxCount", C.c_int32), ("txCount", C.c_int32), ("txMore", c_antenna * (MAX_TX - 1)),
                ("velocity", C.c_float * 3), ("reserved0", C.c_int32)]


class c_knobs(C.Structure):
    _fields_ = [("onlyIndirect", C.c_int32), ("maxDepth", C.c_int32),
                ("useGeometricalOptics", C.c_int32), ("poDiffraction", C.c_int32),
                ("doGuiding", C.c_int32), ("clearBeforeIteration", C.c_int32),
                ("correctPhase", C.c_int32), ("filterGuiding", C.c_int32),
                ("doFiltering", C.c_int32), ("filteringSphere", C.c_int32),
                ("filteringMin", C.c_float), ("filteringMax", C.c_float),
                ("filteringRadius", C.c_float), ("primaryOrder", C.c_int32),
                ("produceDebugImage", C.c_int32)]


class c_run_stats(C.Structure):
    _fields_ = [("trace_ms", C.c_double), ("reduce_ms", C.c_double), ("total_ms", C.c_double),
                ("paths", C.c_int64), ("launches", C.c_int64), ("nodes_visited", C.c_int64),
                ("tris_tested", C.c_int64), ("rays_closest", C.c_int64),
                ("rays_shadow", C.c_int64), ("events", C.c_int64), ("bvh_build_ms", C.c_double),
                ("closest_ms", C.c_double), ("shadow_ms", C.c_double), ("closest_launches", C.c_int64),
                ("shadow_launches", C.c_int64), ("nodes_shadow", C.c_int64), ("tris_shadow", C.c_int64),
                ("graph_launches", C.c_int64), ("graph_captures", C.c_int64), ("nodes_tail", C.c_int64),
                ("tris_tail", C.c_int64), ("tail_batches", C.c_int64 * 8),
                ("primary_builds", C.c_int64), ("primary_ms", C.c_double)]


class c_device_info(C.Structure):
    _fields_ = [("name", C.c_char * 64), ("arch", C.c_char * 32), ("compute_units", C.c_int32),
                ("xcds", C.c_int32), ("hbm_bytes", C.c_int64), ("lds_bytes_per_cu", C.c_int32),
                ("clock_mhz", C.c_int32), ("lds_bytes_per_block", C.c_int32), ("reserved0", C.c_int32)]


class c_server_config(C.Structure):
    _fields_ = [("device", C.c_int32), ("slots", C.c_int32), ("queueCapacity", C.c_int32),
                ("maxDepth", C.c_int32), ("hbmBudget", C.c_int64), ("pathsPerBatch", C.c_int64),
                ("frameReserveBytes", C.c_int64), ("graphs", C.c_int32), ("reserved0", C.c_int32)]


class c_request(C.Structure):
    _fields_ = [("sceneId", C.c_int32), ("guideMode", C.c_int32), ("paths", C.c_int64), ("scene", c_scene),
                ("knobs", c_knobs), ("frame", c_frame_config), ("reserved1", C.c_int32)]


class c_request_info(C.Structure):
    _fields_ = [("status", C.c_int32), ("slot", C.c_int32), ("queue_ms", C.c_double), ("run_ms", C.c_double),
                ("latency_ms", C.c_double), ("totalWeight", C.c_double)]


class c_server_stats(C.Structure):
    _fields_ = [("submitted", C.c_int64), ("completed", C.c_int64), ("failed", C.c_int64),
                ("queued", C.c_int64), ("running", C.c_int64), ("p50_ms", C.c_double), ("p99_ms", C.c_double),
                ("mean_ms", C.c_double), ("max_ms", C.c_double), ("frames_per_s", C.c_double),
                ("hbm_used", C.c_int64), ("hbm_budget", C.c_int64), ("scenes", C.c_int32), ("slots", C.c_int32),
                ("rejected", C.c_int64), ("memory_waits", C.c_int64)]


@dataclass
class RFConfig:
    """radar::RFConfig (radar.h:113-136)."""
    startFreq: float = 77 * GHz
    freqSlope: float = 60 * MHz / us
    adcRate: float = 10 * MHz
    idleTime: float = 100 * us
    rampTime: float = 25.6 * us
    antennaDelay: float = 0.0

    def bandwidth(self) -> float:  # radar.h:128-130
        return float(np.float32(self.freqSlope) * np.float32(self.rampTime))

    def chirpFrequency(self) -> float:  # radar.h:133-135
        return 1.0 / (self.idleTime + self.rampTime)

    def to_c(self) -> c_rf_config:
        return c_rf_config(self.startFreq, self.freqSlope, self.adcRate, self.idleTime,
                           self.rampTime, self.antennaDelay)


@dataclass
class FrameConfig:
    """radar::FrameConfig (radar.h:141-164); frame layout [chirp][sample][channel]."""
    chirpCount: int = 1
    samplesPerChirp: int = 256
    channelCount: int = 1

    def sampleCount(self) -> int:
        return self.chirpCount * self.samplesPerChirp * self.channelCount

    def to_c(self) -> c_frame_config:
        return c_frame_config(self.chirpCount, self.samplesPerChirp, self.channelCount)


class TriangleMesh:
    """hussar::TriangleMesh (mesh.h:14-55): float32 vertices, int32 index triplets."""

    def __init__(self, vertices: np.ndarray | None = None, indices: np.ndarray | None = None):
        self._v: List[np.ndarray] = []
        self._i: List[np.ndarray] = []
        self._nv = 0
        if vertices is not None:
            self.append(vertices, indices)

    def append(self, vertices: np.ndarray, indices: np.ndarray) -> None:
        v = np.ascontiguousarray(vertices, dtype=np.float32).reshape(-1, 3)
        i = np.ascontiguousarray(indices, dtype=np.int32).reshape(-1, 3) + self._nv
        self._v.append(v)
        self._i.append(i)
        self._nv += v.shape[0]

    def addQuad(self, a, b, c) -> None:
        """mesh.h:30-40: vertices a, a+b, a+b+c, a+c; triangles (0,1,2), (0,2,3)."""
        a = np.asarray(a, np.float32)
        b = np.asarray(b, np.float32)
        c = np.asarray(c, np.float32)
        ab = a + b
        v = np.stack([a, ab, ab + c, a + c]).astype(np.float32)
        self.append(v, np.array([[0, 1, 2], [0, 2, 3]], np.int32))

    def addBox(self, lo, hi) -> None:
        """mesh.h:42-54: six quads, the first three from min, the last three from max."""
        lo = np.asarray(lo, np.float32)
        hi = np.asarray(hi, np.float32)
        d = hi - lo
        z = np.float32(0)
        self.addQuad(lo, [z, d[1], z], [z, z, d[2]])
        self.addQuad(lo, [z, z, d[2]], [d[0], z, z])
        self.addQuad(lo, [d[0], z, z], [z, d[1], z])
        d = lo - hi
        self.addQuad(hi, [z, d[1], z], [d[0], z, z])
        self.addQuad(hi, [z, z, d[2]], [z, d[1], z])
        self.addQuad(hi, [d[0], z, z], [z, z, d[2]])

    @property
    def vertexBuffer(self) -> np.ndarray:
        return np.concatenate(self._v) if self._v else np.zeros((0, 3), np.float32)

    @property
    def indexBuffer(self) -> np.ndarray:
        return np.concatenate(self._i) if self._i else np.zeros((0, 3), np.int32)

    def triangle_count(self) -> int:
        return sum(i.shape[0] for i in self._i)


def rotation_y(angle_rad: float) -> np.ndarray:
    """Eigen::AngleAxisf(angle, UnitY).toRotationMatrix(), evaluated in double, as float32."""
    c, s = np.cos(angle_rad), np.sin(angle_rad)
    return np.array([[c, 0, s], [0, 1, 0], [-s, 0, c]], np.float64)


@dataclass
class NFAntenna:
    """hussar::NFAntenna with the AWR radiation pattern (emitter.h:159-214).

    ``rotation``'s columns are the antenna's local axes in world space; local +z is the
    boresight. The reference asserts |det R| ≈ 1 (emitter.h:168-171)."""
    position: Sequence[float] = (0.0, 0.0, 0.0)
    rotation: np.ndarray = field(default_factory=lambda: np.eye(3))

    def __post_init__(self):
        self.position = np.asarray(self.position, np.float32).reshape(3)
        self.rotation = np.asarray(self.rotation, np.float32).reshape(3, 3)
        det = float(np.linalg.det(self.rotation.astype(np.float64)))
        if abs(abs(det) - 1.0) >= 1e-2:
            raise ValueError("supplied rotation matrix is - in fact - not a rotation matrix!")

    def to_c(self) -> c_antenna:
        a = c_antenna()
        for k in range(3):
            a.position[k] = float(self.position[k])
        for k in range(9):
            a.rotation[k] = float(self.rotation.reshape(-1)[k])
        return a


@dataclass
class Scene:
    """hussar::Scene (scene.h:10-16): rfConfig, tx and rx, with the extensions of the C ABI:
    ``rxs`` lists the receivers (when empty, the reference's one receiver ``rx``), ``txs``
    the MIMO transmitters (when empty, ``tx``; pair (t, r) feeds channel t*len(rxs) + r),
    and ``velocity`` is the platform velocity of all antennas [m/s]."""
    rfConfig: RFConfig = field(default_factory=RFConfig)
    tx: NFAntenna = field(default_factory=NFAntenna)
    rx: NFAntenna = field(default_factory=NFAntenna)
    rxs: List[NFAntenna] = field(default_factory=list)
    txs: List[NFAntenna] = field(default_factory=list)
    velocity: Sequence[float] = (0.0, 0.0, 0.0)

    def receivers(self) -> List[NFAntenna]:
        return list(self.rxs) if self.rxs else [self.rx]

    def transmitters(self) -> List[NFAntenna]:
        return list(self.txs) if self.txs else [self.tx]

    def to_c(self) -> c_scene:
        sc = c_scene()
        sc.rfConfig = self.rfConfig.to_c()
        rx = self.receivers()
        tx = self.transmitters()
        if len(rx) > MAX_CHANNELS:
            raise ValueError(f"at most {MAX_CHANNELS} receivers")
        if len(tx) > MAX_TX:
            raise ValueError(f"at most {MAX_TX} transmitters")
        for k, a in enumerate(rx):
            sc.rx[k] = a.to_c()
        sc.rxCount = len(rx)
        sc.tx = tx[0].to_c()
        for k, a in enumerate(tx[1:]):
            sc.txMore[k] = a.to_c()
        sc.txCount = len(tx)
        v = np.asarray(self.velocity, np.float32).reshape(3)
        for k in range(3):
            sc.velocity[k] = float(v[k])
        return sc


@dataclass
class Knobs:
    """PathTracer's public knobs with the reference defaults (path.h:153-169)."""
    onlyIndirect: bool = True
    maxDepth: int = 10
    useGeometricalOptics: bool = True
    poDiffraction: bool = False
    doGuiding: bool = True
    clearBeforeIteration: bool = True
    correctPhase: bool = False
    filterGuiding: bool = True
    doFiltering: bool = True
    filteringSphere: bool = True
    filteringMin: float = 600.0
    filteringMax: float = 900.0
    filteringRadius: float = 160.0
    primaryOrder: int = 0  # 0: g++ evaluation order of get2D (halton.h:24-27), 1: clang's
    produceDebugImage: bool = False  # only its effect on the splat skip rule (integrator.h:121-124)

    def to_c(self) -> c_knobs:
        return c_knobs(int(self.onlyIndirect), int(self.maxDepth), int(self.useGeometricalOptics),
                       int(self.poDiffraction), int(self.doGuiding), int(self.clearBeforeIteration),
                       int(self.correctPhase), int(self.filterGuiding), int(self.doFiltering),
                       int(self.filteringSphere), float(self.filteringMin), float(self.filteringMax),
                       float(self.filteringRadius), int(self.primaryOrder), int(self.produceDebugImage))
