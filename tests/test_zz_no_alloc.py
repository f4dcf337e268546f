"""The allocation-free peak search (hsr_hip_frame_peak_noalloc) called through the C ABI gives
the results of hsr_hip_frame_peak bit for bit, on two calls in a row whose second frame is
smaller than the first (the shared static scratch keeps the first call's contents beyond the
second frame), in the frame and in its DFT; frames beyond the scratch's limit are refused."""
import ctypes as C

import numpy as np
import pytest

import hussar_amd as H
from hussar_amd import scenes
from hussar_amd._lib import lib
from hussar_amd.types import RFConfig

pytestmark = pytest.mark.gpu

HSR_OK, HSR_ERR_INVALID = 0, 1
NOALLOC_CELLS = 1 << 19


def peaks(b, fourier):
    out = []
    for fn in (lib().hsr_hip_frame_peak, lib().hsr_hip_frame_peak_noalloc):
        idx = (C.c_int32 * 3)()
        ref = (C.c_double * 3)()
        mag = C.c_double()
        rc = fn(b._h, int(fourier), idx, ref, C.byref(mag))
        assert rc == HSR_OK, lib().hsr_last_error()
        out.append((tuple(idx), tuple(ref), mag.value))
    return out


def frame_of(mesh, scene, fc, n, vel=None):
    b = H.HipBackend(mesh)
    b._configure(fc)
    if vel is not None:
        b.set_velocities(vel)
    b.run_raw(scene, H.Knobs(maxDepth=3, doGuiding=False), 0, n)
    return b


def test_noalloc_peak_equals_peak_on_two_calls_second_smaller():
    # first call: a 16-chirp x 256-sample x 4-channel cube of moving clutter (16384 cells)
    mesh, _ = scenes.vehicle_mesh(seed=5, triangles=20_000, range_lo=8.0, range_hi=12.0)
    rf = RFConfig(startFreq=77e9, freqSlope=60e12, adcRate=10e6, idleTime=100e-6, rampTime=25.6e-6)
    sc = scenes.radar_array(rf, channels=4, position=(0.0, 0.75, 0.0))
    u = scenes.uniforms(11, 3 * mesh.triangle_count()).reshape(-1, 3)
    vel = ((u - 0.5) * 40.0).astype(np.float32)
    print("no-alloc peak: 16 x 256 x 4 cube", flush=True)
    big = frame_of(mesh, sc, H.FrameConfig(16, 256, 4), 1 << 16, vel)
    for fourier in (0, 1):
        a, n = peaks(big, fourier)
        assert a == n, (fourier, a, n)
        assert a[2] > 0.0
    # second call, smaller input: the dihedral's one-channel frame of 128 samples (128 cells)
    print("no-alloc peak: dihedral 1 x 128 x 1", flush=True)
    small = frame_of(scenes.dihedral_mesh(), scenes.dihedral_pose(0.0), H.FrameConfig(1, 128, 1), 1 << 16)
    for fourier in (0, 1):
        a, n = peaks(small, fourier)
        assert a == n, (fourier, a, n)
        assert a[2] > 0.0
    # and the first backend again after the smaller call
    a, n = peaks(big, 0)
    assert a == n


def test_noalloc_peak_refuses_frames_beyond_its_scratch():
    b = H.HipBackend(scenes.dihedral_mesh())
    b._configure(H.FrameConfig(64, 1024, 16))  # 2^20 cells > 2^19
    idx = (C.c_int32 * 3)()
    ref = (C.c_double * 3)()
    mag = C.c_double()
    rc = lib().hsr_hip_frame_peak_noalloc(b._h, 0, idx, ref, C.byref(mag))
    assert rc == HSR_ERR_INVALID
    assert b"exceeds" in lib().hsr_last_error()
    # exactly at the limit it runs (an empty frame: every value is zero, the first cell wins)
    b2 = H.HipBackend(scenes.dihedral_mesh())
    b2._configure(H.FrameConfig(64, 512, 16))  # 2^19 cells
    assert 64 * 512 * 16 == NOALLOC_CELLS
    rc = lib().hsr_hip_frame_peak_noalloc(b2._h, 0, idx, ref, C.byref(mag))
    assert rc == HSR_OK, lib().hsr_last_error()
    assert tuple(idx) == (0, 0, 0) and mag.value == 0.0
