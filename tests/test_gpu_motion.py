"""GPU parity of the motion (Doppler chirp axis) and MIMO extensions against the oracle,
through the C ABI. The oracle's own pins for these (closed-form cube taps, static and
relative-motion identities, Doppler bin of an approaching plate, MIMO blocks) are in
test_oracle_motion.py.

Bar: hit ids and shadow outcomes bit-exact against the oracle's mixed-precision instance
(the path-length rate is FP64 on both sides, in the same order), the frame cube within 1e-5
relative L2 of it and its peak cells identical.
"""
import os

import numpy as np
import pytest

import hussar_amd as H
from hussar_amd import scenes
from hussar_amd._lib import HsrError, OPT_PIPELINE, OPT_RECORD_HITS
from hussar_amd.types import RFConfig
from conftest import rel_l2

pytestmark = pytest.mark.gpu
THREADS = min(16, os.cpu_count() or 1)


def rf_motion():
    return RFConfig(startFreq=77e9, freqSlope=60e12, adcRate=10e6, idleTime=100e-6, rampTime=25.6e-6)


@pytest.fixture(scope="module")
def clutter():
    mesh, _ = scenes.vehicle_mesh(seed=5, triangles=20_000, range_lo=8.0, range_hi=12.0)
    # every box / panel moves: a per-triangle velocity up to 20 m/s, seeded
    u = scenes.uniforms(11, 3 * mesh.triangle_count()).reshape(-1, 3)
    vel = ((u - 0.5) * 40.0).astype(np.float32)
    return mesh, vel


def run_both(oracle, mesh, vel, scene, knobs, fc, n, record):
    print(f"motion parity: chirps={fc.chirpCount} channels={fc.channelCount} n={n}", flush=True)
    b = H.HipBackend(mesh)
    b._configure(fc)
    b.set_velocities(vel)
    b.clear()
    b.set_option(OPT_RECORD_HITS, record)
    b.run_raw(scene, knobs, 0, n)
    z, tw = b.fetch("complex128")
    hits, nee = b.hit_record(record, max(1, knobs.maxDepth), fc.channelCount)
    om = oracle.OracleMesh.from_mesh(mesh)
    om.set_velocities(vel)
    zo, two, reco, _ = oracle.render(om, scene, knobs, fc, prec=oracle.PREC_MIXED, count=n, record_paths=record,
                                     threads=THREADS)
    assert tw == two == n
    np.testing.assert_array_equal(hits, reco[0])
    np.testing.assert_array_equal(nee, reco[1])
    zo = zo / two
    return z.reshape(zo.shape), zo


@pytest.mark.parametrize("chirps", [1, 16, 64])
def test_moving_clutter_cube_parity(oracle, clutter, chirps):
    mesh, vel = clutter
    rf = rf_motion()
    sc = scenes.radar_array(rf, channels=4, position=(0.0, 0.75, 0.0))
    sc.velocity = (0.0, 0.0, 8.0)  # the platform drives towards the clutter too
    kn = H.Knobs(maxDepth=3, doGuiding=False)
    fc = H.FrameConfig(chirps, 64, 4)
    z, zo = run_both(oracle, mesh, vel, sc, kn, fc, 1 << 15, record=1 << 12)
    assert np.abs(zo).max() > 0
    assert rel_l2(z, zo) < 1e-5
    assert int(np.abs(z).argmax()) == int(np.abs(zo).argmax())
    if chirps > 1:
        assert np.count_nonzero(np.abs(zo).sum(axis=(1, 2)) > 0) > chirps // 2  # energy spread over Doppler


def test_mimo_2tx_4rx_parity(oracle, clutter):
    mesh, _ = clutter
    rf = scenes.dihedral_rf()
    sc = scenes.radar_array(rf, channels=4, position=(0.0, 0.75, 0.0))
    tx1 = H.NFAntenna(np.asarray(sc.tx.position) + np.float32([0.0078, 0.0, 0.0]), sc.tx.rotation)
    mimo = H.Scene(rfConfig=rf, tx=sc.tx, rx=sc.rx, rxs=sc.rxs, txs=[sc.tx, tx1])
    kn = H.Knobs(maxDepth=3, doGuiding=False)
    fc = H.FrameConfig(1, 256, 8)
    z, zo = run_both(oracle, mesh, None, mimo, kn, fc, 1 << 15, record=1 << 12)
    assert rel_l2(z, zo) < 1e-5
    assert np.all(np.abs(zo).max(axis=(0, 1)) > 0)  # all eight virtual channels carry signal


def test_zero_velocities_match_the_static_run(clutter):
    mesh, vel = clutter
    sc = scenes.radar_array(rf_motion(), channels=4, position=(0.0, 0.75, 0.0))
    kn = H.Knobs(maxDepth=3, doGuiding=False)
    fc = H.FrameConfig(1, 256, 4)
    b = H.HipBackend(mesh)
    b._configure(fc)
    b.run_raw(sc, kn, 0, 1 << 16)
    z0, _ = b.fetch("complex128")
    b.set_velocities(np.zeros_like(vel))
    b.clear()
    b.run_raw(sc, kn, 0, 1 << 16)
    z1, _ = b.fetch("complex128")
    assert rel_l2(z1, z0) < 1e-12


def test_motion_errors(clutter):
    mesh, vel = clutter
    sc = scenes.radar_array(rf_motion(), channels=4)
    kn = H.Knobs(maxDepth=2, doGuiding=False)
    b = H.HipBackend(mesh)
    b._configure(H.FrameConfig(8, 64, 4))
    with pytest.raises(HsrError, match="INVALID"):
        b.set_velocities(vel[:-1])
    b.set_option(OPT_PIPELINE, 0)
    with pytest.raises(HsrError, match="UNSUPPORTED"):
        b.run_raw(sc, kn, 0, 1024)  # the fused kernel has no chirp axis
    b._configure(H.FrameConfig(1, 64, 4))
    b.set_velocities(vel)
    with pytest.raises(HsrError, match="UNSUPPORTED"):
        b.run_raw(sc, kn, 0, 1024)
    b.set_option(OPT_PIPELINE, 1)
    b.run_raw(sc, kn, 0, 1024)
    three = H.Scene(rfConfig=sc.rfConfig, tx=sc.tx, rx=sc.rx, rxs=sc.rxs, txs=[sc.tx, sc.tx])
    with pytest.raises(HsrError, match="INVALID"):
        b.run_raw(three, kn, 0, 1024)  # 2 x 4 virtual channels > 4 frame channels


def test_cube_event_list_overflow(clutter):
    """Cube events beyond the event list of the LDS tile splat go straight to the frame with
    device atomics: the same run with a list of 16 events, with none at all, and with the
    default list (every event through LDS). Same taps, so the cubes agree to rounding."""
    from hussar_amd._lib import OPT_COUNTERS, OPT_CUBE_EVENTS
    mesh, vel = clutter
    sc = scenes.radar_array(rf_motion(), channels=4, position=(0.0, 0.75, 0.0))
    sc.velocity = (0.0, 0.0, 8.0)
    kn = H.Knobs(maxDepth=3, doGuiding=False)
    fc = H.FrameConfig(16, 64, 4)
    n = 1 << 15
    frames = []
    for cap in (-1, 16, 0):
        print(f"cube event list: capacity {cap}", flush=True)
        b = H.HipBackend(mesh)
        b._configure(fc)
        b.set_velocities(vel)
        if cap >= 0:
            b.set_option(OPT_CUBE_EVENTS, cap)
        b.set_option(OPT_COUNTERS, 1)
        b.run_raw(sc, kn, 0, n)
        z, tw = b.fetch("complex128")
        assert tw == n
        frames.append(z)
        if cap == 16:
            assert b.stats()["events"] > 16  # the list overflowed
    assert np.abs(frames[0]).max() > 0
    assert rel_l2(frames[1], frames[0]) < 1e-9
    assert rel_l2(frames[2], frames[0]) < 1e-9
    b = H.HipBackend(mesh)
    with pytest.raises(HsrError, match="INVALID"):
        b.set_option(OPT_CUBE_EVENTS, (1 << 24) + 1)
