"""The serving runtime (hsr_server_*): resident scenes, slots with their own streams, FIFO
queue, admission, statistics. Served frames must equal those of a direct PathTracer::run of
the same request on a backend of its own (same kernels, same arithmetic; only the order of
the frame's floating-point sums may differ)."""
import time

import numpy as np
import pytest

import hussar_amd as H
from hussar_amd import scenes
from hussar_amd._lib import HsrError
from hussar_amd.serving import Server
from conftest import rel_l2

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def meshes():
    return scenes.dihedral_mesh(), scenes.vehicle_mesh(seed=10, triangles=10_000, range_lo=8.0, range_hi=14.0)[0]


def direct(mesh, scene, knobs, fc, paths, vel=None):
    pt = H.PathTracer(**{k: getattr(knobs, k) for k in knobs.__dataclass_fields__})
    pt.configureFrame(fc)
    b = H.HipBackend(mesh, pt)
    if vel is not None:
        b.set_velocities(vel)
    pt.run(b, scene, paths)
    z, tw = b.fetch("complex64")
    return z.reshape(max(1, fc.chirpCount), fc.samplesPerChirp, fc.channelCount), tw


def requests(meshes):
    dihedral, clutter = meshes
    rf = scenes.dihedral_rf()
    arr = lambda ch: scenes.radar_array(rf, channels=ch, position=(0.5, 0.75, 0.0))
    return [
        (0, scenes.dihedral_pose(10.0), H.Knobs(maxDepth=3, doGuiding=False), H.FrameConfig(1, 256, 1), 1 << 18),
        (1, arr(4), H.Knobs(maxDepth=4, doGuiding=False), H.FrameConfig(1, 128, 4), 1 << 17),
        (1, arr(12), H.Knobs(maxDepth=3, doGuiding=False), H.FrameConfig(1, 512, 12), 1 << 16),  # cube frame
        (0, scenes.dihedral_pose(-20.0), H.Knobs(maxDepth=3, doGuiding=True), H.FrameConfig(1, 256, 1), 1 << 18),
        (1, arr(2), H.Knobs(maxDepth=4, doGuiding=True), H.FrameConfig(1, 256, 2), 1 << 18),
    ]


def test_served_frames_equal_direct_runs(meshes):
    print("serving: 2 scenes x 5 requests on 3 slots", flush=True)
    srv = Server(slots=3, pathsPerBatch=1 << 20)
    ids = [srv.open_scene(m) for m in meshes]
    reqs = requests(meshes)
    tickets = [srv.submit(ids[s], sc, kn, fc, n) for s, sc, kn, fc, n in reqs]
    for t, (s, sc, kn, fc, n) in zip(tickets, reqs):
        z, info = srv.wait(t)
        ref, tw = direct(meshes[s], sc, kn, fc, n)
        # the guiding schedule's final iteration alone carries the weight (clearBeforeIteration)
        assert info.status == 0 and info.totalWeight == tw == (n if not kn.doGuiding else 147456)
        assert np.abs(ref).max() > 0
        # guided requests train their tree with FP64 atomics in a run-dependent order
        assert rel_l2(z, ref) < (1e-4 if kn.doGuiding else 1e-6), (s, fc)
    st = srv.stats()
    assert st["completed"] == len(reqs) and st["failed"] == 0 and st["queued"] == 0 and st["running"] == 0
    assert 0 < st["p50_ms"] <= st["p99_ms"] <= st["max_ms"]
    assert st["scenes"] == 2 and st["slots"] == 3 and st["hbm_used"] <= st["hbm_budget"]
    srv.close()


def test_concurrent_slots_and_queue(meshes):
    srv = Server(slots=4, pathsPerBatch=1 << 20)
    sid = srv.open_scene(meshes[1])
    sc = scenes.radar_array(scenes.dihedral_rf(), channels=4, position=(0.0, 0.75, 0.0))
    kn = H.Knobs(maxDepth=3, doGuiding=False)
    fc = H.FrameConfig(1, 256, 4)
    tickets = [srv.submit(sid, sc, kn, fc, 1 << 18) for _ in range(16)]
    frames, slots = [], set()
    for t in tickets:
        z, info = srv.wait(t)
        frames.append(z)
        slots.add(info.slot)
    assert len(slots) > 1  # requests ran on several slots
    for z in frames[1:]:
        assert rel_l2(z, frames[0]) < 1e-6  # identical requests, identical frames
    assert srv.stats()["completed"] == 16
    srv.close()


def test_admission_and_errors(meshes):
    srv = Server(slots=1, queueCapacity=1, pathsPerBatch=1 << 20, maxDepth=4, frameReserveBytes=1 << 20)
    sid = srv.open_scene(meshes[1])
    sc = scenes.radar_array(scenes.dihedral_rf(), channels=4, position=(0.0, 0.75, 0.0))
    kn = H.Knobs(maxDepth=4, doGuiding=False)
    with pytest.raises(HsrError, match="OOM"):
        srv.submit(sid, sc, kn, H.FrameConfig(64, 512, 4), 1024)  # 2 MiB x 2 > 1 MiB reservation
    with pytest.raises(HsrError, match="UNSUPPORTED"):
        srv.submit(sid, sc, H.Knobs(maxDepth=8, doGuiding=False), H.FrameConfig(1, 256, 4), 1024)
    with pytest.raises(HsrError, match="INVALID"):
        srv.submit(sid + 7, sc, kn, H.FrameConfig(1, 256, 4), 1024)
    # one running + one queued, the third submit finds the queue full
    got_busy = False
    tickets = []
    for _ in range(3):
        try:
            tickets.append(srv.submit(sid, sc, kn, H.FrameConfig(1, 256, 4), 1 << 23))
        except HsrError as e:
            assert "BUSY" in str(e)
            got_busy = True
    assert got_busy
    srv.close_scene(sid)  # deferred until its requests are done
    with pytest.raises(HsrError, match="INVALID"):
        srv.submit(sid, sc, kn, H.FrameConfig(1, 256, 4), 1024)
    for t in tickets:
        srv.wait(t, fetch=False)
    time.sleep(0.05)
    assert srv.stats()["scenes"] == 0
    srv.close()
    small = Server(slots=1, pathsPerBatch=1 << 16, hbmBudget=1 << 29)
    with pytest.raises(HsrError, match="OOM"):
        small.open_scene(scenes.vehicle_mesh(seed=12, triangles=2_000_000)[0])
    small.close()


def test_destroy_with_queued_requests(meshes):
    srv = Server(slots=1, pathsPerBatch=1 << 20)
    sid = srv.open_scene(meshes[1])
    sc = scenes.radar_array(scenes.dihedral_rf(), channels=4, position=(0.0, 0.75, 0.0))
    for _ in range(4):
        srv.submit(sid, sc, H.Knobs(maxDepth=3, doGuiding=False), H.FrameConfig(1, 256, 4), 1 << 22)
    srv.close()  # interrupts the running request, drops the queued ones
