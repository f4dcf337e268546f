"""The serving runtime (hsr_server_*): resident scenes and their guiding trees, slots with their
own streams, FIFO queue, admission by footprint, statistics. Served frames must equal those of
a direct run of the same request on a backend of its own (same kernels, same arithmetic, and
training statistics summed in a fixed order; only the order of the frame's floating-point sums
may differ), and a request served with its scene's resident tree must give the same frame
alone and in a mix of other requests."""
import time

import numpy as np
import pytest

import hussar_amd as H
from hussar_amd import scenes
from hussar_amd._lib import HsrError
from hussar_amd.serving import GUIDE_PER_REQUEST, GUIDE_SCENE, GUIDE_TRAIN, Server
from conftest import rel_l2

THREADS = 16

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
        # guided requests train the same tree as the direct run (fixed-order statistics)
        assert rel_l2(z, ref) < 1e-6, (s, fc)
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
    small = Server(slots=1, pathsPerBatch=1 << 16, hbmBudget=1 << 28)
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


def clutter_pose(channels, x=0.5):
    return scenes.radar_array(scenes.dihedral_rf(), channels=channels, position=(x, 0.75, 0.0))


def test_guided_training_is_deterministic(meshes):
    """PathTracer::run's training statistics are summed in a fixed order: the same guided run
    twice gives the same sampling tree bit for bit, and the same frame up to the order of its
    splat sums."""
    sc = clutter_pose(4)
    kn = H.Knobs(maxDepth=4, doGuiding=True)
    fc = H.FrameConfig(1, 256, 4)
    pt = H.PathTracer(**{k: getattr(kn, k) for k in kn.__dataclass_fields__})
    pt.configureFrame(fc)
    b = H.HipBackend(meshes[1], pt)
    trees, frames = [], []
    for _ in range(3):
        pt.run(b, sc, 1 << 20)
        trees.append(pt.guiding.sampling())
        frames.append(b.fetch("complex128")[0])
    assert len(trees[0][0]) > 3
    for child, p0 in trees[1:]:
        np.testing.assert_array_equal(child, trees[0][0])
        np.testing.assert_array_equal(p0, trees[0][1])
    for z in frames[1:]:
        assert rel_l2(z, frames[0]) < 1e-12


def direct_scene_pass(mesh, scene, knobs, fc, paths, tree):
    """One final pass with a given sampling tree on a backend of its own."""
    b = H.HipBackend(mesh)
    b._configure(fc)
    g = H.Guide(b)
    g.set_sampling(*tree)
    b.run_raw(scene, knobs, 0, paths, guide=g, is_final=True)
    z, tw = b.fetch("complex64")
    return z.reshape(max(1, fc.chirpCount), fc.samplesPerChirp, fc.channelCount), tw


def test_scene_guide_alone_and_in_the_mix(oracle, meshes):
    """A scene's resident tree: a training request keeps its tree (the same tree a direct
    PathTracer::run of it trains); a request served with the tree is one final pass of its whole
    budget, equal to that pass run directly, within 1e-5 of the oracle with the same tree, and
    the same frame alone and in a mix of other requests on three slots."""
    dihedral, clutter = meshes
    srv = Server(slots=3, pathsPerBatch=1 << 20)
    ids = [srv.open_scene(m) for m in meshes]
    kn = H.Knobs(maxDepth=4, doGuiding=True)
    fc = H.FrameConfig(1, 256, 4)
    train_pose = clutter_pose(4, x=0.0)
    print("serving: training the clutter scene's tree", flush=True)
    _, info = srv.wait(srv.submit(ids[1], train_pose, kn, fc, 1 << 20, GUIDE_TRAIN))
    assert info.status == 0
    pt = H.PathTracer(**{k: getattr(kn, k) for k in kn.__dataclass_fields__})
    pt.configureFrame(fc)
    pt.run(H.HipBackend(clutter, pt), train_pose, 1 << 20)
    tree = pt.guiding.sampling()
    req = (ids[1], clutter_pose(4, x=0.3), kn, fc, 1 << 18)
    z_alone, info = srv.wait(srv.submit(*req, guide_mode=GUIDE_SCENE))
    assert info.status == 0 and info.totalWeight == 1 << 18  # the whole budget is the final pass
    ref, tw = direct_scene_pass(clutter, req[1], kn, fc, 1 << 18, tree)
    assert tw == 1 << 18 and np.abs(ref).max() > 0
    assert rel_l2(z_alone, ref) < 1e-6
    om = oracle.OracleMesh.from_mesh(clutter)
    og = oracle.OracleGuide()
    og.set_sampling(*tree)
    zo, two, _, _ = oracle.render(om, req[1], kn, fc, prec=oracle.PREC_MIXED, count=1 << 18, is_final=True,
                                  guide=og, threads=THREADS)
    assert rel_l2(z_alone, (zo / two).reshape(z_alone.shape)) < 1e-5
    print("serving: the same request inside a mix", flush=True)
    others = [(ids[0], scenes.dihedral_pose(5.0), H.Knobs(maxDepth=3, doGuiding=True), H.FrameConfig(1, 256, 1),
               1 << 18, GUIDE_PER_REQUEST),
              (ids[1], clutter_pose(8, x=-0.4), kn, H.FrameConfig(1, 512, 8), 1 << 19, GUIDE_SCENE),
              (ids[1], clutter_pose(2), H.Knobs(maxDepth=3, doGuiding=False), H.FrameConfig(1, 128, 2), 1 << 18, 0)]
    tickets = []
    for o in others:
        tickets.append(srv.submit(*o[:5], guide_mode=o[5]))
    mine = srv.submit(*req, guide_mode=GUIDE_SCENE)
    for o in others:
        tickets.append(srv.submit(*o[:5], guide_mode=o[5]))
    z_mix, info = srv.wait(mine)
    assert info.status == 0
    for t in tickets:
        assert srv.wait(t, fetch=False)[1].status == 0
    assert rel_l2(z_mix, z_alone) < 1e-6
    # an untrained scene trains its tree from its first GUIDE_SCENE request (the reference's
    # schedule), and serves the next ones from it
    _, info = srv.wait(srv.submit(ids[0], scenes.dihedral_pose(0.0), H.Knobs(maxDepth=3, doGuiding=True),
                                  H.FrameConfig(1, 256, 1), 1 << 18, guide_mode=GUIDE_SCENE))
    assert info.totalWeight == 147456
    _, info = srv.wait(srv.submit(ids[0], scenes.dihedral_pose(3.0), H.Knobs(maxDepth=3, doGuiding=True),
                                  H.FrameConfig(1, 256, 1), 1 << 18, guide_mode=GUIDE_SCENE))
    assert info.totalWeight == 1 << 18
    srv.close()


def test_admission_by_footprint(meshes):
    """Requests start when their own footprint fits beside the resident scenes: with room for
    one request at a time, a second one waits for memory (and then gives the same frame), and a
    request that could never fit is refused at submit."""
    clutter = meshes[1]
    kn = H.Knobs(maxDepth=4, doGuiding=False)
    fc = H.FrameConfig(1, 256, 4)
    sc = clutter_pose(4)
    probe = Server(slots=1, pathsPerBatch=1 << 20)
    pid = probe.open_scene(clutter)
    scene_bytes = probe.stats()["hbm_used"]
    fp = probe.request_bytes(pid, sc, kn, fc, 1 << 20)
    big = probe.request_bytes(pid, clutter_pose(12), kn, H.FrameConfig(1, 512, 12), 1 << 20)
    probe.close()
    assert scene_bytes > 0 and fp > 0 and big > 1.5 * fp
    srv = Server(slots=2, pathsPerBatch=1 << 20, hbmBudget=scene_bytes + int(1.5 * fp))
    sid = srv.open_scene(clutter)
    print(f"admission: footprint {fp >> 20} MiB, budget {(scene_bytes + int(1.5 * fp)) >> 20} MiB", flush=True)
    tickets = [srv.submit(sid, sc, kn, fc, 1 << 20) for _ in range(3)]
    frames = [srv.wait(t) for t in tickets]
    assert all(info.status == 0 for _, info in frames)
    for z, _ in frames[1:]:
        assert rel_l2(z, frames[0][0]) < 1e-6
    st = srv.stats()
    assert st["memory_waits"] >= 1 and st["completed"] == 3
    assert st["hbm_used"] <= st["hbm_budget"]
    with pytest.raises(HsrError, match="OOM"):
        srv.submit(sid, clutter_pose(12), kn, H.FrameConfig(1, 512, 12), 1 << 20)
    assert srv.stats()["rejected"] == 1
    srv.close()
