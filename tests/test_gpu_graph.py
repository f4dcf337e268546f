"""Graph replay of the wavefront batches (HSR_OPT_GRAPH), through the C ABI.

A batch's launch sequence (set-up, generation, per depth the traversals, shading and gather
stages with their stream fork / join and timing events) is built into a HIP graph at the
second batch of its launch shape (the first one launches stage by stage) and replayed from
then on, each kernel node given the batch's own arguments before the replay. Bar, for every replay: hit
ids and shadow outcomes equal those of the
same batch launched stage by stage, the frame within 1e-12 relative L2 of it (only the order
of the LDS tile's FP64 additions differs), the same total weight; between replays the
program does eager work of its own (stage-by-stage runs of another shape, fetches,
synchronisations). The replays must really happen: the run statistics count them.
"""
import numpy as np
import pytest

import hussar_amd as H
from hussar_amd import scenes
from hussar_amd._lib import OPT_GRAPH, OPT_RECORD_HITS
from conftest import rel_l2

pytestmark = pytest.mark.gpu


def run(b, sc, kn, n, graph, record=0, offset=0, guide=None, is_final=False):
    b.set_option(OPT_GRAPH, 1 if graph else 0)
    b.set_option(OPT_RECORD_HITS, record)
    b.clear()
    b.run_raw(sc, kn, offset, n, guide=guide, is_final=is_final)
    z, tw = b.fetch("complex128")
    rec = b.hit_record(record, max(1, kn.maxDepth), b.frame_config.channelCount) if record else None
    return z, tw, rec, b.stats()


@pytest.fixture(scope="module")
def vehicle():
    mesh, _ = scenes.vehicle_mesh(seed=3, triangles=50_000)
    b = H.HipBackend(mesh)
    fc = H.FrameConfig(1, 256, 4)
    b._configure(fc)
    sc = scenes.radar_array(scenes.dihedral_rf(), channels=4, position=(0.0, 0.75, 0.0))
    return mesh, b, fc, sc


def test_replays_equal_stage_launches(vehicle):
    mesh, b, fc, sc = vehicle
    kn = H.Knobs(maxDepth=4, doGuiding=False)
    n, rec = 1 << 18, 1 << 14
    print("graph: stage-by-stage reference", flush=True)
    z0, tw0, rec0, st0 = run(b, sc, kn, n, graph=False, record=rec)
    assert st0["graph_launches"] == 0
    assert np.abs(z0).max() > 0
    other = scenes.radar_array(scenes.dihedral_rf(), channels=2, position=(0.5, 0.75, 0.0))
    for rep in range(5):
        print(f"graph: replay {rep}", flush=True)
        z, tw, r, st = run(b, sc, kn, n, graph=True, record=rec)
        # the shape's first batch launches stage by stage, the second builds its graph
        assert st["graph_launches"] == (0 if rep == 0 else 1)
        assert st["graph_captures"] == (1 if rep == 1 else 0)
        assert tw == tw0
        np.testing.assert_array_equal(r[0], rec0[0])
        np.testing.assert_array_equal(r[1], rec0[1])
        assert rel_l2(z, z0) < 1e-12
        assert st["closest_launches"] == st0["closest_launches"] and st["closest_ms"] > 0
        # eager work of another shape between the replays
        ze, _, _, ste = run(b, other, kn, 1 << 16, graph=False)
        assert ste["graph_launches"] == 0 and np.isfinite(ze).all()


def test_replays_with_new_arguments(vehicle):
    """One shape, other arguments at each replay (sample offsets, poses): each replay equals
    the stage launches of its own arguments."""
    mesh, b, fc, sc = vehicle
    kn = H.Knobs(maxDepth=3, doGuiding=False)
    n = 1 << 17
    run(b, sc, kn, n, graph=True, record=4096)  # the shape's first batch
    for rep, (off, x) in enumerate([(0, 0.0), (n, 0.3), (3 * n, -0.4), (0, 0.0)]):
        sc_r = scenes.radar_array(scenes.dihedral_rf(), channels=4, position=(x, 0.75, 0.0))
        z0, tw0, r0, _ = run(b, sc_r, kn, n, graph=False, record=4096, offset=off)
        z, tw, r, st = run(b, sc_r, kn, n, graph=True, record=4096, offset=off)
        print(f"graph: arguments {rep}: offset {off}, x {x}, |frame| {np.abs(z0).max():.3g}", flush=True)
        assert st["graph_launches"] == 1 and tw == tw0
        np.testing.assert_array_equal(r[0], r0[0])
        np.testing.assert_array_equal(r[1], r0[1])
        assert rel_l2(z, z0) < 1e-12


def test_guided_schedule_and_mimo_cube_replay():
    """A guided schedule (training batches with their statistics, then the final pass) and a
    moving MIMO cube (one batch per transmitter, the cube event list) replayed as graphs give
    the stage launches' frames and trees."""
    mesh, _ = scenes.vehicle_mesh(seed=6, triangles=20_000, range_lo=8.0, range_hi=12.0)
    rf = scenes.dihedral_rf()
    sc = scenes.radar_array(rf, channels=4, position=(0.0, 0.75, 0.0))
    out = {}
    for graph, backend_new, replays in ((False, True, 0), (True, True, 2), (True, False, 4), (True, False, 4)):
        if backend_new:
            pt = H.PathTracer(maxDepth=3, doGuiding=True)
            pt.configureFrame(H.FrameConfig(1, 256, 4))
            b = H.HipBackend(mesh, pt)
        b.set_option(OPT_GRAPH, 1 if graph else 0)
        pt.run_schedule(b, sc, 3, 1 << 16, 1 << 18)
        z = pt.fetchFrame("complex128")
        st = b.stats()
        child, p0 = pt.guiding.sampling()
        print(f"graph: guided schedule, graph={graph}, replays {st['graph_launches']}", flush=True)
        # three training batches of one shape (the second and third replayed at the first
        # schedule), then the final batch of another (replayed from the second schedule on)
        assert st["graph_launches"] == replays
        if graph:
            np.testing.assert_array_equal(child, out["child"])
            np.testing.assert_array_equal(p0, out["p0"])
            assert rel_l2(z, out["z"]) < 1e-12
        else:
            out.update(z=z, child=child, p0=p0)
    u = scenes.uniforms(17, 3 * mesh.triangle_count()).reshape(-1, 3)
    vel = ((u - 0.5) * 30.0).astype(np.float32)
    mimo = scenes.mimo_array(rf, tx=3, rx=4, position=(0.0, 1.0, 0.0))
    kn = H.Knobs(maxDepth=3, doGuiding=False)
    fc = H.FrameConfig(16, 256, 12)
    b = H.HipBackend(mesh)
    b._configure(fc)
    b.set_velocities(vel)
    z0, tw0, r0, _ = run(b, mimo, kn, 1 << 16, graph=False, record=2048)
    for rep in range(3):
        z, tw, r, st = run(b, mimo, kn, 1 << 16, graph=True, record=2048)
        print(f"graph: MIMO cube replay {rep}: replays {st['graph_launches']}", flush=True)
        # one batch per transmitter, one shape: the first run's first batch launches stage by stage
        assert st["graph_launches"] == (2 if rep == 0 else 3) and tw == tw0
        np.testing.assert_array_equal(r[0], r0[0])
        np.testing.assert_array_equal(r[1], r0[1])
        assert rel_l2(z, z0) < 1e-12
