"""The fused tail (HSR_OPT_TAIL), through the C ABI: one kernel that carries every path still
going after the depth-K extension rays to its end equals the stage launches it replaces.

Bar, against the same batch with every depth's stages (HSR_OPT_TAIL 0): hit ids and shadow
outcomes bit-exact on every recorded path, the frame within 1e-12 relative L2 (only the order
of FP64 sums differs), guided schedules train identical trees; and the tail really ran (its
node visits are counted). The stage launches themselves are held to the oracle by the other
GPU tests, which run with the default tail.
"""
import numpy as np
import pytest

import hussar_amd as H
from hussar_amd import scenes
from hussar_amd._lib import OPT_CHAIN, OPT_COUNTERS, OPT_RECORD_HITS, OPT_TAIL
from conftest import rel_l2

pytestmark = pytest.mark.gpu


def run(b, sc, kn, n, tail, record, chain=0, guide=None):
    b.set_option(OPT_TAIL, tail)
    b.set_option(OPT_CHAIN, chain)
    b.set_option(OPT_COUNTERS, 1)
    b.set_option(OPT_RECORD_HITS, record)
    b.clear()
    b.run_raw(sc, kn, 0, n, guide=guide, is_final=guide is not None)
    z, tw = b.fetch("complex128")
    rec = b.hit_record(record, max(1, kn.maxDepth), b.frame_config.channelCount)
    st = b.stats()
    b.set_option(OPT_COUNTERS, 0)
    b.set_option(OPT_CHAIN, 0)
    return z, tw, rec, st


@pytest.fixture(scope="module")
def vehicle():
    """The vehicle clutter with a guiding tree trained on the device, so that most paths hit
    it and many bounce several times."""
    mesh, _ = scenes.vehicle_mesh(seed=3, triangles=50_000)
    sc = scenes.radar_array(scenes.dihedral_rf(), channels=4, position=(0.0, 0.75, 0.0))
    pt = H.PathTracer(maxDepth=5, doGuiding=True)
    pt.configureFrame(H.FrameConfig(1, 256, 4))
    b = H.HipBackend(mesh, pt)
    pt.run_schedule(b, sc, 4, 1 << 18, 1 << 16)
    return b, sc, pt.guiding


@pytest.mark.parametrize("chain", [0, 1])
def test_tail_equals_stage_launches(vehicle, chain):
    b, sc, guide = vehicle
    kn = H.Knobs(maxDepth=5, doGuiding=True)
    n = rec = 1 << 18
    z0, tw0, r0, s0 = run(b, sc, kn, n, 0, rec, chain, guide)
    assert s0["nodes_tail"] == 0 and s0["tail_batches"][0] == 1
    print(f"paths with hits at depths 0..4: {(r0[0] >= 0).sum(0).tolist()}", flush=True)
    assert (r0[0][:, 3] >= 0).sum() > 0  # some paths bounce a fourth time
    for tail in (1, 2, 3, 4):
        z, tw, r, st = run(b, sc, kn, n, tail, rec, chain, guide)
        print(f"tail from depth {tail} (chain {chain}): tail visits {st['nodes_tail']} of {st['nodes_visited']}",
              flush=True)
        assert st["nodes_tail"] > 0 and st["tail_batches"][tail] == 1
        assert tw == tw0
        np.testing.assert_array_equal(r[0], r0[0])
        np.testing.assert_array_equal(r[1], r0[1])
        assert rel_l2(z, z0) < 1e-12
        # the same traversal work, distributed differently
        assert st["nodes_visited"] - st["nodes_shadow"] > 0


def test_tail_guided_schedule_trains_the_same_tree():
    mesh, _ = scenes.vehicle_mesh(seed=6, triangles=20_000)
    sc = scenes.radar_array(scenes.dihedral_rf(), channels=4, position=(0.0, 0.75, 0.0))
    out = []
    for tail in (0, 2):
        pt = H.PathTracer(maxDepth=4, doGuiding=True)
        pt.configureFrame(H.FrameConfig(1, 256, 4))
        b = H.HipBackend(mesh, pt)
        b.set_option(OPT_TAIL, tail)
        pt.run_schedule(b, sc, 3, 1 << 16, 1 << 18)
        out.append((pt.fetchFrame("complex128"), *pt.guiding.sampling()))
        print(f"guided schedule, tail {tail}: tree of {len(out[-1][1])} nodes", flush=True)
    np.testing.assert_array_equal(out[0][1], out[1][1])
    np.testing.assert_array_equal(out[0][2], out[1][2])
    assert rel_l2(out[1][0], out[0][0]) < 1e-12


def test_tail_mimo_cube_with_motion():
    """Cube frames splat through the event list (the tail's NEE calls it wave by wave)."""
    mesh, _ = scenes.vehicle_mesh(seed=5, triangles=20_000, range_lo=8.0, range_hi=12.0)
    u = scenes.uniforms(11, 3 * mesh.triangle_count()).reshape(-1, 3)
    vel = ((u - 0.5) * 40.0).astype(np.float32)
    b = H.HipBackend(mesh)
    b._configure(H.FrameConfig(16, 256, 12))
    b.set_velocities(vel)
    sc = scenes.mimo_array(scenes.dihedral_rf(), tx=3, rx=4, position=(0.0, 1.0, 0.0))
    kn = H.Knobs(maxDepth=4, doGuiding=False)
    z0, tw0, r0, _ = run(b, sc, kn, 1 << 16, 0, 4096)
    z, tw, r, st = run(b, sc, kn, 1 << 16, 2, 4096)
    print(f"MIMO cube: tail visits {st['nodes_tail']}", flush=True)
    assert st["nodes_tail"] > 0 and tw == tw0
    np.testing.assert_array_equal(r[0], r0[0])
    np.testing.assert_array_equal(r[1], r0[1])
    assert rel_l2(z, z0) < 1e-12


def test_adaptive_tail_depth(vehicle):
    """The default tail depth (-1): 2 at a backend's first batch, then the first depth >= 2 whose
    extension rays in the last batch, scaled to this one, were at most 2^17; a larger batch of
    the same scene moves the tail deeper, and the results are those of the stages."""
    b, sc, guide = vehicle
    kn = H.Knobs(maxDepth=5, doGuiding=True)
    n = 1 << 18
    z0, tw0, r0, _ = run(b, sc, kn, n, 0, 4096, 0, guide)
    depths = []
    for rep in range(2):
        z, tw, r, st = run(b, sc, kn, n, -1, 4096, 0, guide)
        depths.append(st["tail_batches"].index(1))
        np.testing.assert_array_equal(r[0], r0[0])
        np.testing.assert_array_equal(r[1], r0[1])
        assert rel_l2(z, z0) < 1e-12
    print(f"adaptive tail depths: {depths}", flush=True)
    assert 2 <= depths[1] < 5

