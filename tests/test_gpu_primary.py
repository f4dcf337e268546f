"""Primary visibility lists (HSR_OPT_PRIMARY) through the C ABI: the depth-0 closest hits of the
wavefront pipeline from the transmitter's lists (a cube map of direction cells around the
transmitter, each listing its triangles in order of a distance lower bound) against the BVH
traversal of the same runs and against the oracle.

Bar: hit triangle ids and shadow-ray outcomes of every recorded path identical with the lists
on and off, the frames identical bit for bit (the same hits give the same arithmetic
downstream), and against the oracle's instance of the product's precision the usual parity
bar. Scenes: the dihedral, the 200k-triangle vehicle clutter, the 1M-triangle ground of the
precision study, the 2M-triangle urban canyon with three transmitters (one list build per
transmitter), a mesh with triangles within the near distance of the transmitter, meshes of one
to four triangles; cell resolutions 4 to 80; a list capacity too small (the BVH takes over, the
next build grows the capacity); the reference's FP32 chain; a guided schedule (identical trees
and frames)."""
import os

import numpy as np
import pytest

import hussar_amd as H
from hussar_amd import scenes
from hussar_amd._lib import (OPT_CHAIN, OPT_COUNTERS, OPT_PRIMARY, OPT_PRIMARY_CAP, OPT_PRIMARY_MIN,
                             OPT_PRIMARY_RES, OPT_RECORD_HITS)
from hussar_amd.types import TriangleMesh
from conftest import rel_l2

pytestmark = [pytest.mark.gpu,
              pytest.mark.skip(reason="primary visibility lists disabled: their first device run (the vehicle "
                                      "case) faulted the GPU and the cause is not found; the library refuses "
                                      "HSR_OPT_PRIMARY = 1")]
THREADS = min(16, os.cpu_count() or 1)


def run(b, sc, kn, n, primary, record=0, res=64, cap=0, guide=None, final=False, counters=False):
    b.set_option(OPT_PRIMARY, primary)
    b.set_option(OPT_PRIMARY_MIN, 0)
    b.set_option(OPT_PRIMARY_RES, res)
    b.set_option(OPT_PRIMARY_CAP, cap)
    b.set_option(OPT_COUNTERS, 1 if counters else 0)
    b.clear()
    b.set_option(OPT_RECORD_HITS, record)
    b.run_raw(sc, kn, 0, n, guide=guide, is_final=final)
    z, tw = b.fetch("complex128")
    rec = b.hit_record(record, max(1, kn.maxDepth), len(sc.receivers())) if record else (None, None)
    return z, tw, rec[0], rec[1], b.stats()


def same_as_bvh(b, sc, kn, n, record, label, **kw):
    z0, tw0, h0, e0, _ = run(b, sc, kn, n, 0, record)
    z1, tw1, h1, e1, st = run(b, sc, kn, n, 1, record, **kw)
    print(f"primary lists {label}: builds {st['primary_builds']}, {st['primary_ms']:.3f} ms", flush=True)
    assert st["primary_builds"] == 1
    assert tw0 == tw1
    np.testing.assert_array_equal(h1, h0)
    np.testing.assert_array_equal(e1, e0)
    assert np.array_equal(z1, z0)
    return z1, h1, e1


@pytest.fixture(scope="module")
def vehicle():
    mesh, _ = scenes.vehicle_mesh(seed=1, triangles=200_000)
    b = H.HipBackend(mesh)
    fc = H.FrameConfig(1, 256, 4)
    b._configure(fc)
    sc = scenes.radar_array(scenes.dihedral_rf(), channels=4)
    return mesh, b, fc, sc


def test_vehicle_lists_equal_bvh_and_oracle(oracle, vehicle):
    mesh, b, fc, sc = vehicle
    kn = H.Knobs(maxDepth=4, doGuiding=False)
    n = 1 << 18
    z, hits, nee = same_as_bvh(b, sc, kn, n, n, "vehicle")
    om = oracle.OracleMesh.from_mesh(mesh)
    zo, two, reco, _ = oracle.render(om, sc, kn, fc, prec=oracle.PREC_MIXED, count=n, record_paths=n, threads=THREADS)
    np.testing.assert_array_equal(hits, reco[0])
    np.testing.assert_array_equal(nee, reco[1])
    assert rel_l2(z, zo / two) < 1e-5
    assert (hits[:, 0] >= 0).mean() > 0.5  # most primary rays hit the clutter


@pytest.mark.parametrize("res", [4, 17, 80])
def test_vehicle_resolutions(vehicle, res):
    mesh, b, fc, sc = vehicle
    same_as_bvh(b, sc, H.Knobs(maxDepth=2, doGuiding=False), 1 << 18, 1 << 16, f"vehicle R={res}", res=res)


def test_capacity_fallback_and_growth(vehicle):
    mesh, b, fc, sc = vehicle
    kn = H.Knobs(maxDepth=2, doGuiding=False)
    n = 1 << 17
    z0, _, h0, _, st0 = run(b, sc, kn, n, 0, n, counters=True)
    # lists far over capacity: the depth-0 rays traverse the BVH (the same node visits)
    z1, _, h1, _, st1 = run(b, sc, kn, n, 1, n, cap=1000, counters=True)
    assert st1["primary_builds"] == 1
    np.testing.assert_array_equal(h1, h0)
    assert np.array_equal(z1, z0)
    assert st1["nodes_visited"] == st0["nodes_visited"]
    # the next run rebuilds with the grown capacity and takes its depth-0 hits from the lists:
    # no node visits at depth 0, the same hits
    b.set_option(OPT_RECORD_HITS, n)
    b.clear()
    b.run_raw(sc, kn, 0, n)
    z2, _ = b.fetch("complex128")
    st2 = b.stats()
    h2, _ = b.hit_record(n, 2, 4)
    assert st2["primary_builds"] == 1
    np.testing.assert_array_equal(h2, h0)
    assert np.array_equal(z2, z0)
    assert st2["nodes_visited"] < 0.8 * st0["nodes_visited"]
    # a plain run at the same position reuses the lists
    b.clear()
    b.run_raw(sc, kn, 0, n)
    assert b.stats()["primary_builds"] == 0
    b.set_option(OPT_COUNTERS, 0)


@pytest.mark.parametrize("yaw", [-45.0, 0.0, 30.0])
def test_dihedral_lists(oracle, yaw):
    mesh = scenes.dihedral_mesh()
    b = H.HipBackend(mesh)
    fc = H.FrameConfig(1, 256, 1)
    b._configure(fc)
    sc = scenes.dihedral_pose(yaw)
    kn = H.Knobs(maxDepth=3, doGuiding=False)
    n = 1 << 18
    z, hits, nee = same_as_bvh(b, sc, kn, n, n, f"dihedral {yaw}")
    om = oracle.OracleMesh.from_mesh(mesh)
    zo, two, reco, _ = oracle.render(om, sc, kn, fc, prec=oracle.PREC_MIXED, count=n, record_paths=n, threads=THREADS)
    np.testing.assert_array_equal(hits, reco[0])
    np.testing.assert_array_equal(nee, reco[1])
    assert rel_l2(z, zo / two) < 1e-5


def test_near_triangles():
    """A small panel 5 mm in front of the transmitter and one that touches it: both are tested
    by every ray before its cell."""
    mesh, _ = scenes.vehicle_mesh(seed=3, triangles=20_000)
    sc = scenes.radar_array(scenes.dihedral_rf(), channels=2)
    p = np.asarray(sc.tx.position, np.float64)
    m = TriangleMesh(mesh.vertexBuffer.copy(), mesh.indexBuffer.copy())
    m.addQuad((p + [-0.002, -0.002, 0.005]).astype(np.float32), [0.004, 0.0, 0.0], [0.0, 0.004, 0.0])
    m.addQuad((p + [0.0, -0.01, -0.01]).astype(np.float32), [0.0, 0.02, 0.0], [0.0, 0.0, 0.02])
    b = H.HipBackend(m)
    b._configure(H.FrameConfig(1, 256, 2))
    z, hits, _ = same_as_bvh(b, sc, H.Knobs(maxDepth=3, doGuiding=False), 1 << 18, 1 << 18, "near panels")
    assert (hits[:, 0] >= 0).any()


@pytest.mark.parametrize("kind", ["triangle", "quad", "coplanar"])
def test_tiny_meshes(kind):
    m = TriangleMesh()
    if kind == "triangle":
        m = TriangleMesh(np.array([[-0.5, 0, 2], [0.5, 0, 2], [0, 1, 2]], np.float32), np.array([[0, 1, 2]], np.int32))
    else:
        m.addQuad([-0.5, 0.0, 2.0], [1.0, 0.0, 0.0], [0.0, 1.0, 0.0])
        if kind == "coplanar":
            m.addQuad([-0.25, 0.25, 2.0], [1.0, 0.0, 0.0], [0.0, 1.0, 0.0])
    b = H.HipBackend(m)
    b._configure(H.FrameConfig(1, 256, 1))
    sc = scenes.radar_array(scenes.dihedral_rf(), channels=1, position=(0.0, 0.5, 0.0))
    same_as_bvh(b, sc, H.Knobs(maxDepth=3, doGuiding=False), 1 << 16, 1 << 16, kind)


def test_reference_chain_lists(vehicle):
    mesh, b, fc, sc = vehicle
    b.set_option(OPT_CHAIN, 1)
    try:
        same_as_bvh(b, sc, H.Knobs(maxDepth=3, doGuiding=False), 1 << 17, 1 << 17, "reference chain")
    finally:
        b.set_option(OPT_CHAIN, 0)


def test_guided_schedule_lists(vehicle):
    """The bench's schedule at reduced size: every pass of a frame shares one build; trees and
    frames identical with the lists on and off."""
    mesh, b, fc, sc = vehicle
    kn = H.Knobs(maxDepth=4, doGuiding=True)
    out = []
    for primary in (0, 1):
        b.set_option(OPT_PRIMARY, primary)
        b.set_option(OPT_PRIMARY_MIN, 1 << 16)
        b.set_option(OPT_RECORD_HITS, 0)
        pt = H.PathTracer(maxDepth=4, doGuiding=True)
        pt.configureFrame(fc)
        pt._bind(b)
        pt.run_schedule(b, sc, 3, 1 << 16, 1 << 18)
        z = pt.fetchFrame("complex128")
        g = pt._guide(b)
        out.append((z, g.sampling(), b.stats()))
    (z0, s0, _), (z1, s1, st1) = out
    assert st1["primary_builds"] == 1
    assert np.array_equal(z1, z0)
    for a_, b_ in zip(s0, s1):
        np.testing.assert_array_equal(a_, b_)
    b.set_option(OPT_PRIMARY_MIN, 0)


def test_precision_ground_lists():
    print("primary lists: building the 1M-triangle precision scene", flush=True)
    mesh = scenes.precision_mesh(seed=2)
    b = H.HipBackend(mesh)
    b._configure(H.FrameConfig(1, 512, 4))
    sc = scenes.radar_array(scenes.dihedral_rf(), channels=4, position=(0.0, 0.0, 0.0))
    same_as_bvh(b, sc, H.Knobs(maxDepth=4, doGuiding=False), 1 << 18, 1 << 17, "precision ground")


def test_urban_mimo_lists():
    print("primary lists: building the 2M-triangle urban canyon", flush=True)
    mesh, vel = scenes.urban_mesh(seed=4)
    b = H.HipBackend(mesh)
    b.set_velocities(vel)
    b._configure(H.FrameConfig(8, 256, 12))
    rf = scenes.dihedral_rf()
    sc = scenes.mimo_array(rf, tx=3, rx=4)
    kn = H.Knobs(maxDepth=5, doGuiding=False)
    z0, tw0, h0, e0, _ = run(b, sc, kn, 1 << 16, 0, 1 << 14)
    z1, tw1, h1, e1, st = run(b, sc, kn, 1 << 16, 1, 1 << 14)
    print(f"primary lists urban: builds {st['primary_builds']}, {st['primary_ms']:.3f} ms", flush=True)
    assert st["primary_builds"] == 3  # one per transmitter
    np.testing.assert_array_equal(h1, h0)
    np.testing.assert_array_equal(e1, e0)
    assert rel_l2(z1, z0) < 1e-12  # the cube's event lists may overflow into atomics in another order
