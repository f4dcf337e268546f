"""Geometrical optics off (useGeometricalOptics = false: a uniform sphere sample per bounce,
libhussar/include/hussar/integrators/path.h:277-279 and 413-427, emitter.h:53-82) on the GPU
through the C ABI, against the oracle.

Bar, as for the default branch: hit triangle ids and shadow-ray outcomes bit-exact against the
oracle instance with the same precision model, the frame within 1e-5 relative L2 of it and
within 1e-3 of the FP64 oracle, the peak range bin equal. Scenes: the two-plate corner of the
oracle's brute-force pin (bistatic, depths 1-3), the dihedral at two poses with the plane
filter, the 200k-triangle vehicle clutter with a 4-receiver radar inside it, moving clutter
with a 2 Tx x 4 Rx radar and 16 chirps, and a guided pass with a sampling tree shared with the
oracle. Both pipelines. Also: the reference's FP32 chain
(bit-exact hits against the oracle's dual instance); the fused tail from depth 1 (identical
records, the frame equal up to summation order); the knobs the reference asserts against
refused with HSR_ERR_INVALID."""
import os

import numpy as np
import pytest

import hussar_amd as H
from hussar_amd import scenes
from hussar_amd._lib import HsrError, OPT_CHAIN, OPT_PIPELINE, OPT_RECORD_HITS, OPT_TAIL
from conftest import rel_l2

pytestmark = pytest.mark.gpu
THREADS = min(16, os.cpu_count() or 1)
TX, RX = (-0.3, 0.0, 0.0), (0.3, 0.0, 0.0)


def corner_mesh():
    m = H.TriangleMesh()
    m.addQuad([-3.0, -3.0, 3.0], [6.0, 0.0, 0.0], [0.0, 6.0, 0.0])
    m.addQuad([3.0, -3.0, 0.5], [0.0, 0.0, 2.5], [0.0, 6.0, 0.0])
    return m


def corner_scene():
    eye = np.eye(3)
    return H.Scene(rfConfig=scenes.dihedral_rf(), tx=H.NFAntenna(TX, eye), rx=H.NFAntenna(RX, eye))


def knobs(depth, **kw):
    return H.Knobs(**{"maxDepth": depth, "doGuiding": False, "useGeometricalOptics": False,
                      "filteringSphere": False, **kw})


def make_backend(mesh, fc, pipeline):
    b = H.HipBackend(mesh)
    b.set_option(OPT_PIPELINE, pipeline)
    b._configure(fc)
    return b


def gpu_run(b, sc, kn, n, record, guide=None, channels=None):
    b.clear()
    b.set_option(OPT_RECORD_HITS, record)
    b.run_raw(sc, kn, 0, n, guide=guide)
    z, tw = b.fetch("complex128")
    return z, tw, b.hit_record(record, max(1, kn.maxDepth), channels or len(sc.receivers()))


def check(oracle, om, b, sc, kn, fc, n, guide=None, oguide=None, prec="PREC_MIXED", fp64=True):
    print(f"geometrical optics off: depth {kn.maxDepth}, {n} paths", flush=True)
    z, tw, rec = gpu_run(b, sc, kn, n, n, guide, channels=fc.channelCount)
    zo, two, reco, _ = oracle.render(om, sc, kn, fc, prec=getattr(oracle, prec), count=n, record_paths=n,
                                     guide=oguide, threads=THREADS)
    z = z.reshape(zo.shape)
    assert tw == two == n
    np.testing.assert_array_equal(rec[0], reco[0])
    np.testing.assert_array_equal(rec[1], reco[1])
    zo = zo / two
    assert np.abs(zo).max() > 0
    assert rel_l2(z, zo) < 1e-5
    c = z.shape[-1]
    assert np.abs(z).reshape(-1, c).argmax(0).tolist() == np.abs(zo).reshape(-1, c).argmax(0).tolist()
    if fp64:
        zf, twf, _, _ = oracle.render(om, sc, kn, fc, prec=oracle.PREC_FP64, count=n, guide=oguide,
                                      threads=THREADS)
        assert rel_l2(z, zf / twf) < 1e-3
    return z, rec


@pytest.fixture(scope="module", params=[0, 1], ids=["fused", "wavefront"])
def corner(request):
    mesh = corner_mesh()
    fc = H.FrameConfig(1, 256, 1)
    return mesh, make_backend(mesh, fc, request.param), fc


@pytest.mark.parametrize("depth", [1, 2, 3])
def test_corner_parity(oracle, corner, depth):
    mesh, b, fc = corner
    om = oracle.OracleMesh.from_mesh(mesh)
    z, rec = check(oracle, om, b, corner_scene(), knobs(depth, doFiltering=False), fc, 1 << 18)
    if depth >= 2:
        assert (rec[0][:, 1] >= 0).sum() > 1000  # second bounces from sampled directions
        assert (rec[1][:, 1, 0] == 3).sum() > 1000


@pytest.mark.parametrize("yaw", [0.0, 30.0])
def test_dihedral_parity(oracle, yaw):
    mesh = scenes.dihedral_mesh()
    fc = H.FrameConfig(1, 256, 1)
    om = oracle.OracleMesh.from_mesh(mesh)
    for pipeline in (0, 1):
        b = make_backend(mesh, fc, pipeline)
        check(oracle, om, b, scenes.dihedral_pose(yaw), knobs(3), fc, 1 << 18)


def test_vehicle_parity(oracle):
    mesh, centre = scenes.vehicle_mesh(seed=1, triangles=200_000)
    fc = H.FrameConfig(1, 256, 4)
    # the radar inside the clutter: sampled bounces meet clutter in every direction
    sc = scenes.radar_array(scenes.dihedral_rf(), channels=4, position=(float(centre[0]), 0.75, float(centre[2])))
    om = oracle.OracleMesh.from_mesh(mesh)
    b = make_backend(mesh, fc, 1)
    z, rec = check(oracle, om, b, sc, knobs(3), fc, 1 << 16)
    hits = [(rec[0][:, k] >= 0).sum() for k in range(3)]
    print(f"vehicle without geometrical optics: hits per depth {hits}", flush=True)
    assert hits[2] > 1000  # sampled bounces inside the clutter
    assert np.all(np.abs(z).max(axis=(0, 1)) > 0)


def test_moving_mimo_cube_parity(oracle):
    """Moving clutter (per-triangle velocities up to 20 m/s, the platform at 8 m/s), 2 Tx x 4 Rx,
    16 chirps: the Doppler terms of the sampled bounces' legs."""
    mesh, centre = scenes.vehicle_mesh(seed=5, triangles=20_000, range_lo=8.0, range_hi=12.0)
    u = scenes.uniforms(11, 3 * mesh.triangle_count()).reshape(-1, 3)
    vel = ((u - 0.5) * 40.0).astype(np.float32)
    rf = scenes.dihedral_rf()
    sc = scenes.radar_array(rf, channels=4, position=(float(centre[0]), 0.75, float(centre[2])))
    tx1 = H.NFAntenna(np.asarray(sc.tx.position) + np.float32([0.0078, 0.0, 0.0]), sc.tx.rotation)
    mimo = H.Scene(rfConfig=rf, tx=sc.tx, rx=sc.rx, rxs=sc.rxs, txs=[sc.tx, tx1])
    mimo.velocity = (0.0, 0.0, 8.0)
    fc = H.FrameConfig(16, 64, 8)
    b = make_backend(mesh, fc, 1)
    b.set_velocities(vel)
    om = oracle.OracleMesh.from_mesh(mesh)
    om.set_velocities(vel)
    z, rec = check(oracle, om, b, mimo, knobs(3), fc, 1 << 15, fp64=False)
    assert (rec[0][:, 1] >= 0).sum() > 1000
    assert np.count_nonzero(np.abs(z).sum(axis=(1, 2)) > 0) > 8  # energy spread over Doppler


def test_guided_pass_parity(oracle):
    """A sampling tree trained on the oracle with geometrical optics off, shared by both sides."""
    mesh = corner_mesh()
    fc = H.FrameConfig(1, 256, 1)
    om = oracle.OracleMesh.from_mesh(mesh)
    sc = corner_scene()
    kn = knobs(3, doGuiding=True)
    og = oracle.OracleGuide()
    for _ in range(2):
        oracle.render(om, sc, kn, fc, prec=oracle.PREC_MIXED, count=1 << 16, guide=og, threads=THREADS)
        og.step()
    child, p0, _ = og.get(0)
    assert len(child) > 3
    b = make_backend(mesh, fc, 1)
    g = H.Guide(b)
    g.set_sampling(child, p0)
    check(oracle, om, b, sc, kn, fc, 1 << 16, guide=g, oguide=og, fp64=False)


def test_reference_chain(oracle):
    mesh = corner_mesh()
    fc = H.FrameConfig(1, 256, 1)
    om = oracle.OracleMesh.from_mesh(mesh)
    b = make_backend(mesh, fc, 1)
    b.set_option(OPT_CHAIN, 1)
    try:
        check(oracle, om, b, corner_scene(), knobs(3, doFiltering=False), fc, 1 << 18, prec="PREC_DUAL")
    finally:
        b.set_option(OPT_CHAIN, 0)


def test_tail_from_depth_one(oracle):
    """The fused tail carries every path from depth 1: the same records, the frame equal up to
    summation order."""
    mesh = corner_mesh()
    fc = H.FrameConfig(1, 256, 1)
    b = make_backend(mesh, fc, 1)
    sc, kn, n = corner_scene(), knobs(3, doFiltering=False), 1 << 18
    out = []
    for tail in (0, 1):
        b.set_option(OPT_TAIL, tail)
        out.append(gpu_run(b, sc, kn, n, n))
    b.set_option(OPT_TAIL, -1)
    (z0, _, r0), (z1, _, r1) = out
    np.testing.assert_array_equal(r1[0], r0[0])
    np.testing.assert_array_equal(r1[1], r0[1])
    assert rel_l2(z1, z0) < 1e-12


@pytest.mark.parametrize("kw", [{}, {"correctPhase": True, "filteringSphere": True}, {"maxDepth": 4}])
def test_asserted_knobs_refused(kw):
    mesh = corner_mesh()
    b = make_backend(mesh, H.FrameConfig(1, 256, 1), 1)
    kn = knobs(3, **kw) if kw else H.Knobs(maxDepth=3, doGuiding=False, useGeometricalOptics=False)
    with pytest.raises(HsrError, match="INVALID"):
        b.run_raw(corner_scene(), kn, 0, 16)
