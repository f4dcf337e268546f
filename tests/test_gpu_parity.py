"""GPU parity of the gfx950 backend against the oracle, through the C ABI.

Bar: hit triangle ids and shadow-ray outcomes bit-exact against the oracle instance with the
same precision model (FP32 traversal, FP64 geometry chain / phase / sums); the frame within
1e-5 relative L2 of that instance and within 1e-3 of the FP64 oracle;
the peak range bin identical.
"""
import ctypes as C
import json
import os

import numpy as np
import pytest

import hussar_amd as H
from hussar_amd import scenes
from hussar_amd._lib import (HsrError, OPT_COUNTERS, OPT_CUBE_EVENTS, OPT_MARKERS, OPT_PATHS_PER_LAUNCH,
                             OPT_PIPELINE, OPT_RECORD_HITS, OPT_SHADOW_SPLAT, OPT_TAIL)
from conftest import rel_l2

pytestmark = pytest.mark.gpu

THREADS = min(16, os.cpu_count() or 1)
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


PIPELINES = [0, 1]  # fused path kernel, wavefront stages


def make_backend(mesh, fc, pipeline, integrator=None):
    b = H.HipBackend(mesh, integrator)
    b.set_option(OPT_PIPELINE, pipeline)
    if fc is not None and integrator is None:
        b._configure(fc)
    return b


def gpu_run(backend, scene, knobs, n, offset=0, record=0, guide=None, is_final=False, clear=True):
    if clear:
        backend.clear()
    backend.set_option(OPT_RECORD_HITS, record)
    backend.run_raw(scene, knobs, offset, n, guide=guide, is_final=is_final)
    z, tw = backend.fetch("complex128")
    rec = backend.hit_record(record, max(1, knobs.maxDepth), len(scene.receivers())) if record else None
    return z, tw, rec


def check_parity(oracle, om, backend, scene, knobs, fc, n, record, guide=None, oguide=None, fp64=True,
                 tol_mixed=1e-5, tol_fp64=1e-3):
    z, tw, rec = gpu_run(backend, scene, knobs, n, record=record, guide=guide)
    zo, two, reco, _ = oracle.render(om, scene, knobs, fc, prec=oracle.PREC_MIXED, count=n, record_paths=record,
                                     guide=oguide, threads=THREADS)
    assert tw == two == n
    np.testing.assert_array_equal(rec[0], reco[0])
    np.testing.assert_array_equal(rec[1], reco[1])
    zo = zo / two
    assert np.abs(zo).max() > 0
    assert rel_l2(z, zo) < tol_mixed
    assert np.abs(z).reshape(-1, z.shape[-1]).argmax(0).tolist() == np.abs(zo).reshape(-1, z.shape[-1]).argmax(0).tolist()
    if fp64:
        zf, twf, _, _ = oracle.render(om, scene, knobs, fc, prec=oracle.PREC_FP64, count=n, guide=oguide,
                                      threads=THREADS)
        assert rel_l2(z, zf / twf) < tol_fp64
    return z


@pytest.fixture(scope="module", params=PIPELINES, ids=["fused", "wavefront"])
def dihedral(request):
    mesh = scenes.dihedral_mesh()
    fc = H.FrameConfig(1, 256, 1)
    return mesh, make_backend(mesh, fc, request.param), fc


@pytest.mark.parametrize("yaw", [-45.0, -17.0, 0.0, 30.0, 45.0])
def test_dihedral_pose_parity(oracle, dihedral, yaw):
    mesh, b, fc = dihedral
    om = oracle.OracleMesh.from_mesh(mesh)
    kn = H.Knobs(maxDepth=3, doGuiding=False)
    check_parity(oracle, om, b, scenes.dihedral_pose(yaw), kn, fc, 1 << 18, record=1 << 18)


def test_dihedral_full_pose_size_vs_fp64(oracle, dihedral):
    """The dihedral example's per-pose size: 2^20 paths, depth 3, vs the FP64 oracle."""
    mesh, b, fc = dihedral
    om = oracle.OracleMesh.from_mesh(mesh)
    kn = H.Knobs(maxDepth=3, doGuiding=False)
    sc = scenes.dihedral_pose(7.0)
    z, tw, _ = gpu_run(b, sc, kn, 1 << 20)
    zf, twf, _, _ = oracle.render(om, sc, kn, fc, prec=oracle.PREC_FP64, count=1 << 20, threads=THREADS)
    assert rel_l2(z, zf / twf) < 1e-5
    assert np.abs(z).argmax() == np.abs(zf).argmax() == 15


@pytest.fixture(scope="module")
def vehicle_mesh():
    return scenes.vehicle_mesh(seed=1, triangles=200_000)


@pytest.fixture(scope="module", params=PIPELINES, ids=["fused", "wavefront"])
def vehicle(request, vehicle_mesh):
    mesh, centre = vehicle_mesh
    fc = H.FrameConfig(1, 256, 4)
    b = make_backend(mesh, fc, request.param)
    rf = scenes.dihedral_rf()
    sc = scenes.radar_array(rf, channels=4)
    return mesh, b, fc, sc, centre


def test_vehicle_4rx_parity(oracle, vehicle):
    """200k-triangle clutter, 1 Tx x 4 Rx, depth 4: BVH traversal, multi-receiver NEE."""
    mesh, b, fc, sc, _ = vehicle
    om = oracle.OracleMesh.from_mesh(mesh)
    kn = H.Knobs(maxDepth=4, doGuiding=False)
    z = check_parity(oracle, om, b, sc, kn, fc, 1 << 17, record=1 << 17)
    # all four channels carry signal
    assert np.all(np.abs(z).max(axis=(0, 1)) > 0)


def test_vehicle_guided_final_pass_parity(oracle, vehicle):
    """A trained sampling tree shared by both sides: guiding lookup on the GPU warps the same
    primary samples as the oracle (bit-exact hits), and the guided frame matches."""
    mesh, b, fc, sc, _ = vehicle
    om = oracle.OracleMesh.from_mesh(mesh)
    kn = H.Knobs(maxDepth=4, doGuiding=True)
    og = oracle.OracleGuide()
    for _ in range(2):
        oracle.render(om, sc, kn, fc, prec=oracle.PREC_MIXED, count=1 << 16, guide=og, threads=THREADS)
        og.step()
    child, p0, _ = og.get(0)
    assert len(child) > 3
    g = H.Guide(b)
    g.set_sampling(child, p0)
    check_parity(oracle, om, b, sc, kn, fc, 1 << 16, record=1 << 16, guide=g, oguide=og, fp64=False,
                 tol_mixed=1e-5)


def test_guiding_record_and_step_match_oracle(oracle, dihedral):
    """Training statistics recorded on the GPU equal the oracle's (FP64 sums in another order),
    and the host step then builds identical trees."""
    mesh, b, fc = dihedral
    om = oracle.OracleMesh.from_mesh(mesh)
    kn = H.Knobs(maxDepth=3, doGuiding=True)
    sc = scenes.dihedral_pose(0.0)
    g = H.Guide(b)
    og = oracle.OracleGuide()
    for it in range(3):
        gpu_run(b, sc, kn, 1 << 16, guide=g)
        oracle.render(om, sc, kn, fc, prec=oracle.PREC_MIXED, count=1 << 16, guide=og, threads=THREADS)
        gc, gs = g.training()
        oc, _, os_ = og.get(1)
        np.testing.assert_array_equal(gc, oc)
        np.testing.assert_allclose(gs, os_, rtol=1e-9, atol=1e-12 * max(1.0, np.abs(os_).max()))
        assert os_.sum() > 0
        g.step()
        og.step()
        sc_, sp_ = g.sampling()
        oc_, op_, _ = og.get(0)
        np.testing.assert_array_equal(sc_, oc_)
        np.testing.assert_allclose(sp_, op_, rtol=1e-6)


def test_path_tracer_run_doubling_schedule(oracle, dihedral):
    """PathTracer::run with guiding (path.h:171-224): same final totalWeight and frame."""
    mesh, b, fc = dihedral
    om = oracle.OracleMesh.from_mesh(mesh)
    pt = H.PathTracer(maxDepth=3, doGuiding=True)
    pt.configureFrame(fc)
    b2 = make_backend(mesh, None, b.pipeline if hasattr(b, "pipeline") else 0, pt)
    sc = scenes.dihedral_pose(10.0)
    pt.run(b2, sc, 1 << 20)
    z = pt.fetchFrame("complex128")
    _, tw = b2.fetch()
    # the training iterations return once their statistics are downloaded; their frame
    # reduction's timing is settled by the next iteration, the last one's here
    st = b2.stats()
    assert st["total_ms"] >= st["trace_ms"] > 0 and st["reduce_ms"] >= 0
    zo, two = oracle.path_tracer_run(om, sc, pt, fc, 1 << 20, prec=oracle.PREC_MIXED, guide=oracle.OracleGuide(),
                                     threads=THREADS)
    assert tw == two == 540672
    assert rel_l2(z, zo / two) < 1e-5


def test_sample_offset_and_launch_split_are_additive(dihedral):
    mesh, b, fc = dihedral
    kn = H.Knobs(maxDepth=3, doGuiding=False)
    sc = scenes.dihedral_pose(0.0)
    whole, tw, _ = gpu_run(b, sc, kn, 300_000)
    b.clear()
    b.run_raw(sc, kn, 0, 123_457)
    b.run_raw(sc, kn, 123_457, 300_000 - 123_457)
    parts, tw2 = b.fetch("complex128")
    assert tw == tw2
    np.testing.assert_allclose(parts, whole, rtol=1e-12, atol=1e-15 * np.abs(whole).max())
    b.set_option(OPT_PATHS_PER_LAUNCH, 50_000)
    try:
        split, _, _ = gpu_run(b, sc, kn, 300_000)
    finally:
        b.set_option(OPT_PATHS_PER_LAUNCH, 1 << 24)
    np.testing.assert_allclose(split, whole, rtol=1e-12, atol=1e-15 * np.abs(whole).max())


def test_repeatable(dihedral):
    mesh, b, fc = dihedral
    kn = H.Knobs(maxDepth=3, doGuiding=False)
    sc = scenes.dihedral_pose(20.0)
    a, _, _ = gpu_run(b, sc, kn, 1 << 20)
    c, _, _ = gpu_run(b, sc, kn, 1 << 20)
    np.testing.assert_allclose(a, c, rtol=1e-13, atol=1e-15 * np.abs(a).max())


@pytest.mark.parametrize("pipeline", PIPELINES, ids=["fused", "wavefront"])
@pytest.mark.parametrize("case", ["direct", "po_diffraction", "plane_filter", "no_filter", "clang_order",
                                  "small_frame", "odd_budget", "correct_phase", "correct_phase_cut"])
def test_knob_and_shape_variants(oracle, case, pipeline):
    mesh = scenes.dihedral_mesh()
    sc = scenes.dihedral_pose(5.0)
    kn = H.Knobs(maxDepth=3, doGuiding=False)
    fc = H.FrameConfig(1, 256, 1)
    n = 1 << 16
    if case == "direct":
        sc = H.Scene(rfConfig=sc.rfConfig, tx=sc.tx,
                     rx=H.NFAntenna(sc.tx.position + np.float32([0, 0.2, 0]), sc.tx.rotation))
        kn.onlyIndirect = False
        kn.produceDebugImage = True
    elif case == "po_diffraction":
        kn.poDiffraction = True
    elif case == "plane_filter":
        kn.filteringSphere = False
    elif case == "no_filter":
        kn.doFiltering = False
    elif case == "clang_order":
        kn.primaryOrder = 1
    elif case == "small_frame":
        fc = H.FrameConfig(1, 24, 1)
    elif case == "odd_budget":
        n = 100_003
    elif case.startswith("correct_phase"):
        # path.h:316-329, 360-368; the corrected values are real for these antenna patterns, so
        # the debug image keeps them from the splat skip rule; "cut": a radius of 20 wavelengths
        # cuts the contributions whose reflected ray passes the receiver further away
        kn.correctPhase = True
        kn.produceDebugImage = True
        kn.filteringRadius = 20.0 if case == "correct_phase_cut" else 1e4
    b = make_backend(mesh, fc, pipeline)
    om = oracle.OracleMesh.from_mesh(mesh)
    check_parity(oracle, om, b, sc, kn, fc, n, record=min(n, 1 << 16), fp64=True)


def few_triangle_mesh(kind):
    """Meshes of one to four triangles (a plate 2 m ahead of a radar at the origin looking down
    +z): whose best split by SAH is a single leaf, and, for the coplanar quads, exact t ties
    that the smaller triangle id wins."""
    if kind == "triangle":
        return H.TriangleMesh(np.array([[-0.3, -0.3, 2.0], [0.3, -0.3, 2.0], [-0.3, 0.3, 2.0]], np.float32),
                              np.array([[0, 1, 2]], np.int32))
    m = scenes.plate_mesh(size=0.5, distance=2.0)
    if kind == "two_coplanar_quads":
        m.addQuad([-0.1, -0.1, 2.0], [0.5, 0.0, 0.0], [0.0, 0.5, 0.0])
    return m


@pytest.mark.parametrize("kind", ["triangle", "quad", "two_coplanar_quads"])
@pytest.mark.parametrize("pipeline", PIPELINES, ids=["fused", "wavefront"])
def test_few_triangle_meshes_parity(oracle, kind, pipeline):
    mesh = few_triangle_mesh(kind)
    fc = H.FrameConfig(1, 256, 1)
    sc = scenes.radar_array(scenes.dihedral_rf(), channels=1, position=(0.0, 0.0, 0.0))
    b = make_backend(mesh, fc, pipeline)
    om = oracle.OracleMesh.from_mesh(mesh)
    n = 1 << 18
    print("few-triangle mesh", kind, "pipeline", pipeline, flush=True)
    z = check_parity(oracle, om, b, sc, H.Knobs(maxDepth=3, doGuiding=False), fc, n, record=n)
    assert np.abs(z).argmax() > 0


@pytest.mark.parametrize("pipeline", PIPELINES, ids=["fused", "wavefront"])
def test_empty_mesh_and_zero_depth(pipeline):
    fc = H.FrameConfig(1, 256, 1)
    b = make_backend(H.TriangleMesh(), fc, pipeline)
    sc = scenes.dihedral_pose(0.0)
    z, tw, rec = gpu_run(b, sc, H.Knobs(maxDepth=3, doGuiding=False), 4096, record=4096)
    assert tw == 4096 and np.all(z == 0) and np.all(rec[0] == -1)
    z, tw, _ = gpu_run(b, sc, H.Knobs(maxDepth=0, doGuiding=False), 4096)
    assert np.all(z == 0)


def test_errors_and_interrupt(dihedral):
    mesh, b, fc = dihedral
    sc = scenes.dihedral_pose(0.0)
    # geometrical optics off with the sphere filter (path.h:344), correctPhase (path.h:317) or
    # more bounces than the Halton dimensions (halton.h:58-67) allow
    for kw in ({}, {"filteringSphere": False, "correctPhase": True}, {"filteringSphere": False, "maxDepth": 4}):
        with pytest.raises(HsrError, match="INVALID"):
            b.run_raw(sc, H.Knobs(**{"useGeometricalOptics": False, "doGuiding": False, "maxDepth": 3, **kw}), 0, 10)
    with pytest.raises(HsrError, match="INVALID"):  # path.h:318 asserts the sphere filter
        b.run_raw(sc, H.Knobs(correctPhase=True, filteringSphere=False, doGuiding=False), 0, 10)
    with pytest.raises(HsrError, match="INVALID"):
        b.run_raw(sc, H.Knobs(doGuiding=True), 0, 10)  # guiding without a guide
    flag = C.c_bool(True)
    with pytest.raises(HsrError, match="INTERRUPTED"):
        b.run_raw(sc, H.Knobs(maxDepth=3, doGuiding=False), 0, 1 << 20, interruptFlag=flag)
    bad = H.Scene(rfConfig=sc.rfConfig, tx=sc.tx, rx=sc.rx, rxs=[sc.rx, sc.rx])
    with pytest.raises(HsrError, match="INVALID"):
        b.run_raw(bad, H.Knobs(maxDepth=3, doGuiding=False), 0, 10)


def test_counters_are_deterministic(dihedral):
    mesh, b, fc = dihedral
    sc = scenes.dihedral_pose(0.0)
    kn = H.Knobs(maxDepth=3, doGuiding=False)
    b.set_option(OPT_COUNTERS, 1)
    try:
        gpu_run(b, sc, kn, 1 << 18)
        s1 = b.stats()
        gpu_run(b, sc, kn, 1 << 18)
        s2 = b.stats()
    finally:
        b.set_option(OPT_COUNTERS, 0)
    for k in ("nodes_visited", "tris_tested", "rays_closest", "rays_shadow", "events"):
        assert s1[k] == s2[k] > 0, k


def test_traversal_visits_per_ray_stay_small():
    """Node visits per ray of the closest-hit and packet traversals on the vehicle clutter
    (deterministic counters, stage launches). A box test that turned too conservative would
    still return the right hits while opening most of the tree: this bounds the visits, at a few
    times what the 50k-triangle tree needs (1.1 per closest-hit ray, most of the unguided primary
    rays missing the clutter, and 12 per shadow packet)."""
    mesh, _ = scenes.vehicle_mesh(seed=3, triangles=50_000)
    b = H.HipBackend(mesh)
    b._configure(H.FrameConfig(1, 256, 4))
    sc = scenes.radar_array(scenes.dihedral_rf(), channels=4, position=(0.0, 0.75, 0.0))
    b.set_option(OPT_TAIL, 0)
    b.set_option(OPT_COUNTERS, 1)
    try:
        gpu_run(b, sc, H.Knobs(maxDepth=4, doGuiding=False), 1 << 18)
        st = b.stats()
    finally:
        b.set_option(OPT_COUNTERS, 0)
        b.set_option(OPT_TAIL, -1)
    closest = (st["nodes_visited"] - st["nodes_shadow"]) / st["rays_closest"]
    shadow = st["nodes_shadow"] / (st["rays_shadow"] / 4)
    print(f"node visits per closest-hit ray {closest:.2f}, per shadow packet {shadow:.2f}", flush=True)
    assert 1.0 < closest < 5.0
    assert 1.0 < shadow < 30.0


def test_markers_leave_results_alone(dihedral):
    """ROCTx ranges around passes, batches and stage launches (HSR_OPT_MARKERS) only mark the
    host's work: the same run gives the same frame and records with and without them."""
    mesh, b, fc = dihedral
    sc = scenes.dihedral_pose(10.0)
    kn = H.Knobs(maxDepth=3, doGuiding=False)
    z0, tw0, r0 = gpu_run(b, sc, kn, 1 << 16, record=4096)
    b.set_option(OPT_MARKERS, 1)
    try:
        z1, tw1, r1 = gpu_run(b, sc, kn, 1 << 16, record=4096)
    finally:
        b.set_option(OPT_MARKERS, 0)
    assert tw0 == tw1
    np.testing.assert_array_equal(r0[0], r1[0])
    np.testing.assert_array_equal(r0[1], r1[1])
    assert rel_l2(z1, z0) < 1e-12
    with pytest.raises(HsrError, match="INVALID"):
        b.set_option(OPT_MARKERS, 2)


@pytest.mark.parametrize("scene", ["vehicle", "dihedral_direct", "cube", "cube_overflow"])
def test_shadow_splat_equals_gather(scene):
    """Batches without guiding records splat in their shadow stages (HSR_OPT_SHADOW_SPLAT, the
    default) instead of a gather stage after each: the same hits and NEE records, and the same
    frame up to the order of its sums; also with the direct paths traced (depth-0 NEE), and for
    a moving 16-chirp cube, whose events go to the event lists (and, past a list's capacity, to
    the cube with device atomics)."""
    vel = None
    if scene.startswith("cube"):
        mesh, _ = scenes.vehicle_mesh(seed=5, triangles=20_000, range_lo=8.0, range_hi=12.0)
        u = scenes.uniforms(11, 3 * mesh.triangle_count()).reshape(-1, 3)
        vel = ((u - 0.5) * 40.0).astype(np.float32)
        sc = scenes.radar_array(scenes.dihedral_rf(), channels=4)
        sc.velocity = (0.0, 0.0, 8.0)
        fc = H.FrameConfig(16, 64, 4)
        kn = H.Knobs(maxDepth=3, doGuiding=False)
    elif scene == "vehicle":
        mesh, _ = scenes.vehicle_mesh(seed=5, triangles=50_000)
        sc = scenes.radar_array(scenes.dihedral_rf(), channels=4)
        fc = H.FrameConfig(1, 256, 4)
        kn = H.Knobs(maxDepth=4, doGuiding=False)
    else:
        # the direct path needs the receiver away from the transmitter (a path of length 0 has
        # an infinite Green's function, as in the reference)
        mesh = scenes.dihedral_mesh()
        sc = scenes.dihedral_pose(20.0)
        sc = H.Scene(rfConfig=sc.rfConfig, tx=sc.tx,
                     rx=H.NFAntenna(sc.tx.position + np.float32([0, 0.2, 0]), sc.tx.rotation))
        fc = H.FrameConfig(1, 256, 1)
        kn = H.Knobs(maxDepth=3, doGuiding=False, onlyIndirect=False)
    b = make_backend(mesh, fc, 1)
    if vel is not None:
        b.set_velocities(vel)
    if scene == "cube_overflow":
        b.set_option(OPT_CUBE_EVENTS, 4096)
    n = 1 << 18
    out = []
    for fused in (0, 1):
        b.set_option(OPT_SHADOW_SPLAT, fused)
        out.append(gpu_run(b, sc, kn, n, record=n))
    (z0, tw0, r0), (z1, tw1, r1) = out
    assert tw0 == tw1
    np.testing.assert_array_equal(r1[0], r0[0])
    np.testing.assert_array_equal(r1[1], r0[1])
    assert (r1[1] == 3).sum() > 100  # visible and splatted contributions were recorded
    assert rel_l2(z1, z0) < 1e-12
    with pytest.raises(HsrError, match="INVALID"):
        b.set_option(OPT_SHADOW_SPLAT, 2)


@pytest.mark.parametrize("size_cm", [5, 15])
def test_reference_dihedral_sweep_fixture_on_gpu(size_cm):
    """The reference's own simulated sweep (evaluation/simulation/dihedral-*cm): same peak bin
    and peak magnitude within 10 % at in-beam poses."""
    from test_oracle_integrator import mesh_1dihedral, scene_1dihedral
    fx = json.load(open(os.path.join(GOLDEN, "dihedral_fixture.json")))["sizes_cm"][str(size_cm)]
    b = make_backend(mesh_1dihedral(size_cm / 100), H.FrameConfig(1, 256, 1), 1)
    kn = H.Knobs(maxDepth=10, doGuiding=False)
    for yaw in np.arange(-30.0, 30.01, 2.5):
        i = int(round((yaw + 75.0) / 0.25))
        z, tw, _ = gpu_run(b, scene_1dihedral(float(yaw)), kn, 1 << 22)
        a = np.abs(z[0, :, 0])
        assert int(a.argmax()) == fx["peak_bin"][i], yaw
        assert a.max() == pytest.approx(fx["peak_abs"][i], rel=0.10), yaw
