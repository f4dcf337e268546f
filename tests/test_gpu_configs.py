"""GPU parity on the stated scenes of the precision study, the urban MIMO frame and the serving mix, through
the C ABI, against the oracle on the same Halton indices (bounded path samples; the frame
itself is what the benchmark times at full path counts).

Precision study: plates and trihedrals at 5, 50 and 150 m over a 1M-triangle ground (0.05 m
RMS heights), 1 Tx x 4 Rx, 512 samples, depth 4, 2^16 paths per IF-phase variant:
  * hit ids and shadow outcomes bit-exact against the oracle instance of the product's
    precision (FP32 traversal, FP64 geometry chain) on every recorded path;
  * FP64 and double-float phase: frame within 1e-5 relative L2 of the FP64 oracle (bar 1e-3),
    phase within 1e-5 rad at the six targets' range bins; FP32 phase (the reference's):
    within 1e-2 and 0.05 rad (the study's point: it misses the 1e-3 bar);
  * the peak range bin of every channel equal to the FP64 oracle's in every variant.
Urban MIMO frame: a 2M-triangle street canyon with four moving vehicle blocks, 3 Tx x 4 Rx,
64 chirps x 256 samples x 12 channels, depth 5, guiding on (a tree trained on the device with
four passes of 2^18 paths, then shared with the oracle), 2^14 final-pass paths: hits and shadow
outcomes bit-exact against the mixed oracle, the cube within 1e-5 of it and within 1e-3 of the
FP64 oracle, the peak cell identical.
Serving mix: the 10k / 300k / 2M-triangle scenes with resident guiding trees, eight requests of
the mix served together, each equal to itself served alone and within 1e-5 of the oracle."""
import os

import numpy as np
import pytest

import hussar_amd as H
from hussar_amd import scenes
from hussar_amd._lib import OPT_PHASE_MODE, OPT_RECORD_HITS
from conftest import rel_l2

pytestmark = pytest.mark.gpu
THREADS = min(16, os.cpu_count() or 1)


def target_bins(points, rf, N):
    """N·frac((2|p|/c)·S/f_s) of single-bounce returns from points seen from the origin."""
    c = float(np.float32(299792458.0))
    return sorted({int(np.floor(N * (((2.0 * np.linalg.norm(p) / c) * rf.freqSlope / rf.adcRate) % 1.0) + 0.5)) % N
                   for p in points})


@pytest.fixture(scope="module")
def precision():
    print("precision study: building the 1M-triangle scene", flush=True)
    mesh = scenes.precision_mesh(seed=2)
    assert mesh.triangle_count() == 1_000_024
    b = H.HipBackend(mesh)
    fc = H.FrameConfig(1, 512, 4)
    b._configure(fc)
    rf = scenes.dihedral_rf()
    sc = scenes.radar_array(rf, channels=4, position=(0.0, 0.0, 0.0))
    return mesh, b, fc, sc, rf


def test_precision_study_phase_variants(oracle, precision):
    mesh, b, fc, sc, rf = precision
    kn = H.Knobs(maxDepth=4, doGuiding=False)
    n = 1 << 16
    om = oracle.OracleMesh.from_mesh(mesh)
    print("precision study: oracle FP64 and mixed", flush=True)
    zf, twf, _, _ = oracle.render(om, sc, kn, fc, prec=oracle.PREC_FP64, count=n, threads=THREADS)
    zm, twm, recm, _ = oracle.render(om, sc, kn, fc, prec=oracle.PREC_MIXED, count=n, record_paths=n,
                                     threads=THREADS)
    zf = (zf / twf).reshape(512, 4)
    zm = (zm / twm).reshape(512, 4)
    pts = []
    for rng in (5.0, 50.0, 150.0):
        pts += [np.array([-1.5, 0.0, rng]), np.array([1.5, -0.25, rng + 0.25])]
    bins = target_bins(pts, rf, 512)
    assert np.all(np.abs(zf[bins]) > 0)
    peaks_ref = np.abs(zf).argmax(0).tolist()
    for mode, tol, ptol in ((0, 1e-5, 1e-5), (2, 1e-5, 1e-5), (1, 1e-2, 0.05)):
        print(f"precision study: phase mode {mode}", flush=True)
        b.set_option(OPT_PHASE_MODE, mode)
        b.clear()
        b.set_option(OPT_RECORD_HITS, n)
        b.run_raw(sc, kn, 0, n)
        z, tw = b.fetch("complex128")
        hits, nee = b.hit_record(n, 4, 4)
        z = z.reshape(512, 4)
        assert tw == twf == n
        np.testing.assert_array_equal(hits, recm[0])
        np.testing.assert_array_equal(nee, recm[1])
        assert (hits[:, 0] >= 0).sum() > n // 8
        err = rel_l2(z, zf)
        assert err < tol, (mode, err)
        if mode == 0:
            assert rel_l2(z, zm) < 1e-6
        assert np.abs(z).argmax(0).tolist() == peaks_ref
        dphi = np.abs(np.angle(z[bins] / zf[bins])).max()
        assert dphi < ptol, (mode, dphi)
        if mode == 1:
            assert err > 1e-4  # the FP32 phase is measurably worse
    b.set_option(OPT_PHASE_MODE, 0)
    b.set_option(OPT_RECORD_HITS, 0)


def test_urban_mimo_guided_cube(oracle):
    print("urban MIMO: building the 2M-triangle canyon", flush=True)
    mesh, vel = scenes.urban_mesh(seed=4)
    assert mesh.triangle_count() >= 2_000_000
    rf = scenes.dihedral_rf()
    sc = scenes.mimo_array(rf, tx=3, rx=4, position=(0.0, 1.0, 0.0))
    kn = H.Knobs(maxDepth=5, doGuiding=True)
    fc = H.FrameConfig(64, 256, 12)
    pt = H.PathTracer(**{k: getattr(kn, k) for k in kn.__dataclass_fields__})
    pt.configureFrame(fc)
    b = H.HipBackend(mesh, pt)
    b.set_velocities(vel)
    print("urban MIMO: training the guide on the device", flush=True)
    pt.run_schedule(b, sc, 4, 1 << 18, 1 << 16)
    child, p0 = pt.guiding.sampling()
    assert len(child) > 3
    n = 1 << 14
    b.clear()
    b.set_option(OPT_RECORD_HITS, n)
    b.run_raw(sc, kn, 0, n, guide=pt.guiding, is_final=True)
    z, tw = b.fetch("complex128")
    hits, nee = b.hit_record(n, 5, 12)
    om = oracle.OracleMesh.from_mesh(mesh)
    om.set_velocities(vel)
    og = oracle.OracleGuide()
    og.set_sampling(child, p0)
    print("urban MIMO: oracle mixed and FP64", flush=True)
    zm, twm, recm, _ = oracle.render(om, sc, kn, fc, prec=oracle.PREC_MIXED, count=n, is_final=True, guide=og,
                                     record_paths=n, threads=THREADS)
    zf, twf, _, _ = oracle.render(om, sc, kn, fc, prec=oracle.PREC_FP64, count=n, is_final=True, guide=og,
                                  threads=THREADS)
    assert tw == twm == twf == n
    np.testing.assert_array_equal(hits, recm[0])
    np.testing.assert_array_equal(nee, recm[1])
    assert (hits[:, 0] >= 0).sum() > n // 2  # guided paths hit the canyon
    assert (hits[:, 2] >= 0).sum() > 0       # and some bounce twice or more
    z = z.reshape(zm.shape)
    zm, zf = zm / twm, zf / twf
    assert np.all(np.abs(zf).max(axis=(0, 1)) > 0)  # all twelve virtual channels
    assert np.count_nonzero(np.abs(zf).sum(axis=(1, 2)) > 0) > 32  # energy spread over Doppler
    assert rel_l2(z, zm) < 1e-5
    assert rel_l2(z, zf) < 1e-3
    assert int(np.abs(z).argmax()) == int(np.abs(zf).argmax())


def test_serving_mix_stated_scenes(oracle):
    """The serving mix on its stated scenes (vehicle clutter of 10k, 300k and 2M triangles,
    seeds 10-12) with resident guiding trees: each scene's tree trained by one request (the same
    tree a direct PathTracer::run of it trains), then eight requests of the mix (poses, bins and
    receivers as drawn, 2^15 paths each) served together on four slots. Every served frame equals
    the same request served alone and is within 1e-5 of the oracle's final pass with the
    scene's tree."""
    from hussar_amd.serving import GUIDE_SCENE, GUIDE_TRAIN, Server, request_mix
    print("serving mix: building the 10k / 300k / 2M scenes", flush=True)
    meshes = [scenes.vehicle_mesh(seed=10 + k, triangles=n)[0] for k, n in enumerate((10_000, 300_000, 2_000_000))]
    rf = scenes.dihedral_rf()
    kn = H.Knobs(maxDepth=4, doGuiding=True)
    srv = Server(slots=4, pathsPerBatch=1 << 20)
    ids = [srv.open_scene(m) for m in meshes]
    nominal = scenes.radar_array(rf, channels=4, position=(0.0, 0.75, 0.0))
    trees = []
    for sid, m in zip(ids, meshes):
        _, info = srv.wait(srv.submit(sid, nominal, kn, H.FrameConfig(1, 256, 4), 1 << 18, guide_mode=GUIDE_TRAIN))
        assert info.status == 0
        pt = H.PathTracer(**{k: getattr(kn, k) for k in kn.__dataclass_fields__})
        pt.configureFrame(H.FrameConfig(1, 256, 4))
        pt.run(H.HipBackend(m, pt), nominal, 1 << 18)
        trees.append(pt.guiding.sampling())
    mix = request_mix(8, 100.0)
    assert {q["scene"] for q in mix} == {0, 1, 2}
    n = 1 << 15
    reqs = [(ids[q["scene"]], scenes.radar_array(rf, channels=q["channels"], position=q["pos"]),
             H.FrameConfig(1, q["bins"], q["channels"])) for q in mix]
    print("serving mix: eight requests together", flush=True)
    tickets = [srv.submit(sid, sc, kn, fc, n, guide_mode=GUIDE_SCENE) for sid, sc, fc in reqs]
    served = [srv.wait(t) for t in tickets]
    oms = [oracle.OracleMesh.from_mesh(m) for m in meshes]
    nonzero = 0
    for (sid, sc, fc), (z, info), q in zip(reqs, served, mix):
        assert info.status == 0 and info.totalWeight == n
        za, _ = srv.wait(srv.submit(sid, sc, kn, fc, n, guide_mode=GUIDE_SCENE))
        og = oracle.OracleGuide()
        og.set_sampling(*trees[q["scene"]])
        zo, two, _, _ = oracle.render(oms[q["scene"]], sc, kn, fc, prec=oracle.PREC_MIXED, count=n, is_final=True,
                                      guide=og, threads=THREADS)
        zo = (zo / two).reshape(z.shape)
        print(f"serving mix: scene {q['scene']}, {q['bins']} bins x {q['channels']} channels, "
              f"|frame| {np.abs(zo).max():.3g}", flush=True)
        if np.abs(zo).max() > 0:
            nonzero += 1
            assert rel_l2(za, z) < 1e-6
            assert rel_l2(z, zo) < 1e-5
        else:
            assert np.abs(z).max() == 0 and np.abs(za).max() == 0
    assert nonzero >= 4
    srv.close()
