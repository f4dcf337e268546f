"""Hit-id parity against the reference's own FP32 arithmetic (HSR_OPT_CHAIN = 1), through the
C ABI.

The oracle's ref32 instance restates the reference's FP32 geometry chain
(libhussar/include/hussar/integrators/path.h:236-442, libhussar/src/arch/cpu.cpp:9-18,73-86):
sphere warp, rotation, hit point o + t·d, face-forwarded normal, reflection and receiver legs
in float. The product's default chain is FP64 (its rays are rounded from it), which follows
the exact geometry more closely, so on a rare grazing or edge-aimed ray it hits another
triangle than the FP32 chain does. With HSR_OPT_CHAIN = 1 the traversal and shadow rays and
every discrete decision come from the FP32 chain while lengths, time, phase and weights stay
FP64 (the oracle's dual instance is the same model). Bar:
  * hit ids and shadow outcomes bit-exact against ref32 (and the dual instance) on every
    recorded path, with the same Halton indices;
  * the frame within 1e-5 relative L2 of the dual instance; the peak range bin of every channel
    equal to the FP64 oracle's.
The FP32 chain's own decisions cost accuracy: on the guided vehicle a path whose FP32 chain
takes another triangle than the exact geometry is typically a grazing bounce with a large GO
weight r/cosθ, so one such path in 32k moves the frame by about 3e-3 relative L2 against the
FP64 oracle (without one, the dual instance is within 2e-6 of it). The frame is therefore held
to 1e-2 against FP64 here (the reference's full FP32 arithmetic, ref32, is at about 8e-3), while
the default FP64 chain meets the 1e-3 bar (about 3e-7). The default chain's disagreement with
ref32 is measured too (a small fraction of paths)."""
import os

import numpy as np
import pytest

import hussar_amd as H
from hussar_amd import scenes
from hussar_amd._lib import HsrError, OPT_CHAIN, OPT_PIPELINE, OPT_RECORD_HITS
from conftest import rel_l2

pytestmark = pytest.mark.gpu
THREADS = min(16, os.cpu_count() or 1)


@pytest.fixture(scope="module")
def vehicle_guided(oracle):
    """The 200k-triangle vehicle workload with a sampling tree trained on the device (four
    passes of 2^20 paths), shared with the oracle."""
    mesh, _ = scenes.vehicle_mesh(seed=1, triangles=200_000)
    sc = scenes.radar_array(scenes.dihedral_rf(), channels=4, position=(0.0, 0.75, 0.0))
    kn = H.Knobs(maxDepth=4, doGuiding=True)
    fc = H.FrameConfig(1, 256, 4)
    pt = H.PathTracer(**{k: getattr(kn, k) for k in kn.__dataclass_fields__})
    pt.configureFrame(fc)
    b = H.HipBackend(mesh, pt)
    print("reference chain: training the vehicle guide", flush=True)
    pt.run_schedule(b, sc, 4, 1 << 20, 1 << 16)
    child, p0 = pt.guiding.sampling()
    og = oracle.OracleGuide()
    og.set_sampling(child, p0)
    return mesh, oracle.OracleMesh.from_mesh(mesh), b, pt, og, sc, fc


def gpu_run(b, sc, kn, n, guide=None, chain=0):
    b.set_option(OPT_CHAIN, chain)
    b.clear()
    b.set_option(OPT_RECORD_HITS, n)
    b.run_raw(sc, kn, 0, n, guide=guide, is_final=True)
    b.set_option(OPT_RECORD_HITS, 0)
    z, tw = b.fetch("complex128")
    hits, nee = b.hit_record(n, max(1, kn.maxDepth), len(sc.receivers()))
    b.set_option(OPT_CHAIN, 0)
    return z, tw, hits, nee


def test_vehicle_reference_chain_hits_bitexact(oracle, vehicle_guided):
    mesh, om, b, pt, og, sc, fc = vehicle_guided
    kn = H.Knobs(maxDepth=4, doGuiding=True)
    n = 1 << 18
    print("reference chain: vehicle, GPU both chains", flush=True)
    z1, tw, hits1, nee1 = gpu_run(b, sc, kn, n, pt.guiding, chain=1)
    z0, _, hits0, nee0 = gpu_run(b, sc, kn, n, pt.guiding, chain=0)
    print("reference chain: vehicle, oracle ref32 / dual / FP64", flush=True)
    _, _, rec32, _ = oracle.render(om, sc, kn, fc, prec=oracle.PREC_REF32, count=n, is_final=True, guide=og,
                                   record_paths=n, threads=THREADS)
    zd, twd, recd, _ = oracle.render(om, sc, kn, fc, prec=oracle.PREC_DUAL, count=n, is_final=True, guide=og,
                                     record_paths=n, threads=THREADS)
    zf, twf, _, _ = oracle.render(om, sc, kn, fc, prec=oracle.PREC_FP64, count=n, is_final=True, guide=og,
                                  threads=THREADS)
    assert tw == twd == twf == n
    assert (hits1[:, 0] >= 0).sum() > n // 2  # guided paths hit the vehicle
    np.testing.assert_array_equal(recd[0], rec32[0])
    np.testing.assert_array_equal(recd[1], rec32[1])
    np.testing.assert_array_equal(hits1, rec32[0])
    np.testing.assert_array_equal(nee1, rec32[1])
    zd, zf = zd / twd, zf / twf
    assert rel_l2(z1, zd) < 1e-5
    assert rel_l2(z1, zf) < 1e-2
    assert rel_l2(z0, zf) < 1e-5
    assert np.abs(z1).reshape(-1, 4).argmax(0).tolist() == np.abs(zf).reshape(-1, 4).argmax(0).tolist()
    # the default FP64 chain agrees with ref32 on all but a small fraction of paths
    other = int((hits0 != rec32[0]).any(1).sum())
    print(f"reference chain: FP64-chain paths with other hit ids than ref32: {other} of {n}", flush=True)
    assert other < n // 1000
    assert rel_l2(z0, zf) < rel_l2(z1, zf) + 1e-6  # and follows the exact geometry at least as well


@pytest.mark.parametrize("case", ["plane_filter", "plane_filter_guiding", "no_filter", "po_diffraction", "clang_order",
                                  "correct_phase"])
def test_reference_chain_knob_variants(oracle, vehicle_guided, case):
    mesh, om, b, pt, og, sc, fc = vehicle_guided
    kw = dict(maxDepth=4, doGuiding=True)
    kw.update({"plane_filter": dict(filteringSphere=False, filteringMin=40.0, filteringMax=300.0),
               "plane_filter_guiding": dict(filteringSphere=False, filteringMin=40.0, filteringMax=300.0,
                                            filterGuiding=True),
               "no_filter": dict(doFiltering=False),
               "po_diffraction": dict(poDiffraction=True), "clang_order": dict(primaryOrder=1),
               # the hard sphere of the phase correction at the default radius; the debug image
               # keeps the real corrected values from the splat skip rule
               "correct_phase": dict(correctPhase=True, filterGuiding=True, produceDebugImage=True)}[case])
    kn = H.Knobs(**kw)
    n = 1 << 16
    print(f"reference chain: {case}", flush=True)
    z, tw, hits, nee = gpu_run(b, sc, kn, n, pt.guiding, chain=1)
    _, _, rec32, _ = oracle.render(om, sc, kn, fc, prec=oracle.PREC_REF32, count=n, is_final=True, guide=og,
                                   record_paths=n, threads=THREADS)
    zd, twd, recd, _ = oracle.render(om, sc, kn, fc, prec=oracle.PREC_DUAL, count=n, is_final=True, guide=og,
                                     record_paths=n, threads=THREADS)
    np.testing.assert_array_equal(hits, rec32[0])
    np.testing.assert_array_equal(nee, rec32[1])
    np.testing.assert_array_equal(hits, recd[0])
    np.testing.assert_array_equal(nee, recd[1])
    zd = zd / twd
    assert np.abs(zd).max() > 0
    assert rel_l2(z, zd) < 1e-5


@pytest.mark.parametrize("yaw", [-45.0, 0.0, 30.0])
def test_dihedral_reference_chain(oracle, yaw):
    mesh = scenes.dihedral_mesh()
    fc = H.FrameConfig(1, 256, 1)
    b = H.HipBackend(mesh)
    b._configure(fc)
    sc = scenes.dihedral_pose(yaw)
    kn = H.Knobs(maxDepth=3, doGuiding=False)
    n = 1 << 18
    z, tw, hits, nee = gpu_run(b, sc, kn, n, chain=1)
    om = oracle.OracleMesh.from_mesh(mesh)
    _, _, rec32, _ = oracle.render(om, sc, kn, fc, prec=oracle.PREC_REF32, count=n, record_paths=n, threads=THREADS)
    zf, twf, _, _ = oracle.render(om, sc, kn, fc, prec=oracle.PREC_FP64, count=n, threads=THREADS)
    np.testing.assert_array_equal(hits, rec32[0])
    np.testing.assert_array_equal(nee, rec32[1])
    assert rel_l2(z, zf / twf) < 1e-4
    assert np.abs(z).argmax() == np.abs(zf).argmax()


def test_reference_chain_direct_paths(oracle):
    """onlyIndirect off: the depth-0 NEE legs Tx -> Rx of the FP32 chain (a receiver 0.5 m
    beside the transmitter; a co-located pair has a zero-length leg, NaN in every arithmetic)."""
    mesh = scenes.plate_mesh(size=0.5, distance=2.0)
    fc = H.FrameConfig(1, 256, 2)
    rf = scenes.dihedral_rf()
    sc = scenes.radar_array(rf, channels=1, position=(0.0, 0.0, 0.0))
    rx = [H.NFAntenna(np.float32([0.5, 0.0, 0.0]), np.eye(3, dtype=np.float32)),
          H.NFAntenna(np.float32([0.0, 0.3, 0.1]), np.eye(3, dtype=np.float32))]
    sc = H.Scene(rfConfig=rf, tx=sc.tx, rx=rx[0], rxs=rx)
    kn = H.Knobs(maxDepth=3, doGuiding=False, onlyIndirect=False)
    b = H.HipBackend(mesh)
    b._configure(fc)
    n = 1 << 16
    z, tw, hits, nee = gpu_run(b, sc, kn, n, chain=1)
    om = oracle.OracleMesh.from_mesh(mesh)
    _, _, rec32, _ = oracle.render(om, sc, kn, fc, prec=oracle.PREC_REF32, count=n, record_paths=n, threads=THREADS)
    zd, twd, _, _ = oracle.render(om, sc, kn, fc, prec=oracle.PREC_DUAL, count=n, threads=THREADS)
    np.testing.assert_array_equal(hits, rec32[0])
    np.testing.assert_array_equal(nee, rec32[1])
    assert np.isfinite(z).all() and np.abs(zd).max() > 0
    assert rel_l2(z, zd / twd) < 1e-5


def test_reference_chain_needs_the_wavefront_pipeline():
    b = H.HipBackend(scenes.dihedral_mesh())
    b._configure(H.FrameConfig(1, 256, 1))
    b.set_option(OPT_PIPELINE, 0)
    b.set_option(OPT_CHAIN, 1)
    with pytest.raises(HsrError, match="UNSUPPORTED"):
        b.run_raw(scenes.dihedral_pose(0.0), H.Knobs(maxDepth=3, doGuiding=False), 0, 1024)
    with pytest.raises(HsrError, match="UNSUPPORTED"):
        b.set_option(OPT_CHAIN, 2)
