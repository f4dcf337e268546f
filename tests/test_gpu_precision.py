"""The IF-phase precision variants (HSR_OPT_PHASE_MODE) on targets at 5, 50 and 150 m over a
rough ground: FP64 (the default) and double-float stay within the 1e-3 relative-L2 bar of
the FP64 oracle by orders of magnitude, the reference's FP32 evaluation does not at range
(its phase argument reaches 10^5 rad). Hit ids and shadow outcomes do not depend on the
mode."""
import os

import numpy as np
import pytest

import hussar_amd as H
from hussar_amd import scenes
from hussar_amd._lib import OPT_PHASE_MODE, OPT_RECORD_HITS, PHASE_DOUBLE_FLOAT, PHASE_FP32, PHASE_FP64
from conftest import rel_l2

pytestmark = pytest.mark.gpu
THREADS = min(16, os.cpu_count() or 1)


def test_phase_modes_against_fp64(oracle):
    mesh = scenes.precision_mesh(seed=2, ground_nx=100, ground_nz=200)
    rf = scenes.dihedral_rf()
    sc = scenes.radar_array(rf, channels=4, position=(0.0, 0.0, 0.0))
    kn = H.Knobs(maxDepth=3, doGuiding=False)
    fc = H.FrameConfig(1, 512, 4)
    n = 1 << 17
    om = oracle.OracleMesh.from_mesh(mesh)
    zf, twf, _, _ = oracle.render(om, sc, kn, fc, prec=oracle.PREC_FP64, count=n, threads=THREADS)
    zf = zf / twf
    b = H.HipBackend(mesh)
    b._configure(fc)
    err, recs = {}, {}
    for mode in (PHASE_FP64, PHASE_DOUBLE_FLOAT, PHASE_FP32):
        print("phase mode", mode, flush=True)
        b.set_option(OPT_PHASE_MODE, mode)
        b.set_option(OPT_RECORD_HITS, 4096)
        b.clear()
        b.run_raw(sc, kn, 0, n)
        z, _ = b.fetch("complex128")
        err[mode] = rel_l2(z, zf)
        recs[mode] = b.hit_record(4096, 3, 4)
    assert err[PHASE_FP64] < 1e-5 and err[PHASE_DOUBLE_FLOAT] < 1e-5
    assert err[PHASE_FP32] > 10 * max(err[PHASE_FP64], err[PHASE_DOUBLE_FLOAT])
    for mode in (PHASE_DOUBLE_FLOAT, PHASE_FP32):
        assert np.array_equal(recs[mode][0], recs[PHASE_FP64][0]) and np.array_equal(recs[mode][1], recs[PHASE_FP64][1])
