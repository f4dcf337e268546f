"""Frame post-processing on the device (hsr_hip_fetch_frame_fft, hsr_hip_frame_peak):
Frame::fft (radar.h:420-435, FFTW forward 3-D, unnormalised) against numpy's FFT of the same
fetched frame, and Frame::argmax + frequencyEstimation against the oracle's restatement (itself
pinned by the closed form of a single splat)."""
import numpy as np
import pytest

import hussar_amd as H
from hussar_amd import scenes
from hussar_amd.types import RFConfig
from conftest import rel_l2

pytestmark = pytest.mark.gpu


def frame_of(mesh, scene, knobs, fc, n, vel=None):
    b = H.HipBackend(mesh)
    b._configure(fc)
    if vel is not None:
        b.set_velocities(vel)
    b.run_raw(scene, knobs, 0, n)
    return b


@pytest.mark.parametrize("shape", [(1, 256, 4), (1, 211, 3), (16, 64, 4), (64, 128, 2)])
def test_device_fft_matches_numpy(shape):
    M, N, C = shape
    mesh, _ = scenes.vehicle_mesh(seed=5, triangles=20_000, range_lo=8.0, range_hi=12.0)
    rf = RFConfig(startFreq=77e9, freqSlope=60e12, adcRate=10e6, idleTime=100e-6, rampTime=25.6e-6)
    sc = scenes.radar_array(rf, channels=C, position=(0.0, 0.75, 0.0))
    vel = None
    if M > 1:
        u = scenes.uniforms(11, 3 * mesh.triangle_count()).reshape(-1, 3)
        vel = ((u - 0.5) * 40.0).astype(np.float32)
    print("device fft", shape, flush=True)
    b = frame_of(mesh, sc, H.Knobs(maxDepth=3, doGuiding=False), H.FrameConfig(M, N, C), 1 << 16, vel)
    z, _ = b.fetch("complex128")
    f = b.fetch_fft("complex128")
    ref = np.fft.fftn(z.reshape(M, N, C))
    assert np.abs(ref).max() > 0
    assert rel_l2(f, ref) < 1e-12
    # the device peak search on the transformed and the plain frame
    for fourier, zz in ((True, ref), (False, z.reshape(M, N, C))):
        idx, refined, mag = b.peak(fourier)
        assert idx == tuple(int(i) for i in np.unravel_index(np.abs(zz).argmax(), zz.shape))
        assert abs(mag - np.abs(zz).max()) <= 1e-12 * mag


def test_peak_matches_oracle_and_dihedral_range(oracle):
    """The dihedral at 1.5 m: the device peak refinement equals the oracle's restatement on
    the same frame, and lands at the corner's range bin 2R/c·S/f_s·N (closed form, 15.37)."""
    mesh = scenes.dihedral_mesh()
    fc = H.FrameConfig(1, 256, 1)
    b = frame_of(mesh, scenes.dihedral_pose(0.0), H.Knobs(maxDepth=3, doGuiding=False), fc, 1 << 20)
    z, _ = b.fetch("complex128")
    idx, refined, _ = b.peak(False)
    oidx, oref = oracle.frame_peak(z.reshape(1, 256, 1))
    assert idx == oidx
    assert np.allclose(refined, oref, rtol=0, atol=1e-12)
    c = float(np.float32(299792458.0))
    # the double bounce returns in the plane normal to the corner line: 2 x 1.5 m
    x = 2 * 1.5 / c * 60e12 / 10e6 * 256
    assert idx[1] == 15 and abs(refined[1] - x) < 0.15
