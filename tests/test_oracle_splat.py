"""Oracle pins for Integrator::splat -> Frame::splat<16> (integrator.h:113-135,
radar.h:232-241, 564-627): the fractional range bin, the IF phase, the 33-tap leakage kernel
and the frame layout, against closed forms and a DFT of a tone."""
import numpy as np
import pytest

from hussar_amd.types import RFConfig, SPEED_OF_LIGHT


def splat(oracle, rf, N, C, ch, dt, v, w, prec=2):
    import ctypes
    frame = np.zeros(2 * N * C, np.float64)
    oracle.lib().orc_splat(prec, ctypes.byref(rf.to_c()), 1, N, C, ch, dt, v.real, v.imag, w, 0, 0.0,
                           frame.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
    return (frame[0::2] + 1j * frame[1::2]).reshape(N, C)


def closed_form(rf, N, dt, v, w):
    """tap(m) = w·v·exp(j·2π(f0 - Δt·S/2)Δt)·e^{jπd}·sinc(d), d = x - m, |m - round(x)| <= 16."""
    f0, S, fs = float(np.float32(rf.startFreq)), float(np.float32(rf.freqSlope)), float(np.float32(rf.adcRate))
    dt = dt + float(np.float32(rf.antennaDelay))
    q = dt * S / fs
    x = N * (q - np.floor(q))
    c = int(np.floor(x + 0.5))
    val = w * v * np.exp(2j * np.pi * (f0 - dt * S / 2) * dt)
    out = np.zeros(N, complex)
    if abs(x - c) < 1e-4:
        out[c % N] += val
        return out, x
    for k in range(-16, 17):
        m = c + k
        d = x - m
        out[m % N] += val * np.exp(1j * np.pi * d) * np.sinc(d)
    return out, x


@pytest.mark.parametrize("dist", [1.5, 3.7, 12.0, 40.0])
def test_splat_closed_form(oracle, dist):
    rf = RFConfig()
    N = 256
    dt = 2 * dist / SPEED_OF_LIGHT
    v, w = 0.3 - 1.7j, 2.0
    got = splat(oracle, rf, N, 1, 0, dt, v, w)[:, 0]
    want, x = closed_form(rf, N, dt, v, w)
    np.testing.assert_allclose(got, want, rtol=1e-9, atol=1e-12 * np.abs(want).max())
    assert np.abs(got).argmax() == int(np.floor(x + 0.5)) % N


def test_splat_delta_peak_and_wrap(oracle):
    rf = RFConfig(antennaDelay=0.0)
    N = 64
    # choose dt so that x is (almost) an integer: x = N * frac(dt S / fs) = 10
    dt = (10.0 / N) * float(np.float32(rf.adcRate)) / float(np.float32(rf.freqSlope))
    got = splat(oracle, rf, N, 1, 0, dt, 1.0 + 0j, 1.0)[:, 0]
    assert np.count_nonzero(got) == 1 and np.abs(got[10]) == pytest.approx(1.0)
    # near the top of the axis the taps wrap modulo N
    dt2 = (62.6 / N) * float(np.float32(rf.adcRate)) / float(np.float32(rf.freqSlope))
    got2 = splat(oracle, rf, N, 1, 0, dt2, 1.0 + 0j, 1.0)[:, 0]
    want2, _ = closed_form(rf, N, dt2, 1.0 + 0j, 1.0)
    np.testing.assert_allclose(got2, want2, rtol=1e-9, atol=1e-12)
    assert np.count_nonzero(np.abs(got2) > 0) == 33


def test_splat_frame_layout_channels(oracle):
    rf = RFConfig()
    N, C = 128, 4
    got = splat(oracle, rf, N, C, 2, 2 * 5.0 / SPEED_OF_LIGHT, 1.0 + 0j, 1.0)
    assert np.all(got[:, [0, 1, 3]] == 0) and np.count_nonzero(got[:, 2]) == 33


def test_splat_approximates_dft_of_a_tone(oracle):
    """The splat is the large-N Dirichlet kernel truncated to ±16 bins: near the peak it
    equals (1/N)·DFT of the IF tone e^{j2π x n / N} (FourierSplatting.ipynb cells 5-6)."""
    rf = RFConfig(antennaDelay=0.0)
    N = 8192
    dt = 2 * 4.321 / SPEED_OF_LIGHT
    got = splat(oracle, rf, N, 1, 0, dt, 1.0 + 0j, 1.0)[:, 0]
    _, x = closed_form(rf, N, dt, 1.0 + 0j, 1.0)
    f0, S = float(np.float32(rf.startFreq)), float(np.float32(rf.freqSlope))
    n = np.arange(N)
    tone = np.exp(2j * np.pi * (f0 - dt * S / 2) * dt) * np.exp(2j * np.pi * x * n / N)
    X = np.fft.fft(tone) / N
    c = int(np.floor(x + 0.5))
    win = np.arange(c - 16, c + 17) % N
    np.testing.assert_allclose(got[win], X[win], rtol=0, atol=2.5e-3)
    # outside the window the splat is zero while the DFT leaks ~ 1/(π|d|)
    mask = np.ones(N, bool)
    mask[win] = False
    assert np.all(got[mask] == 0)


def test_splat_fp32_instance_close(oracle):
    rf = RFConfig()
    dt = 2 * 1.5 / SPEED_OF_LIGHT
    a = splat(oracle, rf, 256, 1, 0, dt, 1.0 + 0j, 1.0, prec=0)
    b = splat(oracle, rf, 256, 1, 0, dt, 1.0 + 0j, 1.0, prec=2)
    # FP32 phase of a ~4.8e3 rad argument: error of order 1e-3 relative
    assert np.linalg.norm(a - b) / np.linalg.norm(b) < 5e-3


@pytest.mark.parametrize("dist", [1.5, 3.7, 12.0])
def test_frequency_estimation_recovers_the_fractional_bin(oracle, dist):
    """Frame::frequencyEstimation (radar.h:631-643) of a single splat: the neighbours of the
    peak hold |sinc(d)| and |sinc(1-d)| times the same amplitude, so b + r/(r+m) (or b -
    l/(l+m)) is exactly the fractional bin x of the closed form; axes of length one give 0.5."""
    rf = RFConfig()
    N = 256
    dt = 2 * dist / SPEED_OF_LIGHT
    z = splat(oracle, rf, N, 1, 0, dt, 1.0 + 0.5j, 1.0)
    _, x = closed_form(rf, N, dt, 1.0 + 0.5j, 1.0)
    idx, ref = oracle.frame_peak(z.reshape(1, N, 1))
    assert idx == (0, int(np.floor(x + 0.5)) % N, 0)
    assert abs(ref[1] - x) < 1e-9
    assert ref[0] == 0.5 and ref[2] == 0.5
