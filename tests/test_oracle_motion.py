"""Oracle pins for the motion and MIMO extensions of the frame (no reference counterpart: the
reference splats every contribution at velocity 0, integrator.h:130, with one transmitter).

* The chirp axis uses the reference's own PIndex::setVelocity (radar.h:248-251) and the
  separable Frame::splat kernel (radar.h:564-627): each tap equals
  v·e^{jπd_c}·sinc(d_c)·e^{jπd_s}·sinc(d_s) (closed form, both axes, with wrap-around).
* A static scene through the motion code (all velocities zero) gives the static frame bit for
  bit; radar motion +v towards a plate equals plate motion -v bit for bit (only relative
  velocities enter the path-length rate).
* A plate approaching at v puts the frame's energy at the chirp bin of the closed-form Doppler
  shift M·frac(2·f0·v/c·T_chirp).
* A MIMO frame's channel block of transmitter t equals a single-transmitter frame of t, and
  the total weight counts each path index once.
"""
import ctypes

import numpy as np
import pytest

import hussar_amd as H
from hussar_amd import scenes
from hussar_amd.types import RFConfig, SPEED_OF_LIGHT


def splat2(oracle, rf, M, N, dt, v, w, dv, prec=2):
    frame = np.zeros(2 * M * N, np.float64)
    oracle.lib().orc_splat(prec, ctypes.byref(rf.to_c()), M, N, 1, 0, dt, v.real, v.imag, w, 1, dv,
                           frame.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
    return (frame[0::2] + 1j * frame[1::2]).reshape(M, N)


def axis_taps(x, n):
    """±16 taps of e^{jπd}·sinc(d), d = x - m, folded modulo n (or the delta peak)."""
    c = int(np.floor(x + 0.5))
    out = np.zeros(n, complex)
    if abs(x - c) < 1e-4:
        out[c % n] += 1.0
        return out
    for k in range(-16, 17):
        d = x - (c + k)
        out[(c + k) % n] += np.exp(1j * np.pi * d) * np.sinc(d)
    return out


@pytest.mark.parametrize("M,dv", [(32, 3.1), (64, -7.3), (8, 11.9), (16, 0.0)])
def test_cube_splat_closed_form(oracle, M, dv):
    rf = RFConfig(startFreq=77e9, freqSlope=60e12, adcRate=10e6, idleTime=100e-6, rampTime=25.6e-6)
    N = 128
    dt, v, w = 2 * 7.31 / SPEED_OF_LIGHT, 0.3 - 0.8j, 0.7
    z = splat2(oracle, rf, M, N, dt, v, w, dv)
    f0, S, fs = [float(np.float32(x)) for x in (rf.startFreq, rf.freqSlope, rf.adcRate)]
    q = dt * S / fs
    xs = N * (q - np.floor(q))
    Tc = float(np.float32(rf.idleTime)) + float(np.float32(rf.rampTime))
    qc = (2 * f0 * dv / SPEED_OF_LIGHT) * Tc
    xc = M * (qc - np.floor(qc))
    val = w * v * np.exp(2j * np.pi * (f0 - dt * S / 2) * dt)
    ref = val * np.outer(axis_taps(xc, M), axis_taps(xs, N))
    assert np.abs(z - ref).max() < 1e-9 * np.abs(ref).max()
    if dv == 0.0:
        assert np.count_nonzero(np.abs(z).sum(1)) == 1  # the chirp axis is a delta at chirp 0


def motion_setup(v_plate=(0.0, 0.0, 0.0), v_radar=(0.0, 0.0, 0.0), chirps=32, channels=1):
    rf = RFConfig(startFreq=77e9, freqSlope=60e12, adcRate=10e6, idleTime=100e-6, rampTime=25.6e-6)
    mesh = scenes.plate_mesh(0.5, 5.0)
    sc = scenes.radar_array(rf, channels=channels, position=(0.0, 0.0, 0.0))
    sc.velocity = v_radar
    kn = H.Knobs(maxDepth=2, doGuiding=False)
    fc = H.FrameConfig(chirps, 128, channels)
    vel = np.tile(np.asarray(v_plate, np.float32), (mesh.triangle_count(), 1))
    return mesh, sc, kn, fc, vel


def test_zero_velocities_give_the_static_frame(oracle):
    mesh, sc, kn, fc, vel = motion_setup(chirps=4)
    m = oracle.OracleMesh.from_mesh(mesh)
    z0, tw0, _, _ = oracle.render(m, sc, kn, fc, prec=1, count=1 << 14, threads=1)
    m.set_velocities(vel)  # all zero, but the motion path runs
    z1, tw1, _, _ = oracle.render(m, sc, kn, fc, prec=1, count=1 << 14, threads=1)
    assert np.array_equal(z0, z1) and tw0 == tw1
    assert np.abs(z0[1:]).max() == 0.0 and np.abs(z0[0]).max() > 0


def test_radar_motion_equals_opposite_plate_motion(oracle):
    v = 4.2
    mesh, sc_still, kn, fc, vel = motion_setup(v_plate=(0.0, 0.0, -v))
    m = oracle.OracleMesh.from_mesh(mesh)
    m.set_velocities(vel)
    za, _, _, _ = oracle.render(m, sc_still, kn, fc, prec=1, count=1 << 14, threads=1)
    _, sc_moving, _, _, _ = motion_setup(v_radar=(0.0, 0.0, v))
    m2 = oracle.OracleMesh.from_mesh(mesh)
    zb, _, _, _ = oracle.render(m2, sc_moving, kn, fc, prec=1, count=1 << 14, threads=1)
    assert np.array_equal(za, zb)


@pytest.mark.parametrize("v", [3.0, 9.5, -6.0])
def test_approaching_plate_lands_on_the_doppler_bin(oracle, v):
    M = 32
    mesh, sc, kn, fc, vel = motion_setup(v_plate=(0.0, 0.0, -v), chirps=M)
    m = oracle.OracleMesh.from_mesh(mesh)
    m.set_velocities(vel)
    z, tw, _, _ = oracle.render(m, sc, kn, fc, prec=2, count=1 << 15)
    f0 = float(np.float32(77e9))
    Tc = float(np.float32(100e-6)) + float(np.float32(25.6e-6))
    q = 2 * f0 * v / SPEED_OF_LIGHT * Tc
    xc = M * (q - np.floor(q))
    energy = (np.abs(z) ** 2).sum(axis=(1, 2))
    assert int(energy.argmax()) == int(np.floor(xc + 0.5)) % M
    # the same contributions without motion all sit in chirp 0, at the same range bin
    m0 = oracle.OracleMesh.from_mesh(mesh)
    z0, _, _, _ = oracle.render(m0, sc, kn, fc, prec=2, count=1 << 15)
    assert np.abs(z0[1:]).max() == 0.0
    assert int(np.abs(z0[0, :, 0]).argmax()) == int(np.abs(z[int(energy.argmax()), :, 0]).argmax())


def test_mimo_blocks_equal_single_transmitter_frames(oracle):
    mesh = scenes.vehicle_mesh(seed=3, triangles=4000, range_lo=6.0, range_hi=9.0)[0]
    rf = scenes.dihedral_rf()
    sc = scenes.radar_array(rf, channels=2, position=(0.0, 0.75, 0.0))
    tx1 = H.NFAntenna(np.asarray(sc.tx.position) + np.float32([0.02, 0.0, 0.0]), sc.tx.rotation)
    mimo = H.Scene(rfConfig=rf, tx=sc.tx, rx=sc.rx, rxs=sc.rxs, txs=[sc.tx, tx1])
    kn = H.Knobs(maxDepth=3, doGuiding=False)
    fc = H.FrameConfig(1, 256, 4)
    m = oracle.OracleMesh.from_mesh(mesh)
    n = 1 << 13
    z, tw, rec, _ = oracle.render(m, mimo, kn, fc, prec=1, count=n, threads=1, record_paths=256)
    assert tw == n
    fc2 = H.FrameConfig(1, 256, 2)
    for t, tx in enumerate([sc.tx, tx1]):
        single = H.Scene(rfConfig=rf, tx=tx, rx=sc.rx, rxs=sc.rxs)
        zs, tws, recs, _ = oracle.render(m, single, kn, fc2, prec=1, count=n, threads=1, record_paths=256)
        assert np.array_equal(z[..., 2 * t:2 * t + 2], zs)
        assert np.array_equal(rec[1][..., 2 * t:2 * t + 2], recs[1])
    assert np.abs(z).max() > 0
