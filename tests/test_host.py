"""Host-side types and scene generators: TriangleMesh::addQuad / addBox semantics
(mesh.h:30-54), NFAntenna's rotation check (emitter.h:168-171) and the seeded workloads."""
import numpy as np
import pytest

import hussar_amd as H
from hussar_amd import scenes


def test_add_quad_vertex_order():
    m = H.TriangleMesh()
    m.addQuad([1, 2, 3], [1, 0, 0], [0, 2, 0])
    np.testing.assert_array_equal(m.vertexBuffer, [[1, 2, 3], [2, 2, 3], [2, 4, 3], [1, 4, 3]])
    np.testing.assert_array_equal(m.indexBuffer, [[0, 1, 2], [0, 2, 3]])


def test_add_box_is_closed_and_outward_consistent():
    m = H.TriangleMesh()
    m.addBox([0, 0, 0], [1, 2, 3])
    V, I = m.vertexBuffer.astype(np.float64), m.indexBuffer
    assert I.shape == (12, 3) and V.shape == (24, 3)
    # total signed volume of a closed surface is ±(box volume)
    vol = np.einsum("ij,ij->i", V[I[:, 0]], np.cross(V[I[:, 1]], V[I[:, 2]])).sum() / 6
    assert abs(abs(vol) - 6.0) < 1e-9
    # every edge appears twice
    e = np.concatenate([I[:, [0, 1]], I[:, [1, 2]], I[:, [2, 0]]])
    keys = sorted(tuple(sorted((tuple(V[a]), tuple(V[b])))) for a, b in e)
    _, counts = np.unique(np.array(keys).reshape(len(keys), -1), axis=0, return_counts=True)
    assert np.all(counts == 2)


def test_nfantenna_rejects_non_rotations():
    H.NFAntenna([0, 0, 0], np.eye(3))
    with pytest.raises(ValueError):
        H.NFAntenna([0, 0, 0], 2 * np.eye(3))


def test_splitmix64_reference_values():
    # SplitMix64 with seed 0: published first outputs
    z = scenes.splitmix64(0, 3)
    assert [int(v) for v in z] == [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F]


def test_vehicle_mesh_deterministic_and_exact():
    a, ca = scenes.vehicle_mesh(seed=1, triangles=20000)
    b, cb = scenes.vehicle_mesh(seed=1, triangles=20000)
    assert a.triangle_count() == 20000
    np.testing.assert_array_equal(a.vertexBuffer, b.vertexBuffer)
    V = a.vertexBuffer
    assert 10.0 <= ca[2] <= 30.0
    lo = np.array([-1.0, 0.0, ca[2] - 2.5]) - 1e-5
    hi = np.array([1.0, 1.5, ca[2] + 2.5]) + 1e-5
    assert np.all(V >= lo) and np.all(V <= hi)
    c, _ = scenes.vehicle_mesh(seed=2, triangles=20000)
    assert not np.array_equal(c.vertexBuffer, V)


def test_dihedral_pose_geometry():
    sc = scenes.dihedral_pose(0.0)
    p = sc.tx.position.astype(np.float64)
    # on the bisector of the two plates, 1.5 m from the fold line, boresight towards the corner
    assert abs(np.hypot(p[0], p[2]) - 1.5) < 1e-6 and abs(p[0] - p[2]) < 1e-6
    bore = sc.tx.rotation.astype(np.float64)[:, 2]
    np.testing.assert_allclose(bore, -np.array([1, 0, 1]) / np.sqrt(2), atol=1e-6)
    assert len(scenes.dihedral_sweep()) == 91


def test_radar_array_spacing():
    rf = H.RFConfig()
    sc = scenes.radar_array(rf, channels=4)
    rx = np.array([a.position for a in sc.receivers()], np.float64)
    lam = H.SPEED_OF_LIGHT / rf.startFreq
    np.testing.assert_allclose(np.diff(rx[:, 0]), lam / 2, rtol=1e-4)
    assert sc.to_c().rxCount == 4
