"""Oracle pins for ray queries: the BVH closest hit and occlusion equal brute force over all
triangles (the Embree replacement of cpu.cpp:53-87), including rays aimed at vertices and
edges, in FP32 and FP64."""
import numpy as np
import pytest

from hussar_amd import scenes


def random_soup(n, seed):
    rng = np.random.default_rng(seed)
    c = rng.uniform(-1, 1, (n, 1, 3))
    V = (c + rng.normal(scale=0.08, size=(n, 3, 3))).reshape(-1, 3).astype(np.float32)
    I = np.arange(3 * n, dtype=np.int32).reshape(-1, 3)
    return V, I


def rays_for(V, I, n, seed):
    rng = np.random.default_rng(seed)
    o = rng.uniform(-2, 2, (n, 3))
    # half random directions, half aimed at vertices / edge midpoints / centroids
    tgt = V[I[rng.integers(0, I.shape[0], n)]].astype(np.float64)
    w = rng.choice(4, n)
    bary = np.array([[1, 0, 0], [0.5, 0.5, 0], [1 / 3, 1 / 3, 1 / 3], [0.2, 0.3, 0.5]])[w]
    aim = np.einsum("nk,nkj->nj", bary, tgt)
    d = np.where(rng.random((n, 1)) < 0.5, rng.normal(size=(n, 3)), aim - o)
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    return np.concatenate([o, d], axis=1).astype(np.float32)


@pytest.mark.parametrize("prec", [0, 2])
def test_bvh_closest_equals_brute_force(oracle, prec):
    V, I = random_soup(3000, 1)
    m = oracle.OracleMesh(V, I)
    R = rays_for(V, I, 20000, 2)
    ib, tb, _ = m.intersect(R, prec=prec, brute=False)
    ia, ta, _ = m.intersect(R, prec=prec, brute=True)
    np.testing.assert_array_equal(ib, ia)
    hit = ia >= 0
    assert hit.mean() > 0.3
    np.testing.assert_array_equal(tb[hit], ta[hit])


@pytest.mark.parametrize("prec", [0, 2])
def test_bvh_occlusion_equals_brute_force(oracle, prec):
    V, I = random_soup(3000, 3)
    m = oracle.OracleMesh(V, I)
    R = rays_for(V, I, 20000, 4)
    for tfar in (0.5, 2.0, np.inf):
        ob_, _, _ = m.intersect(R, prec=prec, brute=False, any_hit=True, tfar=tfar)
        oa, _, _ = m.intersect(R, prec=prec, brute=True, any_hit=True, tfar=tfar)
        np.testing.assert_array_equal(ob_, oa)
        # occluded within tfar  <=>  the closest hit lies before tfar
        ic, tc, _ = m.intersect(R, prec=prec, brute=True)
        np.testing.assert_array_equal(oa.astype(bool), (ic >= 0) & (tc < tfar))


def test_bvh_on_vehicle_clutter(oracle):
    mesh, centre = scenes.vehicle_mesh(seed=1, triangles=20000)
    V, I = mesh.vertexBuffer, mesh.indexBuffer
    m = oracle.OracleMesh(V, I)
    rng = np.random.default_rng(5)
    n = 4000
    o = np.tile(np.array([0.0, 0.75, 0.0]), (n, 1)) + rng.normal(scale=0.1, size=(n, 3))
    tgt = centre + rng.uniform(-1, 1, (n, 3)) * np.array([1.0, 0.75, 2.5])
    d = tgt - o
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    R = np.concatenate([o, d], axis=1).astype(np.float32)
    ib, tb, cnt = m.intersect(R, prec=0)
    ia, ta, _ = m.intersect(R, prec=0, brute=True)
    np.testing.assert_array_equal(ib, ia)
    assert (ia >= 0).mean() > 0.5
    # traversal visits far fewer triangles than brute force
    assert cnt[1] < 0.01 * n * I.shape[0]


def test_tnear_excludes_the_origin_surface(oracle):
    # a ray leaving a plate from its surface must not hit the plate again (Epsilon = 5e-5)
    V = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], np.float32)
    I = np.array([[0, 1, 2]], np.int32)
    m = oracle.OracleMesh(V, I)
    R = np.array([[0.2, 0.2, 0.0, 0, 0, 1], [0.2, 0.2, 1e-4, 0, 0, -1], [0.2, 0.2, 4e-5, 0, 0, -1]], np.float32)
    ids, t, _ = m.intersect(R)
    assert ids.tolist() == [-1, 0, -1]
    assert t[1] == pytest.approx(1e-4, rel=1e-3)
