"""Closed-form and hand-worked pins of the two oracle pieces that parity against the GPU alone
cannot pin: the NEE weighting at a bounce vertex (GO density correction, phase error and the
plane / soft-sphere filters, libhussar/include/hussar/integrators/path.h:330-394) and the guide
tree's build and refine (GuidingWrapper::step, path.h:119-126). The guide step is checked on the
oracle and on the product's host code (hussar_amd/csrc/guide.cpp, compiled here with g++).

Geometry of the filter pins: a ray from the transmitter at the origin along +z hits a plate
facing it at distance r (vertex (0, 0, r), normal -z, reflected direction -z); the receiver
sits at distance R from the vertex, at angle phi off the mirror direction:
rx = (0, 0, r) + R (sin phi, 0, -cos phi). Then
  * the sphere filter's distance from the receiver to the reflected ray is R sin phi;
  * the plane filter's half vector normalize(dn - d_in) makes the angle phi/2 with the normal,
    so its phase error is r tan(phi/2).
Both in wavelengths. No GPU needed."""
import math
import os
import subprocess

import numpy as np
import pytest

import hussar_amd as H

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LAM = 299792458.0 / 77e9  # the wavelength passed in; any positive value works


def vertex_geometry(r, R, phi):
    o = np.array([0.0, 0.0, r])
    d_in = np.array([0.0, 0.0, 1.0])
    n_in = np.array([0.0, 0.0, -1.0])
    d = np.array([0.0, 0.0, -1.0])
    dn = np.array([math.sin(phi), 0.0, -math.cos(phi)])
    return dict(o=o, d=d, dn=dn, rr=R, d_in=d_in, n_in=n_in, r=r)


def weigh(oracle, knobs, g, v=1.0 + 0.5j, cos_theta=0.8, prec=None):
    prec = oracle.PREC_FP64 if prec is None else prec
    return oracle.nee_weight(knobs, g["o"], g["d"], g["dn"], g["rr"], g["d_in"], g["n_in"], g["r"], cos_theta,
                             LAM, v, prec=prec)


def go_factor(r, cos_theta):
    # path.h:334-336: sampled with density cos/r^2, wanted 1/(4 pi r)
    return r / cos_theta / (4.0 * math.pi)


@pytest.mark.parametrize("phi_deg", [5.0, 30.0, 60.0])
def test_phase_error_closed_forms(oracle, phi_deg):
    phi = math.radians(phi_deg)
    g = vertex_geometry(r=7.0, R=3.0, phi=phi)
    off = H.Knobs(doFiltering=False)
    _, _, _, dp_sphere = weigh(oracle, H.Knobs(doFiltering=False, filteringSphere=True), g)
    _, _, _, dp_plane = weigh(oracle, H.Knobs(doFiltering=False, filteringSphere=False), g)
    assert dp_sphere == pytest.approx(3.0 * math.sin(phi) / LAM, rel=1e-12)
    assert dp_plane == pytest.approx(7.0 * math.tan(phi / 2) / LAM, rel=1e-9)
    # without filtering only the GO density correction applies, and gw gains v
    kept, v, gw, _ = weigh(oracle, off, g, v=2.0 - 1.0j, cos_theta=0.6)
    assert kept
    assert v == pytest.approx((2.0 - 1.0j) * go_factor(7.0, 0.6), rel=1e-12)
    assert gw == pytest.approx(v, rel=1e-12)


def test_grazing_vertex_is_skipped(oracle):
    g = vertex_geometry(r=7.0, R=3.0, phi=0.3)
    kept, _, gw, _ = weigh(oracle, H.Knobs(), g, cos_theta=0.999e-3)
    assert not kept and gw == 0
    kept, _, _, _ = weigh(oracle, H.Knobs(), g, cos_theta=1e-3)
    assert kept


# (dphase / filteringRadius, expected factor): 1 inside the sphere, then 1 - 0.2 (x - 1), then 0
SPHERE_CASES = [(0.5, 1.0), (1.0, 1.0), (1.5, 0.9), (3.5, 0.5), (5.5, 0.1), (7.0, 0.0)]


@pytest.mark.parametrize("x,factor", SPHERE_CASES)
def test_soft_sphere_falloff(oracle, x, factor):
    radius = 160.0
    phi = math.radians(30.0)
    R = x * radius * LAM / math.sin(phi)  # so that R sin(phi) / lambda = x * radius
    g = vertex_geometry(r=12.0, R=R, phi=phi)
    kept, v, gw, dp = weigh(oracle, H.Knobs(filteringSphere=True, filteringRadius=radius), g, v=1.0 + 0.5j,
                            cos_theta=0.8)
    assert kept
    assert dp == pytest.approx(x * radius, rel=1e-12)
    expect = (1.0 + 0.5j) * go_factor(12.0, 0.8) * factor
    assert v == pytest.approx(expect, rel=1e-12, abs=1e-15)
    assert gw == pytest.approx(expect, rel=1e-12, abs=1e-15)


# (dphase, factor, guiding term as a multiple of the corrected v) for filteringMin 10,
# filteringMax 50, filterGuiding on: below min 1; between a ramp (50 - dphase) / 40; beyond max
# zero, with v / ((dphase / 50)^2 + 1) going to the guiding weight
PLANE_CASES = [(5.0, 1.0, 1.0), (30.0, 0.5, 0.5), (45.0, 0.125, 0.125), (60.0, 0.0, 1.0 / 2.44),
               (100.0, 0.0, 1.0 / 5.0)]


@pytest.mark.parametrize("dphase,factor,gterm", PLANE_CASES)
def test_plane_filter(oracle, dphase, factor, gterm):
    r = 9.0
    phi = 2.0 * math.atan(dphase * LAM / r)  # so that r tan(phi/2) / lambda = dphase
    g = vertex_geometry(r=r, R=4.0, phi=phi)
    kn = H.Knobs(filteringSphere=False, filteringMin=10.0, filteringMax=50.0, filterGuiding=True)
    kept, v, gw, dp = weigh(oracle, kn, g, v=1.0 + 0.5j, cos_theta=0.8)
    assert kept
    # sqrt(1 - c^2) of a cosine near 1 keeps about 11 digits
    assert dp == pytest.approx(dphase, rel=1e-9)
    base = (1.0 + 0.5j) * go_factor(r, 0.8)
    assert v == pytest.approx(base * factor, rel=1e-9, abs=1e-15)
    assert gw == pytest.approx(base * gterm, rel=1e-9, abs=1e-15)
    # without filterGuiding the cut contribution leaves the guiding weight alone
    kn.filterGuiding = False
    _, v2, gw2, _ = weigh(oracle, kn, g, v=1.0 + 0.5j, cos_theta=0.8)
    assert v2 == pytest.approx(base * factor, rel=1e-9, abs=1e-15)
    assert gw2 == pytest.approx(base * factor, rel=1e-9, abs=1e-15)


def test_float_instance_agrees(oracle):
    g = vertex_geometry(r=12.0, R=2.0, phi=0.2)
    for kn in (H.Knobs(), H.Knobs(filteringSphere=False, filteringMin=10.0, filteringMax=50.0)):
        a = weigh(oracle, kn, g)
        b = weigh(oracle, kn, g, prec=oracle.PREC_REF32)
        assert a[0] == b[0]
        assert b[1] == pytest.approx(a[1], rel=1e-4)
        assert b[3] == pytest.approx(a[3], rel=1e-4)


# ---------------------------------------------------------------------------------------
# Guide step. A tree is (child, stat): child[i] = 0 for a leaf, else the index of the first of
# two consecutive children; node i's cell is halved along axis depth(i) % 2.

@pytest.fixture(scope="module")
def product_step(tmp_path_factory):
    exe = str(tmp_path_factory.mktemp("guide") / "guide_step")
    subprocess.run(["g++", "-std=c++17", "-O2", "-ffp-contract=off", os.path.join(ROOT, "tests", "cpp", "guide_step.cpp"),
                    os.path.join(ROOT, "hussar_amd", "csrc", "guide.cpp"), "-o", exe], check=True)

    def run(child, stat, threshold, second_moment=True):
        inp = f"{threshold!r} {int(second_moment)} {len(child)}\n" + " ".join(map(str, child)) + "\n" + \
            " ".join(repr(float(s)) for s in stat) + "\n"
        out = subprocess.run([exe], input=inp, capture_output=True, text=True, check=True).stdout.split("\n")
        m = int(out[0].split()[1])
        s_child = np.array(out[1].split(), np.int32)
        s_p0 = np.array(out[2].split(), np.float32)
        k = int(out[3].split()[1])
        t_child = np.array(out[4].split(), np.int32)
        t_depth = np.array(out[5].split(), np.int32)
        t_stat = np.array(out[6].split(), np.float64)
        assert len(s_child) == m and len(t_child) == k
        return s_child, s_p0, t_child, t_depth, t_stat
    return run


def oracle_step(oracle, child, stat, threshold, second_moment=True):
    g = oracle.OracleGuide(uniformProb=0.1, splitThreshold=threshold, secondMoment=second_moment)
    g.set_training(child, stat)
    g.step()
    s_child, s_p0, _ = g.get(0)
    t_child, _, t_stat = g.get(1)
    return s_child, s_p0, t_child, depths(t_child), t_stat


def depths(child):
    d = np.zeros(len(child), np.int32)
    for i, c in enumerate(child):
        if c:
            d[c] = d[c + 1] = d[i] + 1
    return d


def cells(child):
    """Leaf cells [x0, x1) x [y0, y1) per node, by walking the tree."""
    out = {0: (0.0, 1.0, 0.0, 1.0, 0)}
    for i, c in enumerate(child):
        if c:
            x0, x1, y0, y1, dep = out[i]
            if dep % 2 == 0:
                m = 0.5 * (x0 + x1)
                out[c], out[c + 1] = (x0, m, y0, y1, dep + 1), (m, x1, y0, y1, dep + 1)
            else:
                m = 0.5 * (y0 + y1)
                out[c], out[c + 1] = (x0, x1, y0, m, dep + 1), (x0, x1, m, y1, dep + 1)
    return out


def leaf_pdf(child, p0, i):
    """Sampling density on leaf i: the product of 2 p0 / 2 (1 - p0) along its root path."""
    parent = {}
    for j, c in enumerate(child):
        if c:
            parent[c], parent[c + 1] = (j, 0), (j, 1)
    pdf = 1.0
    while i in parent:
        j, side = parent[i]
        pdf *= 2.0 * (p0[j] if side == 0 else 1.0 - p0[j])
        i = j
    return pdf


# A hand-worked tree: root split in x; its lower half split in y. Leaves 2 = [.5,1)x[0,1)
# (area 1/2), 3 = [0,.5)x[0,.5) and 4 = [0,.5)x[.5,1) (area 1/4), statistics 0.5, 4 and 1.
HAND_CHILD = [1, 3, 0, 0, 0]
HAND_STAT = [0.0, 0.0, 0.5, 4.0, 1.0]


@pytest.mark.parametrize("impl", ["oracle", "product"])
def test_guide_build_hand_worked(oracle, product_step, impl):
    step = (lambda *a, **k: oracle_step(oracle, *a, **k)) if impl == "oracle" else product_step
    # second moment: leaf masses sqrt(stat * area) = 0.5, 1, 0.5; inner 1.5 and 2
    s_child, p0, *_ = step(HAND_CHILD, HAND_STAT, 0.9, True)
    assert s_child.tolist() == HAND_CHILD
    assert p0[0] == np.float32(0.75) and p0[1] == np.float32(2.0 / 3.0)
    # the density on each leaf is sqrt(stat / area) / sum(sqrt(stat * area)): 1 / 2, 4 / 2, 2 / 2
    assert [leaf_pdf(s_child, p0, i) for i in (2, 3, 4)] == pytest.approx([0.5, 2.0, 1.0], rel=1e-6)
    # first moment: masses = statistics 0.5, 4, 1; the density is stat / area / sum(stat)
    s_child, p0, *_ = step(HAND_CHILD, HAND_STAT, 0.9, False)
    assert p0[0] == np.float32(5.0 / 5.5) and p0[1] == np.float32(0.8)
    assert [leaf_pdf(s_child, p0, i) for i in (2, 3, 4)] == pytest.approx([1 / 5.5, 16 / 5.5, 4 / 5.5], rel=1e-6)
    # a node without mass keeps p0 = 0.5
    s_child, p0, *_ = step(HAND_CHILD, [0, 0, 0.0, 0.0, 0.0], 0.9, True)
    assert p0.tolist() == [0.5] * 5


@pytest.mark.parametrize("impl", ["oracle", "product"])
def test_guide_refine_hand_worked(oracle, product_step, impl):
    step = (lambda *a, **k: oracle_step(oracle, *a, **k)) if impl == "oracle" else product_step
    # shares of the root mass 2: leaf 2 -> 0.25, leaf 3 -> 0.5, leaf 4 -> 0.25
    # threshold 0.2: leaves 2 and 4 split once (children at half share, 0.125); leaf 3 splits
    # into two children of share 0.25, which split again (0.125). In node order:
    # 2 -> 5, 6; 3 -> 7, 8; 4 -> 9, 10; 7 -> 11, 12; 8 -> 13, 14.
    _, _, t_child, t_depth, t_stat = step(HAND_CHILD, HAND_STAT, 0.2, True)
    assert t_child.tolist() == [1, 3, 5, 7, 9, 0, 0, 11, 13, 0, 0, 0, 0, 0, 0]
    assert t_depth.tolist() == [0, 1, 1, 2, 2, 2, 2, 3, 3, 3, 3, 4, 4, 4, 4]
    assert np.all(t_stat == 0.0)  # statistics cleared
    # threshold 0.25: a share equal to the threshold does not split (strictly greater splits):
    # only leaf 3 (0.5) splits, and its children (0.25) do not
    _, _, t_child, _, _ = step(HAND_CHILD, HAND_STAT, 0.25, True)
    assert t_child.tolist() == [1, 3, 0, 5, 0, 0, 0]
    # no mass: no refinement
    _, _, t_child, _, _ = step(HAND_CHILD, [0.0] * 5, 0.001, True)
    assert t_child.tolist() == HAND_CHILD


def random_tree(rng, splits):
    child = [0]
    for _ in range(splits):
        leaves = [i for i, c in enumerate(child) if c == 0]
        i = leaves[rng.integers(len(leaves))]
        child[i] = len(child)
        child += [0, 0]
    stat = np.where(np.array(child) == 0, rng.exponential(1.0, len(child)) * (rng.random(len(child)) < 0.8), 0.0)
    return child, stat


@pytest.mark.parametrize("seed", range(6))
@pytest.mark.parametrize("second_moment", [True, False])
def test_guide_step_properties_random_trees(oracle, product_step, seed, second_moment):
    """On random trees and statistics: the built density on every leaf is proportional to
    sqrt(stat / area) (second moment) or stat / area, and the refined tree replaces each old leaf
    of share s > t by a complete subtree of depth k = ceil(log2(s / t)) (halving shares); the
    product's step equals the oracle's."""
    rng = np.random.default_rng(seed)
    child, stat = random_tree(rng, 12)
    thr = float(np.float32(rng.uniform(0.01, 0.08)))
    o = oracle_step(oracle, child, stat, thr, second_moment)
    p = product_step(child, stat, thr, second_moment)
    for a, b in zip(o, p):
        np.testing.assert_array_equal(a, b)
    s_child, p0, t_child, t_depth, _ = o
    cl = cells(child)
    area = {i: (c[1] - c[0]) * (c[3] - c[2]) for i, c in cl.items()}
    leaves = [i for i, c in enumerate(child) if c == 0]
    mass = {i: (math.sqrt(stat[i] * area[i]) if second_moment else stat[i]) for i in leaves}
    total = sum(mass.values())
    for i in leaves:
        expect = mass[i] / area[i] / total
        assert leaf_pdf(s_child, p0, i) == pytest.approx(expect, rel=2e-6, abs=1e-12)
    # refine: leaves of the new tree below each old leaf
    new_cells = cells(t_child)
    for i in leaves:
        share = mass[i] / total
        k = 0
        while share / 2 ** k > np.float32(thr):
            k += 1
        x0, x1, y0, y1, dep = cl[i]
        under = [j for j, c in enumerate(t_child) if c == 0 and new_cells[j][0] >= x0 and new_cells[j][1] <= x1
                 and new_cells[j][2] >= y0 and new_cells[j][3] <= y1]
        assert len(under) == 2 ** k, (i, share, thr)
        assert all(t_depth[j] == dep + k for j in under)
