"""The geometry of the primary visibility lists (hussar_amd/csrc/primary.h) compiled on the host
with g++: lists built from its distance keys, face footprints, face grids and cell rectangles
give, for rays from the origin through them (near triangles first, then the ray's cell, up to
the first key beyond the best hit), the closest hits of brute force over every triangle slot
with the same float triangle test and tie rule. Origins: the radar in front of the vehicle
clutter, a point inside the clutter (triangles all around, one within the near distance), the
dihedral's radar; cell resolutions 4 to 80. No GPU needed. (The device path that uses these
lists is disabled in this build; these checks pin its geometry.)"""
import os
import subprocess

import numpy as np
import pytest

from hussar_amd import scenes

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "cpp", "primary_lists.cpp")
BVH = os.path.join(ROOT, "hussar_amd", "csrc", "bvh.cpp")


@pytest.fixture(scope="module")
def checker(tmp_path_factory):
    exe = str(tmp_path_factory.mktemp("pv") / "primary_lists")
    subprocess.run(["g++", "-std=c++17", "-O2", "-ffp-contract=off", SRC, BVH, "-o", exe, "-lpthread"], check=True)
    return exe


def write_mesh(path, mesh):
    v, i = mesh.vertexBuffer, mesh.indexBuffer
    with open(path, "wb") as f:
        np.array([len(v), len(i)], np.int64).tofile(f)
        v.tofile(f)
        i.tofile(f)


def run(checker, path, origin, res, rays):
    out = subprocess.run([checker, path, *map(str, origin), str(res), str(rays)], check=True, capture_output=True,
                         text=True, timeout=600)
    f = out.stdout.split()
    return {k: float(v) for k, v in zip(f[0::2], f[1::2])}, out.stdout + out.stderr


@pytest.mark.parametrize("origin,res", [((0.0, 0.75, 0.0), 64), ((0.0, 0.75, 0.0), 4), ((0.0, 0.75, 0.0), 80),
                                        ((0.3, 0.8, 20.0), 64)])
def test_lists_equal_brute_force_vehicle(checker, tmp_path, origin, res):
    mesh = scenes.vehicle_mesh(seed=1, triangles=10_000)[0]
    path = str(tmp_path / "mesh.bin")
    write_mesh(path, mesh)
    vals, text = run(checker, path, origin, res, 12000)
    assert vals["mismatches"] == 0, text
    assert vals["hits"] > 6000, text
    assert vals["tested"] < 40, text  # the early exit keeps the candidates per ray few


def test_lists_near_triangle(checker, tmp_path):
    """The origin 5 mm in front of a panel: that panel is on the near list of every ray."""
    mesh = scenes.vehicle_mesh(seed=3, triangles=4_000)[0]
    mesh.addQuad([-0.002, 0.748, 0.005], [0.004, 0.0, 0.0], [0.0, 0.004, 0.0])
    path = str(tmp_path / "mesh.bin")
    write_mesh(path, mesh)
    vals, text = run(checker, path, (0.0, 0.75, 0.0), 64, 9000)
    assert vals["mismatches"] == 0, text
    assert vals["near"] >= 2, text


def test_lists_dihedral(checker, tmp_path):
    path = str(tmp_path / "mesh.bin")
    write_mesh(path, scenes.dihedral_mesh())
    vals, text = run(checker, path, (1.0, 0.05, 1.0), 64, 9000)
    assert vals["mismatches"] == 0, text
    assert vals["hits"] > 1000, text
