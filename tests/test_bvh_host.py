"""The product's BVH builder (hussar_amd/csrc/bvh.cpp: SAH with clipped references of large
triangles and spatial splits) compiled on the host with g++: closest hits through its
quantised four-wide nodes equal brute force over every slot on the vehicle clutter mesh, and
every triangle is reachable (its slots cover it). No GPU needed."""
import os
import subprocess

import numpy as np
import pytest

from hussar_amd import scenes
from hussar_amd.types import TriangleMesh

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "cpp", "bvh_brute_force.cpp")
BVH = os.path.join(ROOT, "hussar_amd", "csrc", "bvh.cpp")


@pytest.fixture(scope="module")
def checker(tmp_path_factory):
    exe = str(tmp_path_factory.mktemp("bvh") / "bvh_brute_force")
    subprocess.run(["g++", "-std=c++17", "-O2", "-ffp-contract=off", "-I/opt/rocm/include", "-D__HIP_PLATFORM_AMD__",
                    SRC, BVH, "-o", exe, "-lpthread"], check=True)
    return exe


def write_mesh(path, mesh):
    v, i = mesh.vertexBuffer, mesh.indexBuffer
    with open(path, "wb") as f:
        np.array([len(v), len(i)], np.int64).tofile(f)
        v.tofile(f)
        i.tofile(f)
    return len(i)


@pytest.mark.parametrize("triangles", [2_000, 40_000])
def test_bvh_equals_brute_force(checker, tmp_path, triangles):
    mesh = scenes.vehicle_mesh(seed=1, triangles=triangles)[0]
    path = str(tmp_path / "mesh.bin")
    nt = write_mesh(path, mesh)
    out = subprocess.run([checker, path, "6000"], check=True, capture_output=True, text=True, timeout=300).stdout
    f = out.split()
    vals = dict(zip(f[0::2], map(int, f[1::2])))
    assert vals["mismatches"] == 0, out
    assert vals["missing"] == 0, out
    assert vals["hits"] > 3000, out
    assert vals["slots"] > nt, out  # large triangles are split into clipped references here


def small_mesh(kind):
    """Meshes of one to four triangles whose best split by SAH is no split at all: the root
    of the spatial-split build is then a leaf, which the builder must still wrap in an inner
    root node."""
    if kind == "triangle":
        return TriangleMesh(np.array([[0, 0, 1], [1, 0, 1], [0, 1, 1]], np.float32), np.array([[0, 1, 2]], np.int32))
    m = TriangleMesh()
    m.addQuad([-0.5, -0.5, 1.0], [1.0, 0.0, 0.0], [0.0, 1.0, 0.0])
    if kind == "two_coplanar_quads":  # overlapping, in one plane
        m.addQuad([-0.25, -0.25, 1.0], [1.0, 0.0, 0.0], [0.0, 1.0, 0.0])
    elif kind == "dihedral":
        m = scenes.dihedral_mesh()
    return m


@pytest.mark.parametrize("kind", ["triangle", "quad", "two_coplanar_quads", "dihedral"])
def test_bvh_small_meshes(checker, tmp_path, kind):
    mesh = small_mesh(kind)
    path = str(tmp_path / "mesh.bin")
    nt = write_mesh(path, mesh)
    r = subprocess.run([checker, path, "3000"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    f = r.stdout.split()
    vals = dict(zip(f[0::2], map(int, f[1::2])))
    assert vals["mismatches"] == 0, r.stdout
    assert vals["missing"] == 0, r.stdout
    assert vals["hits"] > 0, r.stdout
    assert vals["slots"] >= nt, r.stdout


def test_bvh_empty_mesh(checker, tmp_path):
    path = str(tmp_path / "empty.bin")
    with open(path, "wb") as f:
        np.array([0, 0], np.int64).tofile(f)
    out = subprocess.run([checker, path, "10"], check=True, capture_output=True, text=True).stdout
    assert "mismatches 0" in out
