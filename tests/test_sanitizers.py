"""ThreadSanitizer runs, on the CPU, of the host code that starts threads:
  * the oracle's thread pool (the restatement of Backend::run's pool, cpu.h:21-41): its
    workers share the batch counter under a mutex and fill their own frames, guiding
    statistics, records and counters, which are summed after the join. oracle/tsan_check.cpp
    drives it through a guided schedule and a MIMO render with records on 8 workers;
  * the product's BVH builder (hussar_amd/csrc/bvh.cpp), whose spatial-split build hands the
    larger subtrees to a second thread (from 2^15 references on): built with the host BVH
    check on the 40k-triangle clutter mesh, which takes that path at the root.
Either run fails on a ThreadSanitizer report (halt_on_error). No GPU needed."""
import os
import subprocess

import numpy as np
import pytest

from hussar_amd import scenes

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ENV = dict(os.environ, TSAN_OPTIONS="halt_on_error=1")


def tsan_available():
    probe = subprocess.run(["g++", "-fsanitize=thread", "-x", "c++", "-", "-o", os.devnull],
                           input="int main(){return 0;}", capture_output=True, text=True)
    return probe.returncode == 0


pytestmark = pytest.mark.skipif(not tsan_available(), reason="no ThreadSanitizer runtime for g++ here")


def test_oracle_thread_pool_under_tsan():
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "tsan_check"], check=True)
    r = subprocess.run([os.path.join(ROOT, "oracle", "tsan_check")], capture_output=True, text=True, timeout=600,
                       env=ENV)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "ThreadSanitizer" not in r.stderr, r.stderr
    assert "schedule" in r.stdout and "nodes" in r.stdout


def test_bvh_builder_threads_under_tsan(tmp_path):
    exe = str(tmp_path / "bvh_tsan")
    subprocess.run(["g++", "-std=c++17", "-O1", "-g", "-ffp-contract=off", "-fsanitize=thread",
                    "-I/opt/rocm/include", "-D__HIP_PLATFORM_AMD__",
                    os.path.join(ROOT, "tests", "cpp", "bvh_brute_force.cpp"),
                    os.path.join(ROOT, "hussar_amd", "csrc", "bvh.cpp"), "-o", exe, "-lpthread"], check=True)
    mesh = scenes.vehicle_mesh(seed=1, triangles=40_000)[0]
    path = str(tmp_path / "mesh.bin")
    with open(path, "wb") as f:
        np.array([len(mesh.vertexBuffer), len(mesh.indexBuffer)], np.int64).tofile(f)
        mesh.vertexBuffer.tofile(f)
        mesh.indexBuffer.tofile(f)
    r = subprocess.run([exe, path, "500"], capture_output=True, text=True, timeout=900, env=ENV)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "ThreadSanitizer" not in r.stderr, r.stderr
    vals = dict(zip(r.stdout.split()[0::2], map(int, r.stdout.split()[1::2])))
    assert vals["mismatches"] == 0 and vals["missing"] == 0
