"""Host-only checks of the per-mesh setup under AddressSanitizer and
UndefinedBehaviorSanitizer (tests/host/setup_check.cpp, built with g++; no
GPU code): build_setup's products pass the structural validator (pattern =
the bending stencils' union, indices in range, level schedules that respect
every dependency, LDS ring offsets inside the ring, well-formed BVH), planted
mistakes are rejected, and the on-disk setup cache rejects a file with one
flipped payload byte and rebuilds it."""
import os
import shutil
import subprocess

import numpy as np
import pytest

from cloth_sim_self_collision_amd import scene

HERE = os.path.dirname(os.path.abspath(__file__))
HOST = os.path.join(HERE, "host")


@pytest.fixture(scope="module")
def checker():
    if shutil.which("g++") is None or shutil.which("make") is None:
        pytest.skip("g++/make not available")
    r = subprocess.run(["make", "-s", "-C", HOST], capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr
    return os.path.join(HOST, "bin", "setup_check")


def _write_mesh(path, cl):
    with open(path, "wb") as f:
        f.write(np.array([cl.np_, cl.nt, cl.nq], np.int32).tobytes())
        f.write(np.ascontiguousarray(cl.xyz, np.float64).tobytes())
        f.write(np.ascontiguousarray(cl.tri, np.int32).tobytes())
        f.write(np.ascontiguousarray(cl.quad, np.int32).tobytes())


def _irregular():
    """A grid with jittered vertices and a ragged boundary (triangles of a
    few cells removed), so the pattern has rows of unequal length."""
    cl = scene.drape_over_sphere(12)
    keep = np.ones(cl.nt, bool)
    keep[[0, 1, 7, 30, 31]] = False
    tri = cl.tri.reshape(-1, 3)[keep]
    used = np.unique(tri)
    remap = -np.ones(cl.np_, np.int64)
    remap[used] = np.arange(len(used))
    tri = remap[tri].astype(np.int32)
    xyz = cl.xyz.reshape(-1, 3)[used].reshape(-1)
    rng = np.random.default_rng(3)
    xyz = xyz + rng.uniform(-5e-3, 5e-3, xyz.shape)
    quad = scene.bending_quads(tri.reshape(-1))

    class M:
        pass
    m = M()
    m.np_, m.nt, m.nq = len(used), len(tri), len(quad) // 4
    m.xyz, m.tri, m.quad = xyz, tri.reshape(-1), quad
    return m


@pytest.mark.parametrize("name", ["drape32", "fold40", "stack3x16", "irregular"])
def test_setup_validates_and_cache_rejects_corruption(checker, tmp_path, name):
    cl = {"drape32": lambda: scene.drape_over_sphere(32),
          "fold40": lambda: scene.fold_sweep(40, size=0.5),
          "stack3x16": lambda: scene.stacked_sheets(3, 16),
          "irregular": _irregular}[name]()
    mesh = tmp_path / "mesh.bin"
    _write_mesh(mesh, cl)
    cache = tmp_path / "cache"
    cache.mkdir()
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0",
               UBSAN_OPTIONS="print_stacktrace=1")
    r = subprocess.run([checker, str(mesh), str(cache)], capture_output=True, text=True,
                       timeout=300, env=env)
    assert r.returncode == 0 and r.stdout.startswith("ok"), r.stdout + r.stderr
