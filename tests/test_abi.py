"""Host-side checks without a GPU: the C ABI library loads and exports every declared symbol, the scene
generator is deterministic and follows the reference's storage conventions, the Python view has the
reference's memory layout, and the product path refuses to run without its native library."""
import ctypes as C
import os
import re
import subprocess

import numpy as np
import pytest

import cuvslam_amd as ca
from cuvslam_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = "".join(open(os.path.join(ROOT, "include", h)).read() for h in ("cuvslam_sba.h", "cuvslam_sba_imu.h",
                                                                    "cuvslam_sba_mono.h"))
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(sba_[a-z_0-9]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    L = _lib.lib()
    syms = declared_symbols()
    assert len(syms) >= 20
    for s in syms:
        assert hasattr(L, s), s
    # every declared symbol has a ctypes signature, and nothing undeclared is bound
    assert set(syms) == set(_lib.SIGNATURES)
    out = subprocess.run(["nm", "-D", "--defined-only", _lib.LIB_PATH], capture_output=True, text=True).stdout
    exported = set(re.findall(r" T (sba_[a-z_0-9]+)$", out, flags=re.M))
    assert set(syms) <= exported
    assert _lib.lib().sba_abi_version() == 1


def test_cpp_dropin_symbols_exported():
    out = subprocess.run(["nm", "-DC", "--defined-only", _lib.LIB_PATH], capture_output=True, text=True).stdout
    assert "cuvslam::sba::SchurComplementBundlerGpu::solve(cuvslam::sba::BundleAdjustmentProblem&)" in out
    assert "cuvslam::sba::SchurComplementBundlerGpu::SchurComplementBundlerGpu()" in out


def test_create_without_gpu_fails_cleanly():
    if os.path.exists("/dev/kfd"):
        pytest.skip("a GPU is visible")
    with pytest.raises(ca.SbaError) as e:
        ca.SchurComplementBundlerGpu()
    assert e.value.code == _lib.SBA_E_DEVICE


def test_missing_library_fails_loudly(monkeypatch):
    monkeypatch.setattr(_lib, "_lib", None)
    monkeypatch.setattr(_lib, "LIB_PATH", "/nonexistent/libcuvslam_sba.so")
    with pytest.raises(RuntimeError, match="missing"):
        _lib.lib()


def test_scene_is_deterministic_and_seeded():
    a, ga, qa = ca.scene.generate("c1", seed=7)
    b, gb, qb = ca.scene.generate("c1", seed=7)
    c, _, _ = ca.scene.generate("c1", seed=8)
    for x, y in [(a.points, b.points), (a.observation_xys, b.observation_xys), (a.rig_from_world, b.rig_from_world),
                 (a.point_ids, b.point_ids), (ga, gb), (qa, qb)]:
        assert np.array_equal(x, y)
    assert not np.array_equal(a.observation_xys[:100], c.observation_xys[:100])


@pytest.mark.parametrize("nf", [1, 2])
def test_scene_storage_conventions(nf):
    p, g, q = ca.scene.generate("c1", seed=3, num_fixed_keyframes=nf)
    n = p.rig_from_world.shape[0]
    assert p.num_fixed_key_frames == nf and n == 10
    # observations ordered by key frame, then by point id (bundle_adjustment_problem.h:31-33)
    key = p.pose_ids.astype(np.int64) * 100000 + p.point_ids
    assert np.all(np.diff(key) > 0)
    # fixed (gauge) key frames are stored last and unperturbed; the first generated key frame sits at the origin
    np.testing.assert_array_equal(p.rig_from_world[n - nf:], q[n - nf:])
    np.testing.assert_allclose(q[n - nf][:3, 3], 0, atol=1e-7)
    assert np.all(np.abs(p.rig_from_world[: n - nf] - q[: n - nf]).max(axis=(1, 2)) > 0)
    # the information maps 1 px to 1 (Huber scale 1 = 1 px): diag(fx^2, fy^2)
    np.testing.assert_allclose(p.observation_infos[0], np.diag([450.0 ** 2, 450.0 ** 2]))
    # observation count of the benchmark geometry (points projecting in-image in 10 key frames)
    assert 15000 < p.num_observations < 20000


def test_scene_outliers_and_noise_statistics():
    p, g, q = ca.scene.generate("c1", seed=11, num_points=4000)
    pts = g[p.point_ids].astype(np.float64)
    T = q[p.pose_ids].astype(np.float64)
    pc = np.einsum("nij,nj->ni", T[:, :3, :3], pts) + T[:, :3, 3]
    err_px = (p.observation_xys - pc[:, :2] / pc[:, 2:3]) * 450.0
    big = np.linalg.norm(err_px, axis=1) > 5.0
    assert 0.03 < big.mean() < 0.07  # 5 % outliers (a few land within 5 px)
    assert np.std(err_px[~big]) == pytest.approx(0.5, rel=0.1)


def test_batched_windows_shapes():
    ws = ca.scene.batched_windows(6, seed=3)
    for w in ws:
        nk = w.rig_from_world.shape[0]
        assert 5 <= nk <= 40 and 500 <= w.points.shape[0] <= 8000
        counts = np.bincount(w.point_ids, minlength=w.points.shape[0])
        seen = counts[counts > 0]
        assert seen.min() >= 2 and seen.max() <= nk


def test_problem_view_layout_is_the_reference_memory_layout():
    p, _, _ = ca.scene.generate(seed=1, num_points=50, num_keyframes=3)
    p.observation_infos[0] = [[1.0, 2.0], [3.0, 4.0]]
    v = p.view()
    info = np.ctypeslib.as_array(v.observation_infos, shape=(4,))
    np.testing.assert_array_equal(info, [1.0, 3.0, 2.0, 4.0])  # column-major Matrix2f
    m = np.ctypeslib.as_array(v.rig_from_world, shape=(16,))
    np.testing.assert_array_equal(m[12:15], p.rig_from_world[0, :3, 3])  # column-major Isometry3f translation
    assert v.num_fixed_key_frames == p.num_fixed_key_frames and v.max_iterations == p.max_iterations


def test_cpp_dropin_example_builds():
    exe = os.path.join(ROOT, "tests", "cpp", "dropin_solve")
    assert os.path.exists(exe), "built by __graft_entry__.build()"
