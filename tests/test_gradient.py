"""Pins on the oracle's reference gradients (CPU only).

The oracle differentiates the frame's detached-sampling estimator with
forward-mode dual numbers (FP64, and FP32 with the restatement's value
arithmetic).  These tests pin that machinery against what the mathematics
fixes, independently of any hand-derived derivative:
* central differences of the oracle's own FP64 objective J = sum(d_image * I)
  (the estimator is smooth along fixed paths, so FD converges to the
  derivative; a wrong dual rule, seed index or channel fails it);
* parameters no path evaluates leave J bit-identical under perturbation and
  get a zero gradient;
* J is linear in the environment radiance (every contribution carries exactly
  one environment texel and throughputs do not depend on it), so
  J = sum_e dJ/dE_e * E_e over all environment texels (Euler's theorem);
* base color is never a sampling input: writing it to the evaluation atlas
  alone renders bit-identically to writing it to both atlases, while
  roughness (a sampling input) changes the FP32 image only through sampling
  when written to both;
* the objective value equals sum(d_image * image) of the FP64 and FP32 renders.
"""
import numpy as np
import pytest

import pike_amd
from pike_amd import _ffi, scenes
from oracle.oracle import OracleScene

ATLAS, ENV = _ffi.PK_PARAM_ATLAS, _ffi.PK_PARAM_ENV
W, H, DEPTH, SPP = 20, 16, 3, 2
TEX, EW, EH = 16, 32, 16


def _atlas_off(layer, x, y, c):
    return ((layer * TEX + y) * TEX + x) * 4 + c


def _env_off(x, y, c):
    return (y * EW + x) * 4 + c


@pytest.fixture(scope="module")
def small():
    data = scenes.config5(cells=16, num_spheres=5, tex_res=TEX, env_size=(EW, EH))
    o = OracleScene(data, threads=8)
    o.build_bvh()
    b = pike_amd.BVH(data.positions, data.triangles)
    cam = scenes.normalized_camera(data.camera, b.normalizer())
    d_image = np.random.default_rng(3).uniform(-1.0, 1.0, (H, W, 4)).astype(np.float32)
    return data, o, cam, d_image


def _candidates(data):
    rng = np.random.default_rng(11)
    out = []
    for _ in range(24):
        x, y = rng.integers(0, TEX, 2)
        out.append((ATLAS, _atlas_off(0, x, y, int(rng.integers(0, 3)))))   # base color
    for _ in range(12):
        x, y = rng.integers(0, TEX, 2)
        out.append((ATLAS, _atlas_off(1, x, y, 1)))                          # roughness
    env = data.env.reshape(EH, EW, 3)
    lum = env @ np.array([0.2126, 0.7152, 0.0722])
    bright = np.argsort(lum.ravel())[::-1][:8]
    for t in bright:
        out.append((ENV, _env_off(int(t % EW), int(t // EW), int(rng.integers(0, 3)))))
    for _ in range(12):
        x, y = rng.integers(0, EW), rng.integers(0, EH)
        out.append((ENV, _env_off(int(x), int(y), int(rng.integers(0, 3)))))
    return out


def _value(o, cam, d_image):
    return o.gradient(cam, W, H, DEPTH, d_image, [], num_frames=SPP)[1]


def _read(data, which, off):
    if which == ATLAS:
        return float(data.tex.ravel()[off])
    e = data.env.reshape(EH, EW, 3)
    t, c = off // 4, off % 4
    return float(e[t // EW, t % EW, c])


def test_fp64_gradient_matches_central_differences(small):
    data, o, cam, d_image = small
    params = _candidates(data)
    g, J = o.gradient(cam, W, H, DEPTH, d_image, params, num_frames=SPP)
    assert np.count_nonzero(g) >= 20
    scale = np.abs(g).max()
    eps = 1e-4
    for (which, off), gi in zip(params, g):
        v = _read(data, which, off)
        hi, lo = np.float32(v + eps), np.float32(v - eps)
        o.write_params(which, [hi], off)
        jp = _value(o, cam, d_image)
        o.write_params(which, [lo], off)
        jm = _value(o, cam, d_image)
        o.write_params(which, [np.float32(v)], off)
        if gi == 0.0:
            assert jp == jm, (which, off)
            continue
        fd = (jp - jm) / (float(hi) - float(lo))
        assert abs(fd - gi) <= 1e-5 * abs(gi) + 1e-9 * scale, (which, off, gi, fd)
    assert _value(o, cam, d_image) == J


def test_fp64_env_gradient_is_euler_complete(small):
    data, o, cam, d_image = small
    env = data.env.reshape(EH, EW, 3)
    params = [(ENV, _env_off(x, y, c)) for y in range(EH) for x in range(EW) for c in range(3)]
    g, J = o.gradient(cam, W, H, DEPTH, d_image, params, num_frames=SPP)
    E = np.array([env[off // 4 // EW, off // 4 % EW, off % 4] for _, off in params], np.float64)
    assert abs(float(g @ E) - J) <= 1e-9 * np.abs(g * E).sum()


def test_objective_equals_rendered_image(small):
    data, o, cam, d_image = small
    for precision in (64, 32):
        img, _, _ = o.render(cam, W, H, DEPTH, 0, SPP, precision=precision)
        ref = float((d_image[..., :3].astype(np.float64) * img[..., :3]).sum())
        J = o.gradient(cam, W, H, DEPTH, d_image, [], num_frames=SPP, precision=precision)[1]
        assert abs(J - ref) <= 1e-5 * np.abs(d_image[..., :3] * img[..., :3]).sum(), precision


def test_fp32_and_fp64_gradients_agree(small):
    data, o, cam, d_image = small
    params = _candidates(data)
    g64, _ = o.gradient(cam, W, H, DEPTH, d_image, params, num_frames=SPP)
    g32, _ = o.gradient(cam, W, H, DEPTH, d_image, params, num_frames=SPP, precision=32)
    # common random numbers: the two precisions follow the same paths except
    # for rare rounding flips, so the gradients agree to f32 accuracy overall
    rel = np.abs(g32 - g64).sum() / np.abs(g64).sum()
    assert rel < 1e-3, rel


def test_base_color_is_not_a_sampling_input(small):
    data, _, cam, _ = small
    rng = np.random.default_rng(5)
    base = data.tex.copy()
    base[0, ..., :3] = np.clip(base[0, ..., :3] + rng.normal(0, 0.1, base[0, ..., :3].shape), 0.02, 1.0)
    imgs = []
    for also in (False, True):
        o = OracleScene(data, threads=8)
        o.build_bvh()
        o.write_params(ATLAS, base.ravel(), 0, also_sampling=also)
        imgs.append(o.render(cam, W, H, DEPTH, 0, SPP, precision=32)[0])
        o.close()
    np.testing.assert_array_equal(imgs[0].view(np.uint32), imgs[1].view(np.uint32))
    rough = data.tex.copy()
    rough[1, ..., 1] = np.clip(rough[1, ..., 1] + 0.2, 0.05, 1.0)
    imgs = []
    for also in (False, True):
        o = OracleScene(data, threads=8)
        o.build_bvh()
        o.write_params(ATLAS, rough.ravel(), 0, also_sampling=also)
        imgs.append(o.render(cam, W, H, DEPTH, 0, SPP, precision=32)[0])
        o.close()
    assert not np.array_equal(imgs[0], imgs[1])
