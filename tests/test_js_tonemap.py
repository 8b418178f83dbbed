"""PostProcessPass.tonemap (pike_amd/js/passes/postprocess_pass.js), the
host-side restatement of render/shader/postprocess.wgsl:36-99: exposure,
the ACES fit (input matrix, RRT+ODT rational fit, output matrix, clamp),
saturation through mat3x3(red, green, blue) * color, sRGB encoding, and the
8-bit unorm store of the canvas target.

Checked against a float64 statement of the same WGSL (mat3x3 built from 9
scalars is column-major, so `v * M` dots v with M's columns) and against
what the definitions fix: saturation 1 is the identity and 0 the Rec.601-
style grey of the shader's luminance weights, sRGB maps 0 -> 0 and 1 -> 1
continuously at its 0.0031308 knee, black stays black, large inputs clip to
white, and the curve is monotone in exposure."""
import json
import os
import shutil
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
POSTPROCESS = os.path.join(ROOT, "pike_amd", "js", "passes", "postprocess_pass.js")
NODE = shutil.which("node")
pytestmark = pytest.mark.skipif(NODE is None, reason="node is not installed")

ACES_IN = np.array([0.59719, 0.35458, 0.04823, 0.07600, 0.90834, 0.01566, 0.02840, 0.13383, 0.83777])
ACES_OUT = np.array([1.60475, -0.53108, -0.07367, -0.10208, 1.10813, -0.00605, -0.00327, -0.07276, 1.07602])
LUM = np.array([0.3086, 0.6094, 0.0820])


def js_tonemap(rgba, w, h, exposure, saturation):
    mod = POSTPROCESS
    code = ("const {PostProcessPass}=require(%r);let s='';process.stdin.on('data',d=>s+=d);"
            "process.stdin.on('end',()=>{const a=JSON.parse(s);const p=new PostProcessPass();"
            "p.setExposure(a.e);p.setSaturation(a.s);"
            "console.log(JSON.stringify(Array.from(p.tonemap(Float32Array.from(a.px),a.w,a.h))));});" % mod)
    r = subprocess.run([NODE, "-e", code], input=json.dumps(dict(px=rgba.ravel().tolist(), w=w, h=h, e=exposure,
                                                                 s=saturation)),
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    return np.array(json.loads(r.stdout), np.uint8).reshape(h, w, 4)


def wgsl_tonemap(rgb, exposure, saturation):
    v = rgb.astype(np.float64) * exposure
    v = v @ ACES_IN.reshape(3, 3).T          # dot with each column of mat3x3(9 scalars)
    v = (v * (v + 0.0245786) - 0.000090537) / (v * (0.983729 * v + 0.4329510) + 0.238081)
    v = np.clip(v @ ACES_OUT.reshape(3, 3).T, 0.0, 1.0)
    s = saturation
    M = np.column_stack([LUM[0] * (1 - s) + np.array([s, 0, 0]), LUM[1] * (1 - s) + np.array([0, s, 0]),
                         LUM[2] * (1 - s) + np.array([0, 0, s])])  # columns red, green, blue
    v = v @ M.T
    v = np.where(v <= 0.0031308, v * 12.92, 1.055 * np.power(np.maximum(v, 0), 1 / 2.4) - 0.055)
    return np.floor(np.clip(v, 0, 1) * 255 + 0.5)


def test_tonemap_equals_wgsl_statement():
    rng = np.random.default_rng(3)
    w, h = 32, 16
    img = np.zeros((h, w, 4), np.float32)
    img[..., :3] = np.exp(rng.normal(-1.0, 1.5, (h, w, 3)))
    img[0, :8, :3] = 0.0
    for exposure, sat in ((1.0, 1.0), (0.35, 1.3), (2.5, 0.0), (1.0, 0.6)):
        out = js_tonemap(img, w, h, exposure, sat)
        ref = wgsl_tonemap(img[..., :3], exposure, sat)
        assert np.all(out[..., 3] == 255)
        # the float64 statement and JS doubles agree except where a value
        # sits within rounding of a .5 step
        assert np.abs(out[..., :3].astype(int) - ref).max() <= 1
        assert np.mean(out[..., :3] != ref) < 0.01


def test_tonemap_closed_forms():
    w, h = 64, 1
    img = np.zeros((h, w, 4), np.float32)
    ramp = np.linspace(0, 4, w)
    img[0, :, 0] = img[0, :, 1] = img[0, :, 2] = ramp
    out = js_tonemap(img, w, h, 1.0, 1.0)[0, :, :3].astype(int)
    assert np.all(out[0] == 0)                      # black stays black (the fit clamps its -4e-4 at 0)
    assert np.all(np.diff(out[:, 1]) >= 0)          # monotone
    big = js_tonemap(np.full((1, 1, 4), 1e4, np.float32), 1, 1, 1.0, 1.0)[0, 0, :3]
    assert np.all(big == 255)                       # clips to white
    # saturation 0: every channel equals the luminance-weighted grey
    col = np.zeros((1, 3, 4), np.float32)
    col[0, :, :3] = [[0.9, 0.1, 0.2], [0.1, 0.6, 0.05], [0.3, 0.3, 0.9]]
    grey = js_tonemap(col, 3, 1, 1.0, 0.0)[0, :, :3]
    assert np.all(grey.max(axis=1) - grey.min(axis=1) == 0)
    # saturation 1 leaves the ACES output untouched; 0.5 lies between
    full = js_tonemap(col, 3, 1, 1.0, 1.0)[0, :, :3].astype(int)
    half = js_tonemap(col, 3, 1, 1.0, 0.5)[0, :, :3].astype(int)
    g = grey.astype(int)
    assert np.all((half - np.minimum(full, g) >= -1) & (np.maximum(full, g) - half >= -1))
    # sRGB: continuous at the knee, 1 -> 255
    knee = 0.0031308
    assert abs(knee * 12.92 - (1.055 * knee ** (1 / 2.4) - 0.055)) < 1e-6
