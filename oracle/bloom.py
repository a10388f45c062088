"""Numpy restatement of the reference's bloom and composite passes
(src/engine/render/passes/createBloomPass.ts).

Test infrastructure only: the oracle of the HIP bloom kernels. Each render pass
draws a full-screen triangle whose fragment at target pixel (i, j) has
uv = ((i + 0.5) / W, (j + 0.5) / H) (fullscreenVert :193-204, row 0 at the top),
and reads its sources through a linear, clamp-to-edge sampler (:50). The
half-resolution targets are rgba16float (:148-154), so values are rounded to
binary16 after each of the first three passes; the canvas target is 8-bit unorm.
Sampler filtering precision is implementation-defined in WebGPU (hardware often
quantises the weights to 8 bits); this restatement uses exact weights.
"""
import numpy as np


def sample_linear(img, u, v):
    """Bilinear, clamp-to-edge sample of img [h, w, 4] at uv arrays."""
    h, w = img.shape[:2]
    tx = u * w - 0.5
    ty = v * h - 0.5
    x0 = np.floor(tx)
    y0 = np.floor(ty)
    ax = (tx - x0)[..., None]
    ay = (ty - y0)[..., None]
    x0 = x0.astype(np.int64)
    y0 = y0.astype(np.int64)
    xa, xb = np.clip(x0, 0, w - 1), np.clip(x0 + 1, 0, w - 1)
    ya, yb = np.clip(y0, 0, h - 1), np.clip(y0 + 1, 0, h - 1)
    c00, c10, c01, c11 = img[ya, xa], img[ya, xb], img[yb, xa], img[yb, xb]
    top = c00 * (1 - ax) + c10 * ax
    bot = c01 * (1 - ax) + c11 * ax
    return top * (1 - ay) + bot * ay


def _uv(w, h):
    v, u = np.meshgrid((np.arange(h) + 0.5) / h, (np.arange(w) + 0.5) / w, indexing="ij")
    return u, v


def _f16(x):
    return x.astype(np.float16).astype(np.float64)


def threshold(src, hw, hh, thr):
    """thresholdShader (:206-223): downsample + soft luminance threshold."""
    u, v = _uv(hw, hh)
    c = sample_linear(src, u, v)
    lum = c[..., 0] * 0.2126 + c[..., 1] * 0.7152 + c[..., 2] * 0.0722
    knee = thr * 0.5
    rq = np.clip(lum - thr + knee, 0.0, 2.0 * knee)
    wgt = (rq * rq) / (4.0 * knee * knee + 0.00001)
    mask = np.maximum(wgt, (thr <= lum).astype(np.float64))
    out = np.empty_like(c)
    out[..., :3] = c[..., :3] * mask[..., None] * c[..., 3:4]
    out[..., 3] = c[..., 3] * mask
    return _f16(out)


WEIGHTS = (0.2270270270, 0.1945945946, 0.1216216216, 0.0540540541, 0.0162162162)


def blur(src, radius, horizontal):
    """blurShader (:225-262): 9 taps spaced (1 + 2 radius) texels along one axis."""
    hh, hw = src.shape[:2]
    u, v = _uv(hw, hh)
    scale = 1.0 + radius * 2.0
    su, sv = ((1.0 / hw) * scale, 0.0) if horizontal else (0.0, (1.0 / hh) * scale)
    col = sample_linear(src, u, v) * WEIGHTS[0]
    for k in range(1, 5):
        col = col + sample_linear(src, u + su * k, v + sv * k) * WEIGHTS[k]
        col = col + sample_linear(src, u - su * k, v - sv * k) * WEIGHTS[k]
    return _f16(col)


def unorm8(x):
    return np.rint(np.clip(x, 0.0, 1.0) * 255.0).astype(np.uint8)


def bloom(src_rgba_f16, enabled=True, threshold_=0.6, strength=0.9, radius=0.4):
    """BloomPass.render (:99-129): rgba float16 [h, w, 4] -> canvas uint8 [h, w, 4]."""
    src = np.asarray(src_rgba_f16, dtype=np.float16).astype(np.float64)
    h, w = src.shape[:2]
    hw, hh = max(1, w // 2), max(1, h // 2)
    u, v = _uv(w, h)
    base = sample_linear(src, u, v)
    if not enabled:  # blitShader (:264-274)
        hdr = np.maximum(base[..., :3], 0.0)
        alpha = np.clip(base[..., 3], 0.0, 1.0)
    else:  # compositeShader (:276-293)
        bright = threshold(src, hw, hh, threshold_)
        glow = sample_linear(blur(blur(bright, radius, True), radius, False), u, v)
        hdr = np.maximum(base[..., :3] + np.maximum(glow[..., :3], 0.0) * strength, 0.0)
        ga = np.clip(glow[..., 3] * strength, 0.0, 1.0)
        alpha = np.clip(base[..., 3] + ga * (1.0 - base[..., 3]), 0.0, 1.0)
    out = np.empty((h, w, 4), dtype=np.uint8)
    out[..., :3] = unorm8(np.power(hdr, 0.92))
    out[..., 3] = unorm8(alpha)
    return out
