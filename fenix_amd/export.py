"""Offline export: deterministic-time frames of step + raymarch + bloom/composite,
and the flipbook (sprite sheet) layout.

References: the export loops of src/ui/components/toolbar/ExportPopover.tsx
(exportSpritesheet :336-397, exportVideo :399-451: pause, reset, then frame i
at elapsed (startFrame + i) / rate with step 1 / rate) and renderOffscreenFrame
(src/engine/render/renderer/createViewportRuntime.ts:205-232: step, raymarch into
the HDR image, bloom/composite into the canvas). The viewport overlays and the
video encoder are outside this library.
"""
from typing import Iterator, Optional, Sequence, Tuple

import numpy as np

from .render import BloomPass


def export_frames(sim, raymarch, camera, start_frame: int, end_frame: int, rate: float, width: int, height: int,
                  display_mode: str = "temperature", bloom: Optional[BloomPass] = None,
                  bloom_params=None) -> Iterator[np.ndarray]:
    """Resets the simulation and yields the canvas image (uint8 [height, width, 4]) of
    every frame from start_frame to end_frame inclusive. rate is the export's frame
    rate (video fps or flipbook sample rate)."""
    if end_frame < start_frame:
        raise ValueError("end_frame must not precede start_frame")
    if not rate > 0:
        raise ValueError("rate must be positive")
    bloom = bloom or BloomPass()
    params = bloom_params or bloom.params_struct()
    frame_duration = 1.0 / rate
    sim.reset()
    for i in range(end_frame - start_frame + 1):
        t = (start_frame + i) * frame_duration
        sim.step(None, t, frame_duration)
        hdr = raymarch.render(sim, camera, display_mode, width, height, t)
        yield bloom.render(sim, hdr, params)


def flipbook(frames: Sequence[np.ndarray], columns: int, background: Optional[Tuple[int, int, int, int]] = None) -> np.ndarray:
    """exportSpritesheet's layout: frame i at column i % columns, row i // columns, on a
    sheet of ceil(n / columns) rows; a background colour (RGBA8) fills the sheet first
    and frames are drawn over it with source-over on straight alpha, else the sheet is
    transparent and frames are copied."""
    if not frames:
        raise ValueError("no frames")
    if columns < 1:
        raise ValueError("columns must be >= 1")
    h, w = frames[0].shape[:2]
    rows = (len(frames) + columns - 1) // columns
    sheet = np.zeros((rows * h, columns * w, 4), dtype=np.uint8)
    if background is not None:
        sheet[...] = np.asarray(background, dtype=np.uint8)
    for i, f in enumerate(frames):
        r, c = divmod(i, columns)
        dst = sheet[r * h:(r + 1) * h, c * w:(c + 1) * w]
        if background is None:
            dst[...] = f
        else:
            a = f[..., 3:4].astype(np.float64) / 255.0
            out = f[..., :3].astype(np.float64) * a + dst[..., :3].astype(np.float64) * (1.0 - a)
            dst[..., :3] = np.rint(out).astype(np.uint8)
            dst[..., 3] = np.rint(255.0 * (a[..., 0] + dst[..., 3] / 255.0 * (1.0 - a[..., 0]))).astype(np.uint8)
    return sheet
