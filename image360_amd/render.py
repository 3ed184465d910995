"""Host-side chunk render driver over the HIP backend.

``render_chunk_u16`` is RenderExecutor::Execute's compute sequence
(WindowsApp.Core/SourceFiles/RenderExecutor.cpp:59-140) with every buffer
resident on the device: for each contributor (capped at 32) invert its
forward homography in binary32 (HomographyMath.cpp:94-116), warp it into the
chunk (IComputeBackend::WarpPerspective), apply its gain when
|gain-1| > 1e-6, then blend the warped contributors
(Kernels::SeamBlend::BlendChunkContributors).  Contributors whose
homography does not invert are skipped, as the reference does.
"""
from __future__ import annotations

import numpy as np

MAX_CONTRIBUTORS = 32


def invert_homography_f32(h):
    """HomographyMath.cpp:94-116: binary32 adjugate inverse; None if singular."""
    f = np.float32
    a, b, c, d, e, ff, g, i, j = (f(v) for v in np.asarray(h, dtype=np.float64).reshape(9).astype(np.float32))
    c0 = e * j - ff * i
    c1 = d * j - ff * g
    c2 = d * i - e * g
    det = (a * c0 - b * c1) + c * c2
    if abs(det) < f(1e-12):
        return None
    s = f(1.0) / det
    return np.array([c0 * s, (c * i - b * j) * s, (b * ff - c * e) * s,
                     (ff * g - d * j) * s, (a * j - c * g) * s, (c * d - a * ff) * s,
                     c2 * s, (b * g - a * i) * s, (a * e - b * d) * s], np.float32)


def render_chunk_u16(hip, geometry, chunk, frame_ptrs, warped_ptrs, out_ptr, num_bands=5, stream=None,
                     owner_ptr=None):
    """Renders one chunk into the RGB48 buffer at `out_ptr` (device).

    frame_ptrs[i]: device RGB48 of frame i; warped_ptrs: at least
    len(chunk.contributors) device buffers of chunk.w*chunk.h*3 u16."""
    used = []
    for k, fi in enumerate(chunk.contributors[:MAX_CONTRIBUTORS]):
        inv = invert_homography_f32(geometry.H[fi])
        if inv is None:
            continue
        dst = warped_ptrs[len(used)]
        hip.check(hip.warp_dev(frame_ptrs[fi], int(geometry.frame_w[fi]), int(geometry.frame_h[fi]), dst, chunk.w,
                               chunk.h, inv, chunk.x, chunk.y, stream), "WarpPerspective")
        gain = float(geometry.gains[fi])
        if abs(gain - 1.0) > 1e-6:
            hip.check(hip.gain_dev(dst, chunk.w * chunk.h * 3, gain, stream), "ApplyGain")
        used.append(dst)
    hip.check(hip.blend_dev(used, chunk.w, chunk.h, out_ptr, num_bands, owner_ptr, stream), "BlendChunkContributors")
    return len(used)
