"""Image comparison (the reference build lists tools/image_compare.cpp,
build/image_compare.vcxproj:367, but its source is not in the checkout; this
is written from the definitions).

    psnr(a, b)        10*log10(peak^2 / MSE) over RGB, linear radiance clamped
                      to [0, 1] (peak 1), or 8-bit data (peak 255)
    rmse_rel(a, b)    sqrt(mean((a-b)^2)) / mean(b) over RGB (relative RMSE)
    max_abs(a, b)     largest per-channel absolute difference

CLI:  python -m aviatorengine2_amd.imagecmp a b   (.npy, .png, .ppm, .pfm or .exr)
"""
from __future__ import annotations

import math
import sys

import numpy as np


def _rgb(x):
    x = np.asarray(x)
    return x[..., :3] if x.ndim == 3 and x.shape[-1] == 4 else x


def psnr(a, b, peak: float | None = None) -> float:
    a, b = _rgb(a), _rgb(b)
    if a.dtype == np.uint8 and b.dtype == np.uint8:
        peak = 255.0 if peak is None else peak
        a64, b64 = a.astype(np.float64), b.astype(np.float64)
    else:
        peak = 1.0 if peak is None else peak
        a64 = np.clip(a.astype(np.float64), 0.0, peak)
        b64 = np.clip(b.astype(np.float64), 0.0, peak)
    mse = float(np.mean((a64 - b64) ** 2))
    if mse == 0.0:
        return math.inf
    return 10.0 * math.log10(peak * peak / mse)


def rmse_rel(a, b) -> float:
    a, b = _rgb(a).astype(np.float64), _rgb(b).astype(np.float64)
    mean = float(np.mean(np.abs(b)))
    return float(np.sqrt(np.mean((a - b) ** 2))) / (mean if mean > 0 else 1.0)


def max_abs(a, b) -> float:
    a, b = _rgb(a).astype(np.float64), _rgb(b).astype(np.float64)
    return float(np.max(np.abs(a - b))) if a.size else 0.0


def exact_fraction(a, b) -> float:
    a, b = _rgb(a), _rgb(b)
    if a.size == 0:
        return 1.0
    return float(np.mean(np.all(a == b, axis=-1)))


def report(a, b) -> dict:
    return {"psnr_db": psnr(a, b), "rmse_rel": rmse_rel(a, b), "max_abs": max_abs(a, b),
            "exact_pixels": exact_fraction(a, b)}


def main(argv):
    if len(argv) != 3:
        print(__doc__)
        return 2
    from .imagefile import load
    a, b = load(argv[1]), load(argv[2])
    print(report(a, b))
    return 0


if __name__ == "__main__":
    sys.exit(main(sys.argv))
