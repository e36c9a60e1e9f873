"""The served job stream and its host-side helpers (CPU)."""
import math

import numpy as np

from aviatorengine2_amd import serve


def test_poisson_stream_shape():
    jobs = serve.poisson_jobs(4000, rate_hz=20.0, seed=3)
    arr = np.array([j[0] for j in jobs])
    assert arr[0] == 0.0 and np.all(np.diff(arr) >= 0)
    # exponential gaps: mean 1/rate, coefficient of variation 1
    gaps = np.diff(arr)
    assert abs(gaps.mean() - 0.05) < 0.005
    assert abs(gaps.std() / gaps.mean() - 1.0) < 0.08
    assert {(j[2], j[3]) for j in jobs} == set(serve.RESOLUTIONS)
    spp = np.array([j[4] for j in jobs])
    assert spp.min() == 1 and spp.max() == 64
    # log-uniform: about as many jobs in [1,8) as in [8,64]
    lo, hi = (spp < 8).mean(), (spp >= 8).mean()
    assert abs(lo - hi) < 0.1
    depth = np.array([j[5] for j in jobs])
    assert depth.min() == 4 and depth.max() == 8
    assert set(j[1] for j in jobs) == set(range(8))


def test_stream_is_seeded():
    assert serve.poisson_jobs(50, 5.0, seed=9) == serve.poisson_jobs(50, 5.0, seed=9)
    assert serve.poisson_jobs(50, 5.0, seed=9) != serve.poisson_jobs(50, 5.0, seed=10)
