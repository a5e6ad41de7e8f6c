"""Host logic of the batched / served workloads (no GPU): request streams
and scene generation."""
import numpy as np

from eulersf_amd.serve import poisson_requests, random_scene, summarize


def test_poisson_stream_statistics():
    rng = np.random.default_rng(1)
    reqs = poisson_requests(rng, rate=20.0, seconds=200.0, n_range=(32, 160), k_range=(10, 100))
    t = np.array([r["t"] for r in reqs])
    assert np.all(np.diff(t) > 0) and t[-1] <= 200.0
    assert abs(len(reqs) / 200.0 - 20.0) < 2.0          # rate
    gaps = np.diff(t)
    assert abs(gaps.std() / gaps.mean() - 1.0) < 0.15   # exponential gaps
    ns = np.array([r["n"] for r in reqs])
    ks = np.array([r["k"] for r in reqs])
    assert ns.min() >= 32 and ns.max() <= 160 and ks.min() >= 10 and ks.max() <= 100


def test_random_scene_has_one_to_three_spheres():
    rng = np.random.default_rng(2)
    for _ in range(20):
        text = random_scene(rng, 48)
        k = text.count("solid sphere radius")
        assert 1 <= k <= 3
        assert "number of solids = %d" % k in text


def test_summary_of_records():
    recs = [{"id": 0, "n": 32, "k": 2, "arrival": 0.0, "start": 0.1, "end": 0.5,
             "lat": [0.1, 0.2], "error": None},
            {"id": 1, "n": 64, "k": 1, "arrival": 0.2, "start": 0.5, "end": 0.9,
             "lat": [0.3], "error": None}]
    s = summarize(recs, 1.0)
    assert s["steps"] == 3 and s["steps_per_s"] == 3.0
    assert s["cells_steps_per_s"] == 2 * 32 ** 3 + 64 ** 3
    assert s["completion_latency_ms"]["p50"] == 600.0
