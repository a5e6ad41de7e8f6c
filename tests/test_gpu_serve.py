"""Scenes served together: every scene's result is the same alone and in
the mix.  Each scene owns a context (its own HIP stream and device memory)
and every kernel reduces in a fixed order, so concurrency must not change a
single bit of any field."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def esf():
    import eulersf_amd
    eulersf_amd.load()
    return eulersf_amd


def fields(esf, g, nsolids):
    A = esf._abi
    out = [g.download(A.F_VELOCITY), g.download(A.F_INTENSITY), g.download(A.F_LAMBDA)]
    out += [g.download(A.S_U, b) for b in range(nsolids)]
    return out


def test_scene_results_same_alone_and_in_the_mix(esf):
    from eulersf_amd.serve import BatchRunner, random_scene, step_scene
    rng = np.random.default_rng(11)
    ns = [16, 20, 24, 20]
    texts = [random_scene(rng, n, eps=1e-10) for n in ns]
    nsol = [t.count("solid sphere radius") for t in texts]
    alone = []
    for t, k in zip(texts, nsol):
        g = esf.CoupledIntegrator3D.from_cfg_text(t, esf.ESF_MIXED)
        step_scene(g, 3)
        alone.append(fields(esf, g, k))
        g.close()
    br = BatchRunner(texts, ns, esf.ESF_MIXED, warmup=0)
    wall, lat = br.run_concurrent(3, workers=len(texts))
    assert len(lat) == 3 * len(texts)
    for i, (g, k) in enumerate(zip(br.scenes, nsol)):
        for a, b in zip(alone[i], fields(esf, g, k)):
            np.testing.assert_array_equal(a, b)
    br.close()


def test_server_admits_and_completes_every_request(esf):
    """a short served stream on a small budget: every request completes,
    never more than the slots run at once, and each takes its K steps"""
    from eulersf_amd.serve import Server, poisson_requests
    rng = np.random.default_rng(5)
    reqs = poisson_requests(rng, rate=40.0, seconds=0.5, n_range=(12, 24), k_range=(2, 4),
                            eps=1e-8)
    assert reqs
    srv = Server(budget_bytes=2e9, max_active=3, precision=esf.ESF_MIXED)
    recs, span, info = srv.run(reqs)
    assert len(recs) == len(reqs)
    assert all(r["error"] is None for r in recs), [r["error"] for r in recs]
    assert all(len(r["lat"]) == r["k"] for r in recs)
    assert info["peak_active"] <= 3
    assert all(r["start"] >= r["arrival"] - 1e-3 for r in recs)


def test_batched_scenes_at_workload_sizes(esf):
    """the batched workload's mix: 64 scenes with N drawn from
    {32, 48, 64, 80, 96} (1-3 spheres, mixed precision at 1e-8), three steps
    each alone, then all 64 together on 16 workers: every scene bitwise the
    same"""
    from eulersf_amd.serve import BatchRunner, random_scene, step_scene
    rng = np.random.default_rng(2026)
    ns = [int(rng.choice([32, 48, 64, 80, 96])) for _ in range(64)]
    texts = [random_scene(rng, n) for n in ns]
    nsol = [t.count("solid sphere radius") for t in texts]
    print(f"batched: N {sorted(set(ns))}, {sum(nsol)} solids", flush=True)
    alone = []
    for t, k in zip(texts, nsol):
        g = esf.CoupledIntegrator3D.from_cfg_text(t, esf.ESF_MIXED)
        step_scene(g, 3)
        alone.append([x.view(np.int64) for x in fields(esf, g, k)])
        g.close()
    print("batched: alone done", flush=True)
    br = BatchRunner(texts, ns, esf.ESF_MIXED, warmup=0)
    wall, lat = br.run_concurrent(3, workers=16)
    assert len(lat) == 3 * len(texts)
    for i, (g, k) in enumerate(zip(br.scenes, nsol)):
        for a, b in zip(alone[i], fields(esf, g, k)):
            assert np.array_equal(a, b.view(np.int64)), (i, ns[i])
    br.close()


def test_served_stream_at_workload_sizes(esf):
    """the serving workload's requests (N uniform in [32, 160], 1-3 spheres,
    mixed precision at 1e-8), with K cut to 3-6 steps: served on 4 slots
    under the 200 GB budget, and every request's fields bitwise the same as
    its scene run alone for the same steps"""
    from eulersf_amd.serve import Server, poisson_requests, step_scene
    rng = np.random.default_rng(7)
    reqs = poisson_requests(rng, rate=40.0, seconds=0.25, n_range=(32, 160), k_range=(3, 6))
    reqs = reqs[:8]
    assert len(reqs) >= 4
    print(f"served: N {[r['n'] for r in reqs]}", flush=True)
    served = {}

    def keep(req, g):
        k = req["text"].count("solid sphere radius")
        served[req["id"]] = [x.view(np.int64) for x in fields(esf, g, k)]

    srv = Server(200e9, 4, esf.ESF_MIXED, on_done=keep)
    recs, span, info = srv.run(reqs)
    assert all(r["error"] is None for r in recs), [r["error"] for r in recs]
    assert info["peak_active"] <= 4
    assert sorted(served) == sorted(r["id"] for r in reqs)
    print("served: stream done", flush=True)
    for r in reqs:
        k = r["text"].count("solid sphere radius")
        g = esf.CoupledIntegrator3D.from_cfg_text(r["text"], esf.ESF_MIXED)
        step_scene(g, r["k"])
        for a, b in zip(served[r["id"]], fields(esf, g, k)):
            assert np.array_equal(a, b.view(np.int64)), (r["id"], r["n"])
        g.close()
