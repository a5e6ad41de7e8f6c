"""esf_step_batch: scenes stepped together, every kernel of the step
launched once over the scenes that reach it.  Each scene's fields and
solve reports must be bitwise those of the same scene stepped alone."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def esf():
    import eulersf_amd
    eulersf_amd.load()
    return eulersf_amd


def _say(msg):
    print(msg, flush=True)


def fields(esf, g, nsolids):
    A = esf._abi
    out = [g.download(A.F_VELOCITY), g.download(A.F_INTENSITY), g.download(A.F_LAMBDA)]
    for b in range(nsolids):
        out += [g.download(A.S_U, b), g.download(A.S_MASS, b)]
    return [x.view(np.int64) for x in out]


def run_alone(esf, texts, precision, steps):
    out = []
    for t in texts:
        g = esf.CoupledIntegrator3D.from_cfg_text(t, precision)
        reps = [g.stepImplicit(-1.0) for _ in range(steps)]
        out.append((fields(esf, g, t.count("solid sphere radius")), reps))
        g.close()
    return out


def check_batched(esf, texts, precision, steps, alone):
    gs = [esf.CoupledIntegrator3D.from_cfg_text(t, precision) for t in texts]
    reps = [[] for _ in gs]
    for _ in range(steps):
        for i, r in enumerate(esf.step_batch(gs)):
            reps[i].append(r)
    for i, (g, t) in enumerate(zip(gs, texts)):
        fa, ra = alone[i]
        for a, b in zip(fa, fields(esf, g, t.count("solid sphere radius"))):
            np.testing.assert_array_equal(a, b)
        for x, y in zip(ra, reps[i]):
            assert x == y, (i, x, y)
    for g in gs:
        g.close()


@pytest.mark.parametrize("precision", ["mixed", "fp64"])
def test_batch_bitwise_alone(esf, precision):
    """six uneven scenes (N 12-24, 1-3 spheres), three steps: batched and
    alone give the same bits and the same solve reports"""
    from eulersf_amd.serve import random_scene
    prec = esf.ESF_MIXED if precision == "mixed" else esf.ESF_PLAIN_FP64
    rng = np.random.default_rng(31)
    ns = [16, 20, 24, 12, 20, 16]
    texts = [random_scene(rng, n, eps=1e-10) for n in ns]
    _say(f"batch {precision}: N {ns}")
    alone = run_alone(esf, texts, prec, 3)
    check_batched(esf, texts, prec, 3, alone)


def test_batch_of_one_and_of_moving_scenes(esf):
    """a batch of one scene, and a batch mixing a thrown ball, a FLIP scene
    and two resting ones: the same bits as alone"""
    from eulersf_amd.cfg import unit_cube_ball_cfg
    from eulersf_amd.serve import random_scene
    rng = np.random.default_rng(5)
    thrown = unit_cube_ball_cfg(20, eps=1e-10, velocity=(0.3, -1.0, 0.2))
    flip = unit_cube_ball_cfg(16, eps=1e-10, extra={"flip particles": 1})
    texts = [thrown, random_scene(rng, 18, eps=1e-10), flip, random_scene(rng, 14, eps=1e-10)]
    alone = run_alone(esf, texts, esf.ESF_MIXED, 3)
    check_batched(esf, texts[:1], esf.ESF_MIXED, 3, alone[:1])
    _say("batch of one: same")
    check_batched(esf, texts, esf.ESF_MIXED, 3, alone)


def test_batch_at_workload_sizes(esf):
    """the batched workload's mix: 64 scenes with N drawn from
    {32, 48, 64, 80, 96} (1-3 spheres, mixed precision at 1e-8), two steps
    alone, then two batched steps of all 64: every scene bitwise the same"""
    from eulersf_amd.serve import random_scene
    rng = np.random.default_rng(2026)
    ns = [int(rng.choice([32, 48, 64, 80, 96])) for _ in range(64)]
    texts = [random_scene(rng, n) for n in ns]
    _say(f"batch workload: N {sorted(set(ns))}")
    alone = run_alone(esf, texts, esf.ESF_MIXED, 2)
    _say("batch workload: alone done")
    check_batched(esf, texts, esf.ESF_MIXED, 2, alone)


def test_batched_server_results_same_as_alone(esf):
    """a served stream through the batched serving loop (one esf_step_batch
    per step over the scenes in flight, scenes built beside the stepping):
    every request completes with its K steps, never more than the slots in
    flight, and each result is bitwise that of the scene stepped alone"""
    from eulersf_amd.serve import BatchedServer, poisson_requests
    rng = np.random.default_rng(9)
    reqs = poisson_requests(rng, rate=40.0, seconds=0.5, n_range=(12, 24), k_range=(2, 5),
                            eps=1e-10)
    assert len(reqs) >= 5
    got = {}

    def keep(req, g):
        got[req["id"]] = fields(esf, g, next(r for r in reqs if r["id"] == req["id"])["text"]
                                .count("solid sphere radius"))

    srv = BatchedServer(budget_bytes=4e9, max_active=6, precision=esf.ESF_MIXED, builders=2,
                        on_done=keep)
    recs, span, info = srv.run(reqs)
    _say(f"batched server: {len(recs)} requests, {info}")
    assert len(recs) == len(reqs)
    assert all(r["error"] is None for r in recs), [r["error"] for r in recs]
    assert all(len(r["lat"]) == r["k"] for r in recs)
    assert info["peak_active"] <= 6
    assert info["mean_scenes_per_batch_step"] > 1
    for r in reqs:
        g = esf.CoupledIntegrator3D.from_cfg_text(r["text"], esf.ESF_MIXED)
        for _ in range(r["k"]):
            g.stepImplicit(-1.0)
        alone = fields(esf, g, r["text"].count("solid sphere radius"))
        g.close()
        for a, b in zip(alone, got[r["id"]]):
            np.testing.assert_array_equal(a, b)


def test_lists_grown_after_their_scan_same_bits(esf, monkeypatch):
    """A step reads the solid DOF and valid-cell counts only at its end: a
    list whose scan ran past its capacity is grown there and scanned again.
    ESF_TIGHT_LISTS=1 sizes the lists to exactly their last counts, so every
    step whose count rises takes that path.  A thrown ball and a random
    scene over 9 steps, alone and batched: the same bits and reports as with
    the usual headroom, and the counts do rise."""
    from eulersf_amd.cfg import unit_cube_ball_cfg
    from eulersf_amd.serve import random_scene
    rng = np.random.default_rng(17)
    thrown = unit_cube_ball_cfg(20, eps=1e-10, velocity=(0.5, -1.4, 0.3))
    texts = [thrown, random_scene(rng, 22, eps=1e-10)]
    steps = 9
    usual = run_alone(esf, texts, esf.ESF_MIXED, steps)
    monkeypatch.setenv("ESF_TIGHT_LISTS", "1")
    tight = run_alone(esf, texts, esf.ESF_MIXED, steps)
    rises = 0
    for (fu, ru), (ft, rt) in zip(usual, tight):
        for a, b in zip(fu, ft):
            np.testing.assert_array_equal(a, b)
        assert ru == rt
        for key in ("n_solid_nodes", "n_solid_cells"):
            seq = [r[key] for r in rt]
            rises += sum(b > a for a, b in zip(seq, seq[1:]))
    _say(f"tight lists: same bits alone, {rises} count rises")
    assert rises > 0
    check_batched(esf, texts, esf.ESF_MIXED, steps, usual)
    _say("tight lists: same bits batched")
