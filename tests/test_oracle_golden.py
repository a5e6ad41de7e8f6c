"""The oracle against the reference's own code: golden vectors written by
tests/golden/gen_golden.py from oracle/_ref (the reference's COROTATION,
CONJUGATE_RESIDUAL, FIELD_3D interpolation, SIMPLE_PARSER), and -- where
oracle/_ref is built -- live comparisons on fresh random inputs."""
import ctypes
import json
import os

import numpy as np
import pytest

from oracle import oracle as O

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _nan_equal_close(a, b, atol):
    a, b = np.asarray(a), np.asarray(b)
    both_nan = np.isnan(a) & np.isnan(b)
    assert np.array_equal(np.isnan(a), np.isnan(b))
    d = np.where(both_nan, 0.0, np.abs(np.nan_to_num(a) - np.nan_to_num(b)))
    assert d.max() <= atol, d.max()


def corot_tolerances(kind, lam, mu, F, dF):
    """absolute tolerances per output; the two-small-singular-value branch
    builds U from an arbitrary basis of a degenerate eigenspace
    (COROTATION.cpp:177-226), so R and P are pinned only to 1e-6 there"""
    sc = max(1.0, np.abs(F).max())
    k = lam + mu
    loose = 1e-6 if kind == "two_small" else 1e-10
    return {"R": max(loose, 1e-11), "S": 1e-10 * sc, "L": 1e-10 * k, "P": max(loose, 1e-10) * k * sc,
            "e": 1e-10 * k * sc * sc, "dP": max(loose, 1e-10) * k * max(1.0, np.abs(dF).max())}


def test_corotation_golden_vectors():
    g = np.load(os.path.join(GOLD, "corotation.npz"))
    lam, mu = float(g["lam"]), float(g["mu"])
    for i in range(len(g["F"])):
        kind = str(g["kind"][i])
        R, S, L, P, e, dP = O.corotation(lam, mu, g["F"][i], g["dF"][i])
        ref = g["out"][i].reshape(4, 3, 3)
        tol = corot_tolerances(kind, lam, mu, g["F"][i], g["dF"][i])
        _nan_equal_close(R, ref[0], tol["R"])
        _nan_equal_close(S, ref[1], tol["S"])
        _nan_equal_close(L, ref[2], tol["L"])
        _nan_equal_close(P, ref[3], tol["P"])
        _nan_equal_close([e], [g["energy"][i]], tol["e"])
        _nan_equal_close(dP.ravel(), g["dP"][i], tol["dP"])


def test_corotation_covers_every_branch():
    kinds = set(np.load(os.path.join(GOLD, "corotation.npz"))["kind"].tolist())
    for k in ("near_identity", "general", "inverted", "one_small", "two_small", "all_small",
              "zero", "reflection", "diag_inverted"):
        assert k in kinds


def test_cr_golden_iterations_and_solutions(oracle_lib):
    g = np.load(os.path.join(GOLD, "cr.npz"))
    meta = json.loads(str(g["meta"]))
    for k, m in enumerate(meta):
        A, b, x0, xref = g[f"A{k}"], g[f"b{k}"], g[f"x0{k}"], g[f"x{k}"]
        n = len(b)
        D = g[f"D{k}"] if f"D{k}" in g.files else np.diag(A).copy()
        x = x0.copy()
        it = ctypes.c_int()
        ok = oracle_lib.orc_pcr_dense(n, O.dp(np.ascontiguousarray(A)), O.dp(D), O.dp(b), O.dp(x),
                                      m["eps"], 5000, ctypes.byref(it))
        assert ok == m["ok"] == 1
        # same algorithm and cadence: the reference's iteration count; the
        # dot products sum in another order, so allow one check interval
        assert abs(it.value - m["iters"]) <= 5, (k, it.value, m["iters"])
        if m["variant"] == "matfree":  # singular S: compare S x, not x
            np.testing.assert_allclose(A @ x, A @ xref, atol=1e-9 * np.abs(b).max())
        else:
            np.testing.assert_allclose(x, xref, rtol=0, atol=1e-8 * np.abs(xref).max())


def test_interpolation_golden_vectors_bit_exact(oracle_lib):
    g = np.load(os.path.join(GOLD, "interp.npz"))
    X, Y, Z = (int(v) for v in g["dims"])
    dh, o = float(g["dh"]), np.ascontiguousarray(g["o"])
    f, fv = np.ascontiguousarray(g["f"]), np.ascontiguousarray(g["fv"])
    for i, p in enumerate(g["pts"]):
        p = np.ascontiguousarray(p)
        s = oracle_lib.orc_interp_scalar(O.dp(f), X, Y, Z, dh, O.dp(o), O.dp(p))
        assert s == g["out_s"][i]
        v = np.zeros(3)
        oracle_lib.orc_interp_vec3(O.dp(fv), X, Y, Z, dh, O.dp(o), O.dp(p), O.dp(v))
        assert np.array_equal(v, g["out_v"][3 * i:3 * i + 3])


# ---- live comparisons with the reference build (this container only) ----

def test_corotation_live_random(ref_lib):
    rng = np.random.default_rng(7)
    lam, mu = 1e3, 4e5
    for _ in range(300):
        F = rng.standard_normal((3, 3)) * rng.uniform(0.1, 3)
        dF = rng.standard_normal((3, 3))
        a = O.corotation(lam, mu, F, dF)
        b = O.corotation(lam, mu, F, dF, which="ref")
        tol = corot_tolerances("general", lam, mu, F, dF)
        for x, y, k in zip(a, b, ("R", "S", "L", "P", "e", "dP")):
            _nan_equal_close(np.ravel(x), np.ravel(y), tol[k] * 10)


def test_interpolation_live_out_of_range(ref_lib):
    rng = np.random.default_rng(8)
    X, Y, Z, dh = 4, 5, 6, 0.1
    o = np.array([0.5, -0.25, 0.0])
    f = rng.standard_normal(X * Y * Z)
    pts = rng.uniform(-1, 2, (300, 3))
    out = np.zeros(300)
    ref_lib.ref_interp_scalar(O.dp(f), X, Y, Z, dh, O.dp(o), 300,
                              O.dp(np.ascontiguousarray(pts.reshape(-1))), O.dp(out))
    for i in range(300):
        p = np.ascontiguousarray(pts[i])
        assert O.lib().orc_interp_scalar(O.dp(f), X, Y, Z, dh, O.dp(o), O.dp(p)) == out[i]
