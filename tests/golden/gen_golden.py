"""Write the golden vectors under tests/golden/ from the reference's own code.

Runs only where oracle/_ref/libesfref.so could be built (the reference tree
is present): it calls the reference's COROTATION, CONJUGATE_RESIDUAL,
FIELD_3D / VEC3F_FIELD_3D interpolation and SIMPLE_PARSER through
oracle/ref_driver.cpp and stores inputs and outputs as .npz / .json data.

    python tests/golden/gen_golden.py
"""
import ctypes
import json
import os
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import oracle as O  # noqa: E402

LAM, MU = 5769.2307692307695, 3846.153846153846


def deformation_gradients(rng):
    """F cases covering every branch of COROTATION::svd/orthogonalizeU/
    removeUReflection and the computeF fallback."""
    cases = []

    def rot():
        q, _ = np.linalg.qr(rng.standard_normal((3, 3)))
        if np.linalg.det(q) < 0:
            q[:, 0] *= -1
        return q

    for _ in range(60):
        cases.append(("near_identity", np.eye(3) + 0.05 * rng.standard_normal((3, 3))))
    for _ in range(60):
        cases.append(("general", rng.standard_normal((3, 3))))
    for _ in range(40):
        cases.append(("rot_stretch", rot() @ np.diag(rng.uniform(0.3, 2.0, 3)) @ rot().T))
    for _ in range(40):
        s = rng.uniform(0.3, 2.0, 3)
        s[rng.integers(3)] *= -1
        cases.append(("inverted", rot() @ np.diag(s) @ rot().T))
    for _ in range(20):
        s = rng.uniform(0.5, 1.5, 3)
        s[rng.integers(3)] = rng.uniform(0, 5e-5)
        cases.append(("one_small", rot() @ np.diag(s) @ rot().T))
    for _ in range(10):
        s = rng.uniform(0.5, 1.5, 3)
        k = rng.permutation(3)[:2]
        s[k] = rng.uniform(0, 5e-5, 2)
        cases.append(("two_small", rot() @ np.diag(s) @ rot().T))
    for _ in range(4):
        cases.append(("all_small", rot() @ np.diag(rng.uniform(0, 5e-5, 3)) @ rot().T))
    cases.append(("identity", np.eye(3)))
    cases.append(("zero", np.zeros((3, 3))))
    cases.append(("exact_rotation", rot()))
    cases.append(("reflection", np.diag([1.0, 1.0, -1.0])))
    cases.append(("diag_inverted", np.diag([2.0, 1.5, -0.5])))
    return cases


def gen_corotation(rng):
    cases = deformation_gradients(rng)
    F = np.array([c[1] for c in cases]).reshape(-1, 9)
    dF = rng.standard_normal(F.shape)
    out = np.zeros((len(cases), 36))
    energy = np.zeros(len(cases))
    dP = np.zeros((len(cases), 9))
    for i in range(len(cases)):
        R, S, L, P, e, d = O.corotation(LAM, MU, F[i], dF[i], which="ref")
        out[i] = np.concatenate([R.ravel(), S.ravel(), L.ravel(), P.ravel()])
        energy[i] = e
        dP[i] = d.ravel()
    np.savez_compressed(os.path.join(HERE, "corotation.npz"), F=F, dF=dF, out=out,
                        energy=energy, dP=dP, kind=np.array([c[0] for c in cases]),
                        lam=LAM, mu=MU)


def gen_cr(rng):
    """CONJUGATE_RESIDUAL::solvePCR on small systems, both variants:
    SPD (assembled CSR variant, x = 0 start) and the singular-consistent
    pressure operator of a tiny scene (matrix-free variant, warm start)."""
    R = O.ref()
    systems = []
    for n, cond in ((30, 1e2), (40, 1e4), (25, 1e6)):
        q, _ = np.linalg.qr(rng.standard_normal((n, n)))
        A = q @ np.diag(np.logspace(0, np.log10(cond), n)) @ q.T
        A = 0.5 * (A + A.T)
        b = rng.standard_normal(n)
        rows, cols = np.nonzero(np.ones((n, n)))
        vals = A[rows, cols]
        x = np.zeros(n)
        it = ctypes.c_int()
        ok = R.ref_pcr_csr(n, n * n, np.ascontiguousarray(rows, dtype=np.int32).ctypes.data_as(O._I),
                           np.ascontiguousarray(cols, dtype=np.int32).ctypes.data_as(O._I),
                           O.dp(np.ascontiguousarray(vals)), O.dp(b), O.dp(x), 1e-10, 2000,
                           ctypes.byref(it))
        systems.append({"A": A, "b": b, "x0": np.zeros(n), "x": x.copy(), "iters": it.value,
                        "ok": ok, "eps": 1e-10, "variant": "csr"})
    # pressure operator of a 6^3-cell scene with real masses, dense
    from eulersf_amd.cfg import unit_cube_ball_cfg
    sc = O.OracleScene.from_cfg_text(unit_cube_ball_cfg(6, radius=0.3, center=(0.5, 0.5, 0.5)))
    sc.L.orc_forces(sc.h)
    nc = sc.cells
    S = np.zeros((nc, nc))
    for j in range(nc):
        e = np.zeros(nc)
        e[j] = 1.0
        S[:, j] = sc.pressure_apply(e)
    D = sc.pressure_diag()
    v = rng.standard_normal(nc)
    b = S @ v  # consistent right-hand side
    for warm in (0, 1):
        x0 = 0.1 * rng.standard_normal(nc) if warm else np.zeros(nc)
        x = x0.copy()
        it = ctypes.c_int()
        ok = R.ref_pcr_matfree_dense(nc, O.dp(np.ascontiguousarray(S)), O.dp(D), O.dp(b), O.dp(x),
                                     warm, 1e-12, 5000, ctypes.byref(it))
        systems.append({"A": S, "b": b, "x0": x0, "x": x.copy(), "iters": it.value, "ok": ok,
                        "eps": 1e-12, "variant": "matfree", "D": D})
    sc.close()
    arrays = {}
    meta = []
    for k, s in enumerate(systems):
        for key in ("A", "b", "x0", "x"):
            arrays[f"{key}{k}"] = s[key]
        if "D" in s:
            arrays[f"D{k}"] = s["D"]
        meta.append({"iters": s["iters"], "ok": s["ok"], "eps": s["eps"], "variant": s["variant"]})
    arrays["meta"] = np.array(json.dumps(meta))
    np.savez_compressed(os.path.join(HERE, "cr.npz"), **arrays)


def gen_interp(rng):
    R = O.ref()
    X, Y, Z, dh = 5, 6, 7, 0.25
    o = np.array([-0.3, 0.1, 0.2])
    f = rng.standard_normal(X * Y * Z)
    fv = rng.standard_normal(3 * X * Y * Z)
    lo = o - 2 * dh
    hi = o + dh * np.array([X + 1, Y + 1, Z + 1])
    pts = rng.uniform(lo, hi, (400, 3))
    pts[:20] = o + dh * rng.integers(0, 4, (20, 3))  # exact nodes
    out_s = np.zeros(len(pts))
    out_v = np.zeros(3 * len(pts))
    p = np.ascontiguousarray(pts.reshape(-1))
    R.ref_interp_scalar(O.dp(f), X, Y, Z, dh, O.dp(o), len(pts), O.dp(p), O.dp(out_s))
    R.ref_interp_vec3(O.dp(fv), X, Y, Z, dh, O.dp(o), len(pts), O.dp(p), O.dp(out_v))
    np.savez_compressed(os.path.join(HERE, "interp.npz"), dims=np.array([X, Y, Z]), dh=dh, o=o,
                        f=f, fv=fv, pts=pts, out_s=out_s, out_v=out_v)


CFG_TEXT = """# comment line
xRes = 33
  YRes=  17
zres = 9.9
dh = 0.03125
Origin = -0.5, 0.25 ,1e-1
solid path 0 = ./Meshes/Sphere/
Material Lambda 0 = 5.5e3
cr max iterations = 12.7

vorticity eps = 1.5abc
empty value =
dh = 0.0625
"""
# a line without '=' that is not blank stops the parse (SIMPLE_PARSER.cpp:96-103)
CFG_BAD = """xRes = 12
 # not a comment: column 1
yRes = 40
"""
KEYS_F = ["dh", "material lambda 0", "vorticity eps", "empty value", "missing key", "zres"]
KEYS_I = ["xres", "yres", "zres", "cr max iterations", "missing key"]
KEYS_S = ["solid path 0", "missing key"]
KEYS_V = ["origin", "missing key"]


def parse_with_ref(text):
    R = O.ref()
    fd, path = tempfile.mkstemp(suffix=".cfg")
    with os.fdopen(fd, "w") as fh:
        fh.write(text)
    ok = R.ref_parse(path.encode())
    os.unlink(path)
    return ok


def gen_parser():
    R = O.ref()
    bad_ok = parse_with_ref(CFG_BAD)
    bad = {"text": CFG_BAD, "parse_ok": bad_ok, "xres": R.ref_get_int(b"xres", -7),
           "yres": R.ref_get_int(b"yres", -7)}
    # SIMPLE_PARSER's store is process-global: later keys override, absent
    # keys of the second text may still hold values of the first; the keys
    # queried below all appear in CFG_TEXT or in neither text.
    ok = parse_with_ref(CFG_TEXT)
    res = {"text": CFG_TEXT, "parse_ok": ok, "float": {}, "int": {}, "string": {}, "vec3": {},
           "bad": bad}
    for k in KEYS_F:
        res["float"][k] = R.ref_get_float(k.encode(), -7.0)
    for k in KEYS_I:
        res["int"][k] = R.ref_get_int(k.encode(), -7)
    for k in KEYS_S:
        buf = ctypes.create_string_buffer(256)
        R.ref_get_string(k.encode(), b"default", buf, 256)
        res["string"][k] = buf.value.decode()
    for k in KEYS_V:
        out = np.zeros(3)
        R.ref_get_vec3(k.encode(), O.dp(np.array([9.0, 8.0, 7.0])), O.dp(out))
        res["vec3"][k] = out.tolist()
    with open(os.path.join(HERE, "parser.json"), "w") as fh:
        json.dump(res, fh, indent=1, sort_keys=True)


def main():
    if not O.ref_available():
        print("oracle/_ref/libesfref.so is not built; nothing to do", file=sys.stderr)
        return 1
    rng = np.random.default_rng(20261202)
    gen_corotation(rng)
    gen_cr(rng)
    gen_interp(rng)
    gen_parser()
    print("golden vectors written to", HERE)
    return 0


if __name__ == "__main__":
    sys.exit(main())
