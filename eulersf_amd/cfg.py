"""The reference's cfg format (src/util/SIMPLE_PARSER.cpp:64-236) and the
scene cfgs used for the unit-cube ball workloads.

Semantics kept from the reference: lines ``key = value``; a ``#`` in column 0
comments the line; keys and values are lowercased; floats go through
``atof``; integers are ``(int)atof`` (truncation); vectors split on ``,``.
A line without ``=`` that is not blank is a parse error.
"""
import math
import re

import numpy as np

_FLOAT_PREFIX = re.compile(r"^\s*[+-]?((\d+\.?\d*)|(\.\d+))([eE][+-]?\d+)?")


def atof(s):
    """C atof: the longest numeric prefix, 0.0 when there is none."""
    m = _FLOAT_PREFIX.match(s)
    if not m:
        low = s.strip().lower()
        for w, v in (("inf", math.inf), ("+inf", math.inf), ("-inf", -math.inf),
                     ("nan", math.nan)):
            if low.startswith(w):
                return v
        return 0.0
    return float(m.group(0))


class Cfg:
    def __init__(self, text=None, path=None):
        self.vals = {}
        if path is not None:
            with open(path, "r") as f:
                text = f.read()
        if text is not None:
            self.parse_text(text)

    def parse_text(self, text):
        for lineno, line in enumerate(text.split("\n"), 1):
            line = line[:511]
            if not line:
                continue
            if line[0] == "#":
                continue
            pos = line.find("=")
            if pos >= 0:
                self.vals[line[:pos].strip().lower()] = line[pos + 1:].strip().lower()
            elif line.strip():
                raise ValueError(f"unable to parse line {lineno}: {line!r}")

    def has(self, k):
        return k.lower() in self.vals

    def get_float(self, k, default):
        v = self.vals.get(k.lower())
        return default if v is None else atof(v)

    def get_int(self, k, default):
        v = self.vals.get(k.lower())
        return default if v is None else int(atof(v))

    def get_string(self, k, default):
        v = self.vals.get(k.lower())
        return default if v is None else v

    def get_vec3(self, k, default):
        v = self.vals.get(k.lower())
        if v is None:
            return list(default)
        toks = v.split(",")
        return [atof(toks[i]) if i < len(toks) else 0.0 for i in range(3)]


def params_from_cfg(cfg, dh):
    """cfg keys and defaults of the integrator, solver and solid
    (COUPLED_INTEGRATOR_3D.cpp:43-74, :445-446, :499-501;
    CONJUGATE_RESIDUAL.inl:26-27; DEFOR_SOLID_3D.cpp:62, :243-244)."""
    return {
        "dt": cfg.get_float("dt", 0.01),
        "vorticity_eps": cfg.get_float("vorticity eps", 2.0),
        "gravity": cfg.get_float("gravity", 0.0),
        "buoyancy": cfg.get_float("buoyancy", 0.1),
        "rayleigh_alpha": cfg.get_float("rayleigh alpha", 0.0),
        "rayleigh_beta": cfg.get_float("rayleigh beta", 0.0),
        "fluid_density": cfg.get_float("fluid density", 1.0),
        "implicit_eps": cfg.get_float("implicit solver eps", 1e-8),
        "divfree_eps": cfg.get_float("div free solver eps", 1e-8),
        "cr_max_iterations": cfg.get_int("cr max iterations", 10),
        "mass_sample_dh": cfg.get_float("mass sample dh", dh / 10),
        "tight_border": cfg.get_int("tight border", 3),
        "extrapolation_cells": cfg.get_int("extrapolation cells", 20),
        # FLIP particles (DEFOR_SOLID_3D.cpp:336-337, 901, 948-952, 1101); the
        # round trip runs in the step only when "flip particles" is set
        "particles_per_cell": cfg.get_int("number of particles per material cell", 8),
        "flip_alpha": cfg.get_float("flip alpha", 0.95),
        "flip_cutoff": cfg.get_int("flip cutoff", 3),
        "wall_thickness": cfg.get_float("wall thickness", dh * 1.5),
        "spring_stiffness": cfg.get_float("contact spring stiffness", 10000.0),
        "friction": cfg.get_float("friction", 0.5),
        "flip": cfg.get_int("flip particles", 0),
    }


def lame_from_young(E, nu):
    """(lambda, mu) of Young's modulus E and Poisson ratio nu."""
    return E * nu / ((1 + nu) * (1 - 2 * nu)), E / (2 * (1 + nu))


def mass_sample_dh(n):
    """sample spacing of the solid mass rasterisation on an n^3-cell unit
    cube: dh / k with k = min(8, floor(512 dh)) samples per cell edge, so
    dh / sdh is an exact integer (the reference truncates it) and
    density * sdh^3 stays above the reference's absolute 1e-6 sample cut
    (DEFOR_SOLID_3D.cpp:629-630) for density 1000"""
    dh = 1.0 / n
    k = max(1, min(8, int(512 * dh)))
    return dh / k


def scene_cfg(n, spheres, fill=0.3, eps=1e-10, max_iterations=20000, fluid_density=1000.0,
              dt=1.0 / 240.0, vorticity_eps=1.5, gravity=9.8, frames=250, buoyancy=0.0,
              rayleigh_alpha=0.0, rayleigh_beta=0.0, extra=None):
    """cfg text of a unit-cube scene on an n^3-cell grid ((n+1)^3 nodes,
    dh = 1/n): the smoke carpet (fluid pool) fills the bottom `fill` of the
    domain; `spheres` is a list of dicts with keys center, radius, E, nu,
    density, velocity (corotational spheres rasterised analytically)."""
    dh = 1.0 / n
    sdh = mass_sample_dh(n)
    lines = [
        "# unit-cube scene, %d^3 cells, %d sphere(s)" % (n, len(spheres)),
        "xRes = %d" % (n + 1), "yRes = %d" % (n + 1), "zRes = %d" % (n + 1),
        "dh = %r" % dh, "origin = 0,0,0",
        "number of solids = %d" % len(spheres),
    ]
    for i, sp in enumerate(spheres):
        lam, mu = lame_from_young(sp.get("E", 1e4), sp.get("nu", 0.3))
        lines += [
            "solid path %d = ./" % i, "solid name %d = analytic_sphere" % i,
            "solid sphere radius %d = %r" % (i, float(sp["radius"])),
            "sphere mass dh %d = %r" % (i, sdh),
            "solid mass density %d = %r" % (i, float(sp.get("density", 1000.0))),
            "material lambda %d = %r" % (i, lam), "material mu %d = %r" % (i, mu),
            "solid plastic yield %d = -1" % i,
            "solid initial velocity %d = %r, %r, %r" % ((i,) + tuple(float(v) for v in sp.get("velocity", (0, 0, 0)))),
            "solid translation %d = %r, %r, %r" % ((i,) + tuple(float(v) for v in sp["center"])),
            "solid scaling %d = 1,1,1" % i, "solid rotation %d = 0,0,0" % i,
        ]
    lines += [
        "fluid density = %r" % fluid_density, "fluid viscocity = 0",
        "vorticity eps = %r" % vorticity_eps, "buoyancy = %r" % float(buoyancy),
        "smoke carpet thickness = %r" % float(fill),
        "dt = %r" % dt, "gravity = %r" % gravity,
        "extrapolation cells = 20", "tight border = 5",
        "rayleigh alpha = %r" % float(rayleigh_alpha),
        "rayleigh beta = %r" % float(rayleigh_beta),
        "cr max iterations = %d" % max_iterations,
        "implicit solver eps = %r" % eps, "div free solver eps = %r" % eps,
        "mass sample dh = %r" % sdh,
        "start frame = 0", "end frame = %d" % frames,
        "save data = 0",
    ]
    lines += ["%s = %s" % (k, v) for k, v in (extra or {}).items()]
    return "\n".join(lines) + "\n"


def unit_cube_ball_cfg(n, eps=1e-10, max_iterations=20000, radius=0.15,
                       center=(0.5, 0.65, 0.5), E=1e4, nu=0.3, density=1000.0,
                       fluid_density=1000.0, carpet=0.3, dt=1.0 / 240.0,
                       vorticity_eps=1.5, gravity=9.8, frames=250, velocity=(0, 0, 0),
                       buoyancy=0.0, rayleigh_alpha=0.0, rayleigh_beta=0.0, extra=None):
    """cfg text of the unit-cube ball scene on an n^3-cell grid: (n+1)^3 nodes,
    dh = 1/n, one corotational sphere, the smoke carpet as the fluid pool;
    `extra`: further cfg keys and values"""
    sphere = {"center": center, "radius": radius, "E": E, "nu": nu, "density": density,
              "velocity": velocity}
    return scene_cfg(n, [sphere], fill=carpet, eps=eps, max_iterations=max_iterations,
                     fluid_density=fluid_density, dt=dt, vorticity_eps=vorticity_eps,
                     gravity=gravity, frames=frames, buoyancy=buoyancy,
                     rayleigh_alpha=rayleigh_alpha, rayleigh_beta=rayleigh_beta, extra=extra)


def random_spheres(rng, k, radius_range, log10_E_range, center_range=None, margin=0.02,
                   tries=10000):
    """k non-overlapping spheres (rejection sampling): radius uniform in
    radius_range, E log-uniform over 10^log10_E_range, nu = 0.3, centres
    uniform in center_range^3 (default: inside the unit cube with the
    sphere clear of the walls by `margin`)"""
    out = []
    for _ in range(tries):
        if len(out) == k:
            break
        r = float(rng.uniform(*radius_range))
        lo, hi = center_range if center_range else (r + margin, 1.0 - r - margin)
        c = rng.uniform(lo, hi, 3)
        if any(((c - np.asarray(o["center"])) ** 2).sum() < (r + o["radius"] + margin) ** 2
               for o in out):
            continue
        out.append({"center": tuple(float(v) for v in c), "radius": r,
                    "E": float(10 ** rng.uniform(*log10_E_range)), "nu": 0.3,
                    "density": 1000.0})
    if len(out) < k:
        raise ValueError("could not place %d non-overlapping spheres" % k)
    return out


def random_deformations(rng, n, stretch_sd=0.05):
    """n deformation gradients R diag(s): R a uniform random rotation (from a
    uniform unit quaternion, Shoemake's method), s_i ~ N(1, stretch_sd);
    returns (n, 3, 3), row-major"""
    u1, u2, u3 = rng.uniform(0, 1, (3, n))
    q = np.stack([np.sqrt(1 - u1) * np.sin(2 * np.pi * u2), np.sqrt(1 - u1) * np.cos(2 * np.pi * u2),
                  np.sqrt(u1) * np.sin(2 * np.pi * u3), np.sqrt(u1) * np.cos(2 * np.pi * u3)], 1)
    x, y, z, w = q.T
    R = np.stack([1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w),
                  2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w),
                  2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)], 1)
    sv = rng.normal(1.0, stretch_sd, (n, 3))
    return R.reshape(n, 3, 3) * sv[:, None, :]
