"""Shared scene builders: the same cfg text drives libesf (the product,
through esf_scene_load_cfg) and the CPU oracle."""
import numpy as np

from eulersf_amd.cfg import unit_cube_ball_cfg


def ball_cfg(n, **kw):
    return unit_cube_ball_cfg(n, **kw)


def interior_mask(X, Y, Z):
    m = np.zeros((Z, Y, X), dtype=bool)
    m[1:-1, 1:-1, 1:-1] = True
    return m.reshape(-1)


def rel_err(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    nb = np.linalg.norm(b)
    d = np.linalg.norm(a - b)
    return d / nb if nb > 0 else d


def solution_order(v_nodes, X, Y, Z):
    """AoS node field (3*nodes) -> the reference's solution order over
    interior nodes (3*(X-2)(Y-2)(Z-2))."""
    v = np.asarray(v_nodes).reshape(Z, Y, X, 3)[1:-1, 1:-1, 1:-1]
    return v.reshape(-1)
