"""CPU-side checks of the product: libesf.so loads and exports every entry
point include/esf.h declares; the host cfg reader (C++, esf_cfg_read) and
its Python mirror read cfgs exactly as the reference's SIMPLE_PARSER does."""
import ctypes
import json
import os
import tempfile
import re
import subprocess

import numpy as np
import pytest

import eulersf_amd
from eulersf_amd import _abi
from eulersf_amd.cfg import Cfg, params_from_cfg, unit_cube_ball_cfg

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = os.path.join(ROOT, "tests", "golden")


def header_symbols():
    text = open(os.path.join(ROOT, "include", "esf.h")).read()
    return sorted(set(re.findall(r"^(?:esf_status|void|const char\*)\s+(esf_\w+)\(", text, re.M)))


def test_library_exports_every_header_symbol():
    lib = _abi.load()
    out = subprocess.run(["nm", "-D", "--defined-only", _abi.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    exported = set(re.findall(r" T (esf_\w+)", out))
    hdr = header_symbols()
    assert len(hdr) > 30
    missing = [s for s in hdr if s not in exported]
    assert not missing, missing
    for s in hdr:
        getattr(lib, s)
    # the Python bindings cover the header exactly
    assert sorted(_abi.exported_symbols()) == hdr


def test_library_is_gfx950_code():
    """the device code object in libesf.so (built here) targets gfx950 only"""
    data = open(_abi.LIB_PATH, "rb").read()
    targets = set(re.findall(rb"amdgcn-amd-amdhsa--(gfx\w+)", data))
    assert targets == {b"gfx950"}, targets


def read_with_host(text):
    lib = _abi.load()
    with tempfile.TemporaryDirectory() as tmp:
        path = os.path.join(tmp, "host_cfg_test.cfg")
        with open(path, "w") as f:
            f.write(text)
        g, p, info = _abi.Grid(), _abi.Params(), _abi.SceneInfo()
        st = lib.esf_cfg_read(path.encode(), ctypes.byref(g), ctypes.byref(p),
                              ctypes.byref(info))
    return st, g, p, info


def test_host_cfg_reader_matches_reference_parser():
    gold = json.load(open(os.path.join(GOLD, "parser.json")))
    st, g, p, info = read_with_host(gold["text"])
    assert st == 0
    assert (g.xres, g.yres, g.zres) == (gold["int"]["xres"], gold["int"]["yres"], gold["int"]["zres"])
    assert g.dh == gold["float"]["dh"]
    assert list(g.origin) == gold["vec3"]["origin"]
    assert p.vorticity_eps == gold["float"]["vorticity eps"]
    assert p.cr_max_iterations == gold["int"]["cr max iterations"]
    c = Cfg(text=gold["text"])
    for k, v in gold["float"].items():
        assert c.get_float(k, -7.0) == v, k
    for k, v in gold["int"].items():
        assert c.get_int(k, -7) == v, k
    for k, v in gold["string"].items():
        assert c.get_string(k, "default") == v, k
    for k, v in gold["vec3"].items():
        assert c.get_vec3(k, (9.0, 8.0, 7.0)) == v, k


def test_cfg_parse_error_stops_like_the_reference():
    gold = json.load(open(os.path.join(GOLD, "parser.json")))["bad"]
    assert gold["parse_ok"] == 0
    st, *_ = read_with_host(gold["text"])
    assert st == _abi.ESF_ERR_IO
    c = Cfg()
    with pytest.raises(ValueError):
        c.parse_text(gold["text"])
    # keys before the bad line are kept, keys after it are not
    assert c.get_int("xres", -7) == gold["xres"]
    assert c.get_int("yres", -7) == gold["yres"]


def test_host_defaults_are_the_reference_defaults():
    st, g, p, info = read_with_host("xRes = 10\n")
    assert st == 0
    assert (g.xres, g.yres, g.zres, g.dh) == (10, 100, 100, 0.01)
    assert p.dt == 0.01 and p.vorticity_eps == 2.0 and p.gravity == 0.0
    assert p.buoyancy == 0.1 and p.fluid_density == 1.0
    assert p.implicit_eps == 1e-8 and p.divfree_eps == 1e-8 and p.cr_max_iterations == 10
    assert p.mass_sample_dh == 0.01 / 10 and p.tight_border == 3 and p.extrapolation_cells == 20
    assert info.frame_dt == 0.005 and info.carpet_thickness == 0.06


def test_shipped_reference_cfgs_read_identically():
    """the reference's own cfg files (read in place where the reference tree
    exists), by the C++ host and the Python mirror"""
    for name in ("ball.lowres.cfg", "ball.highres.cfg"):
        path = os.path.join("/root/reference/cfg", name)
        if not os.path.exists(path):
            pytest.skip("reference tree not present")
        text = open(path).read()
        st, g, p, info = read_with_host(text)
        c = Cfg(text=text)
        assert st == 0
        assert (g.xres, g.yres, g.zres) == (c.get_int("xRes", 100), c.get_int("yRes", 100),
                                            c.get_int("zRes", 100))
        pp = params_from_cfg(c, g.dh)
        for k, v in pp.items():
            assert getattr(p, k) == v, k


def test_unit_cube_cfg_round_trip():
    text = unit_cube_ball_cfg(64, eps=1e-10)
    st, g, p, info = read_with_host(text)
    assert st == 0
    assert (g.xres, g.dh) == (65, 1.0 / 64)
    assert p.divfree_eps == 1e-10 and p.implicit_eps == 1e-10
    assert p.mass_sample_dh == 1.0 / 512
    lam, mu = 5769.2307692307695, 3846.153846153846
    c = Cfg(text=text)
    assert abs(c.get_float("material lambda 0", 0) - lam) < 1e-9
    assert abs(c.get_float("material mu 0", 0) - mu) < 1e-9
    # dh / mass sample dh is an exact integer (the reference truncates it)
    assert (1.0 / 64) / p.mass_sample_dh == 8.0


def test_no_gpu_calls_fail_cleanly():
    """creating a context without a GPU is an error, not a crash"""
    lib = _abi.load()
    if os.path.exists("/dev/kfd"):
        pytest.skip("a GPU driver is present")
    ctx = ctypes.c_void_p()
    g = _abi.Grid(9, 9, 9, 0.125, (ctypes.c_double * 3)(0, 0, 0))
    p = _abi.Params()
    st = lib.esf_ctx_create(ctypes.byref(ctx), ctypes.byref(g), ctypes.byref(p))
    assert st == _abi.ESF_ERR_DEVICE
