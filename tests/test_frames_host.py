"""The reference's gz field format (FIELD_3D / VEC3F_FIELD_3D writeGz and
readGz, FIELD_3D.h:286-365, VEC3F_FIELD_3D.h:199-267) on the host:
libesf.so's writer is read back by the reference's own readGz and the
reference's writeGz output by libesf.so's reader, payloads and headers
identical, decompressed streams byte-identical.  Host only (no GPU)."""
import ctypes
import gzip
import os

import numpy as np
import pytest

from eulersf_amd import _abi
from oracle import oracle as O

needs_ref = pytest.mark.skipif(not O.ref_available(), reason="oracle/_ref not built here")


def ref_write(path, grid, data, comps):
    X, Y, Z, dh, o = grid
    oo = np.array(o, dtype=np.float64)
    a = np.ascontiguousarray(data, dtype=np.float64)
    O.ref().ref_field_write_gz(path.encode(), X, Y, Z, dh, O.dp(oo), comps, O.dp(a))


def ref_read(path, comps, n):
    dims = (ctypes.c_int * 3)()
    hdr = np.zeros(4)
    out = np.zeros(comps * n)
    got = O.ref().ref_field_read_gz(path.encode(), comps, dims, O.dp(hdr), O.dp(out), out.size)
    return (dims[0], dims[1], dims[2], hdr[0], tuple(hdr[1:])), out, got


GRIDS = [(5, 7, 3, 0.125, (0.0, 0.0, 0.0)), (9, 4, 6, 1.0 / 64, (-0.25, 0.5, 1e-3))]


@needs_ref
@pytest.mark.parametrize("grid", GRIDS)
@pytest.mark.parametrize("comps", [1, 3])
def test_writer_read_by_reference(tmp_path, grid, comps):
    X, Y, Z = grid[:3]
    rng = np.random.default_rng(X * 100 + comps)
    data = rng.standard_normal(comps * X * Y * Z)
    data[::7] = 0.0
    data[1] = -0.0
    p = str(tmp_path / "ours.gz")
    _abi.field_write_gz(p, grid, data)
    g, back, n = ref_read(p, comps, X * Y * Z)
    assert n == X * Y * Z
    assert g == grid
    np.testing.assert_array_equal(back.view(np.uint64), data.view(np.uint64))


@needs_ref
@pytest.mark.parametrize("grid", GRIDS)
@pytest.mark.parametrize("comps", [1, 3])
def test_reference_writer_read_back(tmp_path, grid, comps):
    X, Y, Z = grid[:3]
    rng = np.random.default_rng(X * 10 + comps)
    data = rng.standard_normal(comps * X * Y * Z)
    pr = str(tmp_path / "ref.gz")
    po = str(tmp_path / "ours.gz")
    ref_write(pr, grid, data, comps)
    g, back = _abi.field_read_gz(pr, comps)
    assert g == grid
    np.testing.assert_array_equal(back.view(np.uint64), data.view(np.uint64))
    # the same bytes once decompressed: header layout and payload order
    _abi.field_write_gz(po, grid, data)
    assert gzip.open(pr).read() == gzip.open(po).read()


def test_header_layout(tmp_path):
    """int32 X, Y, Z; double dh, origin[3]; then the payload, x fastest"""
    grid = (3, 2, 2, 0.5, (1.0, 2.0, 3.0))
    data = np.arange(12, dtype=np.float64)
    p = str(tmp_path / "f.gz")
    _abi.field_write_gz(p, grid, data)
    raw = gzip.open(p).read()
    assert len(raw) == 12 + 32 + 8 * 12
    assert np.frombuffer(raw[:12], dtype=np.int32).tolist() == [3, 2, 2]
    assert np.frombuffer(raw[12:44], dtype=np.float64).tolist() == [0.5, 1.0, 2.0, 3.0]
    np.testing.assert_array_equal(np.frombuffer(raw[44:], dtype=np.float64), data)


def test_read_errors(tmp_path):
    lib = _abi.load()
    g = _abi.Grid()
    assert lib.esf_field_read_gz(str(tmp_path / "absent.gz").encode(), ctypes.byref(g), 1,
                                 None, 0) == _abi.ESF_ERR_IO
    p = str(tmp_path / "f.gz")
    _abi.field_write_gz(p, (4, 4, 4, 0.25, (0.0, 0.0, 0.0)), np.ones(64))
    small = np.zeros(10)
    assert lib.esf_field_read_gz(p.encode(), ctypes.byref(g), 1, _abi.dptr(small),
                                 small.size) == _abi.ESF_ERR_INVALID
    # a truncated payload is an I/O error, not a partial field
    raw = gzip.open(p).read()
    with gzip.open(p, "wb") as f:
        f.write(raw[:-8])
    full = np.zeros(64)
    assert lib.esf_field_read_gz(p.encode(), ctypes.byref(g), 1, _abi.dptr(full),
                                 full.size) == _abi.ESF_ERR_IO
    assert lib.esf_field_write_gz(p.encode(), ctypes.byref(g), 2, _abi.dptr(full)) == \
        _abi.ESF_ERR_INVALID
