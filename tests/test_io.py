"""On-disk formats: Mat<T> files (util/mat.h:178-215) and the normal-map
encoding (normal_map.cc:133-176).  The library's encoder is checked bit for
bit against the numpy restatement in oracle/io_oracle.py, which is itself
pinned by closed forms (the axis normal, the clamp of a NaN coordinate) and
by the decode(encode(n)) ~= n round trip; the file layout is checked byte
for byte."""
import struct

import numpy as np
import pytest

from oracle import io_oracle
from sensemap_amd import io as pmio


def test_mat_file_layout_is_text_header_then_little_endian_values(tmp_path):
    p = tmp_path / "d.bin"
    a = np.arange(6, dtype=np.float32).reshape(2, 3) * 0.5
    pmio.write_depth_map(str(p), a)
    raw = p.read_bytes()
    assert raw.startswith(b"3&2&1&")
    body = raw[len(b"3&2&1&"):]
    assert body == struct.pack("<6f", *a.ravel())


@pytest.mark.parametrize("dtype", [np.float32, np.uint32, np.uint16, np.uint8, np.float64])
def test_mat_round_trip(tmp_path, dtype):
    rng = np.random.default_rng(3)
    a = (rng.random((3, 17, 29)) * 200).astype(dtype)
    p = str(tmp_path / "m.bin")
    pmio.write_mat(p, a)
    b = pmio.read_mat(p, dtype)
    assert b.shape == a.shape and np.array_equal(a, b)


def test_header_with_leading_whitespace_and_wrong_size(tmp_path):
    p = tmp_path / "w.bin"
    p.write_bytes(b" 2& 1&1&" + struct.pack("<2f", 1.0, 2.0))
    assert np.array_equal(pmio.read_depth_map(str(p)), np.array([[1.0, 2.0]], np.float32))
    q = tmp_path / "short.bin"
    q.write_bytes(b"4&4&1&" + b"\0" * 8)
    with pytest.raises(Exception):
        pmio.read_depth_map(str(q))


def unit_normals(n, seed=0):
    rng = np.random.default_rng(seed)
    v = rng.standard_normal((3, n)).astype(np.float64)
    v /= np.linalg.norm(v, axis=0)
    return v.astype(np.float32)


def test_encoder_matches_restatement_bitwise():
    n = unit_normals(20000)
    special = np.array([[1, 0, 0], [-1, 0, 0], [0, 0, 0], [0, 1, 0], [0, 0, -1], [np.nan, 0, 0],
                        [0.6, 0.8, 0.0], [1e-8, -1.0, 0.0]], np.float32).T
    n = np.concatenate([n, special], axis=1).reshape(3, 1, -1)
    got = pmio.encode_normals(n)
    want = io_oracle.encode(n)
    assert np.array_equal(got, want)


def test_decoder_matches_restatement_and_round_trips():
    n = unit_normals(5000, seed=1).reshape(3, 50, 100)
    codes = pmio.encode_normals(n)
    dec = pmio.decode_normals(codes)
    want = io_oracle.decode(codes)
    assert np.array_equal(dec.view(np.uint32), want.view(np.uint32))
    # 16-bit quantisation of the parameterisation; it degrades towards
    # n = (-1, 0, 0), where the reference decoder can even return NaN
    # (quantised q1^2 + q2^2 above 1)
    ok = np.isfinite(dec).all(axis=0) & (n[0] > -0.99)
    assert ok.mean() > 0.99
    assert np.max(np.abs(dec - n)[:, ok]) < 1e-2
    assert np.median(np.abs(dec - n)) < 1e-4


def test_closed_forms():
    # the normal (1, 0, 0) maps to the centre code
    c = io_oracle.encode(np.array([[1.0], [0.0], [0.0]], np.float32))
    assert int(c[0]) == (32767 << 16) | 32767
    # a NaN coordinate clamps to the upper bound (std::min(0.999999f, NaN))
    c = io_oracle.encode(np.array([[0.0], [np.nan], [0.0]], np.float32))
    assert int(c[0]) >> 16 == int((np.float32(0.999999) + np.float32(1)) * np.float32(32767))
    d = io_oracle.decode(np.array([(32767 << 16) | 32767], np.uint32))
    assert abs(d[0, 0] - 1.0) < 1e-6 and abs(d[1, 0]) < 1e-4 and abs(d[2, 0]) < 1e-4


def test_normal_map_file_round_trip(tmp_path):
    n = unit_normals(64 * 48, seed=2).reshape(3, 48, 64)
    p = str(tmp_path / "n.bin")
    pmio.write_normal_map(p, n)
    raw = open(p, "rb").read()
    assert raw.startswith(b"64&48&1&") and len(raw) == len(b"64&48&1&") + 64 * 48 * 4
    back = pmio.read_normal_map(p)
    want = pmio.decode_normals(pmio.encode_normals(n))
    assert np.array_equal(back.view(np.uint32), want.view(np.uint32))
