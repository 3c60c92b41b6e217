"""psi snapshot files (qusim_amd/snapshot.py): round trip and rejection of
damaged files.  Host only."""
import os
import struct

import numpy as np
import pytest

from qusim_amd import snapshot

HEADER = {"dim": 2, "extent": [-150.0, 150.0, -150.0, 150.0], "dt": [0.2, 0.0],
          "mass": 1.0, "hbar": 1.0, "method": "O2", "boundary": "Period", "step": 17,
          "precision": 64}


def _psi(shape=(64, 32), seed=3):
    r = np.random.default_rng(seed)
    return r.standard_normal(shape) + 1j * r.standard_normal(shape)


def test_round_trip_is_exact(tmp_path):
    p = str(tmp_path / "a.qpsi")
    psi = _psi()
    # values whose decimal forms do not round-trip in 17 digits still must
    psi[0, 0] = complex(np.nextafter(0.1, 1.0), -np.finfo(float).tiny)
    snapshot.write(p, HEADER, psi)
    h, got = snapshot.read(p)
    assert got.dtype == np.complex128 and got.shape == psi.shape
    assert np.array_equal(got.view(np.uint64), psi.view(np.uint64))
    for k, v in HEADER.items():
        assert h[k] == v
    assert not os.path.exists(p + ".tmp")


def test_one_dimensional_and_empty_header_fields(tmp_path):
    p = str(tmp_path / "b.qpsi")
    psi = _psi((1024,), seed=5)
    snapshot.write(p, {"dim": 1, "step": 0}, psi)
    h, got = snapshot.read(p)
    assert h["shape"] == [1024] and np.array_equal(got, psi)


def test_flipped_data_byte_fails_the_checksum(tmp_path):
    p = str(tmp_path / "c.qpsi")
    snapshot.write(p, HEADER, _psi())
    raw = bytearray(open(p, "rb").read())
    raw[-100] ^= 0x01
    open(p, "wb").write(bytes(raw))
    with pytest.raises(ValueError, match="checksum"):
        snapshot.read(p)


def test_truncated_and_overlong_files_are_rejected(tmp_path):
    p = str(tmp_path / "d.qpsi")
    snapshot.write(p, HEADER, _psi())
    raw = open(p, "rb").read()
    open(p, "wb").write(raw[:-16])
    with pytest.raises(ValueError, match="data bytes"):
        snapshot.read(p)
    open(p, "wb").write(raw + b"\0" * 16)
    with pytest.raises(ValueError, match="data bytes"):
        snapshot.read(p)
    open(p, "wb").write(raw[:12])
    with pytest.raises(ValueError, match="truncated"):
        snapshot.read(p)


def test_foreign_files_are_rejected(tmp_path):
    p = str(tmp_path / "e.qpsi")
    open(p, "wb").write(b"(0.1, 0.2)\n" * 8)   # a text dump, not a snapshot
    with pytest.raises(ValueError, match="not a psi snapshot"):
        snapshot.read(p)
    open(p, "wb").write(snapshot.MAGIC + struct.pack("<Q", 1 << 40))
    with pytest.raises(ValueError, match="out of range"):
        snapshot.read(p)
    hb = b'{"format": 1, "dtype": "<c8", "shape": [1], "crc32": 0}'
    open(p, "wb").write(snapshot.MAGIC + struct.pack("<Q", len(hb)) + hb + b"\0" * 8)
    with pytest.raises(ValueError, match="unsupported"):
        snapshot.read(p)


def test_headers_without_shape_or_checksum_are_rejected(tmp_path):
    p = str(tmp_path / "f.qpsi")
    for hb in (b'{"format": 1, "dtype": "<c16", "crc32": 0}',
               b'{"format": 1, "dtype": "<c16", "shape": [0], "crc32": 0}',
               b'{"format": 1, "dtype": "<c16", "shape": [2, "x"], "crc32": 0}',
               b'{"format": 1, "dtype": "<c16", "shape": [1]}',
               b'[1, 2, 3]'):
        open(p, "wb").write(snapshot.MAGIC + struct.pack("<Q", len(hb)) + hb + b"\0" * 16)
        with pytest.raises(ValueError):
            snapshot.read(p)
