"""psi snapshots of split-operator evolutions (checkpoint / resume).

The state of a reference evolution is psi (`fPsi`, mirrored in `fPsiHost`)
and the step count `fStep` (src/EvolverImpl.h:20-22); the reference has no
checkpoint format, and its text dumpers (src/Dump.h:5-14) are unused.  A
snapshot here holds that same state, psi and the step count, with the grid
and step parameters it belongs to, so an evolver built with the same
parameters resumes bit for bit.

File layout (little-endian):
    8 bytes    magic b"QSIMPSI1"
    8 bytes    u64 length H of the header
    H bytes    UTF-8 JSON header: format, dim, shape, extent, dt [re, im],
               mass, hbar, method, boundary, precision, step, dtype "<c16",
               crc32 of the data bytes
    16*N bytes psi, complex128 in the grid's row-major order (idx = i*ny + j)
"""
import json
import os
import struct
import zlib

import numpy as np

MAGIC = b"QSIMPSI1"
FORMAT = 1
_MAX_HEADER = 1 << 20


def write(path, header, psi):
    """Write `psi` (complex128, C order) with `header` (a dict of the grid
    and step parameters) to `path`, replacing it atomically."""
    psi = np.ascontiguousarray(psi, dtype="<c16")
    data = psi.tobytes()
    h = dict(header)
    h.update(format=FORMAT, shape=list(psi.shape), dtype="<c16",
             crc32=zlib.crc32(data) & 0xFFFFFFFF)
    hb = json.dumps(h, sort_keys=True).encode("utf-8")
    tmp = path + ".tmp"
    with open(tmp, "wb") as f:
        f.write(MAGIC)
        f.write(struct.pack("<Q", len(hb)))
        f.write(hb)
        f.write(data)
    os.replace(tmp, path)


def read(path):
    """Return (header, psi) from a snapshot file; ValueError if the file is
    not a snapshot, is truncated or fails its checksum."""
    with open(path, "rb") as f:
        if f.read(8) != MAGIC:
            raise ValueError("%s: not a psi snapshot" % path)
        raw = f.read(8)
        if len(raw) != 8:
            raise ValueError("%s: truncated header" % path)
        (hl,) = struct.unpack("<Q", raw)
        if hl > _MAX_HEADER:
            raise ValueError("%s: header length %d out of range" % (path, hl))
        hb = f.read(hl)
        if len(hb) != hl:
            raise ValueError("%s: truncated header" % path)
        try:
            h = json.loads(hb.decode("utf-8"))
        except ValueError as e:
            raise ValueError("%s: bad header (%s)" % (path, e)) from None
        if not isinstance(h, dict):
            raise ValueError("%s: unsupported header (not a JSON object)" % path)
        if h.get("format") != FORMAT or h.get("dtype") != "<c16":
            raise ValueError("%s: unsupported format %r / dtype %r"
                             % (path, h.get("format"), h.get("dtype")))
        shape = h.get("shape")
        if (not isinstance(shape, list) or not shape
                or not all(isinstance(s, int) and s >= 1 for s in shape)
                or not isinstance(h.get("crc32"), int)):
            raise ValueError("%s: header lacks a valid shape or crc32" % path)
        shape = tuple(shape)
        n = int(np.prod(shape))
        data = f.read(16 * n + 1)
    if len(data) != 16 * n:
        raise ValueError("%s: %d data bytes, expected %d" % (path, len(data), 16 * n))
    if zlib.crc32(data) & 0xFFFFFFFF != h["crc32"]:
        raise ValueError("%s: checksum mismatch" % path)
    return h, np.frombuffer(data, dtype="<c16").astype(np.complex128).reshape(shape)
