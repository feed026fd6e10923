"""Readers for the files overlapInCore writes: .ovb (ovFile full records), .oc (counts), .stats.

Formats as canu writes them:
- .ovb: buffers of uint32 (a_iid, b_iid, hi32(w0), lo32(w0), hi32(w1), lo32(w1)), each a raw
  snappy block preceded by its uint64 compressed length (stores/ovStoreFile.C:190-283);
- .oc: uint64 overlaps, uint32 reads + 1, uint32 overlaps per read (stores/ovStoreFile.H:50-110);
- .stats: 8 "name = value" lines (overlapInCore/overlapInCore.C:539-560).

The snappy decoder follows the published raw-block format (varint length, then literal and
copy elements); it needs no snappy library.
"""
import struct

import numpy as np

RECORD_DTYPE = np.dtype([("a_iid", "<u4"), ("b_iid", "<u4"), ("w0", "<u8"), ("w1", "<u8")])


def snappy_uncompress(buf):
    """Decode one raw snappy block."""
    mv = memoryview(buf)
    n, shift, i = 0, 0, 0
    while True:
        b = mv[i]
        i += 1
        n |= (b & 0x7F) << shift
        if b < 0x80:
            break
        shift += 7
    out = bytearray()
    while i < len(mv):
        tag = mv[i]
        i += 1
        kind = tag & 3
        if kind == 0:
            ln = tag >> 2
            if ln >= 60:
                nb = ln - 59
                ln = int.from_bytes(mv[i:i + nb], "little")
                i += nb
            ln += 1
            out += mv[i:i + ln]
            i += ln
            continue
        if kind == 1:
            ln = ((tag >> 2) & 7) + 4
            off = ((tag >> 5) << 8) | mv[i]
            i += 1
        elif kind == 2:
            ln = (tag >> 2) + 1
            off = mv[i] | (mv[i + 1] << 8)
            i += 2
        else:
            ln = (tag >> 2) + 1
            off = int.from_bytes(mv[i:i + 4], "little")
            i += 4
        if off == 0 or off > len(out):
            raise ValueError("bad snappy copy offset")
        start = len(out) - off
        if off >= ln:
            out += out[start:start + ln]
        else:
            for k in range(ln):
                out.append(out[start + k])
    if len(out) != n:
        raise ValueError(f"snappy block decodes to {len(out)} bytes, header says {n}")
    return bytes(out)


def read_ovb(path):
    """All records of a full-format overlap file, in file order."""
    words = []
    with open(path, "rb") as f:
        while True:
            h = f.read(8)
            if not h:
                break
            (bl,) = struct.unpack("<Q", h)
            words.append(np.frombuffer(snappy_uncompress(f.read(bl)), dtype="<u4"))
    w = np.concatenate(words) if words else np.zeros(0, "<u4")
    if len(w) % 6:
        raise ValueError("truncated .ovb")
    w = w.reshape(-1, 6).astype(np.uint64)
    rec = np.zeros(len(w), RECORD_DTYPE)
    rec["a_iid"] = w[:, 0]
    rec["b_iid"] = w[:, 1]
    rec["w0"] = (w[:, 2] << np.uint64(32)) | w[:, 3]
    rec["w1"] = (w[:, 4] << np.uint64(32)) | w[:, 5]
    return rec


def read_oc(path):
    """(overlaps, per-read counts) of a .oc file."""
    with open(path, "rb") as f:
        (n,) = struct.unpack("<Q", f.read(8))
        (m,) = struct.unpack("<I", f.read(4))
        opr = np.frombuffer(f.read(4 * m), dtype="<u4")
    return n, opr


def read_stats(path):
    """The 8 counters of a .stats file, in file order."""
    return [int(line.split("=")[1]) for line in open(path) if "=" in line]
