"""canu seqStore directories: a writer, and a reader that hands reads over as overlapInCore gets them.

The overlapInCore command line reads stores with its own C++ reader (csrc/oic_seqstore.h); this
module writes stores for it and states the same hand-over rules in Python, so each reader can be
checked against the other.

Store layout (version 9), as canu's store code reads and writes it:
- info: tagged objects, each a 4-byte tag, a uint32 payload length and the payload
  (stores/sqStoreInfo.C:265-319);
- libraries: sqLibrary[NLIB + 1], 176 bytes each (stores/sqLibrary.H:86-138);
- reads: sqReadMeta[NREA + 1], two uint64 bit-field words (stores/sqRead.H:306-323);
- reads-rawu, -rawc, -coru, -corc: sqReadSeq[NREA + 1], three uint32 bit-field words
  (stores/sqRead.H:246-256);
- blobs.0001, ...: per read a BLOB chunk holding NAME and sequence chunks
  (stores/sqReadDataWriter.C:226-243, stores/sqReadData.C:48-91);
- homopolymerCompression: an empty marker file (stores/sqStoreCreate.C:342).

Not pinned by canu's own sources (they live in meryl-utility, which canu does not ship): the
chunk framing beyond what canu's chunk walk reads (no padding is written here), the 2-bit code
(A=0, C=1, G=2, T=3, first base in the high bits), the 3-bit code (never written here; reads with
N use the 8-bit code, which holds the bases as bytes), and the raw-array form of the metadata files.
"""
import os
import struct

import numpy as np

MAGIC = 0x5145533A756E6163          # "canu:SEQ"
VERSION = 9
RAW, CORRECTED, COMPRESSED, TRIMMED = 0x1, 0x2, 0x4, 0x8
TWO_BIT = "ACGT"
LIBRARY_SIZE, READ_SIZE = 176, 152   # sizeof(sqLibrary), sizeof(sqRead) on x86-64
PARAMS = [("LSIZ", LIBRARY_SIZE), ("RSIZ", READ_SIZE), ("MLB ", 6), ("LNS ", 128), ("MRB ", 37), ("MRLB", 21)]


def homopoly_compress(s):
    """First base of each maximal run of equal bases."""
    return "".join(c for i, c in enumerate(s) if i == 0 or s[i - 1] != c)


def homopoly_map(s):
    """ntoc[i]: position in the compressed sequence of the run holding base i; ntoc[len] = compressed length."""
    ntoc, c = [], -1
    for i, ch in enumerate(s):
        if i == 0 or s[i - 1] != ch:
            c += 1
        ntoc.append(c)
    ntoc.append(c + 1)
    return ntoc


def _chunk(tag, payload):
    return tag.encode() + struct.pack("<I", len(payload)) + payload


def _encode(bases):
    """(tag letter, payload): 2-bit when only ACGT, else 8-bit."""
    if all(c in TWO_BIT for c in bases):
        codes = np.frombuffer(bases.encode(), np.uint8)
        lut = np.zeros(256, np.uint8)
        for v, c in enumerate(TWO_BIT):
            lut[ord(c)] = v
        v = lut[codes]
        pad = (-len(v)) % 4
        v = np.concatenate([v, np.zeros(pad, np.uint8)]).reshape(-1, 4)
        packed = (v[:, 0] << 6) | (v[:, 1] << 4) | (v[:, 2] << 2) | v[:, 3]
        return "2", packed.astype(np.uint8).tobytes()
    return "U", bases.encode()


def _normalize(s):
    """Bases as canu stores them: uppercase, U as T (stores/sqRead.H:532-538)."""
    s = s.upper().replace("U", "T")
    bad = set(s) - set("ACGTN")
    if bad:
        raise ValueError(f"bases other than ACGTNU: {sorted(bad)}")
    return s


def _seq_words(valid, length, ignore_u=False, ignore_t=False, clear=None, trimmed=False):
    """sqReadSeq as three uint32 words (least significant bit first)."""
    bgn, end = (0, length) if clear is None else clear
    return [(1 if valid else 0) | (length << 2),
            (1 if ignore_u else 0) | (2 if (ignore_t or ignore_u) else 0) | (bgn << 2),
            (1 if trimmed else 0) | (end << 2)]


def write_store(path, raw=None, corrected=None, names=None, libraries=None, hpc_marker=False, clear=None,
                ignore=()):
    """Write a version-9 store of reads 1..n.

    raw, corrected: lists of sequences (None entries for absent ones); names: read names;
    libraries: per-read library IDs (1-based; default all 1); clear: {read id: (bgn, end)} clear
    ranges in uncompressed coordinates of the latest version, which marks it trimmed; ignore: read IDs
    whose sequences are flagged ignored.
    """
    n = len(raw if raw is not None else corrected)
    raw = [None] * n if raw is None else raw
    corrected = [None] * n if corrected is None else corrected
    names = names or [f"read{i + 1}" for i in range(n)]
    libraries = libraries or [1] * n
    clear = clear or {}
    nlib = max(libraries)
    os.makedirs(path)
    meta = np.zeros((n + 1, 2), np.uint64)
    seqs = {k: np.zeros((n + 1, 3), np.uint32) for k in ("rawu", "rawc", "coru", "corc")}
    blob = bytearray()
    latest = "cor" if any(c is not None for c in corrected) else "raw"
    for i in range(n):
        rid = i + 1
        body = _chunk("NAME", names[i].encode())
        meta[rid, 0] = rid | (libraries[i] << 30)
        meta[rid, 1] = (1 << 8) | (len(blob) << 24)          # blob file 1, byte offset
        for kind, s, suffix in (("raw", raw[i], "R"), ("cor", corrected[i], "C")):
            if s is None:
                continue
            s = _normalize(s)
            letter, payload = _encode(s)
            body += _chunk(letter + "SQ" + suffix, payload)
            hs = homopoly_compress(s)
            ign = rid in ignore
            u = dict(valid=True, length=len(s), ignore_u=ign)
            c = dict(valid=True, length=len(hs), ignore_u=ign, clear=(0, len(hs)))
            if kind == latest and rid in clear:
                bgn, end = clear[rid]
                ntoc = homopoly_map(s)
                u.update(clear=(bgn, end), trimmed=True)
                c.update(clear=(ntoc[bgn], ntoc[end]), trimmed=True)
            seqs[kind + "u"][rid] = _seq_words(**u)
            seqs[kind + "c"][rid] = _seq_words(**c)
        blob += _chunk("BLOB", bytes(body))
    with open(os.path.join(path, "blobs.0001"), "wb") as f:
        f.write(blob)
    with open(os.path.join(path, "reads"), "wb") as f:
        f.write(meta.tobytes())
    for k, v in seqs.items():
        with open(os.path.join(path, "reads-" + k), "wb") as f:
            f.write(v.tobytes())
    libs = bytearray(LIBRARY_SIZE * (nlib + 1))
    for lid in range(1, nlib + 1):
        name = f"lib{lid}".encode()
        libs[LIBRARY_SIZE * lid:LIBRARY_SIZE * lid + len(name)] = name
        struct.pack_into("<II", libs, LIBRARY_SIZE * lid + 128, lid, 0)
    with open(os.path.join(path, "libraries"), "wb") as f:
        f.write(bytes(libs))
    reads, bases = _count(seqs, n)
    info = _chunk("MAGC", struct.pack("<Q", MAGIC)) + _chunk("VERS", struct.pack("<Q", VERSION))
    for tag, v in PARAMS:
        info += _chunk(tag, struct.pack("<I", v))
    info += _chunk("NLIB", struct.pack("<I", nlib)) + _chunk("NREA", struct.pack("<I", n))
    info += _chunk("NBLO", struct.pack("<I", 1))
    info += _chunk("READ", struct.pack("<16Q", *reads)) + _chunk("BASE", struct.pack("<16Q", *bases))
    with open(os.path.join(path, "info"), "wb") as f:
        f.write(info)
    if hpc_marker:
        open(os.path.join(path, "homopolymerCompression"), "w").close()


def _count(seqs, n):
    """Reads and bases per version, as the store's info file counts them (stores/sqStoreInfo.C:121-186)."""
    reads, bases = [0] * 16, [0] * 16
    for key, w in (("rawu", RAW), ("rawc", RAW | COMPRESSED), ("coru", CORRECTED), ("corc", CORRECTED | COMPRESSED)):
        for rid in range(1, n + 1):
            s = _Seq(seqs[key][rid])
            length = s.length if s.valid and not s.ignore_u else 0
            trimmed = s.clear_end - s.clear_bgn if s.trimmed and not s.ignore_t else 0
            if length:
                reads[w] += 1
                bases[w] += length
            if trimmed:
                reads[w | TRIMMED] += 1
                bases[w | TRIMMED] += trimmed
    reads[0] = n
    return reads, bases


class _Seq:
    def __init__(self, w):
        w = [int(x) for x in w]
        self.valid, self.length = bool(w[0] & 1), w[0] >> 2
        self.ignore_u, self.ignore_t, self.clear_bgn = bool(w[1] & 1), bool(w[1] & 3), w[1] >> 2
        self.trimmed, self.clear_end = bool(w[2] & 1), w[2] >> 2


def _objects(data):
    out, p = {}, 0
    while p + 8 <= len(data):
        (ln,) = struct.unpack_from("<I", data, p + 4)
        out[data[p:p + 4].decode("latin-1")] = data[p + 8:p + 8 + ln]
        p += 8 + ln
    return out


def read_store(path):
    """(version flags, [(library, sequence)] for reads 1..n) as overlapInCore is handed them.

    Version: the last of raw, raw-trimmed, corrected, corrected-trimmed with reads, compressed when the
    marker exists (stores/sqStoreConstructor.C:141-166).  Sequence: '' when the read's length is 0
    (stores/sqStore.H:407-420), else the decoded bases, compressed, then cut to the clear range
    (stores/sqRead.H:399-460).
    """
    obj = _objects(open(os.path.join(path, "info"), "rb").read())
    if struct.unpack("<Q", obj["MAGC"])[0] != MAGIC or struct.unpack("<Q", obj["VERS"])[0] != VERSION:
        raise ValueError("not a version-9 canu seqStore")
    n = struct.unpack("<I", obj["NREA"])[0]
    counts = struct.unpack("<16Q", obj["READ"])
    which = 0
    for w in (RAW, RAW | TRIMMED, CORRECTED, CORRECTED | TRIMMED):
        if counts[w] > 0:
            which = w
    if os.path.exists(os.path.join(path, "homopolymerCompression")):
        which |= COMPRESSED
    meta = np.fromfile(os.path.join(path, "reads"), np.uint64).reshape(-1, 2)
    key = ("cor" if which & CORRECTED else "raw") + ("c" if which & COMPRESSED else "u")
    full_key = ("cor" if which & CORRECTED else "raw") + "u"
    seq = np.fromfile(os.path.join(path, "reads-" + key), np.uint32).reshape(-1, 3)
    full = np.fromfile(os.path.join(path, "reads-" + full_key), np.uint32).reshape(-1, 3)
    out = []
    blobs = {}
    for rid in range(1, n + 1):
        s = _Seq(seq[rid])
        lib = int(meta[rid, 0] >> np.uint64(30)) & 0xFFF
        if which & TRIMMED:
            ln = 0 if (not s.trimmed or not s.valid or s.ignore_t) else s.clear_end - s.clear_bgn
        else:
            ln = 0 if (not s.valid or s.ignore_u) else s.length
        if ln == 0:
            out.append((lib, ""))
            continue
        seg = int(meta[rid, 1] >> np.uint64(8)) & 0xFFFF
        pos = int(meta[rid, 1] >> np.uint64(24))
        if seg not in blobs:
            blobs[seg] = open(os.path.join(path, f"blobs.{seg:04d}"), "rb").read()
        b = blobs[seg]
        assert b[pos:pos + 4] == b"BLOB"
        (bl,) = struct.unpack_from("<I", b, pos + 4)
        body, p, bases = b[pos + 8:pos + 8 + bl], 0, None
        nfull = _Seq(full[rid]).length
        want = "C" if which & CORRECTED else "R"
        while p + 8 <= len(body):
            tag = body[p:p + 4].decode("latin-1")
            (cl,) = struct.unpack_from("<I", body, p + 4)
            data = body[p + 8:p + 8 + cl]
            if tag[1:3] == "SQ" and tag[3] == want:
                if tag[0] == "2":
                    v = np.frombuffer(data, np.uint8)
                    codes = np.stack([(v >> 6) & 3, (v >> 4) & 3, (v >> 2) & 3, v & 3], 1).reshape(-1)[:nfull]
                    bases = "".join(TWO_BIT[c] for c in codes)
                elif tag[0] == "U":
                    bases = data[:nfull].decode()
                else:
                    raise ValueError(f"read {rid}: unsupported chunk {tag}")
            p += 8 + cl
        if which & COMPRESSED:
            bases = homopoly_compress(bases)
        if which & TRIMMED:
            bases = bases[s.clear_bgn:s.clear_end]
        out.append((lib, bases))
    return which, out
