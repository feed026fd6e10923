"""seqStore input: the command line's C++ reader against canu_amd/seqstore.py, and both against the
hand-over rules worked by hand.

The rules (stores/sqStoreConstructor.C:141-166, stores/sqStore.H:407-420, stores/sqRead.H:399-460):
the latest version with reads is handed over, compressed when the store carries the marker; an
ignored or absent read has length 0; bases are decoded, compressed, then cut to the clear range.
The 2-bit code's base order and the chunk framing come from meryl-utility, which canu does not
ship: parity there is unpinned, and these tests pin only that the two readers and the writer agree.
"""
import os
import struct
import subprocess

import numpy as np
import pytest

from canu_amd import seqstore

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "canu_amd", "bin", "overlapInCore")


def dump(path):
    """[(library, sequence)] from overlapInCore --dump-store."""
    out = subprocess.run([CLI, "--dump-store", str(path)], check=True, capture_output=True, text=True).stdout
    lines = out.split("\n")
    got = []
    for i in range(0, len(lines) - 1, 2):
        rid, lib, ln = map(int, lines[i][1:].split())
        assert rid == len(got) + 1 and ln == len(lines[i + 1])
        got.append((lib, lines[i + 1]))
    return got


def random_reads(seed, n=40):
    r = np.random.default_rng(seed)
    reads = []
    for i in range(n):
        s = "".join(r.choice(list("ACGT"), int(r.integers(1, 300))))
        if i % 7 == 3:   # homopolymer runs
            s = "".join(c * int(r.integers(1, 5)) for c in s)
        reads.append(s)
    reads[5] = reads[5][:10] + "N" + reads[5][11:] if len(reads[5]) > 11 else "ACGTN"
    reads[6] = "acgtu" * 9   # lowercase and U, stored as ACGTT
    return reads


def test_hand_over_rules_by_hand(tmp_path):
    raw = ["AACCCGTTA", "ACGTACGT", "GGGA", "ACNNGT"]
    p = tmp_path / "raw"
    seqstore.write_store(p, raw=raw, ignore={3})
    which, got = seqstore.read_store(p)
    assert which == seqstore.RAW
    assert [s for _, s in got] == ["AACCCGTTA", "ACGTACGT", "", "ACNNGT"]
    assert dump(p) == got

    p = tmp_path / "hpc"
    seqstore.write_store(p, raw=raw, hpc_marker=True)
    which, got = seqstore.read_store(p)
    assert which == seqstore.RAW | seqstore.COMPRESSED
    assert [s for _, s in got] == ["ACGTA", "ACGTACGT", "GA", "ACNGT"]
    assert dump(p) == got

    # clear range [1, 7) of AACCCGTTA; the compressed range is [ntoc[1], ntoc[7]) (stores/sqStore.C:341-342)
    # with ntoc[i] the run of base i: runs AA CCC G TT A, so [0, 3) of ACGTA = ACG
    p = tmp_path / "trim"
    seqstore.write_store(p, raw=raw, clear={1: (1, 7), 2: (2, 6)}, hpc_marker=True)
    which, got = seqstore.read_store(p)
    assert which == seqstore.RAW | seqstore.TRIMMED | seqstore.COMPRESSED
    assert [s for _, s in got] == ["ACG", "GTAC", "", ""]
    assert dump(p) == got

    p = tmp_path / "cor"
    seqstore.write_store(p, raw=raw, corrected=[None, "TTTTGCA", None, "ACGG"], clear={2: (1, 6)})
    which, got = seqstore.read_store(p)
    assert which == seqstore.CORRECTED | seqstore.TRIMMED
    assert [s for _, s in got] == ["", "TTTGC", "", ""]
    assert dump(p) == got


@pytest.mark.parametrize("variant", ["raw", "hpc", "libraries", "ignored", "trimmed_hpc", "corrected"])
def test_cli_reader_matches_python_reader(tmp_path, variant):
    reads = random_reads(hash(variant) % 1000)
    n = len(reads)
    kw = {}
    if variant == "hpc":
        kw["hpc_marker"] = True
    if variant == "libraries":
        kw["libraries"] = [1 + (i % 3) for i in range(n)]
    if variant == "ignored":
        kw["ignore"] = {2, 9, n}
    if variant == "trimmed_hpc":
        kw["hpc_marker"] = True
        kw["clear"] = {i + 1: (len(s) // 4, len(s) - len(s) // 5) for i, s in enumerate(reads) if len(s) > 8}
    if variant == "corrected":
        kw["corrected"] = [s[::-1] if i % 2 else None for i, s in enumerate(reads)]
    p = tmp_path / "s"
    seqstore.write_store(p, raw=reads, **kw)
    which, got = seqstore.read_store(p)
    assert dump(p) == got
    if variant in ("raw", "libraries"):
        assert [s for _, s in got] == [seqstore._normalize(s) for s in reads]
    if variant == "libraries":
        assert [lib for lib, _ in got] == kw["libraries"]
    if variant == "ignored":
        assert all((s == "") == (i + 1 in kw["ignore"]) for i, (_, s) in enumerate(got))


def test_two_bit_packing(tmp_path):
    """Four bases a byte, first base in the two high bits, A=0 C=1 G=2 T=3."""
    p = tmp_path / "s"
    seqstore.write_store(p, raw=["CGTAC"])
    blob = (p / "blobs.0001").read_bytes()
    i = blob.index(b"2SQR")
    (ln,) = struct.unpack_from("<I", blob, i + 4)
    assert ln == 2 and blob[i + 8:i + 10] == bytes([0b01101100, 0b01000000])
    assert dump(p) == [(1, "CGTAC")]


def test_store_metadata_words(tmp_path):
    """Bit fields of sqReadMeta and sqReadSeq, least significant bit first."""
    p = tmp_path / "s"
    seqstore.write_store(p, raw=["ACGT", "AAAC"], libraries=[2, 1], ignore={2})
    meta = np.fromfile(p / "reads", np.uint64).reshape(-1, 2)
    assert int(meta[1, 0]) == 1 | (2 << 30) and int(meta[2, 0]) == 2 | (1 << 30)
    assert int(meta[1, 1]) & 0xFF == 0 and (int(meta[1, 1]) >> 8) & 0xFFFF == 1 and int(meta[1, 1]) >> 24 == 0
    rawc = np.fromfile(p / "reads-rawc", np.uint32).reshape(-1, 3)
    assert list(rawc[2]) == [1 | (2 << 2), 3, 2 << 2]   # valid, length 2; ignoreU+ignoreT; clearEnd 2
    info = seqstore._objects((p / "info").read_bytes())
    reads = struct.unpack("<16Q", info["READ"])
    assert reads[0] == 2 and reads[seqstore.RAW] == 1 and reads[seqstore.RAW | seqstore.COMPRESSED] == 1


def patch(path, old, new):
    data = path.read_bytes()
    assert old in data
    path.write_bytes(data.replace(old, new, 1))


@pytest.mark.parametrize("case", ["three_bit", "version8", "read_size", "missing_blob"])
def test_refused_stores(tmp_path, case):
    p = tmp_path / "s"
    seqstore.write_store(p, raw=["ACGTACGT", "TTGCA"])
    if case == "three_bit":
        patch(p / "blobs.0001", b"2SQR", b"3SQR")
        want = "3-bit"
    elif case == "version8":
        patch(p / "info", b"MAGC", b"XXXX")
        (p / "info").write_bytes(struct.pack("<Q", 0x504B473A756E6162) + (p / "info").read_bytes()[8:])
        want = "version 8"
    elif case == "read_size":
        patch(p / "info", b"RSIZ" + struct.pack("<II", 4, 152), b"RSIZ" + struct.pack("<II", 4, 160))
        want = "RSIZ"
    else:
        os.unlink(p / "blobs.0001")
        want = "blobs.0001"
    r = subprocess.run([CLI, "--dump-store", str(p)], capture_output=True, text=True)
    assert r.returncode == 1 and want in r.stderr, r.stderr
