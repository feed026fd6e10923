"""The .ovb codec and the overlapInCore command line.

CPU: the engine's snappy block encoder (overlapInCore --snappy-block) against Google's snappy as
pyarrow ships it, both ways, and the pure-Python decoder of canu_amd/ovfile.py.
GPU: the command line end to end -- FASTA or seqStore in; .ovb, .oc and .stats out -- against the
oracle, and store input against FASTA input.
"""
import os
import subprocess

import numpy as np
import pytest

from canu_amd import ovfile, readsim

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "canu_amd", "bin", "overlapInCore")


def samples():
    r = np.random.default_rng(5)
    recs = np.zeros(5000, ovfile.RECORD_DTYPE)
    recs["a_iid"] = np.sort(r.integers(1, 4000, 5000))
    recs["b_iid"] = recs["a_iid"] + r.integers(1, 50, 5000)
    recs["w0"] = r.integers(0, 1 << 62, 5000, dtype=np.uint64)
    recs["w1"] = r.integers(0, 1 << 62, 5000, dtype=np.uint64) & np.uint64(0xFFFFFFFFFF)
    return {"empty": b"", "short": b"acgta", "random": r.bytes(70_000), "runs": bytes(200_000),
            "periodic": bytes(range(7)) * 30_000, "records": recs.tobytes()}


@pytest.mark.parametrize("name", list(samples()))
def test_snappy_block_decodes_with_google_snappy(tmp_path, name):
    pa = pytest.importorskip("pyarrow")
    data = samples()[name]
    src, dst = tmp_path / "in", tmp_path / "out"
    src.write_bytes(data)
    subprocess.run([CLI, "--snappy-block", str(src), str(dst)], check=True)
    block = dst.read_bytes()
    assert ovfile.snappy_uncompress(block) == data
    if data:
        got = pa.Codec("snappy").decompress(block, decompressed_size=len(data)).to_pybytes()
        assert got == data
    if name in ("runs", "periodic"):
        assert len(block) < len(data) // 4


@pytest.mark.parametrize("name", list(samples()))
def test_python_decoder_reads_google_snappy(name):
    pa = pytest.importorskip("pyarrow")
    data = samples()[name]
    if not data:
        return
    block = pa.Codec("snappy").compress(data).to_pybytes()
    assert ovfile.snappy_uncompress(block) == data


def write_fasta(path, R, hpc=False):
    with open(path, "w") as f:
        for i in range(1, R.n + 1):
            s = R.read(i).decode()
            f.write(f">r{i}\n")
            for k in range(0, len(s), 80):
                f.write(s[k:k + 80] + "\n")


def hp_compress(s):
    out = [s[0]] if s else []
    for ch in s[1:]:
        if ch.lower() != out[-1].lower():
            out.append(ch)
    return "".join(out)


def run_cli(tmp_path, fasta, *args):
    out = tmp_path / "1.ovb.WORKING"
    stats = tmp_path / "1.stats"
    subprocess.run([CLI, "-k", "22", "-t", "4", "-o", str(out), "-s", str(stats), *args, str(fasta)], check=True,
                   timeout=600)
    return out, tmp_path / "1.oc", stats


def check_outputs(ovb, oc, stats, R, ref):
    from oracle import oracle
    recs = ovfile.read_ovb(ovb)
    a, b = oracle.sort_records(recs), oracle.sort_records(ref["records"])
    assert len(a) == len(b)
    for f in ("a_iid", "b_iid", "w0", "w1"):
        assert np.array_equal(a[f], b[f]), f
    n, opr = ovfile.read_oc(oc)
    assert n == len(recs) and len(opr) == R.n + 1
    want = np.bincount(np.concatenate([recs["a_iid"], recs["b_iid"]]).astype(np.int64), minlength=R.n + 1)
    assert np.array_equal(opr, want)
    assert ovfile.read_stats(stats) == list(ref["stats"])


@pytest.mark.gpu
def test_cli_matches_oracle(tmp_path):
    from oracle import oracle
    R = readsim.simulate(readsim.PRESETS["c1"].scaled(200_000, seed=51))
    fasta = tmp_path / "reads.fasta"
    write_fasta(fasta, R)
    ovb, oc, stats = run_cli(tmp_path, fasta, "--maxerate", "0.045", "--minlength", "500")
    ref = oracle.run(R, kmer=22, maxErate=0.045, minOlap=500, nThreads=4, ompThreads=16)
    assert len(ref["records"]) > 1000
    check_outputs(ovb, oc, stats, R, ref)


@pytest.mark.gpu
def test_cli_hpc_ranges_and_options(tmp_path):
    """--hpc compresses homopolymers on load; -h/-r ranges, -m, --minkmers pass through."""
    from oracle import oracle
    cfg = readsim.PRESETS["c5"].scaled(250_000, seed=52)
    cfg.hpc = 0
    R = readsim.simulate(cfg)
    fasta = tmp_path / "reads.fasta"
    write_fasta(fasta, R)
    n = R.n
    ovb, oc, stats = run_cli(tmp_path, fasta, "--hpc", "--maxerate", "0.01", "--minlength", "1000", "-m", "--minkmers",
                             "-h", f"2-{n - 1}", "-r", f"3-{n}")
    Rh = readsim.from_strings([hp_compress(R.read(i).decode()) for i in range(1, n + 1)])
    ref = oracle.run(Rh, kmer=22, maxErate=0.01, minOlap=1000, unique=0, minkmers=1, nThreads=4, ompThreads=16,
                     bgnHash=2, endHash=n - 1, bgnRef=3, endRef=n)
    assert len(ref["records"]) > 100
    check_outputs(ovb, oc, stats, Rh, ref)


def write_fasta_strings(path, reads):
    with open(path, "w") as f:
        for i, s in enumerate(reads):
            f.write(f">r{i + 1}\n{s}\n")


@pytest.mark.gpu
def test_cli_reads_seqstore(tmp_path):
    """A seqStore and a FASTA of the same reads give identical .ovb records, .oc and .stats; ignored
    reads match empty FASTA records, and the store's homopolymerCompression marker matches --hpc."""
    from canu_amd import seqstore
    from oracle import oracle
    R = readsim.simulate(readsim.PRESETS["c1"].scaled(150_000, seed=53))
    reads = [R.read(i).decode() for i in range(1, R.n + 1)]
    ignore = {5, 17, R.n}
    fasta = tmp_path / "reads.fasta"
    write_fasta_strings(fasta, ["" if i + 1 in ignore else s for i, s in enumerate(reads)])
    for marker in (False, True):
        outs = []
        for kind in ("store", "fasta"):
            d = tmp_path / f"{kind}{int(marker)}"
            d.mkdir()
            src = fasta
            extra = ["--hpc"] if (marker and kind == "fasta") else []
            if kind == "store":
                src = d / "asm.seqStore"
                seqstore.write_store(src, raw=reads, hpc_marker=marker, ignore=ignore)
            print(f"cli {kind} marker={marker}", flush=True)
            outs.append(run_cli(d, src, "--maxerate", "0.045", "--minlength", "500", *extra))
        (ovb_s, oc_s, st_s), (ovb_f, oc_f, st_f) = outs
        a, b = oracle.sort_records(ovfile.read_ovb(ovb_s)), oracle.sort_records(ovfile.read_ovb(ovb_f))
        assert len(a) > 300 and len(a) == len(b)
        for f in ("a_iid", "b_iid", "w0", "w1"):
            assert np.array_equal(a[f], b[f]), f
        assert not np.isin(a["a_iid"], list(ignore)).any() and not np.isin(a["b_iid"], list(ignore)).any()
        assert oc_s.read_bytes() == oc_f.read_bytes()
        assert st_s.read_bytes() == st_f.read_bytes()
