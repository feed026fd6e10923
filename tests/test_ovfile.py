"""The .ovb codec and the overlapInCore command line.

CPU: the engine's snappy block encoder (overlapInCore --snappy-block) against Google's snappy as
pyarrow ships it, both ways, and the pure-Python decoder of canu_amd/ovfile.py.
GPU: the command line end to end -- FASTA in; .ovb, .oc and .stats out -- against the oracle.
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
