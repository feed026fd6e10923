"""Synthetic genomes and reads for the overlapInCore workloads (ctypes over libreadsim.so).

Workload presets (seeds fixed):

  c1  2 Mbp uniform genome, 15x, lognormal 8k/3k (min 1 kbp), 0.5/0.5/0.5 % sub/ins/del
  c2  50 Mbp uniform genome, 30x, lognormal 12k/5k (max 60 kbp), 1.5/1.5/1.0 %
  c3  c2 plus 25 % repeats: 300 families (3-8 kbp, copies 1-5 % diverged) and
      20 tandem arrays of a 171 bp unit (50-200 kbp, 0.5 % unit divergence)
  c5  100 Mbp with the c3 repeat mix, 30x, lognormal 18k/4k, 0.2 % errors
      (mostly homopolymer-length indels), homopolymer-compressed reads

Reads are 1-based: read i is ``bases[start[i]:start[i]+length[i]]``, NUL-terminated.
"""
import ctypes
import dataclasses
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "libreadsim.so")


class _Config(ctypes.Structure):
    _fields_ = [
        ("genome_len", ctypes.c_uint64), ("seed", ctypes.c_uint64), ("coverage", ctypes.c_double),
        ("ln_mu", ctypes.c_double), ("ln_sigma", ctypes.c_double),
        ("min_len", ctypes.c_uint32), ("max_len", ctypes.c_uint32),
        ("p_sub", ctypes.c_double), ("p_ins", ctypes.c_double), ("p_del", ctypes.c_double),
        ("hp_errors", ctypes.c_int32),
        ("repeat_frac", ctypes.c_double),
        ("n_families", ctypes.c_uint32), ("fam_min", ctypes.c_uint32), ("fam_max", ctypes.c_uint32),
        ("div_min", ctypes.c_double), ("div_max", ctypes.c_double),
        ("n_sat", ctypes.c_uint32), ("sat_unit", ctypes.c_uint32), ("sat_min", ctypes.c_uint32),
        ("sat_max", ctypes.c_uint32), ("sat_div", ctypes.c_double),
        ("hpc", ctypes.c_int32),
    ]


class _Reads(ctypes.Structure):
    _fields_ = [
        ("bases", ctypes.POINTER(ctypes.c_char)), ("start", ctypes.POINTER(ctypes.c_uint64)),
        ("len", ctypes.POINTER(ctypes.c_uint32)), ("pos", ctypes.POINTER(ctypes.c_uint64)),
        ("tlen", ctypes.POINTER(ctypes.c_uint32)), ("strand", ctypes.POINTER(ctypes.c_uint8)),
        ("nreads", ctypes.c_uint32), ("total", ctypes.c_uint64),
        ("genome", ctypes.POINTER(ctypes.c_char)), ("genome_len", ctypes.c_uint64),
    ]


@dataclasses.dataclass
class SimConfig:
    genome_len: int
    seed: int
    coverage: float
    ln_mu: float
    ln_sigma: float
    min_len: int = 1000
    max_len: int = 0
    p_sub: float = 0.005
    p_ins: float = 0.005
    p_del: float = 0.005
    hp_errors: int = 0
    repeat_frac: float = 0.0
    n_families: int = 0
    fam_min: int = 3000
    fam_max: int = 8000
    div_min: float = 0.01
    div_max: float = 0.05
    n_sat: int = 0
    sat_unit: int = 171
    sat_min: int = 50000
    sat_max: int = 200000
    sat_div: float = 0.005
    hpc: int = 0

    def scaled(self, genome_len, seed=None):
        """Same read/error model on a smaller genome (repeat sizes scaled with it)."""
        f = genome_len / self.genome_len
        kw = dataclasses.asdict(self)
        kw["genome_len"] = int(genome_len)
        if seed is not None:
            kw["seed"] = seed
        if self.n_families:
            kw["n_families"] = max(1, int(round(self.n_families * f)))
        if self.n_sat:
            kw["n_sat"] = max(1, int(round(self.n_sat * f)))
            kw["sat_min"] = max(self.sat_unit * 4, int(self.sat_min * f))
            kw["sat_max"] = max(kw["sat_min"], int(self.sat_max * f))
        return SimConfig(**kw)


PRESETS = {
    "c1": SimConfig(2_000_000, 1, 15.0, 8.92141, 0.36273, 1000, 0, 0.005, 0.005, 0.005),
    "c2": SimConfig(50_000_000, 2, 30.0, 9.31262, 0.40011, 1000, 60000, 0.015, 0.015, 0.010),
    "c3": SimConfig(50_000_000, 3, 30.0, 9.31262, 0.40011, 1000, 60000, 0.015, 0.015, 0.010,
                    repeat_frac=0.25, n_families=300, n_sat=20),
    "c5": SimConfig(100_000_000, 5, 30.0, 9.77403, 0.21955, 1000, 0, 0.0004, 0.0008, 0.0008,
                    hp_errors=1, repeat_frac=0.25, n_families=300, n_sat=20, hpc=1),
}


@dataclasses.dataclass
class Reads:
    """1-based read set: read i is bases[start[i]:start[i]+length[i]] (uppercase ACGT[N])."""
    bases: np.ndarray      # uint8, NUL after every read
    start: np.ndarray      # uint64 [n+1]
    length: np.ndarray     # uint32 [n+1]
    pos: np.ndarray        # uint64 [n+1] genome start of the template
    strand: np.ndarray     # uint8  [n+1]
    genome_len: int
    tlen: np.ndarray = None  # uint32 [n+1] template length on the genome (the read covers [pos, pos + tlen))

    @property
    def n(self):
        return len(self.length) - 1

    def read(self, i):
        s = int(self.start[i])
        return self.bases[s:s + int(self.length[i])].tobytes()

    def subset(self, n):
        """The first n reads as their own 1-based set."""
        end = int(self.start[n] + self.length[n] + 1)
        return Reads(self.bases[:end].copy(), self.start[:n + 1].copy(), self.length[:n + 1].copy(),
                     self.pos[:n + 1].copy(), self.strand[:n + 1].copy(), self.genome_len,
                     None if self.tlen is None else self.tlen[:n + 1].copy())


_lib = None


def _load():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            raise RuntimeError(f"{_LIB} is missing: run __graft_entry__.build()")
        _lib = ctypes.CDLL(_LIB)
        _lib.rs_generate.argtypes = [ctypes.POINTER(_Config), ctypes.POINTER(_Reads)]
        _lib.rs_free.argtypes = [ctypes.POINTER(_Reads)]
    return _lib


def simulate(cfg: SimConfig) -> Reads:
    lib = _load()
    c = _Config(**dataclasses.asdict(cfg))
    r = _Reads()
    lib.rs_generate(ctypes.byref(c), ctypes.byref(r))
    try:
        n = r.nreads
        nb = int(r.start[n]) + int(r.len[n]) + 1 if n else 0
        bases = np.ctypeslib.as_array(ctypes.cast(r.bases, ctypes.POINTER(ctypes.c_uint8)), (max(nb, 1),))[:nb].copy()
        start = np.ctypeslib.as_array(r.start, (n + 1,)).copy()
        length = np.ctypeslib.as_array(r.len, (n + 1,)).copy()
        pos = np.ctypeslib.as_array(r.pos, (n + 1,)).copy()
        strand = np.ctypeslib.as_array(r.strand, (n + 1,)).copy()
        tlen = np.ctypeslib.as_array(r.tlen, (n + 1,)).copy()
        glen = int(r.genome_len)
    finally:
        lib.rs_free(ctypes.byref(r))
    return Reads(bases, start, length, pos, strand, glen, tlen)


def from_strings(seqs):
    """Build a 1-based Reads set from a list of strings/bytes."""
    seqs = [s.encode() if isinstance(s, str) else bytes(s) for s in seqs]
    n = len(seqs)
    start = np.zeros(n + 1, np.uint64)
    length = np.zeros(n + 1, np.uint32)
    chunks, off = [], 0
    for i, s in enumerate(seqs, 1):
        start[i] = off
        length[i] = len(s)
        chunks.append(s + b"\0")
        off += len(s) + 1
    bases = np.frombuffer(b"".join(chunks), np.uint8).copy() if chunks else np.zeros(0, np.uint8)
    return Reads(bases, start, length, np.zeros(n + 1, np.uint64), np.zeros(n + 1, np.uint8), 0)
