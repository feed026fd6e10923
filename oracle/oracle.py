"""ctypes binding of the CPU oracle (liboic_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg, and only as the reference results are compared with.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "liboic_oracle.so")

MAX_READLEN = (1 << 21) - 1


class Params(ctypes.Structure):
    _fields_ = [
        ("kmer", ctypes.c_uint32), ("maxErate", ctypes.c_double), ("alignNoise", ctypes.c_double),
        ("minOlap", ctypes.c_int32), ("partial", ctypes.c_int32), ("unique", ctypes.c_int32),
        ("hopeless", ctypes.c_int32), ("minkmers", ctypes.c_int32), ("fragOlapLimit", ctypes.c_uint64),
        ("hashBits", ctypes.c_uint32), ("hashLoad", ctypes.c_double), ("hashDataLen", ctypes.c_uint64),
        ("bgnHash", ctypes.c_uint32), ("endHash", ctypes.c_uint32),
        ("bgnRef", ctypes.c_uint32), ("endRef", ctypes.c_uint32),
        ("nThreads", ctypes.c_uint32), ("ompThreads", ctypes.c_uint32),
        ("skipKmers", ctypes.c_char_p), ("nSkip", ctypes.c_uint64), ("restrictIndex", ctypes.c_int32),
        ("refList", ctypes.c_void_p), ("nRefList", ctypes.c_uint64),
    ]


class Record(ctypes.Structure):
    _fields_ = [("a_iid", ctypes.c_uint32), ("b_iid", ctypes.c_uint32),
                ("w0", ctypes.c_uint64), ("w1", ctypes.c_uint64)]


RECORD_DTYPE = np.dtype([("a_iid", "<u4"), ("b_iid", "<u4"), ("w0", "<u8"), ("w1", "<u8")])

TRACE_DTYPE = np.dtype([
    ("s_id", "<u4"), ("t_id", "<u4"), ("dir", "<i4"), ("s_len", "<i4"), ("t_len", "<i4"),
    ("n_hits", "<i4"), ("n_nodes", "<i4"), ("diag_min", "<i4"), ("diag_max", "<i4"),
    ("consistent", "<i4"), ("longest", "<i4"), ("n_ext", "<i4"), ("n_olaps", "<i4"),
    ("_pad", "<i4"), ("cells", "<u8"), ("max_mult", "<i4"), ("exit_d", "<i4")])


class Tables(ctypes.Structure):
    _fields_ = [
        ("maxErate", ctypes.c_double), ("alignNoise", ctypes.c_double), ("partial", ctypes.c_int),
        ("MAX_ERRORS", ctypes.c_uint32), ("MIN_BRANCH_END_DIST", ctypes.c_uint32),
        ("MIN_BRANCH_TAIL_SLOPE", ctypes.c_double), ("Branch_Match_Value", ctypes.c_double),
        ("Error_Bound", ctypes.POINTER(ctypes.c_int32)), ("Edit_Match_Limit", ctypes.POINTER(ctypes.c_int32)),
    ]


class Result(ctypes.Structure):
    _fields_ = [
        ("rec", ctypes.POINTER(Record)), ("nrec", ctypes.c_uint64), ("stats", ctypes.c_uint64 * 8),
        ("cells", ctypes.c_uint64), ("bases", ctypes.c_uint64), ("dp_calls", ctypes.c_uint64),
        ("extensions", ctypes.c_uint64), ("hits", ctypes.c_uint64), ("pairs", ctypes.c_uint64),
        ("cells_used", ctypes.c_uint64),
        ("trace", ctypes.c_void_p), ("ntrace", ctypes.c_uint64),
        ("hashEnd", ctypes.c_uint32 * 64), ("nblocks", ctypes.c_uint32),
        ("t_index", ctypes.c_double), ("t_search", ctypes.c_double),
    ]


class Match(ctypes.Structure):
    _fields_ = [("Offset", ctypes.c_int32), ("Len", ctypes.c_int32), ("Start", ctypes.c_int32), ("Next", ctypes.c_int32)]


_lib = None
_P32 = ctypes.POINTER(ctypes.c_int32)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            raise RuntimeError(f"{_LIB} is missing: run `make -C oracle` or __graft_entry__.build()")
        L = ctypes.CDLL(_LIB)
        L.oo_tables_new.restype = ctypes.POINTER(Tables)
        L.oo_tables_new.argtypes = [ctypes.c_double, ctypes.c_double, ctypes.c_int]
        L.oo_tables_free.argtypes = [ctypes.POINTER(Tables)]
        L.oo_cli_rate.restype = ctypes.c_double
        L.oo_cli_rate.argtypes = [ctypes.c_double]
        L.oo_binomial_bound.restype = ctypes.c_int
        L.oo_binomial_bound.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_int]
        L.oo_dp_new.restype = ctypes.c_void_p
        L.oo_dp_new.argtypes = [ctypes.POINTER(Tables)]
        L.oo_dp_free.argtypes = [ctypes.c_void_p]
        L.oo_dp_set_poison.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.oo_dp_poison_hits.restype = ctypes.c_int
        L.oo_dp_poison_hits.argtypes = [ctypes.c_void_p]
        L.oo_forward.restype = ctypes.c_int32
        L.oo_forward.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int32, ctypes.c_void_p, ctypes.c_int32,
                                 ctypes.c_int32, _P32, _P32, _P32]
        L.oo_reverse.restype = ctypes.c_int32
        L.oo_reverse.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int32, ctypes.c_void_p, ctypes.c_int32,
                                 ctypes.c_int32, _P32, _P32, _P32, _P32]
        L.oo_extend.restype = ctypes.c_int32
        L.oo_extend.argtypes = [ctypes.c_void_p, ctypes.POINTER(Match), ctypes.c_void_p, ctypes.c_uint32, ctypes.c_int32,
                                ctypes.c_void_p, ctypes.c_uint32, ctypes.c_int32, _P32, _P32, _P32, _P32, _P32]
        L.oo_dp_left_delta.restype = ctypes.c_int32
        L.oo_dp_left_delta.argtypes = [ctypes.c_void_p, _P32, ctypes.c_int32]
        L.oo_dp_right_delta.restype = ctypes.c_int32
        L.oo_dp_right_delta.argtypes = [ctypes.c_void_p, _P32, ctypes.c_int32]
        L.oo_dp_counters.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_uint64)]
        L.oo_dp_reset_counters.argtypes = [ctypes.c_void_p]
        L.oo_params_default.argtypes = [ctypes.POINTER(Params)]
        L.oo_run.restype = ctypes.c_int
        L.oo_run.argtypes = [ctypes.POINTER(Params), ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                             ctypes.c_uint32, ctypes.c_int, ctypes.POINTER(Result)]
        L.oo_result_free.argtypes = [ctypes.POINTER(Result)]
        L.oo_encode_evalue.restype = ctypes.c_uint32
        L.oo_encode_evalue.argtypes = [ctypes.c_double]
        L.oo_match_list.restype = ctypes.c_int32
        L.oo_match_list.argtypes = [ctypes.c_uint32, _P32, _P32, ctypes.c_int32, _P32, ctypes.c_int32, _P32]
        L.oo_lies_on_alignment.restype = ctypes.c_int32
        L.oo_lies_on_alignment.argtypes = [ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, _P32, ctypes.c_int32]
        L.oo_distinct_olaps.restype = ctypes.c_int32
        L.oo_distinct_olaps.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int32, _P32,
                                        ctypes.POINTER(ctypes.c_double), _P32, ctypes.POINTER(ctypes.c_double)]
        _lib = L
    return _lib


def match_list(kmer, offsets, toffs):
    """Add_Match over one pair's seed hits in arrival order: the final match list, front first, as
    (Offset, Len, Start) rows, and the consistent flag."""
    o = np.ascontiguousarray(offsets, np.int32)
    t = np.ascontiguousarray(toffs, np.int32)
    out = np.zeros((max(1, len(o)), 3), np.int32)
    cons = ctypes.c_int32(0)
    k = lib().oo_match_list(kmer, o.ctypes.data_as(_P32), t.ctypes.data_as(_P32), len(o), out.ctypes.data_as(_P32),
                            len(out), ctypes.byref(cons))
    assert k >= 0
    return out[:k].copy(), bool(cons.value)


MAX_DISTINCT_OLAPS = 3


def distinct_olaps(boxes, quals, partial=False, unique=False):
    """Add_Overlap over extension results (rows s_lo, s_hi, t_lo, t_hi with their qualities), then the
    per-pair selection: rows (s_lo, s_hi, t_lo, t_hi, min_diag, max_diag, s_left, s_right, t_left,
    t_right, deleted) and the qualities of the distinct overlaps."""
    bx = np.ascontiguousarray(boxes, np.int32).reshape(-1, 4)
    q = np.ascontiguousarray(quals, np.float64)
    assert len(q) == len(bx)
    out = np.zeros((MAX_DISTINCT_OLAPS, 11), np.int32)
    oq = np.zeros(MAX_DISTINCT_OLAPS, np.float64)
    pd = ctypes.POINTER(ctypes.c_double)
    k = lib().oo_distinct_olaps(int(partial), int(unique), len(bx), bx.ctypes.data_as(_P32), q.ctypes.data_as(pd),
                                out.ctypes.data_as(_P32), oq.ctypes.data_as(pd))
    return out[:k].copy(), oq[:k].copy()


def lies_on_alignment(start, offset, s_lo, t_lo, delta):
    """Lies_On_Alignment: whether the match node at (start, offset) lies on the alignment from
    (s_lo, t_lo) whose Left_Delta is delta."""
    d = np.ascontiguousarray(delta if len(delta) else [0], np.int32)
    return bool(lib().oo_lies_on_alignment(start, offset, s_lo, t_lo, d.ctypes.data_as(_P32), len(delta)))


def cli_rate(x):
    return lib().oo_cli_rate(x)


class OracleTables:
    """Error_Bound / Edit_Match_Limit exactly as prefixEditDistance's constructor builds them."""

    def __init__(self, maxErate, alignNoise=1.0, partial=False, cli=True):
        L = lib()
        me = cli_rate(maxErate) if cli else maxErate
        an = cli_rate(alignNoise) if cli else alignNoise
        self._t = L.oo_tables_new(me, an, int(partial))
        t = self._t.contents
        self.maxErate = t.maxErate
        self.MAX_ERRORS = t.MAX_ERRORS
        self.Branch_Match_Value = t.Branch_Match_Value
        self.MIN_BRANCH_TAIL_SLOPE = t.MIN_BRANCH_TAIL_SLOPE
        self.error_bound = np.ctypeslib.as_array(t.Error_Bound, (MAX_READLEN + 1,)).copy()
        self.match_limit = np.ctypeslib.as_array(t.Edit_Match_Limit, (t.MAX_ERRORS,)).copy()

    def __del__(self):
        try:
            lib().oo_tables_free(self._t)
        except Exception:
            pass


class OracleDP:
    """One prefixEditDistance object: forward / reverse / Extend_Alignment on byte strings."""

    def __init__(self, maxErate, alignNoise=1.0, partial=False, match_limit=None, cli=True):
        self.tables = OracleTables(maxErate, alignNoise, partial, cli)
        if match_limit is not None:   # test hook: override Edit_Match_Limit (e.g. all zeros)
            ml = np.asarray(match_limit, np.int32)
            ptr = self.tables._t.contents.Edit_Match_Limit
            for i in range(min(len(ml), self.tables.MAX_ERRORS)):
                ptr[i] = int(ml[i])
        self._w = lib().oo_dp_new(self.tables._t)
        self._keep = []

    def __del__(self):
        try:
            lib().oo_dp_free(self._w)
        except Exception:
            pass

    def poison(self, on=True):
        lib().oo_dp_set_poison(self._w, int(on))

    def poison_hits(self):
        return lib().oo_dp_poison_hits(self._w)

    @staticmethod
    def _buf(s):
        b = ctypes.create_string_buffer(bytes(s), len(s) + 1)
        return b

    def forward(self, A, T, limit):
        """forward(A[0..m), T[0..n)) -> (e, A_End, T_End, match_to_end, Right_Delta)"""
        a, t = self._buf(A), self._buf(T)
        ae, te, mte = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int32()
        e = lib().oo_forward(self._w, a, len(A), t, len(T), limit, ctypes.byref(ae), ctypes.byref(te), ctypes.byref(mte))
        return e, ae.value, te.value, bool(mte.value), self.right_delta()

    def reverse(self, A, T, limit):
        """reverse on A[m-1 .. 0], T[n-1 .. 0] (pointers at the last base) -> (e, A_End, T_End, leftover, mte, Left_Delta)"""
        a, t = self._buf(A), self._buf(T)
        ae, te, lo, mte = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int32()
        pa = ctypes.cast(a, ctypes.c_void_p).value + len(A) - 1
        pt = ctypes.cast(t, ctypes.c_void_p).value + len(T) - 1
        e = lib().oo_reverse(self._w, pa, len(A), pt, len(T), limit, ctypes.byref(ae), ctypes.byref(te),
                             ctypes.byref(lo), ctypes.byref(mte))
        return e, ae.value, te.value, lo.value, bool(mte.value), self.left_delta()

    def extend(self, S, T, start, offset, length, s_id=1, t_id=2):
        """Extend_Alignment from seed (S pos start, T pos offset, len) -> (kind, S_Lo, S_Hi, T_Lo, T_Hi, Errors, delta)"""
        s, t = self._buf(S), self._buf(T)
        m = Match(offset, length, start, 0)
        v = [ctypes.c_int32() for _ in range(5)]
        kind = lib().oo_extend(self._w, ctypes.byref(m), s, s_id, len(S), t, t_id, len(T),
                               *[ctypes.byref(x) for x in v])
        return (kind,) + tuple(x.value for x in v) + (self.left_delta(),)

    def left_delta(self):
        n = lib().oo_dp_left_delta(self._w, None, 0)
        out = np.zeros(max(n, 1), np.int32)
        lib().oo_dp_left_delta(self._w, out.ctypes.data_as(_P32), n)
        return out[:n].copy()

    def right_delta(self):
        n = lib().oo_dp_right_delta(self._w, None, 0)
        out = np.zeros(max(n, 1), np.int32)
        lib().oo_dp_right_delta(self._w, out.ctypes.data_as(_P32), n)
        return out[:n].copy()

    def counters(self):
        c = (ctypes.c_uint64 * 4)()
        lib().oo_dp_counters(self._w, c)
        return dict(cells=c[0], bases=c[1], calls=c[2], rows=c[3])

    def reset_counters(self):
        lib().oo_dp_reset_counters(self._w)


def default_params(**kw):
    p = Params()
    lib().oo_params_default(ctypes.byref(p))
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def parse_skip_file(text, kmer):
    """The k-mers of a -k skip file, as Mark_Skip_Kmers reads them (overlapInCore-Build_Hash_Index.C:185-257):
    each line, or the line after a '>' header, cut at its first blank; every k-mer must be kmer long."""
    lines = text.split("\n")
    out, i = [], 0
    while i < len(lines):
        line = lines[i]
        if i == len(lines) - 1 and line == "":
            break
        if line.startswith(">"):
            i += 1
            line = lines[i] if i < len(lines) else ""
        cut = len(line)
        for j, ch in enumerate(line):
            if ch in " \t\r":
                cut = j
                break
        km = line[:cut]
        if len(km) != kmer:
            raise ValueError(f"skip k-mer '{km}' at line {i + 1} has length {len(km)}, expecting {kmer}")
        out.append(km.lower())
        i += 1
    return out


def run(reads, collect_trace=False, **kw):
    """Run the whole overlapInCore restatement on a 1-based read set (canu_amd.readsim.Reads).

    Returns dict(records=structured array, stats=[8], counters...)."""
    skip = kw.pop("skip", None)
    refs = kw.pop("refs", None)          # explicit ref IDs (searched exactly, no -t chunk rule)
    p = default_params(**kw)
    if refs is not None:
        ref_arr = np.ascontiguousarray(refs, np.uint32)
        if len(ref_arr) == 0 or ref_arr.min() < 1 or ref_arr.max() > reads.n:
            raise ValueError("ref IDs out of range")
        p.refList = ref_arr.ctypes.data
        p.nRefList = len(ref_arr)
    if skip:
        buf = b"".join(x.encode() if isinstance(x, str) else bytes(x) for x in skip)
        p.skipKmers = buf
        p.nSkip = len(skip)
    res = Result()
    bases = np.ascontiguousarray(reads.bases)
    start = np.ascontiguousarray(reads.start, np.uint64)
    length = np.ascontiguousarray(reads.length, np.uint32)
    rc = lib().oo_run(ctypes.byref(p), bases.ctypes.data, start.ctypes.data, length.ctypes.data,
                      int(reads.n), int(collect_trace), ctypes.byref(res))
    if rc != 0:
        raise RuntimeError(f"oracle oo_run failed: {rc}")
    try:
        n = res.nrec
        if n:
            buf = ctypes.string_at(ctypes.cast(res.rec, ctypes.c_void_p).value, n * RECORD_DTYPE.itemsize)
            recs = np.frombuffer(buf, RECORD_DTYPE).copy()
        else:
            recs = np.zeros(0, RECORD_DTYPE)
        trace = None
        if collect_trace and res.ntrace:
            buf = ctypes.string_at(res.trace, res.ntrace * TRACE_DTYPE.itemsize)
            trace = np.frombuffer(buf, TRACE_DTYPE).copy()
        out = dict(records=recs, stats=list(res.stats), cells=res.cells, cells_used=res.cells_used, bases=res.bases,
                   dp_calls=res.dp_calls,
                   extensions=res.extensions, hits=res.hits, pairs=res.pairs, trace=trace,
                   hash_end=list(res.hashEnd[:min(res.nblocks, 64)]), t_index=res.t_index, t_search=res.t_search)
    finally:
        lib().oo_result_free(ctypes.byref(res))
    return out


def sort_records(recs):
    """Canonical order of ovOverlap::operator< (ovOverlap.H:263-277)."""
    if len(recs) == 0:
        return recs
    idx = np.lexsort((recs["w1"], recs["w0"], recs["b_iid"], recs["a_iid"]))
    return recs[idx]
