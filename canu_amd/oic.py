"""Python host mirror of the overlapInCore C ABI (include/oic.h, liboic_hip.so).

The engine runs on the GPU only: if the HIP library is missing or no device is visible,
constructing an Engine raises.  Names follow canu's overlapInCore surface:

    Engine(kmer=22, maxerate=0.045, minlength=500, ...)   # oicParameters / argv (overlapInCore.C:293-411)
    eng.load_reads(reads)                                  # sqStore / sqCache reads
    eng.build_hash_index(bgn, end) -> last loaded ID       # Build_Hash_Index
    eng.search(bgn_ref, end_ref) -> Result                 # Process_Overlaps over a ref range
    eng.run(...) -> Result                                 # OverlapDriver: hash-block loop
    eng.extend_alignment(tasks)                            # prefixEditDistance::Extend_Alignment
    eng.forward(tasks) / eng.reverse(tasks)                # prefixEditDistance::forward / ::reverse
"""
import ctypes
import dataclasses
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liboic_hip.so")

OIC_OK, OIC_EINVAL, OIC_EHIP, OIC_ENOMEM, OIC_ESTATE, OIC_ECAPACITY, OIC_EUNSUPPORTED = 0, -1, -2, -3, -4, -5, -6
SCHED_PLAIN, SCHED_BATCHED, SCHED_LEARNED = 0, 1, 2

EXPORTS = [
    "oic_params_default", "oic_create", "oic_destroy", "oic_last_error", "oic_load_reads", "oic_index_build",
    "oic_search", "oic_run", "oic_fetch_pending", "oic_get_counters", "oic_extend_batch", "oic_prefix_dp_batch",
    "oic_prefix_dp_batch_noalloc", "oic_tables",
    "oic_dp_tables", "oic_set_tracing", "oic_get_traces", "oic_model_train", "oic_model_get", "oic_model_set",
    "oic_model_predict", "oic_model_grad", "oic_get_tuning", "oic_set_tuning", "oic_autotune", "oic_set_skip_kmers",
    "oic_distinct_olaps_batch",
]

NFEAT, NOUT = 12, 2
NPARAM = 64 * NFEAT + 64 + 64 * 64 + 64 + NOUT * 64 + NOUT


class Tuning(ctypes.Structure):
    _fields_ = [("list_wave_min_hits", ctypes.c_uint32), ("fwd_lanes", ctypes.c_uint32), ("fwd_tile", ctypes.c_uint32),
                ("fwd_order", ctypes.c_uint32), ("fwd_graph", ctypes.c_uint32), ("fwd_counters", ctypes.c_uint32),
                ("fwd_loop", ctypes.c_uint32)]


class TuneTrial(ctypes.Structure):
    _fields_ = [("tuning", Tuning), ("stage", ctypes.c_int32), ("ms", ctypes.c_double)]


class TrainParams(ctypes.Structure):
    _fields_ = [("epochs", ctypes.c_uint32), ("batch", ctypes.c_uint32), ("lr", ctypes.c_float), ("beta1", ctypes.c_float),
                ("beta2", ctypes.c_float), ("eps", ctypes.c_float), ("w_exit", ctypes.c_float), ("seed", ctypes.c_uint64)]

RECORD_DTYPE = np.dtype([("a_iid", "<u4"), ("b_iid", "<u4"), ("w0", "<u8"), ("w1", "<u8")])


class OicError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"oic error {code}: {msg}")
        self.code = code


class Params(ctypes.Structure):
    _fields_ = [
        ("kmer", ctypes.c_uint32), ("max_erate", ctypes.c_double), ("align_noise", ctypes.c_double),
        ("min_olap", ctypes.c_int32), ("partial", ctypes.c_int32), ("unique", ctypes.c_int32),
        ("hopeless", ctypes.c_int32), ("minkmers", ctypes.c_int32), ("frag_olap_limit", ctypes.c_uint64),
        ("hash_bits", ctypes.c_uint32), ("hash_load", ctypes.c_double), ("hash_data_len", ctypes.c_uint64),
        ("n_threads", ctypes.c_uint32), ("device", ctypes.c_int32), ("schedule", ctypes.c_int32),
    ]


class ReadsT(ctypes.Structure):
    _fields_ = [("bases", ctypes.c_void_p), ("start", ctypes.c_void_p), ("len", ctypes.c_void_p), ("nreads", ctypes.c_uint32)]


class Sink(ctypes.Structure):
    _fields_ = [("records", ctypes.c_void_p), ("capacity", ctypes.c_uint64), ("count", ctypes.c_uint64),
                ("records_needed", ctypes.c_uint64), ("stats", ctypes.c_uint64 * 8)]


class Counters(ctypes.Structure):
    _fields_ = [(n, ctypes.c_uint64) for n in ("kmer_probes", "hits", "pairs", "nodes", "extensions", "dp_calls",
                                               "dp_cells", "dp_lane_slots", "dp_bases", "overflow_retries",
                                               "spec_tasks", "spec_cells", "spec_rows", "spec_lane_slots", "spec_wave_rows",
                                               "spec_bases", "spec_slide_iters", "spec_wave_slide_iters", "fwd_launches", "fwd_cells", "fwd_bases", "graphs_launched", "list_wave_pairs",
                                               "list_handed_back", "list_second_level")] + \
               [(n, ctypes.c_double) for n in ("ms_index", "ms_gather", "ms_lists", "ms_extend", "ms_total", "ms_dp_kernel",
                                               "ms_fwd_kernel", "ms_fwd_narrow", "ms_fwd_graph", "ms_pp_kernel")] + \
               [(n, ctypes.c_uint64) for n in ("pp_cells", "pp_bases", "learned_pairs", "pp_launches", "spec_wave_chunks", "allocs")] + \
               [("ms_alloc", ctypes.c_double)]


EXT_TASK = np.dtype([("s_id", "<u4"), ("t_id", "<u4"), ("s_dir", "<i4"), ("start", "<i4"), ("offset", "<i4"), ("len", "<i4")])
EXT_RESULT = np.dtype([("kind", "<i4"), ("s_lo", "<i4"), ("s_hi", "<i4"), ("t_lo", "<i4"), ("t_hi", "<i4"),
                       ("errors", "<i4"), ("delta_len", "<i4"), ("status", "<i4")])
OLAP_IN = np.dtype([("s_lo", "<i4"), ("s_hi", "<i4"), ("t_lo", "<i4"), ("t_hi", "<i4"), ("quality", "<f8")])
OLAP_OUT = np.dtype([("s_lo", "<i4"), ("s_hi", "<i4"), ("t_lo", "<i4"), ("t_hi", "<i4"), ("min_diag", "<i4"),
                     ("max_diag", "<i4"), ("s_left", "<i4"), ("s_right", "<i4"), ("t_left", "<i4"), ("t_right", "<i4"),
                     ("deleted", "<i4"), ("pad", "<i4"), ("quality", "<f8")])
MAX_DISTINCT_OLAPS = 3
DP_TASK = np.dtype([("a_id", "<u4"), ("b_id", "<u4"), ("a_strand", "<i4"), ("b_strand", "<i4"), ("a_pos", "<i4"),
                    ("b_pos", "<i4"), ("m", "<i4"), ("n", "<i4"), ("error_limit", "<i4"), ("dir", "<i4")])
DP_RESULT = np.dtype([("errors", "<i4"), ("a_end", "<i4"), ("t_end", "<i4"), ("match_to_end", "<i4"),
                      ("leftover", "<i4"), ("delta_len", "<i4"), ("status", "<i4"), ("pad", "<i4"), ("cells", "<u8")])

_lib = None


def lib():
    """Load liboic_hip.so (fails loudly when it is missing)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: run __graft_entry__.build()")
        L = ctypes.CDLL(LIB_PATH)
        vp = ctypes.c_void_p
        L.oic_params_default.argtypes = [ctypes.POINTER(Params)]
        L.oic_create.argtypes = [ctypes.POINTER(Params), ctypes.POINTER(vp)]
        L.oic_destroy.argtypes = [vp]
        L.oic_last_error.restype = ctypes.c_char_p
        L.oic_last_error.argtypes = [vp]
        L.oic_load_reads.argtypes = [vp, ctypes.POINTER(ReadsT)]
        L.oic_index_build.argtypes = [vp, ctypes.c_uint32, ctypes.c_uint32, ctypes.POINTER(ctypes.c_uint32)]
        L.oic_search.argtypes = [vp, ctypes.c_uint32, ctypes.c_uint32, ctypes.POINTER(Sink)]
        L.oic_run.argtypes = [vp, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, ctypes.POINTER(Sink)]
        L.oic_fetch_pending.argtypes = [vp, ctypes.POINTER(Sink)]
        L.oic_get_counters.argtypes = [vp, ctypes.POINTER(Counters)]
        L.oic_extend_batch.argtypes = [vp, vp, ctypes.c_uint32, vp, vp, ctypes.c_uint32]
        L.oic_prefix_dp_batch.argtypes = [vp, vp, ctypes.c_uint32, vp, vp, ctypes.c_uint32]
        L.oic_distinct_olaps_batch.argtypes = [vp, vp, ctypes.c_uint32, vp, vp, vp]
        L.oic_prefix_dp_batch_noalloc.argtypes = [vp, vp, ctypes.c_uint32, vp, vp, ctypes.c_uint32]
        L.oic_tables.argtypes = [vp, ctypes.POINTER(ctypes.POINTER(ctypes.c_int32)),
                                 ctypes.POINTER(ctypes.POINTER(ctypes.c_int32)), ctypes.POINTER(ctypes.c_uint32)]
        L.oic_dp_tables.argtypes = [ctypes.c_double, ctypes.c_double, vp, vp, ctypes.POINTER(ctypes.c_uint32)]
        L.oic_set_tracing.argtypes = [vp, ctypes.c_int]
        L.oic_get_traces.argtypes = [vp, vp, vp, ctypes.c_uint64, ctypes.POINTER(ctypes.c_uint64)]
        L.oic_model_train.argtypes = [vp, vp, vp, ctypes.c_uint64, ctypes.POINTER(TrainParams), ctypes.POINTER(ctypes.c_float)]
        L.oic_model_get.argtypes = [vp, vp, vp, vp]
        L.oic_model_set.argtypes = [vp, vp, vp, vp]
        L.oic_model_predict.argtypes = [vp, vp, ctypes.c_uint64, vp]
        L.oic_model_grad.argtypes = [vp, vp, vp, vp, ctypes.c_uint64, ctypes.c_float, ctypes.POINTER(ctypes.c_float), vp]
        L.oic_set_skip_kmers.argtypes = [vp, ctypes.c_char_p, ctypes.c_uint64]
        L.oic_get_tuning.argtypes = [vp, ctypes.POINTER(Tuning)]
        L.oic_set_tuning.argtypes = [vp, ctypes.POINTER(Tuning)]
        L.oic_autotune.argtypes = [vp, ctypes.c_uint32, ctypes.c_uint32, ctypes.POINTER(Tuning), ctypes.POINTER(TuneTrial),
                                   ctypes.c_uint32, ctypes.POINTER(ctypes.c_uint32)]
        _lib = L
    return _lib


def dp_tables(maxerate, alignnoise=1.0):
    """(Error_Bound, Edit_Match_Limit) as the engine builds them; needs no device."""
    me = ctypes.c_uint32()
    rc = lib().oic_dp_tables(maxerate, alignnoise, None, None, ctypes.byref(me))
    if rc != 0:
        raise OicError(rc, "oic_dp_tables")
    eb = np.zeros(1 << 21, np.int32)
    ml = np.zeros(me.value, np.int32)
    lib().oic_dp_tables(maxerate, alignnoise, eb.ctypes.data, ml.ctypes.data, ctypes.byref(me))
    return eb, ml


@dataclasses.dataclass
class Result:
    records: np.ndarray           # RECORD_DTYPE, in output order
    stats: list                   # the 8 .stats counters
    counters: dict

    def stats_text(self):
        """The .stats file overlapInCore writes (overlapInCore.C:550-558)."""
        names = [" Kmer hits without olaps", "    Kmer hits with olaps", "  Multiple overlaps/pair",
                 " Total overlaps produced", "      Contained overlaps", "       Dovetail overlaps",
                 "Rejected by short window", " Rejected by long window"]
        return "".join(f"{n} = {v}\n" for n, v in zip(names, self.stats))


class Engine:
    def __init__(self, kmer=22, maxerate=0.06, minlength=0, alignnoise=1.0, partial=False, unique=True,
                 hopeless=True, minkmers=False, hashbits=22, hashload=0.6, hashdatalen=100000000, threads=1,
                 device=0, schedule=SCHED_PLAIN):
        L = lib()
        p = Params()
        L.oic_params_default(ctypes.byref(p))
        p.kmer, p.max_erate, p.align_noise, p.min_olap = kmer, maxerate, alignnoise, minlength
        p.partial, p.unique, p.hopeless, p.minkmers = int(partial), int(unique), int(hopeless), int(minkmers)
        p.hash_bits, p.hash_load, p.hash_data_len, p.n_threads = hashbits, hashload, hashdatalen, threads
        p.device, p.schedule = device, schedule
        self._ctx = ctypes.c_void_p()
        rc = L.oic_create(ctypes.byref(p), ctypes.byref(self._ctx))
        if rc != 0:
            raise OicError(rc, "oic_create failed (no HIP device, or invalid parameters)")
        self.params = p
        self._reads = None

    def close(self):
        if self._ctx:
            lib().oic_destroy(self._ctx)
            self._ctx = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc):
        if rc != 0:
            raise OicError(rc, lib().oic_last_error(self._ctx).decode())

    def load_reads(self, reads):
        self._reads = (np.ascontiguousarray(reads.bases), np.ascontiguousarray(reads.start, np.uint64),
                       np.ascontiguousarray(reads.length, np.uint32))
        r = ReadsT(self._reads[0].ctypes.data, self._reads[1].ctypes.data, self._reads[2].ctypes.data, int(reads.n))
        self._check(lib().oic_load_reads(self._ctx, ctypes.byref(r)))
        self.nreads = int(reads.n)

    def build_hash_index(self, bgn, end):
        last = ctypes.c_uint32()
        self._check(lib().oic_index_build(self._ctx, bgn, end, ctypes.byref(last)))
        return last.value

    def _collect(self, call, cap=1 << 16):
        """Run a search into a sink of cap records; when they do not fit, take the records the
        engine holds into a larger buffer (oic_fetch_pending) instead of searching again."""
        buf = np.zeros(cap, RECORD_DTYPE)
        s = Sink(buf.ctypes.data, cap, 0, 0)
        rc = call(s)
        if rc == OIC_ECAPACITY:
            need = int(s.records_needed)
            big = np.zeros(need, RECORD_DTYPE)
            big[:s.count] = buf[:s.count]
            buf = big
            s.records, s.capacity = buf.ctypes.data, need
            rc = lib().oic_fetch_pending(self._ctx, ctypes.byref(s))
        self._check(rc)
        return Result(buf[:s.count].copy(), list(s.stats), self.counters())

    def search(self, bgn_ref, end_ref, cap=1 << 16):
        return self._collect(lambda s: lib().oic_search(self._ctx, bgn_ref, end_ref, ctypes.byref(s)), cap)

    def run(self, bgn_hash=1, end_hash=0xFFFFFFFF, bgn_ref=1, end_ref=0xFFFFFFFF, cap=1 << 16):
        return self._collect(lambda s: lib().oic_run(self._ctx, bgn_hash, end_hash, bgn_ref, end_ref, ctypes.byref(s)), cap)

    def counters(self):
        c = Counters()
        self._check(lib().oic_get_counters(self._ctx, ctypes.byref(c)))
        return {n: getattr(c, n) for n, _ in Counters._fields_}

    def extend_alignment(self, tasks, delta_stride=4096):
        tasks = np.ascontiguousarray(tasks, EXT_TASK)
        res = np.zeros(len(tasks), EXT_RESULT)
        delta = np.zeros((len(tasks), delta_stride), np.int32)
        self._check(lib().oic_extend_batch(self._ctx, tasks.ctypes.data, len(tasks), res.ctypes.data, delta.ctypes.data, delta_stride))
        return res, delta

    def distinct_olaps(self, groups):
        """Add_Overlap over each pair's extension results, then the pair's selection under this
        engine's partial/unique parameters.  groups: one sequence of OLAP_IN rows per pair.  Returns
        one OLAP_OUT array per pair (its distinct overlaps, deleted ones flagged)."""
        first = np.zeros(len(groups) + 1, np.uint32)
        first[1:] = np.cumsum([len(g) for g in groups])
        rows = np.concatenate([np.asarray(g, OLAP_IN) for g in groups]) if first[-1] else np.zeros(1, OLAP_IN)
        rows = np.ascontiguousarray(rows, OLAP_IN)
        out = np.zeros((max(1, len(groups)), MAX_DISTINCT_OLAPS), OLAP_OUT)
        count = np.zeros(max(1, len(groups)), np.int32)
        self._check(lib().oic_distinct_olaps_batch(self._ctx, first.ctypes.data, len(groups), rows.ctypes.data,
                                                   out.ctypes.data, count.ctypes.data))
        return [out[g, :count[g]].copy() for g in range(len(groups))]

    def prefix_dp(self, tasks, delta_stride=4096):
        tasks = np.ascontiguousarray(tasks, DP_TASK)
        res = np.zeros(len(tasks), DP_RESULT)
        delta = np.zeros((len(tasks), delta_stride), np.int32)
        self._check(lib().oic_prefix_dp_batch(self._ctx, tasks.ctypes.data, len(tasks), res.ctypes.data, delta.ctypes.data, delta_stride))
        return res, delta

    def set_skip_kmers(self, text):
        """-k <file>: the file's text; its k-mers (both strands) give no seed hits and screen read ends."""
        b = text.encode() if isinstance(text, str) else bytes(text)
        self._check(lib().oic_set_skip_kmers(self._ctx, b, len(b)))

    def tuning(self):
        t = Tuning()
        self._check(lib().oic_get_tuning(self._ctx, ctypes.byref(t)))
        return {n: getattr(t, n) for n, _ in Tuning._fields_}

    def set_tuning(self, **kw):
        """Change launch-shape knobs (oic_tuning); results never depend on them."""
        t = Tuning()
        self._check(lib().oic_get_tuning(self._ctx, ctypes.byref(t)))
        for k, v in kw.items():
            if not hasattr(t, k):
                raise KeyError(k)
            setattr(t, k, v)
        self._check(lib().oic_set_tuning(self._ctx, ctypes.byref(t)))

    def autotune(self, bgn_ref, end_ref):
        """Pick the fastest launch shapes on a sample of ref reads (oic_autotune); returns
        (chosen tuning, list of trials)."""
        best = Tuning()
        cap = 64
        trials = (TuneTrial * cap)()
        n = ctypes.c_uint32()
        self._check(lib().oic_autotune(self._ctx, bgn_ref, end_ref, ctypes.byref(best), trials, cap, ctypes.byref(n)))
        fields = [f for f, _ in Tuning._fields_]
        out = [dict({f: getattr(t.tuning, f) for f in fields}, stage=("fwd", "lists", "extend")[t.stage], ms=t.ms)
               for t in trials[:min(n.value, cap)]]
        return {f: getattr(best, f) for f in fields}, out

    # ---- learned cost model
    def set_tracing(self, on=True):
        self._check(lib().oic_set_tracing(self._ctx, int(on)))

    def traces(self):
        """(features n x 12, targets n x 2) of the pairs searched since set_tracing(True)."""
        n = ctypes.c_uint64()
        self._check(lib().oic_get_traces(self._ctx, None, None, 0, ctypes.byref(n)))
        f = np.zeros((n.value, NFEAT), np.float32)
        t = np.zeros((n.value, NOUT), np.float32)
        self._check(lib().oic_get_traces(self._ctx, f.ctypes.data, t.ctypes.data, n.value, ctypes.byref(n)))
        return f, t

    def train_model(self, feat, targ, epochs=12, batch=4096, lr=2e-3, w_exit=0.05, seed=1):
        feat = np.ascontiguousarray(feat, np.float32)
        targ = np.ascontiguousarray(targ, np.float32)
        tp = TrainParams(epochs, batch, lr, 0.9, 0.999, 1e-8, w_exit, seed)
        loss = ctypes.c_float()
        self._check(lib().oic_model_train(self._ctx, feat.ctypes.data, targ.ctypes.data, len(feat), ctypes.byref(tp), ctypes.byref(loss)))
        return loss.value

    def model(self):
        w = np.zeros(NPARAM, np.float32)
        mu = np.zeros(NFEAT, np.float32)
        sd = np.zeros(NFEAT, np.float32)
        self._check(lib().oic_model_get(self._ctx, w.ctypes.data, mu.ctypes.data, sd.ctypes.data))
        return w, mu, sd

    def set_model(self, w, mu, sd):
        w, mu, sd = (np.ascontiguousarray(a, np.float32) for a in (w, mu, sd))
        self._check(lib().oic_model_set(self._ctx, w.ctypes.data, mu.ctypes.data, sd.ctypes.data))

    def predict(self, feat):
        feat = np.ascontiguousarray(feat, np.float32)
        out = np.zeros((len(feat), NOUT), np.float32)
        self._check(lib().oic_model_predict(self._ctx, feat.ctypes.data, len(feat), out.ctypes.data))
        return out

    def model_grad(self, w, z, t, w_exit=0.05):
        w, z, t = (np.ascontiguousarray(a, np.float32) for a in (w, z, t))
        g = np.zeros(NPARAM, np.float32)
        loss = ctypes.c_float()
        self._check(lib().oic_model_grad(self._ctx, w.ctypes.data, z.ctypes.data, t.ctypes.data, len(z), w_exit,
                                         ctypes.byref(loss), g.ctypes.data))
        return loss.value, g

    def tables(self):
        eb = ctypes.POINTER(ctypes.c_int32)()
        ml = ctypes.POINTER(ctypes.c_int32)()
        me = ctypes.c_uint32()
        self._check(lib().oic_tables(self._ctx, ctypes.byref(eb), ctypes.byref(ml), ctypes.byref(me)))
        return (np.ctypeslib.as_array(eb, ((1 << 21),)).copy(), np.ctypeslib.as_array(ml, (me.value,)).copy())


def sort_records(recs):
    """Canonical order of ovOverlap::operator< (ovOverlap.H:263-277)."""
    if len(recs) == 0:
        return recs
    return recs[np.lexsort((recs["w1"], recs["w0"], recs["b_iid"], recs["a_iid"]))]
