// oic.hip -- liboic_hip.so: the C ABI of include/oic.h over the gfx950 kernels.
//
// Host orchestration mirrors overlapInCore's driver (src/overlapInCore/overlapInCore.C:161-277):
// a hash-block loop around Build_Hash_Index, then the ref range searched in both orientations.
// Ref reads are processed in batches sized to bound device memory; every stage of a batch runs
// on the context's stream.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/oic.h"
#include "oic_tables.h"
#include "oic_sort.h"
#include "oic_search.h"
#include "oic_sched.h"
#include "oic_batch.h"
#include "oic_flat.h"
#include "oic_lane.h"
#include "oic_mlp.h"
#include <random>
#include <numeric>

namespace {

// Owns one device allocation.  Not copyable: a copy's destructor would free the buffer under the
// original.  Exchange ownership with swap().
struct DevBuf {
  void *p = nullptr;
  size_t bytes = 0;
  DevBuf() = default;
  DevBuf(const DevBuf &) = delete;
  DevBuf &operator=(const DevBuf &) = delete;
  void swap(DevBuf &o) {
    std::swap(p, o.p);
    std::swap(bytes, o.bytes);
  }
  hipError_t ensure(size_t b) {
    if (b <= bytes && p) return hipSuccess;
    release();
    size_t nb = b < 256 ? 256 : b;
    hipError_t e = hipMalloc(&p, nb);
    if (e != hipSuccess) { p = nullptr; bytes = 0; (void)hipGetLastError(); return e; }
    bytes = nb;
    return hipSuccess;
  }
  void release() {
    if (p) hipFree(p);
    p = nullptr;
    bytes = 0;
  }
  template <class T> T *as() const { return (T *)p; }
  ~DevBuf() { release(); }
};

double now_ms() {
  using namespace std::chrono;
  return duration<double, std::milli>(steady_clock::now().time_since_epoch()).count();
}

int ceil_log2(uint64_t x) {
  int b = 0;
  while ((1ull << b) < x) b++;
  return b;
}

constexpr int kCapSmall = 512;
constexpr uint64_t kPassBytesPerHit = 140;   // device bytes a search pass holds per seed hit (hits, sort, lists, nodes, outcomes)
constexpr int kCapLarge = 8192;

}  // namespace

struct oic_ctx {
  oic_params P{};
  oic::Tables T;
  std::string err;
  hipStream_t st = nullptr;
  hipStream_t st2 = nullptr;      // second stream: the 512-hit list level beside the 128-hit one
  hipEvent_t evl[2] = {};
  int ncu = 0;

  DevBuf d_eb, d_ml, d_minkmers;
  // reads
  uint32_t nreads = 0;
  int32_t maxlen = 0;
  std::vector<int32_t> h_len;
  DevBuf d_words, d_nmask, d_b0f, d_b0r, d_len, d_hasN;
  // index of the current hash block
  bool have_index = false;
  uint32_t hbgn = 0, hend = 0, nblock = 0;
  uint64_t nu = 0, nvalid = 0, npos = 0;
  int dshift = 0;
  DevBuf x_ukey, x_ustart, x_dir, x_occ;
  // sort / scan temporaries
  DevBuf s_k1, s_v1, s_k2, s_v2, s_scan, s_flag, s_idx, s_small, s_rp;
  // search
  DevBuf g_cnt, g_off, g_pbase, g_searched, g_sdsum, g_phi, g_pval;
  DevBuf l_scr, l_done, l_big, l_cnt, l_crowd;
  std::vector<uint64_t> skip;           // -k skip k-mers, both strands, sorted, unique
  DevBuf d_skip, x_tscreen, g_sscreen;
  oic_tuning tune{};
  uint64_t hit_cap = 0;                 // seed hits per search pass
  uint64_t h_alloc = 0;                 // seed hits the per-hit pass buffers are sized for
  int wallclock_khz = 0;                // rate of wall_clock64()
  DevBuf p_pairs, p_tmp, p_nodes, p_recs, p_pout, p_order, p_queue, p_stats, p_nout, p_noff, p_outrec, p_retry;
  DevBuf sc_rowL, sc_rowR, sc_rowOff, sc_rowVal, sc_dbuf, sc_ld, sc_rd;
  DevBuf b_outc, b_probe, b_spec, b_cnt, b_off, b_stats, b_order, b_ovf, b_novf;
  // learned cost model
  bool tracing = false;
  std::vector<float> tr_feat, tr_targ;
  float mparam[MLP_NPARAM] = {0}, mmu[MLP_IN] = {0}, msd[MLP_IN] = {0};
  DevBuf m_param, m_mu, m_sd, m_feat, m_pred;
  int32_t sc_rows = 0;
  int64_t sc_val = 0;
  uint32_t sc_waves = 0;
  oic_counters ctr{};
  bool have_model = false;
  bool batch_pred = false;              // m_pred holds finite predictions for the current batch's pairs
  // records of the last oic_search / oic_run that did not fit the caller's sink, held until
  // oic_fetch_pending or the next search
  std::vector<oic_overlap> pending;
  void *static_scratch = nullptr;       // this device's copy of the no-allocation entry point's scratch
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  hipEvent_t evf[12] = {};
};

static int fail(oic_ctx *c, int code, const char *fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  if (c) c->err = buf;
  return code;
}

static int lookup_scratch(oic_ctx *c, int dev);

#define HIPCHK(c, x)                                                                      \
  do {                                                                                    \
    hipError_t _e = (x);                                                                  \
    if (_e != hipSuccess) return fail((c), OIC_EHIP, "%s: %s (%s:%d)", #x, hipGetErrorString(_e), __FILE__, __LINE__); \
  } while (0)
#define ALLOC(c, buf, nbytes)                                                             \
  do {                                                                                    \
    const size_t _b = (nbytes);                                                           \
    if (_b > (buf).bytes || !(buf).p) {                                                   \
      const double _t = now_ms();                                                         \
      if ((buf).ensure(_b) != hipSuccess)                                                 \
        return fail((c), OIC_ENOMEM, "device allocation of %zu bytes failed (%s:%d)", _b, __FILE__, __LINE__); \
      (c)->ctr.allocs++;                                                                  \
      (c)->ctr.ms_alloc += now_ms() - _t;                                                 \
    }                                                                                     \
  } while (0)

static DPTables dev_tables(oic_ctx *c) {
  DPTables tb;
  tb.error_bound = c->d_eb.as<int32_t>();
  tb.match_limit = c->d_ml.as<int32_t>();
  tb.bmv = c->T.bmv;
  tb.slope_min = c->T.slope_min;
  tb.min_branch_end_dist = 20;
  tb.partial = c->P.partial;
  return tb;
}

static ReadsDev dev_reads(oic_ctx *c) {
  ReadsDev R;
  R.words = c->d_words.as<uint32_t>();
  R.nmask = c->d_nmask.as<uint32_t>();
  R.b0f = c->d_b0f.as<int64_t>();
  R.b0r = c->d_b0r.as<int64_t>();
  R.len = c->d_len.as<int32_t>();
  R.hasN = c->d_hasN.as<uint8_t>();
  return R;
}

extern "C" void oic_params_default(oic_params *p) {
  memset(p, 0, sizeof(*p));
  p->kmer = 22;
  p->max_erate = 0.06;
  p->align_noise = 1.0;
  p->min_olap = 0;
  p->unique = 1;
  p->hopeless = 1;
  p->frag_olap_limit = UINT64_MAX;
  p->hash_bits = 22;
  p->hash_load = 0.6;
  p->hash_data_len = 100000000;
  p->n_threads = 1;
  p->device = 0;
  p->schedule = OIC_SCHED_PLAIN;
}

extern "C" const char *oic_last_error(const oic_ctx *c) { return c ? c->err.c_str() : "null context"; }

extern "C" int oic_create(const oic_params *p, oic_ctx **out) {
  if (!p || !out) return OIC_EINVAL;
  *out = nullptr;
  oic_ctx *c = new oic_ctx();
  c->P = *p;
  if (p->kmer < 1 || p->kmer > 31) { int r = fail(c, OIC_EINVAL, "kmer %u out of range 1..31", p->kmer); delete c; return r; }
  if (!(p->max_erate > 0.0) || p->max_erate > 0.5) { delete c; return OIC_EINVAL; }
  if (p->frag_olap_limit != UINT64_MAX) { delete c; return OIC_EUNSUPPORTED; }
  if (p->n_threads == 0) c->P.n_threads = 1;
  if (c->P.max_erate > 0.06) {}
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
    (void)hipGetLastError();
    delete c;
    return OIC_EHIP;
  }
  if (hipSetDevice(p->device) != hipSuccess) { delete c; return OIC_EHIP; }
  if (lookup_scratch(c, p->device) != OIC_OK) { delete c; return OIC_EHIP; }
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, p->device) != hipSuccess) { delete c; return OIC_EHIP; }
  c->ncu = prop.multiProcessorCount;
  {
    int khz = 0;
    if (hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, p->device) == hipSuccess) c->wallclock_khz = khz;
  }
  c->tune.list_wave_min_hits = 2048;
  {
    // seed hits per search pass: 45 % of the device's memory for pass buffers (about 9e8 on a 288 GB MI355X)
    size_t fr = 0, tot = 0;
    if (hipMemGetInfo(&fr, &tot) != hipSuccess) tot = (size_t)64 << 30;
    c->hit_cap = std::min<uint64_t>(1200000000ull, std::max<uint64_t>(50000000ull, (uint64_t)(0.45 * (double)tot) / kPassBytesPerHit));
  }
  c->tune.fwd_lanes = 4;
  c->tune.fwd_tile = 64;
  c->tune.fwd_loop = 2;
  if (hipStreamCreateWithFlags(&c->st, hipStreamNonBlocking) != hipSuccess) { delete c; return OIC_EHIP; }
  hipEventCreate(&c->ev0);
  hipEventCreate(&c->ev1);
  for (auto &e : c->evf) hipEventCreate(&e);
  if (hipStreamCreateWithFlags(&c->st2, hipStreamNonBlocking) != hipSuccess) { delete c; return OIC_EHIP; }
  for (auto &e : c->evl) hipEventCreateWithFlags(&e, hipEventDisableTiming);
  c->tune.fwd_graph = 1;
  c->T.build(p->max_erate, p->align_noise);
  // overlapInCore.C:400-401
  if (c->T.maxErate > 0.06) c->P.hopeless = 0;
  int rc = OIC_OK;
  if (c->d_eb.ensure(c->T.error_bound.size() * 4) != hipSuccess || c->d_ml.ensure(c->T.match_limit.size() * 4) != hipSuccess)
    rc = OIC_ENOMEM;
  if (rc == OIC_OK) {
    hipMemcpy(c->d_eb.p, c->T.error_bound.data(), c->T.error_bound.size() * 4, hipMemcpyHostToDevice);
    hipMemcpy(c->d_ml.p, c->T.match_limit.data(), c->T.match_limit.size() * 4, hipMemcpyHostToDevice);
  }
  if (rc == OIC_OK && c->P.minkmers) {
    // overlapInCore.C:410-411 and Process_String_Overlaps.C:22-33, in the reference's types
    const uint64_t K = c->P.kmer;
    const uint64_t span = (uint64_t)(int64_t)c->P.min_olap - K + 1;
    const uint64_t filt = (uint64_t)(int)(std::floor(std::exp(-1.0 * (double)K * c->T.maxErate) * (double)span));
    std::vector<uint32_t> tab(oic::kMaxReadLen + 1);
    for (uint32_t o = 0; o <= oic::kMaxReadLen; o++) {
      const double ovl = (double)o;
      uint64_t ex = 0;
      if (!(ovl < (double)K)) ex = (uint64_t)(int)(std::floor(std::exp(-1.0 * (double)K * c->T.maxErate) * (ovl - (double)K + 1)));
      const uint64_t v = filt ? std::max(filt, ex) : 0;
      tab[o] = (uint32_t)std::min<uint64_t>(v, 0xffffffffu);
    }
    if (filt == 0) c->P.minkmers = 0;
    else if (c->d_minkmers.ensure(tab.size() * 4) != hipSuccess) rc = OIC_ENOMEM;
    else hipMemcpy(c->d_minkmers.p, tab.data(), tab.size() * 4, hipMemcpyHostToDevice);
  }
  if (rc != OIC_OK) { delete c; return rc; }
  *out = c;
  return OIC_OK;
}

extern "C" void oic_destroy(oic_ctx *c) {
  if (!c) return;
  if (c->st) hipStreamSynchronize(c->st);
  if (c->ev0) hipEventDestroy(c->ev0);
  if (c->ev1) hipEventDestroy(c->ev1);
  for (auto &e : c->evf)
    if (e) hipEventDestroy(e);
  if (c->st2) hipStreamSynchronize(c->st2);
  for (auto &e : c->evl)
    if (e) hipEventDestroy(e);
  if (c->st2) hipStreamDestroy(c->st2);
  if (c->st) hipStreamDestroy(c->st);
  delete c;
}

extern "C" int oic_tables(const oic_ctx *c, const int32_t **eb, const int32_t **ml, uint32_t *max_errors) {
  if (!c) return OIC_EINVAL;
  if (eb) *eb = c->T.error_bound.data();
  if (ml) *ml = c->T.match_limit.data();
  if (max_errors) *max_errors = c->T.max_errors;
  return OIC_OK;
}

extern "C" int oic_dp_tables(double max_erate, double align_noise, int32_t *eb, int32_t *ml, uint32_t *max_errors) {
  if (!(max_erate > 0.0) || max_erate > 0.5 || !max_errors) return OIC_EINVAL;
  oic::Tables T;
  T.build(max_erate, align_noise);
  if (eb) memcpy(eb, T.error_bound.data(), T.error_bound.size() * 4);
  if (ml) memcpy(ml, T.match_limit.data(), (size_t)T.max_errors * 4);
  *max_errors = T.max_errors;
  return OIC_OK;
}

// Mark_Skip_Kmers' reading of a -k file (overlapInCore-Build_Hash_Index.C:185-257): each line -- or,
// after a '>' line, the next one -- is cut at its first blank and must hold exactly k bases; the
// k-mer and its reverse complement are skipped.
extern "C" int oic_set_skip_kmers(oic_ctx *c, const char *text, uint64_t len) {
  if (!c || (!text && len)) return OIC_EINVAL;
  const uint32_t K = c->P.kmer;
  std::vector<uint64_t> keys;
  uint64_t i = 0;
  int32_t line_no = 0;
  auto next_line = [&](std::string &out) -> bool {
    if (i >= len) return false;
    const uint64_t b = i;
    while (i < len && text[i] != '\n') i++;
    out.assign(text + b, i - b);
    if (i < len) i++;
    line_no++;
    return true;
  };
  std::string line;
  while (next_line(line)) {
    if (!line.empty() && line[0] == '>') {
      if (!next_line(line)) line.clear();
    }
    const size_t cut = line.find_first_of(" \t\r\n");
    if (cut != std::string::npos) line.resize(cut);
    if (line.size() != K)
      return fail(c, OIC_EINVAL, "Short kmer skip kmer '%s' at line %d, expecting length %u got length %zu.", line.c_str(), line_no, K,
                  line.size());
    uint64_t f = 0, r = 0;
    bool ok = true;
    for (uint32_t j = 0; j < K; j++) {
      const char ch = (char)tolower((unsigned char)line[j]);
      const int code = ch == 'a' ? 0 : ch == 'c' ? 1 : ch == 'g' ? 2 : ch == 't' ? 3 : -1;
      if (code < 0) { ok = false; break; }
      f |= (uint64_t)code << (2 * j);                    // base j at bits 2j, as the index keys
      r |= (uint64_t)(3 - code) << (2 * (K - 1 - j));    // reverse complement
    }
    if (!ok) return fail(c, OIC_EUNSUPPORTED, "skip k-mer '%s' at line %d holds a base other than ACGT", line.c_str(), line_no);
    keys.push_back(f);
    keys.push_back(r);
  }
  std::sort(keys.begin(), keys.end());
  keys.erase(std::unique(keys.begin(), keys.end()), keys.end());
  c->skip = keys;
  if (!keys.empty()) {
    ALLOC(c, c->d_skip, keys.size() * 8);
    HIPCHK(c, hipMemcpy(c->d_skip.p, keys.data(), keys.size() * 8, hipMemcpyHostToDevice));
  }
  c->have_index = false;   // takes effect at the next index build
  return OIC_OK;
}

extern "C" int oic_get_tuning(const oic_ctx *c, oic_tuning *out) {
  if (!c || !out) return OIC_EINVAL;
  *out = c->tune;
  return OIC_OK;
}

extern "C" int oic_set_tuning(oic_ctx *c, const oic_tuning *t) {
  if (!c || !t) return OIC_EINVAL;
  const bool lanes_ok = t->fwd_lanes == 1 || t->fwd_lanes == 4 || t->fwd_lanes == 8 || t->fwd_lanes == 16 || t->fwd_lanes == 32;
  const bool tile_ok = t->fwd_lanes == 1
                           ? (t->fwd_tile == 16 || t->fwd_tile == 32 || t->fwd_tile == 64)
                           : (t->fwd_tile == 128 || t->fwd_tile == 256 || (t->fwd_tile == 512 && t->fwd_lanes != 4) ||
                              ((t->fwd_tile == 64 || t->fwd_tile == 32) && (t->fwd_lanes == 8 || t->fwd_lanes == 16)));
  if (t->fwd_order > 1) return fail(c, OIC_EINVAL, "fwd_order must be 0 or 1");
  if (t->fwd_loop > 2) return fail(c, OIC_EINVAL, "fwd_loop must be 0, 1 or 2");
  if (t->fwd_loop == 2) {
    if ((t->fwd_lanes != 4 && t->fwd_lanes != 8) || (t->fwd_tile != 32 && t->fwd_tile != 64 && t->fwd_tile != 128))
      return fail(c, OIC_EINVAL, "fwd_loop 2 (flat rows) runs 4 or 8 lanes per DP with fwd_tile 32, 64 or 128");
    c->tune = *t;
    return OIC_OK;
  }
  if (!lanes_ok || !tile_ok)
    return fail(c, OIC_EINVAL, "fwd_lanes must be 1 (tile 16, 32 or 64), or 4, 8, 16 or 32 with fwd_tile 32 or 64 (8 or 16 lanes), "
                               "128, 256 or 512 (not with 4 lanes)");
  c->tune = *t;
  return OIC_OK;
}

extern "C" int oic_get_counters(const oic_ctx *c, oic_counters *out) {
  if (!c || !out) return OIC_EINVAL;
  *out = c->ctr;
  return OIC_OK;
}

// ------------------------------------------------------------------------------------ reads

extern "C" int oic_load_reads(oic_ctx *c, const oic_reads *rd) {
  if (!c || !rd || !rd->bases || !rd->start || !rd->len) return OIC_EINVAL;
  const uint32_t n = rd->nreads;
  c->nreads = n;
  c->have_index = false;
  c->h_len.assign((size_t)n + 1, 0);
  std::vector<int64_t> b0f((size_t)n + 1, 0), b0r((size_t)n + 1, 0);
  std::vector<uint64_t> start((size_t)n + 1, 0);
  uint64_t lo = UINT64_MAX, hi = 0;
  int64_t acc = 0;
  c->maxlen = 0;
  for (uint32_t r = 1; r <= n; r++) {
    const uint32_t L = rd->len[r];
    if (L > oic::kMaxReadLen) return fail(c, OIC_EINVAL, "read %u has length %u > 2^21-1", r, L);
    c->h_len[r] = (int32_t)L;
    c->maxlen = std::max<int32_t>(c->maxlen, (int32_t)L);
    const int64_t span = ((int64_t)(L + 31) / 32 + 2) * 32;
    b0f[r] = acc; acc += span;
    b0r[r] = acc; acc += span;
    lo = std::min<uint64_t>(lo, rd->start[r]);
    hi = std::max<uint64_t>(hi, rd->start[r] + L);
  }
  if (n == 0) { lo = hi = 0; }
  for (uint32_t r = 1; r <= n; r++) start[r] = rd->start[r] - lo;
  // slack after the last strand: the packed forward-DP kernels stage OIC_SEQ_WIN words from a
  // tail's start whatever its length
  const uint64_t nwords = (uint64_t)acc / 16 + 4 + 2 * OIC_SEQ_WIN, nmw = (uint64_t)acc / 32 + 4;
  ALLOC(c, c->d_words, nwords * 4);
  ALLOC(c, c->d_nmask, nmw * 4);
  ALLOC(c, c->d_b0f, ((size_t)n + 1) * 8);
  ALLOC(c, c->d_b0r, ((size_t)n + 1) * 8);
  ALLOC(c, c->d_len, ((size_t)n + 1) * 4);
  ALLOC(c, c->d_hasN, (size_t)n + 1);
  DevBuf raw, dstart;
  ALLOC(c, raw, (hi - lo) + 1);
  ALLOC(c, dstart, ((size_t)n + 1) * 8);
  HIPCHK(c, hipMemsetAsync(c->d_words.p, 0, nwords * 4, c->st));
  HIPCHK(c, hipMemsetAsync(c->d_nmask.p, 0, nmw * 4, c->st));
  HIPCHK(c, hipMemsetAsync(c->d_hasN.p, 0, (size_t)n + 1, c->st));
  HIPCHK(c, hipMemcpyAsync(raw.p, rd->bases + lo, hi - lo, hipMemcpyHostToDevice, c->st));
  HIPCHK(c, hipMemcpyAsync(dstart.p, start.data(), ((size_t)n + 1) * 8, hipMemcpyHostToDevice, c->st));
  HIPCHK(c, hipMemcpyAsync(c->d_b0f.p, b0f.data(), ((size_t)n + 1) * 8, hipMemcpyHostToDevice, c->st));
  HIPCHK(c, hipMemcpyAsync(c->d_b0r.p, b0r.data(), ((size_t)n + 1) * 8, hipMemcpyHostToDevice, c->st));
  HIPCHK(c, hipMemcpyAsync(c->d_len.p, c->h_len.data(), ((size_t)n + 1) * 4, hipMemcpyHostToDevice, c->st));
  if (n) {
    hipLaunchKernelGGL(k_pack, dim3(n), dim3(256), 0, c->st, raw.as<uint8_t>(), dstart.as<uint64_t>(), c->d_len.as<int32_t>(),
                       c->d_b0f.as<int64_t>(), c->d_b0r.as<int64_t>(), c->d_words.as<uint32_t>(), c->d_nmask.as<uint32_t>(),
                       c->d_hasN.as<uint8_t>(), n);
  }
  HIPCHK(c, hipGetLastError());
  HIPCHK(c, hipStreamSynchronize(c->st));
  return OIC_OK;
}

// ------------------------------------------------------------------------------------ index

// alternate key / value buffers for sort_pairs, and scan scratch, for n elements
static int ensure_sort_temp(oic_ctx *c, uint64_t n) {
  ALLOC(c, c->s_k2, n * 8);
  ALLOC(c, c->s_v2, n * 8);
  ALLOC(c, c->s_scan, scan_scratch_words(n) * 8);
  return OIC_OK;
}

// Stable sort of n (key, value) pairs on key bits [0, nbits): LSD passes of 8 to 11-bit digits
// (oic_sort.h; as few passes as 11-bit digits allow, the digit width then evened out), ping-ponging
// with the alternate buffers (s_k2 / s_v2); *k / *v point at the result.
template <int BITS>
static int sort_pass(oic_ctx *c, const uint64_t *ki, const uint64_t *vi, uint64_t *ko, uint64_t *vo, uint64_t n, int shift) {
  const uint64_t ntiles = (n + RS_TILE - 1) / RS_TILE;
  const uint64_t ncount = ((uint64_t)1 << BITS) * ntiles;
  uint32_t *counts = (uint32_t *)c->s_rp.p;
  uint64_t *offs = (uint64_t *)((uint8_t *)c->s_rp.p + ncount * 4);
  hipLaunchKernelGGL(k_rs_hist<BITS>, dim3((unsigned)ntiles), dim3(RS_THREADS), 0, c->st, ki, n, shift, counts, (uint32_t)ntiles);
  HIPCHK(c, scan_excl<uint32_t>(counts, offs, ncount, c->s_small.as<uint64_t>(), c->st, nullptr));
  hipLaunchKernelGGL(k_rs_scatter<BITS>, dim3((unsigned)ntiles), dim3(RS_THREADS), 0, c->st, ki, vi, ko, vo, n, shift, offs,
                     (uint32_t)ntiles);
  HIPCHK(c, hipGetLastError());
  return OIC_OK;
}

static int sort_pairs(oic_ctx *c, uint64_t **k, uint64_t **v, uint64_t n, int nbits) {
  if (n < 2 || nbits <= 0) return OIC_OK;
  const uint64_t ntiles = (n + RS_TILE - 1) / RS_TILE;
  if (ntiles > 0xffffffffull) return fail(c, OIC_EINVAL, "sort of %lu keys has too many tiles", (unsigned long)n);
  const int passes = (nbits + 10) / 11;
  const int bits = std::max(8, (nbits + passes - 1) / passes);
  const uint64_t ncount = ((uint64_t)1 << bits) * ntiles;
  ALLOC(c, c->s_rp, ncount * 12);
  ALLOC(c, c->s_small, scan_scratch_words(ncount) * 8);
  uint64_t *ki = *k, *vi = *v, *ko = c->s_k2.as<uint64_t>(), *vo = c->s_v2.as<uint64_t>();
  for (int shift = 0; shift < nbits; shift += bits) {
    int rc = OIC_OK;
    switch (bits) {
      case 8: rc = sort_pass<8>(c, ki, vi, ko, vo, n, shift); break;
      case 9: rc = sort_pass<9>(c, ki, vi, ko, vo, n, shift); break;
      case 10: rc = sort_pass<10>(c, ki, vi, ko, vo, n, shift); break;
      default: rc = sort_pass<11>(c, ki, vi, ko, vo, n, shift); break;
    }
    if (rc) return rc;
    std::swap(ki, ko);
    std::swap(vi, vo);
  }
  *k = ki;
  *v = vi;
  return OIC_OK;
}

// builds the k-mer table over hash reads [bgn, end]; fills first_t (distinct k-mers first seen per read)
static int build_table(oic_ctx *c, uint32_t bgn, uint32_t end, std::vector<uint32_t> *first_t) {
  const int K = (int)c->P.kmer;
  const uint32_t nb = end - bgn + 1;
  std::vector<uint8_t> loaded(nb);
  std::vector<uint64_t> posbase(nb + 1);
  uint64_t N = 0;
  for (uint32_t t = 0; t < nb; t++) {
    const int32_t L = c->h_len[bgn + t];
    loaded[t] = (L >= c->P.min_olap && L >= K) ? 1 : 0;
    posbase[t] = N;
    if (loaded[t]) N += (uint64_t)(L - K + 1);
  }
  posbase[nb] = N;
  c->npos = N;
  DevBuf d_loaded, d_posbase, d_first;
  ALLOC(c, d_loaded, nb);
  ALLOC(c, d_posbase, (nb + 1) * 8);
  ALLOC(c, d_first, nb * 4);
  HIPCHK(c, hipMemcpyAsync(d_loaded.p, loaded.data(), nb, hipMemcpyHostToDevice, c->st));
  HIPCHK(c, hipMemcpyAsync(d_posbase.p, posbase.data(), (nb + 1) * 8, hipMemcpyHostToDevice, c->st));
  HIPCHK(c, hipMemsetAsync(d_first.p, 0, nb * 4, c->st));
  ALLOC(c, c->s_k1, std::max<uint64_t>(N, 1) * 8);
  ALLOC(c, c->x_occ, std::max<uint64_t>(N, 1) * 8);
  if (ensure_sort_temp(c, std::max<uint64_t>(N, 1)) != OIC_OK) return OIC_ENOMEM;
  const uint64_t invalid = 1ull << (2 * K);
  ReadsDev R = dev_reads(c);
  uint64_t *keys = c->s_k1.as<uint64_t>(), *vals = c->x_occ.as<uint64_t>();
  if (N) {
    hipLaunchKernelGGL(k_kmer_emit, dim3(nb), dim3(256), 0, c->st, R, bgn, nb, d_posbase.as<uint64_t>(), d_loaded.as<uint8_t>(), K,
                       invalid, keys, vals);
    const int rc = sort_pairs(c, &keys, &vals, N, 2 * K + 1);   // stable: positions stay in (read, offset) order
    if (rc) return rc;
    // the sorted arrays may have moved into the temp buffers: swap ownership so x_occ holds vals
    if (keys != c->s_k1.as<uint64_t>()) c->s_k1.swap(c->s_k2);
    if (vals != c->x_occ.as<uint64_t>()) c->x_occ.swap(c->s_v2);
    keys = c->s_k1.as<uint64_t>();
    vals = c->x_occ.as<uint64_t>();
  }
  ALLOC(c, c->s_flag, std::max<uint64_t>(N, 1) * 4);
  ALLOC(c, c->s_idx, std::max<uint64_t>(N, 1) * 8);
  uint64_t nu = 0;
  if (N) {
    const unsigned g = (unsigned)((N + 255) / 256);
    hipLaunchKernelGGL(k_unique_flags, dim3(g), dim3(256), 0, c->st, keys, N, invalid, c->s_flag.as<uint32_t>());
    HIPCHK(c, scan_excl<uint32_t>(c->s_flag.as<uint32_t>(), c->s_idx.as<uint64_t>(), N, c->s_scan.as<uint64_t>(), c->st, &nu));
  }
  ALLOC(c, c->x_ukey, std::max<uint64_t>(nu, 1) * 8);
  ALLOC(c, c->x_ustart, (nu + 1) * 8);
  // number of valid positions: invalid keys sort last
  uint64_t nvalid = N;
  if (N) {
    std::vector<uint64_t> tail(1);
    // binary search on host over device keys would need many copies; count invalid windows instead
    uint64_t ninv = 0;
    for (uint32_t t = 0; t < nb; t++) (void)t;
    // invalid windows only exist in reads with N; find the first invalid key by bisection on device data
    uint64_t lo = 0, hi = N;
    while (lo < hi) {
      const uint64_t mid = (lo + hi) / 2;
      uint64_t kv = 0;
      HIPCHK(c, hipMemcpyAsync(&kv, keys + mid, 8, hipMemcpyDeviceToHost, c->st));
      HIPCHK(c, hipStreamSynchronize(c->st));
      if (kv < invalid) lo = mid + 1; else hi = mid;
    }
    nvalid = lo;
    (void)ninv;
  }
  if (nu) {
    const unsigned g = (unsigned)((N + 255) / 256);
    hipLaunchKernelGGL(k_unique_write, dim3(g), dim3(256), 0, c->st, keys, vals, N, c->s_flag.as<uint32_t>(), c->s_idx.as<uint64_t>(),
                       c->x_ukey.as<uint64_t>(), c->x_ustart.as<uint64_t>(), d_first.as<uint32_t>());
  }
  HIPCHK(c, hipMemcpyAsync(c->x_ustart.as<uint64_t>() + nu, &nvalid, 8, hipMemcpyHostToDevice, c->st));
  int dbits = std::max(8, std::min(ceil_log2(std::max<uint64_t>(nu, 1)), std::min(2 * K, 28)));
  if (dbits > 2 * K) dbits = 2 * K;
  c->dshift = 2 * K - dbits;
  const uint64_t nbk = 1ull << dbits;
  ALLOC(c, c->x_dir, (nbk + 2) * 8);
  {
    const unsigned g = (unsigned)((nu + 1 + 255) / 256);
    hipLaunchKernelGGL(k_build_dir, dim3(g), dim3(256), 0, c->st, c->x_ukey.as<uint64_t>(), nu, c->dshift, nbk, c->x_dir.as<uint64_t>());
  }
  HIPCHK(c, hipGetLastError());
  if (first_t) {
    first_t->resize(nb);
    HIPCHK(c, hipMemcpyAsync(first_t->data(), d_first.p, nb * 4, hipMemcpyDeviceToHost, c->st));
  }
  if (!c->skip.empty()) {
    ALLOC(c, c->x_tscreen, (uint64_t)nb * 4);
    HIPCHK(c, hipMemsetAsync(c->x_tscreen.p, 0, (uint64_t)nb * 4, c->st));
    KmerTable X;
    X.ukey = c->x_ukey.as<uint64_t>();
    X.ustart = c->x_ustart.as<uint64_t>();
    X.occ = c->x_occ.as<uint64_t>();
    X.dir = c->x_dir.as<uint64_t>();
    X.nu = nu;
    X.shift = c->dshift;
    const uint32_t ns = (uint32_t)c->skip.size();
    hipLaunchKernelGGL(k_mark_screens, dim3((ns + 255) / 256), dim3(256), 0, c->st, c->d_skip.as<uint64_t>(), ns, X,
                       c->d_len.as<int32_t>(), bgn, K, c->x_tscreen.as<uint32_t>());
    HIPCHK(c, hipGetLastError());
  }
  HIPCHK(c, hipStreamSynchronize(c->st));
  c->nu = nu;
  c->nvalid = nvalid;
  c->hbgn = bgn;
  c->hend = end;
  c->nblock = nb;
  c->have_index = true;
  return OIC_OK;
}

extern "C" int oic_index_build(oic_ctx *c, uint32_t bgn, uint32_t end, uint32_t *last_loaded) {
  if (!c) return OIC_EINVAL;
  if (c->nreads == 0) return fail(c, OIC_ESTATE, "no reads loaded");
  if (bgn < 1) bgn = 1;
  if (end > c->nreads) end = c->nreads;
  if (bgn > end) return fail(c, OIC_EINVAL, "empty hash range %u-%u", bgn, end);
  const double t0 = now_ms();
  // Build_Hash_Index.C:504-506: a read is loaded while total_len < Max_Hash_Data_Len and
  // Hash_Entries (distinct k-mers so far) < Max_Hash_Load * 2^hashbits * 21.
  const uint64_t maxData = c->P.hash_data_len + oic::kMaxReadLen;
  const uint64_t limit = (uint64_t)(c->P.hash_load * (double)(1ull << c->P.hash_bits) * 21.0);
  uint64_t total = 0;
  uint32_t cand = bgn - 1;
  for (uint32_t r = bgn; r <= end; r++) {
    if (total >= maxData) break;
    cand = r;
    if (c->h_len[r] >= c->P.min_olap) total += (uint64_t)c->h_len[r] + 1;
  }
  std::vector<uint32_t> first;
  int rc = build_table(c, bgn, cand, &first);
  if (rc) return rc;
  uint64_t distinct = 0;
  uint32_t blk = cand;
  for (uint32_t r = bgn; r <= cand; r++) {
    if (distinct >= limit) { blk = r - 1; break; }
    distinct += first[r - bgn];
  }
  if (blk < cand) {
    rc = build_table(c, bgn, blk, nullptr);
    if (rc) return rc;
  }
  if (last_loaded) *last_loaded = blk;
  c->ctr.ms_index += now_ms() - t0;
  return OIC_OK;
}

// ------------------------------------------------------------------------------------ search

static int ensure_wave_scratch(oic_ctx *c, uint32_t waves, int cap, int64_t val_cap) {
  const int32_t rows = c->T.error_bound[std::min<int32_t>(c->maxlen, (int32_t)oic::kMaxReadLen)] + 4;
  ALLOC(c, c->sc_rowL, (size_t)waves * rows * 4);
  ALLOC(c, c->sc_rowR, (size_t)waves * rows * 4);
  ALLOC(c, c->sc_rowOff, (size_t)waves * rows * 8);
  ALLOC(c, c->sc_rowVal, (size_t)waves * val_cap * 4);
  ALLOC(c, c->sc_dbuf, (size_t)waves * (rows + 4) * 4);
  ALLOC(c, c->sc_ld, (size_t)waves * (rows + 4) * 4);
  ALLOC(c, c->sc_rd, (size_t)waves * (rows + 4) * 4);
  c->sc_rows = rows;
  c->sc_val = val_cap;
  c->sc_waves = waves;
  (void)cap;
  return OIC_OK;
}

static DPScratch scratch0(oic_ctx *c) {
  DPScratch s;
  s.rowL = c->sc_rowL.as<int32_t>();
  s.rowR = c->sc_rowR.as<int32_t>();
  s.rowOff = c->sc_rowOff.as<int64_t>();
  s.rowVal = c->sc_rowVal.as<int32_t>();
  s.val_cap = c->sc_val;
  s.max_rows = c->sc_rows - 2;
  s.dbuf = c->sc_dbuf.as<int32_t>();
  s.ldelta = c->sc_ld.as<int32_t>();
  s.rdelta = c->sc_rd.as<int32_t>();
  return s;
}

static SearchConst search_const(oic_ctx *c, bool elide) {
  SearchConst C;
  C.tb = dev_tables(c);
  C.min_olap = c->P.min_olap;
  C.unique = c->P.unique;
  C.partial = c->P.partial;
  C.hopeless = c->P.hopeless;
  C.K = (int32_t)c->P.kmer;
  C.min_kmers = c->P.minkmers ? c->d_minkmers.as<uint32_t>() : nullptr;
  C.elide = elide ? 1 : 0;
  C.sscreen = nullptr;
  C.tscreen = c->skip.empty() ? nullptr : c->x_tscreen.as<uint32_t>();
  C.s0 = 0;
  C.hbgn = c->hbgn;
  return C;
}

// runs k_process_pairs over order[0..n) with the given tile; returns the number of overflowed pairs
template <int CAP>
static int run_pairs(oic_ctx *c, PairKernelArgs A, uint64_t n, uint32_t waves, int64_t val_cap, double *ms) {
  if (n == 0) return OIC_OK;
  waves = (uint32_t)std::min<uint64_t>(waves, n);
  int rc = ensure_wave_scratch(c, waves, CAP, val_cap);
  if (rc) return rc;
  A.sc0 = scratch0(c);
  A.row_stride = c->sc_rows;
  A.val_stride = c->sc_val;
  A.delta_stride = c->sc_rows + 4;
  A.norder = n;
  HIPCHK(c, hipMemsetAsync(A.queue, 0, 4, c->st));
  HIPCHK(c, hipEventRecord(c->ev0, c->st));
  hipLaunchKernelGGL(k_process_pairs<CAP>, dim3(waves), dim3(OIC_WAVE), 0, c->st, A);
  HIPCHK(c, hipGetLastError());
  c->ctr.pp_launches++;
  HIPCHK(c, hipEventRecord(c->ev1, c->st));
  HIPCHK(c, hipEventSynchronize(c->ev1));
  float f = 0;
  hipEventElapsedTime(&f, c->ev0, c->ev1);
  *ms += f;
  return OIC_OK;
}

// which ref reads a -t T run searches: OverlapDriver hands out chunks of perThread reads from
// bgnRefID and a chunk starting exactly at endRefID is never run (overlapInCore.C:225,
// overlapInCore-Process_Overlaps.C:33)
static uint32_t ref_end_searched(uint32_t bgn, uint32_t end, uint32_t nthreads) {
  if (end < bgn) return end;
  const uint32_t per = 1 + (end - bgn) / nthreads / 8;
  return ((end - bgn) % per == 0) ? end - 1 : end;
}

// model prediction (log1p cells) for every pair of the batch, on the device
static int predict_pairs(oic_ctx *c, uint64_t np, std::vector<float> &pred) {
  ALLOC(c, c->m_feat, np * MLP_IN * 4);
  ALLOC(c, c->m_pred, np * MLP_OUT * 4);
  const unsigned g = (unsigned)((np + 255) / 256);
  hipLaunchKernelGGL(k_pair_features, dim3(g), dim3(256), 0, c->st, c->p_pairs.as<PairRec>(), np, c->d_len.as<int32_t>(),
                     c->m_feat.as<float>());
  hipLaunchKernelGGL(k_mlp_forward, dim3(g), dim3(256), 0, c->st, c->m_feat.as<float>(), np, c->m_param.as<float>(),
                     c->m_mu.as<float>(), c->m_sd.as<float>(), c->m_pred.as<float>());
  HIPCHK(c, hipGetLastError());
  std::vector<float> y(np * MLP_OUT);
  HIPCHK(c, hipMemcpyAsync(y.data(), c->m_pred.p, np * MLP_OUT * 4, hipMemcpyDeviceToHost, c->st));
  HIPCHK(c, hipStreamSynchronize(c->st));
  pred.resize(np);
  for (uint64_t p = 0; p < np; p++) pred[p] = y[p * MLP_OUT];
  return OIC_OK;
}

// the cost model's per-pair predictions (on the device) when they order this batch, else null
static const float *spec_pred(oic_ctx *c) { return c->batch_pred ? c->m_pred.as<float>() : nullptr; }

// shapes that only the flat-row form runs (4 lanes per DP with a tile of 64 cells or fewer)
template <int G, int CAP>
constexpr bool flat_only() { return G == 4 && CAP <= 64; }
template <int G, int CAP>
constexpr bool has_flat() { return (G == 4 || G == 8) && CAP <= 128; }

// the packed kernel in the loop form the tuning selects (one DP per wavefront: the per-task loop)
template <int G, int CAP, bool STATS>
static void launch_fwd(oic_ctx *c, const FwdArgs &FA, uint32_t waves) {
  if constexpr (has_flat<G, CAP>()) {
    if (c->tune.fwd_loop == 2 || flat_only<G, CAP>()) {
      hipLaunchKernelGGL((k_fwd_flat<G, CAP, STATS>), dim3(waves), dim3(OIC_WAVE), 0, c->st, FA);
      return;
    }
  }
  if constexpr (flat_only<G, CAP>()) {
    return;
  } else if constexpr (G == 1) {
    hipLaunchKernelGGL((k_fwd_lane<CAP, STATS>), dim3(waves), dim3(OIC_WAVE), 0, c->st, FA);
    return;
  } else if constexpr (G < 64) {
    if (c->tune.fwd_loop == 0) {
      hipLaunchKernelGGL((k_fwd_sm<G, CAP, STATS>), dim3(waves), dim3(OIC_WAVE), 0, c->st, FA);
      return;
    }
  }
  if constexpr (!flat_only<G, CAP>())
    hipLaunchKernelGGL((k_fwd_tasks<G, CAP, STATS>), dim3(waves), dim3(OIC_WAVE), 0, c->st, FA);
}

// persistent wavefronts for a packed launch: as many as fit on every CU at once (registers and LDS
// decide), but no more than the tasks can use
template <int G, int CAP, bool STATS>
static uint32_t fwd_waves(oic_ctx *c, uint64_t max_tasks) {
  static int cached[3] = {0, 0, 0};   // per loop form; the same on every device of one architecture
  const int sm = (flat_only<G, CAP>() || (has_flat<G, CAP>() && c->tune.fwd_loop == 2)) ? 2
                 : (G < 64 && c->tune.fwd_loop == 0) ? 1 : 0;
  int per_cu = cached[sm];
  if (per_cu < 1) {
    hipError_t e = hipSuccess;
    if constexpr (flat_only<G, CAP>()) {
      e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_fwd_flat<G, CAP, STATS>, OIC_WAVE, 0);
    } else if constexpr (G == 1) {
      e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_fwd_lane<CAP, STATS>, OIC_WAVE, 0);
    } else {
      if constexpr (has_flat<G, CAP>()) {
        if (sm == 2) e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_fwd_flat<G, CAP, STATS>, OIC_WAVE, 0);
      }
      if constexpr (G < 64) {
        if (sm) e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_fwd_sm<G, CAP, STATS>, OIC_WAVE, 0);
      }
      if (sm == 0) e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_fwd_tasks<G, CAP, STATS>, OIC_WAVE, 0);
    }
    if (e != hipSuccess || per_cu < 1) { (void)hipGetLastError(); per_cu = 16; }
    cached[sm] = per_cu;
  }
  const uint64_t need = (max_tasks + (OIC_WAVE / G) - 1) / (OIC_WAVE / G);
  return (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>((uint64_t)c->ncu * (uint64_t)per_cu, need));
}

template <int G, int CAP, bool STATS>
static int run_fwd(oic_ctx *c, FwdArgs FA, uint64_t n, double *ms) {
  if (n == 0) return OIC_OK;
  FA.ntasks = n;
  HIPCHK(c, hipMemsetAsync(FA.queue, 0, 4, c->st));
  HIPCHK(c, hipMemsetAsync(FA.stats, 0, 64, c->st));
  const uint32_t waves = fwd_waves<G, CAP, STATS>(c, n);
  HIPCHK(c, hipEventRecord(c->ev0, c->st));
  launch_fwd<G, CAP, STATS>(c, FA, waves);
  HIPCHK(c, hipGetLastError());
  HIPCHK(c, hipEventRecord(c->ev1, c->st));
  HIPCHK(c, hipEventSynchronize(c->ev1));
  float f = 0;
  hipEventElapsedTime(&f, c->ev0, c->ev1);
  *ms += f;
  unsigned long long st[8];
  HIPCHK(c, hipMemcpy(st, FA.stats, sizeof st, hipMemcpyDeviceToHost));
  c->ctr.spec_cells += st[0];
  c->ctr.spec_rows += st[1];
  c->ctr.spec_lane_slots += st[2];
  c->ctr.spec_wave_rows += st[3];
  c->ctr.spec_bases += st[4];
  c->ctr.spec_slide_iters += st[5];
  c->ctr.spec_wave_slide_iters += st[6];
  c->ctr.spec_wave_chunks += st[7];
  if (G < 64) {   // the packed narrow-band kernel
    c->ctr.fwd_launches++;
    c->ctr.fwd_cells += st[0];
    c->ctr.fwd_bases += st[4];
    c->ctr.ms_fwd_narrow += f;
  }
  return OIC_OK;
}

// packed narrow-band pass, then the tasks whose band outgrew its tile, one per wavefront
static int run_fwd_narrow(oic_ctx *c, const FwdArgs &FA, uint64_t n, double *ms) {
  const uint32_t g = c->tune.fwd_lanes, t = c->tune.fwd_tile;
#define FWD_CASE(G_, T_) \
  if (g == G_ && t == T_) return c->tune.fwd_counters ? run_fwd<G_, T_, true>(c, FA, n, ms) : run_fwd<G_, T_, false>(c, FA, n, ms);
  FWD_CASE(1, 16) FWD_CASE(1, 32) FWD_CASE(1, 64)
  FWD_CASE(4, 32) FWD_CASE(4, 64) FWD_CASE(4, 128) FWD_CASE(4, 256)
  FWD_CASE(8, 32) FWD_CASE(8, 64) FWD_CASE(8, 128) FWD_CASE(8, 256) FWD_CASE(8, 512)
  FWD_CASE(16, 32) FWD_CASE(16, 64) FWD_CASE(16, 128) FWD_CASE(16, 256) FWD_CASE(16, 512)
  FWD_CASE(32, 128) FWD_CASE(32, 256) FWD_CASE(32, 512)
#undef FWD_CASE
  return fail(c, OIC_EINVAL, "no forward-DP kernel for %u lanes x %u-cell tile", g, t);
}

static int run_fwd_binned(oic_ctx *c, FwdArgs FA, uint64_t n, double *ms) {
  HIPCHK(c, hipMemsetAsync(FA.novf, 0, 4, c->st));
  int rc = run_fwd_narrow(c, FA, n, ms);
  if (rc) return rc;
  uint32_t nov = 0;
  HIPCHK(c, hipMemcpy(&nov, FA.novf, 4, hipMemcpyDeviceToHost));
  if (nov == 0) return OIC_OK;
  c->ctr.overflow_retries += nov;
  // one per wavefront with a 512-cell tile, then what is wider still with an 8192-cell tile
  FA.tasks = FA.ovf;
  FA.ovf = FA.ovf + n;
  HIPCHK(c, hipMemsetAsync(FA.novf, 0, 4, c->st));
  rc = c->tune.fwd_counters ? run_fwd<64, 512, true>(c, FA, nov, ms) : run_fwd<64, 512, false>(c, FA, nov, ms);
  if (rc) return rc;
  HIPCHK(c, hipMemcpy(&nov, FA.novf, 4, hipMemcpyDeviceToHost));
  if (nov == 0) return OIC_OK;
  FA.tasks = FA.ovf;
  FA.ovf = nullptr;
  return c->tune.fwd_counters ? run_fwd<64, 8192, true>(c, FA, nov, ms) : run_fwd<64, 8192, false>(c, FA, nov, ms);
}

// probe pass over every pair's longest node, then every other node of the pairs whose probe
// ended at a branch point; results land in the per-node outcome words
static int run_forward_passes(oic_ctx *c, uint64_t np, uint64_t H) {
  ALLOC(c, c->b_outc, H * 8);
  ALLOC(c, c->b_probe, np * sizeof(FwdTask));
  ALLOC(c, c->b_cnt, np * 4);
  ALLOC(c, c->b_off, np * 8);
  ALLOC(c, c->b_stats, 64);
  HIPCHK(c, hipMemsetAsync(c->b_outc.p, 0, H * 8, c->st));
  const unsigned g = (unsigned)((np + 255) / 256);
  hipLaunchKernelGGL(k_probe_tasks, dim3(g), dim3(256), 0, c->st, c->p_pairs.as<PairRec>(), np, c->b_probe.as<FwdTask>());
  FwdArgs FA;
  FA.R = dev_reads(c);
  FA.tb = dev_tables(c);
  FA.pairs = c->p_pairs.as<PairRec>();
  FA.nodes = c->p_nodes.as<int4>();
  FA.queue = c->p_queue.as<uint32_t>();
  FA.outc = c->b_outc.as<uint64_t>();
  FA.stats = c->b_stats.as<unsigned long long>();
  FA.ntasks_dev = nullptr;
  FA.tstamp = nullptr;
  double ms = 0;
  ALLOC(c, c->b_ovf, 2 * np * sizeof(FwdTask));
  ALLOC(c, c->b_novf, 16);
  FA.ovf = c->b_ovf.as<FwdTask>();
  FA.novf = c->b_novf.as<uint32_t>();
  FA.tasks = c->b_probe.as<FwdTask>();
  int rc = run_fwd_binned(c, FA, np, &ms);
  if (rc) return rc;
  hipLaunchKernelGGL(k_spec_count, dim3(g), dim3(256), 0, c->st, c->p_pairs.as<PairRec>(), np, c->b_probe.as<FwdTask>(),
                     c->b_outc.as<uint64_t>(), spec_pred(c), c->p_order.as<uint32_t>(),
                     c->b_cnt.as<uint32_t>());
  uint64_t nt = 0;
  ALLOC(c, c->s_scan, scan_scratch_words(np) * 8);
  HIPCHK(c, scan_excl<uint32_t>(c->b_cnt.as<uint32_t>(), c->b_off.as<uint64_t>(), np, c->s_scan.as<uint64_t>(), c->st, &nt));
  if (nt) {
    ALLOC(c, c->b_spec, nt * sizeof(FwdTask));
    hipLaunchKernelGGL(k_spec_write, dim3(g), dim3(256), 0, c->st, c->p_pairs.as<PairRec>(), np, c->b_probe.as<FwdTask>(),
                       c->p_order.as<uint32_t>(), c->b_cnt.as<uint32_t>(), c->b_off.as<uint64_t>(), c->b_spec.as<FwdTask>());
    FA.tasks = c->b_spec.as<FwdTask>();
    if (c->tune.fwd_order == 1) {
      // the hit buffers are free once the match lists exist
      if (ensure_sort_temp(c, nt) != OIC_OK) return OIC_ENOMEM;
      ALLOC(c, c->s_k1, nt * 8);
      ALLOC(c, c->s_v1, nt * 8);
      uint64_t *tk = c->s_k1.as<uint64_t>(), *tv = c->s_v1.as<uint64_t>();
      hipLaunchKernelGGL(k_task_keys, dim3((unsigned)((nt + 255) / 256)), dim3(256), 0, c->st, c->b_spec.as<FwdTask>(), nt,
                         c->p_pairs.as<PairRec>(), c->p_nodes.as<int4>(), tk, tv);
      const int src = sort_pairs(c, &tk, &tv, nt, 16);
      if (src) return src;
      FA.tasks = (const FwdTask *)tv;
    }
    ALLOC(c, c->b_ovf, 2 * nt * sizeof(FwdTask));
    FA.ovf = c->b_ovf.as<FwdTask>();
    rc = run_fwd_binned(c, FA, nt, &ms);
    if (rc) return rc;
  }
  c->ctr.spec_tasks += np + nt;
  c->ctr.ms_fwd_kernel += ms;
  c->ctr.ms_dp_kernel += ms;
  return OIC_OK;
}

// ------------------------------------------------------------ forward passes as one hipGraph

// e0 / e1 (recorded only outside graph capture) time the launch
template <int G, int CAP, bool STATS>
static void enqueue_fwd(oic_ctx *c, const FwdArgs &FA, uint64_t max_tasks, hipEvent_t e0, hipEvent_t e1) {
  const uint32_t waves = fwd_waves<G, CAP, STATS>(c, max_tasks);
  if (e0) hipEventRecord(e0, c->st);
  launch_fwd<G, CAP, STATS>(c, FA, waves);
  if (e1) hipEventRecord(e1, c->st);
}

static int enqueue_fwd_narrow(oic_ctx *c, const FwdArgs &FA, uint64_t max_tasks, hipEvent_t e0, hipEvent_t e1) {
  const uint32_t g = c->tune.fwd_lanes, t = c->tune.fwd_tile;
#define FWD_CASE(G_, T_)                                                                                   \
  if (g == G_ && t == T_) {                                                                                \
    if (c->tune.fwd_counters) enqueue_fwd<G_, T_, true>(c, FA, max_tasks, e0, e1);                         \
    else enqueue_fwd<G_, T_, false>(c, FA, max_tasks, e0, e1);                                             \
    return OIC_OK;                                                                                         \
  }
  FWD_CASE(1, 16) FWD_CASE(1, 32) FWD_CASE(1, 64)
  FWD_CASE(4, 32) FWD_CASE(4, 64) FWD_CASE(4, 128) FWD_CASE(4, 256)
  FWD_CASE(8, 32) FWD_CASE(8, 64) FWD_CASE(8, 128) FWD_CASE(8, 256) FWD_CASE(8, 512)
  FWD_CASE(16, 32) FWD_CASE(16, 64) FWD_CASE(16, 128) FWD_CASE(16, 256) FWD_CASE(16, 512)
  FWD_CASE(32, 128) FWD_CASE(32, 256) FWD_CASE(32, 512)
#undef FWD_CASE
  return fail(c, OIC_EINVAL, "no forward-DP kernel for %u lanes x %u-cell tile", g, t);
}

// The probe and speculation passes of one search batch, captured into a hipGraph and launched once:
// task counts stay on the device (the spec count from the scan, overflow counts from each level),
// buffers are sized by the batch's node count, and counters are cleared by a fill kernel rather
// than memset nodes.  Each pass runs in three levels: the packed narrow kernel; the tasks whose
// band outgrew its tile, one per wavefront with a 512-cell tile; those wider still with an
// 8192-cell tile.  Results are the same as run_forward_passes'.
static int run_forward_passes_graph(oic_ctx *c, uint64_t np, uint64_t H, uint64_t nnodes) {
  ALLOC(c, c->b_outc, std::max<uint64_t>(H, 1) * 8);
  ALLOC(c, c->b_probe, np * sizeof(FwdTask));
  ALLOC(c, c->b_cnt, np * 4);
  ALLOC(c, c->b_off, np * 8);
  ALLOC(c, c->b_spec, std::max<uint64_t>(nnodes, 1) * sizeof(FwdTask));
  ALLOC(c, c->b_ovf, 2 * (np + nnodes) * sizeof(FwdTask));
  ALLOC(c, c->b_stats, 6 * 64 + 32);
  ALLOC(c, c->b_novf, 64);
  ALLOC(c, c->s_scan, scan_scratch_words(np) * 8);
  // cnt: [0..5] queues of the six launches, [6] spec tasks, [8..11] overflow counts
  uint32_t *cnt = c->b_novf.as<uint32_t>();
  unsigned long long *stats = c->b_stats.as<unsigned long long>();
  const unsigned g = (unsigned)((np + 255) / 256);
  const bool graph = c->tune.fwd_graph != 0;
  hipEvent_t *ev = graph ? nullptr : c->evf;
  if (graph) HIPCHK(c, hipStreamBeginCapture(c->st, hipStreamCaptureModeThreadLocal));
  // everything up to the end of the capture; an error inside it still ends the capture below, so
  // the stream never stays in capture mode
  auto enqueue_all = [&]() -> int {
    const uint64_t cw = 2 * H;
    hipLaunchKernelGGL(k_fill_u32, dim3((unsigned)((cw + 255) / 256)), dim3(256), 0, c->st, (uint32_t *)c->b_outc.p, cw, 0u);
    hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(256), 0, c->st, cnt, (uint64_t)16, 0u);
    unsigned long long *tst = stats + 48;      // wall-clock [start, end] of the two packed launches
    hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(256), 0, c->st, (uint32_t *)stats, (uint64_t)(6 * 16), 0u);
    hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(256), 0, c->st, (uint32_t *)tst, (uint64_t)8, 0u);
    hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, c->st, (uint32_t *)tst, (uint64_t)2, 0xffffffffu);
    hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, c->st, (uint32_t *)(tst + 2), (uint64_t)2, 0xffffffffu);
    hipLaunchKernelGGL(k_probe_tasks, dim3(g), dim3(256), 0, c->st, c->p_pairs.as<PairRec>(), np, c->b_probe.as<FwdTask>());
    FwdArgs FA;
    FA.R = dev_reads(c);
    FA.tb = dev_tables(c);
    FA.pairs = c->p_pairs.as<PairRec>();
    FA.nodes = c->p_nodes.as<int4>();
    FA.outc = c->b_outc.as<uint64_t>();
    FwdTask *ovf = c->b_ovf.as<FwdTask>();
    // one pass: launch slot k (0 probe, 1 spec) uses queues/stats 3k..3k+2 and overflow counts 8+2k, 9+2k
    auto pass = [&](int k, const FwdTask *tasks, uint64_t n, const uint32_t *n_dev, uint64_t cap, FwdTask *ov1, FwdTask *ov2) -> int {
      FA.tasks = tasks; FA.ntasks = n; FA.ntasks_dev = n_dev;
      FA.queue = cnt + 3 * k; FA.stats = stats + 8 * (3 * k); FA.ovf = ov1; FA.novf = cnt + 8 + 2 * k;
      FA.tstamp = tst + 2 * k;
      int rc = enqueue_fwd_narrow(c, FA, cap, ev ? ev[6 * k] : nullptr, ev ? ev[6 * k + 1] : nullptr);
      if (rc) return rc;
      FA.tstamp = nullptr;
      FA.tasks = ov1; FA.ntasks = 0; FA.ntasks_dev = cnt + 8 + 2 * k;
      FA.queue = cnt + 3 * k + 1; FA.stats = stats + 8 * (3 * k + 1); FA.ovf = ov2; FA.novf = cnt + 9 + 2 * k;
      if (c->tune.fwd_counters) enqueue_fwd<64, 512, true>(c, FA, cap, ev ? ev[6 * k + 2] : nullptr, ev ? ev[6 * k + 3] : nullptr);
      else enqueue_fwd<64, 512, false>(c, FA, cap, ev ? ev[6 * k + 2] : nullptr, ev ? ev[6 * k + 3] : nullptr);
      FA.tasks = ov2; FA.ntasks = 0; FA.ntasks_dev = cnt + 9 + 2 * k;
      FA.queue = cnt + 3 * k + 2; FA.stats = stats + 8 * (3 * k + 2); FA.ovf = nullptr; FA.novf = nullptr;
      if (c->tune.fwd_counters) enqueue_fwd<64, 8192, true>(c, FA, cap, ev ? ev[6 * k + 4] : nullptr, ev ? ev[6 * k + 5] : nullptr);
      else enqueue_fwd<64, 8192, false>(c, FA, cap, ev ? ev[6 * k + 4] : nullptr, ev ? ev[6 * k + 5] : nullptr);
      return OIC_OK;
    };
    int rc = pass(0, c->b_probe.as<FwdTask>(), np, nullptr, np, ovf, ovf + np);
    if (rc) return rc;
    hipLaunchKernelGGL(k_spec_count, dim3(g), dim3(256), 0, c->st, c->p_pairs.as<PairRec>(), np, c->b_probe.as<FwdTask>(),
                       c->b_outc.as<uint64_t>(), spec_pred(c), c->p_order.as<uint32_t>(),
                       c->b_cnt.as<uint32_t>());
    HIPCHK(c, scan_excl<uint32_t>(c->b_cnt.as<uint32_t>(), c->b_off.as<uint64_t>(), np, c->s_scan.as<uint64_t>(), c->st, nullptr));
    hipLaunchKernelGGL(k_scan_total, dim3(1), dim3(64), 0, c->st, c->b_cnt.as<uint32_t>(), c->b_off.as<uint64_t>(), np, cnt + 6);
    hipLaunchKernelGGL(k_spec_write, dim3(g), dim3(256), 0, c->st, c->p_pairs.as<PairRec>(), np, c->b_probe.as<FwdTask>(),
                       c->p_order.as<uint32_t>(), c->b_cnt.as<uint32_t>(), c->b_off.as<uint64_t>(), c->b_spec.as<FwdTask>());
    rc = pass(1, c->b_spec.as<FwdTask>(), 0, cnt + 6, nnodes, ovf + 2 * np, ovf + 2 * np + nnodes);
    if (rc) return rc;
    HIPCHK(c, hipGetLastError());
    return OIC_OK;
  };
  const int erc = enqueue_all();
  if (graph) {
    hipGraph_t gr = nullptr;
    const hipError_t ce = hipStreamEndCapture(c->st, &gr);
    if (erc != OIC_OK || ce != hipSuccess) {
      if (gr) hipGraphDestroy(gr);
      (void)hipGetLastError();
      return erc != OIC_OK ? erc : fail(c, OIC_EHIP, "hipStreamEndCapture: %s", hipGetErrorString(ce));
    }
    hipGraphExec_t ge = nullptr;
    const hipError_t ie = hipGraphInstantiate(&ge, gr, nullptr, nullptr, 0);
    if (ie != hipSuccess) { hipGraphDestroy(gr); return fail(c, OIC_EHIP, "hipGraphInstantiate: %s", hipGetErrorString(ie)); }
    hipEventRecord(c->ev0, c->st);
    const hipError_t le = hipGraphLaunch(ge, c->st);
    hipEventRecord(c->ev1, c->st);
    const hipError_t se = hipStreamSynchronize(c->st);
    hipGraphExecDestroy(ge);
    hipGraphDestroy(gr);
    if (le != hipSuccess || se != hipSuccess) return fail(c, OIC_EHIP, "forward-pass graph: %s", hipGetErrorString(le != hipSuccess ? le : se));
    c->ctr.graphs_launched++;
  } else {
    if (erc != OIC_OK) return erc;
    HIPCHK(c, hipStreamSynchronize(c->st));
  }
  unsigned long long st[52];
  uint32_t cn[16];
  HIPCHK(c, hipMemcpy(st, stats, sizeof st, hipMemcpyDeviceToHost));
  HIPCHK(c, hipMemcpy(cn, cnt, sizeof cn, hipMemcpyDeviceToHost));
  const uint64_t nt = cn[6];
  // launch times: per kernel when issued directly; in a graph, the whole graph counts, shared by
  // the two narrow launches in proportion to their cells (its fill, count and scan kernels and the
  // retries of the few wide bands take a few percent at most)
  float f[6] = {0, 0, 0, 0, 0, 0};
  if (graph) {
    // the graph's time (HIP events around its launch); the packed launches' own spans from the
    // wall-clock stamps of their first and last wavefronts; the rest of the graph is the retries
    // and the small fill / count / scan kernels
    float gms = 0;
    HIPCHK(c, hipEventElapsedTime(&gms, c->ev0, c->ev1));
    double narrow = 0;
    for (int k = 0; k < 2; k++) {
      const unsigned long long a = st[48 + 2 * k], b = st[49 + 2 * k];
      f[3 * k] = (b > a && c->wallclock_khz > 0) ? (float)((double)(b - a) / (double)c->wallclock_khz) : 0.0f;
      narrow += f[3 * k];
    }
    f[1] = (float)std::max(0.0, (double)gms - narrow);
    c->ctr.ms_fwd_graph += gms;
  } else {
    for (int i = 0; i < 6; i++) HIPCHK(c, hipEventElapsedTime(&f[i], c->evf[2 * i], c->evf[2 * i + 1]));
  }
  for (int i = 0; i < 6; i++) {
    const unsigned long long *x = st + 8 * i;
    c->ctr.spec_cells += x[0];
    c->ctr.spec_rows += x[1];
    c->ctr.spec_lane_slots += x[2];
    c->ctr.spec_wave_rows += x[3];
    c->ctr.spec_bases += x[4];
    c->ctr.spec_slide_iters += x[5];
    c->ctr.spec_wave_slide_iters += x[6];
    c->ctr.spec_wave_chunks += x[7];
  }
  for (int i : {0, 3}) {   // the packed narrow-band launches
    if (i == 3 && nt == 0) continue;
    c->ctr.fwd_launches++;
    c->ctr.fwd_cells += st[8 * i];
    c->ctr.fwd_bases += st[8 * i + 4];
    c->ctr.ms_fwd_narrow += f[i];
  }
  c->ctr.overflow_retries += cn[8] + cn[10];
  c->ctr.spec_tasks += np + nt;
  double ms = 0;
  for (float x : f) ms += x;
  c->ctr.ms_fwd_kernel += ms;
  c->ctr.ms_dp_kernel += ms;
  return OIC_OK;
}

static int search_batch(oic_ctx *c, uint32_t s0, uint32_t ns, const std::vector<uint8_t> &searched, oic_sink *sink) {
  const int K = (int)c->P.kmer;
  const double t0 = now_ms();
  std::vector<uint64_t> pbase(2 * (size_t)ns + 1);
  uint64_t P = 0;
  for (uint32_t i = 0; i < ns; i++) {
    const uint64_t np = searched[i] ? (uint64_t)(c->h_len[s0 + i] - K + 1) : 0;
    pbase[2 * i] = P; P += np;
    pbase[2 * i + 1] = P; P += np;
  }
  pbase[2 * (size_t)ns] = P;
  if (P == 0) return OIC_OK;
  ALLOC(c, c->g_pbase, pbase.size() * 8);
  ALLOC(c, c->g_searched, ns);
  ALLOC(c, c->g_cnt, P * 4);
  ALLOC(c, c->g_off, P * 8);
  HIPCHK(c, hipMemcpyAsync(c->g_pbase.p, pbase.data(), pbase.size() * 8, hipMemcpyHostToDevice, c->st));
  HIPCHK(c, hipMemcpyAsync(c->g_searched.p, searched.data(), ns, hipMemcpyHostToDevice, c->st));
  GatherArgs GA;
  GA.R = dev_reads(c);
  GA.X.ukey = c->x_ukey.as<uint64_t>();
  GA.X.ustart = c->x_ustart.as<uint64_t>();
  GA.X.occ = c->x_occ.as<uint64_t>();
  GA.X.dir = c->x_dir.as<uint64_t>();
  GA.X.nu = c->nu;
  GA.X.shift = c->dshift;
  GA.s0 = s0;
  GA.ns = ns;
  GA.hbgn = c->hbgn;
  GA.k = K;
  GA.min_olap = c->P.min_olap;
  GA.searched = c->g_searched.as<uint8_t>();
  GA.tbits = std::max(1, ceil_log2(c->nblock));
  GA.skip = c->skip.empty() ? nullptr : c->d_skip.as<uint64_t>();
  GA.nskip = (uint32_t)c->skip.size();
  ALLOC(c, c->g_sscreen, 2 * (size_t)ns * 4);
  GA.sscreen = c->g_sscreen.as<uint32_t>();
  const int sbits = std::max(1, ceil_log2(2ull * ns));
  if (GA.tbits + sbits > 64) return fail(c, OIC_EINVAL, "pair key needs %d bits", GA.tbits + sbits);
  if (c->nu == 0) return OIC_OK;
  ALLOC(c, c->g_sdsum, 2 * (size_t)ns * 8);
  ALLOC(c, c->g_phi, P * 8);
  ALLOC(c, c->g_pval, P * 8);
  hipLaunchKernelGGL(k_hit_count, dim3(2 * ns), dim3(256), 0, c->st, GA, c->g_cnt.as<uint32_t>(), c->g_pbase.as<uint64_t>(),
                     c->g_sdsum.as<unsigned long long>(), c->g_phi.as<uint64_t>(), c->g_pval.as<uint64_t>());
  HIPCHK(c, hipGetLastError());
  ALLOC(c, c->s_scan, scan_scratch_words(P) * 8);
  uint64_t H = 0;
  HIPCHK(c, scan_excl<uint32_t>(c->g_cnt.as<uint32_t>(), c->g_off.as<uint64_t>(), P, c->s_scan.as<uint64_t>(), c->st, &H));
  const uint64_t hit_cap = c->hit_cap;
  if (H > hit_cap && ns > 1) {
    // too many seed hits for one pass (repeat reads): split at read boundaries
    std::vector<uint64_t> sds(2 * (size_t)ns);
    HIPCHK(c, hipMemcpy(sds.data(), c->g_sdsum.p, sds.size() * 8, hipMemcpyDeviceToHost));
    uint32_t a = 0;
    uint64_t acc = 0;
    for (uint32_t i = 0; i < ns; i++) {
      const uint64_t h = sds[2 * i] + sds[2 * i + 1];
      if (i > a && acc + h > hit_cap) {
        std::vector<uint8_t> sub(searched.begin() + a, searched.begin() + i);
        int rc = search_batch(c, s0 + a, i - a, sub, sink);
        if (rc) return rc;
        a = i;
        acc = 0;
      }
      acc += h;
    }
    std::vector<uint8_t> sub(searched.begin() + a, searched.end());
    return search_batch(c, s0 + a, ns - a, sub, sink);
  }
  c->ctr.kmer_probes += P;
  c->ctr.hits += H;
  if (H == 0) { c->ctr.ms_gather += now_ms() - t0; return OIC_OK; }
  // The per-hit buffers grow to at least twice the largest pass so far, capped at the pass limit:
  // a pass slightly larger than every earlier one then reuses them instead of freeing and
  // reallocating some 100 bytes per hit (tens of GB at the limit).
  if (H > c->h_alloc) c->h_alloc = std::max<uint64_t>(H, std::min<uint64_t>(2 * c->h_alloc, c->hit_cap));
  const uint64_t HA = c->h_alloc;
  ALLOC(c, c->s_k1, HA * 8);
  ALLOC(c, c->s_v1, HA * 8);
  if (ensure_sort_temp(c, HA) != OIC_OK) return OIC_ENOMEM;
  uint64_t *hk = c->s_k1.as<uint64_t>(), *hv = c->s_v1.as<uint64_t>();
  hipLaunchKernelGGL(k_hit_emit_tiles, dim3((unsigned)((H + EMIT_TILE - 1) / EMIT_TILE)), dim3(256), 0, c->st, c->g_off.as<uint64_t>(),
                     c->g_phi.as<uint64_t>(), c->g_pval.as<uint64_t>(), P, H, GA.X.occ, GA.tbits, hk, hv);
  HIPCHK(c, hipGetLastError());
  {
    // Hits arrive in (S read, orientation) order (windows laid out by pbase).  A stable sort on the
    // digits holding the T bits keeps equal sorted digits in that order, so each (S, orientation, T)
    // pair ends up contiguous, its hits in generation order, without sorting the S bits above them.
    const int rc = sort_pairs(c, &hk, &hv, H, GA.tbits);
    if (rc) return rc;
  }
  ALLOC(c, c->s_flag, HA * 4);
  ALLOC(c, c->s_idx, HA * 8);
  uint64_t np = 0;
  {
    const unsigned g = (unsigned)((H + 255) / 256);
    hipLaunchKernelGGL(k_pair_flags, dim3(g), dim3(256), 0, c->st, hk, H, c->s_flag.as<uint32_t>());
    HIPCHK(c, scan_excl<uint32_t>(c->s_flag.as<uint32_t>(), c->s_idx.as<uint64_t>(), H, c->s_scan.as<uint64_t>(), c->st, &np));
    ALLOC(c, c->p_pairs, np * sizeof(PairRec));
    hipLaunchKernelGGL(k_pair_write, dim3(g), dim3(256), 0, c->st, hk, H, c->s_flag.as<uint32_t>(), c->s_idx.as<uint64_t>(),
                       c->p_pairs.as<PairRec>(), s0, c->hbgn, GA.tbits);
    hipLaunchKernelGGL(k_pair_ends, dim3((unsigned)((np + 255) / 256)), dim3(256), 0, c->st, c->p_pairs.as<PairRec>(), np, H);
  }
  c->ctr.pairs += np;
  c->ctr.ms_gather += now_ms() - t0;
  const double t1 = now_ms();
  ALLOC(c, c->p_tmp, HA * 16);
  ALLOC(c, c->p_nodes, HA * 16);
  uint64_t batch_nodes = 0;
  ALLOC(c, c->l_scr, HA * (8 * 4 + 1));
  {
    ListScratch LS;
    int32_t *b = c->l_scr.as<int32_t>();
    LS.sA = b; LS.sB = b + H; LS.oA = b + 2 * H; LS.oB = b + 3 * H; LS.fr = b + 4 * H; LS.dq = b + 5 * H; LS.db = b + 6 * H;
    LS.pos = b + 7 * H;
    LS.fl = (uint8_t *)(b + 8 * H);
    // pairs with many hits: one wavefront each; the rest (and any handed back): one thread each
    ALLOC(c, c->l_done, np);
    ALLOC(c, c->l_big, np * 4 * (LW_BUCKETS + 2));
    ALLOC(c, c->l_crowd, np);
    hipLaunchKernelGGL(k_fill_u8, dim3((unsigned)((np + 255) / 256)), dim3(256), 0, c->st, c->l_crowd.as<uint8_t>(), np);
    hipLaunchKernelGGL(k_crowded_pairs, dim3((unsigned)((H + 255) / 256)), dim3(256), 0, c->st, hk, hv, c->s_flag.as<uint32_t>(),
                       c->s_idx.as<uint64_t>(), H, c->l_crowd.as<uint8_t>());
    ALLOC(c, c->l_cnt, 64);
    HIPCHK(c, hipMemsetAsync(c->l_cnt.p, 0, 64, c->st));
    uint32_t *cnt = c->l_cnt.as<uint32_t>();
    uint32_t *lists = c->l_big.as<uint32_t>();
    hipLaunchKernelGGL(k_big_select, dim3((unsigned)((np + 255) / 256)), dim3(256), 0, c->st, c->p_pairs.as<PairRec>(), np,
                       (uint64_t)c->tune.list_wave_min_hits, c->l_done.as<uint8_t>(), lists, cnt, cnt + 12,
                       c->l_crowd.as<uint8_t>(), lists + LW_BUCKETS * np, cnt + 8);
    BigListArgs BA;
    BA.pairs = c->p_pairs.as<PairRec>();
    BA.big = c->l_big.as<uint32_t>();
    BA.nbig = cnt;
    BA.bcnt = cnt + 12;
    BA.bstride = np;
    BA.next = cnt + 1;
    BA.handed_back = cnt + 2;
    BA.done = c->l_done.as<uint8_t>();
    BA.nnodes_total = (unsigned long long *)(cnt + 4);
    BA.hval = hv;
    BA.tmp = c->p_tmp.as<int4>();
    BA.nodes = c->p_nodes.as<int4>();
    BA.K = K;
    const uint64_t ncu = (uint64_t)std::max(c->ncu, 1);
    // crowded pairs (an offset of more than 128 hits): the 512-hit level, on the second stream,
    // beside the first level; the two take disjoint pairs
    BigListArgs B2 = BA;
    B2.big = lists + LW_BUCKETS * np;
    B2.nbig = cnt + 8;
    B2.bcnt = nullptr;
    B2.next = cnt + 9;
    B2.up = nullptr;
    B2.nup = nullptr;
    // Second stream, beside the 128-hit level: the one-thread builder for the pairs no wavefront
    // takes, then (when there are any) the crowded pairs at 512 hits, one wavefront each.
    uint32_t ncrowd = 0;
    HIPCHK(c, hipMemcpyAsync(&ncrowd, cnt + 8, 4, hipMemcpyDeviceToHost, c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
    HIPCHK(c, hipEventRecord(c->evl[0], c->st));
    HIPCHK(c, hipStreamWaitEvent(c->st2, c->evl[0], 0));
    hipLaunchKernelGGL(k_build_lists, dim3((unsigned)((np + 127) / 128)), dim3(128), 0, c->st2, c->p_pairs.as<PairRec>(), np, hv,
                       c->p_tmp.as<int4>(), c->p_nodes.as<int4>(), K, LS, c->l_done.as<uint8_t>(), (uint8_t)0, BA.nnodes_total);
    if (ncrowd)
      hipLaunchKernelGGL(k_build_lists_wave<LW_CAP>, dim3((unsigned)std::min<uint64_t>(ncrowd, 4 * ncu)), dim3(64), 0, c->st2, B2);
    HIPCHK(c, hipEventRecord(c->evl[1], c->st2));
    // first level: up to 128 hits per offset; a pair it still finds crowded runs at 512 hits after it
    BA.up = lists + (LW_BUCKETS + 1) * np;
    BA.nup = cnt + 10;
    hipLaunchKernelGGL(k_build_lists_wave<LW_CAP_FIRST>, dim3((unsigned)std::min<uint64_t>(np, 16 * ncu)), dim3(64), 0, c->st, BA);
    BigListArgs B3 = B2;
    B3.big = BA.up;
    B3.nbig = BA.nup;
    B3.next = cnt + 11;
    hipLaunchKernelGGL(k_build_lists_wave<LW_CAP>, dim3((unsigned)std::min<uint64_t>(np, 4 * ncu)), dim3(64), 0, c->st, B3);
    HIPCHK(c, hipStreamWaitEvent(c->st, c->evl[1], 0));
    // pairs a wavefront handed back (an offset of more than 512 hits)
    hipLaunchKernelGGL(k_build_lists, dim3((unsigned)((np + 127) / 128)), dim3(128), 0, c->st, c->p_pairs.as<PairRec>(), np, hv,
                       c->p_tmp.as<int4>(), c->p_nodes.as<int4>(), K, LS, c->l_done.as<uint8_t>(), (uint8_t)2, BA.nnodes_total);
  }
  HIPCHK(c, hipGetLastError());
  {
    uint32_t lc[16];
    HIPCHK(c, hipMemcpyAsync(lc, c->l_cnt.p, 64, hipMemcpyDeviceToHost, c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
    c->ctr.list_wave_pairs += lc[0];
    c->ctr.list_handed_back += lc[2];
    c->ctr.list_second_level += lc[8] + lc[10];
    memcpy(&batch_nodes, lc + 4, 8);
    c->ctr.nodes += batch_nodes;
  }
  c->ctr.ms_lists += now_ms() - t1;

  // processing order
  const double t2 = now_ms();
  ALLOC(c, c->p_order, np * 4);
  ALLOC(c, c->p_pout, np * sizeof(PairOut));
  ALLOC(c, c->p_recs, np * OIC_MAX_DISTINCT * sizeof(OutRec));
  ALLOC(c, c->p_queue, 256);
  ALLOC(c, c->p_stats, STAT_NSTATS * 8);
  HIPCHK(c, hipMemsetAsync(c->p_stats.p, 0, STAT_NSTATS * 8, c->st));
  HIPCHK(c, hipMemsetAsync(c->p_pout.p, 0, np * sizeof(PairOut), c->st));
  std::vector<uint32_t> order;
  c->batch_pred = false;
  {
    std::vector<float> pred;
    if (c->P.schedule == OIC_SCHED_LEARNED && c->have_model) {
      int rc = predict_pairs(c, np, pred);
      if (rc) return rc;
      bool finite = true;
      for (float v : pred) finite = finite && std::isfinite(v);
      if (finite) c->ctr.learned_pairs += np;
      c->batch_pred = finite;
    }
    int rc = make_order(c->P.schedule, c->p_pairs.as<PairRec>(), np, c->h_len.data(), pred.empty() ? nullptr : pred.data(), c->st, order);
    if (rc) return fail(c, OIC_EHIP, "scheduling failed");
  }
  HIPCHK(c, hipMemcpyAsync(c->p_order.p, order.data(), np * 4, hipMemcpyHostToDevice, c->st));

  // batched schedules: forward DPs of probe and speculative nodes, packed 4 per wavefront
  const bool batched = c->P.schedule != OIC_SCHED_PLAIN && !c->P.partial;
  if (batched) {
    int rc = (c->tune.fwd_order == 0) ? run_forward_passes_graph(c, np, H, batch_nodes) : run_forward_passes(c, np, H);
    if (rc) return rc;
  }

  PairKernelArgs A;
  A.R = dev_reads(c);
  A.C = search_const(c, true);
  if (!c->skip.empty()) {
    A.C.sscreen = c->g_sscreen.as<uint32_t>();
    A.C.s0 = s0;
  }
  A.outc = batched ? c->b_outc.as<uint64_t>() : nullptr;
  A.pairs = c->p_pairs.as<PairRec>();
  A.order = c->p_order.as<uint32_t>();
  A.nodes = c->p_nodes.as<int4>();
  A.recs = c->p_recs.as<OutRec>();
  A.pout = c->p_pout.as<PairOut>();
  A.queue = c->p_queue.as<uint32_t>();
  A.stats = c->p_stats.as<unsigned long long>();
  double ms = 0;
  const uint32_t waves = (uint32_t)c->ncu * 8;
  int rc = run_pairs<kCapSmall>(c, A, np, waves, (int64_t)1 << 20, &ms);
  if (rc) return rc;
  // pairs whose band outgrew the small tile: rerun with the wide tile
  std::vector<PairOut> pout(np);
  HIPCHK(c, hipMemcpyAsync(pout.data(), c->p_pout.p, np * sizeof(PairOut), hipMemcpyDeviceToHost, c->st));
  HIPCHK(c, hipStreamSynchronize(c->st));
  std::vector<uint32_t> retry;
  for (uint64_t p = 0; p < np; p++)
    if (pout[p].status == 1) retry.push_back((uint32_t)p);
  if (!retry.empty()) {
    c->ctr.overflow_retries += retry.size();
    ALLOC(c, c->p_retry, retry.size() * 4);
    HIPCHK(c, hipMemcpyAsync(c->p_retry.p, retry.data(), retry.size() * 4, hipMemcpyHostToDevice, c->st));
    A.order = c->p_retry.as<uint32_t>();
    const int32_t rows = c->T.error_bound[std::min<int32_t>(c->maxlen, (int32_t)oic::kMaxReadLen)] + 4;
    const int64_t vcap = std::min<int64_t>((int64_t)rows * kCapLarge, (int64_t)1 << 25);
    rc = run_pairs<kCapLarge>(c, A, retry.size(), std::min<uint32_t>(waves, 256), vcap, &ms);
    if (rc) return rc;
    HIPCHK(c, hipMemcpyAsync(pout.data(), c->p_pout.p, np * sizeof(PairOut), hipMemcpyDeviceToHost, c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
    for (uint64_t p = 0; p < np; p++)
      if (pout[p].status == 1) {
        PairRec pr;
        HIPCHK(c, hipMemcpy(&pr, c->p_pairs.as<PairRec>() + p, sizeof pr, hipMemcpyDeviceToHost));
        return fail(c, OIC_EUNSUPPORTED, "pair S %u (%s) x T %u needs a DP band wider than the widest tile (%d cells)", pr.s_id,
                    pr.dir ? "reverse" : "forward", pr.t_id, kCapLarge);
      }
  }
  c->ctr.ms_dp_kernel += ms;
  c->ctr.ms_pp_kernel += ms;
  c->ctr.ms_extend += now_ms() - t2;

  // records
  ALLOC(c, c->p_nout, np * 4);
  ALLOC(c, c->p_noff, np * 8);
  uint64_t nrec = 0;
  hipLaunchKernelGGL(k_nout, dim3((unsigned)((np + 255) / 256)), dim3(256), 0, c->st, c->p_pout.as<PairOut>(), np, c->p_nout.as<uint32_t>());
  ALLOC(c, c->s_scan, scan_scratch_words(np) * 8);
  HIPCHK(c, scan_excl<uint32_t>(c->p_nout.as<uint32_t>(), c->p_noff.as<uint64_t>(), np, c->s_scan.as<uint64_t>(), c->st, &nrec));
  if (nrec) {
    ALLOC(c, c->p_outrec, nrec * sizeof(OutRec));
    hipLaunchKernelGGL(k_compact_recs, dim3((unsigned)((np + 255) / 256)), dim3(256), 0, c->st, c->p_recs.as<OutRec>(), c->p_pout.as<PairOut>(),
                       c->p_noff.as<uint64_t>(), np, c->p_outrec.as<OutRec>());
  }
  uint64_t stats[STAT_NSTATS];
  HIPCHK(c, hipMemcpyAsync(stats, c->p_stats.p, sizeof stats, hipMemcpyDeviceToHost, c->st));
  HIPCHK(c, hipStreamSynchronize(c->st));
  for (int i = 0; i < 6; i++) sink->stats[i] += stats[i];
  c->ctr.dp_cells += stats[STAT_CELLS];
  c->ctr.dp_lane_slots += stats[STAT_SLOTS];
  c->ctr.extensions += stats[STAT_EXT];
  c->ctr.dp_calls += stats[STAT_DPCALLS];
  c->ctr.dp_bases += stats[STAT_BASES];
  c->ctr.pp_cells += stats[STAT_PPCELLS];
  c->ctr.pp_bases += stats[STAT_BASES];
  // records that fit go straight into the caller's sink, the rest are held by the context
  // (oic_fetch_pending hands them over without searching again)
  const uint64_t room = (sink->records && sink->capacity > sink->count) ? sink->capacity - sink->count : 0;
  const uint64_t direct = c->pending.empty() ? std::min<uint64_t>(room, nrec) : 0;
  if (direct)
    HIPCHK(c, hipMemcpy(sink->records + sink->count, c->p_outrec.p, direct * sizeof(OutRec), hipMemcpyDeviceToHost));
  if (nrec > direct) {
    const size_t o = c->pending.size();
    c->pending.resize(o + (nrec - direct));
    HIPCHK(c, hipMemcpy(c->pending.data() + o, c->p_outrec.as<OutRec>() + direct, (nrec - direct) * sizeof(OutRec),
                        hipMemcpyDeviceToHost));
  }
  sink->count += direct;
  sink->records_needed = sink->count + c->pending.size();
  if (c->tracing) {
    ALLOC(c, c->m_feat, np * MLP_IN * 4);
    hipLaunchKernelGGL(k_pair_features, dim3((unsigned)((np + 255) / 256)), dim3(256), 0, c->st, c->p_pairs.as<PairRec>(), np,
                       c->d_len.as<int32_t>(), c->m_feat.as<float>());
    const size_t o = c->tr_feat.size();
    c->tr_feat.resize(o + np * MLP_IN);
    HIPCHK(c, hipMemcpy(c->tr_feat.data() + o, c->m_feat.p, np * MLP_IN * 4, hipMemcpyDeviceToHost));
    for (uint64_t p = 0; p < np; p++) {
      c->tr_targ.push_back((float)std::log1p((double)pout[p].cells));
      c->tr_targ.push_back((float)(pout[p].exit_d / 1000.0));
    }
  }
  return OIC_OK;
}

static_assert(sizeof(OutRec) == sizeof(oic_overlap), "device records are ovOverlap records");

static int search_range(oic_ctx *c, uint32_t bgn, uint32_t end, oic_sink *sink) {
  if (!c->have_index) return fail(c, OIC_ESTATE, "no index built");
  const int K = (int)c->P.kmer;
  const uint64_t POS_CAP = (uint64_t)1 << 27;
  uint32_t r = bgn;
  while (r <= end) {
    uint64_t pos = 0;
    uint32_t ns = 0;
    std::vector<uint8_t> searched;
    while (r + ns <= end) {
      const int32_t L = c->h_len[r + ns];
      const bool s = L >= c->P.min_olap && L >= K;
      const uint64_t np = s ? 2ull * (uint64_t)(L - K + 1) : 0;
      if (ns > 0 && pos + np > POS_CAP) break;
      searched.push_back(s ? 1 : 0);
      pos += np;
      ns++;
      if (ns >= (1u << 20)) break;
    }
    int rc = search_batch(c, r, ns, searched, sink);
    if (rc) return rc;
    r += ns;
  }
  return OIC_OK;
}

extern "C" int oic_search(oic_ctx *c, uint32_t bgn_ref, uint32_t end_ref, oic_sink *sink) {
  if (!c || !sink) return OIC_EINVAL;
  if (bgn_ref < 1) bgn_ref = 1;
  if (end_ref > c->nreads) end_ref = c->nreads;
  c->pending.clear();
  if (sink->count > sink->capacity) return fail(c, OIC_EINVAL, "sink count %lu exceeds its capacity %lu", (unsigned long)sink->count,
                                                (unsigned long)sink->capacity);
  const uint32_t e = ref_end_searched(bgn_ref, end_ref, c->P.n_threads);
  const double t0 = now_ms();
  int rc = (e >= bgn_ref) ? search_range(c, bgn_ref, e, sink) : OIC_OK;
  c->ctr.ms_total += now_ms() - t0;
  if (rc == OIC_OK && !c->pending.empty())
    return fail(c, OIC_ECAPACITY, "sink holds %lu records, %lu needed", (unsigned long)sink->capacity,
                (unsigned long)sink->records_needed);
  return rc;
}

extern "C" int oic_fetch_pending(oic_ctx *c, oic_sink *sink) {
  if (!c || !sink) return OIC_EINVAL;
  const uint64_t n = c->pending.size();
  if (n == 0) return OIC_OK;
  if (!sink->records || sink->count > sink->capacity || sink->capacity - sink->count < n) {
    sink->records_needed = sink->count + n;
    return fail(c, OIC_ECAPACITY, "%lu held records need a sink of %lu", (unsigned long)n, (unsigned long)sink->records_needed);
  }
  memcpy(sink->records + sink->count, c->pending.data(), n * sizeof(oic_overlap));
  sink->count += n;
  sink->records_needed = sink->count;
  c->pending.clear();
  c->pending.shrink_to_fit();
  return OIC_OK;
}

extern "C" int oic_run(oic_ctx *c, uint32_t bgn_hash, uint32_t end_hash, uint32_t bgn_ref, uint32_t end_ref, oic_sink *sink) {
  if (!c || !sink) return OIC_EINVAL;
  const double t0 = now_ms();
  c->ctr = oic_counters{};
  c->pending.clear();
  sink->count = 0;
  sink->records_needed = 0;
  memset(sink->stats, 0, sizeof sink->stats);
  // overlapInCore.C:179-189
  if (bgn_hash < 1) bgn_hash = 1;
  if (c->nreads < end_hash) end_hash = c->nreads;
  if (bgn_ref < 1) bgn_ref = 1;
  if (end_ref > c->nreads) end_ref = c->nreads;
  const uint32_t e = ref_end_searched(bgn_ref, end_ref, c->P.n_threads);
  uint32_t bh = bgn_hash, eh = end_hash;
  while (bh < end_hash) {   // overlapInCore.C:204
    if (eh > end_hash) eh = end_hash;
    uint32_t last = 0;
    int rc = oic_index_build(c, bh, eh, &last);
    if (rc) return rc;
    if (e >= bgn_ref) {
      rc = search_range(c, bgn_ref, e, sink);
      if (rc) return rc;
    }
    bh = last + 1;
    eh = end_hash;
  }
  c->ctr.ms_total = now_ms() - t0;
  if (!c->pending.empty())
    return fail(c, OIC_ECAPACITY, "sink holds %lu records, %lu needed", (unsigned long)sink->capacity,
                (unsigned long)sink->records_needed);
  return OIC_OK;
}

// ------------------------------------------------------------------------------------ autotuner

// Times the search of a sample of ref reads under each launch shape and keeps the fastest: the
// packed forward-DP kernel over lanes per DP x band tile, then its loop form (device time of that
// kernel), the task order (time of the whole extension stage), then the list-builder threshold
// (device + host time of the list stage).
// Results never depend on the shape, so the sample's records are discarded.
extern "C" int oic_autotune(oic_ctx *c, uint32_t bgn_ref, uint32_t end_ref, oic_tuning *best, oic_tune_trial *trials,
                            uint32_t cap, uint32_t *ntrials) {
  if (!c) return OIC_EINVAL;
  if (!c->have_index) return fail(c, OIC_ESTATE, "no index built");
  if (bgn_ref < 1) bgn_ref = 1;
  if (end_ref > c->nreads) end_ref = c->nreads;
  if (end_ref < bgn_ref) return fail(c, OIC_EINVAL, "empty sample range");
  const oic_counters saved = c->ctr;
  oic_tuning cur = c->tune;
  uint32_t nt = 0;
  auto trial = [&](const oic_tuning &t, int what, double *ms) -> int {
    c->tune = t;
    c->ctr = oic_counters{};
    oic_sink sk{};
    const int rc = search_range(c, bgn_ref, end_ref, &sk);
    if (rc) return rc;
    *ms = (what == 0) ? c->ctr.ms_fwd_kernel : (what == 1) ? c->ctr.ms_lists : c->ctr.ms_extend;
    if (trials && nt < cap) {
      trials[nt].tuning = t;
      trials[nt].stage = what;
      trials[nt].ms = *ms;
    }
    nt++;
    return OIC_OK;
  };
  double ms = 0;
  int rc = trial(cur, 1, &ms);   // first run pays for buffer growth; not compared
  if (rc) { c->ctr = saved; return rc; }
  // Each knob in turn: the incumbent (the current setting) is timed first and a candidate replaces
  // it only when it is kTuneMargin faster, so timing noise does not flip the choice between runs.
  constexpr double kTuneMargin = 0.03;
  auto better = [&](double x, double incumbent) { return x < incumbent * (1.0 - kTuneMargin); };
  if (c->P.schedule != OIC_SCHED_PLAIN && !c->P.partial) {
    oic_tuning b = cur;
    double bms = 0;
    rc = trial(cur, 0, &bms);
    if (rc) { c->ctr = saved; return rc; }
    for (uint32_t g : {1u, 4u, 8u, 16u, 32u})
      for (uint32_t t : {16u, 32u, 64u, 128u, 256u, 512u}) {
        if (g == 1 ? t > 64 : (t == 16 || (g == 4 && t == 512) || (t <= 64 && g != 8 && g != 16))) continue;
        if (g == cur.fwd_lanes && t == cur.fwd_tile) continue;
        oic_tuning x = cur;
        x.fwd_lanes = g;
        x.fwd_tile = t;
        if (x.fwd_loop == 2) x.fwd_loop = 0;   // the flat form is tried with the loop forms below
        rc = trial(x, 0, &ms);
        if (rc) { c->ctr = saved; return rc; }
        if (better(ms, bms)) { bms = ms; b = x; }
      }
    cur = b;
  }
  if (c->P.schedule != OIC_SCHED_PLAIN && !c->P.partial) {
    // the loop form: the other of row-stepped and per-task for the chosen shape (one DP per lane
    // has a single form), and flat rows with 4 and 8 lanes per DP and 64-cell tiles
    oic_tuning b = cur;
    double bms = 0;
    rc = trial(cur, 0, &bms);
    if (rc) { c->ctr = saved; return rc; }
    std::vector<oic_tuning> cand;
    if (cur.fwd_lanes > 1 && cur.fwd_loop < 2) {
      oic_tuning x = cur;
      x.fwd_loop = 1 - cur.fwd_loop;
      cand.push_back(x);
    }
    for (uint32_t g : {4u, 8u}) {
      oic_tuning x = cur;
      x.fwd_lanes = g;
      x.fwd_tile = 64;
      x.fwd_loop = 2;
      if (!(cur.fwd_loop == 2 && cur.fwd_lanes == g && cur.fwd_tile == 64)) cand.push_back(x);
    }
    for (const oic_tuning &x : cand) {
      rc = trial(x, 0, &ms);
      if (rc) { c->ctr = saved; return rc; }
      if (better(ms, bms)) { bms = ms; b = x; }
    }
    cur = b;
  }
  if (c->P.schedule != OIC_SCHED_PLAIN && !c->P.partial) {
    // seed-length order runs without the per-batch graph: the whole extension stage is timed
    oic_tuning b = cur;
    double bms = 0;
    rc = trial(cur, 2, &bms);
    if (rc) { c->ctr = saved; return rc; }
    oic_tuning x = cur;
    x.fwd_order = 1 - cur.fwd_order;
    rc = trial(x, 2, &ms);
    if (rc) { c->ctr = saved; return rc; }
    if (better(ms, bms)) b = x;
    cur = b;
  }
  {
    // The list stage of the sample lasts tens of milliseconds and its thresholds differ there by
    // less than its run-to-run noise, while a sample without the heaviest repeat pairs can favour
    // a threshold that the whole run pays for (C3: 1024 against 2048 cost 2.3x over 2500 refs):
    // a candidate threshold must be kListMargin faster.
    constexpr double kListMargin = 0.10;
    oic_tuning b = cur;
    double bms = 0;
    rc = trial(cur, 1, &bms);
    if (rc) { c->ctr = saved; return rc; }
    for (uint32_t h : {256u, 1024u, 2048u, 4096u, 16384u}) {
      if (h == cur.list_wave_min_hits) continue;
      oic_tuning x = cur;
      x.list_wave_min_hits = h;
      rc = trial(x, 1, &ms);
      if (rc) { c->ctr = saved; return rc; }
      if (ms < bms * (1.0 - kListMargin)) { bms = ms; b = x; }
    }
    cur = b;
  }
  c->tune = cur;
  c->ctr = saved;
  if (best) *best = cur;
  if (ntrials) *ntrials = nt;
  return OIC_OK;
}

// ------------------------------------------------------------------------------------ DP units

struct BatchArgs {
  ReadsDev R;
  DPTables tb;
  SearchConst C;
  const oic_dp_task *dtasks;
  const oic_ext_task *etasks;
  uint32_t n;
  oic_dp_result *dres;
  oic_ext_result *eres;
  int32_t *delta;
  uint32_t delta_stride;
  uint32_t *queue;
  DPScratch sc0;
  int64_t row_stride, val_stride, delta_stride_sc;
};

template <int CAP>
__device__ __forceinline__ DPScratch batch_scratch(const BatchArgs &A, uint32_t w) {
  DPScratch s = A.sc0;
  s.rowL = A.sc0.rowL + (int64_t)w * A.row_stride;
  s.rowR = A.sc0.rowR + (int64_t)w * A.row_stride;
  s.rowOff = A.sc0.rowOff + (int64_t)w * A.row_stride;
  s.rowVal = A.sc0.rowVal + (int64_t)w * A.val_stride;
  s.dbuf = A.sc0.dbuf + (int64_t)w * A.delta_stride_sc;
  s.ldelta = A.sc0.ldelta + (int64_t)w * A.delta_stride_sc;
  s.rdelta = A.sc0.rdelta + (int64_t)w * A.delta_stride_sc;
  return s;
}

template <int CAP>
__global__ void __launch_bounds__(OIC_WAVE) k_dp_batch(BatchArgs A) {
  __shared__ int32_t bufA[CAP], bufB[CAP];
  __shared__ uint32_t item;
  const int lane = lane_id();
  const DPScratch sc = batch_scratch<CAP>(A, blockIdx.x);
  while (true) {
    if (lane == 0) item = atomicAdd(A.queue, 1u);
    __syncthreads();
    const uint32_t i = item;
    __syncthreads();
    if (i >= A.n) break;
    const oic_dp_task t = A.dtasks[i];
    SeqV a, b;
    if (t.dir == 0) {
      a = read_view(A.R, t.a_id, t.a_strand); a.b0 += t.a_pos;
      b = read_view(A.R, t.b_id, t.b_strand); b.b0 += t.b_pos;
    } else {
      a = read_view(A.R, t.a_id, 1 - t.a_strand); a.b0 += A.R.len[t.a_id] - 1 - t.a_pos;
      b = read_view(A.R, t.b_id, 1 - t.b_strand); b.b0 += A.R.len[t.b_id] - 1 - t.b_pos;
    }
    uint64_t cells = 0, slots = 0, lb = 0;
    DPOut o = wave_prefix_dp<CAP>(a, t.m, b, t.n, t.error_limit, t.dir == 0, true, A.tb, bufA, bufB, sc, cells, slots, lb);
    const int32_t *src = (t.dir == 0) ? sc.rdelta : sc.ldelta;
    if (A.delta && !o.overflow) {
      const int32_t nd = o.ndelta < (int32_t)A.delta_stride ? o.ndelta : (int32_t)A.delta_stride;
      for (int32_t j = lane; j < nd; j += OIC_WAVE) A.delta[(uint64_t)i * A.delta_stride + j] = src[j];
    }
    if (lane == 0) {
      oic_dp_result r;
      r.errors = o.e; r.a_end = o.a_end; r.t_end = o.t_end; r.match_to_end = o.mte; r.leftover = o.leftover;
      r.delta_len = o.ndelta; r.status = o.overflow ? OIC_ECAPACITY : 0; r.pad = 0; r.cells = cells;
      A.dres[i] = r;
    }
    __syncthreads();
  }
}

template <int CAP>
__global__ void __launch_bounds__(OIC_WAVE) k_ext_batch(BatchArgs A) {
  __shared__ int32_t bufA[CAP], bufB[CAP];
  __shared__ uint32_t item;
  const int lane = lane_id();
  const DPScratch sc = batch_scratch<CAP>(A, blockIdx.x);
  while (true) {
    if (lane == 0) item = atomicAdd(A.queue, 1u);
    __syncthreads();
    const uint32_t i = item;
    __syncthreads();
    if (i >= A.n) break;
    const oic_ext_task t = A.etasks[i];
    const SeqV S = read_view(A.R, t.s_id, t.s_dir), Src = read_view(A.R, t.s_id, 1 - t.s_dir);
    const SeqV T = read_view(A.R, t.t_id, 0), Trc = read_view(A.R, t.t_id, 1);
    uint64_t cells = 0, slots = 0, lb = 0;
    uint32_t calls = 0;
    ExtOut x = wave_extend<CAP>(t.start, t.offset, t.len, S, Src, A.R.len[t.s_id], T, Trc, A.R.len[t.t_id], A.C, bufA, bufB, sc, cells, slots,
                                calls, lb);
    if (A.delta && !x.overflow) {
      const int32_t nd = x.ndelta < (int32_t)A.delta_stride ? x.ndelta : (int32_t)A.delta_stride;
      for (int32_t j = lane; j < nd; j += OIC_WAVE) A.delta[(uint64_t)i * A.delta_stride + j] = sc.ldelta[j];
    }
    if (lane == 0) {
      oic_ext_result r;
      r.kind = x.kind; r.s_lo = x.s_lo; r.s_hi = x.s_hi; r.t_lo = x.t_lo; r.t_hi = x.t_hi; r.errors = x.errors;
      r.delta_len = x.ndelta; r.status = x.overflow ? OIC_ECAPACITY : 0;
      A.eres[i] = r;
    }
    __syncthreads();
  }
}

template <int CAP>
static int launch_batch(oic_ctx *c, BatchArgs A, bool ext, uint32_t waves, int64_t vcap) {
  int rc = ensure_wave_scratch(c, waves, CAP, vcap);
  if (rc) return rc;
  A.sc0 = scratch0(c);
  A.row_stride = c->sc_rows;
  A.val_stride = c->sc_val;
  A.delta_stride_sc = c->sc_rows + 4;
  HIPCHK(c, hipMemsetAsync(A.queue, 0, 4, c->st));
  if (ext) hipLaunchKernelGGL(k_ext_batch<CAP>, dim3(waves), dim3(OIC_WAVE), 0, c->st, A);
  else     hipLaunchKernelGGL(k_dp_batch<CAP>, dim3(waves), dim3(OIC_WAVE), 0, c->st, A);
  HIPCHK(c, hipGetLastError());
  HIPCHK(c, hipStreamSynchronize(c->st));
  return OIC_OK;
}

extern "C" int oic_prefix_dp_batch(oic_ctx *c, const oic_dp_task *tasks, uint32_t n, oic_dp_result *res, int32_t *delta,
                                   uint32_t stride) {
  if (!c || (!tasks && n) || (!res && n)) return OIC_EINVAL;
  if (n == 0) return OIC_OK;
  for (uint32_t i = 0; i < n; i++) {
    const oic_dp_task &t = tasks[i];
    if (t.a_id < 1 || t.a_id > c->nreads || t.b_id < 1 || t.b_id > c->nreads) return fail(c, OIC_EINVAL, "task %u: read id", i);
    if (t.m < 0 || t.n < 0 || t.m > t.n || t.error_limit < 0 || t.error_limit >= (int32_t)c->T.max_errors)
      return fail(c, OIC_EINVAL, "task %u: need 0 <= m <= n, 0 <= limit < max errors", i);
    const int32_t la = c->h_len[t.a_id], lb = c->h_len[t.b_id];
    if (t.dir == 0) {
      if (t.a_pos < 0 || t.a_pos + t.m > la || t.b_pos < 0 || t.b_pos + t.n > lb) return fail(c, OIC_EINVAL, "task %u: window", i);
    } else {
      if (t.a_pos >= la || t.a_pos - t.m + 1 < 0 || t.b_pos >= lb || t.b_pos - t.n + 1 < 0) return fail(c, OIC_EINVAL, "task %u: window", i);
    }
  }
  DevBuf dt, dr, dd, q;
  ALLOC(c, dt, (size_t)n * sizeof(oic_dp_task));
  ALLOC(c, dr, (size_t)n * sizeof(oic_dp_result));
  ALLOC(c, q, 256);
  if (delta) ALLOC(c, dd, (size_t)n * stride * 4);
  HIPCHK(c, hipMemcpyAsync(dt.p, tasks, (size_t)n * sizeof(oic_dp_task), hipMemcpyHostToDevice, c->st));
  BatchArgs A{};
  A.R = dev_reads(c);
  A.tb = dev_tables(c);
  A.dtasks = dt.as<oic_dp_task>();
  A.n = n;
  A.dres = dr.as<oic_dp_result>();
  A.delta = delta ? dd.as<int32_t>() : nullptr;
  A.delta_stride = stride;
  A.queue = q.as<uint32_t>();
  const uint32_t waves = std::min<uint32_t>(n, (uint32_t)c->ncu * 4);
  const int32_t rows = c->T.error_bound[std::min<int32_t>(c->maxlen, (int32_t)oic::kMaxReadLen)] + 4;
  int rc = launch_batch<kCapLarge>(c, A, false, waves, std::min<int64_t>((int64_t)rows * kCapLarge, (int64_t)1 << 24));
  if (rc) return rc;
  HIPCHK(c, hipMemcpy(res, dr.p, (size_t)n * sizeof(oic_dp_result), hipMemcpyDeviceToHost));
  if (delta) HIPCHK(c, hipMemcpy(delta, dd.p, (size_t)n * stride * 4, hipMemcpyDeviceToHost));
  return OIC_OK;
}

// ------------------------------------------------------------ DP units without allocation
//
// oic_prefix_dp_batch_noalloc runs oic_prefix_dp_batch's kernel on memory the library holds from
// load time: one file-scope device array per GPU, carved into the task and result staging areas,
// the delta area, the work-queue counter and the per-wavefront row store of the traceback.  Its
// bounds are fixed at compile time (na::kMax* below).  Every call on a device shares that device's
// array, so calls that overlap on one GPU race; each call clears the counter and the delta area it
// uses before its kernel runs.

namespace na {
constexpr size_t up256(size_t b) { return (b + 255) / 256 * 256; }
constexpr uint32_t kMaxTasks = 16384;                  // tasks per call
constexpr uint64_t kMaxDeltaInts = (uint64_t)1 << 20;  // tasks x delta_stride per call
constexpr uint32_t kWaves = 64;                        // persistent wavefronts of the kernel
constexpr int32_t kRows = 1024;                        // rows stored per DP: error limits up to kRows - 3
constexpr int64_t kVals = (int64_t)1 << 17;            // band cells stored per DP (else status OIC_ECAPACITY)
constexpr int32_t kMaxLimit = kRows - 3;
constexpr size_t kQueueOff = 0;
constexpr size_t kTaskOff = 256;
constexpr size_t kResOff = kTaskOff + up256((size_t)kMaxTasks * sizeof(oic_dp_task));
constexpr size_t kDeltaOff = kResOff + up256((size_t)kMaxTasks * sizeof(oic_dp_result));
constexpr size_t kRowLOff = kDeltaOff + up256((size_t)kMaxDeltaInts * 4);
constexpr size_t kRowROff = kRowLOff + up256((size_t)kWaves * kRows * 4);
constexpr size_t kRowOffOff = kRowROff + up256((size_t)kWaves * kRows * 4);
constexpr size_t kRowValOff = kRowOffOff + up256((size_t)kWaves * kRows * 8);
constexpr size_t kDbufOff = kRowValOff + up256((size_t)kWaves * kVals * 4);
constexpr size_t kLdOff = kDbufOff + up256((size_t)kWaves * (kRows + 4) * 4);
constexpr size_t kRdOff = kLdOff + up256((size_t)kWaves * (kRows + 4) * 4);
constexpr size_t kBytes = kRdOff + up256((size_t)kWaves * (kRows + 4) * 4);   // about 41 MB
}  // namespace na

__device__ __align__(256) uint8_t scratch[na::kBytes];

// scratch's device address on each GPU, looked up once per device (oic_create)
static std::mutex g_scratch_mu;
static void *g_scratch_ptr[256] = {};

static int lookup_scratch(oic_ctx *c, int dev) {
  if (dev < 0 || dev >= 256) return fail(c, OIC_EINVAL, "device %d out of range", dev);
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) return fail(c, OIC_EHIP, "hipGetSymbolAddress(scratch): %s", hipGetErrorString(e));
    g_scratch_ptr[dev] = p;
  }
  c->static_scratch = g_scratch_ptr[dev];
  return OIC_OK;
}

extern "C" int oic_prefix_dp_batch_noalloc(oic_ctx *c, const oic_dp_task *tasks, uint32_t n, oic_dp_result *res, int32_t *delta,
                                           uint32_t stride) {
  if (!c || (!tasks && n) || (!res && n)) return OIC_EINVAL;
  if (n == 0) return OIC_OK;
  if (!c->static_scratch) return fail(c, OIC_ESTATE, "no static scratch on this device");
  if (n > na::kMaxTasks) return fail(c, OIC_EUNSUPPORTED, "%u tasks: at most %u per call without allocation", n, na::kMaxTasks);
  if (delta && (uint64_t)n * stride > na::kMaxDeltaInts)
    return fail(c, OIC_EUNSUPPORTED, "%u tasks x delta stride %u: at most %lu delta entries per call without allocation", n, stride,
                (unsigned long)na::kMaxDeltaInts);
  for (uint32_t i = 0; i < n; i++) {
    const oic_dp_task &t = tasks[i];
    if (t.a_id < 1 || t.a_id > c->nreads || t.b_id < 1 || t.b_id > c->nreads) return fail(c, OIC_EINVAL, "task %u: read id", i);
    if (t.m < 0 || t.n < 0 || t.m > t.n || t.error_limit < 0 || t.error_limit >= (int32_t)c->T.max_errors)
      return fail(c, OIC_EINVAL, "task %u: need 0 <= m <= n, 0 <= limit < max errors", i);
    if (t.error_limit > na::kMaxLimit)
      return fail(c, OIC_EUNSUPPORTED, "task %u: error limit %d, at most %d without allocation", i, t.error_limit, na::kMaxLimit);
    const int32_t la = c->h_len[t.a_id], lb = c->h_len[t.b_id];
    if (t.dir == 0) {
      if (t.a_pos < 0 || t.a_pos + t.m > la || t.b_pos < 0 || t.b_pos + t.n > lb) return fail(c, OIC_EINVAL, "task %u: window", i);
    } else {
      if (t.a_pos >= la || t.a_pos - t.m + 1 < 0 || t.b_pos >= lb || t.b_pos - t.n + 1 < 0) return fail(c, OIC_EINVAL, "task %u: window", i);
    }
  }
  uint8_t *base = (uint8_t *)c->static_scratch;
  oic_dp_task *dt = (oic_dp_task *)(base + na::kTaskOff);
  oic_dp_result *dr = (oic_dp_result *)(base + na::kResOff);
  int32_t *dd = (int32_t *)(base + na::kDeltaOff);
  uint32_t *queue = (uint32_t *)(base + na::kQueueOff);
  HIPCHK(c, hipMemcpyAsync(dt, tasks, (size_t)n * sizeof(oic_dp_task), hipMemcpyHostToDevice, c->st));
  hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, c->st, queue, (uint64_t)1, 0u);
  if (delta) {
    const uint64_t nd = (uint64_t)n * stride;
    hipLaunchKernelGGL(k_fill_u32, dim3((unsigned)((nd + 255) / 256)), dim3(256), 0, c->st, (uint32_t *)dd, nd, 0u);
  }
  BatchArgs A{};
  A.R = dev_reads(c);
  A.tb = dev_tables(c);
  A.dtasks = dt;
  A.n = n;
  A.dres = dr;
  A.delta = delta ? dd : nullptr;
  A.delta_stride = stride;
  A.queue = queue;
  A.sc0.rowL = (int32_t *)(base + na::kRowLOff);
  A.sc0.rowR = (int32_t *)(base + na::kRowROff);
  A.sc0.rowOff = (int64_t *)(base + na::kRowOffOff);
  A.sc0.rowVal = (int32_t *)(base + na::kRowValOff);
  A.sc0.val_cap = na::kVals;
  A.sc0.max_rows = na::kRows - 2;
  A.sc0.dbuf = (int32_t *)(base + na::kDbufOff);
  A.sc0.ldelta = (int32_t *)(base + na::kLdOff);
  A.sc0.rdelta = (int32_t *)(base + na::kRdOff);
  A.row_stride = na::kRows;
  A.val_stride = na::kVals;
  A.delta_stride_sc = na::kRows + 4;
  const uint32_t waves = std::min<uint32_t>(n, na::kWaves);
  hipLaunchKernelGGL(k_dp_batch<kCapLarge>, dim3(waves), dim3(OIC_WAVE), 0, c->st, A);
  HIPCHK(c, hipGetLastError());
  HIPCHK(c, hipMemcpyAsync(res, dr, (size_t)n * sizeof(oic_dp_result), hipMemcpyDeviceToHost, c->st));
  if (delta) HIPCHK(c, hipMemcpyAsync(delta, dd, (size_t)n * stride * 4, hipMemcpyDeviceToHost, c->st));
  HIPCHK(c, hipStreamSynchronize(c->st));
  return OIC_OK;
}

extern "C" int oic_extend_batch(oic_ctx *c, const oic_ext_task *tasks, uint32_t n, oic_ext_result *res, int32_t *delta, uint32_t stride) {
  if (!c || (!tasks && n) || (!res && n)) return OIC_EINVAL;
  if (n == 0) return OIC_OK;
  for (uint32_t i = 0; i < n; i++) {
    const oic_ext_task &t = tasks[i];
    if (t.s_id < 1 || t.s_id > c->nreads || t.t_id < 1 || t.t_id > c->nreads) return fail(c, OIC_EINVAL, "task %u: read id", i);
    const int32_t ls = c->h_len[t.s_id], lt = c->h_len[t.t_id];
    if (t.len < 1 || t.start < 0 || t.offset < 0 || t.start + t.len > ls || t.offset + t.len > lt || (t.s_dir != 0 && t.s_dir != 1))
      return fail(c, OIC_EINVAL, "task %u: seed outside the reads", i);
  }
  DevBuf dt, dr, dd, q;
  ALLOC(c, dt, (size_t)n * sizeof(oic_ext_task));
  ALLOC(c, dr, (size_t)n * sizeof(oic_ext_result));
  ALLOC(c, q, 256);
  if (delta) ALLOC(c, dd, (size_t)n * stride * 4);
  HIPCHK(c, hipMemcpyAsync(dt.p, tasks, (size_t)n * sizeof(oic_ext_task), hipMemcpyHostToDevice, c->st));
  BatchArgs A{};
  A.R = dev_reads(c);
  A.tb = dev_tables(c);
  A.C = search_const(c, false);
  A.etasks = dt.as<oic_ext_task>();
  A.n = n;
  A.eres = dr.as<oic_ext_result>();
  A.delta = delta ? dd.as<int32_t>() : nullptr;
  A.delta_stride = stride;
  A.queue = q.as<uint32_t>();
  const uint32_t waves = std::min<uint32_t>(n, (uint32_t)c->ncu * 4);
  const int32_t rows = c->T.error_bound[std::min<int32_t>(c->maxlen, (int32_t)oic::kMaxReadLen)] + 4;
  int rc = launch_batch<kCapLarge>(c, A, true, waves, std::min<int64_t>((int64_t)rows * kCapLarge, (int64_t)1 << 24));
  if (rc) return rc;
  HIPCHK(c, hipMemcpy(res, dr.p, (size_t)n * sizeof(oic_ext_result), hipMemcpyDeviceToHost));
  if (delta) HIPCHK(c, hipMemcpy(delta, dd.p, (size_t)n * stride * 4, hipMemcpyDeviceToHost));
  return OIC_OK;
}

// Add_Overlap and the pair's selection over explicit extension results, one pair per lane
__global__ void k_distinct_olaps(SearchConst C, const uint32_t *__restrict__ first, uint32_t n,
                                 const oic_olap_in *__restrict__ in, oic_olap_out *__restrict__ out,
                                 int32_t *__restrict__ count) {
  const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= n) return;
  OlapDev olap[OIC_MAX_DISTINCT];
  int32_t ct = 0;
  for (uint32_t k = first[g]; k < first[g + 1]; k++) {
    const oic_olap_in x = in[k];
    add_overlap(C, x.s_lo, x.s_hi, x.t_lo, x.t_hi, x.quality, 0, olap, &ct);
  }
  int32_t deleted[OIC_MAX_DISTINCT] = {0, 0, 0};
  select_distinct(C, olap, ct, deleted);
  for (int32_t i = 0; i < ct; i++) {
    const OlapDev &o = olap[i];
    oic_olap_out r;
    r.s_lo = o.s_lo; r.s_hi = o.s_hi; r.t_lo = o.t_lo; r.t_hi = o.t_hi;
    r.min_diag = o.min_diag; r.max_diag = o.max_diag;
    r.s_left = o.slb; r.s_right = o.srb; r.t_left = o.tlb; r.t_right = o.trb;
    r.deleted = deleted[i];
    r.pad = 0;
    r.quality = o.quality;
    out[(size_t)g * OIC_MAX_DISTINCT + i] = r;
  }
  count[g] = ct;
}

extern "C" int oic_distinct_olaps_batch(oic_ctx *c, const uint32_t *first, uint32_t n, const oic_olap_in *in,
                                        oic_olap_out *out, int32_t *count) {
  if (!c || (n && (!first || !out || !count))) return OIC_EINVAL;
  if (n == 0) return OIC_OK;
  if (first[0] != 0) return fail(c, OIC_EINVAL, "first[0] is %u, not 0", first[0]);
  for (uint32_t g = 0; g < n; g++)
    if (first[g + 1] < first[g]) return fail(c, OIC_EINVAL, "first[] decreases at pair %u", g);
  const uint32_t total = first[n];
  if (total && !in) return OIC_EINVAL;
  DevBuf df, di, dout, dc;
  ALLOC(c, df, (size_t)(n + 1) * 4);
  ALLOC(c, di, (size_t)std::max<uint32_t>(total, 1) * sizeof(oic_olap_in));
  ALLOC(c, dout, (size_t)n * OIC_MAX_DISTINCT * sizeof(oic_olap_out));
  ALLOC(c, dc, (size_t)n * 4);
  HIPCHK(c, hipMemcpyAsync(df.p, first, (size_t)(n + 1) * 4, hipMemcpyHostToDevice, c->st));
  if (total) HIPCHK(c, hipMemcpyAsync(di.p, in, (size_t)total * sizeof(oic_olap_in), hipMemcpyHostToDevice, c->st));
  HIPCHK(c, hipMemsetAsync(dout.p, 0, (size_t)n * OIC_MAX_DISTINCT * sizeof(oic_olap_out), c->st));
  const SearchConst C = search_const(c, false);
  k_distinct_olaps<<<(n + 255) / 256, 256, 0, c->st>>>(C, df.as<uint32_t>(), n, di.as<oic_olap_in>(),
                                                       dout.as<oic_olap_out>(), dc.as<int32_t>());
  HIPCHK(c, hipGetLastError());
  HIPCHK(c, hipMemcpyAsync(out, dout.p, (size_t)n * OIC_MAX_DISTINCT * sizeof(oic_olap_out), hipMemcpyDeviceToHost, c->st));
  HIPCHK(c, hipMemcpyAsync(count, dc.p, (size_t)n * 4, hipMemcpyDeviceToHost, c->st));
  HIPCHK(c, hipStreamSynchronize(c->st));
  return OIC_OK;
}

// ------------------------------------------------------------------------------------ cost model

extern "C" int oic_set_tracing(oic_ctx *c, int on) {
  if (!c) return OIC_EINVAL;
  c->tracing = on != 0;
  c->tr_feat.clear();
  c->tr_targ.clear();
  return OIC_OK;
}

extern "C" int oic_get_traces(oic_ctx *c, float *feat, float *targ, uint64_t cap, uint64_t *count) {
  if (!c || !count) return OIC_EINVAL;
  const uint64_t n = c->tr_targ.size() / MLP_OUT;
  *count = n;
  if (!feat && !targ) return OIC_OK;
  if (cap < n) return fail(c, OIC_ECAPACITY, "%lu traces, capacity %lu", (unsigned long)n, (unsigned long)cap);
  if (feat) memcpy(feat, c->tr_feat.data(), n * MLP_IN * 4);
  if (targ) memcpy(targ, c->tr_targ.data(), n * MLP_OUT * 4);
  return OIC_OK;
}

static int upload_model(oic_ctx *c) {
  ALLOC(c, c->m_param, MLP_NPARAM * 4);
  ALLOC(c, c->m_mu, MLP_IN * 4);
  ALLOC(c, c->m_sd, MLP_IN * 4);
  HIPCHK(c, hipMemcpy(c->m_param.p, c->mparam, MLP_NPARAM * 4, hipMemcpyHostToDevice));
  HIPCHK(c, hipMemcpy(c->m_mu.p, c->mmu, MLP_IN * 4, hipMemcpyHostToDevice));
  HIPCHK(c, hipMemcpy(c->m_sd.p, c->msd, MLP_IN * 4, hipMemcpyHostToDevice));
  c->have_model = true;
  return OIC_OK;
}

extern "C" int oic_model_get(const oic_ctx *c, float *w, float *mu, float *sd) {
  if (!c || !c->have_model) return OIC_ESTATE;
  if (w) memcpy(w, c->mparam, MLP_NPARAM * 4);
  if (mu) memcpy(mu, c->mmu, MLP_IN * 4);
  if (sd) memcpy(sd, c->msd, MLP_IN * 4);
  return OIC_OK;
}

extern "C" int oic_model_set(oic_ctx *c, const float *w, const float *mu, const float *sd) {
  if (!c || !w || !mu || !sd) return OIC_EINVAL;
  memcpy(c->mparam, w, MLP_NPARAM * 4);
  memcpy(c->mmu, mu, MLP_IN * 4);
  memcpy(c->msd, sd, MLP_IN * 4);
  return upload_model(c);
}

extern "C" int oic_model_predict(oic_ctx *c, const float *feat, uint64_t n, float *out) {
  if (!c || (!feat && n) || (!out && n)) return OIC_EINVAL;
  if (!c->have_model) return fail(c, OIC_ESTATE, "no model");
  if (n == 0) return OIC_OK;
  DevBuf x, y;
  ALLOC(c, x, n * MLP_IN * 4);
  ALLOC(c, y, n * MLP_OUT * 4);
  HIPCHK(c, hipMemcpy(x.p, feat, n * MLP_IN * 4, hipMemcpyHostToDevice));
  hipLaunchKernelGGL(k_mlp_forward, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->st, x.as<float>(), n, c->m_param.as<float>(),
                     c->m_mu.as<float>(), c->m_sd.as<float>(), y.as<float>());
  HIPCHK(c, hipGetLastError());
  HIPCHK(c, hipStreamSynchronize(c->st));
  HIPCHK(c, hipMemcpy(out, y.p, n * MLP_OUT * 4, hipMemcpyDeviceToHost));
  return OIC_OK;
}

static uint32_t grad_waves(oic_ctx *c, uint64_t n) {
  return (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>((n + 63) / 64, (uint64_t)c->ncu * 2));
}

extern "C" int oic_model_grad(oic_ctx *c, const float *w, const float *z, const float *t, uint64_t n, float w_exit, float *loss,
                              float *grad) {
  if (!c || !w || !z || !t || !n || !grad) return OIC_EINVAL;
  DevBuf dp, dz, dt, dg, dl;
  ALLOC(c, dp, MLP_NPARAM * 4);
  ALLOC(c, dz, n * MLP_IN * 4);
  ALLOC(c, dt, n * MLP_OUT * 4);
  ALLOC(c, dg, MLP_NPARAM * 4);
  ALLOC(c, dl, 4);
  HIPCHK(c, hipMemcpy(dp.p, w, MLP_NPARAM * 4, hipMemcpyHostToDevice));
  HIPCHK(c, hipMemcpy(dz.p, z, n * MLP_IN * 4, hipMemcpyHostToDevice));
  HIPCHK(c, hipMemcpy(dt.p, t, n * MLP_OUT * 4, hipMemcpyHostToDevice));
  hipLaunchKernelGGL(k_fill_f, dim3((MLP_NPARAM + 255) / 256), dim3(256), 0, c->st, dg.as<float>(), (uint64_t)MLP_NPARAM, 0.f);
  hipLaunchKernelGGL(k_fill_f, dim3(1), dim3(64), 0, c->st, dl.as<float>(), (uint64_t)1, 0.f);
  hipLaunchKernelGGL(k_mlp_grad, dim3(grad_waves(c, n)), dim3(64), 0, c->st, dz.as<float>(), dt.as<float>(), (const uint32_t *)nullptr, n,
                     dp.as<float>(), w_exit, 1.0f / (float)n, dg.as<float>(), dl.as<float>());
  HIPCHK(c, hipGetLastError());
  HIPCHK(c, hipStreamSynchronize(c->st));
  HIPCHK(c, hipMemcpy(grad, dg.p, MLP_NPARAM * 4, hipMemcpyDeviceToHost));
  float l = 0;
  HIPCHK(c, hipMemcpy(&l, dl.p, 4, hipMemcpyDeviceToHost));
  if (loss) *loss = l / (float)n;
  return OIC_OK;
}

extern "C" int oic_model_train(oic_ctx *c, const float *feat, const float *targ, uint64_t n, const oic_train_params *tp, float *final_loss) {
  if (!c || !feat || !targ || n < 2 || !tp || tp->batch == 0 || tp->epochs == 0) return OIC_EINVAL;
  // standardisation
  for (int i = 0; i < MLP_IN; i++) {
    double s = 0, s2 = 0;
    for (uint64_t r = 0; r < n; r++) { const double v = feat[r * MLP_IN + i]; s += v; s2 += v * v; }
    const double m = s / n, var = s2 / n - m * m;
    c->mmu[i] = (float)m;
    c->msd[i] = (float)(var > 1e-12 ? std::sqrt(var) : 1.0);
  }
  std::vector<float> z(n * MLP_IN);
  for (uint64_t r = 0; r < n; r++)
    for (int i = 0; i < MLP_IN; i++) z[r * MLP_IN + i] = (feat[r * MLP_IN + i] - c->mmu[i]) / c->msd[i];
  // Glorot-uniform initial weights from a fixed seed, zero biases
  std::mt19937_64 rng(tp->seed);
  auto init = [&](int off, int fan_out, int fan_in) {
    const double a = std::sqrt(6.0 / (fan_in + fan_out));
    std::uniform_real_distribution<double> u(-a, a);
    for (int i = 0; i < fan_out * fan_in; i++) c->mparam[off + i] = (float)u(rng);
  };
  memset(c->mparam, 0, sizeof c->mparam);
  init(MLP_OFF_W1, MLP_H, MLP_IN);
  init(MLP_OFF_W2, MLP_H, MLP_H);
  init(MLP_OFF_W3, MLP_OUT, MLP_H);
  int rc = upload_model(c);
  if (rc) return rc;
  DevBuf dz, dt, dg, dm, dv, dl, didx;
  ALLOC(c, dz, n * MLP_IN * 4);
  ALLOC(c, dt, n * MLP_OUT * 4);
  ALLOC(c, dg, MLP_NPARAM * 4);
  ALLOC(c, dm, MLP_NPARAM * 4);
  ALLOC(c, dv, MLP_NPARAM * 4);
  ALLOC(c, dl, 4);
  ALLOC(c, didx, n * 4);
  HIPCHK(c, hipMemcpy(dz.p, z.data(), n * MLP_IN * 4, hipMemcpyHostToDevice));
  HIPCHK(c, hipMemcpy(dt.p, targ, n * MLP_OUT * 4, hipMemcpyHostToDevice));
  hipLaunchKernelGGL(k_fill_f, dim3((MLP_NPARAM + 255) / 256), dim3(256), 0, c->st, dm.as<float>(), (uint64_t)MLP_NPARAM, 0.f);
  hipLaunchKernelGGL(k_fill_f, dim3((MLP_NPARAM + 255) / 256), dim3(256), 0, c->st, dv.as<float>(), (uint64_t)MLP_NPARAM, 0.f);
  std::vector<uint32_t> perm(n);
  std::iota(perm.begin(), perm.end(), 0u);
  int step = 0;
  float last = 0;
  for (uint32_t ep = 0; ep < tp->epochs; ep++) {
    std::shuffle(perm.begin(), perm.end(), rng);
    HIPCHK(c, hipMemcpyAsync(didx.p, perm.data(), n * 4, hipMemcpyHostToDevice, c->st));
    hipLaunchKernelGGL(k_fill_f, dim3(1), dim3(64), 0, c->st, dl.as<float>(), (uint64_t)1, 0.f);
    for (uint64_t b0 = 0; b0 < n; b0 += tp->batch) {
      const uint64_t nb = std::min<uint64_t>(tp->batch, n - b0);
      hipLaunchKernelGGL(k_fill_f, dim3((MLP_NPARAM + 255) / 256), dim3(256), 0, c->st, dg.as<float>(), (uint64_t)MLP_NPARAM, 0.f);
      hipLaunchKernelGGL(k_mlp_grad, dim3(grad_waves(c, nb)), dim3(64), 0, c->st, dz.as<float>(), dt.as<float>(),
                         didx.as<uint32_t>() + b0, nb, c->m_param.as<float>(), tp->w_exit, 1.0f / (float)nb, dg.as<float>(),
                         dl.as<float>());
      step++;
      const float c1 = 1.f - std::pow(tp->beta1, (float)step), c2 = 1.f - std::pow(tp->beta2, (float)step);
      hipLaunchKernelGGL(k_adam, dim3((MLP_NPARAM + 255) / 256), dim3(256), 0, c->st, c->m_param.as<float>(), dg.as<float>(),
                         dm.as<float>(), dv.as<float>(), MLP_NPARAM, tp->lr, tp->beta1, tp->beta2, tp->eps, c1, c2);
    }
    HIPCHK(c, hipGetLastError());
    float l = 0;
    HIPCHK(c, hipMemcpyAsync(&l, dl.p, 4, hipMemcpyDeviceToHost, c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
    last = l / (float)n;
  }
  HIPCHK(c, hipMemcpy(c->mparam, c->m_param.p, MLP_NPARAM * 4, hipMemcpyDeviceToHost));
  if (final_loss) *final_loss = last;
  return OIC_OK;
}
