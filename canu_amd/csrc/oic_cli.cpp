// oic_cli.cpp -- overlapInCore's command line on the MI355X engine.
//
//   overlapInCore [-partial] [-t T] -k K [-k skip.kmers] [--hashbits B] [--hashload L] [--hashdatalen N]
//                 [--maxerate E] [--alignnoise A] [--minlength M] [--minkmers] [-m|-u] [-z]
//                 [-l L] [-h a-b] [-r c-d] [--hpc] -o out.ovb[.WORKING] [-s out.stats] <seqStore|reads.fasta|.fastq>
//
// Options and their defaults follow overlapInCore.C:293-411 (main's argv loop) and
// overlapInCore.H:367-473 (oicParameters::initialize).  Output, as ovFile(ovFileFullWrite) and
// main() write it:
//   - the -o file: buffers of full records (a_iid, b_iid, the two ovOverlapDAT words as high and low
//     32-bit halves), each buffer a raw snappy block preceded by its uint64 compressed length
//     (stores/ovStoreFile.C:98-283);
//   - <prefix>.oc: uint64 overlaps, uint32 reads + 1, uint32 overlaps per read (each overlap counts
//     for both reads) (stores/ovStoreFile.H:50-110); the prefix is the -o name without a trailing
//     ".WORKING" and ".ovb", which is how canu's pipeline names it (OverlapInCore.pm:188,197);
//   - the -s file (stderr if absent): the 8 stats lines (overlapInCore.C:539-560).
//
// Reads come from a seqStore directory, in the version and form the store hands over (read by
// oic_seqstore.h), or from a FASTA or FASTQ file, numbered 1.. in file order.  --hpc
// homopolymer-compresses FASTA/FASTQ reads, as a store with the homopolymerCompression marker hands
// them over (stores/sqRead.H:211, :454).
//
//   overlapInCore --dump-store <seqStore>   prints the reads a run would see, as FASTA (">id lib len").
#include <cctype>
#include <cerrno>
#include <cinttypes>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/oic.h"
#include "oic_seqstore.h"
#include "oic_snappy.h"

static void die(const char *msg, const char *arg = nullptr) {
  if (arg) fprintf(stderr, "overlapInCore: %s '%s'\n", msg, arg);
  else     fprintf(stderr, "overlapInCore: %s\n", msg);
  exit(1);
}

// "a-b" or "a" (both ends inclusive)
static void decode_range(const char *s, uint32_t &bgn, uint32_t &end) {
  char *e = nullptr;
  const unsigned long a = strtoul(s, &e, 10);
  if (e == s) die("bad range", s);
  bgn = (uint32_t)a;
  end = (uint32_t)a;
  if (*e == '-') {
    const char *t = e + 1;
    const unsigned long b = strtoul(t, &e, 10);
    if (e == t) die("bad range", s);
    end = (uint32_t)b;
  }
}

struct ReadSet {
  std::string bases;                 // all reads, each followed by a NUL
  std::vector<uint64_t> start{0};    // 1-based
  std::vector<uint32_t> len{0};
};

static void add_read(ReadSet &R, std::string &seq, bool hpc) {
  if (hpc) {   // keep the first base of each run of equal bases (case-insensitive)
    std::string c;
    c.reserve(seq.size());
    for (char ch : seq)
      if (c.empty() || tolower((unsigned char)c.back()) != tolower((unsigned char)ch)) c.push_back(ch);
    seq.swap(c);
  }
  R.start.push_back(R.bases.size());
  R.len.push_back((uint32_t)seq.size());
  R.bases += seq;
  R.bases.push_back('\0');
  seq.clear();
}

static ReadSet load_reads(const char *path, bool hpc) {
  FILE *F = fopen(path, "r");
  if (!F) die("cannot open reads", path);
  ReadSet R;
  std::string seq;
  char *line = nullptr;
  size_t cap = 0;
  ssize_t n;
  bool fastq = false, in_read = false;
  int fq_state = 0;   // FASTQ: 0 header, 1 sequence, 2 plus, 3 quality
  while ((n = getline(&line, &cap, F)) > 0) {
    while (n > 0 && (line[n - 1] == '\n' || line[n - 1] == '\r')) line[--n] = 0;
    if (!in_read && !fastq && n > 0 && line[0] == '@') fastq = true;
    if (fastq) {
      if (fq_state == 0) { if (n == 0) continue; fq_state = 1; }
      else if (fq_state == 1) { seq.assign(line, n); fq_state = 2; }
      else if (fq_state == 2) { fq_state = 3; }
      else { add_read(R, seq, hpc); fq_state = 0; }
      continue;
    }
    if (n > 0 && line[0] == '>') {
      if (in_read) add_read(R, seq, hpc);
      in_read = true;
      continue;
    }
    if (in_read) seq.append(line, n);
  }
  if (in_read) add_read(R, seq, hpc);
  free(line);
  fclose(F);
  if (R.len.size() < 2) die("no reads in", path);
  return R;
}

// every read of a store, 1..last, with length 0 where the store hands over nothing
static ReadSet load_store(const char *path) {
  oic_store::Store S;
  std::string err, seq;
  if (!S.open(path, err)) die(err.c_str());
  ReadSet R;
  for (uint32_t id = 1; id <= S.nreads; id++) {
    seq.clear();
    if (S.length(id) > 0) {
      if (!S.sequence(id, seq, err)) die(err.c_str());
      if (seq.size() != S.length(id)) die("read length differs from the store's metadata in", path);
    }
    add_read(R, seq, false);
  }
  if (R.len.size() < 2) die("no reads in", path);
  return R;
}

static int dump_store(const char *path) {
  oic_store::Store S;
  std::string err;
  if (!S.open(path, err)) die(err.c_str());
  ReadSet R = load_store(path);
  for (uint32_t id = 1; id < R.len.size(); id++)
    printf(">%u %u %u\n%s\n", id, S.library(id), R.len[id], R.bases.c_str() + R.start[id]);
  return 0;
}

// ovFile(ovFileFullWrite): buffers of (bufferSize / (lcm * 4)) * lcm uint32 with lcm = 20 * 24 and a
// 1 MiB bufferSize (ovStoreFile.C:107-121), each snappy-compressed (:190-207)
class OvbWriter {
 public:
  explicit OvbWriter(const char *path) : F_(fopen(path, "wb")) {
    if (!F_) die("cannot write", path);
  }
  ~OvbWriter() {
    flush();
    fclose(F_);
  }
  void add(const oic_overlap &o) {
    if (buf_.size() >= kMax) flush();
    buf_.push_back(o.a_iid);
    buf_.push_back(o.b_iid);
    for (int i = 0; i < 2; i++) {
      buf_.push_back((uint32_t)(o.dat[i] >> 32));
      buf_.push_back((uint32_t)(o.dat[i] & 0xffffffffu));
    }
  }

 private:
  static constexpr size_t kLcm = (4 * 1 + 16) * (4 * 2 + 16);
  static constexpr size_t kMax = (1024 * 1024 / (kLcm * 4)) * kLcm;
  void flush() {
    if (buf_.empty()) return;
    const std::string c = oic_snappy::compress((const uint8_t *)buf_.data(), buf_.size() * 4);
    const uint64_t bl = c.size();
    if (fwrite(&bl, 8, 1, F_) != 1 || fwrite(c.data(), 1, c.size(), F_) != c.size()) die("write failed");
    buf_.clear();
  }
  FILE *F_;
  std::vector<uint32_t> buf_;
};

static std::string counts_prefix(std::string name) {
  auto strip = [&](const char *suf) {
    const size_t k = strlen(suf);
    if (name.size() > k && name.compare(name.size() - k, k, suf) == 0) name.resize(name.size() - k);
  };
  strip(".WORKING");
  strip(".ovb");
  return name;
}

// --snappy-block IN OUT: compress IN into one raw snappy block (the .ovb buffer codec), no device
static int snappy_block(const char *in, const char *out) {
  FILE *F = fopen(in, "rb");
  if (!F) die("cannot open", in);
  std::string data;
  char buf[65536];
  size_t n;
  while ((n = fread(buf, 1, sizeof buf, F)) > 0) data.append(buf, n);
  fclose(F);
  const std::string c = oic_snappy::compress((const uint8_t *)data.data(), data.size());
  FILE *G = fopen(out, "wb");
  if (!G) die("cannot write", out);
  if (!c.empty() && fwrite(c.data(), 1, c.size(), G) != c.size()) die("write failed");
  fclose(G);
  return 0;
}

int main(int argc, char **argv) {
  if (argc == 4 && !strcmp(argv[1], "--snappy-block")) return snappy_block(argv[2], argv[3]);
  if (argc == 3 && !strcmp(argv[1], "--dump-store")) return dump_store(argv[2]);
  oic_params p;
  oic_params_default(&p);
  p.kmer = 0;
  p.schedule = OIC_SCHED_LEARNED;
  uint32_t bgn_hash = 1, end_hash = UINT32_MAX, bgn_ref = 1, end_ref = UINT32_MAX;
  const char *out_name = nullptr, *stats_name = nullptr, *reads_path = nullptr, *skip_path = nullptr;
  bool hpc = false;
  int err = 0;
  for (int arg = 1; arg < argc; arg++) {
    const char *a = argv[arg];
    auto next = [&]() -> const char * {
      if (arg + 1 >= argc) die("missing value after", a);
      return argv[++arg];
    };
    if (!strcmp(a, "-partial")) p.partial = 1;
    else if (!strcmp(a, "-h")) decode_range(next(), bgn_hash, end_hash);
    else if (!strcmp(a, "-r")) decode_range(next(), bgn_ref, end_ref);
    else if (!strcmp(a, "-H") || !strcmp(a, "-R")) { next(); die("library ranges are not supported:", a); }
    else if (!strcmp(a, "-k")) {             // a k-mer size, or a skip-k-mer file if one exists by that name
      const char *v = next();
      FILE *f = fopen(v, "r");
      if (f) { fclose(f); skip_path = v; }
      else p.kmer = (uint32_t)strtoul(v, nullptr, 10);
    } else if (!strcmp(a, "-l")) {
      const long v = strtol(next(), nullptr, 10);
      p.frag_olap_limit = (v < 1) ? UINT64_MAX : (uint64_t)v;
    } else if (!strcmp(a, "-m")) p.unique = 0;
    else if (!strcmp(a, "-u")) p.unique = 1;
    else if (!strcmp(a, "--hashbits")) p.hash_bits = (uint32_t)strtoull(next(), nullptr, 10);
    else if (!strcmp(a, "--hashdatalen")) p.hash_data_len = strtoull(next(), nullptr, 10);
    else if (!strcmp(a, "--hashload")) p.hash_load = atof(next());
    else if (!strcmp(a, "-o")) out_name = next();
    else if (!strcmp(a, "-s")) stats_name = next();
    else if (!strcmp(a, "-t")) p.n_threads = (uint32_t)strtoul(next(), nullptr, 10);
    else if (!strcmp(a, "--minlength")) p.min_olap = (int32_t)strtol(next(), nullptr, 10);
    else if (!strcmp(a, "--minkmers")) p.minkmers = 1;
    else if (!strcmp(a, "--maxerate")) p.max_erate = strtod(next(), nullptr);     // rounded to float by the engine
    else if (!strcmp(a, "--alignnoise")) p.align_noise = strtod(next(), nullptr);
    else if (!strcmp(a, "-z")) p.hopeless = 0;
    else if (!strcmp(a, "--hpc")) hpc = true;
    else if (!strcmp(a, "--schedule")) {
      const char *v = next();
      p.schedule = !strcmp(v, "plain") ? OIC_SCHED_PLAIN : !strcmp(v, "batched") ? OIC_SCHED_BATCHED : OIC_SCHED_LEARNED;
    } else if (!reads_path) reads_path = a;
    else { fprintf(stderr, "Unknown option '%s'\n", a); err++; }
  }
  if (p.kmer == 0) fprintf(stderr, "* No kmer length supplied; -k needed!\n"), err++;
  if (!out_name) fprintf(stderr, "ERROR:  No output file name specified\n"), err++;
  if (err || !reads_path) {
    fprintf(stderr, "USAGE:  %s [options] <seqStore|reads.fasta|reads.fastq>\n", argv[0]);
    return 1;
  }

  const bool store = oic_store::is_dir(reads_path);
  if (store && hpc) die("--hpc applies to FASTA/FASTQ input; a seqStore's homopolymerCompression marker decides:", reads_path);
  ReadSet R = store ? load_store(reads_path) : load_reads(reads_path, hpc);
  const uint32_t nreads = (uint32_t)(R.len.size() - 1);
  oic_ctx *ctx = nullptr;
  int rc = oic_create(&p, &ctx);
  if (rc != OIC_OK) {
    fprintf(stderr, "overlapInCore: oic_create failed (%d): %s\n", rc, ctx ? oic_last_error(ctx) : "no usable GPU or bad parameters");
    return 1;
  }
  if (skip_path) {
    FILE *F = fopen(skip_path, "rb");
    if (!F) die("cannot open", skip_path);
    std::string text;
    char buf[65536];
    size_t n;
    while ((n = fread(buf, 1, sizeof buf, F)) > 0) text.append(buf, n);
    fclose(F);
    if ((rc = oic_set_skip_kmers(ctx, text.data(), text.size())) != OIC_OK) die(oic_last_error(ctx));
  }
  const oic_reads rd = {R.bases.data(), R.start.data(), R.len.data(), nreads};
  if ((rc = oic_load_reads(ctx, &rd)) != OIC_OK) die(oic_last_error(ctx));

  // A learned schedule needs a model; without one the engine orders work by its static estimate.
  std::vector<oic_overlap> recs(1 << 20);
  oic_sink sink = {};
  sink.records = recs.data();
  sink.capacity = recs.size();
  rc = oic_run(ctx, bgn_hash, end_hash, bgn_ref, end_ref, &sink);
  if (rc == OIC_ECAPACITY) {
    // the job ran to its end; the engine holds the records that did not fit
    recs.resize(sink.records_needed);
    sink.records = recs.data();
    sink.capacity = recs.size();
    rc = oic_fetch_pending(ctx, &sink);
  }
  if (rc != OIC_OK) die(oic_last_error(ctx));

  {
    OvbWriter W(out_name);
    for (uint64_t i = 0; i < sink.count; i++) W.add(recs[i]);
  }
  {
    std::vector<uint32_t> opr((size_t)nreads + 1, 0);
    for (uint64_t i = 0; i < sink.count; i++) {
      opr[recs[i].a_iid]++;
      opr[recs[i].b_iid]++;
    }
    const std::string oc = counts_prefix(out_name) + ".oc";
    FILE *F = fopen(oc.c_str(), "wb");
    if (!F) die("cannot write", oc.c_str());
    const uint64_t n = sink.count;
    const uint32_t m = nreads + 1;
    fwrite(&n, 8, 1, F);
    fwrite(&m, 4, 1, F);
    fwrite(opr.data(), 4, opr.size(), F);
    fclose(F);
  }
  FILE *S = stderr;
  if (stats_name) {
    errno = 0;
    S = fopen(stats_name, "w");
    if (!S) {
      fprintf(stderr, "WARNING: failed to open '%s' for writing: %s\n", stats_name, strerror(errno));
      S = stderr;
    }
  }
  const char *names[8] = {" Kmer hits without olaps", "    Kmer hits with olaps", "  Multiple overlaps/pair",
                          " Total overlaps produced", "      Contained overlaps", "       Dovetail overlaps",
                          "Rejected by short window", " Rejected by long window"};
  for (int i = 0; i < 8; i++) fprintf(S, "%s = %" PRId64 "\n", names[i], (int64_t)sink.stats[i]);
  if (S != stderr) fclose(S);
  oic_destroy(ctx);
  fprintf(stderr, "Bye.\n");
  return 0;
}
