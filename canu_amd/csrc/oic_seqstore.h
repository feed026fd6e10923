// oic_seqstore.h -- the reads of a canu seqStore, as overlapInCore receives them.
//
// Layout of a store directory (version 9), as canu's store code reads it:
//   info           tagged objects: 4-byte tag, uint32 payload length, payload (stores/sqStore.C:
//                  sqStoreInfo::readInfo, stores/sqStoreInfo.C:265-298): MAGC, VERS, LSIZ, RSIZ,
//                  MLB , LNS , MRB , MRLB, NLIB, NREA, NBLO, READ[16], BASE[16];
//   reads          sqReadMeta[NREA + 1], 16 bytes each (stores/sqRead.H:306-323): word 0 = readID:30
//                  | libraryID:12 | assignment:15 | assignmentScore:7, word 1 = unused:8 | mSegm:16 |
//                  mByte:40, least significant bits first;
//   reads-rawu, reads-rawc, reads-coru, reads-corc
//                  sqReadSeq[NREA + 1], 12 bytes each (stores/sqRead.H:246-256): valid:1 | unused:1 |
//                  length:30, ignoreU:1 | ignoreT:1 | clearBgn:30, trimmed:1 | unused:1 | clearEnd:30;
//   blobs.NNNN     per read at byte mByte of blob file mSegm: a "BLOB" chunk whose payload is a run of
//                  chunks (4-byte tag, uint32 length, data), NAME then 2SQR/3SQR/USQR (raw bases) and
//                  2SQC/3SQC/USQC (corrected bases) (stores/sqReadData.C:48-91,
//                  stores/sqReadDataWriter.C:226-243);
//   homopolymerCompression   an empty marker file: reads are handed over homopolymer-compressed.
//
// Which version of the reads overlapInCore sees (stores/sqStoreConstructor.C:141-166, with no
// version set by the caller): the last of raw, raw-trimmed, corrected, corrected-trimmed that the
// info file counts any reads of, compressed when the marker exists.  A read's length is 0 when its
// sequence is absent or ignored (stores/sqStore.H:407-420); its bases are the decoded raw or
// corrected sequence, homopolymer-compressed if so selected, then cut to the clear range if trimmed
// (stores/sqRead.H:399-460, stores/sqCache.C:262-318).
//
// Three pieces live in meryl-utility, which canu does not ship: the tagged-chunk framing, the 2-bit
// code and the layout of the 3-bit code.  Framing is read exactly as canu's own chunk walk reads it
// (tag, length, payload; the next chunk starts 8 + length bytes on).  The 2-bit code is read as
// A=0, C=1, G=2, T=3 with the first base in the two high bits of a byte; nothing in canu pins that
// order.  3-bit chunks (reads with N) are refused with an error; 8-bit chunks hold the bases as bytes.
#pragma once

#include <sys/stat.h>

#include <cstdint>
#include <cstdio>
#include <cstring>
#include <map>
#include <string>
#include <vector>

namespace oic_store {

constexpr uint64_t kMagic = 0x5145533a756e6163ull;    // "canu:SEQ", version 9 stores
constexpr uint64_t kMagic8 = 0x504b473a756e6162ull;   // "banu:GKP", version 8 and older
constexpr uint64_t kVersion = 9;

constexpr uint32_t kRaw = 0x1, kCorrected = 0x2, kCompressed = 0x4, kTrimmed = 0x8;

// the 2-bit code's bases, by value
constexpr char kTwoBitBases[4] = {'A', 'C', 'G', 'T'};

struct SeqMeta {   // sqReadSeq
  uint32_t w[3];
  bool valid() const { return w[0] & 1u; }
  uint32_t length() const { return w[0] >> 2; }
  bool ignore_u() const { return w[1] & 1u; }
  bool ignore_t() const { return (w[1] & 3u) != 0; }   // ignoring the untrimmed read ignores the trimmed one
  uint32_t clear_bgn() const { return w[1] >> 2; }
  bool trimmed() const { return w[2] & 1u; }
  uint32_t clear_end() const { return w[2] >> 2; }
};
static_assert(sizeof(SeqMeta) == 12, "sqReadSeq is three 32-bit words");

struct ReadMeta {   // sqReadMeta
  uint64_t w0, w1;
  uint32_t read_id() const { return (uint32_t)(w0 & ((1ull << 30) - 1)); }
  uint32_t library_id() const { return (uint32_t)((w0 >> 30) & 0xfff); }
  uint32_t segment() const { return (uint32_t)((w1 >> 8) & 0xffff); }
  uint64_t byte() const { return w1 >> 24; }
};
static_assert(sizeof(ReadMeta) == 16, "sqReadMeta is two 64-bit words");

inline bool is_dir(const std::string &p) {
  struct stat st;
  return stat(p.c_str(), &st) == 0 && S_ISDIR(st.st_mode);
}

inline bool is_file(const std::string &p) {
  struct stat st;
  return stat(p.c_str(), &st) == 0 && S_ISREG(st.st_mode);
}

inline bool slurp(const std::string &path, std::string &out) {
  FILE *F = fopen(path.c_str(), "rb");
  if (!F) return false;
  out.clear();
  char buf[1 << 16];
  size_t n;
  while ((n = fread(buf, 1, sizeof buf, F)) > 0) out.append(buf, n);
  fclose(F);
  return true;
}

// keep the first base of each maximal run of equal bases
inline void homopoly_compress(std::string &s) {
  size_t o = 0;
  for (size_t i = 0; i < s.size(); i++)
    if (o == 0 || s[o - 1] != s[i]) s[o++] = s[i];
  s.resize(o);
}

class Store {
 public:
  Store() = default;
  Store(const Store &) = delete;              // owns the open blob files
  Store &operator=(const Store &) = delete;

  std::string path;
  uint32_t nreads = 0;     // last read ID
  uint32_t nlibs = 0;
  uint32_t nblobs = 0;
  uint32_t which = 0;      // version handed over: kRaw or kCorrected, plus kCompressed / kTrimmed
  uint64_t counts[16] = {};
  std::vector<ReadMeta> meta;
  std::vector<SeqMeta> seq[4];   // rawu, rawc, coru, corc

  bool open(const std::string &dir, std::string &err) {
    path = dir;
    if (!is_dir(dir)) return fail(err, "not a seqStore directory: " + dir);
    std::string info;
    if (!slurp(dir + "/info", info)) return fail(err, "cannot read " + dir + "/info");
    uint64_t magic = 0, version = 0;
    if (info.size() >= 8) memcpy(&magic, info.data(), 8);   // version 8 info files open with the bare magic
    if (magic == kMagic8) return fail(err, "seqStore version 8 or older is not supported: " + dir);
    std::map<std::string, std::string> obj;
    for (size_t p = 0; p + 8 <= info.size();) {
      uint32_t len;
      memcpy(&len, info.data() + p + 4, 4);
      if (p + 8 + len > info.size()) return fail(err, "truncated object in " + dir + "/info");
      obj[info.substr(p, 4)] = info.substr(p + 8, len);
      p += 8 + (size_t)len;
    }
    if (!get(obj, "MAGC", magic) || magic != kMagic) return fail(err, "not a canu seqStore (magic): " + dir);
    if (!get(obj, "VERS", version) || version != kVersion)
      return fail(err, "seqStore version " + std::to_string(version) + " is not supported (need 9): " + dir);
    // the parameters canu checks before using a store (stores/sqStoreInfo.C:58-108)
    const struct { const char *tag; uint32_t want; } checks[] = {
        {"LSIZ", 176}, {"RSIZ", 152}, {"MLB ", 6}, {"LNS ", 128}, {"MRB ", 37}, {"MRLB", 21}};
    for (const auto &c : checks) {
      uint32_t v = 0;
      if (!get(obj, c.tag, v) || v != c.want)
        return fail(err, std::string("seqStore parameter ") + c.tag + " = " + std::to_string(v) + ", expected " +
                             std::to_string(c.want) + ": " + dir);
    }
    if (!get(obj, "NLIB", nlibs) || !get(obj, "NREA", nreads) || !get(obj, "NBLO", nblobs))
      return fail(err, "seqStore info lacks NLIB/NREA/NBLO: " + dir);
    auto rd = obj.find("READ");
    if (rd == obj.end() || rd->second.size() != sizeof counts) return fail(err, "seqStore info lacks READ: " + dir);
    memcpy(counts, rd->second.data(), sizeof counts);
    if (!load_array(dir + "/reads", meta, err)) return false;
    const char *names[4] = {"/reads-rawu", "/reads-rawc", "/reads-coru", "/reads-corc"};
    for (int i = 0; i < 4; i++)
      if (!load_array(dir + names[i], seq[i], err)) return false;
    // the most recent version with reads, then compression if the store is marked for it
    uint32_t mr = 0;
    if (counts[kRaw] > 0) mr = kRaw;
    if (counts[kRaw | kTrimmed] > 0) mr = kRaw | kTrimmed;
    if (counts[kCorrected] > 0) mr = kCorrected;
    if (counts[kCorrected | kTrimmed] > 0) mr = kCorrected | kTrimmed;
    if (mr == 0) return fail(err, "seqStore has neither raw nor corrected reads: " + dir);
    which = mr;
    if (is_file(dir + "/homopolymerCompression")) which |= kCompressed;
    return true;
  }

  const SeqMeta &seq_meta(uint32_t id, uint32_t w) const {
    const int k = ((w & kCorrected) ? 2 : 0) + ((w & kCompressed) ? 1 : 0);
    return seq[k][id];
  }

  // sqStore_getReadLength(id, which)
  uint32_t length(uint32_t id) const {
    if (id == 0 || id > nreads) return 0;
    const SeqMeta &s = seq_meta(id, which);
    if (which & kTrimmed)
      return (!s.trimmed() || !s.valid() || s.ignore_t()) ? 0 : s.clear_end() - s.clear_bgn();
    return (!s.valid() || s.ignore_u()) ? 0 : s.length();
  }

  uint32_t library(uint32_t id) const { return meta[id].library_id(); }

  // the bases overlapInCore indexes and searches for read id (uppercase, as stored); length(id) must be > 0
  bool sequence(uint32_t id, std::string &out, std::string &err) {
    const ReadMeta &m = meta[id];
    std::string blob;
    if (!fetch_blob(m, blob, err)) return false;
    const bool cor = (which & kCorrected) != 0;
    const uint32_t full = seq_meta(id, cor ? kCorrected : kRaw).length();   // untrimmed, uncompressed
    bool found = false;
    for (size_t p = 0; p + 8 <= blob.size();) {
      uint32_t len;
      memcpy(&len, blob.data() + p + 4, 4);
      if (p + 8 + len > blob.size()) return fail(err, "truncated chunk in read " + std::to_string(id));
      const char *tag = blob.data() + p;
      const uint8_t *data = (const uint8_t *)blob.data() + p + 8;
      if (tag[1] == 'S' && tag[2] == 'Q' && tag[3] == (cor ? 'C' : 'R')) {
        if (tag[0] == '2') {
          if ((uint64_t)len * 4 < full) return fail(err, "short 2-bit chunk in read " + std::to_string(id));
          out.resize(full);
          for (uint32_t i = 0; i < full; i++) out[i] = kTwoBitBases[(data[i >> 2] >> (6 - 2 * (i & 3))) & 3];
        } else if (tag[0] == 'U') {
          if (len < full) return fail(err, "short 8-bit chunk in read " + std::to_string(id));
          out.assign((const char *)data, full);
        } else if (tag[0] == '3') {
          return fail(err, "read " + std::to_string(id) +
                               " is stored with the 3-bit code (bases other than ACGT); not supported");
        } else {
          return fail(err, "unknown chunk in read " + std::to_string(id));
        }
        found = true;
      }
      p += 8 + (size_t)len;
    }
    if (!found) return fail(err, "read " + std::to_string(id) + " has no sequence chunk");
    if (which & kCompressed) homopoly_compress(out);
    if (which & kTrimmed) {
      const SeqMeta &s = seq_meta(id, which);
      if (s.clear_end() > out.size() || s.clear_bgn() > s.clear_end())
        return fail(err, "clear range outside read " + std::to_string(id));
      out = out.substr(s.clear_bgn(), s.clear_end() - s.clear_bgn());
    }
    return true;
  }

 private:
  std::map<uint32_t, FILE *> blobs_;

  static bool fail(std::string &err, const std::string &msg) {
    err = msg;
    return false;
  }

  template <typename T>
  static bool get(const std::map<std::string, std::string> &obj, const char *tag, T &v) {
    auto it = obj.find(tag);
    if (it == obj.end() || it->second.size() != sizeof(T)) return false;
    memcpy(&v, it->second.data(), sizeof(T));
    return true;
  }

  template <typename T>
  bool load_array(const std::string &file, std::vector<T> &v, std::string &err) {
    std::string raw;
    if (!slurp(file, raw)) return fail(err, "cannot read " + file);
    if (raw.size() < sizeof(T) * ((size_t)nreads + 1))
      return fail(err, file + " holds " + std::to_string(raw.size() / sizeof(T)) + " records, expected " +
                           std::to_string((size_t)nreads + 1));
    v.resize((size_t)nreads + 1);
    memcpy(v.data(), raw.data(), sizeof(T) * v.size());
    return true;
  }

  bool fetch_blob(const ReadMeta &m, std::string &blob, std::string &err) {
    const uint32_t seg = m.segment();
    FILE *&F = blobs_[seg];
    if (!F) {
      char name[32];
      snprintf(name, sizeof name, "/blobs.%04u", seg);
      F = fopen((path + name).c_str(), "rb");
      if (!F) return fail(err, "cannot open " + path + name);
    }
    char head[8];
    uint32_t len = 0;
    if (fseeko(F, (off_t)m.byte(), SEEK_SET) != 0 || fread(head, 1, 8, F) != 8)
      return fail(err, "cannot read blob of read " + std::to_string(m.read_id()));
    memcpy(&len, head + 4, 4);
    if (memcmp(head, "BLOB", 4) != 0) return fail(err, "expected BLOB for read " + std::to_string(m.read_id()));
    blob.resize(len);
    if (len && fread(&blob[0], 1, len, F) != len)
      return fail(err, "truncated blob of read " + std::to_string(m.read_id()));
    return true;
  }

 public:
  ~Store() {
    for (auto &kv : blobs_)
      if (kv.second) fclose(kv.second);
  }
};

}  // namespace oic_store
