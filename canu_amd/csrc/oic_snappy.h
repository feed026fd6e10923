// oic_snappy.h -- a compressor for snappy's raw block format, as ovFile writes its buffers
// (stores/ovStoreFile.C:190-207: snappy::RawCompress of each buffer, preceded by its uint64 length).
//
// Written from the published format description: a varint of the uncompressed length, then
// elements whose tag's two low bits select a literal (00) or a copy with a 1-, 2- or 4-byte offset
// (01, 10, 11).  Any sequence of valid elements decodes, so the bytes differ from Google's
// compressor; readers (ovFile's snappy::RawUncompress, canu_amd/ovfile.py) see the same data.
//
// Matching: greedy, 4-byte hashes into a 16384-entry table, within 64 KiB blocks (offsets fit a
// 2-byte copy), matches extended up to the block end.
#pragma once

#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

namespace oic_snappy {

inline void put_varint(std::string &out, uint32_t v) {
  while (v >= 0x80) {
    out.push_back((char)(v | 0x80));
    v >>= 7;
  }
  out.push_back((char)v);
}

inline void emit_literal(std::string &out, const uint8_t *p, size_t n) {
  while (n > 0) {
    const size_t len = n < 65536 ? n : 65536;
    const uint32_t l1 = (uint32_t)(len - 1);
    if (l1 < 60) {
      out.push_back((char)(l1 << 2));
    } else {
      const int nb = l1 < 0x100 ? 1 : l1 < 0x10000 ? 2 : l1 < 0x1000000 ? 3 : 4;
      out.push_back((char)((59 + nb) << 2));
      for (int i = 0; i < nb; i++) out.push_back((char)((l1 >> (8 * i)) & 0xff));
    }
    out.append((const char *)p, len);
    p += len;
    n -= len;
  }
}

inline void emit_copy1_or_2(std::string &out, uint32_t offset, uint32_t len) {   // 4 <= len <= 64
  if (len <= 11 && offset < 2048) {
    out.push_back((char)(1 | ((len - 4) << 2) | ((offset >> 8) << 5)));
    out.push_back((char)(offset & 0xff));
  } else {
    out.push_back((char)(2 | ((len - 1) << 2)));
    out.push_back((char)(offset & 0xff));
    out.push_back((char)((offset >> 8) & 0xff));
  }
}

inline void emit_copy(std::string &out, uint32_t offset, uint32_t len) {
  while (len >= 68) { emit_copy1_or_2(out, offset, 64); len -= 64; }
  if (len > 64) { emit_copy1_or_2(out, offset, 60); len -= 60; }
  emit_copy1_or_2(out, offset, len);
}

inline uint32_t load32(const uint8_t *p) {
  uint32_t v;
  memcpy(&v, p, 4);
  return v;
}

// compress n bytes of in into a raw snappy block
inline std::string compress(const uint8_t *in, size_t n) {
  std::string out;
  out.reserve(n + n / 6 + 32);
  put_varint(out, (uint32_t)n);
  constexpr size_t kBlock = 65536;
  constexpr int kHashBits = 14;
  std::vector<int32_t> table(1u << kHashBits);
  for (size_t b0 = 0; b0 < n; b0 += kBlock) {
    const size_t bn = (n - b0) < kBlock ? (n - b0) : kBlock;
    const uint8_t *blk = in + b0;
    std::fill(table.begin(), table.end(), -1);
    size_t lit = 0, i = 0;
    while (i + 4 <= bn) {
      const uint32_t v = load32(blk + i);
      const uint32_t h = (v * 0x1e35a7bdu) >> (32 - kHashBits);
      const int32_t cand = table[h];
      table[h] = (int32_t)i;
      if (cand >= 0 && load32(blk + cand) == v) {
        size_t len = 4;
        while (i + len < bn && blk[cand + len] == blk[i + len]) len++;
        if (i > lit) emit_literal(out, blk + lit, i - lit);
        emit_copy(out, (uint32_t)(i - (size_t)cand), (uint32_t)len);
        i += len;
        lit = i;
      } else {
        i++;
      }
    }
    if (bn > lit) emit_literal(out, blk + lit, bn - lit);
  }
  return out;
}

}  // namespace oic_snappy
