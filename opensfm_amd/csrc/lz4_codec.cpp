// LZ4 block codec for the depthmap files (the reference writes
// np.savez(...) wrapped in lz4.frame, opensfm/dataset.py:1080-1149).  The
// frame layer lives in opensfm_amd/depthmap_io.py; these two functions are
// the block format of the public LZ4 specification: sequences of a token
// (literal length << 4 | match length - 4, 15 meaning "more bytes follow"),
// the literals, a 16-bit little-endian offset and the match; the last
// sequence carries literals only.
#include <stdint.h>
#include <string.h>

#include "../../include/opensfm_amd.h"
#include "pm_math.h"

using namespace pmhost;

namespace {

inline uint32_t read32(const uint8_t* p) {
  uint32_t v;
  memcpy(&v, p, 4);
  return v;
}

inline uint32_t hash4(uint32_t v) { return (v * 2654435761u) >> 16; }  // 16-bit table index

bool put_length(uint8_t*& op, const uint8_t* oend, int64_t len) {
  while (len >= 255) {
    if (op >= oend) return false;
    *op++ = 255;
    len -= 255;
  }
  if (op >= oend) return false;
  *op++ = (uint8_t)len;
  return true;
}

}  // namespace

extern "C" {

int pm_lz4_block_decompress(const uint8_t* src, int64_t src_size, uint8_t* out, int64_t out_pos,
                            int64_t out_cap, int64_t* new_pos) {
  if (!src || !out || !new_pos || src_size < 0 || out_pos < 0 || out_pos > out_cap)
    return fail(PM_ERR_INVALID, "lz4: bad arguments");
  const uint8_t* ip = src;
  const uint8_t* const iend = src + src_size;
  int64_t op = out_pos;
  while (ip < iend) {
    const uint8_t token = *ip++;
    int64_t lit = token >> 4;
    if (lit == 15) {
      uint8_t b;
      do {
        if (ip >= iend) return fail(PM_ERR_INVALID, "lz4: truncated literal length");
        b = *ip++;
        lit += b;
      } while (b == 255);
    }
    if (lit > iend - ip) return fail(PM_ERR_INVALID, "lz4: literals past the block");
    if (lit > out_cap - op) return fail(PM_ERR_INVALID, "lz4: output too small");
    if (lit > 0) memcpy(out + op, ip, (size_t)lit);
    ip += lit;
    op += lit;
    if (ip == iend) break;  // last sequence: literals only
    if (iend - ip < 2) return fail(PM_ERR_INVALID, "lz4: truncated offset");
    const int64_t off = (int64_t)ip[0] | ((int64_t)ip[1] << 8);
    ip += 2;
    if (off == 0 || off > op) return fail(PM_ERR_INVALID, "lz4: offset outside the output");
    int64_t mlen = (token & 15);
    if (mlen == 15) {
      uint8_t b;
      do {
        if (ip >= iend) return fail(PM_ERR_INVALID, "lz4: truncated match length");
        b = *ip++;
        mlen += b;
      } while (b == 255);
    }
    mlen += 4;
    if (mlen > out_cap - op) return fail(PM_ERR_INVALID, "lz4: output too small");
    const uint8_t* m = out + op - off;
    uint8_t* d = out + op;
    if (off >= mlen) {
      memcpy(d, m, (size_t)mlen);
    } else {
      for (int64_t k = 0; k < mlen; ++k) d[k] = m[k];  // overlapping copy
    }
    op += mlen;
  }
  *new_pos = op;
  return PM_OK;
}

int pm_lz4_block_compress(const uint8_t* src, int64_t n, uint8_t* dst, int64_t cap,
                          int64_t* written) {
  if ((!src && n) || !dst || !written || n < 0 || cap < 0)
    return fail(PM_ERR_INVALID, "lz4: bad arguments");
  if (n > 0x7fffffff) return fail(PM_ERR_INVALID, "lz4: blocks are limited to 2^31 - 1 bytes");
  static thread_local int32_t table[1 << 16];
  for (auto& t : table) t = -1;
  uint8_t* op = dst;
  const uint8_t* const oend = dst + cap;
  int64_t ip = 0, anchor = 0;
  const int64_t mflimit = n - 12;   // a match starts at least 12 bytes before the end
  const int64_t matchlimit = n - 5; // the last 5 bytes are literals
  auto emit = [&](int64_t lit_from, int64_t lit_len, int64_t offset, int64_t mlen) -> bool {
    if (op >= oend) return false;
    uint8_t* token = op++;
    const int64_t ml = mlen > 0 ? mlen - 4 : 0;
    *token = (uint8_t)(((lit_len >= 15 ? 15 : lit_len) << 4) | (ml >= 15 ? 15 : ml));
    if (lit_len >= 15 && !put_length(op, oend, lit_len - 15)) return false;
    if (lit_len > oend - op) return false;
    if (lit_len > 0) memcpy(op, src + lit_from, (size_t)lit_len);
    op += lit_len;
    if (mlen == 0) return true;
    if (oend - op < 2) return false;
    *op++ = (uint8_t)(offset & 255);
    *op++ = (uint8_t)(offset >> 8);
    if (ml >= 15 && !put_length(op, oend, ml - 15)) return false;
    return true;
  };
  while (ip < mflimit) {
    const uint32_t seq = read32(src + ip);
    const uint32_t h = hash4(seq);
    const int64_t ref = table[h];
    table[h] = (int32_t)ip;
    if (ref >= 0 && ip - ref <= 65535 && read32(src + ref) == seq) {
      int64_t len = 4;
      while (ip + len < matchlimit && src[ref + len] == src[ip + len]) ++len;
      if (!emit(anchor, ip - anchor, ip - ref, len))
        return fail(PM_ERR_INVALID, "lz4: output too small");
      ip += len;
      anchor = ip;
    } else {
      ++ip;
    }
  }
  if (!emit(anchor, n - anchor, 0, 0)) return fail(PM_ERR_INVALID, "lz4: output too small");
  *written = op - dst;
  return PM_OK;
}

}  // extern "C"
