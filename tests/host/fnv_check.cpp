// fnv_check.cpp -- test infrastructure: the block split of FNV-1a 64 (csrc/host/fnv_split.hpp,
// the arithmetic of the hash_kernels.hip kernels) run on the CPU step by step as the kernels run
// it -- 256-entry low-byte maps per 4 KiB block (two entries per packed word), composition 64 at
// a time, the walk down the levels, block polynomials, the final sum -- against the sequential
// FNV-1a over the same bytes.
//   fnv_check N SEED   (N bytes; prints {"split": "...", "sequential": "...", "levels": L})
//   fnv_check --file PATH   (the sequential FNV-1a of a file's bytes; prints {"sequential": "..."})
#include <cinttypes>
#include <cstdio>
#include <cstdlib>
#include <algorithm>
#include <cstring>
#include <vector>

#include "fnv_split.hpp"

using namespace bwg;

int main(int argc, char** argv) {
  if (argc < 3) return 2;
  if (!std::strcmp(argv[1], "--file")) {
    FILE* f = std::fopen(argv[2], "rb");
    if (!f) return 4;
    uint64_t h = kFnvStart;
    int c;
    while ((c = std::fgetc(f)) != EOF) h = (h ^ uint64_t(c)) * kFnvPrime;
    std::fclose(f);
    std::printf("{\"sequential\": \"%016" PRIx64 "\"}\n", h);
    return 0;
  }
  const size_t n = std::strtoull(argv[1], nullptr, 10);
  uint64_t z = std::strtoull(argv[2], nullptr, 10) * 0x9E3779B97F4A7C15ULL + 1;
  std::vector<unsigned char> b(n);
  for (size_t i = 0; i < n; ++i) {
    z ^= z << 13; z ^= z >> 7; z ^= z << 17;
    b[i] = (unsigned char)(i % 97 == 0 ? 0 : z >> 56);
  }
  uint64_t seq = kFnvStart;
  for (size_t i = 0; i < n; ++i) seq = (seq ^ b[i]) * kFnvPrime;

  const size_t nblk = (n + kFnvBlock - 1) / kFnvBlock;
  // 1. maps, two entry values per packed word as the kernel does
  std::vector<std::vector<unsigned char>> lv(1, std::vector<unsigned char>(nblk * 256));
  for (size_t j = 0; j < nblk; ++j) {
    const size_t b0 = j * kFnvBlock, len = std::min(kFnvBlock, n - b0);
    for (uint32_t t = 0; t < 128; ++t) {
      uint32_t x = t | ((t + 128) << 16);
      for (size_t k = 0; k < len; ++k) x = fnv_low_step2(x, b[b0 + k]);
      lv[0][j * 256 + t] = (unsigned char)(x & 0xFF);
      lv[0][j * 256 + t + 128] = (unsigned char)(x >> 16);
    }
  }
  // 2. compose 64 at a time until one group is left
  std::vector<size_t> cnt(1, nblk);
  while (cnt.back() > 64) {
    const size_t c = (cnt.back() + 63) / 64;
    std::vector<unsigned char> out(c * 256);
    for (size_t g = 0; g < c; ++g)
      for (uint32_t v = 0; v < 256; ++v) {
        uint32_t x = v;
        for (size_t j = g * 64; j < std::min(cnt.back(), g * 64 + 64); ++j) x = lv.back()[j * 256 + x];
        out[g * 256 + v] = (unsigned char)x;
      }
    lv.push_back(out);
    cnt.push_back(c);
  }
  // 3. walk down
  std::vector<unsigned char> parent;
  for (int k = int(lv.size()) - 1; k >= 0; --k) {
    std::vector<unsigned char> st(cnt[size_t(k)]);
    for (size_t g = 0; g * 64 < cnt[size_t(k)]; ++g) {
      uint32_t x = parent.empty() ? uint32_t(kFnvStart & 0xFF) : parent[g];
      for (size_t j = g * 64; j < std::min(cnt[size_t(k)], g * 64 + 64); ++j) {
        st[j] = (unsigned char)x;
        x = lv[size_t(k)][j * 256 + x];
      }
    }
    parent = st;
  }
  // 4. block polynomials (words of 4 bytes, then the tail)
  std::vector<uint64_t> S(nblk);
  for (size_t j = 0; j < nblk; ++j) {
    const size_t b0 = j * kFnvBlock, len = std::min(kFnvBlock, n - b0);
    uint32_t l = parent[j];
    uint64_t s = 0;
    size_t k = 0;
    for (; k + 4 <= len; k += 4) {
      const uint32_t w = b[b0 + k] | (b[b0 + k + 1] << 8) | (b[b0 + k + 2] << 16) | (uint32_t(b[b0 + k + 3]) << 24);
      fnv_poly_word(w, 4, l, s);
    }
    for (; k < len; ++k) fnv_poly_word(b[b0 + k], 1, l, s);
    S[j] = s;
  }
  // 5. the sum as the finishing workgroup forms it: 256 ranges of the full blocks, then the last
  uint64_t h = kFnvStart * fnv_pow(n);
  if (nblk > 0) {
    const int64_t m = int64_t(nblk) - 1, per = (m + 255) / 256;
    uint64_t inner = 0;
    for (int64_t t = 0; t < 256; ++t) inner += fnv_range_sum(S.data(), t * per, std::min(t * per + per, m), m);
    h += S[size_t(m)] + fnv_pow(n - size_t(m) * kFnvBlock) * inner;
    // and the direct form, block by block
    uint64_t d = kFnvStart * fnv_pow(n);
    for (size_t j = 0; j < nblk; ++j) d += S[j] * fnv_pow(n - std::min(n, (j + 1) * kFnvBlock));
    if (d != h) h = ~h;  // the two forms of the sum must agree
  }
  std::printf("{\"split\": \"%016" PRIx64 "\", \"sequential\": \"%016" PRIx64 "\", \"levels\": %zu}\n", h, seq,
              lv.size());
  return 0;
}
