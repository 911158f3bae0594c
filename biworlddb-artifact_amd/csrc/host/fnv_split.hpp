// fnv_split.hpp -- FNV-1a 64 split into independent blocks (shared by the device kernels in
// hash_kernels.hip and the CPU check of them).
//
// FNV-1a: h <- (h ^ b) * P per byte, P = 0x100000001b3.  XOR with a byte changes only the low 8
// bits of h, so h ^ b = h + d with d = (l ^ b) - l, l = h & 0xFF, a small signed number; and the
// low byte of (h ^ b) * P depends only on the low bytes: l' = ((l ^ b) * 0xB3) & 0xFF.  Hence:
//   * the low-byte sequence is an 8-bit chain, and over a block of bytes it is a map of the 256
//     possible entry values -- computable for all 256 at once, composable across blocks;
//   * once a block's entry low byte is known, its d values are, and
//       h_end = h_start * P^len + S,  S = sum over the block's bytes of d_i * P^(end - i)
//     (Horner from 0: S <- (S + d_i) * P), so the whole hash is
//       h_n = h_0 * P^n + sum over blocks j of S_j * P^(n - end_j)   (mod 2^64),
//     where every block but the last is kFnvBlock long, so the sum is a polynomial in
//     Q = P^kFnvBlock (evaluated by Horner over ranges of blocks, fnv_range_sum).
#pragma once
#include <stddef.h>
#include <stdint.h>

#ifndef BWT_HD
#ifdef __HIPCC__
#define BWT_HD __host__ __device__
#else
#define BWT_HD
#endif
#endif

namespace bwg {

constexpr uint64_t kFnvPrime = 0x100000001b3ULL;
// The start value of the reference's FNV-1a (src/common.hpp:396); the published offset basis
// 14695981039346656037 has one more digit.
constexpr uint64_t kFnvStart = 1469598103934665603ULL;
// Bytes per block of the split.
constexpr size_t kFnvBlock = 4096;

// One byte through the low-byte chain.
BWT_HD inline uint32_t fnv_low_step(uint32_t l, uint32_t b) { return ((l ^ b) * 0xB3u) & 0xFFu; }

// Two low-byte chains packed in one word (bits 0-7 and 16-23), one byte b through both: an 8-bit
// value times 0xB3 stays below 2^16, so one 32-bit multiply serves both halves.
BWT_HD inline uint32_t fnv_low_step2(uint32_t x, uint32_t b) { return ((x ^ (b * 0x10001u)) * 0xB3u) & 0x00FF00FFu; }

// base^k mod 2^64.
BWT_HD inline uint64_t pow64(uint64_t base, uint64_t k) {
  uint64_t r = 1;
  while (k) {
    if (k & 1) r *= base;
    base *= base;
    k >>= 1;
  }
  return r;
}

// P^k mod 2^64.
BWT_HD inline uint64_t fnv_pow(uint64_t k) { return pow64(kFnvPrime, k); }

// The block sum of blocks [a, b) among the m full blocks before the last block, weighted for the
// total: with Q = P^kFnvBlock, sum_{j<m} S_j P^(n - end_j) = P^len_last * sum_{j<m} S_j Q^(m-1-j),
// and a range's share of the inner sum is (Horner over the range) * Q^(m - b).
BWT_HD inline uint64_t fnv_range_sum(const uint64_t* S, int64_t a, int64_t b, int64_t m) {
  if (a >= b) return 0;
  const uint64_t Q = fnv_pow(kFnvBlock);
  uint64_t h = 0;
  for (int64_t j = a; j < b; ++j) h = h * Q + S[j];
  return h * pow64(Q, (uint64_t)(m - b));
}

// The nb low bytes of w (little-endian order) through the chain from entry low byte l, adding
// their terms to the block polynomial s.
BWT_HD inline void fnv_poly_word(uint32_t w, int nb, uint32_t& l, uint64_t& s) {
  for (int k = 0; k < nb; ++k) {
    const uint32_t t = l ^ ((w >> (8 * k)) & 0xFFu);
    s = (s + (uint64_t)((int64_t)t - (int64_t)l)) * kFnvPrime;
    l = (t * 0xB3u) & 0xFFu;
  }
}

}  // namespace bwg
