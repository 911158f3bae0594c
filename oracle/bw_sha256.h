/* bw_sha256.h -- FIPS 180-4 SHA-256, written for the oracle (test infrastructure only).
 *
 * The reference hashes every complete answer vector with SHA-256 over the little-endian
 * int64 bytes and prints lowercase hex (src/sha_fast.hpp:27-43; software path src/sha256.hpp;
 * the oracle's own copy src/oracle.cpp:43-165).  This is an independent implementation from
 * the FIPS 180-4 text; it is pinned by the NIST known-answer vectors in tests/test_oracle.py
 * and by SHA equality with OpenSSL (the product's extract path).
 */
#ifndef BW_SHA256_H
#define BW_SHA256_H
#include <stddef.h>
#include <stdint.h>
#include <string.h>

typedef struct {
  uint32_t h[8];
  uint64_t nbytes;
  uint8_t blk[64];
  size_t used;
} bw_sha256_ctx;

static const uint32_t bw_sha256_k[64] = {
    0x428a2f98u, 0x71374491u, 0xb5c0fbcfu, 0xe9b5dba5u, 0x3956c25bu, 0x59f111f1u, 0x923f82a4u,
    0xab1c5ed5u, 0xd807aa98u, 0x12835b01u, 0x243185beu, 0x550c7dc3u, 0x72be5d74u, 0x80deb1feu,
    0x9bdc06a7u, 0xc19bf174u, 0xe49b69c1u, 0xefbe4786u, 0x0fc19dc6u, 0x240ca1ccu, 0x2de92c6fu,
    0x4a7484aau, 0x5cb0a9dcu, 0x76f988dau, 0x983e5152u, 0xa831c66du, 0xb00327c8u, 0xbf597fc7u,
    0xc6e00bf3u, 0xd5a79147u, 0x06ca6351u, 0x14292967u, 0x27b70a85u, 0x2e1b2138u, 0x4d2c6dfcu,
    0x53380d13u, 0x650a7354u, 0x766a0abbu, 0x81c2c92eu, 0x92722c85u, 0xa2bfe8a1u, 0xa81a664bu,
    0xc24b8b70u, 0xc76c51a3u, 0xd192e819u, 0xd6990624u, 0xf40e3585u, 0x106aa070u, 0x19a4c116u,
    0x1e376c08u, 0x2748774cu, 0x34b0bcb5u, 0x391c0cb3u, 0x4ed8aa4au, 0x5b9cca4fu, 0x682e6ff3u,
    0x748f82eeu, 0x78a5636fu, 0x84c87814u, 0x8cc70208u, 0x90befffau, 0xa4506cebu, 0xbef9a3f7u,
    0xc67178f2u};

static inline uint32_t bw_ror32(uint32_t x, int n) { return (x >> n) | (x << (32 - n)); }

static void bw_sha256_block(bw_sha256_ctx* c, const uint8_t* p) {
  uint32_t w[64];
  for (int i = 0; i < 16; ++i)
    w[i] = ((uint32_t)p[4 * i] << 24) | ((uint32_t)p[4 * i + 1] << 16) |
           ((uint32_t)p[4 * i + 2] << 8) | (uint32_t)p[4 * i + 3];
  for (int i = 16; i < 64; ++i) {
    const uint32_t a = w[i - 15], b = w[i - 2];
    const uint32_t s0 = bw_ror32(a, 7) ^ bw_ror32(a, 18) ^ (a >> 3);
    const uint32_t s1 = bw_ror32(b, 17) ^ bw_ror32(b, 19) ^ (b >> 10);
    w[i] = w[i - 16] + s0 + w[i - 7] + s1;
  }
  uint32_t v[8];
  for (int i = 0; i < 8; ++i) v[i] = c->h[i];
  for (int i = 0; i < 64; ++i) {
    const uint32_t e = v[4], a = v[0];
    const uint32_t t1 = v[7] + (bw_ror32(e, 6) ^ bw_ror32(e, 11) ^ bw_ror32(e, 25)) +
                        ((e & v[5]) ^ (~e & v[6])) + bw_sha256_k[i] + w[i];
    const uint32_t t2 = (bw_ror32(a, 2) ^ bw_ror32(a, 13) ^ bw_ror32(a, 22)) +
                        ((a & v[1]) ^ (a & v[2]) ^ (v[1] & v[2]));
    v[7] = v[6]; v[6] = v[5]; v[5] = v[4]; v[4] = v[3] + t1;
    v[3] = v[2]; v[2] = v[1]; v[1] = v[0]; v[0] = t1 + t2;
  }
  for (int i = 0; i < 8; ++i) c->h[i] += v[i];
}

static void bw_sha256_init(bw_sha256_ctx* c) {
  static const uint32_t iv[8] = {0x6a09e667u, 0xbb67ae85u, 0x3c6ef372u, 0xa54ff53au,
                                 0x510e527fu, 0x9b05688cu, 0x1f83d9abu, 0x5be0cd19u};
  memcpy(c->h, iv, sizeof(iv));
  c->nbytes = 0;
  c->used = 0;
}

static void bw_sha256_update(bw_sha256_ctx* c, const void* data, size_t n) {
  const uint8_t* p = (const uint8_t*)data;
  c->nbytes += n;
  while (n > 0) {
    if (c->used == 0 && n >= 64) {
      bw_sha256_block(c, p);
      p += 64;
      n -= 64;
      continue;
    }
    size_t take = 64 - c->used;
    if (take > n) take = n;
    memcpy(c->blk + c->used, p, take);
    c->used += take;
    p += take;
    n -= take;
    if (c->used == 64) {
      bw_sha256_block(c, c->blk);
      c->used = 0;
    }
  }
}

static void bw_sha256_final_hex(bw_sha256_ctx* c, char out[65]) {
  const uint64_t bits = c->nbytes * 8u;
  uint8_t pad = 0x80;
  uint64_t keep = c->nbytes;
  bw_sha256_update(c, &pad, 1);
  pad = 0;
  while (c->used != 56) bw_sha256_update(c, &pad, 1);
  uint8_t len[8];
  for (int i = 0; i < 8; ++i) len[i] = (uint8_t)(bits >> (56 - 8 * i));
  bw_sha256_update(c, len, 8);
  c->nbytes = keep;
  static const char hx[] = "0123456789abcdef";
  for (int i = 0; i < 8; ++i)
    for (int j = 0; j < 4; ++j) {
      const uint8_t b = (uint8_t)(c->h[i] >> (24 - 8 * j));
      out[8 * i + 2 * j] = hx[b >> 4];
      out[8 * i + 2 * j + 1] = hx[b & 15];
    }
  out[64] = 0;
}

static void bw_sha256_hex(const void* data, size_t n, char out[65]) {
  bw_sha256_ctx c;
  bw_sha256_init(&c);
  bw_sha256_update(&c, data, n);
  bw_sha256_final_hex(&c, out);
}

#endif
