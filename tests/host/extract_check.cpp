// extract_check.cpp -- test infrastructure: the overlapped extract (csrc/host/extract.cpp) built
// on the CPU against a stand-in for the two HIP runtime calls it makes (a device-to-host copy is
// a host memcpy here, a stream synchronisation is a no-op), so its chunking, threads, file output
// and both checksums can be compared with one-shot SHA-256 / FNV-1a over the same bytes.
//
//   extract_check N FILE   (N answers, a fixed pattern; prints one JSON line)
#include <hip/hip_runtime_api.h>
#include <openssl/sha.h>

#include <cinttypes>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "bwgpu_extract.h"

extern "C" hipError_t hipStreamSynchronize(hipStream_t) { return hipSuccess; }
// the device FNV path is not taken here (flags 0); its calls only need to link
extern "C" hipError_t hipMalloc(void**, size_t) { return hipErrorOutOfMemory; }
extern "C" hipError_t hipFree(void*) { return hipSuccess; }
extern "C" hipError_t hipStreamCreate(hipStream_t*) { return hipErrorNotSupported; }
extern "C" hipError_t hipStreamDestroy(hipStream_t) { return hipSuccess; }
extern "C" int bw_fnv1a_workspace(size_t, size_t* b) { *b = 256; return 0; }
extern "C" int bw_fnv1a_device(const void*, size_t, uint64_t, void*, size_t, uint64_t*, void*) { return 1; }
extern "C" hipError_t hipMemcpyAsync(void* dst, const void* src, size_t n, hipMemcpyKind, hipStream_t) {
  std::memcpy(dst, src, n);
  return hipSuccess;
}

int main(int argc, char** argv) {
  if (argc < 3) return 2;
  const size_t n = std::strtoull(argv[1], nullptr, 10);
  std::vector<int64_t> src(n), dst(n, 0);
  uint64_t z = 0x9E3779B97F4A7C15ULL;
  for (size_t i = 0; i < n; ++i) {
    z ^= z << 13; z ^= z >> 7; z ^= z << 17;
    src[i] = int64_t(z) >> (i % 40);
  }
  BwExtractResult r;
  const int rc = bw_extract_answers(src.data(), n, nullptr, argv[2], dst.data(), 0, &r);
  unsigned char d[32];
  SHA256(reinterpret_cast<const unsigned char*>(src.data()), n * 8, d);
  char hex[65];
  for (int i = 0; i < 32; ++i) std::snprintf(hex + 2 * i, 3, "%02x", d[i]);
  uint64_t h = 1469598103934665603ULL;
  const unsigned char* p = reinterpret_cast<const unsigned char*>(src.data());
  for (size_t i = 0; i < n * 8; ++i) h = (h ^ p[i]) * 0x100000001b3ULL;
  const bool copied = std::memcmp(src.data(), dst.data(), n * 8) == 0;
  std::printf("{\"rc\": %d, \"sha\": \"%s\", \"sha_one_shot\": \"%s\", \"fnv\": \"%016" PRIx64
              "\", \"fnv_one_shot\": \"%016" PRIx64 "\", \"copied\": %s, \"total_ms\": %.3f}\n",
              rc, rc == 0 ? r.sha256_hex : "", hex, r.fnv1a, h, copied ? "true" : "false", r.total_ms);
  return 0;
}
