// extract.cpp -- the overlapped extract (include/bwgpu_extract.h).  The calling thread copies the
// answers off the device chunk by chunk; threads follow it chunk by chunk: the file writer,
// SHA-256 and (unless the device computes it, bw_fnv1a_device, beside the copies) FNV-1a.  Each
// checksum is one sequential chain over the bytes on the host, so the extract takes as long as
// its slowest chain instead of the sum of copy, write and both hashes.
#include "bwgpu_extract.h"

#include <hip/hip_runtime_api.h>
#include <openssl/evp.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstdio>
#include <memory>
#include <mutex>
#include <thread>

namespace {

using clk = std::chrono::steady_clock;

constexpr size_t kChunkBytes = size_t(4) << 20;

double ms_since(clk::time_point t0) { return std::chrono::duration<double, std::milli>(clk::now() - t0).count(); }

// Chunks published by the copying thread; followers wait for chunk k, or for a failed copy.
class Progress {
 public:
  void publish(size_t k) {
    {
      std::lock_guard<std::mutex> g(m_);
      ready_ = k;
    }
    cv_.notify_all();
  }
  void fail() {
    {
      std::lock_guard<std::mutex> g(m_);
      failed_ = true;
    }
    cv_.notify_all();
  }
  // false when the copy failed before chunk k (exclusive count) arrived
  bool wait_for(size_t k) {
    std::unique_lock<std::mutex> l(m_);
    cv_.wait(l, [&] { return ready_ >= k || failed_; });
    return ready_ >= k;
  }

 private:
  std::mutex m_;
  std::condition_variable cv_;
  size_t ready_ = 0;
  bool failed_ = false;
};

}  // namespace

extern "C" int bw_extract_answers(const int64_t* d_answers, size_t n, void* stream, const char* path,
                                  int64_t* host_out, int flags, BwExtractResult* result) {
  if (!result || (n && !d_answers)) return int(hipErrorInvalidValue);
  const auto t0 = clk::now();
  *result = BwExtractResult{};
  hipStream_t st = hipStream_t(stream);
  hipError_t e = hipStreamSynchronize(st);
  if (e != hipSuccess) return int(e);
  FILE* f = nullptr;
  if (path) {
    f = std::fopen(path, "wb");
    if (!f) return -1;
  }
  std::unique_ptr<int64_t[]> own;
  int64_t* host = host_out;
  if (!host) {
    own.reset(new int64_t[std::max<size_t>(n, 1)]);
    host = own.get();
  }
  const unsigned char* bytes = reinterpret_cast<const unsigned char*>(host);
  const size_t total = n * sizeof(int64_t);
  const size_t nchunk = (total + kChunkBytes - 1) / kChunkBytes;
  auto chunk_len = [&](size_t k) { return std::min(kChunkBytes, total - k * kChunkBytes); };

  uint64_t fnv = 1469598103934665603ULL;  // the reference's start value (src/common.hpp:396)
  // FNV-1a on the device: its kernels run on a stream of their own, beside the copies
  const bool dev_fnv = (flags & BW_EXTRACT_FNV_ON_DEVICE) != 0;
  hipStream_t fs = nullptr;
  void* fwork = nullptr;
  uint64_t* d_fnv = nullptr;
  if (dev_fnv) {
    size_t wb = 0;
    e = hipError_t(bw_fnv1a_workspace(total, &wb));
    if (e == hipSuccess) e = hipMalloc(&fwork, wb + 256);
    if (e == hipSuccess) e = hipStreamCreate(&fs);
    if (e == hipSuccess) {
      d_fnv = reinterpret_cast<uint64_t*>(static_cast<unsigned char*>(fwork) + wb);
      e = hipError_t(bw_fnv1a_device(d_answers, total, fnv, fwork, wb, d_fnv, fs));
    }
    if (e != hipSuccess) {
      if (fs) (void)hipStreamDestroy(fs);
      if (fwork) (void)hipFree(fwork);
      if (f) std::fclose(f);
      return int(e);
    }
  }

  Progress prog;
  bool write_ok = true, sha_ok = true;
  unsigned char digest[32] = {};
  std::thread writer, sha, fnv_thread;
  if (f)
    writer = std::thread([&] {
      for (size_t k = 0; k < nchunk && write_ok; ++k) {
        if (!prog.wait_for(k + 1)) return;
        write_ok = std::fwrite(bytes + k * kChunkBytes, 1, chunk_len(k), f) == chunk_len(k);
      }
      if (std::fclose(f) != 0) write_ok = false;
      f = nullptr;
      result->write_ms = ms_since(t0);
    });
  sha = std::thread([&] {
    EVP_MD_CTX* c = EVP_MD_CTX_new();
    sha_ok = c && EVP_DigestInit_ex(c, EVP_sha256(), nullptr) == 1;
    for (size_t k = 0; k < nchunk && sha_ok; ++k) {
      if (!prog.wait_for(k + 1)) break;
      sha_ok = EVP_DigestUpdate(c, bytes + k * kChunkBytes, chunk_len(k)) == 1;
    }
    unsigned int len = 0;
    if (sha_ok) sha_ok = EVP_DigestFinal_ex(c, digest, &len) == 1 && len == 32;
    if (c) EVP_MD_CTX_free(c);
    result->sha_ms = ms_since(t0);
  });
  if (!dev_fnv)
    fnv_thread = std::thread([&] {
      uint64_t h = fnv;
      for (size_t k = 0; k < nchunk; ++k) {
        if (!prog.wait_for(k + 1)) break;
        const unsigned char* p = bytes + k * kChunkBytes;
        for (size_t i = 0, m = chunk_len(k); i < m; ++i) h = (h ^ p[i]) * 0x100000001b3ULL;
      }
      fnv = h;
      result->fnv_ms = ms_since(t0);
    });

  for (size_t k = 0; k < nchunk; ++k) {
    e = hipMemcpyAsync(reinterpret_cast<unsigned char*>(host) + k * kChunkBytes,
                       reinterpret_cast<const unsigned char*>(d_answers) + k * kChunkBytes, chunk_len(k),
                       hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) {
      prog.fail();
      break;
    }
    prog.publish(k + 1);
  }
  result->copy_ms = ms_since(t0);
  if (dev_fnv) {
    hipError_t fe = hipMemcpyAsync(&fnv, d_fnv, sizeof(fnv), hipMemcpyDeviceToHost, fs);
    if (fe == hipSuccess) fe = hipStreamSynchronize(fs);
    result->fnv_ms = ms_since(t0);
    (void)hipStreamDestroy(fs);
    (void)hipFree(fwork);
    if (e == hipSuccess) e = fe;
  }
  if (writer.joinable()) writer.join();
  sha.join();
  if (fnv_thread.joinable()) fnv_thread.join();
  if (f) std::fclose(f);  // only when the writer never ran to its end
  result->total_ms = ms_since(t0);
  if (e != hipSuccess) return int(e);
  if (!write_ok) return -1;
  if (!sha_ok) return int(hipErrorUnknown);
  static const char* hx = "0123456789abcdef";
  for (int i = 0; i < 32; ++i) {
    result->sha256_hex[2 * i] = hx[digest[i] >> 4];
    result->sha256_hex[2 * i + 1] = hx[digest[i] & 15];
  }
  result->sha256_hex[64] = 0;
  result->fnv1a = fnv;
  return 0;
}
