/* bwgpu_extract.h -- C-ABI of the overlapped extract: the answer vector leaves the device in
 * chunks while host threads write it to the answers file and hash it, so the copy, the write and
 * the two checksums run at the same time instead of one after another.
 *
 * Counterpart in the reference: the GPU workers' download and extract phases -- one blocking
 * cudaMemcpy of the whole answer vector, then fwrite, FNV-1a 64 and SHA-256 over it in turn
 * (src/panel_batch_gpu.cpp:223-235, hash src/sha_fast.hpp:27-43, FNV-1a src/common.hpp:396-400).
 * The outputs are the same: the file holds the little-endian int64 answers, `sha256_hex` is the
 * lowercase hex SHA-256 of those bytes and `fnv1a` their FNV-1a 64 as the reference computes it:
 * prime 0x100000001b3 from the start value 1469598103934665603 (0x14650fb0739d0383; the published
 * FNV offset basis has one more digit, 14695981039346656037 -- the reference's value is kept so
 * the checksums match its).  Both checksums are sequential over the bytes, so each runs on one
 * thread; the extract's time becomes that of its slowest part instead of the sum of all parts.
 * Host code only; the copies run on `stream` (a hipStream_t, NULL = default stream), which is
 * synchronised first, so the answers must be complete when the work enqueued on it is. */
#ifndef BWGPU_EXTRACT_H
#define BWGPU_EXTRACT_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct BwExtractResult {
  char sha256_hex[65];
  uint64_t fnv1a;
  double copy_ms;   /* until the last chunk was on the host */
  double write_ms;  /* until the file was written and closed (0 without a file) */
  double sha_ms;    /* until the SHA-256 was final */
  double fnv_ms;    /* until the FNV-1a was final */
  double total_ms;
} BwExtractResult;

/* Flags of bw_extract_answers. */
#define BW_EXTRACT_FNV_ON_DEVICE 1 /* FNV-1a by bw_fnv1a_device (overlapping the copy) */

/* Copies n answers from d_answers (device) to host_out (n int64, or NULL: an internal buffer) in
 * chunks, writing them to `path` (NULL: no file) and hashing them as the chunks arrive.
 * Returns 0, a hipError_t of a failed copy, or -1 when the file cannot be opened or fully
 * written (the reference workers' exit code 4). */
int bw_extract_answers(const int64_t* d_answers, size_t n, void* stream, const char* path, int64_t* host_out,
                       int flags, BwExtractResult* result);

/* FNV-1a 64 of n bytes of device memory, from `start` (the reference's start value is
 * 1469598103934665603), computed on the device: XOR with a byte changes only the low 8 bits of the
 * state, and the low byte of the next state depends only on low bytes, so the low-byte sequence is
 * an 8-bit chain whose map over a 4 KiB block can be tabulated for all 256 entry values at once
 * and composed across blocks; with every block's entry low byte known, the hash is a polynomial
 * in the prime, h = start * P^n + sum_j S_j * P^(n - end_j), whose block terms S_j are
 * independent.  Writes the value to d_result (device, 8 bytes); asynchronous on `stream`.
 * d_bytes: 4-byte aligned.  work: bw_fnv1a_workspace(n) bytes of device memory, 256-byte aligned. */
int bw_fnv1a_workspace(size_t n, size_t* bytes);
int bw_fnv1a_device(const void* d_bytes, size_t n, uint64_t start, void* work, size_t work_bytes, uint64_t* d_result,
                    void* stream);

#ifdef __cplusplus
}
#endif
#endif
