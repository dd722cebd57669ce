// Wave-level reductions (wave64) and the device-wide exclusive scan.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace spray {

__device__ __forceinline__ float waveMin(float v) {
  for (int o = 32; o > 0; o >>= 1) v = fminf(v, __shfl_xor(v, o, 64));
  return v;
}
__device__ __forceinline__ float waveMax(float v) {
  for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o, 64));
  return v;
}
__device__ __forceinline__ float waveSumF(float v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}
__device__ __forceinline__ uint32_t waveSumU(uint32_t v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}
__device__ __forceinline__ unsigned long long waveSumULL(unsigned long long v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// out[i] = sum(in[0..i-1]) for n uint32 values; in == out is allowed.  The
// first form keeps its tile sums in a buffer of the calling host thread; the
// second takes them from the caller (scanSumsWords(n) words), so scans on
// several streams of one thread do not share it.
void exclusiveScan(const uint32_t* in, uint32_t* out, long long n, hipStream_t s);
void exclusiveScan(const uint32_t* in, uint32_t* out, long long n, hipStream_t s, uint32_t* tile_sums);
long long scanSumsWords(long long n);

}  // namespace spray
