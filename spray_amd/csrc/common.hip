// Device-wide exclusive scan of uint32 counts: per-block scans of 2048
// values, a single-block scan of the block sums (chunked, any length), and
// a final add (up to 1024 tiles, each block of the add sums the tile sums
// before it instead).  Used for vertex adjacency offsets, radix-sort digit
// offsets, PLOC merge positions and the marching-cubes edge and cell block
// offsets.
#include <algorithm>

#include "ctx.h"
#include "kernels_common.h"

namespace spray {

constexpr int kScanBlock = 256;
constexpr int kScanItems = 8;
constexpr int kScanTile = kScanBlock * kScanItems;

// block-wide exclusive scan of one value per thread; returns the total
__device__ __forceinline__ uint32_t blockExclusive(uint32_t v, uint32_t* lds, uint32_t* total) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = blockDim.x >> 6;
  uint32_t x = v;
  for (int o = 1; o < 64; o <<= 1) {
    uint32_t y = __shfl_up(x, o, 64);
    if (lane >= o) x += y;
  }
  if (lane == 63) lds[w] = x;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t acc = 0;
    for (int i = 0; i < nw; ++i) {
      uint32_t t = lds[i];
      lds[i] = acc;
      acc += t;
    }
    lds[nw] = acc;
  }
  __syncthreads();
  uint32_t r = x - v + lds[w];
  *total = lds[nw];
  __syncthreads();
  return r;
}

__global__ void kScanTiles(const uint32_t* in, uint32_t* out, long long n, uint32_t* sums) {
  __shared__ uint32_t lds[kScanBlock / 64 + 1];
  long long base = (long long)blockIdx.x * kScanTile + (long long)threadIdx.x * kScanItems;
  uint32_t vals[kScanItems];
  uint32_t local = 0;
  for (int j = 0; j < kScanItems; ++j) {
    long long i = base + j;
    vals[j] = i < n ? in[i] : 0u;
    local += vals[j];
  }
  uint32_t total;
  uint32_t pre = blockExclusive(local, lds, &total);
  for (int j = 0; j < kScanItems; ++j) {
    long long i = base + j;
    if (i < n) out[i] = pre;
    pre += vals[j];
  }
  if (threadIdx.x == 0) sums[blockIdx.x] = total;
}

__global__ void kScanSums(uint32_t* sums, long long m) {
  __shared__ uint32_t lds[1024 / 64 + 1];
  uint32_t carry = 0;
  for (long long base = 0; base < m; base += blockDim.x) {
    long long i = base + threadIdx.x;
    uint32_t v = i < m ? sums[i] : 0u;
    uint32_t total;
    uint32_t pre = blockExclusive(v, lds, &total);
    if (i < m) sums[i] = pre + carry;
    carry += total;
  }
}

// Up to this many tiles, kScanAddSum adds the prefix of the tile sums that
// each block forms itself, and the separate tile-sum scan is skipped.
constexpr long long kScanFuseTiles = 1024;

__global__ void kScanAddSum(uint32_t* out, long long n, const uint32_t* sums) {
  __shared__ uint32_t pre;
  if (threadIdx.x < 64) {
    uint32_t x = 0;
    for (long long q = threadIdx.x; q < blockIdx.x; q += 64) x += sums[q];
    for (int o = 32; o; o >>= 1) x += __shfl_xor(x, o, 64);
    if (threadIdx.x == 0) pre = x;
  }
  __syncthreads();
  const long long base = (long long)blockIdx.x * kScanTile;
  for (int j = threadIdx.x; j < kScanTile; j += blockDim.x) {
    const long long i = base + j;
    if (i < n) out[i] += pre;
  }
}

__global__ void kScanAdd(uint32_t* out, long long n, const uint32_t* sums) {
  long long base = (long long)blockIdx.x * kScanTile;
  uint32_t add = sums[blockIdx.x];
  for (int j = threadIdx.x; j < kScanTile; j += blockDim.x) {
    long long i = base + j;
    if (i < n) out[i] += add;
  }
}

long long scanSumsWords(long long n) { return (n + kScanTile - 1) / kScanTile + 16; }

void exclusiveScan(const uint32_t* in, uint32_t* out, long long n, hipStream_t s, uint32_t* tile_sums) {
  if (n <= 0) return;
  long long nb = (n + kScanTile - 1) / kScanTile;
  kScanTiles<<<(unsigned)nb, kScanBlock, 0, s>>>(in, out, n, tile_sums);
  if (nb > 1 && nb <= kScanFuseTiles) {
    kScanAddSum<<<(unsigned)nb, kScanBlock, 0, s>>>(out, n, tile_sums);
  } else if (nb > 1) {
    kScanSums<<<1, 1024, 0, s>>>(tile_sums, nb);
    kScanAdd<<<(unsigned)nb, kScanBlock, 0, s>>>(out, n, tile_sums);
  }
  SPRAY_HIP(hipGetLastError());
}

void exclusiveScan(const uint32_t* in, uint32_t* out, long long n, hipStream_t s) {
  if (n <= 0) return;
  long long nb = (n + kScanTile - 1) / kScanTile;
  // per host thread; freed when the thread ends
  struct Sums {
    DevBuf b;
    ~Sums() { b.release(); }
  };
  static thread_local Sums sums;
  sums.b.reserve(std::max<size_t>((size_t)nb * 4 + 64, (size_t)4 << 20));
  exclusiveScan(in, out, n, s, sums.b.as<uint32_t>());
}

}  // namespace spray
