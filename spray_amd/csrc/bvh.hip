// HBM domain pool: vertex normals, BVH build (Morton codes, LDS radix sort,
// PLOC clustering or a Karras hierarchy whose inner boxes are leaf-range
// unions, collapse to a 4-wide BVH)
// and the batch ray queries behind spray_trace_rays / spray_domain_list.
//
// Replaces the reference's per-domain Embree scene (rtcCommit in
// render/trimesh_buffer.cc:213-224) and TriMeshBuffer::computeNormals
// (render/trimesh_buffer.cc:267-326).
#include <hip/hip_runtime.h>

#include <mutex>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>

#include "ctx.h"
#include "kernels_common.h"
#include "traverse.h"

namespace spray {

// ---------------------------------------------------------------------------
// Vertex normals: the reference adds each face's unnormalised normal to its
// three vertices in face order.  A vertex-parallel gather over the faces of
// each vertex, visited in ascending face order, reproduces that float
// summation order exactly.
// Face corners are numbered in 32 bits: a domain has fewer than 2^24 faces.
__global__ void kVertCount(const uint32_t* __restrict__ face, long long nf, uint32_t* cnt) {
  const uint32_t nc = (uint32_t)nf * 3u;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < nc; i += gridDim.x * blockDim.x)
    atomicAdd(&cnt[face[i]], 1u);
}
__global__ void kVertFill(const uint32_t* __restrict__ face, long long nf,
                          const uint32_t* __restrict__ off, uint32_t* cursor, uint32_t* adj) {
  const uint32_t nc = (uint32_t)nf * 3u;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < nc; i += gridDim.x * blockDim.x) {
    const uint32_t v = face[i];
    const uint32_t slot = atomicAdd(&cursor[v], 1u);
    adj[off[v] + slot] = i / 3u;
  }
}
__global__ void kVertNormal(const float* __restrict__ vtx, const uint32_t* __restrict__ face,
                            long long nv, const uint32_t* __restrict__ off,
                            const uint32_t* __restrict__ cnt, uint32_t* adj, float* nrm) {
  for (long long v = blockIdx.x * (long long)blockDim.x + threadIdx.x; v < nv;
       v += (long long)gridDim.x * blockDim.x) {
    uint32_t b = off[v], n = cnt[v];
    uint32_t* L = adj + b;
    for (uint32_t i = 1; i < n; ++i) {  // insertion sort: lists are short
      uint32_t x = L[i];
      int j = (int)i - 1;
      while (j >= 0 && L[j] > x) {
        L[j + 1] = L[j];
        --j;
      }
      L[j + 1] = x;
    }
    float ax = 0.0f, ay = 0.0f, az = 0.0f;
    for (uint32_t i = 0; i < n; ++i) {
      const uint32_t* fc = face + (size_t)L[i] * 3;
      const float* p0 = vtx + (size_t)fc[0] * 3;
      const float* p1 = vtx + (size_t)fc[1] * 3;
      const float* p2 = vtx + (size_t)fc[2] * 3;
      float ux = p1[0] - p0[0], uy = p1[1] - p0[1], uz = p1[2] - p0[2];
      float wx = p2[0] - p0[0], wy = p2[1] - p0[1], wz = p2[2] - p0[2];
      ax += uy * wz - uz * wy;
      ay += uz * wx - ux * wz;
      az += ux * wy - uy * wx;
    }
    nrm[v * 3] = ax;
    nrm[v * 3 + 1] = ay;
    nrm[v * 3 + 2] = az;
  }
}

// Shading record per face: the three vertex normals and colours of the
// face in one 48-byte record, so shading reads one record per hit instead
// of the face and then six vertex attributes.
// It also sets the centroid bounds (ordered bits, see kCentroidBounds) to
// their starting values for the Morton pass that follows it on the stream.
__global__ void kFaceShade(const uint32_t* __restrict__ face, long long nf,
                           const float* __restrict__ nrm, const uint32_t* __restrict__ rgb,
                           float4* out, uint32_t* cbounds) {
  if (blockIdx.x == 0 && threadIdx.x < 6) cbounds[threadIdx.x] = threadIdx.x < 3 ? 0xffffffffu : 0u;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < nf;
       i += (long long)gridDim.x * blockDim.x) {
    for (int k = 0; k < 3; ++k) {
      uint32_t v = face[i * 3 + k];
      out[i * 3 + k] = make_float4(nrm[(size_t)v * 3], nrm[(size_t)v * 3 + 1],
                                   nrm[(size_t)v * 3 + 2], __uint_as_float(rgb[v]));
    }
  }
}

// ---------------------------------------------------------------------------
// Centroid bounds and Morton codes
__device__ __forceinline__ uint32_t orderedBits(float f) {
  uint32_t u = __float_as_uint(f);
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float fromOrdered(uint32_t u) {
  return __uint_as_float((u & 0x80000000u) ? (u & 0x7fffffffu) : ~u);
}

// Launched on a small grid (kBoundsGrid blocks): per-block reduction in LDS,
// then one atomic per block and bound, so the six bound words see few
// same-address atomics.
constexpr int kBoundsGrid = 256;
__global__ void __launch_bounds__(256) kCentroidBounds(const float* __restrict__ vtx,
                                                       const uint32_t* __restrict__ face, int nf,
                                                       uint32_t* bmin, uint32_t* bmax) {
  __shared__ float red[6][4];
  float lo[3] = {kInfF, kInfF, kInfF}, hi[3] = {-kInfF, -kInfF, -kInfF};
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < nf; i += gridDim.x * blockDim.x) {
    for (int k = 0; k < 3; ++k) {
      float a = vtx[(size_t)face[i * 3] * 3 + k], b = vtx[(size_t)face[i * 3 + 1] * 3 + k],
            c = vtx[(size_t)face[i * 3 + 2] * 3 + k];
      float cen = 0.5f * (fminf(a, fminf(b, c)) + fmaxf(a, fmaxf(b, c)));
      lo[k] = fminf(lo[k], cen);
      hi[k] = fmaxf(hi[k], cen);
    }
  }
  const int w = threadIdx.x >> 6;
  for (int k = 0; k < 3; ++k) {
    float l = waveMin(lo[k]), h = waveMax(hi[k]);
    if ((threadIdx.x & 63) == 0) {
      red[k][w] = l;
      red[3 + k][w] = h;
    }
  }
  __syncthreads();
  if (threadIdx.x < 6) {
    const int k = threadIdx.x;
    float v = red[k][0];
    for (int i = 1; i < (int)(blockDim.x >> 6); ++i) v = k < 3 ? fminf(v, red[k][i]) : fmaxf(v, red[k][i]);
    if (k < 3)
      atomicMin(&bmin[k], orderedBits(v));
    else
      atomicMax(&bmax[k - 3], orderedBits(v));
  }
}

__device__ __forceinline__ uint32_t expandBits10(uint32_t v) {
  v = (v * 0x00010001u) & 0xFF0000FFu;
  v = (v * 0x00000101u) & 0x0F00F00Fu;
  v = (v * 0x00000011u) & 0xC30C30C3u;
  v = (v * 0x00000005u) & 0x49249249u;
  return v;
}

__global__ void kMorton(const float* __restrict__ vtx, const uint32_t* __restrict__ face, int nf,
                        const uint32_t* bmin, const uint32_t* bmax, uint32_t* keys,
                        uint32_t* vals) {
  float lo[3], ext[3];
  for (int k = 0; k < 3; ++k) {
    lo[k] = fromOrdered(bmin[k]);
    float e = fromOrdered(bmax[k]) - lo[k];
    ext[k] = e > 0.f ? e : 1.f;
  }
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < nf; i += gridDim.x * blockDim.x) {
    uint32_t q[3];
    for (int k = 0; k < 3; ++k) {
      float a = vtx[(size_t)face[i * 3] * 3 + k], b = vtx[(size_t)face[i * 3 + 1] * 3 + k],
            c = vtx[(size_t)face[i * 3 + 2] * 3 + k];
      float cen = 0.5f * (fminf(a, fminf(b, c)) + fmaxf(a, fmaxf(b, c)));
      float x = (cen - lo[k]) / ext[k] * 1024.0f;
      x = fminf(fmaxf(x, 0.0f), 1023.0f);
      q[k] = (uint32_t)x;
    }
    keys[i] = (expandBits10(q[0]) << 2) | (expandBits10(q[1]) << 1) | expandBits10(q[2]);
    vals[i] = (uint32_t)i;
  }
}

// ---------------------------------------------------------------------------
// LSD radix sort of (key, value) pairs, 8-bit digits, 2048-key tiles.
// Ranking inside a tile is stable: each 256-key round is split per wave with
// a ballot-based digit match, waves are ordered through LDS counters and a
// running per-digit count carries across rounds.
constexpr int kRsBlock = 256;
constexpr int kRsItems = 8;
constexpr int kRsTile = kRsBlock * kRsItems;

__global__ void kRsHist(const uint32_t* __restrict__ keys, int n, int shift, uint32_t* hist) {
  __shared__ uint32_t h[256];
  h[threadIdx.x] = 0;
  __syncthreads();
  int base = blockIdx.x * kRsTile;
  for (int j = 0; j < kRsItems; ++j) {
    int idx = base + j * kRsBlock + threadIdx.x;
    if (idx < n) atomicAdd(&h[(keys[idx] >> shift) & 255u], 1u);
  }
  __syncthreads();
  hist[threadIdx.x * gridDim.x + blockIdx.x] = h[threadIdx.x];
}

__global__ void kRsScatter(const uint32_t* __restrict__ kin, const uint32_t* __restrict__ vin,
                           uint32_t* __restrict__ kout, uint32_t* __restrict__ vout, int n,
                           int shift, const uint32_t* __restrict__ hist_scanned) {
  __shared__ uint32_t base[256];
  __shared__ uint32_t run[256];
  __shared__ uint32_t wc[4][256];
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  base[tid] = hist_scanned[tid * gridDim.x + blockIdx.x];
  run[tid] = 0;
  const int tile = blockIdx.x * kRsTile;
  for (int j = 0; j < kRsItems; ++j) {
    wc[0][tid] = 0;
    wc[1][tid] = 0;
    wc[2][tid] = 0;
    wc[3][tid] = 0;
    __syncthreads();
    int idx = tile + j * kRsBlock + tid;
    bool valid = idx < n;
    uint32_t key = valid ? kin[idx] : 0u;
    uint32_t d = (key >> shift) & 255u;
    uint64_t peers = __ballot(valid);
#pragma unroll
    for (int b = 0; b < 8; ++b) {
      bool bit = (d >> b) & 1u;
      uint64_t m = __ballot(bit);
      peers &= bit ? m : ~m;
    }
    uint32_t rank = __popcll(peers & ((1ull << lane) - 1ull));
    if (valid && rank == 0) wc[w][d] = (uint32_t)__popcll(peers);
    __syncthreads();
    {
      uint32_t c0 = wc[0][tid], c1 = wc[1][tid], c2 = wc[2][tid], c3 = wc[3][tid];
      uint32_t r = run[tid];
      wc[0][tid] = r;
      wc[1][tid] = r + c0;
      wc[2][tid] = r + c0 + c1;
      wc[3][tid] = r + c0 + c1 + c2;
      run[tid] = r + c0 + c1 + c2 + c3;
    }
    __syncthreads();
    if (valid) {
      uint32_t dst = base[d] + wc[w][d] + rank;
      kout[dst] = key;
      vout[dst] = vin[idx];
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// Karras 2012 hierarchy over sorted keys; duplicate keys are ordered by index.
__device__ __forceinline__ int deltaK(const uint32_t* keys, int n, int i, int j) {
  if (j < 0 || j >= n) return -1;
  uint32_t a = keys[i], b = keys[j];
  if (a == b) return 32 + __clz((uint32_t)(i ^ j));
  return __clz(a ^ b);
}

// inner node i: children encoded c >= 0 inner, c < 0 leaf ~c (sorted position)
__global__ void kKarras(const uint32_t* __restrict__ keys, int n, int* left, int* right,
                        int* first, int* last, int* parent_inner, int* parent_leaf) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n - 1; i += gridDim.x * blockDim.x) {
    int d = (deltaK(keys, n, i, i + 1) - deltaK(keys, n, i, i - 1)) >= 0 ? 1 : -1;
    int dmin = deltaK(keys, n, i, i - d);
    int lmax = 2;
    while (deltaK(keys, n, i, i + lmax * d) > dmin) lmax *= 2;
    int l = 0;
    for (int t = lmax / 2; t >= 1; t /= 2)
      if (deltaK(keys, n, i, i + (l + t) * d) > dmin) l += t;
    int j = i + l * d;
    int dnode = deltaK(keys, n, i, j);
    int s = 0;
    int t = l;
    do {
      t = (t + 1) >> 1;
      if (deltaK(keys, n, i, i + (s + t) * d) > dnode) s += t;
    } while (t > 1);
    int gamma = i + s * d + (d < 0 ? d : 0);
    int lo = i < j ? i : j, hi = i < j ? j : i;
    int lc = (lo == gamma) ? ~gamma : gamma;
    int rc = (hi == gamma + 1) ? ~(gamma + 1) : gamma + 1;
    left[i] = lc;
    right[i] = rc;
    first[i] = lo;
    last[i] = hi;
    if (lc < 0)
      parent_leaf[~lc] = i;
    else
      parent_inner[lc] = i;
    if (rc < 0)
      parent_leaf[~rc] = i;
    else
      parent_inner[rc] = i;
  }
}

// Leaf boxes in sorted order + the leaf-ordered triangle array of the pool.
// parent_init (may be null): set to -1 for every leaf, the Karras
// hierarchy's parent array before kKarras fills it (the root keeps -1).
__global__ void kLeaves(const float* __restrict__ vtx, const uint32_t* __restrict__ face, int nf,
                        const uint32_t* __restrict__ order, float* leafbox, LTri* out,
                        int* parent_init = nullptr) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < nf; i += gridDim.x * blockDim.x) {
    if (parent_init) parent_init[i] = -1;
    uint32_t p = order[i];
    const float* a = vtx + (size_t)face[p * 3] * 3;
    const float* b = vtx + (size_t)face[p * 3 + 1] * 3;
    const float* c = vtx + (size_t)face[p * 3 + 2] * 3;
    LTri t;
    for (int k = 0; k < 3; ++k) {
      t.v0[k] = a[k];
      t.v1[k] = b[k];
      t.v2[k] = c[k];
      leafbox[(size_t)i * 6 + k] = fminf(a[k], fminf(b[k], c[k]));
      leafbox[(size_t)i * 6 + 3 + k] = fmaxf(a[k], fmaxf(b[k], c[k]));
    }
    t.v0[3] = __int_as_float((int)p);
    t.v1[3] = 0.f;
    t.v2[3] = 0.f;
    out[i] = t;
  }
}

// Inner-node boxes without a bottom-up climb: in a Karras hierarchy inner
// node p covers the sorted leaves [first[p], last[p]], so its box is the
// union of those leaf boxes, a range min/max.  Leaves are grouped in blocks
// of 64; kLeafScan writes, per leaf, the union from its block's start
// (prefix) and to its block's end (suffix), and per block the block's box;
// a sparse table over block boxes (level j: unions of 2^j consecutive
// blocks) answers the whole blocks of a range with two lookups.  min/max
// are exact and order-free, so the boxes equal a bottom-up refit's, with
// no device-scope fences or atomics.
constexpr int kRangeBlock = 64;

struct RangeLevels {
  long long off[32];  // float offset of table level j (level 0: block boxes)
};

__device__ __forceinline__ void boxMerge(float lo[3], float hi[3], const float* b) {
  for (int k = 0; k < 3; ++k) {
    lo[k] = fminf(lo[k], b[k]);
    hi[k] = fmaxf(hi[k], b[3 + k]);
  }
}

__global__ void __launch_bounds__(256) kLeafScan(const float* __restrict__ leafbox, int n,
                                                 float* __restrict__ pre, float* __restrict__ suf,
                                                 float* __restrict__ blockbox) {
  const int lane = threadIdx.x & 63;
  const long long nblk = (n + kRangeBlock - 1) / kRangeBlock;
  const long long nw = ((long long)gridDim.x * blockDim.x) >> 6;
  for (long long b = (blockIdx.x * (long long)blockDim.x + threadIdx.x) >> 6; b < nblk; b += nw) {
    const long long i = b * kRangeBlock + lane;
    const bool in = i < n;
    float plo[3], phi[3], slo[3], shi[3];
    for (int k = 0; k < 3; ++k) {
      plo[k] = slo[k] = in ? leafbox[i * 6 + k] : kInfF;
      phi[k] = shi[k] = in ? leafbox[i * 6 + 3 + k] : -kInfF;
    }
    for (int o = 1; o < 64; o <<= 1)
      for (int k = 0; k < 3; ++k) {
        const float a = __shfl_up(plo[k], o, 64), b2 = __shfl_up(phi[k], o, 64);
        const float c = __shfl_down(slo[k], o, 64), d = __shfl_down(shi[k], o, 64);
        if (lane >= o) {
          plo[k] = fminf(plo[k], a);
          phi[k] = fmaxf(phi[k], b2);
        }
        if (lane + o < 64) {
          slo[k] = fminf(slo[k], c);
          shi[k] = fmaxf(shi[k], d);
        }
      }
    if (in)
      for (int k = 0; k < 3; ++k) {
        pre[i * 6 + k] = plo[k];
        pre[i * 6 + 3 + k] = phi[k];
        suf[i * 6 + k] = slo[k];
        suf[i * 6 + 3 + k] = shi[k];
      }
    if (lane == 0)
      for (int k = 0; k < 3; ++k) {
        blockbox[b * 6 + k] = slo[k];
        blockbox[b * 6 + 3 + k] = shi[k];
      }
  }
}

// table level j from level j - 1: m entries, each the union of entries i
// and i + half of the level below
__global__ void kRangeLevel(const float* __restrict__ src, float* __restrict__ dst, long long m,
                            long long half) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < m;
       i += (long long)gridDim.x * blockDim.x)
    for (int k = 0; k < 3; ++k) {
      dst[i * 6 + k] = fminf(src[i * 6 + k], src[(i + half) * 6 + k]);
      dst[i * 6 + 3 + k] = fmaxf(src[i * 6 + 3 + k], src[(i + half) * 6 + 3 + k]);
    }
}

// Levels j and j + 1 of the sparse table in one launch, from level j - 1
// (h = 2^(j-1) blocks per entry): level j entry i = union(src[i], src[i + h])
// for i < mj, and level j + 1 entry i = union(level j [i], level j [i + 2h])
// for i < mj2, each level-j box formed as the one-level kernel forms it.
__global__ void kRangeLevel2(const float* __restrict__ src, float* __restrict__ dst,
                             float* __restrict__ dst2, long long mj, long long mj2, long long h) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < mj;
       i += (long long)gridDim.x * blockDim.x) {
    float lo[3], hi[3];
    for (int k = 0; k < 3; ++k) {
      lo[k] = fminf(src[i * 6 + k], src[(i + h) * 6 + k]);
      hi[k] = fmaxf(src[i * 6 + 3 + k], src[(i + h) * 6 + 3 + k]);
      dst[i * 6 + k] = lo[k];
      dst[i * 6 + 3 + k] = hi[k];
    }
    if (i < mj2) {
      const long long a = i + 2 * h, b = i + 3 * h;
      for (int k = 0; k < 3; ++k) {
        dst2[i * 6 + k] = fminf(lo[k], fminf(src[a * 6 + k], src[b * 6 + k]));
        dst2[i * 6 + 3 + k] = fmaxf(hi[k], fmaxf(src[a * 6 + 3 + k], src[b * 6 + 3 + k]));
      }
    }
  }
}

__global__ void kRangeBoxes(int ninner, const int* __restrict__ first, const int* __restrict__ last,
                            const float* __restrict__ leafbox, const float* __restrict__ pre,
                            const float* __restrict__ suf, const float* __restrict__ table,
                            RangeLevels L, float* __restrict__ innerbox) {
  for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < ninner; p += gridDim.x * blockDim.x) {
    const int a = first[p], b = last[p];
    const int ba = a / kRangeBlock, bb = b / kRangeBlock;
    float lo[3] = {kInfF, kInfF, kInfF}, hi[3] = {-kInfF, -kInfF, -kInfF};
    if (ba == bb) {
      for (int i = a; i <= b; ++i) boxMerge(lo, hi, leafbox + (size_t)i * 6);
    } else {
      boxMerge(lo, hi, suf + (size_t)a * 6);
      boxMerge(lo, hi, pre + (size_t)b * 6);
      if (bb - ba >= 2) {
        const int x = ba + 1, y = bb - 1;
        const int j = 31 - __clz(y - x + 1);
        const float* lv = table + L.off[j];
        boxMerge(lo, hi, lv + (size_t)x * 6);
        boxMerge(lo, hi, lv + (size_t)(y - (1 << j) + 1) * 6);
      }
    }
    for (int k = 0; k < 3; ++k) {
      innerbox[(size_t)p * 6 + k] = lo[k];
      innerbox[(size_t)p * 6 + 3 + k] = hi[k];
    }
  }
}

// Table levels for n leaves: offsets and total floats.
static long long rangeLevels(int n, RangeLevels* L, int* nlev) {
  const long long nblk = (n + kRangeBlock - 1) / kRangeBlock;
  long long off = 0;
  int j = 0;
  for (long long w = 1; w <= nblk && j < 32; w <<= 1, ++j) {
    L->off[j] = off;
    off += (nblk - w + 1) * 6;
  }
  *nlev = j;
  return off;
}

// 4-wide nodes of a domain of nf triangles, at most.  In the collapsed tree
// every inner node has two or more children and at most nf leaves; a node
// with fewer than four children has only leaf children, so it covers more
// than kLeafMax triangles, and those nodes are disjoint.  With N4 nodes of
// four children and Ns of fewer: 3 N4 + Ns <= nf - 1 and Ns <= nf / (kLeafMax
// + 1), so N4 + Ns <= (nf - 1) / 3 + 2 nf / (3 (kLeafMax + 1)).
__host__ __device__ inline long long nodeCap(long long nf) {
  return nf / 3 + (2 * nf) / (3 * (kLeafMax + 1)) + 16;
}

// 4-wide collapse levels launched per host wait, and at most (deeper trees
// do not fit the traversal stack anyway); a level counter block is the node
// counter, the overflow flag and the item count of every level
constexpr int kLevelBatch = 16;
constexpr int kMaxLevels = 48;
constexpr int kLevelWords = 2 + kMaxLevels + 1;

// Collapse the binary tree into 4-wide nodes, one frontier level per launch.
struct FrontierItem {
  int bnode;  // binary inner node
  int wnode;  // 4-wide node slot
};

__device__ __forceinline__ float boxArea(const float* b) {
  float dx = fmaxf(b[3] - b[0], 0.f), dy = fmaxf(b[4] - b[1], 0.f), dz = fmaxf(b[5] - b[2], 0.f);
  return dx * dy + dy * dz + dz * dx;
}

// Takes cnt consecutive entries of *ctr per lane, in lane order, with one
// atomic per wave; every lane of the wave calls it (cnt may be 0).
__device__ __forceinline__ int waveAlloc(int* ctr, int cnt) {
  const int lane = threadIdx.x & 63;
  int incl = cnt;
  for (int o = 1; o < 64; o <<= 1) {
    const int v = __shfl_up(incl, o, 64);
    if (lane >= o) incl += v;
  }
  const int total = __shfl(incl, 63, 64);
  int base = 0;
  if (lane == 63 && total > 0) base = atomicAdd(ctr, total);
  base = __shfl(base, 63, 64);
  return base + incl - cnt;
}

// One frontier level: *nin items in fin (0: nothing to do), next level's
// items counted in *nout, on blocks blk of nblk.  Launched level after level
// without host round trips; the loop bound is uniform per block so whole
// waves allocate.
__device__ __forceinline__ void collapseLevel(
    int blk, int nblk, const FrontierItem* __restrict__ fin, const int* __restrict__ nin_ptr,
    FrontierItem* fout, int* nout, int* node_counter, int node_cap, const int* __restrict__ left,
    const int* __restrict__ right, const int* __restrict__ first, const int* __restrict__ last,
    const float* __restrict__ leafbox, const float* __restrict__ innerbox, BNode4* nodes,
    int* overflow) {
  const int nin = *nin_ptr;
  for (int base = blk * blockDim.x; base < nin; base += nblk * blockDim.x) {
    const int it = base + threadIdx.x;
    const bool act = it < nin;
    FrontierItem f{0, 0};
    int ch[4] = {0, 0, 0, 0};
    int nc = 0;
    auto count = [&](int c) { return c < 0 ? 1 : last[c] - first[c] + 1; };
    int ninner = 0;
    if (act) {
      f = fin[it];
      nc = 2;
      ch[0] = left[f.bnode];
      ch[1] = right[f.bnode];
      while (nc < 4) {
        int best = -1;
        float barea = -1.f;
        for (int i = 0; i < nc; ++i) {
          if (ch[i] >= 0 && count(ch[i]) > kLeafMax) {
            float a = boxArea(innerbox + (size_t)ch[i] * 6);
            if (a > barea) {
              barea = a;
              best = i;
            }
          }
        }
        if (best < 0) break;
        int c = ch[best];
        ch[best] = left[c];
        ch[nc++] = right[c];
      }
      // the inner children of a node take consecutive slots (siblings share
      // cache lines when a ray enters several of them)
      for (int i = 0; i < nc; ++i) ninner += (ch[i] >= 0 && count(ch[i]) > kLeafMax) ? 1 : 0;
    }
    int slot = waveAlloc(node_counter, ninner);
    if (ninner > 0 && slot + ninner > node_cap) {
      atomicExch(overflow, 1);
      ninner = 0;
    }
    int q = waveAlloc(nout, ninner);
    if (!act) continue;
    BNode4 nd;
    for (int i = 0; i < 4; ++i) {
      float b[6];
      int word;
      if (i >= nc) {
        for (int k = 0; k < 3; ++k) {
          b[k] = kInfF;
          b[3 + k] = -kInfF;
        }
        word = (int)0x80000000u;
      } else {
        int c = ch[i];
        const float* src = c < 0 ? leafbox + (size_t)(~c) * 6 : innerbox + (size_t)c * 6;
        for (int k = 0; k < 6; ++k) b[k] = src[k];
        if (c < 0) {
          word = (int)(0x80000000u | (1u << 24) | (uint32_t)(~c));
        } else if (count(c) <= kLeafMax) {
          word = (int)(0x80000000u | ((uint32_t)count(c) << 24) | (uint32_t)first[c]);
        } else if (ninner == 0) {  // capacity exceeded (reported through overflow)
          word = (int)0x80000000u;
        } else {
          fout[q++] = FrontierItem{c, slot};
          word = slot++;
        }
      }
      nd.lox[i] = b[0];
      nd.loy[i] = b[1];
      nd.loz[i] = b[2];
      nd.hix[i] = b[3];
      nd.hiy[i] = b[4];
      nd.hiz[i] = b[5];
      nd.child[i] = word;
    }
    nd.pad[0] = nd.pad[1] = nd.pad[2] = nd.pad[3] = 0;
    nodes[f.wnode] = nd;
  }
}

__global__ void kCollapse(const FrontierItem* __restrict__ fin, const int* __restrict__ nin_ptr,
                          FrontierItem* fout, int* nout, int* node_counter, int node_cap,
                          const int* __restrict__ left, const int* __restrict__ right,
                          const int* __restrict__ first, const int* __restrict__ last,
                          const float* __restrict__ leafbox, const float* __restrict__ innerbox,
                          BNode4* nodes, int* overflow) {
  collapseLevel(blockIdx.x, gridDim.x, fin, nin_ptr, fout, nout, node_counter, node_cap, left, right,
                first, last, leafbox, innerbox, nodes, overflow);
}

// A group of domains collapsed together: one launch per level for all of
// them, blocksPerDom blocks per domain.  Each domain's binary tree sits in
// its own tree slot; cc is its level counter block ([0] node counter, [1]
// overflow, [2 + l] items of level l).
constexpr int kCollapseGroupMax = 16;
struct CollapseDom {
  const int *left, *right, *first, *last;
  const float *leafbox, *innerbox;
  FrontierItem *fa, *fb;
  BNode4* nodes;
  int* cc;
  int node_cap;
};
struct CollapseGroup {
  CollapseDom dom[kCollapseGroupMax];
  int n;
};

// Collapse state of each domain of the group: node counter 1 (the root),
// no overflow, one frontier item (the binary root) at level 0, empty later
// levels.
__global__ void kCollapseGroupInit(CollapseGroup G) {
  const CollapseDom& D = G.dom[blockIdx.x];
  for (int i = threadIdx.x; i < kLevelWords; i += blockDim.x) D.cc[i] = (i == 0 || i == 2) ? 1 : 0;
  if (threadIdx.x == 0) D.fa[0] = FrontierItem{0, 0};
}

__global__ void kCollapseGroupLevel(CollapseGroup G, int l, int blocks_per_dom) {
  const int j = blockIdx.x / blocks_per_dom;  // uniform per block
  const CollapseDom& D = G.dom[j];
  const FrontierItem* fin = (l & 1) ? D.fb : D.fa;
  FrontierItem* fout = (l & 1) ? D.fa : D.fb;
  collapseLevel(blockIdx.x - j * blocks_per_dom, blocks_per_dom, fin, D.cc + 2 + l, fout,
                D.cc + 3 + l, D.cc, D.node_cap, D.left, D.right, D.first, D.last, D.leafbox,
                D.innerbox, D.nodes, D.cc + 1);
}


// Root for tiny domains (<= kLeafMax triangles): one leaf child.
__global__ void kTinyRoot(int nf, const float* __restrict__ leafbox, BNode4* nodes) {
  BNode4 nd;
  float b[6] = {kInfF, kInfF, kInfF, -kInfF, -kInfF, -kInfF};
  for (int i = 0; i < nf; ++i)
    for (int k = 0; k < 3; ++k) {
      b[k] = fminf(b[k], leafbox[i * 6 + k]);
      b[3 + k] = fmaxf(b[3 + k], leafbox[i * 6 + 3 + k]);
    }
  for (int i = 0; i < 4; ++i) {
    bool used = (i == 0);
    nd.lox[i] = used ? b[0] : kInfF;
    nd.loy[i] = used ? b[1] : kInfF;
    nd.loz[i] = used ? b[2] : kInfF;
    nd.hix[i] = used ? b[3] : -kInfF;
    nd.hiy[i] = used ? b[4] : -kInfF;
    nd.hiz[i] = used ? b[5] : -kInfF;
    nd.child[i] = used ? (int)(0x80000000u | ((uint32_t)nf << 24)) : (int)0x80000000u;
    nd.pad[i] = 0;
  }
  nodes[0] = nd;
}

// ---------------------------------------------------------------------------
// PLOC hierarchy (Meister & Bittner, "Parallel Locally-Ordered Clustering
// for Bounding Volume Hierarchy Construction", IEEE TVCG 2018).  Clusters
// start as the leaves in Morton order.  Each pass pairs every cluster with
// the neighbour within +-kPlocRadius array positions whose union box has the
// least surface area, merges mutual pairs into a new inner node that takes
// the lower position, and compacts the array in order.  Distance ties go to
// the lexicographically smallest (lower, upper) position pair, a strict
// order on pairs, so the globally closest pair is always mutual and every
// pass merges at least one pair.  Inner node ids follow creation order; the
// root is the last one (nf - 2).
constexpr int kPlocRadius = 16;
constexpr int kPlocBlock = 256;

__global__ void kPlocInit(const float* __restrict__ vtx, const uint32_t* __restrict__ face, int nf,
                          const uint32_t* __restrict__ order, int* cid, float* cbox) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < nf; i += gridDim.x * blockDim.x) {
    uint32_t p = order[i];
    const float* a = vtx + (size_t)face[p * 3] * 3;
    const float* b = vtx + (size_t)face[p * 3 + 1] * 3;
    const float* c = vtx + (size_t)face[p * 3 + 2] * 3;
    for (int k = 0; k < 3; ++k) {
      cbox[(size_t)i * 6 + k] = fminf(a[k], fminf(b[k], c[k]));
      cbox[(size_t)i * 6 + 3 + k] = fmaxf(a[k], fmaxf(b[k], c[k]));
    }
    cid[i] = ~i;
  }
}

__device__ __forceinline__ float unionArea(const float* a, const float* b) {
  float dx = fmaxf(a[3], b[3]) - fminf(a[0], b[0]);
  float dy = fmaxf(a[4], b[4]) - fminf(a[1], b[1]);
  float dz = fmaxf(a[5], b[5]) - fminf(a[2], b[2]);
  return dx * dy + dy * dz + dz * dx;
}

// nn[i] = nearest neighbour of cluster i; the block's boxes plus a halo of
// kPlocRadius on each side are staged in LDS.
__global__ void __launch_bounds__(kPlocBlock) kPlocNearest(const float* __restrict__ cbox, int n,
                                                            int* __restrict__ nn) {
  constexpr int kSpan = kPlocBlock + 2 * kPlocRadius;
  __shared__ float sb[kSpan * 6];
  const int lo = blockIdx.x * kPlocBlock - kPlocRadius;
  for (int t = threadIdx.x; t < kSpan * 6; t += kPlocBlock) {
    int g = lo * 6 + t;
    sb[t] = (g >= 0 && g < n * 6) ? cbox[g] : 0.f;
  }
  __syncthreads();
  const int i = blockIdx.x * kPlocBlock + threadIdx.x;
  if (i >= n) return;
  float me[6];
  for (int k = 0; k < 6; ++k) me[k] = sb[(i - lo) * 6 + k];
  const int jlo = max(i - kPlocRadius, 0), jhi = min(i + kPlocRadius, n - 1);
  float best = kInfF;
  int bj = -1;
  // ascending j visits the candidate pairs in ascending (lower, upper) order,
  // so a strict comparison keeps the smallest pair among equal areas
  for (int j = jlo; j <= jhi; ++j) {
    if (j == i) continue;
    float a = unionArea(me, sb + (j - lo) * 6);
    if (a < best || bj < 0) {
      best = a;
      bj = j;
    }
  }
  nn[i] = bj;
}

// mflag[i]: cluster i starts a merge (mutual pair, lower position);
// kflag[i]: cluster i keeps a position after the pass.
__global__ void kPlocFlags(const int* __restrict__ nn, int n, uint32_t* mflag, uint32_t* kflag) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i <= n; i += gridDim.x * blockDim.x) {
    if (i == n) {
      mflag[n] = 0u;
      kflag[n] = 0u;
      continue;
    }
    int j = nn[i];
    bool mutual = nn[j] == i;
    mflag[i] = (mutual && i < j) ? 1u : 0u;
    kflag[i] = (mutual && j < i) ? 0u : 1u;
  }
}

__global__ void kPlocMerge(const int* __restrict__ nn, int n, const int* __restrict__ cid,
                           const float* __restrict__ cbox, const uint32_t* __restrict__ mpos,
                           const uint32_t* __restrict__ kpos, int node_base, int* cid2,
                           float* cbox2, int* left, int* right, float* innerbox, int* icount) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    int j = nn[i];
    bool mutual = nn[j] == i;
    if (mutual && j < i) continue;
    uint32_t o = kpos[i];
    if (mutual) {
      int id = node_base + (int)mpos[i];
      int a = cid[i], b = cid[j];
      left[id] = a;
      right[id] = b;
      icount[id] = (a < 0 ? 1 : icount[a]) + (b < 0 ? 1 : icount[b]);
      for (int k = 0; k < 3; ++k) {
        float l = fminf(cbox[(size_t)i * 6 + k], cbox[(size_t)j * 6 + k]);
        float h = fmaxf(cbox[(size_t)i * 6 + 3 + k], cbox[(size_t)j * 6 + 3 + k]);
        innerbox[(size_t)id * 6 + k] = l;
        innerbox[(size_t)id * 6 + 3 + k] = h;
        cbox2[(size_t)o * 6 + k] = l;
        cbox2[(size_t)o * 6 + 3 + k] = h;
      }
      cid2[o] = id;
    } else {
      for (int k = 0; k < 6; ++k) cbox2[(size_t)o * 6 + k] = cbox[(size_t)i * 6 + k];
      cid2[o] = cid[i];
    }
  }
}

// Depth-first leaf ranges, top-down over the nodes one pass created (their
// parents come from later passes and are already placed).  Leaf children are
// renumbered to their depth-first position q, and perm[q] keeps the Morton
// position they came from.
__global__ void kPlocRanges(int id_lo, int id_hi, int* left, int* right,
                            const int* __restrict__ icount, int* first, int* last, uint32_t* perm) {
  for (int id = id_lo + blockIdx.x * blockDim.x + threadIdx.x; id < id_hi;
       id += gridDim.x * blockDim.x) {
    int f = first[id];
    int a = left[id], b = right[id];
    int fb = f + (a < 0 ? 1 : icount[a]);
    if (a < 0) {
      perm[f] = (uint32_t)(~a);
      left[id] = ~f;
    } else {
      first[a] = f;
    }
    if (b < 0) {
      perm[fb] = (uint32_t)(~b);
      right[id] = ~fb;
    } else {
      first[b] = fb;
    }
    last[id] = f + icount[id] - 1;
  }
}

__global__ void kComposeOrder(const uint32_t* __restrict__ perm, const uint32_t* __restrict__ morton,
                              int n, uint32_t* order) {
  for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < n; q += gridDim.x * blockDim.x)
    order[q] = morton[perm[q]];
}

// ---------------------------------------------------------------------------
static int gridFor(long long n, int block) {
  long long g = (n + block - 1) / block;
  if (g < 1) g = 1;
  if (g > 65535) g = 65535;
  return (int)g;
}

// Scratch of one domain build, sized for the largest domain.
constexpr int kBuildStreams = 4;  // domain pipelines of a whole-scene LBVH build
// tree slots of all pipelines together at most this many bytes
constexpr size_t kTreeSlotBudget = (size_t)3 << 29;  // 1.5 GB

// One domain's binary tree (children, leaf ranges, leaf and inner boxes)
// and its collapse frontiers: what the collapse reads after the build.
struct TreeSlot {
  DevBuf lb, ib, lft, rgt, fst, lst, fa, fb;
  // a frontier level holds disjoint subtrees of more than kLeafMax triangles
  static size_t frontierItems(int maxf) { return (size_t)maxf / (kLeafMax + 1) + 64; }
  static size_t bytesFor(int maxf) {
    return (size_t)maxf * (24 + 24 + 16) + 2 * frontierItems(maxf) * sizeof(FrontierItem);
  }
  void reserve(int maxf) {
    const size_t f4 = (size_t)maxf * 4 + 64;
    for (DevBuf* b : {&lft, &rgt, &fst, &lst}) b->reserve(f4);
    lb.reserve((size_t)maxf * 24 + 64);
    ib.reserve((size_t)maxf * 24 + 64);
    fa.reserve(frontierItems(maxf) * sizeof(FrontierItem));
    fb.reserve(frontierItems(maxf) * sizeof(FrontierItem));
  }
  ~TreeSlot() {
    for (DevBuf* b : {&lb, &ib, &lft, &rgt, &fst, &lst, &fa, &fb}) b->release();
  }
};

struct BuildScratch {
  DevBuf keys, vals, keys2, vals2, hist, pin, plf, cnts, scal, adj, rpre, rsuf, rtab, lvl, sums;
  DevBuf pc, pc2, pb, pb2, pnn, pcnt, pmf, pkf, pmp, pkp, pperm, pord;  // PLOC
  std::vector<TreeSlot> slot;  // trees of the domains a group collapse takes together
  void reserve(int maxf, long long maxv, bool ploc, int nslots = 1) {
    const size_t f4 = (size_t)maxf * 4 + 64;
    for (DevBuf* b : {&keys, &vals, &keys2, &vals2, &pin, &plf}) b->reserve(f4);
    if (slot.size() != (size_t)std::max(1, nslots)) {  // fresh slots (no live buffer is copied)
      slot.clear();
      slot.resize(std::max(1, nslots));
    }
    for (TreeSlot& t : slot) t.reserve(maxf);
    hist.reserve((size_t)256 * ((maxf + kRsTile - 1) / kRsTile) * 4 + 64);
    cnts.reserve((size_t)(maxv + 1) * 12 + 64);
    scal.reserve(64);
    lvl.reserve(kLevelWords * 4 + 64);
    // tile sums of this set's scans (vertex offsets, radix histograms, PLOC flags)
    const long long scan_n = std::max<long long>(
        std::max<long long>(maxv + 1, (long long)maxf + 1), 256ll * ((maxf + kRsTile - 1) / kRsTile));
    sums.reserve((size_t)scanSumsWords(scan_n) * 4 + 64);
    adj.reserve((size_t)maxf * 12 + 64);
    rpre.reserve((size_t)maxf * 24 + 64);
    rsuf.reserve((size_t)maxf * 24 + 64);
    RangeLevels L;
    int nlev;
    rtab.reserve((size_t)rangeLevels(std::max(maxf, 1), &L, &nlev) * 4 + 64);
    if (ploc) {
      for (DevBuf* b : {&pc, &pc2, &pnn, &pcnt, &pmf, &pkf, &pmp, &pkp, &pperm, &pord}) b->reserve(f4);
      pb.reserve((size_t)maxf * 24 + 64);
      pb2.reserve((size_t)maxf * 24 + 64);
    }
  }
  ~BuildScratch() {
    for (DevBuf* b : {&keys, &vals, &keys2, &vals2, &hist, &pin, &plf, &cnts, &scal, &adj, &rpre,
                      &rsuf, &rtab, &lvl, &sums, &pc, &pc2, &pb, &pb2, &pnn, &pcnt, &pmf, &pkf, &pmp,
                      &pkp, &pperm, &pord})
      b->release();
  }
};

struct BuildCache {
  std::vector<BuildScratch> S;  // one scratch set per build pipeline
};

void buildCacheRelease(Ctx* c) {
  delete c->build_cache;  // the scratch sets release their buffers
  c->build_cache = nullptr;
}

void HostMeshSource::fill(Ctx* c, int d, float* v, uint32_t* f, uint32_t* rgb, hipStream_t s) {
  (void)c;
  SPRAY_HIP(hipMemcpyAsync(v, m[d].v.data(), m[d].v.size() * 4, hipMemcpyHostToDevice, s));
  SPRAY_HIP(hipMemcpyAsync(f, m[d].f.data(), m[d].f.size() * 4, hipMemcpyHostToDevice, s));
  SPRAY_HIP(hipMemcpyAsync(rgb, m[d].rgb.data(), m[d].rgb.size() * 4, hipMemcpyHostToDevice, s));
}

void buildPool(Ctx* c, MeshSource& src, bool keep_buffers) {
  auto t_start = std::chrono::steady_clock::now();
  DomainPool& P = c->pool;
  hipStream_t s = c->stream;
  int nd = (int)c->scene.domains.size();
  // a cache smaller than the scene pages domains (raw meshes are staged one
  // domain at a time, images kept in pinned host memory)
  const bool paged = c->cache_size > 0 && c->cache_size < nd;
  if (keep_buffers && !paged && !P.paged) {
    P.desc.clear();  // device buffers stay; they grow when too small
  } else {
    P.release();
  }
  P.ndomains = nd;
  P.desc.resize(nd);
  P.nverts.resize(nd);
  long long tv = 0, tf = 0, tn = 0;
  int maxf = 0;
  long long maxv = 0, max_nodes = 0;
  for (int d = 0; d < nd; ++d) {
    const HostDomain& hd = c->scene.domains[d];
    DomainDesc& D = P.desc[d];
    long long nv = src.numVerts(d), nf = src.numFaces(d);
    P.nverts[d] = nv;
    if (nf >= (1 << 24)) throw StatusError(SPRAY_ERR_CAPACITY, "domain has more than 2^24 faces");
    D.vtx_off = paged ? 0 : tv;
    D.face_off = paged ? 0 : tf;
    D.tri_off = tf;
    D.node_off = tn;
    D.shade_off = D.face_off;
    D.num_tris = (int)nf;
    D.num_nodes = 0;
    for (int k = 0; k < 3; ++k) {
      D.lo[k] = hd.lo[k];
      D.hi[k] = hd.hi[k];
    }
    D.mtl = hd.mtl;
    D.eta_a = hd.eta_a;
    D.eta_b = hd.eta_b;
    tv += nv;
    tf += nf;
    tn += nodeCap(nf);
    maxf = std::max(maxf, (int)nf);
    maxv = std::max(maxv, nv);
    max_nodes = std::max(max_nodes, nodeCap(nf));
  }
  P.total_verts = tv;
  P.total_tris = tf;
  if (paged) {
    P.paged = true;
    P.capacity = c->cache_size;
    P.block_nodes = max_nodes;
    P.block_tris = maxf;
    P.vtx.reserve(maxv * 12 + 64);
    P.face.reserve((size_t)maxf * 12 + 64);
    P.rgb.reserve(maxv * 4 + 64);
    P.nrm.reserve(maxv * 12 + 64);
    P.tris.reserve((size_t)P.capacity * maxf * sizeof(LTri) + 64);
    P.fshade.reserve((size_t)P.capacity * maxf * 48 + 64);
    P.nodes.reserve((size_t)P.capacity * max_nodes * sizeof(BNode4) + 64);
    size_t bytes = 0;
    P.h_image_off.resize(nd);
    for (int d = 0; d < nd; ++d) {
      P.h_image_off[d] = (long long)bytes;
      const size_t nf = (size_t)P.desc[d].num_tris;
      bytes += ((size_t)nodeCap((long long)nf) * sizeof(BNode4) + nf * (sizeof(LTri) + 48) + 255) & ~(size_t)255;
    }
    SPRAY_HIP(hipHostMalloc(&P.h_image, bytes, hipHostMallocDefault));
    P.h_nrm.resize(nd);
    SPRAY_HIP(hipStreamCreate(&P.copy_stream));
  } else {
    P.vtx.reserve(tv * 12);
    P.face.reserve(tf * 12);
    P.rgb.reserve(tv * 4);
    P.nrm.reserve(tv * 12);
    P.tris.reserve(tf * sizeof(LTri));
    P.fshade.reserve(tf * 48 + 64);
    P.nodes.reserve(tn * sizeof(BNode4));
  }
  // A whole-scene LBVH build runs up to kBuildStreams domain pipelines side
  // by side, one per free hardware queue (domain d on stream d % K, each
  // stream with its own scratch set), so the many small per-domain kernels of different domains overlap;
  // each domain's collapse launches a fixed number of levels with no host
  // wait, and the level counts of every domain are checked once after the
  // last.  A source whose fills may run concurrently fills each domain in
  // its pipeline.  PLOC and paged builds stay in one stream with a host
  // check per domain.
  const bool multi = !paged && c->bvh_builder == 0 && nd > 1;
  const bool pipe_fill = multi && src.pipelined();
  if (!paged && !pipe_fill) {
    for (int d = 0; d < nd; ++d) {
      const DomainDesc& D = P.desc[d];
      src.fill(c, d, P.vtx.as<float>() + D.vtx_off * 3, P.face.as<uint32_t>() + D.face_off * 3,
               P.rgb.as<uint32_t>() + D.vtx_off, s);
    }
    src.check(c);
  }
  DevBuf lvl_all;
  std::vector<int> launched(nd, 0);  // collapse levels launched per domain (multi)
  // a remesh of the same domain structure launches the previous depth + 4
  // levels (a domain that needs more is rebuilt with host-checked levels)
  std::vector<int> prev_depth;
  if (keep_buffers && (int)P.bvh_depth.size() == nd) prev_depth = P.bvh_depth;
  struct ReleaseLvl {
    DevBuf& lvl;
    ~ReleaseLvl() { lvl.release(); }
  } release_lvl{lvl_all};
  // one domain's normals, shading records and BVH, issued on stream s, its
  // binary tree in tree slot TS; grouped: the collapse is left to the
  // domain's group (no host wait), else it is host-checked here
  auto buildOne = [&](BuildScratch& S, TreeSlot& TS, int d, hipStream_t s, bool grouped) {
    DomainDesc& D = P.desc[d];
    const int nf = D.num_tris;
    const long long nv = src.numVerts(d);
    if (paged) {  // build into block d % capacity from the staged mesh
      const int b = d % P.capacity;
      D.node_off = (long long)b * P.block_nodes;
      D.tri_off = D.shade_off = (long long)b * P.block_tris;
      src.fill(c, d, P.vtx.as<float>(), P.face.as<uint32_t>(), P.rgb.as<uint32_t>(), s);
      src.check(c);
    }
    const float* vtx = P.vtx.as<float>() + D.vtx_off * 3;
    const uint32_t* face = P.face.as<uint32_t>() + D.face_off * 3;
    if (pipe_fill)
      src.fill(c, d, P.vtx.as<float>() + D.vtx_off * 3, P.face.as<uint32_t>() + D.face_off * 3,
               P.rgb.as<uint32_t>() + D.vtx_off, s);
    // normals
    uint32_t* cnt = S.cnts.as<uint32_t>();
    uint32_t* cur = cnt + (nv + 1);  // cleared with cnt in one fill
    uint32_t* off = cur + (nv + 1);
    SPRAY_HIP(hipMemsetAsync(cnt, 0, (size_t)(nv + 1) * 8, s));
    kVertCount<<<gridFor((long long)nf * 3, 256), 256, 0, s>>>(face, nf, cnt);
    exclusiveScan(cnt, off, nv + 1, s, S.sums.as<uint32_t>());
    kVertFill<<<gridFor((long long)nf * 3, 256), 256, 0, s>>>(face, nf, off, cur,
                                                              S.adj.as<uint32_t>());
    kVertNormal<<<gridFor(nv, 256), 256, 0, s>>>(vtx, face, nv, off, cnt, S.adj.as<uint32_t>(),
                                                  P.nrm.as<float>() + D.vtx_off * 3);
    // Morton codes (kFaceShade sets the centroid bounds' starting values)
    uint32_t* bmm = S.scal.as<uint32_t>();
    kFaceShade<<<gridFor(nf, 256), 256, 0, s>>>(face, nf, P.nrm.as<float>() + D.vtx_off * 3,
                                                 P.rgb.as<uint32_t>() + D.vtx_off,
                                                 P.fshade.as<float4>() + D.shade_off * 3, bmm);
    kCentroidBounds<<<std::min(gridFor(nf, 256), kBoundsGrid), 256, 0, s>>>(vtx, face, nf, bmm,
                                                                            bmm + 3);
    kMorton<<<gridFor(nf, 256), 256, 0, s>>>(vtx, face, nf, bmm, bmm + 3, S.keys.as<uint32_t>(),
                                             S.vals.as<uint32_t>());
    // radix sort, 4 passes of 8 bits
    int nb = (nf + kRsTile - 1) / kRsTile;
    uint32_t *ka = S.keys.as<uint32_t>(), *va = S.vals.as<uint32_t>(), *kb = S.keys2.as<uint32_t>(),
             *vb = S.vals2.as<uint32_t>();
    for (int pass = 0; pass < 4; ++pass) {
      kRsHist<<<nb, kRsBlock, 0, s>>>(ka, nf, pass * 8, S.hist.as<uint32_t>());
      exclusiveScan(S.hist.as<uint32_t>(), S.hist.as<uint32_t>(), (long long)256 * nb, s,
                    S.sums.as<uint32_t>());
      kRsScatter<<<nb, kRsBlock, 0, s>>>(ka, va, kb, vb, nf, pass * 8, S.hist.as<uint32_t>());
      std::swap(ka, kb);
      std::swap(va, vb);
    }
    LTri* tris = P.tris.as<LTri>() + D.tri_off;
    BNode4* nodes = P.nodes.as<BNode4>() + D.node_off;
    float* leafbox = TS.lb.as<float>();
    const int node_cap = (int)nodeCap(nf);
    // paged pool: the finished domain's image goes to pinned host memory
    // (and its normals to the host, for the normals query)
    auto pageOut = [&]() {
      if (!paged) return;
      char* dst = (char*)P.h_image + P.h_image_off[d];
      const size_t nb = (size_t)D.num_nodes * sizeof(BNode4), tb = (size_t)nf * sizeof(LTri);
      SPRAY_HIP(hipMemcpyAsync(dst, nodes, nb, hipMemcpyDeviceToHost, s));
      SPRAY_HIP(hipMemcpyAsync(dst + nb, tris, tb, hipMemcpyDeviceToHost, s));
      SPRAY_HIP(hipMemcpyAsync(dst + nb + tb, P.fshade.as<float4>() + D.shade_off * 3,
                               (size_t)nf * 48, hipMemcpyDeviceToHost, s));
      P.h_nrm[d].resize((size_t)nv * 3);
      SPRAY_HIP(hipMemcpyAsync(P.h_nrm[d].data(), P.nrm.p, (size_t)nv * 12, hipMemcpyDeviceToHost, s));
      SPRAY_HIP(hipStreamSynchronize(s));
    };
    if (nf <= kLeafMax) {
      kLeaves<<<gridFor(nf, 256), 256, 0, s>>>(vtx, face, nf, va, leafbox, tris);
      kTinyRoot<<<1, 1, 0, s>>>(nf, leafbox, nodes);
      D.num_nodes = 1;
      P.bvh_depth[d] = 1;
      pageOut();
      return;
    }
    int* L = TS.lft.as<int>();
    int* R = TS.rgt.as<int>();
    int* F = TS.fst.as<int>();
    int* Lst = TS.lst.as<int>();
    float* inner = TS.ib.as<float>();
    // Collapse the binary tree under `root` into 4-wide nodes; returns the
    // number of 4-wide levels.
    // Levels are launched kLevelBatch at a time; each reads its item count
    // from the device, so the host waits once per batch (a level past the
    // last finds 0 items).  Returns the number of levels with items
    // (kMaxLevels + 1 when deeper than kMaxLevels).
    auto collapse = [&](int root) {
      int* cc = S.lvl.as<int>();  // [0] node counter, [1] overflow, [2 + l] items of level l
      int init[kLevelWords] = {};
      init[0] = 1;
      init[2] = 1;
      SPRAY_HIP(hipMemcpyAsync(cc, init, sizeof(init), hipMemcpyHostToDevice, s));
      FrontierItem item{root, 0};
      SPRAY_HIP(hipMemcpyAsync(TS.fa.p, &item, sizeof(item), hipMemcpyHostToDevice, s));
      const int grid = std::min(gridFor(std::max(1, nf / 8), 256), 2048);
      int done = 0;
      for (;;) {
        const int upto = std::min(done + kLevelBatch, kMaxLevels);
        for (int l = done; l < upto; ++l) {
          FrontierItem* fin = (l & 1) ? TS.fb.as<FrontierItem>() : TS.fa.as<FrontierItem>();
          FrontierItem* fout = (l & 1) ? TS.fa.as<FrontierItem>() : TS.fb.as<FrontierItem>();
          kCollapse<<<grid, 256, 0, s>>>(fin, cc + 2 + l, fout, cc + 3 + l, cc, node_cap, L, R, F,
                                         Lst, leafbox, inner, nodes, cc + 1);
        }
        done = upto;
        int h[kLevelWords];
        SPRAY_HIP(hipMemcpyAsync(h, cc, sizeof(h), hipMemcpyDeviceToHost, s));
        SPRAY_HIP(hipStreamSynchronize(s));
        if (h[1]) throw StatusError(SPRAY_ERR_CAPACITY, "BVH node capacity exceeded");
        D.num_nodes = h[0];
        if (h[2 + done] == 0 || done == kMaxLevels) {
          int depth = 0;
          while (depth < done && h[2 + depth] > 0) ++depth;
          return h[2 + done] > 0 ? kMaxLevels + 1 : depth;
        }
      }
    };
    // worst-case stack need: three pushes per 4-wide level
    auto fitsStack = [](int depth) { return 3 * depth + 1 <= kRing + kOverflow; };
    int depth = -1;
    if (c->bvh_builder == 1) {
      // PLOC passes until one cluster is left
      int* cid = S.pc.as<int>();
      int* cid2 = S.pc2.as<int>();
      float* cbox = S.pb.as<float>();
      float* cbox2 = S.pb2.as<float>();
      int* nn = S.pnn.as<int>();
      int* icount = S.pcnt.as<int>();
      uint32_t* mflag = S.pmf.as<uint32_t>();
      uint32_t* kflag = S.pkf.as<uint32_t>();
      uint32_t* mpos = S.pmp.as<uint32_t>();
      uint32_t* kpos = S.pkp.as<uint32_t>();
      kPlocInit<<<gridFor(nf, 256), 256, 0, s>>>(vtx, face, nf, va, cid, cbox);
      std::vector<int> pass_base{0};
      int n = nf, made = 0;
      bool ok = true;
      while (n > 1) {
        kPlocNearest<<<(n + kPlocBlock - 1) / kPlocBlock, kPlocBlock, 0, s>>>(cbox, n, nn);
        kPlocFlags<<<gridFor(n + 1, 256), 256, 0, s>>>(nn, n, mflag, kflag);
        exclusiveScan(mflag, mpos, n + 1, s, S.sums.as<uint32_t>());
        exclusiveScan(kflag, kpos, n + 1, s, S.sums.as<uint32_t>());
        kPlocMerge<<<gridFor(n, 256), 256, 0, s>>>(nn, n, cid, cbox, mpos, kpos, made, cid2, cbox2,
                                                    L, R, inner, icount);
        uint32_t h[2];
        SPRAY_HIP(hipMemcpyAsync(&h[0], mpos + n, 4, hipMemcpyDeviceToHost, s));
        SPRAY_HIP(hipMemcpyAsync(&h[1], kpos + n, 4, hipMemcpyDeviceToHost, s));
        SPRAY_HIP(hipStreamSynchronize(s));
        if (h[0] == 0 || (int)h[1] != n - (int)h[0]) {  // no progress: non-finite boxes
          ok = false;
          break;
        }
        made += (int)h[0];
        pass_base.push_back(made);
        n = (int)h[1];
        std::swap(cid, cid2);
        std::swap(cbox, cbox2);
      }
      if (ok) {
        // depth-first leaf order: roots of later passes first
        const int root = made - 1;
        uint32_t* perm = S.pperm.as<uint32_t>();
        SPRAY_HIP(hipMemsetAsync(F + root, 0, 4, s));
        for (int p = (int)pass_base.size() - 1; p > 0; --p) {
          int lo = pass_base[p - 1], hi = pass_base[p];
          kPlocRanges<<<gridFor(hi - lo, 256), 256, 0, s>>>(lo, hi, L, R, icount, F, Lst, perm);
        }
        uint32_t* order = S.pord.as<uint32_t>();
        kComposeOrder<<<gridFor(nf, 256), 256, 0, s>>>(perm, va, nf, order);
        kLeaves<<<gridFor(nf, 256), 256, 0, s>>>(vtx, face, nf, order, leafbox, tris);
        depth = collapse(root);
        if (!fitsStack(depth)) depth = -1;
      }
    }
    if (depth < 0) {
      // Karras LBVH: leaves in Morton order, bottom-up refit
      int* PI = S.pin.as<int>();
      int* PL = S.plf.as<int>();
      kLeaves<<<gridFor(nf, 256), 256, 0, s>>>(vtx, face, nf, va, leafbox, tris, PI);
      kKarras<<<gridFor(nf, 256), 256, 0, s>>>(ka, nf, L, R, F, Lst, PI, PL);
      // inner boxes as unions over leaf ranges (replaces a bottom-up refit)
      RangeLevels RL;
      int nlev;
      rangeLevels(nf, &RL, &nlev);
      const long long nblk = (nf + kRangeBlock - 1) / kRangeBlock;
      float* tab = S.rtab.as<float>();
      kLeafScan<<<gridFor(nblk * 64, 256), 256, 0, s>>>(leafbox, nf, S.rpre.as<float>(),
                                                        S.rsuf.as<float>(), tab);
      for (int j = 1; j < nlev; j += 2) {  // two levels per launch
        const long long m = nblk - (1ll << j) + 1;
        if (j + 1 < nlev)
          kRangeLevel2<<<gridFor(m, 256), 256, 0, s>>>(tab + RL.off[j - 1], tab + RL.off[j],
                                                       tab + RL.off[j + 1], m,
                                                       nblk - (1ll << (j + 1)) + 1, 1ll << (j - 1));
        else
          kRangeLevel<<<gridFor(m, 256), 256, 0, s>>>(tab + RL.off[j - 1], tab + RL.off[j], m,
                                                      1ll << (j - 1));
      }
      kRangeBoxes<<<gridFor(nf - 1, 256), 256, 0, s>>>(nf - 1, F, Lst, leafbox, S.rpre.as<float>(),
                                                       S.rsuf.as<float>(), tab, RL, inner);
      if (grouped) return;  // collapsed with its group
      depth = collapse(0);
      if (!fitsStack(depth))
        throw StatusError(SPRAY_ERR_CAPACITY, "BVH too deep for the traversal stack");
    }
    P.bvh_depth[d] = depth;
    pageOut();
  };
  P.bvh_depth.assign(nd, 0);
  if (!multi) {
    BuildScratch S;
    S.reserve(maxf, maxv, c->bvh_builder == 1);
    for (int d = 0; d < nd; ++d) buildOne(S, S.slot[0], d, s, false);
  } else {
    const int K = std::min(pipelineCount(kBuildStreams), nd);
    // each pipeline keeps the binary trees of up to G domains and collapses
    // them together, one launch per level for the group
    const size_t slot_b = std::max<size_t>(TreeSlot::bytesFor(maxf), 1);
    const int G = (int)std::max<size_t>(
        1, std::min<size_t>({(size_t)kCollapseGroupMax, (size_t)((nd + K - 1) / K),
                             kTreeSlotBudget / ((size_t)K * slot_b)}));
    // a remesh keeps the pipelines' scratch sets for the next remesh (they
    // only grow), so repeated remeshes allocate nothing here; other builds
    // hold no scratch afterwards
    std::vector<BuildScratch> local;
    std::vector<BuildScratch>* Sv = &local;
    if (src.rebuilds() && keep_buffers) {
      if (!c->build_cache) c->build_cache = new BuildCache;
      Sv = &c->build_cache->S;
    } else {
      buildCacheRelease(c);
    }
    if ((int)Sv->size() != K) {  // fresh sets (no live buffer is copied)
      Sv->clear();
      Sv->resize(K);
    }
    std::vector<BuildScratch>& S = *Sv;
    for (auto& x : S) x.reserve(maxf, maxv, false, G);
    lvl_all.reserve((size_t)nd * kLevelWords * 4 + 64);
    // pipeline 0 on s, the others on their own streams (declared after the
    // scratch sets: joined before they go)
    StreamSet side(K - 1, s);
    auto pipe = [&](int k) { return k == 0 ? s : side[k - 1]; };
    std::vector<std::vector<int>> pending(K);  // domains of each pipeline's open group
    auto flush = [&](int k) {
      std::vector<int>& g = pending[k];
      if (g.empty()) return;
      CollapseGroup CG{};
      CG.n = (int)g.size();
      int levels = 1, bpd = 1;
      for (int j = 0; j < CG.n; ++j) {
        const int d = g[j];
        const DomainDesc& D = P.desc[d];
        TreeSlot& TS = S[k].slot[j];
        CollapseDom& X = CG.dom[j];
        X.left = TS.lft.as<int>();
        X.right = TS.rgt.as<int>();
        X.first = TS.fst.as<int>();
        X.last = TS.lst.as<int>();
        X.leafbox = TS.lb.as<float>();
        X.innerbox = TS.ib.as<float>();
        X.fa = TS.fa.as<FrontierItem>();
        X.fb = TS.fb.as<FrontierItem>();
        X.nodes = P.nodes.as<BNode4>() + D.node_off;
        X.cc = lvl_all.as<int>() + (size_t)d * kLevelWords;
        X.node_cap = (int)nodeCap(D.num_tris);
        levels = std::max(levels, prev_depth.empty() ? kMaxLevels
                                                     : std::min(kMaxLevels, prev_depth[d] + 4));
        bpd = std::max(bpd, std::min(gridFor(std::max(1, D.num_tris / 8), 256), 2048));
      }
      kCollapseGroupInit<<<CG.n, 64, 0, pipe(k)>>>(CG);
      for (int l = 0; l < levels; ++l) {
        // level l holds at most 4^l items (one root, at most four children per item)
        const int b = l < 12 ? std::min(bpd, (int)(((1ll << (2 * l)) + 255) / 256)) : bpd;
        kCollapseGroupLevel<<<CG.n * b, 256, 0, pipe(k)>>>(CG, l, b);
      }
      SPRAY_HIP(hipGetLastError());
      for (int d : g) launched[d] = levels;
      g.clear();
    };
    for (int d = 0; d < nd; ++d) {
      const int k = d % K;
      const bool grouped = P.desc[d].num_tris > kLeafMax;  // one-leaf roots are built directly
      buildOne(S[k], S[k].slot[pending[k].size()], d, pipe(k), grouped);
      if (grouped) {
        pending[k].push_back(d);
        if ((int)pending[k].size() == G) flush(k);
      }
    }
    for (int k = 0; k < K; ++k) flush(k);
    side.release();
    SPRAY_HIP(hipStreamSynchronize(s));  // pipeline 0
    if (pipe_fill) src.check(c);
    std::vector<int> h((size_t)nd * kLevelWords);
    SPRAY_HIP(hipMemcpy(h.data(), lvl_all.p, h.size() * 4, hipMemcpyDeviceToHost));
    std::vector<int> redo;
    for (int d = 0; d < nd; ++d) {
      if (P.desc[d].num_tris <= kLeafMax) continue;  // one-leaf root, built directly
      const int* cc = &h[(size_t)d * kLevelWords];
      if (cc[1]) throw StatusError(SPRAY_ERR_CAPACITY, "BVH node capacity exceeded");
      if (cc[2 + launched[d]] > 0) {  // items left after the last launched level
        if (launched[d] >= kMaxLevels)
          throw StatusError(SPRAY_ERR_CAPACITY, "BVH too deep for the traversal stack");
        redo.push_back(d);
        continue;
      }
      int depth = 0;
      while (depth < launched[d] && cc[2 + depth] > 0) ++depth;
      if (3 * depth + 1 > kRing + kOverflow)
        throw StatusError(SPRAY_ERR_CAPACITY, "BVH too deep for the traversal stack");
      P.desc[d].num_nodes = cc[0];
      P.bvh_depth[d] = depth;
    }
    for (int d : redo) buildOne(S[0], S[0].slot[0], d, s, false);  // deeper than last time: host-checked
  }
  SPRAY_HIP(hipStreamSynchronize(s));
  long long used = 0;
  for (auto& D : P.desc) used += D.num_nodes;
  P.total_nodes = used;
  if (paged) {
    // the last `capacity` domains built stay in their blocks (least recently
    // built first); the raw-mesh staging arrays are no longer needed
    std::vector<std::pair<int, int>> resident;
    for (int d = std::max(0, nd - P.capacity); d < nd; ++d) resident.push_back({d, d % P.capacity});
    for (DevBuf* b : {&P.vtx, &P.face, &P.rgb, &P.nrm}) b->release();
    P.d_desc.reserve(sizeof(DomainDesc) * nd);
    SPRAY_HIP(hipMemcpy(P.d_desc.p, P.desc.data(), sizeof(DomainDesc) * nd, hipMemcpyHostToDevice));
    poolInitPaging(c, resident);
  } else {
    P.d_desc.reserve(sizeof(DomainDesc) * nd);
    SPRAY_HIP(hipMemcpy(P.d_desc.p, P.desc.data(), sizeof(DomainDesc) * nd, hipMemcpyHostToDevice));
  }
  P.build_ms =
      std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_start).count();
}

// ---------------------------------------------------------------------------
// Batch queries (parity entry points)
constexpr int kQBlock = 256;

__global__ void __launch_bounds__(kQBlock) kTraceRays(
    const BNode4* __restrict__ nodes, const LTri* __restrict__ tris, const float* __restrict__ org,
    const float* __restrict__ dir, const float* __restrict__ tnear, const float* __restrict__ tfar,
    long long n, int mode, uint32_t max_nodes, float* t_o, float* u_o, float* v_o, int* prim_o,
    uint32_t* ovf, int* err) {
  __shared__ uint32_t ring[kRing * kQBlock];
  LaneStack st;
  st.init(ring, ovf);
  TravCount tc;
  TravState s;
  // the loop bound is uniform per block, so every wave runs the phases whole
  for (long long base = blockIdx.x * (long long)blockDim.x; base < n;
       base += (long long)gridDim.x * blockDim.x) {
    const long long i = base + threadIdx.x;
    const bool active = i < n;
    if (active) {
      float o[3] = {org[i * 3], org[i * 3 + 1], org[i * 3 + 2]};
      float d[3] = {dir[i * 3], dir[i * 3 + 1], dir[i * 3 + 2]};
      travInit(s, 0u, 0u, o, d, tnear[i], tfar[i], mode != 0, st);
    }
    while (__ballot(active && !travDone(s)) != 0ull) {
      travNodePhase(s, st, tc, active, nodes);
      travLeafPhase(s, st, tc, active, tris);
    }
    if (active) {
      if (mode == 0) {
        t_o[i] = s.tlim;
        u_o[i] = s.u;
        v_o[i] = s.v;
        prim_o[i] = s.prim;
      } else {
        prim_o[i] = s.prim > 0 ? 1 : 0;
      }
    }
  }
  if (st.overflowed()) atomicExch(err, 1);
}

void traceRaysDevice(Ctx* c, const float* d_org, const float* d_dir, const float* d_tnear,
                     const float* d_tfar, size_t n, int mode, int domain, float* d_t, float* d_u,
                     float* d_v, int* d_prim) {
  const DomainDesc& D = c->pool.desc[domain];
  int grid = std::min<long long>(2048, std::max<long long>(1, ((long long)n + kQBlock - 1) / kQBlock));
  c->q_out.reserve((size_t)grid * kQBlock * kOverflow * 4 + 64);
  int* err = (int*)((char*)c->q_out.p + (size_t)grid * kQBlock * kOverflow * 4);
  SPRAY_HIP(hipMemsetAsync(err, 0, 4, c->stream));
  kTraceRays<<<grid, kQBlock, 0, c->stream>>>(
      c->pool.nodes.as<BNode4>() + D.node_off, c->pool.tris.as<LTri>() + D.tri_off, d_org, d_dir,
      d_tnear, d_tfar, (long long)n, mode, (uint32_t)D.num_nodes, d_t, d_u, d_v, d_prim,
      c->q_out.as<uint32_t>(), err);
  int h = 0;
  SPRAY_HIP(hipMemcpyAsync(&h, err, 4, hipMemcpyDeviceToHost, c->stream));
  SPRAY_HIP(hipStreamSynchronize(c->stream));
  if (h) throw StatusError(SPRAY_ERR_CAPACITY, "traversal stack overflow or corrupt BVH");
}

// ---------------------------------------------------------------------------
// spray_trace_rays_device: the same query on caller-owned device arrays with
// no allocation.  Its scratch (the traversal overflow stacks of at most
// kScratchBlocks blocks, then one error word) is this file-scope array,
// which every GPU the library runs on holds until the process ends.  Host
// code never uses the array's name as an address: scratchInit looks up each
// device's copy once (hipGetSymbolAddress) and keeps the pointer per device.
constexpr int kScratchBlocks = 256;
constexpr size_t kScratchStackBytes = (size_t)kScratchBlocks * kQBlock * kOverflow * 4;
constexpr size_t kScratchBytes = kScratchStackBytes + 256;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch[kMaxDevices] = {};
}  // namespace

__global__ void kClearWord(int* w) { *w = 0; }

void scratchInit(int device) {
  if (device < 0 || device >= kMaxDevices)
    throw StatusError(SPRAY_ERR_DEVICE, "device ordinal beyond the scratch table");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch[device]) return;
  int cur = 0;
  SPRAY_HIP(hipGetDevice(&cur));
  SPRAY_HIP(hipSetDevice(device));
  void* ptr = nullptr;
  const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
  (void)hipSetDevice(cur);
  if (e != hipSuccess || ptr == nullptr)
    throw DeviceError(std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
  g_scratch[device] = ptr;
}

size_t scratchBytes() { return kScratchBytes; }

void traceRaysNoAlloc(Ctx* c, const float* d_org, const float* d_dir, const float* d_tnear,
                      const float* d_tfar, size_t n, int mode, int domain, float* d_t, float* d_u,
                      float* d_v, int* d_prim) {
  void* base = nullptr;
  {
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    base = (c->device >= 0 && c->device < kMaxDevices) ? g_scratch[c->device] : nullptr;
  }
  if (!base) throw StatusError(SPRAY_ERR_STATE, "scratch not set up on this device");
  const DomainDesc& D = c->pool.desc[domain];
  const int grid =
      (int)std::min<long long>(kScratchBlocks, std::max<long long>(1, ((long long)n + kQBlock - 1) / kQBlock));
  uint32_t* ovf = (uint32_t*)base;
  int* err = (int*)((char*)base + kScratchStackBytes);
  kClearWord<<<1, 1, 0, c->stream>>>(err);
  kTraceRays<<<grid, kQBlock, 0, c->stream>>>(
      c->pool.nodes.as<BNode4>() + D.node_off, c->pool.tris.as<LTri>() + D.tri_off, d_org, d_dir,
      d_tnear, d_tfar, (long long)n, mode, (uint32_t)D.num_nodes, d_t, d_u, d_v, d_prim, ovf, err);
  SPRAY_HIP(hipGetLastError());
  int h = 0;
  SPRAY_HIP(hipMemcpyAsync(&h, err, 4, hipMemcpyDeviceToHost, c->stream));
  SPRAY_HIP(hipStreamSynchronize(c->stream));
  if (h) throw StatusError(SPRAY_ERR_CAPACITY, "traversal stack overflow or corrupt BVH");
}

// Domain list per ray: brute force over the domain AABBs staged in LDS,
// insertion into a (t, id)-sorted list of at most maxk entries.
constexpr int kMaxList = 64;
__global__ void __launch_bounds__(256) kDomainList(const DomainDesc* __restrict__ desc, int nd,
                                                   const float* __restrict__ org,
                                                   const float* __restrict__ dir, long long n,
                                                   int maxk, int* counts, int* ids, float* ts) {
  extern __shared__ float boxes[];  // 6 floats per domain
  for (int i = threadIdx.x; i < nd; i += blockDim.x)
    for (int k = 0; k < 3; ++k) {
      boxes[i * 6 + k] = desc[i].lo[k];
      boxes[i * 6 + 3 + k] = desc[i].hi[k];
    }
  __syncthreads();
  for (long long r = blockIdx.x * (long long)blockDim.x + threadIdx.x; r < n;
       r += (long long)gridDim.x * blockDim.x) {
    float o[3] = {org[r * 3], org[r * 3 + 1], org[r * 3 + 2]};
    float d[3] = {dir[r * 3], dir[r * 3 + 1], dir[r * 3 + 2]};
    int cnt = 0, kept = 0;
    int* my_ids = ids + r * maxk;
    float* my_ts = ts + r * maxk;
    for (int i = 0; i < nd; ++i) {
      float tmin;
      if (!slabDomain(&boxes[i * 6], &boxes[i * 6 + 3], o, d, &tmin)) continue;
      ++cnt;
      // insert (tmin, i) keeping the maxk smallest, ordered by (t, id)
      int j = kept;
      if (kept == maxk) {
        float lt = my_ts[maxk - 1];
        if (!(tmin < lt)) continue;  // ids ascend, so equal t keeps the earlier id
        j = maxk - 1;
      } else {
        ++kept;
      }
      while (j > 0 && my_ts[j - 1] > tmin) {
        my_ts[j] = my_ts[j - 1];
        my_ids[j] = my_ids[j - 1];
        --j;
      }
      my_ts[j] = tmin;
      my_ids[j] = i;
    }
    counts[r] = cnt;
  }
}

void domainListDevice(Ctx* c, const float* d_org, const float* d_dir, size_t n, int maxk,
                      int* d_counts, int* d_ids, float* d_ts) {
  int nd = c->pool.ndomains;
  int grid = std::min<long long>(4096, std::max<long long>(1, ((long long)n + 255) / 256));
  kDomainList<<<grid, 256, nd * 6 * sizeof(float), c->stream>>>(
      c->pool.d_desc.as<DomainDesc>(), nd, d_org, d_dir, (long long)n, maxk, d_counts, d_ids, d_ts);
  SPRAY_HIP(hipGetLastError());
  SPRAY_HIP(hipStreamSynchronize(c->stream));
}

}  // namespace spray
