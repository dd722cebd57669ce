// Domain-hit predictor of the learned schedule: a 22-128-128-1 MLP
// (input: ray origin normalised to the scene bound, ray direction, and a
// 16-dim embedding of the domain) predicting whether a domain produces the
// ray's final closest hit.  bf16 weights/activations on MFMA
// (v_mfma_f32_16x16x32_bf16), fp32 accumulation, fp32 master weights, Adam.
//
// Training data are (origin, direction, domain, label) samples logged by
// the tracer.  Inference orders speculative domain visits: per radiance ray
// with several candidate domains, the network is asked about the nearest
// candidate (and the second, when the nearest is predicted to miss); the
// domains entered no later than the first one predicted to hold the hit run
// in phase A of a pass, the rest are deferred to the end of the pass's
// launch (or, in small passes, to a second launch), where the tdom <=
// history filter prunes most of them.  Exact traversal still decides every
// hit, so the image does not depend on the network.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <numeric>
#include <string>
#include <vector>

#include "ctx.h"
#include "tracer_dev.h"

namespace spray {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int kIn = 22, kInPad = 32, kHid = 128, kEmb = 16, kRayFeat = 6;
// flat fp32 parameter layout
constexpr int oW1 = 0;                     // [128][22]
constexpr int oB1 = oW1 + kHid * kIn;      // [128]
constexpr int oW2 = oB1 + kHid;            // [128][128] (out, in)
constexpr int oB2 = oW2 + kHid * kHid;     // [128]
constexpr int oW3 = oB2 + kHid;            // [128]
constexpr int oB3 = oW3 + kHid;            // [1]
constexpr int oEmb = oB3 + 1;              // [nd][16]
constexpr int kDenseParams = oEmb;         // 19585

struct MlpDev {
  const __bf16* w1f;   // inference fragments of W1 and W2 (kPackFrags)
  const __bf16* w2f;
  const __bf16* w1;    // [128][32] bf16, columns 22..31 zero
  const __bf16* w2;    // [128][128]
  const __bf16* w2t;   // [128][128] transposed (in, out)
  const __bf16* w1et;  // [16][128]: w1et[e][o] = W1[o][6 + e]
  const float* p;      // fp32 master (b1, b2, w3, b3, emb read from here)
  float center[3], inv_half[3];
};

__device__ __forceinline__ __bf16 tobf(float x) { return (__bf16)x; }

// LDS written by some lanes and read by others of the same wave
__device__ __forceinline__ void waveSync() {
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
}

// Features of one (ray, domain) sample into a bf16 row of 32.
__device__ __forceinline__ void features(const MlpDev& M, const float o[3], const float d[3],
                                         int dom, float x[kInPad]) {
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    x[k] = (o[k] - M.center[k]) * M.inv_half[k];
    x[3 + k] = d[k];
  }
  const float* e = M.p + oEmb + (size_t)dom * kEmb;
#pragma unroll
  for (int k = 0; k < kEmb; ++k) x[kRayFeat + k] = e[k];
#pragma unroll
  for (int k = kIn; k < kInPad; ++k) x[k] = 0.f;
}

// One wave computes 16 samples.  LDS rows are row-major bf16.
//   A operand (16 x 32 slice): lane l -> row l & 15, k = 8 * (l >> 4) + j
//   B operand from W[out][in]: lane l -> out = 16 * nb + (l & 15), same k
//   D: lane l -> rows 4 * (l >> 4) + i, column 16 * nb + (l & 15)
__device__ __forceinline__ bf16x8 ldsA(const __bf16* rows, int ld, int kb) {
  const int l = __lane_id();
  return *(const bf16x8*)(rows + (l & 15) * ld + kb * 32 + 8 * (l >> 4));
}
__device__ __forceinline__ bf16x8 ldB(const __bf16* w, int ld, int nb, int kb) {
  const int l = __lane_id();
  return *(const bf16x8*)(w + (size_t)(16 * nb + (l & 15)) * ld + kb * 32 + 8 * (l >> 4));
}

// acc[nb] (8 column blocks of 16) = A (16 x K) * W^T, K = 32 * nk
__device__ __forceinline__ void gemmRows(const __bf16* a_lds, int lda, const __bf16* w, int ldw,
                                         int nk, f32x4 acc[8]) {
#pragma unroll
  for (int nb = 0; nb < 8; ++nb) acc[nb] = f32x4{0.f, 0.f, 0.f, 0.f};
  for (int kb = 0; kb < nk; ++kb) {
    bf16x8 a = ldsA(a_lds, lda, kb);
#pragma unroll
    for (int nb = 0; nb < 8; ++nb)
      acc[nb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, ldB(w, ldw, nb, kb), acc[nb], 0, 0, 0);
  }
}

// D-layout values (+bias, ReLU) -> bf16 LDS rows
__device__ __forceinline__ void storeRows(f32x4 acc[8], const float* bias, bool relu,
                                          __bf16* out, int ld) {
  const int l = __lane_id();
#pragma unroll
  for (int nb = 0; nb < 8; ++nb) {
    int col = 16 * nb + (l & 15);
    float b = bias ? bias[col] : 0.f;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      float v = acc[nb][i] + b;
      if (relu) v = fmaxf(v, 0.f);
      acc[nb][i] = v;
      out[(4 * (l >> 4) + i) * ld + col] = tobf(v);
    }
  }
}

// z for the 4 rows a lane owns: sum over columns of h2 * w3, reduced over
// the 16 lanes that share (l >> 4)
__device__ __forceinline__ void logits(const f32x4 h2[8], const float* w3, float b3, float z[4]) {
  const int l = __lane_id();
  float s[4] = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int nb = 0; nb < 8; ++nb) {
    float w = w3[16 * nb + (l & 15)];
#pragma unroll
    for (int i = 0; i < 4; ++i) s[i] += h2[nb][i] * w;
  }
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    float v = s[i];
    for (int o = 1; o < 16; o <<= 1) v += __shfl_xor(v, o, 64);
    z[i] = v + b3;
  }
}

// Forward of 16 samples whose feature rows are already in xs (16 x 32).
// Leaves h1 rows in LDS and returns h2 (D layout) and the logits.
__device__ __forceinline__ void forward16(const MlpDev& M, __bf16* xs, __bf16* h1s, f32x4 h2[8],
                                          f32x4 h1acc[8], float z[4]) {
  gemmRows(xs, kInPad, M.w1, kInPad, 1, h1acc);
  storeRows(h1acc, M.p + oB1, true, h1s, kHid);
  gemmRows(h1s, kHid, M.w2, kHid, kHid / 32, h2);
  const int l = __lane_id();
#pragma unroll
  for (int nb = 0; nb < 8; ++nb) {
    float b = M.p[oB2 + 16 * nb + (l & 15)];
#pragma unroll
    for (int i = 0; i < 4; ++i) h2[nb][i] = fmaxf(h2[nb][i] + b, 0.f);
  }
  logits(h2, M.p + oW3, M.p[oB3], z);
}

// ---------------------------------------------------------------------------
// Inference over the radiance queue of a pass (learned form).
constexpr int kMlpBlock = 256;  // 4 waves, 16 entries per wave
constexpr int kLdsPerWave = 16 * kInPad + 16 * kHid;  // bf16 elements

__device__ __forceinline__ uint32_t orderedF(float f) {
  uint32_t u = __float_as_uint(f);
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}

// Inference over the radiance rays of a pass (learned form).  A ray with
// two or more candidate domains asks the network about its nearest
// candidate and, when that one is predicted to miss and a third candidate
// exists, about the second; the first candidate predicted to hold the hit
// (logit > kHitLogit) bounds phase A: the ray's entries entering no later
// than it.  A ray with no such prediction keeps every entry in phase A.
//
// kCount appends one ask record per asking ray, so the kernel streams a
// dense list: every MFMA column is a real ask, and each wave keeps
// kAskStages tiles of 16 records in flight (independent, coalesced loads).
// The domain embeddings sit in LDS as bf16 layer-1 slots; second asks wait
// in a per-wave LDS ring until 16 of them fill a tile.
//
// The weights stay in registers for the whole persistent launch, laid out
// as v_mfma_f32_16x16x32_bf16 A operands (hidden units as rows); the asks
// are the B operand's 16 columns.  Layer 1's accumulators hold hidden
// units 16 mt + 4 (l / 16) + i of column l % 16 in lane l, and layer 2 takes
// them as its B operand in place: its K-block kb reads slot j of lane group
// g from unit 16 (2 kb + j / 4) + 4 g + j % 4, and the layer-2 weight
// fragments are packed in that unit order (kPackFrags).  No LDS traffic
// per MFMA.
// A deferral costs little when the prediction is wrong (the deferred entries
// still run, later in the same launch), so the bound is taken at p > 0.27: on C3 a
// sweep of the threshold from logit -2 to 2 gave 68.0 M to 76.2 M radiance
// visits and the fewest traversal milliseconds at -1.
constexpr float kHitLogit = -1.0f;
constexpr int kInferWaves = 4;  // waves per block
constexpr int kAskStages = 4;   // ask tiles in flight per wave
constexpr int kEmbLdsMax = 1024;  // domains whose embedding slots fit the LDS copy (48 B each)

__device__ __forceinline__ int l2Unit(int kb, int g, int j) {
  return 16 * (2 * kb + (j >> 2)) + 4 * g + (j & 3);
}

// w1f[mt][lane][j] = W1[16 mt + lane % 16][8 (lane / 16) + j] (features
// padded to 32 with zeros); w2f[mt][kb][lane][j] = W2[16 mt + lane % 16]
// [l2Unit(kb, lane / 16, j)].
__global__ void kPackFrags(const float* p, __bf16* w1f, __bf16* w2f) {
  for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < 8 * 4 * 64 * 8; t += gridDim.x * blockDim.x) {
    const int j = t & 7, lane = (t >> 3) & 63;
    if (t < 8 * 64 * 8) {
      const int mt = t >> 9, o = 16 * mt + (lane & 15), k = 8 * (lane >> 4) + j;
      w1f[t] = tobf(k < kIn ? p[oW1 + o * kIn + k] : 0.f);
    }
    const int kb = (t >> 9) & 3, mt = t >> 11, o = 16 * mt + (lane & 15);
    w2f[t] = tobf(p[oW2 + o * kHid + l2Unit(kb, lane >> 4, j)]);
  }
}

// Embedding part of the layer-1 input slots of domain d for lane groups
// g = 0..2 (group 3 holds only padding): slots 6, 7 of group 0 are e0, e1;
// group 1 holds e2..e9; group 2 e10..e15 and two zeros.
__device__ __forceinline__ bf16x8 embSlotsOf(const float* e, int g) {
  bf16x8 b;
#pragma unroll
  for (int j = 0; j < 8; ++j) b[j] = tobf(0.f);
  if (g == 0) {
    b[6] = tobf(e[0]);
    b[7] = tobf(e[1]);
  } else if (g == 1) {
#pragma unroll
    for (int j = 0; j < 8; ++j) b[j] = tobf(e[2 + j]);
  } else if (g == 2) {
#pragma unroll
    for (int j = 0; j < 6; ++j) b[j] = tobf(e[10 + j]);
  }
  return b;
}

// The LDS table of embedding slots, [domain][group 0..2].
__device__ __forceinline__ void fillEmbLds(const MlpDev& M, int nd, bf16x8* embl) {
  for (int i = threadIdx.x; i < nd * 3; i += blockDim.x) {
    const int d = i / 3, g = i - 3 * d;
    embl[i] = embSlotsOf(M.p + oEmb + (size_t)d * kEmb, g);
  }
}

// Layer-1 input slots 8 g .. 8 g + 7 of one (ray, domain) sample in bf16:
// the normalised origin and the direction (slots 0..5), the domain
// embedding (6..21), zeros (22..31).  The embedding comes from the LDS
// table when there is one, else from the fp32 master weights.
__device__ __forceinline__ bf16x8 askSlots(const MlpDev& M, const bf16x8* embl, int g, float ox,
                                           float oy, float oz, float dx, float dy, float dz, int dom) {
  bf16x8 b;
  if (g == 3) {
#pragma unroll
    for (int j = 0; j < 8; ++j) b[j] = tobf(0.f);
    return b;
  }
  b = embl ? embl[dom * 3 + g] : embSlotsOf(M.p + oEmb + (size_t)dom * kEmb, g);
  if (g == 0) {
    b[0] = tobf((ox - M.center[0]) * M.inv_half[0]);
    b[1] = tobf((oy - M.center[1]) * M.inv_half[1]);
    b[2] = tobf((oz - M.center[2]) * M.inv_half[2]);
    b[3] = tobf(dx);
    b[4] = tobf(dy);
    b[5] = tobf(dz);
  }
  return b;
}

// The network on 16 samples (the B operand's columns): returns the logit of
// column l % 16 in every lane.
__device__ __forceinline__ float inferColumns(const bf16x8 (&w1f)[8], const bf16x8 (&w2f)[8][4],
                                              const float* b1s, const float* b2s, const float* w3s,
                                              float b3, bf16x8 xb) {
  const int g = __lane_id() >> 4;
  f32x4 acc[8];
#pragma unroll
  for (int mt = 0; mt < 8; ++mt)
    acc[mt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(w1f[mt], xb, f32x4{0.f, 0.f, 0.f, 0.f}, 0, 0, 0);
  bf16x8 hb[4];
#pragma unroll
  for (int mt = 0; mt < 8; ++mt) {
    const float4 bb = *(const float4*)(b1s + 16 * mt + 4 * g);
    const float bv[4] = {bb.x, bb.y, bb.z, bb.w};
#pragma unroll
    for (int i = 0; i < 4; ++i) hb[mt >> 1][4 * (mt & 1) + i] = tobf(fmaxf(acc[mt][i] + bv[i], 0.f));
  }
#pragma unroll
  for (int mt = 0; mt < 8; ++mt) {
    f32x4 a = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int kb = 0; kb < 4; ++kb) a = __builtin_amdgcn_mfma_f32_16x16x32_bf16(w2f[mt][kb], hb[kb], a, 0, 0, 0);
    acc[mt] = a;
  }
  float z = 0.f;
#pragma unroll
  for (int mt = 0; mt < 8; ++mt) {
    const float4 bb = *(const float4*)(b2s + 16 * mt + 4 * g);
    const float4 ww = *(const float4*)(w3s + 16 * mt + 4 * g);
    z += fmaxf(acc[mt][0] + bb.x, 0.f) * ww.x + fmaxf(acc[mt][1] + bb.y, 0.f) * ww.y +
         fmaxf(acc[mt][2] + bb.z, 0.f) * ww.z + fmaxf(acc[mt][3] + bb.w, 0.f) * ww.w;
  }
  z += __shfl_xor(z, 16, 64);
  z += __shfl_xor(z, 32, 64);
  return z + b3;
}

__device__ __forceinline__ uint32_t boundKey(float t) { return orderedF(t); }

__global__ void __launch_bounds__(64 * kInferWaves, 1) kPredictRays(TraceParams P, MlpDev M,
                                                                   int emb_lds) {
  extern __shared__ bf16x8 embs[];  // [nd][3] embedding slots (when emb_lds)
  __shared__ __attribute__((aligned(16))) float b1s[kHid], b2s[kHid], w3s[kHid];
  __shared__ float4 ring[kInferWaves][32 * 3];  // second asks of each wave
  const DevCounters* c = P.ctr;
  if (!c->active) return;
  for (int i = threadIdx.x; i < kHid; i += blockDim.x) {
    b1s[i] = M.p[oB1 + i];
    b2s[i] = M.p[oB2 + i];
    w3s[i] = M.p[oW3 + i];
  }
  const bf16x8* embl = emb_lds ? embs : nullptr;
  if (emb_lds) fillEmbLds(M, P.ndomains, embs);
  const float b3 = M.p[oB3];
  const int l = __lane_id(), g = l >> 4, col = l & 15, w = threadIdx.x >> 6;
  bf16x8 w1f[8], w2f[8][4];
#pragma unroll
  for (int mt = 0; mt < 8; ++mt) {
    w1f[mt] = *(const bf16x8*)(M.w1f + ((size_t)mt * 64 + l) * 8);
#pragma unroll
    for (int kb = 0; kb < 4; ++kb) w2f[mt][kb] = *(const bf16x8*)(M.w2f + (((size_t)mt * 4 + kb) * 64 + l) * 8);
  }
  __syncthreads();
  const uint32_t reg = P.ask_region;
  if (reg == 0u) return;
  const uint32_t nwv = gridDim.x * kInferWaves;
  const uint32_t first = blockIdx.x * kInferWaves + (uint32_t)w;
  float4* q = ring[w];
  uint32_t qh = 0, qn = 0;  // the ring's head and fill (wave-uniform)
  // A tile of second asks from the ring: column col takes slot qh + col;
  // columns at or beyond cnt are padding.
  auto secondTile = [&](uint32_t cnt) {
    waveSync();
    const bool valid = (uint32_t)col < cnt;
    const uint32_t sl = ((qh + (uint32_t)col) & 31u) * 3u;
    const float4 b0 = q[sl], b1 = q[sl + 1], b2 = q[sl + 2];
    waveSync();
    const bf16x8 xb = askSlots(M, embl, g, b0.x, b0.y, b0.z, b0.w, b1.x, b1.y,
                               valid ? __float_as_int(b2.y) : 0);
    const float z1 = inferColumns(w1f, w2f, b1s, b2s, w3s, b3, xb);
    if (g == 0 && valid) P.rkey[__float_as_uint(b2.x)] = z1 > kHitLogit ? boundKey(b1.z) : kNoBound;
  };
  // one kCount block region per wave at a time
  for (uint32_t rb = first; rb < P.ask_regions; rb += nwv) {
    const uint32_t n = min(P.ask_cnt[rb], reg);
    if (n == 0u) continue;
    const float4* recs = P.ask + (size_t)rb * reg * 3;
    // unconditional loads (columns past the end read the last record and are
    // masked), so that no load result is merged with an older value and the
    // loads stay in flight until their tile is consumed
    float4 A0[kAskStages], A1[kAskStages], A2[kAskStages];
    auto fetch = [&](int k, uint32_t tile) {
      const float4* rec = recs + (size_t)min(tile * 16u + (uint32_t)col, n - 1u) * 3;
      A0[k] = rec[0];
      A1[k] = rec[1];
      A2[k] = rec[2];
    };
#pragma unroll
    for (int k = 0; k < kAskStages; ++k) fetch(k, (uint32_t)k);
    for (uint32_t t0 = 0; t0 * 16u < n; t0 += kAskStages) {
#pragma unroll
      for (int k = 0; k < kAskStages; ++k) {
        const uint32_t tile = t0 + (uint32_t)k;
        if (tile * 16u >= n) break;  // wave-uniform
        const float4 a0 = A0[k], a1 = A1[k], a2 = A2[k];
        fetch(k, tile + kAskStages);  // the slot's next tile
        const bool valid = tile * 16u + (uint32_t)col < n;
        const bf16x8 xb = askSlots(M, embl, g, a0.x, a0.y, a0.z, a0.w, a1.x, a1.y,
                                   valid ? __float_as_int(a2.y) : 0);
        const float z0 = inferColumns(w1f, w2f, b1s, b2s, w3s, b3, xb);
        const bool hit0 = valid && z0 > kHitLogit;
        const bool need2 = valid && !hit0 && __float_as_int(a2.w) >= 3;
        if (g == 0 && valid && !need2) P.rkey[__float_as_uint(a2.x)] = hit0 ? boundKey(a1.z) : kNoBound;
        // second asks: the ray with its second candidate's entry t and domain
        const unsigned long long m = __ballot(g == 0 && need2);
        if (m != 0ull) {
          if (g == 0 && need2) {
            const uint32_t sl =
                ((qh + qn + (uint32_t)__popcll(m & ((1ull << l) - 1ull))) & 31u) * 3u;
            q[sl] = a0;
            q[sl + 1] = make_float4(a1.x, a1.y, a1.w, 0.f);
            q[sl + 2] = make_float4(a2.x, a2.z, 0.f, 0.f);
          }
          qn += (uint32_t)__popcll(m);
          if (qn >= 16u) {  // at most 31 wait: 15 before this tile's 16
            secondTile(16u);
            qh = (qh + 16u) & 31u;
            qn -= 16u;
          }
        }
      }
    }
  }
  if (qn > 0u) secondTile(qn);
}

// Splits the radiance queue into the phase-A list and the deferred list,
// in queue order (so that phase A keeps the schedule's domain order).
// Entries of rays without a bound (kNoBound: one candidate domain, or no
// candidate predicted to hold the hit) stay in phase A.  Each of
// kAssignWaves waves owns a contiguous span of the queue (a multiple of 64
// entries): kAssignCount writes one deferral bit per entry and the span's
// two counts, kAssignScan turns the counts into list bases, kAssignWrite
// places the entries.  No atomics.
constexpr int kAssignUnroll = 4;

__device__ __forceinline__ uint32_t assignSpan(uint32_t n) {
  return ((n + kAssignWaves - 1) / kAssignWaves + 63u) & ~63u;
}

__global__ void __launch_bounds__(256) kAssignCount(TraceParams P) {
  const DevCounters* c = P.ctr;
  if (!c->active) return;
  const uint32_t n = c->q_total[0], span = assignSpan(n);
  const uint32_t wv = blockIdx.x * 4 + (threadIdx.x >> 6);
  const int l = __lane_id();
  const uint32_t beg = min(wv * span, n), end = min(beg + span, n);
  uint32_t ndef = 0;
  for (uint32_t q0 = beg; q0 < end; q0 += 64 * kAssignUnroll) {  // wave-uniform
    bool d[kAssignUnroll];
#pragma unroll
    for (int k = 0; k < kAssignUnroll; ++k) {
      const uint32_t q = q0 + 64u * k + (uint32_t)l;
      d[k] = false;
      if (q < end) {
        const QEntry e = P.q[0][q];
        d[k] = orderedF(e.tdom) > P.rkey[e.ray];  // beyond the chosen domain
      }
    }
#pragma unroll
    for (int k = 0; k < kAssignUnroll; ++k) {
      const unsigned long long m = __ballot(d[k]);
      const uint32_t w0 = q0 + 64u * k;
      if (w0 < end && l == 0) P.dflag[w0 >> 6] = m;
      ndef += (uint32_t)__popcll(m);
    }
  }
  if (l == 0) {
    P.acnt[wv] = (end - beg) - ndef;  // phase A
    P.acnt[kAssignWaves + wv] = ndef;  // deferred
  }
}

// One block: exclusive scans of both count arrays (list bases in P.acnt's
// place), the list lengths and the deferred-entry statistic.
__global__ void __launch_bounds__(1024) kAssignScan(TraceParams P) {
  DevCounters* c = P.ctr;
  if (!c->active) return;
  constexpr int kPer = kAssignWaves / 1024;
  __shared__ uint32_t wsum[16];
  const int t = threadIdx.x, l = __lane_id(), w = t >> 6;
  for (int list = 0; list < 2; ++list) {
    uint32_t* a = P.acnt + list * kAssignWaves;
    uint32_t v[kPer], sum = 0;
#pragma unroll
    for (int k = 0; k < kPer; ++k) {
      v[k] = a[t * kPer + k];
      sum += v[k];
    }
    uint32_t x = sum;  // inclusive wave scan
    for (int o = 1; o < 64; o <<= 1) {
      const uint32_t y = __shfl_up(x, o, 64);
      if (l >= o) x += y;
    }
    if (l == 63) wsum[w] = x;
    __syncthreads();
    uint32_t pre = x - sum, total = 0;
    for (int j = 0; j < 16; ++j) {
      pre += j < w ? wsum[j] : 0u;
      total += wsum[j];
    }
#pragma unroll
    for (int k = 0; k < kPer; ++k) {
      a[t * kPer + k] = pre;
      pre += v[k];
    }
    if (t == 0) {
      c->n_phase[list] = total;
      if (list == 1) c->deferred += total;
    }
    __syncthreads();
  }
}

__global__ void __launch_bounds__(256) kAssignWrite(TraceParams P) {
  DevCounters* c = P.ctr;
  if (!c->active) return;
  const uint32_t n = c->q_total[0], span = assignSpan(n);
  const uint32_t wv = blockIdx.x * 4 + (threadIdx.x >> 6);
  const int l = __lane_id();
  const unsigned long long below = (1ull << l) - 1ull;
  const uint32_t beg = min(wv * span, n), end = min(beg + span, n);
  uint32_t ba = P.acnt[wv], bb = P.acnt[kAssignWaves + wv];  // wave-uniform
  bool over = false;
  for (uint32_t q0 = beg; q0 < end; q0 += 64 * kAssignUnroll) {
    unsigned long long m[kAssignUnroll];
#pragma unroll
    for (int k = 0; k < kAssignUnroll; ++k) {
      const uint32_t w0 = q0 + 64u * k;
      m[k] = w0 < end ? P.dflag[w0 >> 6] : 0ull;
    }
#pragma unroll
    for (int k = 0; k < kAssignUnroll; ++k) {
      const uint32_t w0 = q0 + 64u * k;
      if (w0 >= end) break;  // wave-uniform
      const uint32_t cnt = min(end - w0, 64u);
      const unsigned long long valid = cnt == 64u ? ~0ull : ((1ull << cnt) - 1ull);
      const unsigned long long da = valid & ~m[k], db = m[k];
      const uint32_t q = w0 + (uint32_t)l;
      if ((valid >> l) & 1ull) {
        const bool d = (db >> l) & 1ull;
        const uint32_t idx = d ? bb + (uint32_t)__popcll(db & below) : ba + (uint32_t)__popcll(da & below);
        if (idx < P.q_cap[0])
          (d ? P.qb : P.qa)[idx] = q;
        else
          over = true;
      }
      ba += (uint32_t)__popcll(da);
      bb += (uint32_t)__popcll(db);
    }
  }
  if (over) atomicOr(&c->err, 1u);
}

// ---------------------------------------------------------------------------
// Training: forward + backward of 16 samples per wave; activations for the
// weight gradients are written transposed ([feature][sample], bf16) so that
// the weight-gradient GEMMs read contiguous K = batch.
struct TrainBufs {
  const float* samples;      // 8 floats per sample (org, dir, dom, label)
  const uint32_t* index;     // batch sample ids
  int batch;
  float inv_batch;
  __bf16 *xt, *h1t, *dh2t, *dh1t;  // [32|128][batch]
  float* grad;               // flat fp32, same layout as the parameters
  float* loss;               // sum of BCE over the batch (and correct count)
};

__global__ void __launch_bounds__(kMlpBlock) kTrainForwardBackward(MlpDev M, TrainBufs B) {
  __shared__ __attribute__((aligned(16))) __bf16 lds[4 * (16 * kInPad + 3 * 16 * kHid)];
  const int w = threadIdx.x >> 6, l = __lane_id();
  __bf16* xs = lds + w * (16 * kInPad + 3 * 16 * kHid);
  __bf16* h1s = xs + 16 * kInPad;
  __bf16* dh2s = h1s + 16 * kHid;
  __bf16* dh1s = dh2s + 16 * kHid;
  const int s0 = (blockIdx.x * 4 + w) * 16;
  if (s0 >= B.batch) return;  // batch is a multiple of 16
  // features
  float label_row = 0.f;
  int dom_row = 0;
  if (l < 16) {
    uint32_t sid = B.index[s0 + l];
    const float* sm = B.samples + (size_t)sid * 8;
    float o[3] = {sm[0], sm[1], sm[2]}, d[3] = {sm[3], sm[4], sm[5]};
    dom_row = (int)sm[6];
    label_row = sm[7];
    float x[kInPad];
    features(M, o, d, dom_row, x);
#pragma unroll
    for (int k = 0; k < kInPad; ++k) {
      xs[l * kInPad + k] = tobf(x[k]);
      B.xt[(size_t)k * B.batch + s0 + l] = tobf(x[k]);
    }
  }
  waveSync();
  f32x4 h1acc[8], h2[8];
  float z[4];
  forward16(M, xs, h1s, h2, h1acc, z);
  // labels of my 4 rows
  float y[4];
  int rbase = 4 * (l >> 4);
#pragma unroll
  for (int i = 0; i < 4; ++i) y[i] = __shfl(label_row, rbase + i, 64);
  // BCE with logits, dz = (sigmoid(z) - y) / batch
  float dz[4];
  float lsum = 0.f, corr = 0.f;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    float p = 1.f / (1.f + __expf(-z[i]));
    dz[i] = (p - y[i]) * B.inv_batch;
    float bce = fmaxf(z[i], 0.f) - z[i] * y[i] + __logf(1.f + __expf(-fabsf(z[i])));
    if ((l & 15) == 0) {
      lsum += bce;
      corr += ((z[i] > 0.f) == (y[i] > 0.5f)) ? 1.f : 0.f;
    }
  }
  lsum = waveSumF(lsum);
  corr = waveSumF(corr);
  if (l == 0) {
    atomicAdd(&B.loss[0], lsum);
    atomicAdd(&B.loss[1], corr);
  }
  // layer 3 and layer-2 output gradients
  const float* w3 = M.p + oW3;
  float db3 = 0.f;
#pragma unroll
  for (int i = 0; i < 4; ++i) db3 += ((l & 15) == 0) ? dz[i] : 0.f;
  db3 = waveSumF(db3);
  if (l == 0) atomicAdd(&B.grad[oB3], db3);
#pragma unroll
  for (int nb = 0; nb < 8; ++nb) {
    int col = 16 * nb + (l & 15);
    float wv = w3[col];
    float gw3 = 0.f, gb2 = 0.f;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      float h = h2[nb][i];
      gw3 += dz[i] * h;
      float g = h > 0.f ? dz[i] * wv : 0.f;
      gb2 += g;
      __bf16 gb = tobf(g);
      dh2s[(rbase + i) * kHid + col] = gb;
      B.dh2t[(size_t)col * B.batch + s0 + rbase + i] = gb;
    }
    gw3 += __shfl_xor(gw3, 16, 64);
    gw3 += __shfl_xor(gw3, 32, 64);
    gb2 += __shfl_xor(gb2, 16, 64);
    gb2 += __shfl_xor(gb2, 32, 64);
    if (l < 16) {
      atomicAdd(&B.grad[oW3 + col], gw3);
      atomicAdd(&B.grad[oB2 + col], gb2);
    }
  }
  // h1 transposed for dW2
#pragma unroll
  for (int nb = 0; nb < 8; ++nb)
#pragma unroll
    for (int i = 0; i < 4; ++i)
      B.h1t[(size_t)(16 * nb + (l & 15)) * B.batch + s0 + rbase + i] = tobf(h1acc[nb][i]);
  waveSync();
  // dh1 = (dh2 * W2) * relu'(h1): B operand W2^T rows (in, out)
  f32x4 g1[8];
  gemmRows(dh2s, kHid, M.w2t, kHid, kHid / 32, g1);
#pragma unroll
  for (int nb = 0; nb < 8; ++nb) {
    int col = 16 * nb + (l & 15);
    float gb1 = 0.f;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      float g = h1acc[nb][i] > 0.f ? g1[nb][i] : 0.f;
      gb1 += g;
      __bf16 gb = tobf(g);
      dh1s[(rbase + i) * kHid + col] = gb;
      B.dh1t[(size_t)col * B.batch + s0 + rbase + i] = gb;
    }
    gb1 += __shfl_xor(gb1, 16, 64);
    gb1 += __shfl_xor(gb1, 32, 64);
    if (l < 16) atomicAdd(&B.grad[oB1 + col], gb1);
  }
  waveSync();
  // embedding gradient: demb = dh1 (16 x 128) * W1[:, 6:22] -> 16 x 16, one
  // column block; lanes hold rows 4 * (l >> 4) + i, dimension l & 15
  f32x4 ge = f32x4{0.f, 0.f, 0.f, 0.f};
  for (int kb = 0; kb < kHid / 32; ++kb)
    ge = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ldsA(dh1s, kHid, kb), ldB(M.w1et, kHid, 0, kb),
                                                 ge, 0, 0, 0);
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    int dom = __shfl(dom_row, rbase + i, 64);
    atomicAdd(&B.grad[oEmb + (size_t)dom * kEmb + (l & 15)], ge[i]);
  }
}

// Weight gradients as GEMMs over the batch: G[o][i] = sum_s A^T[o][s] B^T[i][s]
// (A^T = dh2t or dh1t, B^T = h1t or xt).  One wave per 16 x 16 block.
__global__ void __launch_bounds__(64) kTrainWeightGrad(TrainBufs B, int which) {
  // which 0: dW2 (128 x 128) = dh2t . h1t^T ; which 1: dW1 (128 x 22) = dh1t . xt^T
  const int l = __lane_id();
  const int ncols = which == 0 ? kHid : kInPad;
  const int nbn = ncols / 16;
  const int mb = blockIdx.x / nbn, nbk = blockIdx.x % nbn;
  const __bf16* at = which == 0 ? B.dh2t : B.dh1t;
  const __bf16* bt = which == 0 ? B.h1t : B.xt;
  f32x4 acc = f32x4{0.f, 0.f, 0.f, 0.f};
  for (int kb = 0; kb < B.batch / 32; ++kb) {
    bf16x8 a = *(const bf16x8*)(at + (size_t)(16 * mb + (l & 15)) * B.batch + kb * 32 + 8 * (l >> 4));
    bf16x8 b = *(const bf16x8*)(bt + (size_t)(16 * nbk + (l & 15)) * B.batch + kb * 32 + 8 * (l >> 4));
    acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc, 0, 0, 0);
  }
  int col = 16 * nbk + (l & 15);
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    int row = 16 * mb + 4 * (l >> 4) + i;
    if (which == 0)
      B.grad[oW2 + row * kHid + col] = acc[i];
    else if (col < kIn)
      B.grad[oW1 + row * kIn + col] = acc[i];
  }
}

// Adam (Kingma & Ba 2015) on the flat fp32 parameters, then bf16 packing.
__global__ void kAdam(float* p, float* m, float* v, const float* g, int n, float lr, float b1,
                      float b2, float eps, float c1, float c2) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    float gi = g[i];
    float mi = b1 * m[i] + (1.f - b1) * gi;
    float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
  }
}

__global__ void kPackWeights(const float* p, __bf16* w1, __bf16* w2, __bf16* w2t, __bf16* w1et) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < kHid * kHid; i += gridDim.x * blockDim.x) {
    int o = i / kHid, k = i % kHid;
    w2[i] = tobf(p[oW2 + i]);
    w2t[k * kHid + o] = tobf(p[oW2 + i]);
    if (k < kInPad) w1[o * kInPad + k] = tobf(k < kIn ? p[oW1 + o * kIn + k] : 0.f);
    if (k < kEmb) w1et[k * kHid + o] = tobf(p[oW1 + o * kIn + kRayFeat + k]);
  }
}

__global__ void kFill(float* a, int n, float v) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) a[i] = v;
}

// Forward-only logits for a list of samples (tests and evaluation).
__global__ void __launch_bounds__(kMlpBlock) kLogits(MlpDev M, const float* samples,
                                                     const uint32_t* index, int n, float* out) {
  __shared__ __attribute__((aligned(16))) __bf16 lds[4 * kLdsPerWave];
  const int w = threadIdx.x >> 6, l = __lane_id();
  __bf16* xs = lds + w * kLdsPerWave;
  __bf16* h1s = xs + 16 * kInPad;
  int s0 = (blockIdx.x * 4 + w) * 16;
  if (s0 >= n) return;
  if (l < 16) {
    int s = s0 + l;
    float x[kInPad];
    if (s < n) {
      const float* sm = samples + (size_t)(index ? index[s] : (uint32_t)s) * 8;
      float o[3] = {sm[0], sm[1], sm[2]}, d[3] = {sm[3], sm[4], sm[5]};
      features(M, o, d, (int)sm[6], x);
    } else {
      for (int k = 0; k < kInPad; ++k) x[k] = 0.f;
    }
    for (int k = 0; k < kInPad; ++k) xs[l * kInPad + k] = tobf(x[k]);
  }
  waveSync();
  f32x4 h1acc[8], h2[8];
  float z[4];
  forward16(M, xs, h1s, h2, h1acc, z);
  if ((l & 15) == 0) {
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      int s = s0 + 4 * (l >> 4) + i;
      if (s < n) out[s] = z[i];
    }
  }
}

// ---------------------------------------------------------------------------
// Host side
struct Predictor {
  int nd = 0;
  int nparams = 0;
  DevBuf p, m, v, g, w1, w2, w2t, w1et, w1f, w2f, loss;
  DevBuf xt, h1t, dh2t, dh1t, index;
  int step = 0;
  bool trained = false;
  float center[3], inv_half[3];
  int tbatch = 0;
  int pred_grid = 0;  // persistent inference grid (blocks)
  int emb_lds = 0;    // inference keeps the embedding slots in LDS
  MlpDev dev() const {
    MlpDev M;
    M.w1f = w1f.as<__bf16>();
    M.w2f = w2f.as<__bf16>();
    M.w1 = w1.as<__bf16>();
    M.w2 = w2.as<__bf16>();
    M.w2t = w2t.as<__bf16>();
    M.w1et = w1et.as<__bf16>();
    M.p = p.as<float>();
    for (int k = 0; k < 3; ++k) {
      M.center[k] = center[k];
      M.inv_half[k] = inv_half[k];
    }
    return M;
  }
};

void predictorRelease(Ctx* c) {
  if (!c->pred) return;
  Predictor* R = c->pred;
  for (DevBuf* b : {&R->p, &R->m, &R->v, &R->g, &R->w1, &R->w2, &R->w2t, &R->w1et, &R->w1f, &R->w2f, &R->loss,
                    &R->xt, &R->h1t, &R->dh2t, &R->dh1t, &R->index})
    b->release();
  delete R;
  c->pred = nullptr;
}

static void packWeights(Ctx* c) {
  Predictor* R = c->pred;
  kPackWeights<<<64, 256, 0, c->stream>>>(R->p.as<float>(), R->w1.as<__bf16>(), R->w2.as<__bf16>(),
                                          R->w2t.as<__bf16>(), R->w1et.as<__bf16>());
  SPRAY_HIP(hipGetLastError());
  kPackFrags<<<64, 256, 0, c->stream>>>(R->p.as<float>(), R->w1f.as<__bf16>(), R->w2f.as<__bf16>());
  SPRAY_HIP(hipGetLastError());
}

// Creates the network (He-uniform init from a fixed seed) for the loaded scene.
static Predictor* ensurePredictor(Ctx* c) {
  int nd = c->pool.ndomains;
  if (c->pred && c->pred->nd == nd) return c->pred;
  predictorRelease(c);
  Predictor* R = new Predictor();
  c->pred = R;
  R->nd = nd;
  R->nparams = kDenseParams + nd * kEmb;
  for (int k = 0; k < 3; ++k) {
    float lo = c->scene.lo[k], hi = c->scene.hi[k];
    R->center[k] = 0.5f * (lo + hi);
    float h = 0.5f * (hi - lo);
    R->inv_half[k] = h > 0.f ? 1.0f / h : 1.0f;
  }
  std::vector<float> h(R->nparams, 0.f);
  uint64_t s = 0x5eed1234abcdull;
  auto uni = [&]() {
    s = s * 6364136223846793005ull + 1442695040888963407ull;
    return (double)(s >> 11) * (1.0 / 9007199254740992.0);
  };
  auto fill = [&](int off, int n, double bound) {
    for (int i = 0; i < n; ++i) h[off + i] = (float)((2.0 * uni() - 1.0) * bound);
  };
  fill(oW1, kHid * kIn, std::sqrt(6.0 / kIn));
  fill(oW2, kHid * kHid, std::sqrt(6.0 / kHid));
  fill(oW3, kHid, std::sqrt(6.0 / kHid));
  fill(oEmb, nd * kEmb, 1.0);
  R->p.reserve((size_t)R->nparams * 4);
  R->m.reserve((size_t)R->nparams * 4);
  R->v.reserve((size_t)R->nparams * 4);
  R->g.reserve((size_t)R->nparams * 4);
  R->w1.reserve(kHid * kInPad * 2);
  R->w2.reserve(kHid * kHid * 2);
  R->w2t.reserve(kHid * kHid * 2);
  R->w1et.reserve(kEmb * kHid * 2);
  R->w1f.reserve(8 * 64 * 8 * 2);
  R->w2f.reserve(8 * 4 * 64 * 8 * 2);
  R->loss.reserve(64);
  SPRAY_HIP(hipMemcpy(R->p.p, h.data(), (size_t)R->nparams * 4, hipMemcpyHostToDevice));
  SPRAY_HIP(hipMemset(R->m.p, 0, (size_t)R->nparams * 4));
  SPRAY_HIP(hipMemset(R->v.p, 0, (size_t)R->nparams * 4));
  packWeights(c);
  SPRAY_HIP(hipStreamSynchronize(c->stream));
  return R;
}

static void ensureTrainBufs(Predictor* R, int batch) {
  if (R->tbatch >= batch) return;
  R->xt.reserve((size_t)kInPad * batch * 2);
  R->h1t.reserve((size_t)kHid * batch * 2);
  R->dh2t.reserve((size_t)kHid * batch * 2);
  R->dh1t.reserve((size_t)kHid * batch * 2);
  R->index.reserve((size_t)batch * 4);
  R->tbatch = batch;
}

// One forward/backward over `batch` samples (ids in d_index); gradient in R->g.
static void gradStep(Ctx* c, const float* d_samples, const uint32_t* d_index, int batch) {
  Predictor* R = c->pred;
  hipStream_t s = c->stream;
  ensureTrainBufs(R, batch);
  kFill<<<64, 256, 0, s>>>(R->g.as<float>(), R->nparams, 0.f);
  kFill<<<1, 64, 0, s>>>(R->loss.as<float>(), 2, 0.f);
  TrainBufs B;
  B.samples = d_samples;
  B.index = d_index;
  B.batch = batch;
  B.inv_batch = 1.0f / (float)batch;
  B.xt = R->xt.as<__bf16>();
  B.h1t = R->h1t.as<__bf16>();
  B.dh2t = R->dh2t.as<__bf16>();
  B.dh1t = R->dh1t.as<__bf16>();
  B.grad = R->g.as<float>();
  B.loss = R->loss.as<float>();
  MlpDev M = R->dev();
  kTrainForwardBackward<<<(batch + 63) / 64, kMlpBlock, 0, s>>>(M, B);
  kTrainWeightGrad<<<(kHid / 16) * (kHid / 16), 64, 0, s>>>(B, 0);
  kTrainWeightGrad<<<(kHid / 16) * (kInPad / 16), 64, 0, s>>>(B, 1);
  SPRAY_HIP(hipGetLastError());
}

__global__ void kBatchIndex(uint32_t* idx, int batch, unsigned long long start,
                            unsigned long long stride, unsigned long long n) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < batch) idx[i] = (uint32_t)(((start + (unsigned long long)i) * stride) % n);
}

static size_t embLdsBytes(int nd) { return nd <= kEmbLdsMax ? (size_t)nd * 3 * sizeof(bf16x8) : 0; }

void predictorInfer(Ctx* c, const TraceParams& P, hipStream_t s) {
  Predictor* R = c->pred;  // grid set by predictorPrepare (no allocation here: graph capture)
  kPredictRays<<<R->pred_grid, 64 * kInferWaves, R->emb_lds ? embLdsBytes(P.ndomains) : 0, s>>>(
      P, R->dev(), R->emb_lds);
}

void predictorAssign(Ctx* c, const TraceParams& P) {
  hipStream_t s = c->stream;
  kAssignCount<<<kAssignBlocks, 256, 0, s>>>(P);
  kAssignScan<<<1, 1024, 0, s>>>(P);
  kAssignWrite<<<kAssignBlocks, 256, 0, s>>>(P);
}

bool predictorReady(Ctx* c) { return c->pred && c->pred->trained; }

void predictorPrepare(Ctx* c) {
  if (!c->pred) throw StatusError(SPRAY_ERR_STATE, "learned schedule needs a trained predictor");
  Predictor* R = c->pred;
  const int nd = c->pool.ndomains;
  R->emb_lds = nd <= kEmbLdsMax ? 1 : 0;
  const size_t lds = embLdsBytes(nd);
  hipDeviceProp_t prop;
  SPRAY_HIP(hipGetDeviceProperties(&prop, c->device));
  int per_cu = 0;
  SPRAY_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kPredictRays, 64 * kInferWaves, lds));
  R->pred_grid = prop.multiProcessorCount * std::max(1, per_cu);
}

}  // namespace spray

using namespace spray;

namespace {
template <typename F>
int guard(spray_ctx* ctx, F&& f) {
  try {
    f();
    ctx->c.err.clear();
    return SPRAY_OK;
  } catch (const StatusError& e) {
    ctx->c.err = e.msg;
    return e.code;
  } catch (const DeviceError& e) {
    ctx->c.err = e.msg;
    return SPRAY_ERR_DEVICE;
  } catch (const std::exception& e) {
    ctx->c.err = e.what();
    return SPRAY_ERR_INVALID;
  }
}
}  // namespace

extern "C" {

int spray_predictor_log(spray_ctx* ctx, int enable, long long max_samples) {
  if (!ctx) return SPRAY_ERR_INVALID;
  return guard(ctx, [&] {
    Ctx* c = &ctx->c;
    if (!c->loaded) throw StatusError(SPRAY_ERR_STATE, "no scene loaded");
    tracerSetLogging(c, enable != 0, max_samples);
  });
}

long long spray_predictor_num_samples(const spray_ctx* ctx) {
  if (!ctx) return 0;
  return tracerLogCount(&ctx->c);
}

int spray_predictor_train(spray_ctx* ctx, int epochs, int batch, float lr, float* loss_out) {
  if (!ctx) return SPRAY_ERR_INVALID;
  return guard(ctx, [&] {
    Ctx* c = &ctx->c;
    if (!c->loaded) throw StatusError(SPRAY_ERR_STATE, "no scene loaded");
    if (epochs <= 0 || batch < 64 || (batch % 64) || !(lr > 0.f))
      throw StatusError(SPRAY_ERR_INVALID, "epochs > 0, batch a multiple of 64, lr > 0");
    long long n = tracerLogCount(c);
    long long ntrain = n - n / 10;  // the last 10% are held out
    if (ntrain < batch) throw StatusError(SPRAY_ERR_STATE, "too few logged samples");
    Predictor* R = ensurePredictor(c);
    ensureTrainBufs(R, batch);
    const float* samples = tracerLogBuffer(c);
    hipStream_t s = c->stream;
    long long steps_per_epoch = ntrain / batch;
    // a fixed stride coprime with ntrain walks the samples in scattered order
    unsigned long long stride = 2654435761ull % (unsigned long long)ntrain;
    while (stride == 0 || std::gcd((unsigned long long)ntrain, stride) != 1ull) ++stride;
    float last = 0.f;
    for (int ep = 0; ep < epochs; ++ep) {
      double acc_loss = 0;
      for (long long st = 0; st < steps_per_epoch; ++st) {
        kBatchIndex<<<(batch + 255) / 256, 256, 0, s>>>(R->index.as<uint32_t>(), batch,
                                                         (unsigned long long)st * batch, stride,
                                                         (unsigned long long)ntrain);
        gradStep(c, samples, R->index.as<uint32_t>(), batch);
        R->step++;
        const float b1 = 0.9f, b2 = 0.999f;
        float c1 = 1.f - std::pow(b1, (float)R->step), c2 = 1.f - std::pow(b2, (float)R->step);
        kAdam<<<64, 256, 0, s>>>(R->p.as<float>(), R->m.as<float>(), R->v.as<float>(),
                                 R->g.as<float>(), R->nparams, lr, b1, b2, 1e-8f, c1, c2);
        packWeights(c);
        if (st == steps_per_epoch - 1 || (st % 64) == 0) {
          float h[2];
          SPRAY_HIP(hipMemcpyAsync(h, R->loss.p, 8, hipMemcpyDeviceToHost, s));
          SPRAY_HIP(hipStreamSynchronize(s));
          acc_loss = h[0] / batch;
        }
      }
      last = (float)acc_loss;
    }
    R->trained = true;
    if (loss_out) *loss_out = last;
  });
}

int spray_predictor_eval(spray_ctx* ctx, float* loss_out, float* accuracy_out) {
  if (!ctx) return SPRAY_ERR_INVALID;
  return guard(ctx, [&] {
    Ctx* c = &ctx->c;
    if (!c->pred) throw StatusError(SPRAY_ERR_STATE, "no predictor");
    long long n = tracerLogCount(c);
    long long nval = n / 10;
    if (nval < 16) throw StatusError(SPRAY_ERR_STATE, "too few held-out samples");
    long long start = n - nval;
    const float* samples = tracerLogBuffer(c);
    DevBuf out, idx;
    out.reserve((size_t)nval * 4);
    idx.reserve((size_t)nval * 4);
    std::vector<uint32_t> hidx(nval);
    for (long long i = 0; i < nval; ++i) hidx[i] = (uint32_t)(start + i);
    SPRAY_HIP(hipMemcpy(idx.p, hidx.data(), nval * 4, hipMemcpyHostToDevice));
    kLogits<<<(int)((nval + 63) / 64), kMlpBlock, 0, c->stream>>>(c->pred->dev(), samples,
                                                                   idx.as<uint32_t>(), (int)nval,
                                                                   out.as<float>());
    std::vector<float> z(nval), sm((size_t)nval * 8);
    SPRAY_HIP(hipMemcpy(z.data(), out.p, nval * 4, hipMemcpyDeviceToHost));
    SPRAY_HIP(hipMemcpy(sm.data(), samples + start * 8, (size_t)nval * 32, hipMemcpyDeviceToHost));
    double loss = 0, correct = 0;
    for (long long i = 0; i < nval; ++i) {
      double y = sm[i * 8 + 7], zi = z[i];
      loss += std::max(zi, 0.0) - zi * y + std::log1p(std::exp(-std::fabs(zi)));
      correct += ((zi > 0) == (y > 0.5)) ? 1 : 0;
    }
    if (loss_out) *loss_out = (float)(loss / nval);
    if (accuracy_out) *accuracy_out = (float)(correct / nval);
  });
}

int spray_predictor_save(spray_ctx* ctx, const char* path) {
  if (!ctx || !path) return SPRAY_ERR_INVALID;
  return guard(ctx, [&] {
    Predictor* R = ctx->c.pred;
    if (!R) throw StatusError(SPRAY_ERR_STATE, "no predictor");
    std::vector<float> h(R->nparams);
    SPRAY_HIP(hipMemcpy(h.data(), R->p.p, (size_t)R->nparams * 4, hipMemcpyDeviceToHost));
    FILE* f = std::fopen(path, "wb");
    if (!f) throw StatusError(SPRAY_ERR_IO, std::string("cannot write ") + path);
    int hdr[4] = {0x53505259, 1, R->nd, R->nparams};  // "SPRY", version 1
    std::fwrite(hdr, sizeof(hdr), 1, f);
    std::fwrite(h.data(), 4, h.size(), f);
    std::fclose(f);
  });
}

int spray_predictor_load(spray_ctx* ctx, const char* path) {
  if (!ctx || !path) return SPRAY_ERR_INVALID;
  return guard(ctx, [&] {
    Ctx* c = &ctx->c;
    if (!c->loaded) throw StatusError(SPRAY_ERR_STATE, "no scene loaded");
    FILE* f = std::fopen(path, "rb");
    if (!f) throw StatusError(SPRAY_ERR_IO, std::string("cannot read ") + path);
    int hdr[4];
    bool ok = std::fread(hdr, sizeof(hdr), 1, f) == 1 && hdr[0] == 0x53505259 && hdr[1] == 1;
    Predictor* R = ensurePredictor(c);
    ok = ok && hdr[2] == R->nd && hdr[3] == R->nparams;
    std::vector<float> h(R->nparams);
    ok = ok && std::fread(h.data(), 4, h.size(), f) == h.size();
    std::fclose(f);
    if (!ok) throw StatusError(SPRAY_ERR_FORMAT, "predictor file does not match this scene");
    SPRAY_HIP(hipMemcpy(R->p.p, h.data(), (size_t)R->nparams * 4, hipMemcpyHostToDevice));
    packWeights(c);
    SPRAY_HIP(hipStreamSynchronize(c->stream));
    R->trained = true;
  });
}

// Test hooks: one forward/backward on host-provided samples (8 floats each:
// origin, direction, domain, label) with host-provided parameters (flat fp32,
// layout W1[128][22], b1, W2[128][128], b2, w3, b3, emb[nd][16]); returns
// the summed BCE and the gradient of the mean BCE.
// Host samples name a domain of the loaded scene (their embedding row).
static void checkSampleDomains(const float* samples, int n, int nd) {
  for (int i = 0; i < n; ++i) {
    const float d = samples[(size_t)i * 8 + 6];
    if (!(d >= 0.f && d < (float)nd) || d != (float)(int)d)
      throw StatusError(SPRAY_ERR_INVALID, "sample " + std::to_string(i) + " names no domain of the scene");
  }
}

int spray_mlp_grad(spray_ctx* ctx, const float* samples, int n, const float* params,
                   float* loss_out, float* grad_out) {
  if (!ctx || !samples || !params || !grad_out) return SPRAY_ERR_INVALID;
  return guard(ctx, [&] {
    Ctx* c = &ctx->c;
    if (!c->loaded) throw StatusError(SPRAY_ERR_STATE, "no scene loaded");
    if (n < 32 || n % 32) throw StatusError(SPRAY_ERR_INVALID, "n must be a positive multiple of 32");
    checkSampleDomains(samples, n, c->pool.ndomains);
    Predictor* R = ensurePredictor(c);
    SPRAY_HIP(hipMemcpy(R->p.p, params, (size_t)R->nparams * 4, hipMemcpyHostToDevice));
    packWeights(c);
    DevBuf ds, di;
    ds.reserve((size_t)n * 32);
    di.reserve((size_t)n * 4);
    std::vector<uint32_t> idx(n);
    for (int i = 0; i < n; ++i) idx[i] = (uint32_t)i;
    SPRAY_HIP(hipMemcpy(ds.p, samples, (size_t)n * 32, hipMemcpyHostToDevice));
    SPRAY_HIP(hipMemcpy(di.p, idx.data(), (size_t)n * 4, hipMemcpyHostToDevice));
    gradStep(c, ds.as<float>(), di.as<uint32_t>(), n);
    SPRAY_HIP(hipStreamSynchronize(c->stream));
    SPRAY_HIP(hipMemcpy(grad_out, R->g.p, (size_t)R->nparams * 4, hipMemcpyDeviceToHost));
    float h[2];
    SPRAY_HIP(hipMemcpy(h, R->loss.p, 8, hipMemcpyDeviceToHost));
    if (loss_out) *loss_out = h[0];
  });
}

// Logits of host samples through the inference kernel's arithmetic (the
// register-resident fragments, the chained layers and the bf16 input slots
// of kPredictRays), so that the logits the learned schedule acts on are
// checked against the float64 oracle.  Even 16-sample tiles take the
// embedding slots from the LDS table, odd tiles from the fp32 master (the
// path of scenes with more than kEmbLdsMax domains), so both are checked.
__global__ void __launch_bounds__(64) kLogitsInfer(MlpDev M, int nd, int emb_lds, const float* samples,
                                                   int n, float* out) {
  extern __shared__ bf16x8 embs[];
  __shared__ __attribute__((aligned(16))) float b1s[kHid], b2s[kHid], w3s[kHid];
  for (int i = threadIdx.x; i < kHid; i += blockDim.x) {
    b1s[i] = M.p[oB1 + i];
    b2s[i] = M.p[oB2 + i];
    w3s[i] = M.p[oW3 + i];
  }
  if (emb_lds) fillEmbLds(M, nd, embs);
  __syncthreads();
  const int l = __lane_id(), g = l >> 4;
  bf16x8 w1f[8], w2f[8][4];
#pragma unroll
  for (int mt = 0; mt < 8; ++mt) {
    w1f[mt] = *(const bf16x8*)(M.w1f + ((size_t)mt * 64 + l) * 8);
#pragma unroll
    for (int kb = 0; kb < 4; ++kb) w2f[mt][kb] = *(const bf16x8*)(M.w2f + (((size_t)mt * 4 + kb) * 64 + l) * 8);
  }
  for (int s0 = blockIdx.x * 16; s0 < n; s0 += gridDim.x * 16) {
    const int sidx = s0 + (l & 15);
    float x[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    int dom = 0;
    if (sidx < n) {
      const float* sm = samples + (size_t)sidx * 8;
      for (int k = 0; k < 6; ++k) x[k] = sm[k];
      dom = (int)sm[6];
    }
    const bf16x8* embl = (emb_lds && ((s0 >> 4) & 1) == 0) ? embs : nullptr;
    const bf16x8 xb = askSlots(M, embl, g, x[0], x[1], x[2], x[3], x[4], x[5], dom);
    const float z = inferColumns(w1f, w2f, b1s, b2s, w3s, M.p[oB3], xb);
    if (l < 16 && sidx < n) out[sidx] = z;
  }
}

int spray_mlp_logits(spray_ctx* ctx, const float* samples, int n, const float* params,
                     float* logits_out) {
  if (!ctx || !samples || !logits_out || n <= 0) return SPRAY_ERR_INVALID;
  return guard(ctx, [&] {
    Ctx* c = &ctx->c;
    if (!c->loaded) throw StatusError(SPRAY_ERR_STATE, "no scene loaded");
    checkSampleDomains(samples, n, c->pool.ndomains);
    Predictor* R = ensurePredictor(c);
    if (params) {
      SPRAY_HIP(hipMemcpy(R->p.p, params, (size_t)R->nparams * 4, hipMemcpyHostToDevice));
      packWeights(c);
    }
    DevBuf ds, out;
    ds.reserve((size_t)n * 32);
    out.reserve((size_t)n * 4);
    SPRAY_HIP(hipMemcpy(ds.p, samples, (size_t)n * 32, hipMemcpyHostToDevice));
    const int nd = c->pool.ndomains;
    kLogitsInfer<<<std::min(1024, (n + 15) / 16), 64, embLdsBytes(nd), c->stream>>>(
        R->dev(), nd, nd <= kEmbLdsMax ? 1 : 0, ds.as<float>(), n, out.as<float>());
    SPRAY_HIP(hipStreamSynchronize(c->stream));
    SPRAY_HIP(hipMemcpy(logits_out, out.p, (size_t)n * 4, hipMemcpyDeviceToHost));
  });
}

int spray_mlp_num_params(spray_ctx* ctx) {
  if (!ctx || !ctx->c.loaded) return 0;
  return kDenseParams + ctx->c.pool.ndomains * kEmb;
}

}  // extern "C"
