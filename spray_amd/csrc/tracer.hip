// The per-rank speculative tracer on the GPU (ooc/ooc_tracer.inl,
// ooc/ooc_pcontext.h, ooc/ooc_tcontext.{h,inl}, ooc/ooc_shader_{ao,pt}.h).
//
// One blocking tile at a time (render/tile.cc:317-391).  A window covers
// SPRAY_HISTORY_SIZE = 4 virtual depths; pass k of a window traces the rays
// of virtual depth k plus the shadow rays spawned in pass k-1, so a window
// has at most 5 passes.  Each pass:
//   count    - domain list of every new ray (slab test against the domain
//              AABBs staged in LDS); per-domain entry counts and the
//              DomainStats score sum_d count(d) * max(0, 16 - d)
//   schedule - exclusive offsets, domains ordered by score (ties by id),
//              work items for the batched form
//   scatter  - per-domain queues of (ray, tdom)
//   trace    - per domain: filter (tdom <= history[depth] and VBuf chain
//              correct), closest hit, VBuf update, shade and spawn; shadow
//              entries: filter, any hit, occlusion flag
//   retire   - shadow rays tested in this pass add their weight when
//              unoccluded and still on the final history chain
// The plain form launches one trace kernel per scheduled domain in score
// order; the batched form traces every queue of a pass in one persistent
// launch and captures a whole tile in a hipGraph.
#include <hip/hip_runtime.h>

#include <sys/stat.h>

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <string>

#include "ctx.h"
#include "tracer_dev.h"

namespace spray {

constexpr int kBlock = 256;
// grid of kCount and kScatter (they must match: blocks own queue slots)
constexpr int kQueueGrid = 1024;
constexpr int kShadeRec = 5;  // float4s per shade record
// Ask records per kCount block region: a block takes at most kBlock of
// every kQueueGrid * kBlock consecutive rays of a pass.
inline uint32_t askRegion(uint32_t r_cap) {
  const uint32_t stride = (uint32_t)kQueueGrid * kBlock;
  return (uint32_t)kBlock * ((r_cap + stride - 1) / stride);
}
// Domain grid of the domain-list pass: a uniform grid over the domain boxes
// (at most kGridCells cells, kGridMaxRes per axis), each cell listing the
// domains whose box
// overlaps it with positive extent (cells [floor(lo), ceil(hi) - 1] per
// axis, so a box that fills a cell is listed once); at most
// kGridIdsPerDomain list entries per domain on average (the grid is
// coarsened until that holds).  Contacts on cell faces, edges and corners
// are the walk's business (kGridTol).
constexpr int kGridCells = 32768;
constexpr int kGridMaxRes = 64;
constexpr int kGridIdsPerDomain = 32;
constexpr float kGridTol = 1e-4f;
// dbox bytes: domain boxes, grid header (4 float4), cell offsets, cell ids
__host__ __device__ inline size_t gridHeadBytes(int nd) { return (size_t)nd * 32; }
__host__ __device__ inline size_t gridOffBytes(int nd) { return gridHeadBytes(nd) + 64; }
__host__ __device__ inline size_t gridIdsBytes(int nd) {
  return gridOffBytes(nd) + (size_t)(kGridCells + 1) * 4;
}
__host__ __device__ inline size_t dboxBytes(int nd) {
  return gridIdsBytes(nd) + (size_t)kGridIdsPerDomain * nd * 2 + 64;
}

// kCount's dynamic LDS (byte offsets): domain boxes, counts, scores, the
// block's hit-mask words ([word][thread]).
struct CountLds {
  size_t lc, ls, mask, total;
};
__host__ __device__ inline CountLds countLds(int nd, int block) {
  const int nw = (nd + 63) / 64;
  CountLds L;
  L.lc = (size_t)nd * 32;
  L.ls = L.lc + (size_t)nd * 8;
  L.mask = (L.ls + (size_t)nd * 4 + 7) & ~(size_t)7;
  L.total = L.mask + (size_t)nw * block * 8;
  return L;
}

// Grid header as uploaded: (lo, res x bits), (hi, res y bits), (cell size,
// res z bits), (1 / cell size, 0).
struct DomGrid {
  float lo[3], hi[3], h[3], inv[3];
  int r[3];
};
constexpr int kShadeBlock = 512;  // threads per shading block (one pool atomic per block)
// The trace kernel's waves reserve queue entries kRange at a time (one
// atomic per range, not per refill); the ranges still follow queue order,
// so the waves of a launch work through the domains together.
constexpr uint32_t kRange = 256;
// The persistent launches hand out ranges through kWorkCounters counters
// (a wave starts at its block's, block index mod kWorkCounters, which is its
// XCD under round-robin dispatch), kWorkStride words apart: device-scope
// atomics on one address are served one at a time, and one counter for a
// launch's 250k ranges kept that address busy much of the launch.
constexpr int kWorkCounters = 8;
constexpr int kWorkStride = 1024;
static_assert((kWorkCounters & (kWorkCounters - 1)) == 0, "a wave steps through the counters mod K");
// kCount's lanes take new rays once this many of a wave's lanes wait.
constexpr int kCountRefill = 16;
// Shade slots are taken kShadeChunk at a time per wave.
constexpr uint32_t kShadeChunk = 64;
static_assert(kQueueGrid % kBlock == 0, "kBlockBases splits the grid evenly over its threads");

// Block-wide flush of N wave-uniform counts: one atomic per counter and
// block instead of per wave (waves that end together would otherwise queue
// on the same device-scope atomics).  Every thread of the block calls it.
template <int N>
__device__ __forceinline__ void blockFlush(unsigned long long* const (&dst)[N],
                                           const unsigned long long (&v)[N]) {
  __shared__ unsigned long long part[16][N];
  const int w = threadIdx.x >> 6;
  if (__lane_id() == 0)
    for (int k = 0; k < N; ++k) part[w][k] = v[k];
  __syncthreads();
  if (threadIdx.x < N) {
    unsigned long long sum = 0;
    for (int i = 0; i < (int)(blockDim.x >> 6); ++i) sum += part[i][threadIdx.x];
    if (sum) atomicAdd(dst[threadIdx.x], sum);
  }
}

// Appends to a ray pool with one device-scope atomic per block and call
// (same-address atomics from every wave queue up at the device level).
// Every thread of the block makes the same sequence of calls; the two LDS
// buffers alternate, so a call never overwrites words the previous call's
// readers may still need.
struct BlockAppend {
  uint32_t* sm;  // 2 x 34 words of LDS
  uint32_t flip;
  __device__ uint32_t operator()(bool pred, uint32_t* counter, uint32_t cap, uint32_t* err) {
    uint32_t* b = sm + 34 * flip;
    flip ^= 1u;
    const unsigned long long m = __ballot(pred);
    const int lane = __lane_id(), w = threadIdx.x >> 6;
    if (lane == 0) b[w] = (uint32_t)__popcll(m);
    __syncthreads();
    if (threadIdx.x == 0) {
      uint32_t acc = 0;
      for (int i = 0; i < (int)(blockDim.x >> 6); ++i) {
        const uint32_t x = b[i];
        b[i] = acc;
        acc += x;
      }
      const uint32_t base = acc ? atomicAdd(counter, acc) : 0u;
      b[16] = base;
      if (acc && base + acc > cap) atomicOr(err, 1u);
    }
    __syncthreads();
    const uint32_t idx = b[16] + b[w] + (uint32_t)__popcll(m & ((1ull << lane) - 1ull));
    return (pred && idx < cap) ? idx : 0xffffffffu;
  }
  // Up to two items into each of two pools in one call (the same two
  // barriers): pool A takes every wave's a1 items, then its a2 items; pool B
  // its b1 then b2 items.  idx gets the a1, a2, b1, b2 slots (~0u: none).
  __device__ void pair(bool a1, bool a2, uint32_t* ca, uint32_t capa, bool b1, bool b2,
                       uint32_t* cb, uint32_t capb, uint32_t* err, uint32_t idx[4]) {
    uint32_t* b = sm + 34 * flip;
    flip ^= 1u;
    const unsigned long long ma1 = __ballot(a1), ma2 = __ballot(a2), mb1 = __ballot(b1),
                             mb2 = __ballot(b2);
    const int lane = __lane_id(), w = threadIdx.x >> 6;
    const uint32_t na1 = (uint32_t)__popcll(ma1), nb1 = (uint32_t)__popcll(mb1);
    if (lane == 0) {
      b[w] = na1 + (uint32_t)__popcll(ma2);
      b[17 + w] = nb1 + (uint32_t)__popcll(mb2);
    }
    __syncthreads();
    if (threadIdx.x < 2) {  // lane 0: pool A, lane 1: pool B
      uint32_t* bb = b + 17 * threadIdx.x;
      uint32_t acc = 0;
      for (int i = 0; i < (int)(blockDim.x >> 6); ++i) {
        const uint32_t x = bb[i];
        bb[i] = acc;
        acc += x;
      }
      const uint32_t cap = threadIdx.x ? capb : capa;
      const uint32_t base = acc ? atomicAdd(threadIdx.x ? cb : ca, acc) : 0u;
      bb[16] = base;
      if (acc && base + acc > cap) atomicOr(err, 1u);
    }
    __syncthreads();
    const unsigned long long lt = (1ull << lane) - 1ull;
    const uint32_t ba = b[16] + b[w], bb = b[33] + b[17 + w];
    const uint32_t i0 = ba + (uint32_t)__popcll(ma1 & lt), i1 = ba + na1 + (uint32_t)__popcll(ma2 & lt);
    const uint32_t i2 = bb + (uint32_t)__popcll(mb1 & lt), i3 = bb + nb1 + (uint32_t)__popcll(mb2 & lt);
    idx[0] = (a1 && i0 < capa) ? i0 : 0xffffffffu;
    idx[1] = (a2 && i1 < capa) ? i1 : 0xffffffffu;
    idx[2] = (b1 && i2 < capb) ? i2 : 0xffffffffu;
    idx[3] = (b2 && i3 < capb) ? i3 : 0xffffffffu;
  }
};

// ---------------------------------------------------------------------------
// Tile / window bookkeeping (single-thread kernels keep the whole loop on
// the device so that it can be captured in a graph).
__global__ void kTileStart(TraceParams P) {
  DevCounters* c = P.ctr;
  uint32_t S = (uint32_t)(P.tile_rec[1] * P.image_w * P.spp);
  c->r_count = S;
  c->r_init = S;
  c->r_mark = 0;
  c->s_count = 0;
  c->s_mark = 0;
  c->p_count = 0;
  c->useful += S;  // eye rays
}

__global__ void kGenEyes(TraceParams P) {
  const int tile_y = P.tile_rec[0], tile_h = P.tile_rec[1];
  const long long S = (long long)tile_h * P.image_w * P.spp;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < S;
       i += (long long)gridDim.x * blockDim.x) {
    // samid = index in the batch of tiles (history-buffer slot); the
    // jitter seed is the tile-local sample index, as with one tile at a time
    int bufid = (int)(i % P.tile_samples);
    int pix_local = (int)(i / P.spp);
    int x = pix_local % P.image_w;
    int y = tile_y + pix_local / P.image_w;
    f3 d;
    if (P.spp > 1) {
      uint32_t rng = rngInit(bufid);
      float fx = (float)x + rng1D(rng);
      float fy = (float)y + rng1D(rng);
      d = cameraDir(P.cam, fx, fy);
    } else {
      d = cameraDir(P.cam, (float)x, (float)y);
    }
    P.R.op[i] = make_float4(P.cam.pos.x, P.cam.pos.y, P.cam.pos.z, __int_as_float(y * P.image_w + x));
    P.R.ds[i] = make_float4(d.x, d.y, d.z, __int_as_float((int)i));
    P.R.wd[i] = make_float4(1.f, 1.f, 1.f, __int_as_float(0));
    P.R.h[i] = make_float4(kInfF, kInfF, kInfF, kInfF);
  }
}

// Window restart (ooc_tcontext.h:206-216 procPendingQ, with the pending
// ray's full 4-entry parent history validated: deviation Q1)
__global__ void kWindowStart(TraceParams P) {
  DevCounters* c = P.ctr;
  c->r_count = 0;
  c->r_mark = 0;
  c->s_count = 0;
  c->s_mark = 0;
}
__global__ void kRestartPending(TraceParams P) {
  __shared__ uint32_t app_sm[2 * 34];
  DevCounters* c = P.ctr;
  BlockAppend app{app_sm, 0u};
  const uint32_t n = min(c->p_count, P.p_cap);
  unsigned long long useful = 0;
  for (uint32_t base = blockIdx.x * blockDim.x; base < n; base += gridDim.x * blockDim.x) {
    uint32_t i = base + threadIdx.x;
    bool valid = i < n;
    float4 hh = valid ? P.Pn.h[i] : make_float4(0, 0, 0, 0);
    float h[4] = {hh.x, hh.y, hh.z, hh.w};
    int samid = valid ? __float_as_int(P.Pn.ds[i].w) : 0;
    bool ok = valid && vbufCorrect(P.vbuf, samid, h, kHistory);
    const uint32_t idx = app(ok, &c->r_count, P.r_cap, &c->err);
    if (ok && idx != 0xffffffffu) {
      P.R.op[idx] = P.Pn.op[i];
      P.R.ds[idx] = P.Pn.ds[i];
      float4 wd = P.Pn.wd[i];
      P.R.wd[idx] = make_float4(wd.x, wd.y, wd.z, __int_as_float(0));
      P.R.h[idx] = make_float4(kInfF, kInfF, kInfF, kInfF);
      ++useful;
    }
  }
  blockFlush<1>({&c->useful}, {waveSumULL(useful)});
}
__global__ void kRestartDone(TraceParams P) {
  DevCounters* c = P.ctr;
  c->r_init = min(c->r_count, P.r_cap);
  c->p_count = 0;
}

__global__ void kVbufReset(TraceParams P) {
  const long long n = (long long)P.tile_rec[1] * P.image_w * P.spp * kHistory;
  unsigned long long inf = mkW(kInfF, 0u);
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    P.vbuf[i] = inf;
}

__global__ void kPassBegin(TraceParams P) {
  DevCounters* c = P.ctr;
  // pool levels past capacity (an overflow, reported to the host) are clamped
  // so that no kernel reads beyond a pool
  c->r_beg = c->r_mark;
  c->r_end = min(c->r_count, P.r_cap);
  c->r_mark = c->r_end;
  c->s_beg = c->s_mark;
  c->s_end = min(c->s_count, P.s_cap);
  c->s_mark = c->s_end;
  // the pass's domain hit masks hold at most mask_cap rays: a larger pass is
  // cut to that size and reported as a pool overflow (grow and render again)
  const uint32_t nr = c->r_end - c->r_beg, ns = c->s_end - c->s_beg;
  if ((unsigned long long)nr + ns > P.mask_cap) {
    c->err |= 1u;
    c->r_end = c->r_beg + min(nr, P.mask_cap);
    c->s_end = c->s_beg + min(ns, P.mask_cap - (c->r_end - c->r_beg));
    c->r_mark = c->r_end;
    c->s_mark = c->s_end;
  }
  c->active = (c->r_end > c->r_beg || c->s_end > c->s_beg) ? 1u : 0u;
  if (c->active) c->passes++;
  c->vlog_count = 0;
  // the visit log runs only while the sample log has room
  c->log_live = (P.log_on && *P.log_count < P.log_cap) ? 1u : 0u;
  c->shade_count = 0;
  c->sh_mark = 0;
  c->s_local = c->s_mark;
  c->n_phase[0] = 0;
  c->n_phase[1] = 0;
}

__global__ void kClearQ(TraceParams P) {
  if (!P.ctr->active) return;
  int nd = P.ndomains;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < nd; i += gridDim.x * blockDim.x) {
    P.cnt[i] = 0;
    P.cnt[nd + i] = 0;
    P.score[i] = 0;
  }
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < nd + 3; i += gridDim.x * blockDim.x)
    P.wctr[i] = 0;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < 3 * kWorkCounters; i += gridDim.x * blockDim.x)
    P.wq[i * kWorkStride] = 0;
}

// Ray of the current pass: index into the radiance range, then the shadow range.
__device__ __forceinline__ bool passRay(const TraceParams& P, long long i, int* kind,
                                        uint32_t* ray) {
  const DevCounters* c = P.ctr;
  long long nr = (long long)c->r_end - c->r_beg;
  long long ns = (long long)c->s_end - c->s_beg;
  if (i < nr) {
    *kind = 0;
    *ray = c->r_beg + (uint32_t)i;
    return true;
  }
  if (i < nr + ns) {
    *kind = 1;
    *ray = c->s_beg + (uint32_t)(i - nr);
    return true;
  }
  return false;
}

// One ray's domain list, a walk that a lane resumes one cell step at a
// time: per-domain counts (lc), the 16 nearest (t, id) entries sorted
// (registers), the domains hit as a bit mask (one word per 64 domains, this
// lane's words at mask[w * stride]) so that the scatter re-tests only those,
// and the number of domains hit.  Candidates come from
// a 3D-DDA walk of the domain grid over the ray's t range inside the grid
// (from max(0, entry)).  The walk is a
// supercover where it matters: when the ray crosses two or three cell faces
// at the same t (within kGridTol of t), it also visits the cells of the
// partial steps, and it starts with the neighbours across the faces its
// first point lies on (within kGridTol of a cell), so every cell whose
// closed region the ray touches is visited and a domain touched only on a
// cell face, edge or corner is still a candidate.  Each candidate not
// already marked gets the exact slab test, and the list is kept sorted by
// (t, id), so the result does not depend on the grid or on the visiting
// order.
// What every lane of a block shares: domain boxes, grid, cell lists, the
// block's counts (LDS) and the lanes' hit-mask words (LDS).
struct DomEnv {
  const float4* box;
  DomGrid G;
  const uint32_t* __restrict__ coff;
  const uint16_t* __restrict__ cids;
  uint32_t* lc;
  unsigned long long* mask;  // this lane's words at mask[w * stride]
  int stride, nd;
};

struct DomWalk {
  // the ray and its list
  float o[3], d[3], inv[3];
  int spawn, kind;
  float tt[kTopDepths];
  int ti[kTopDepths];
  int kept, hits;
  // the walk: current cell (8 bits per axis, kGridMaxRes <= 256), step per
  // axis (2 bits each, -1..1 stored + 1), exit t, steps left (0: done)
  uint32_t cell, steps;
  float t1;
  int left;

  // candidates of one cell (cells outside the grid have none)
  __device__ __forceinline__ void visit(const DomEnv& E, int x, int y, int z, bool safe) {
    const DomGrid& G = E.G;
    if (x < 0 || y < 0 || z < 0 || x >= G.r[0] || y >= G.r[1] || z >= G.r[2]) return;
    const int cl = (z * G.r[1] + y) * G.r[0] + x;
    const uint32_t q1 = E.coff[cl + 1];
    for (uint32_t q = E.coff[cl]; q < q1; ++q) {
      const int dd = E.cids[q];
      unsigned long long* mw = &E.mask[(dd >> 6) * E.stride];
      if ((*mw >> (dd & 63)) & 1ull) continue;  // already a hit
      float tmin;
      // isectWithoutCurrentDomain (ooc_isector.h:147-174): a shadow ray's
      // spawning domain was tested right after shading
      if (!slabBox4(E.box[2 * dd], E.box[2 * dd + 1], o, inv, safe, &tmin) || dd == spawn) continue;
      ++hits;
      atomicAdd(&E.lc[kind * E.nd + dd], 1u);
      *mw |= 1ull << (dd & 63);
      // insertion into the (t, id)-sorted top list (unrolled: stays in
      // registers); steps past every active lane's list end are skipped
      if (kept < kTopDepths || tmin < tt[kTopDepths - 1]) {
        float ct = tmin;
        int ci = dd;
#pragma unroll
        for (int j = 0; j < kTopDepths; ++j) {
          if (__ballot(j <= kept) == 0ull) break;
          if (j < kept) {
            if (ct < tt[j] || (ct == tt[j] && ci < ti[j])) {
              float xt = tt[j];
              int xi = ti[j];
              tt[j] = ct;
              ti[j] = ci;
              ct = xt;
              ci = xi;
            }
          } else if (j == kept) {
            tt[j] = ct;
            ti[j] = ci;
          }
        }
        if (kept < kTopDepths) ++kept;
      }
    }
  }

  // Starts the walk: the ray's interval inside the grid box (an axis with an
  // infinite 1/dir does not step and only needs the origin inside the slab),
  // the first cell and the neighbours across the faces its point lies on.
  __device__ __forceinline__ void start(const DomEnv& E, bool safe) {
    const DomGrid& G = E.G;
    kept = 0;
    hits = 0;
    left = 0;
    float t0 = 0.f;
    t1 = kInfF;
    bool in = true, moves[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      moves[k] = fabsf(inv[k]) < kInfF;
      if (moves[k]) {
        const float a = (G.lo[k] - o[k]) * inv[k], b = (G.hi[k] - o[k]) * inv[k];
        t0 = fmaxf(t0, fminf(a, b));
        t1 = fminf(t1, fmaxf(a, b));
      } else if (o[k] < G.lo[k] || o[k] > G.hi[k]) {
        in = false;
      }
    }
    if (!(in && t0 <= t1)) return;
    int nb[3], c[3], st[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      const float p = o[k] + t0 * d[k];
      const float u = (p - G.lo[k]) * G.inv[k];
      c[k] = min(max((int)floorf(u), 0), G.r[k] - 1);
      st[k] = !moves[k] ? 0 : (d[k] > 0.f ? 1 : -1);
      const float fr = u - (float)c[k];
      nb[k] = (fr < kGridTol && c[k] > 0) ? -1 : (fr > 1.f - kGridTol && c[k] < G.r[k] - 1) ? 1 : 0;
    }
    const int nbm = (nb[0] != 0 ? 1 : 0) | (nb[1] != 0 ? 2 : 0) | (nb[2] != 0 ? 4 : 0);
    for (int sub = 1; sub < 8; ++sub)
      if ((sub & ~nbm) == 0)
        visit(E, c[0] + ((sub & 1) ? nb[0] : 0), c[1] + ((sub & 2) ? nb[1] : 0),
              c[2] + ((sub & 4) ? nb[2] : 0), safe);
    cell = (uint32_t)c[0] | ((uint32_t)c[1] << 8) | ((uint32_t)c[2] << 16);
    steps = (uint32_t)(st[0] + 1) | ((uint32_t)(st[1] + 1) << 2) | ((uint32_t)(st[2] + 1) << 4);
    left = G.r[0] + G.r[1] + G.r[2] + 3;
  }

  // One step: the current cell, then the next cell(s) the ray enters.
  __device__ __forceinline__ void step(const DomEnv& E, bool safe) {
    const DomGrid& G = E.G;
    int c[3] = {(int)(cell & 255u), (int)((cell >> 8) & 255u), (int)((cell >> 16) & 255u)};
    const int st[3] = {(int)(steps & 3u) - 1, (int)((steps >> 2) & 3u) - 1,
                       (int)((steps >> 4) & 3u) - 1};
    visit(E, c[0], c[1], c[2], safe);
    --left;
    // next cell: the axes whose cell faces the ray crosses first
    float tn[3];
#pragma unroll
    for (int k = 0; k < 3; ++k)
      tn[k] = st[k] == 0 ? kInfF
                         : (G.lo[k] + (float)(c[k] + (st[k] > 0 ? 1 : 0)) * G.h[k] - o[k]) * inv[k];
    const float m = fminf(fminf(tn[0], tn[1]), tn[2]);
    if (!(m <= t1)) {
      left = 0;
      return;
    }
    const float lim = m + kGridTol * fmaxf(1.f, fabsf(m));
    const int tm = (st[0] != 0 && tn[0] <= lim ? 1 : 0) | (st[1] != 0 && tn[1] <= lim ? 2 : 0) |
                   (st[2] != 0 && tn[2] <= lim ? 4 : 0);
    // crossings at (nearly) the same t: the cells of the partial steps too
    for (int sub = 1; sub < 7; ++sub)
      if ((sub & ~tm) == 0 && sub != tm)
        visit(E, c[0] + ((sub & 1) ? st[0] : 0), c[1] + ((sub & 2) ? st[1] : 0),
              c[2] + ((sub & 4) ? st[2] : 0), safe);
#pragma unroll
    for (int k = 0; k < 3; ++k)
      if (tm & (1 << k)) {
        c[k] += st[k];
        if (c[k] < 0 || c[k] >= G.r[k]) left = 0;
      }
    cell = (uint32_t)(c[0] & 255) | ((uint32_t)(c[1] & 255) << 8) | ((uint32_t)(c[2] & 255) << 16);
  }
};

// Isector::isectAll / isectWithoutCurrentDomain (ooc_isector.h:126-174) as
// counts: entries per domain and the DomainStats score
// (ooc_domain_stats.cc:60-88) weight max(0, 16 - position in the ray's list).
// kLearned: the learned form's ask records are written (a separate
// instantiation, so the other forms keep kCount's smaller register set).
// kList: more than 256 domains, rays with at most kTopDepths hit domains
// write the packed list instead of the mask (an instantiation of its own, so
// scenes with fewer domains keep the mask-only register set).
template <bool kLearned, bool kList>
__global__ void __launch_bounds__(kBlock, 4) kCount(TraceParams P) {
  extern __shared__ uint32_t sm[];
  if (!P.ctr->active) return;
  const int nd = P.ndomains, nw = (nd + 63) >> 6;
  const CountLds L = countLds(nd, blockDim.x);
  unsigned char* smc = (unsigned char*)sm;
  float4* box = (float4*)smc;                                 // 2 * nd: (lo, 0), (hi, 0)
  uint32_t* lc = (uint32_t*)(smc + L.lc);                     // 2 * nd counts
  uint32_t* ls = (uint32_t*)(smc + L.ls);                     // nd score (32-bit per block)
  unsigned long long* lmask = (unsigned long long*)(smc + L.mask) + threadIdx.x;
  __shared__ uint32_t ask_n[1];  // learned form: ask records of this block
  if (threadIdx.x == 0) ask_n[0] = 0;
  for (int i = threadIdx.x; i < 2 * nd; i += blockDim.x) box[i] = P.dbox[i];
  DomGrid G;
  {
    const float4* gh = (const float4*)((const char*)P.dbox + gridHeadBytes(nd));
    const float4 a = gh[0], b = gh[1], h = gh[2], v = gh[3];
    G.lo[0] = a.x, G.lo[1] = a.y, G.lo[2] = a.z;
    G.hi[0] = b.x, G.hi[1] = b.y, G.hi[2] = b.z;
    G.h[0] = h.x, G.h[1] = h.y, G.h[2] = h.z;
    G.inv[0] = v.x, G.inv[1] = v.y, G.inv[2] = v.z;
    G.r[0] = __float_as_int(a.w), G.r[1] = __float_as_int(b.w), G.r[2] = __float_as_int(h.w);
  }
  const uint32_t* coff = (const uint32_t*)((const char*)P.dbox + gridOffBytes(nd));
  const uint16_t* cids = (const uint16_t*)((const char*)P.dbox + gridIdsBytes(nd));
  for (int i = threadIdx.x; i < nd; i += blockDim.x) {
    lc[i] = 0;
    lc[nd + i] = 0;
    ls[i] = 0;
  }
  __syncthreads();
  const DevCounters* c = P.ctr;
  const long long total = (long long)(c->r_end - c->r_beg) + (c->s_end - c->s_beg);
  const long long stride = (long long)gridDim.x * blockDim.x;
  // The wave's rays are i = wbase + k * stride + j (j < 64, k = 0, 1, ...),
  // the rays a grid-stride loop would give it, so every block counts the
  // rays kScatter's block places.  Walks differ in length, so a lane takes
  // the wave's next ray as soon as its own is done (once kCountRefill lanes
  // wait), instead of the wave waiting for its longest walk.
  const long long wbase = blockIdx.x * (long long)blockDim.x + (threadIdx.x & ~63);
  const int lane = __lane_id();
  uint32_t seq = 0;  // wave-uniform: rays handed out
  bool more = wbase < total;
  bool busy = false;
  uint32_t i = 0;  // the lane's ray: pass index (< 2^32; its pool index follows from it)
  DomEnv E;
  E.box = box;
  E.G = G;
  E.coff = coff;
  E.cids = cids;
  E.lc = lc;
  E.mask = lmask;
  E.stride = blockDim.x;
  E.nd = nd;
  DomWalk W;
  W.kind = 0;
  W.left = 0;
  W.kept = W.hits = 0;
  W.spawn = -1;
  W.o[0] = W.o[1] = W.o[2] = 0.f;
  W.d[0] = W.d[1] = W.d[2] = 1.f;
  W.inv[0] = W.inv[1] = W.inv[2] = 1.f;
  while (true) {
    const unsigned long long idle = __ballot(!busy);
    const int nidle = __popcll(idle);
    bool fresh = false;
    if (more && (nidle >= kCountRefill || nidle == 64)) {
      if (!busy) {
        const uint32_t q = seq + (uint32_t)__popcll(idle & ((1ull << lane) - 1ull));
        const long long ii = wbase + (long long)(q >> 6) * stride + (q & 63u);
        i = (uint32_t)ii;
        int kind = 0;
        uint32_t r = 0;
        bool valid = ii < total && passRay(P, ii, &kind, &r);
        // shadow rays occluded in their spawning domain are done (procShads2)
        if (valid && kind == 1 && P.Sx.occluded[r]) valid = false;
        if (valid) {
          const RayPool& pool = kind ? P.S : P.R;
          const float4 op = pool.op[r], ds = pool.ds[r];
          W.o[0] = op.x;
          W.o[1] = op.y;
          W.o[2] = op.z;
          W.d[0] = ds.x;
          W.d[1] = ds.y;
          W.d[2] = ds.z;
          W.inv[0] = 1.0f / ds.x;
          W.inv[1] = 1.0f / ds.y;
          W.inv[2] = 1.0f / ds.z;
          W.spawn = kind ? (__float_as_int(pool.wd[r].w) >> 8) : -1;
          W.kind = kind;
          for (int w = 0; w < nw; ++w) lmask[w * blockDim.x] = 0ull;
          // NaN-checked slab tests only in waves with an axis-parallel ray
          W.start(E, __ballot(hasInfInv(W.inv)) != 0ull);
          busy = fresh = true;
        }
      }
      seq += (uint32_t)nidle;
      more = wbase + (long long)(seq >> 6) * stride + (seq & 63u) < total;
    }
    const bool safe = __ballot(busy && hasInfInv(W.inv)) != 0ull;
    if (busy && W.left > 0 && !fresh) W.step(E, safe);
    // a ray whose walk is done: its mask words, its score, (learned) its ask
    const bool fin = busy && W.left <= 0;
    if (fin) {
      unsigned long long* row = P.dmask + (size_t)i * maskRowWords(nd);
      if (kList && W.hits <= kTopDepths) {
        // the top list holds every hit domain: 32 bytes instead of the mask
        uint32_t pk[kTopDepths / 2];
#pragma unroll
        for (int j = 0; j < kTopDepths / 2; ++j)
          pk[j] = (2 * j < W.hits ? (uint32_t)W.ti[2 * j] : 0xffffu) |
                  ((2 * j + 1 < W.hits ? (uint32_t)W.ti[2 * j + 1] : 0xffffu) << 16);
        uint4* r4 = (uint4*)row;
        r4[0] = make_uint4(pk[0], pk[1], pk[2], pk[3]);
        r4[1] = make_uint4(pk[4], pk[5], pk[6], pk[7]);
        P.dfmt[i] = 1;
      } else {
        for (int w = 0; w < nw; ++w) row[w] = lmask[w * blockDim.x];
        if (kList) P.dfmt[i] = 0;
      }
#pragma unroll
      for (int k = 0; k < kTopDepths; ++k)
        if (k < W.kept) atomicAdd(&ls[W.ti[k]], (uint32_t)(kTopDepths - k));
    }
    // learned form: a radiance ray with two or more candidate domains writes
    // an ask record for the predictor into this block's region (one LDS
    // atomic per wave); the others have no phase-A bound
    if (kLearned) {
      const bool rad = fin && W.kind == 0;
      const uint32_t r = c->r_beg + i;  // a radiance ray's pool index
      const bool ask = rad && W.kept >= 2;
      const unsigned long long am = __ballot(ask);
      uint32_t abase = 0;
      if (am != 0ull) {
        const int leader = __builtin_ctzll(am);
        if (lane == leader) abase = atomicAdd(ask_n, (uint32_t)__popcll(am));
        abase = __builtin_amdgcn_readlane(abase, leader);
      }
      const uint32_t a = abase + (uint32_t)__popcll(am & ((1ull << lane) - 1ull));
      if (ask && a < P.ask_region) {
        float4* rec = P.ask + ((size_t)blockIdx.x * P.ask_region + a) * 3;
        const float4 ds = P.R.ds[r];
        rec[0] = make_float4(W.o[0], W.o[1], W.o[2], ds.x);
        rec[1] = make_float4(ds.y, ds.z, W.tt[0], W.tt[1]);
        rec[2] = make_float4(__uint_as_float(r), __int_as_float(W.ti[0]), __int_as_float(W.ti[1]),
                             __int_as_float(min(W.hits, 255)));
      } else if (rad) {
        P.rkey[r] = kNoBound;
      }
    }
    if (fin) busy = false;
    if (!more && __ballot(busy) == 0ull) break;
  }
  __syncthreads();
  if (kLearned && threadIdx.x == 0) P.ask_cnt[blockIdx.x] = min(*ask_n, P.ask_region);
  // per-block counts, [kind][domain][block]: kScatter's blocks take the same
  // rays and place their entries at the bases these counts give
  for (int j = threadIdx.x; j < 2 * nd; j += blockDim.x)
    P.bcnt[(size_t)j * kQueueGrid + blockIdx.x] = lc[j];
  for (int i = threadIdx.x; i < nd; i += blockDim.x) {
    if (lc[i]) atomicAdd(&P.cnt[i], lc[i]);
    if (lc[nd + i]) atomicAdd(&P.cnt[nd + i], lc[nd + i]);
    if (ls[i]) atomicAdd(&P.score[i], (unsigned long long)ls[i]);
  }
}

// Schedule (DomainStats::schedule, ooc_domain_stats.cc:60-111: descending
// score, ties by id) and queue offsets laid out in schedule order, in one block.
__device__ __forceinline__ uint32_t blockScanExcl(uint32_t v, uint32_t* wsum, uint32_t* total) {
  uint32_t x = v;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  for (int o = 1; o < 64; o <<= 1) {
    uint32_t y = __shfl_up(x, o, 64);
    if (lane >= o) x += y;
  }
  if (lane == 63) wsum[w] = x;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t acc = 0;
    for (int i = 0; i < (int)(blockDim.x >> 6); ++i) {
      uint32_t t = wsum[i];
      wsum[i] = acc;
      acc += t;
    }
    wsum[16] = acc;
  }
  __syncthreads();
  uint32_t r = wsum[w] + x - v;
  *total = wsum[16];
  __syncthreads();
  return r;
}

__global__ void __launch_bounds__(1024) kSchedule(TraceParams P) {
  DevCounters* c = P.ctr;
  if (!c->active) return;
  const int nd = P.ndomains;
  __shared__ uint32_t wsum[16 + 1];
  for (int d = threadIdx.x; d < nd; d += blockDim.x) {
    unsigned long long sd = P.score[d];
    int rank = 0;
    for (int e = 0; e < nd; ++e) {
      unsigned long long se = P.score[e];
      rank += (se > sd) || (se == sd && e < d);
    }
    P.order[rank] = d;
  }
  __syncthreads();
  for (int kind = 0; kind < 2; ++kind) {
    uint32_t carry = 0;
    for (int base = 0; base < nd; base += blockDim.x) {
      int k = base + threadIdx.x;
      int d = k < nd ? P.order[k] : 0;
      uint32_t v = k < nd ? P.cnt[kind * nd + d] : 0u;
      uint32_t total;
      uint32_t pre = blockScanExcl(v, wsum, &total);
      if (k < nd) {
        P.off[kind * (nd + 1) + k] = carry + pre;
        P.dbase[kind * nd + d] = carry + pre;
      }
      carry += total;
    }
    if (threadIdx.x == 0) {
      P.off[kind * (nd + 1) + nd] = carry;
      c->q_total[kind] = min(carry, P.q_cap[kind]);
      c->queue_entries += carry;
      if (carry > P.q_cap[kind]) atomicOr(&c->err, 1u);
    }
    __syncthreads();
  }
}

// The queue entries of one ray, at its block's base plus the block-local
// slot of each domain; only the domains kCount marked are tested again
// (for their entry distance).
// One queue entry of ray r for domain dd, at the domain's entry distance.
__device__ __forceinline__ void scatterOne(const float4* box, const float o[3], const float inv[3],
                                           bool safe, int dd, uint32_t* lc_kind,
                                           const uint32_t* lb_kind, const TraceParams& P, int kind,
                                           uint32_t r) {
  float tmin;
  slabBox4(box[2 * dd], box[2 * dd + 1], o, inv, safe, &tmin);
  const uint32_t dst = lb_kind[dd] + atomicAdd(&lc_kind[dd], 1u);
  if (dst < P.q_cap[kind]) P.q[kind][dst] = QEntry{r, tmin};
}

__device__ __forceinline__ void scatterDomains(const float4* box, int nd, const float o[3],
                                               const float inv[3], bool safe,
                                               const unsigned long long* mask, uint32_t* lc_kind,
                                               const uint32_t* lb_kind, const TraceParams& P,
                                               int kind, uint32_t r) {
  for (int w = 0; w < ((nd + 63) >> 6); ++w) {
    unsigned long long m = mask[w];
    while (m) {
      const int dd = (w << 6) + __builtin_ctzll(m);
      m &= m - 1ull;
      scatterOne(box, o, inv, safe, dd, lc_kind, lb_kind, P, kind, r);
    }
  }
}

// Queue base of every (kind, domain, block of kCount/kScatter): the
// domain's base in schedule order plus the entries of the lower blocks.
// One block per (kind, domain).
__global__ void __launch_bounds__(kBlock) kBlockBases(TraceParams P) {
  __shared__ uint32_t wsum[16 + 1];
  if (!P.ctr->active) return;
  const int j = blockIdx.x;  // kind * nd + domain
  constexpr int kPer = kQueueGrid / kBlock;
  const uint32_t* src = P.bcnt + (size_t)j * kQueueGrid + threadIdx.x * kPer;
  uint32_t v[kPer], sum = 0;
#pragma unroll
  for (int k = 0; k < kPer; ++k) {
    v[k] = src[k];
    sum += v[k];
  }
  uint32_t total;
  uint32_t pre = P.dbase[j] + blockScanExcl(sum, wsum, &total);
  uint32_t* dst = P.bbase + (size_t)j * kQueueGrid + threadIdx.x * kPer;
#pragma unroll
  for (int k = 0; k < kPer; ++k) {
    dst[k] = pre;
    pre += v[k];
  }
}

// Scatter the queue entries: the same grid and ray-to-block mapping as
// kCount, so each block owns the slots kBlockBases reserved for it.
// kList: the rays kCount gave a packed hit list (as kCount's kList).
template <bool kList>
__global__ void __launch_bounds__(kBlock) kScatter(TraceParams P) {
  extern __shared__ uint32_t sm[];
  if (!P.ctr->active) return;
  const int nd = P.ndomains;
  float4* box = (float4*)sm;                 // 2 * nd: (lo, 0), (hi, 0)
  uint32_t* lc = (uint32_t*)(box + 2 * nd);  // 2 * nd block-local slots
  uint32_t* lb = lc + 2 * nd;                // 2 * nd block bases
  for (int i = threadIdx.x; i < 2 * nd; i += blockDim.x) {
    box[i] = P.dbox[i];
    lc[i] = 0;
    lb[i] = P.bbase[(size_t)i * kQueueGrid + blockIdx.x];
  }
  __syncthreads();
  const DevCounters* c = P.ctr;
  const long long total = (long long)(c->r_end - c->r_beg) + (c->s_end - c->s_beg);
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long base = blockIdx.x * (long long)blockDim.x; base < total; base += stride) {
    long long i = base + threadIdx.x;
    int kind = 0;
    uint32_t r = 0;
    bool valid = passRay(P, i, &kind, &r);
    float o[3] = {0, 0, 0}, inv[3] = {1, 1, 1};
    if (valid) {
      const RayPool& pool = kind ? P.S : P.R;
      float4 op = pool.op[r], ds = pool.ds[r];
      o[0] = op.x;
      o[1] = op.y;
      o[2] = op.z;
      inv[0] = 1.0f / ds.x;
      inv[1] = 1.0f / ds.y;
      inv[2] = 1.0f / ds.z;
      if (kind && P.Sx.occluded[r]) valid = false;  // occluded in its spawning domain
    }
    // the same wave-uniform slab variant as kCount (same rays per wave)
    const bool safe = __ballot(valid && hasInfInv(inv)) != 0ull;
    const unsigned long long* row = P.dmask + (size_t)i * maskRowWords(nd);
    if (kList && valid && P.dfmt[i]) {
      // packed ids of the ray's hit domains (kCount's top list)
      const uint4* r4 = (const uint4*)row;
      const uint4 a = r4[0], b = r4[1];
      const uint32_t pk[kTopDepths / 2] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w};
#pragma unroll
      for (int j = 0; j < kTopDepths; ++j) {
        const int dd = (int)((pk[j >> 1] >> (16 * (j & 1))) & 0xffffu);
        if (dd == 0xffff) break;
        scatterOne(box, o, inv, safe, dd, lc + kind * nd, lb + kind * nd, P, kind, r);
      }
    } else if (valid) {
      scatterDomains(box, nd, o, inv, safe, row, lc + kind * nd, lb + kind * nd, P, kind, r);
    }
  }
}

// ---------------------------------------------------------------------------
// Shading (ooc_shader_ao.h:81-197, ooc_shader_pt.h:93-227) and spawning.
__device__ __forceinline__ void writeRay(const RayPool& pool, uint32_t idx, f3 org, int pixid,
                                         f3 dir, int samid, f3 w, int depth_word,
                                         const float h[4]) {
  pool.op[idx] = make_float4(org.x, org.y, org.z, __int_as_float(pixid));
  pool.ds[idx] = make_float4(dir.x, dir.y, dir.z, __int_as_float(samid));
  pool.wd[idx] = make_float4(w.x, w.y, w.z, __int_as_float(depth_word));
  pool.h[idx] = make_float4(h[0], h[1], h[2], h[3]);
}

// Children's history: parent's history[0..depth) then t, the rest infinity.
__device__ __forceinline__ void childHistory(const float ph[4], int depth, float t, float h[4]) {
#pragma unroll
  for (int k = 0; k < kHistory; ++k) h[k] = (k < depth) ? ph[k] : (k == depth ? t : kInfF);
}


__device__ __forceinline__ void spawnShadow(const TraceParams& P, BlockAppend& app, bool pred,
                                            f3 pos, int pixid, f3 wi, int samid, f3 Lr, int depth,
                                            int dom, const float hist[4], float t, float tmax) {
  uint32_t idx = app(pred, &P.ctr->s_count, P.s_cap, &P.ctr->err);
  if (idx == 0xffffffffu) return;
  float h[4];
  childHistory(hist, depth, t, h);
  writeRay(P.S, idx, pos, pixid, wi, samid, Lr, (depth + 1) | (dom << 8), h);
  P.Sx.tmax[idx] = tmax;
  P.Sx.occluded[idx] = 0u;
}

// The hit's bounce rays (reflection b1 along d1, transmission b2 along d2),
// appended together.
__device__ __forceinline__ void spawnBounces(const TraceParams& P, BlockAppend& app, bool b1,
                                             f3 d1, f3 L1, bool b2, f3 d2, f3 L2, bool pending,
                                             f3 pos, int pixid, int samid, int depth,
                                             const float hist[4], float t) {
  // rays whose next virtual depth is 4 go to the pending pool (genR2,
  // ooc_shader_pt.h:73-90) with their parent's full history
  uint32_t ix[4];
  app.pair(b1 && !pending, b2 && !pending, &P.ctr->r_count, P.r_cap, b1 && pending, b2 && pending,
           &P.ctr->p_count, P.p_cap, &P.ctr->err, ix);
  float h[4];
  childHistory(hist, depth, t, h);
  if (ix[0] != 0xffffffffu) writeRay(P.R, ix[0], pos, pixid, d1, samid, L1, depth + 1, h);
  if (ix[1] != 0xffffffffu) writeRay(P.R, ix[1], pos, pixid, d2, samid, L2, depth + 1, h);
  if (ix[2] != 0xffffffffu) writeRay(P.Pn, ix[2], pos, pixid, d1, samid, L1, 0, h);
  if (ix[3] != 0xffffffffu) writeRay(P.Pn, ix[3], pos, pixid, d2, samid, L2, 0, h);
}

// Light sampling (render/light.h:41-89 and the rect light)
__device__ __forceinline__ f3 sampleLight(const DevLight& L, uint32_t& rng, f3 pos, f3 nff, f3* wi,
                                          float* pdf, float* tmax) {
  if (L.type == kPointLight) {
    *wi = normalize(sub(mk(L.pos[0], L.pos[1], L.pos[2]), pos));
    *pdf = 1.0f;
    *tmax = kInfF;
    return mk(L.radiance[0], L.radiance[1], L.radiance[2]);
  }
  float u1 = rng1D(rng);
  float u2 = rng1D(rng);
  if (L.type == kHemisphereLight) {
    cosineSample(u1, u2, nff, wi, pdf);
    *tmax = kInfF;
    return mk(L.radiance[0], L.radiance[1], L.radiance[2]);
  }
  f3 p = add(add(mk(L.pos[0], L.pos[1], L.pos[2]), scl(u1, mk(L.e1[0], L.e1[1], L.e1[2]))),
             scl(u2, mk(L.e2[0], L.e2[1], L.e2[2])));
  f3 d = sub(p, pos);
  float dist2 = dot(d, d);
  float dist = sqrtf(dist2);
  *wi = scl(d, 1.0f / dist);
  float cosl = -dot(mk(L.normal[0], L.normal[1], L.normal[2]), *wi);
  *tmax = dist;
  *pdf = (cosl > 0.0f) ? dist2 / (L.area * cosl) : 0.0f;
  return mk(L.radiance[0], L.radiance[1], L.radiance[2]);
}

__device__ __forceinline__ float frDielectric(float cosI, float etaI, float etaT, f3 wo, f3 n, f3* wt) {
  float sin2I = fmaxf(0.0f, 1.0f - (cosI * cosI));
  float eta = etaI / etaT;
  float sin2T = eta * eta * sin2I;
  if (sin2T >= 1.0f) return 1.0f;
  float cosT = sqrtf(1.0f - sin2T);
  float Rparl = ((etaT * cosI) - (etaI * cosT)) / ((etaT * cosI) + (etaI * cosT));
  float Rperp = ((etaI * cosI) - (etaT * cosT)) / ((etaI * cosI) + (etaT * cosT));
  float fr = (Rparl * Rparl + Rperp * Rperp) / 2.0f;
  *wt = add(scl(eta, neg(wo)), scl(n, eta * cosI - cosT));
  return fr;
}

// All lanes of a wave call this together (the spawn appends are ballots);
// `act` marks lanes whose hit was accepted by the history buffer.
__device__ __forceinline__ void shadeSpawn(const TraceParams& P, BlockAppend& app, bool act,
                                           int dom, float4 op,
                                           float4 ds, float4 wd, const float hist[4], int depth,
                                           float t, float u, float v, int prim, int ray_depth) {
  const DomainDesc& D = P.desc[dom];
  f3 org = mk(op.x, op.y, op.z), dir = mk(ds.x, ds.y, ds.z);
  int pixid = __float_as_int(op.w), samid = __float_as_int(ds.w);
  f3 pos = mk(dir.x * t + org.x, dir.y * t + org.y, dir.z * t + org.z);
  uint32_t color = 0;
  f3 normal = mk(0, 0, 1);
  if (act) {  // render/trimesh_buffer.cc:327-360
    const float4* fs = P.fshade + (size_t)(D.shade_off + prim) * 3;
    const float4 a0 = fs[0], a1 = fs[1], a2 = fs[2];
    uint32_t c0 = __float_as_uint(a0.w), c1 = __float_as_uint(a1.w), c2 = __float_as_uint(a2.w);
    float w = 1.f - u - v;
    uint32_t r = (uint32_t)((float)((c0 >> 16) & 0xff) * w + (float)((c1 >> 16) & 0xff) * u +
                            (float)((c2 >> 16) & 0xff) * v);
    uint32_t g = (uint32_t)((float)((c0 >> 8) & 0xff) * w + (float)((c1 >> 8) & 0xff) * u +
                            (float)((c2 >> 8) & 0xff) * v);
    uint32_t b = (uint32_t)((float)(c0 & 0xff) * w + (float)(c1 & 0xff) * u + (float)(c2 & 0xff) * v);
    color = (r << 16) | (g << 8) | b;
    normal = mk((a0.x * w) + (a1.x * u) + (a2.x * v), (a0.y * w) + (a1.y * u) + (a2.y * v),
                (a0.z * w) + (a1.z * u) + (a2.z * v));
  }
  f3 surf = unpackColor(color);
  f3 wo = neg(dir);
  f3 Lin = mk(wd.x, wd.y, wd.z);
  float cos_theta_i = dot(wo, normal);
  bool entering = (cos_theta_i > 0.0f);
  f3 nff = normalize(entering ? normal : neg(normal));
  const bool delta = (D.mtl != kDiffuse);
  const int next_virtual = depth + 1;
  const int next_actual = ray_depth + next_virtual;
  const bool pending = (next_virtual == kHistory);
  f3 wi;
  float pdf, costheta;
  if (P.ao_mode) {
    const float ao_weight = 1.0f / (float)P.ao_samples;
    for (int l = 0; l < P.ao_samples; ++l) {
      uint32_t rng = rngInit(pixid * (l + 1));
      float u1 = rng1D(rng), u2 = rng1D(rng);
      cosineSample(u1, u2, nff, &wi, &pdf);
      costheta = clamp01(dot(nff, wi));
      f3 Lr = scl(mul(Lin, surf), ((kOneOverPi * costheta) * ao_weight) / pdf);
      spawnShadow(P, app, act && hasPositive(Lr), pos, pixid, wi, samid, Lr, depth, dom, hist, t,
                  P.ao_max_dist);
    }
  } else {
    bool lit = act && !delta;
    uint32_t ls = rngInit((samid % P.tile_samples) * next_actual);
    f3 ks = mk(P.ks[0], P.ks[1], P.ks[2]);
    for (int l = 0; l < P.nlights; ++l) {
      const DevLight& L = P.lights[l];
      int nsamp = (L.type == kPointLight) ? 1 : P.ao_samples;
      for (int s = 0; s < nsamp; ++s) {
        float tmax;
        f3 Li = sampleLight(L, ls, pos, nff, &wi, &pdf, &tmax);
        bool ok = lit && pdf > 0.0f;
        f3 Lr = mk(0, 0, 0);
        if (ok) {
          costheta = clamp01(dot(nff, wi));
          // blinnPhong (render/reflection.h:202-214)
          f3 half = normalize(add(wi, wo));
          float ndh = clamp01(dot(nff, half));
          // with ks = 0 the term ks * pw is exactly 0 (pw lies in [0, 1])
          float pw = P.specular ? (float)pow((double)ndh, (double)P.shininess) : 0.0f;
          f3 bp = mul(Li, add(scl(surf, costheta), scl(ks, pw)));
          float sc = (L.type == kPointLight) ? (1.0f / pdf) : (1.0f / (pdf * (float)nsamp));
          Lr = scl(mul(Lin, bp), sc);
        }
        spawnShadow(P, app, ok && hasPositive(Lr), pos, pixid, wi, samid, Lr, depth, dom, hist, t,
                    tmax);
      }
    }
  }
  // bounce (ooc_shader_pt.h:177-226)
  bool bounce = act && next_actual < P.bounces;
  wo = normalize(wo);
  // up to two bounce rays (reflection, transmission), spawned after the
  // branch so that every thread of the block appends at the same points
  bool b1 = false, b2 = false;
  f3 d1 = mk(0, 0, 0), d2 = mk(0, 0, 0), L1 = mk(0, 0, 0), L2 = mk(0, 0, 0);
  if (delta) {
    bool refl = false, trans = false;
    f3 wr = mk(0, 0, 0), wt = mk(0, 0, 0), Lr1 = mk(0, 0, 0), Lr2 = mk(0, 0, 0);
    if (bounce && cos_theta_i != 0.0f) {
      float ci = fminf(fmaxf(cos_theta_i, -1.0f), 1.0f);
      float ac = fabsf(ci);
      if (!entering) ci = ac;
      float fr = 1.0f;
      if (D.mtl == kMirror) {
        refl = true;
      } else {
        float etaI = entering ? D.eta_a : D.eta_b;
        float etaT = entering ? D.eta_b : D.eta_a;
        if (D.mtl == kGlass) {
          fr = frDielectric(ci, etaI, etaT, wo, nff, &wt);
          refl = (fr != 0.0f);
          trans = (fr != 1.0f);
        } else {
          float sin2I = fmaxf(0.0f, 1.0f - (ci * ci));
          float eta = etaI / etaT;
          float sin2T = eta * eta * sin2I;
          if (sin2T >= 1.0f) {
            refl = true;
            fr = 1.0f;
          } else {
            float cosT = sqrtf(1.0f - sin2T);
            wt = add(scl(eta, neg(wo)), scl(nff, eta * ci - cosT));
            trans = true;
            fr = 0.0f;
          }
        }
      }
      if (refl) {
        wr = normalize(add(neg(wo), scl(2.0f * dot(wo, nff), nff)));
        Lr1 = scl(Lin, fr / ac);
      }
      if (trans) {
        wt = normalize(wt);
        Lr2 = scl(Lin, (1.0f - fr) / ac);
      }
    }
    b1 = refl && hasPositive(Lr1);
    d1 = wr;
    L1 = Lr1;
    b2 = trans && hasPositive(Lr2);
    d2 = wt;
    L2 = Lr2;
  } else {
    f3 Lr = mk(0, 0, 0);
    bool go = false;
    if (bounce) {
      uint32_t rng = rngInit((samid % P.tile_samples) * next_actual);
      float u1 = rng1D(rng), u2 = rng1D(rng);
      cosineSample(u1, u2, nff, &wi, &pdf);
      costheta = clamp01(dot(nff, wi));
      Lr = dvs(scl(scl(mul(Lin, surf), kOneOverPi), costheta), pdf);
      go = hasPositive(Lr);
      if (go && P.rr_start > 0 && next_actual >= P.rr_start) {
        // Russian roulette (new): survive with p = min(1, max(Lr)); the draw
        // continues the bounce sampler's stream
        float p = fmaxf(Lr.x, fmaxf(Lr.y, Lr.z));
        if (p > 1.0f) p = 1.0f;
        float uu = rng1D(rng);
        go = uu < p;
        if (go) Lr = dvs(Lr, p);
      }
    }
    b1 = go;
    d1 = wi;
    L1 = Lr;
  }
  spawnBounces(P, app, b1, d1, L1, b2, d2, L2, pending, pos, pixid, samid, depth, hist, t);
}

// ---------------------------------------------------------------------------
// Trace step of a pass.  Every lane owns one queue entry at a time:
//   radiance (filterRqs + procRads, ooc_tcontext.inl:28-49,137-150): keep the
//     entry if tdom <= history[depth] and the ray's history is still the
//     buffer's chain; closest hit; history-buffer update; accepted hits
//     become shade records;
//   shadow (filterSqs + procShads, ooc_tcontext.inl:51-70,152-168): keep if
//     not yet occluded and still on the chain; any hit sets the flag.
// A wave refetches entries for its idle lanes with one atomic, so long
// traversals do not hold up the rest of the wave.
constexpr int kTravBlock = 256;


// Marks shade slots [cur, end) (fewer than 64) as holding no record.
__device__ __forceinline__ void shadeSlotsEmpty(const TraceParams& P, uint32_t cur, uint32_t end,
                                                int lane) {
  const uint32_t i = cur + (uint32_t)lane;
  if (i < end && i < P.sh_cap)
    P.sh_rec[(size_t)i * kShadeRec + 4] = make_float4(kInfF, 0.f, 0.f, __int_as_float(-1));
}

__global__ void __launch_bounds__(kTravBlock) kTraverse(TraceParams P, int plain_dom, int phase) {
  extern __shared__ uint32_t ring[];
  DevCounters* c = P.ctr;
  if (!c->active) return;
  const int nd = P.ndomains;
  // entry space of this launch: radiance entries [0, nr), shadow entries
  // [nr, nr + ns), then (learned form) deferred radiance entries [nr + ns, total)
  //   plain_dom >= 0: one domain's queues (plain schedule)
  //   plain_dom == -1: every queue of the pass (batched); with the learned
  //     form the phase-A list first and the deferred list last, so that the
  //     deferred entries meet the histories phase A has updated.  When phase
  //     A and the shadow queues hold fewer than 4 entries per lane of the
  //     grid, every wave would start on deferred entries at once: then phase
  //     0 stops after the shadow queues and phase 1 (a second launch) takes
  //     the deferred list; otherwise phase 1 has nothing to do
  //   plain_dom == -2: shadow rays spawned in this pass, in their spawning
  //     domain (procShads2, ooc_tcontext.inl:51-70)
  //   plain_dom == -3: the entry ranges P.rng of both queues (the chunks of
  //     consecutive domains a paged pool traces while they are resident)
  uint32_t nr = 0, ns = 0, nb = 0, rbase = 0, sbase = 0, local_base = 0;
  const uint32_t* rlist = nullptr;
  uint32_t* wc;
  uint32_t* wq = nullptr;  // persistent launches: their kWorkCounters counters
  const bool local = plain_dom == -2;
  if (plain_dom >= 0) {
    rbase = P.dbase[plain_dom];
    sbase = P.dbase[nd + plain_dom];
    nr = P.cnt[plain_dom];
    ns = P.cnt[nd + plain_dom];
    wc = &P.wctr[plain_dom];
  } else if (plain_dom == -3) {
    rbase = P.rng[0];
    nr = P.rng[1];
    sbase = P.rng[2];
    ns = P.rng[3];
    wc = &P.wctr[nd];
    wq = P.wq;
  } else if (local) {
    local_base = c->s_local;
    uint32_t end = min(c->s_count, P.s_cap);
    ns = end > local_base ? end - local_base : 0u;
    wc = &P.wctr[nd + 2];
    wq = P.wq + 2 * kWorkCounters * kWorkStride;
  } else if (!P.learned) {
    nr = c->q_total[0];
    ns = c->q_total[1];
    wc = &P.wctr[nd];
    wq = P.wq;
  } else {
    // learned: the phase-A radiance list, the shadow queues, then the
    // deferred list, filtered against the histories phase A updated
    const uint32_t na = min(c->n_phase[0], P.q_cap[0]), nd2 = min(c->n_phase[1], P.q_cap[0]);
    const bool split = (unsigned long long)na + c->q_total[1] <
                       4ull * gridDim.x * blockDim.x;
    if (phase == 0) {
      nr = na;
      ns = c->q_total[1];
      nb = split ? 0u : nd2;
      rlist = P.qa;
    } else {
      nr = split ? nd2 : 0u;
      rlist = P.qb;
    }
    wc = &P.wctr[nd + phase];
    wq = P.wq + phase * kWorkCounters * kWorkStride;
  }
  // bounded by the queue capacities even after an overflow
  if (!local) {
    nr = rbase < P.q_cap[0] ? min(nr, P.q_cap[0] - rbase) : 0u;
    ns = sbase < P.q_cap[1] ? min(ns, P.q_cap[1] - sbase) : 0u;
  }
  const uint32_t total = nr + ns + nb;
  LaneStack st;
  st.init(ring, P.ovf);
  TravCount tc;
  TravState s;
  unsigned long long rvis = 0, svis = 0, pruned = 0;  // wave-uniform
  const bool log_live = P.log_on && c->log_live;
  const int lane = __lane_id();
  bool active = false, exhausted = false;
  int kind = 0, dom = 0;
  uint32_t ray = 0;
  // wave-uniform: entries left in this wave's range; shade slots left in its chunk
  uint32_t r_cur = 0, r_end = 0, sh_cur = 0, sh_end = 0;
  // per lane: the last queue segment met, [cur_s, cur_e) of queue cur_k - 1
  // (0: none), and its domain
  uint32_t cur_s = 0, cur_e = 0;
  int cur_k = 0, cur_d = 0;
  // small launches keep ranges at one refill's worth so every wave gets work
  const uint32_t wpb = blockDim.x >> 6, nwaves = gridDim.x * wpb;
  const uint32_t range = (unsigned long long)total >= (unsigned long long)nwaves * 1024u ? kRange : 64u;
  // Each wave of the blocks the GPU holds at once takes its first range at a
  // fixed place (wave w: [w * range, (w + 1) * range)), so a launch starts
  // without a burst of atomics on one counter; the counter hands out the
  // entries after those ranges, from dyn0 on.
  const uint32_t sblocks = min(gridDim.x, (uint32_t)max(P.trav_static, 0));
  const uint32_t dyn0 = sblocks * wpb * range;
  bool first = blockIdx.x < sblocks;
  // split counters: counter k hands out the ranges k, k + K, k + 2K, ... of
  // the entries from dyn0 on, so all of them move through the queue together;
  // a wave takes from kc, its block's at first, and moves to the next once
  // kc is spent (a spent counter stays spent), until all K are
  const uint32_t nranges = total > dyn0 ? (total - dyn0 + range - 1) / range : 0u;
  int kc = (int)(blockIdx.x % kWorkCounters), kspent = 0;
  // one launch per domain (plain) takes exactly the slots it fills
  const uint32_t sh_chunk = plain_dom >= 0 ? 0u : kShadeChunk;
  while (true) {
    unsigned long long idle = __ballot(!active);
    int nidle = __popcll(idle);
    if (!exhausted && (nidle >= P.refill || nidle == 64)) {
      if (r_cur >= r_end) {
        uint32_t b;
        if (first) {
          first = false;
          b = (blockIdx.x * wpb + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6)) * range;
        } else if (wq) {
          const int leader = __builtin_ctzll(idle);
          uint32_t j = nranges;
          while (kspent < kWorkCounters) {
            uint32_t i = 0;
            if (lane == leader) i = atomicAdd(wq + kc * kWorkStride, 1u);
            j = (uint32_t)kc + kWorkCounters * __builtin_amdgcn_readfirstlane(
                                                   __builtin_amdgcn_readlane(i, leader));
            if (j < nranges) break;
            kc = (kc + 1) & (kWorkCounters - 1);
            ++kspent;
          }
          b = j < nranges ? dyn0 + j * range : total;
        } else {
          const int leader = __builtin_ctzll(idle);
          uint32_t a = 0;
          if (lane == leader) a = atomicAdd(wc, range);
          a = __builtin_amdgcn_readlane(a, leader);
          b = a < total - min(total, dyn0) ? dyn0 + a : total;
        }
        r_cur = min(b, total);
        r_end = total - r_cur > range ? r_cur + range : total;
        if (r_cur >= total) exhausted = true;
      }
      const uint32_t base = r_cur, limit = r_end;
      r_cur = min(r_cur + (uint32_t)nidle, r_end);
      bool go = false, fetched = false;
      if (!active) {
        uint32_t e = base + (uint32_t)__popcll(idle & ((1ull << lane) - 1ull));
        if (e < limit) {
          fetched = true;
          kind = (e < nr || e >= nr + ns) ? 0 : 1;
          QEntry qe;
          if (local) {
            qe = QEntry{local_base + e, 0.f};
          } else {
            const uint32_t q = kind == 1 ? sbase + (e - nr)
                             : e >= nr + ns ? P.qb[e - nr - ns]
                                            : (rlist ? rlist[e] : rbase + e);
            qe = P.q[kind][q];
            if (plain_dom >= 0) {
              dom = plain_dom;
            } else {
              // The entry's domain: queues are laid out domain by domain in
              // schedule order, so it is the segment of P.off that holds q.
              // A lane's positions rise within a queue, so it keeps the last
              // segment it met (cur_k: kind + 1, 0 none) and searches only
              // for an entry beyond it.
              if (cur_k != kind + 1 || q < cur_s || q >= cur_e) {
                const uint32_t* of = P.off + kind * (nd + 1);
                int rk = 0, hi = nd - 1;
                while (rk < hi) {  // the last segment starting at or before q
                  const int mid = (rk + hi + 1) >> 1;
                  if (of[mid] <= q)
                    rk = mid;
                  else
                    hi = mid - 1;
                }
                cur_k = kind + 1;
                cur_s = of[rk];
                cur_e = of[rk + 1];
                cur_d = P.order[rk];
              }
              dom = cur_d;
            }
          }
          ray = qe.ray;
          const RayPool& pool = kind ? P.S : P.R;
          float4 op = pool.op[ray], ds = pool.ds[ray], wd = pool.wd[ray], hh = pool.h[ray];
          if (local) dom = __float_as_int(wd.w) >> 8;
          float hist[4] = {hh.x, hh.y, hh.z, hh.w};
          const int depth = __float_as_int(wd.w) & 0xff;
          const int samid = __float_as_int(ds.w);
          float tfar = kInfF;
          if (kind == 0) {
            go = qe.tdom <= pick4(hist, depth) && vbufCorrect(P.vbuf, samid, hist, depth);
          } else if (local) {
            go = true;  // spawned from a final-chain hit in this pass
            tfar = P.Sx.tmax[ray];
          } else {
            go = !__hip_atomic_load(&P.Sx.occluded[ray], __ATOMIC_RELAXED,
                                    __HIP_MEMORY_SCOPE_AGENT) &&
                 vbufCorrect(P.vbuf, samid, hist, depth);
            tfar = P.Sx.tmax[ray];
          }
          if (go) {
            const DomainDesc& D = P.desc[dom];
            float o[3] = {op.x, op.y, op.z}, d[3] = {ds.x, ds.y, ds.z};
            travInit(s, (uint32_t)D.node_off, (uint32_t)D.tri_off, o, d, kRayEps, tfar, kind == 1,
                     st);
            active = true;
          }
        }
      }
      rvis += (unsigned long long)__popcll(__ballot(go && kind == 0));
      svis += (unsigned long long)__popcll(__ballot(go && kind == 1));
      pruned += (unsigned long long)__popcll(__ballot(fetched && !go));
    }
    if (__ballot(active) == 0ull) {
      if (exhausted) break;
      continue;
    }
    travNodePhase(s, st, tc, active, P.nodes);
    travLeafPhase(s, st, tc, active, P.tris);
    bool upd = false, logged = false;
    float4 wd = make_float4(0.f, 0.f, 0.f, 0.f), hh = wd;
    if (active && travDone(s)) {
      active = false;
      if (kind == 0) {
        // the history prefix the buffer update checks is fixed for this pass
        wd = P.R.wd[ray];
        hh = P.R.h[ray];
        const int depth = __float_as_int(wd.w) & 0xff;
        const int samid = __float_as_int(P.R.ds[ray].w);
        float hist[4] = {hh.x, hh.y, hh.z, hh.w};
        upd = s.prim >= 0 && vbufUpdate(P.vbuf, samid, hist, depth, s.tlim);
        logged = log_live;
        if (upd) atomicMin((uint32_t*)&P.R.h[ray] + depth, __float_as_uint(s.tlim));
      } else if (s.prim > 0) {
        __hip_atomic_store(&P.Sx.occluded[ray], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      }
    }
    // visit log for the predictor, (ray, domain, t) of every finished
    // radiance visit: one atomic per wave and round
    if (log_live) {
      const unsigned long long lm = __ballot(logged);
      if (lm != 0ull) {
        const int leader = __builtin_ctzll(lm);
        uint32_t b = 0;
        if (lane == leader) b = atomicAdd(&c->vlog_count, (uint32_t)__popcll(lm));
        b = __builtin_amdgcn_readlane(b, leader);
        const uint32_t i = b + (uint32_t)__popcll(lm & ((1ull << lane) - 1ull));
        if (logged && i < P.vlog_cap) {
          P.vlog[(size_t)i * 3] = ray;
          P.vlog[(size_t)i * 3 + 1] = (uint32_t)dom;
          P.vlog[(size_t)i * 3 + 2] = __float_as_uint(s.prim >= 0 ? s.tlim : kInfF);
        }
      }
    }
    // shade records: slots come from the wave's chunk (one atomic per
    // kShadeChunk slots); the rest of a chunk that cannot hold this round's
    // records is marked empty
    const unsigned long long sm = __ballot(upd);
    if (sm != 0ull) {
      const uint32_t n = (uint32_t)__popcll(sm);
      if (sh_cur + n > sh_end) {
        shadeSlotsEmpty(P, sh_cur, sh_end, lane);
        const int leader = __builtin_ctzll(sm);
        uint32_t b = 0;
        const uint32_t take = max(n, sh_chunk);
        if (lane == leader) b = atomicAdd(&c->shade_count, take);
        sh_cur = __builtin_amdgcn_readlane(b, leader);
        sh_end = sh_cur + take;
        if (sh_end > P.sh_cap && lane == leader) atomicOr(&c->err, 1u);
      }
      const uint32_t idx = sh_cur + (uint32_t)__popcll(sm & ((1ull << lane) - 1ull));
      sh_cur += n;
      if (upd && idx < P.sh_cap) {
        // the whole ray goes into the shade record, written in order, so
        // shading reads no ray record at random; the history prefix
        // [0, depth) is fixed for the pass and entry `depth` is replaced
        // by the hit's t in the children
        float4* rec = P.sh_rec + (size_t)idx * kShadeRec;
        rec[0] = P.R.op[ray];
        rec[1] = P.R.ds[ray];
        // radiance rays keep only their depth in wd.w: the domain joins it
        rec[2] = make_float4(wd.x, wd.y, wd.z, __int_as_float(__float_as_int(wd.w) | (dom << 8)));
        rec[3] = hh;
        rec[4] = make_float4(s.tlim, s.u, s.v, __int_as_float(s.prim));
      }
    }
  }
  shadeSlotsEmpty(P, sh_cur, sh_end, lane);
  blockFlush<5>({&c->radiance_visits, &c->shadow_visits, &c->pruned, &c->node_visits,
                 &c->tri_tests},
                {rvis, svis, pruned, tc.nodes, waveSumULL(tc.tris)});
  if (st.overflowed()) atomicOr(&c->err, 4u);
}

// Shading after the trace step: every hit whose history entry is still the
// buffer's final entry for this pass is shaded once (stale speculative hits
// would only spawn rays that can never retire, so they are skipped).
__global__ void __launch_bounds__(kShadeBlock, 4) kShade(TraceParams P) {
  __shared__ uint32_t app_sm[2 * 34];
  DevCounters* c = P.ctr;
  if (!c->active) return;
  BlockAppend app{app_sm, 0u};
  const uint32_t n = min(c->shade_count, P.sh_cap), first = c->sh_mark;
  const int ray_depth = P.tile_rec[2];
  unsigned long long shades = 0;
  for (uint32_t base = first + blockIdx.x * blockDim.x; base < n; base += gridDim.x * blockDim.x) {
    uint32_t i = base + threadIdx.x;
    bool valid = i < n;
    float4 op = make_float4(0, 0, 0, 0), ds = make_float4(1, 0, 0, 0), wd = make_float4(0, 0, 0, 0);
    float4 hh = make_float4(kInfF, kInfF, kInfF, kInfF), tuv = make_float4(kInfF, 0, 0, 0);
    int dom = 0;
    if (valid) {
      const float4* rec = P.sh_rec + (size_t)i * kShadeRec;
      tuv = rec[4];
      valid = __float_as_int(tuv.w) >= 0;  // slots left empty by the trace kernel
    }
    if (valid) {
      const float4* rec = P.sh_rec + (size_t)i * kShadeRec;
      op = rec[0];
      ds = rec[1];
      wd = rec[2];
      hh = rec[3];
      dom = __float_as_int(wd.w) >> 8;
    }
    float hist[4] = {hh.x, hh.y, hh.z, hh.w};
    const int depth = __float_as_int(wd.w) & 0xff;
    const float t = tuv.x, u = tuv.y, v = tuv.z;
    const int prim = __float_as_int(tuv.w);
    const bool act = valid && vbufIsFinal(P.vbuf, __float_as_int(ds.w), hist, depth, t);
    shades += act ? 1 : 0;
    shadeSpawn(P, app, act, dom, op, ds, wd, hist, depth, t, u, v, prim, ray_depth);
  }
  blockFlush<1>({&c->shades}, {waveSumULL(shades)});
}

// After one domain (or chunk of domains) of a paged pool: its shade records
// and spawned shadow rays are done, and the work counters of the range and
// local-shadow launches start again at zero.
__global__ void kDomainDone(TraceParams P) {
  DevCounters* c = P.ctr;
  if (!c->active) return;
  c->sh_mark = min(c->shade_count, P.sh_cap);
  c->s_local = min(c->s_count, P.s_cap);
  P.wctr[P.ndomains] = 0;
  P.wctr[P.ndomains + 2] = 0;
  for (int k = 0; k < kWorkCounters; ++k) {
    P.wq[k * kWorkStride] = 0;
    P.wq[(2 * kWorkCounters + k) * kWorkStride] = 0;
  }
}

// Retire (ooc_tcontext.inl:123-135): shadow rays tested in this pass.
__global__ void kRetire(TraceParams P) {
  DevCounters* c = P.ctr;
  if (!c->active) return;
  const uint32_t b = c->s_beg, e = c->s_end;
  unsigned long long useful = 0;
  for (uint32_t i = b + blockIdx.x * blockDim.x + threadIdx.x; i < e; i += gridDim.x * blockDim.x) {
    float4 hh = P.S.h[i];
    float hist[4] = {hh.x, hh.y, hh.z, hh.w};
    float4 wd = P.S.wd[i];
    int depth = __float_as_int(wd.w) & 0xff;
    int samid = __float_as_int(P.S.ds[i].w);
    if (!vbufCorrect(P.vbuf, samid, hist, depth)) continue;
    ++useful;
    if (P.Sx.occluded[i]) continue;
    int pixid = __float_as_int(P.S.op[i].w);
    float* px = P.image + (size_t)pixid * 4;
    atomicAdd(px + 0, (float)(P.inv_spp * (double)wd.x));
    atomicAdd(px + 1, (float)(P.inv_spp * (double)wd.y));
    atomicAdd(px + 2, (float)(P.inv_spp * (double)wd.z));
  }
  blockFlush<1>({&c->useful}, {waveSumULL(useful)});
}

// Window end: census of non-stale spawned radiance rays and final hits.
__global__ void kWindowEnd(TraceParams P) {
  DevCounters* c = P.ctr;
  unsigned long long useful = 0, hits = 0;
  const uint32_t rend = min(c->r_count, P.r_cap);
  for (uint32_t i = c->r_init + blockIdx.x * blockDim.x + threadIdx.x; i < rend;
       i += gridDim.x * blockDim.x) {
    float4 hh = P.R.h[i];
    float hist[4] = {hh.x, hh.y, hh.z, hh.w};
    int depth = __float_as_int(P.R.wd[i].w) & 0xff;
    if (vbufCorrect(P.vbuf, __float_as_int(P.R.ds[i].w), hist, depth)) ++useful;
  }
  const long long S = (long long)P.tile_rec[1] * P.image_w * P.spp;
  for (long long s = blockIdx.x * (long long)blockDim.x + threadIdx.x; s < S;
       s += (long long)gridDim.x * blockDim.x) {
    const unsigned long long* w = P.vbuf + s * kHistory;
    float prev = 0.f;
    for (int d = 0; d < kHistory; ++d) {
      unsigned long long x = w[d];
      float t = wT(x);
      if (t == kInfF) break;
      if (d > 0 && wP(x) != __float_as_uint(prev)) break;
      ++hits;
      prev = t;
    }
  }
  blockFlush<2>({&c->useful, &c->final_hits}, {waveSumULL(useful), waveSumULL(hits)});
}
__global__ void kWindowAdvance(TraceParams P) {
  int* rec = (int*)P.tile_rec;
  rec[2] += kHistory;
}

// Visit log -> predictor samples: (org, dir, domain, label) where the label
// says whether this domain produced the ray's final closest hit.
__global__ void kLogFinalize(TraceParams P) {
  DevCounters* c = P.ctr;
  if (!P.log_on || !c->log_live) return;
  uint32_t n = min(c->vlog_count, P.vlog_cap);
  const uint32_t* vl = P.vlog;
  const int lane = __lane_id();
  // wave-uniform loop: one atomic per wave and step for the sample slots
  for (uint32_t base = blockIdx.x * blockDim.x; base < n; base += gridDim.x * blockDim.x) {
    const uint32_t i = base + threadIdx.x;
    const bool valid = i < n;
    uint32_t r = 0;
    int dom = 0;
    float label = 0.f;
    if (valid) {
      r = vl[(size_t)i * 3];
      dom = (int)vl[(size_t)i * 3 + 1];
      const float t = __uint_as_float(vl[(size_t)i * 3 + 2]);
      float4 wd = P.R.wd[r];
      int depth = __float_as_int(wd.w) & 0xff;
      float4 hh = P.R.h[r];
      float best = depth == 0 ? hh.x : depth == 1 ? hh.y : depth == 2 ? hh.z : hh.w;
      label = (t != kInfF && t == best) ? 1.f : 0.f;
    }
    const unsigned long long vm = __ballot(valid);
    if (vm == 0ull) continue;
    const int leader = __builtin_ctzll(vm);
    unsigned long long sb = 0;
    if (lane == leader) sb = atomicAdd(P.log_count, (unsigned long long)__popcll(vm));
    sb = ((unsigned long long)__builtin_amdgcn_readlane((uint32_t)(sb >> 32), leader) << 32) |
         (unsigned long long)__builtin_amdgcn_readlane((uint32_t)sb, leader);
    const unsigned long long slot = sb + (unsigned long long)__popcll(vm & ((1ull << lane) - 1ull));
    if (!valid || slot >= P.log_cap) continue;
    float4 op = P.R.op[r], ds = P.R.ds[r];
    float* o = P.log_buf + slot * 8;
    o[0] = op.x;
    o[1] = op.y;
    o[2] = op.z;
    o[3] = ds.x;
    o[4] = ds.y;
    o[5] = ds.z;
    o[6] = (float)dom;
    o[7] = label;
  }
}

// ---------------------------------------------------------------------------
// Host side
struct TracerState {
  long long S = 0;  // samples of the largest tile
  uint32_t r_cap = 0, s_cap = 0, p_cap = 0, q_cap[2] = {0, 0}, sh_cap = 0, mask_cap = 0;
  int ndomains = 0;
  DevBuf R[4], Pn[4], Sh[4], s_tmax, s_occ, vbuf, image, ctr, tile_rec, cnt, off, dbase, bcnt, bbase,
      wctr, wq, score, order, q[2], ask, ask_cnt, rkey, dmask, dfmt, dbox, ovf, vlog, log, sh_rec, qa,
      qb, dflag, acnt;
  unsigned long long log_cap = 0;
  bool logging = false;
  DevBuf log_ctr;
  int grid_persistent = 0;
  int num_cu = 0;
  int trav_res_block = 0, trav_res_blocks = 0;  // block size -> resident kTraverse blocks
  // captured tile graphs by key (schedule, windows, batch parameters,
  // logging, configuration version).  A graph whose buffers were reallocated
  // is retired, never launched again, and destroyed only when the tracer is
  // released, so no graph is destroyed while the context is tracing.
  struct KeyedGraph {
    int key[6];
    hipGraphExec_t exec;
  };
  std::vector<KeyedGraph> graphs;
  std::vector<hipGraphExec_t> retired;
  std::vector<uint32_t> h_cnt;
  std::vector<float> h_box;
  std::vector<int> h_order;
  unsigned long long paged_passes = 0;  // alternates the chunk direction of paged passes
  // schedule log of the plain form (test hook): where the frame is, and per
  // recorded pass {tile, window, pass, entries}, the scores and the order
  int cur_tile = 0, cur_window = 0, cur_pass = 0;
  int slog_max = 0;
  std::vector<int> slog_meta, slog_order;
  std::vector<unsigned long long> slog_scores;
  size_t heldBytes() const {
    size_t b = log_ctr.bytes;
    for (const DevBuf* x : {R, Pn, Sh})
      for (int k = 0; k < 4; ++k) b += x[k].bytes;
    for (const DevBuf* x : {&s_tmax, &s_occ, &vbuf, &image, &ctr, &tile_rec, &cnt, &off, &dbase, &bcnt,
                            &bbase, &wctr, &wq, &score, &order, &q[0], &q[1], &ask, &ask_cnt, &rkey,
                            &dmask, &dfmt, &dbox, &ovf, &vlog, &log, &sh_rec, &qa, &qb, &dflag, &acnt})
      b += x->bytes;
    return b;
  }
  // learned form: the predictor's stream beside the schedule and scatter
  // (forked from and joined back to the tracer's stream, graph capture included)
  hipStream_t aux = nullptr;
  hipEvent_t fork_ev = nullptr, join_ev = nullptr;
  // kernel timing of the trace launches (hipEvents on the tracer's stream)
  std::vector<hipEvent_t> ev;
  size_t ev_used = 0;
  double prof_ms = 0;
  long long prof_launches = 0;
  int evBegin(hipStream_t s) {
    if (ev_used + 2 > ev.size()) {
      for (int k = 0; k < 256; ++k) {
        hipEvent_t e;
        SPRAY_HIP(hipEventCreate(&e));
        ev.push_back(e);
      }
    }
    SPRAY_HIP(hipEventRecord(ev[ev_used], s));
    ev_used += 2;
    return (int)ev_used - 2;
  }
  void evEnd(int i, hipStream_t s) { SPRAY_HIP(hipEventRecord(ev[i + 1], s)); }
  void evCollect() {
    for (size_t i = 0; i + 1 < ev_used; i += 2) {
      float ms = 0;
      SPRAY_HIP(hipEventElapsedTime(&ms, ev[i], ev[i + 1]));
      prof_ms += ms;
      prof_launches++;
    }
    ev_used = 0;
  }
  hipGraphExec_t findGraph(const int key[6]) const {
    for (const KeyedGraph& g : graphs)
      if (std::memcmp(g.key, key, sizeof(g.key)) == 0) return g.exec;
    return nullptr;
  }
  void trimRetired() {  // a bound on what a long-lived context keeps (the oldest go first)
    while (retired.size() > 64) {
      (void)hipGraphExecDestroy(retired.front());
      retired.erase(retired.begin());
    }
  }
  void releaseGraph() {
    for (const KeyedGraph& g : graphs) retired.push_back(g.exec);
    graphs.clear();
    trimRetired();
  }
  void destroyGraphs() {
    releaseGraph();
    for (hipGraphExec_t e : retired) (void)hipGraphExecDestroy(e);
    retired.clear();
  }
};

void tracerRelease(Ctx* c) {
  if (!c->tracer) return;
  TracerState* T = c->tracer;
  T->destroyGraphs();
  for (hipEvent_t e : T->ev) (void)hipEventDestroy(e);
  if (T->aux) {
    (void)hipStreamSynchronize(T->aux);
    (void)hipStreamDestroy(T->aux);
  }
  for (hipEvent_t e : {T->fork_ev, T->join_ev})
    if (e) (void)hipEventDestroy(e);
  for (auto* arr : {T->R, T->Pn, T->Sh})
    for (int k = 0; k < 4; ++k) arr[k].release();
  for (DevBuf* b : {&T->s_tmax, &T->s_occ, &T->vbuf, &T->image, &T->ctr, &T->tile_rec, &T->cnt,
                    &T->off, &T->dbase, &T->bcnt, &T->bbase, &T->wctr, &T->wq, &T->score, &T->order, &T->q[0],
                    &T->q[1], &T->ask, &T->ask_cnt, &T->rkey, &T->dmask, &T->dfmt, &T->dbox, &T->ovf, &T->vlog, &T->log, &T->sh_rec, &T->log_ctr, &T->qa, &T->qb, &T->dflag, &T->acnt})
    b->release();
  delete T;
  c->tracer = nullptr;
}

static void tileList(const spray_config& cfg, std::vector<int>* ys, std::vector<int>* hs) {
  long long samples = (long long)cfg.image_w * cfg.image_h * cfg.pixel_samples;
  long long nt = (samples + cfg.max_samples_per_rank - 1) / cfg.max_samples_per_rank;
  int tile_h = cfg.image_h / (int)nt;
  ys->clear();
  hs->clear();
  for (int y = 0; y < cfg.image_h; y += tile_h) {
    ys->push_back(y);
    hs->push_back(std::min(tile_h, cfg.image_h - y));
  }
}

static int spawnPerVertex(const Ctx* c) {
  if (c->cfg.ao_mode) return c->cfg.ao_samples;
  int n = 0;
  for (auto& L : c->scene.lights) n += (L.type == kPointLight) ? 1 : c->cfg.ao_samples;
  return std::max(n, 1);
}

// Grows a buffer; a reallocation invalidates the captured graph.
static void need(TracerState* T, DevBuf& b, size_t bytes) {
  if (bytes <= b.bytes) return;
  T->releaseGraph();
  b.reserve(bytes);
}

// Blocking tiles traced together (batched forms): consecutive tiles share
// every pass, so each launch holds enough queue entries to keep the GPU busy
// while its longest traversals finish.  Each tile keeps its own sample
// numbering (RNG seeds), so the image does not depend on the batch size.
static int tilesPerBatch(const Ctx* c, const std::vector<int>& hs) {
  const long long per_tile = (long long)c->cfg.image_w * hs[0] * c->cfg.pixel_samples;
  const long long n = (long long)hs.size();
  if (c->tiles_in_flight > 0) return (int)std::min<long long>(c->tiles_in_flight, n);
  // as many tiles as fit the sample budget, spread evenly over the batches
  const long long budget = std::max(kBatchSamples, c->batch_samples);
  const long long kmax = std::max<long long>(1, std::min<long long>(n, budget / per_tile));
  const long long batches = (n + kmax - 1) / kmax;
  return (int)((n + batches - 1) / batches);
}

static size_t countShmem(int nd);

// Device bytes of tracer state per sample in flight: the capacities that
// ensureState sets at grow = 1, per sample.
static double stateBytesPerSample(const Ctx* c) {
  const spray_config& cfg = c->cfg;
  const int nd = c->pool.ndomains;
  const double spec = 2.0, depths = std::min(cfg.bounces, kHistory), spv = spawnPerVertex(c);
  const double per_ray = std::min(nd, 12), words = maskRowWords(nd);
  const double r = 1.0 + spec * depths, s = spv * spec * depths, p = spec;
  const double q0 = spec * per_ray, q1 = spv * spec * per_ray * 0.5, m = spec * (1.0 + spv);
  // radiance queue: entry + phase-list index; shadow queue: entry
  return r * 64 + s * 72 + p * 64 + q0 * (8 + 8) + std::min(q0, r) * kShadeRec * 16 + q1 * 8 +
         m * (8 * words + 1) + 32 + r;
}

static void ensureState(Ctx* c, double grow) {
  const spray_config& cfg = c->cfg;
  std::vector<int> ys, hs;
  tileList(cfg, &ys, &hs);
  {  // automatic batches: samples whose state takes at most 40% of the HBM the
     // tracer can use (free plus what it already holds), so that one grown
     // retry still fits
    size_t free_b = 0, total_b = 0;
    SPRAY_HIP(hipMemGetInfo(&free_b, &total_b));
    const double held = c->tracer ? (double)c->tracer->heldBytes() : 0.0;
    c->batch_samples = (long long)(0.40 * ((double)free_b + held) / stateBytesPerSample(c));
  }
  const int per_batch = cfg.schedule != SPRAY_SCHEDULE_PLAIN ? tilesPerBatch(c, hs) : 1;
  long long rows = 0;
  for (size_t b0 = 0; b0 < hs.size(); b0 += per_batch) {
    long long r = 0;
    for (size_t ti = b0; ti < std::min(hs.size(), b0 + (size_t)per_batch); ++ti) r += hs[ti];
    rows = std::max(rows, r);
  }
  long long S = (long long)cfg.image_w * rows * cfg.pixel_samples;
  if (S > 0x7fffffffLL) throw StatusError(SPRAY_ERR_CAPACITY, "too many samples in flight");
  if (!c->tracer) c->tracer = new TracerState();
  TracerState* T = c->tracer;
  int nd = c->pool.ndomains;
  if (nd > 1024) throw StatusError(SPRAY_ERR_CAPACITY, "the tracer supports at most 1024 domains");
  if (countShmem(nd) > 65536)  // many domains: kCount's LDS beyond the default limit
    for (const void* k : {(const void*)kCount<false, false>, (const void*)kCount<true, false>,
                          (const void*)kCount<false, true>, (const void*)kCount<true, true>})
      SPRAY_HIP(hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)countShmem(nd)));
  int depths = std::min(cfg.bounces, kHistory);
  double spec = 2.0 * grow;  // headroom for speculative re-shading
  auto cap = [](double v) { return (uint32_t)std::min<double>(3.9e9, v + 4096); };
  uint32_t r_cap = cap(S * (1.0 + spec * depths));
  uint32_t p_cap = cap(S * spec);
  uint32_t s_cap = cap((double)S * spawnPerVertex(c) * spec * depths);
  uint32_t per_ray = (uint32_t)std::min(nd, 12);  // domains a ray crosses, typical bound
  uint32_t q0 = cap((double)S * spec * per_ray);
  uint32_t q1 = cap((double)S * spawnPerVertex(c) * spec * per_ray * 0.5);
  // rays of one pass: at most one radiance ray and its shadow rays per
  // sample when each pass shades once, with the same speculative headroom
  uint32_t mask_cap = std::min(r_cap + (uint64_t)s_cap, (uint64_t)cap((double)S * (1.0 + spawnPerVertex(c)) * spec));
  if (T->S != S || T->ndomains != nd) T->releaseGraph();
  T->S = S;
  T->ndomains = nd;
  T->r_cap = std::max(T->r_cap, r_cap);
  T->s_cap = std::max(T->s_cap, s_cap);
  T->p_cap = std::max(T->p_cap, p_cap);
  T->q_cap[0] = std::max(T->q_cap[0], q0);
  T->q_cap[1] = std::max(T->q_cap[1], q1);
  T->mask_cap = std::max(T->mask_cap, mask_cap);
  for (int k = 0; k < 4; ++k) {
    need(T, T->R[k], (size_t)T->r_cap * 16);
    need(T, T->Pn[k], (size_t)T->p_cap * 16);
    need(T, T->Sh[k], (size_t)T->s_cap * 16);
  }
  need(T, T->s_tmax, (size_t)T->s_cap * 4);
  need(T, T->s_occ, (size_t)T->s_cap * 4);
  need(T, T->vbuf, (size_t)S * kHistory * 8);
  if (cfg.schedule == SPRAY_SCHEDULE_LEARNED) {
    need(T, T->ask, (size_t)kQueueGrid * askRegion(T->r_cap) * 48 + 64);
    need(T, T->ask_cnt, (size_t)kQueueGrid * 4 + 64);
    need(T, T->rkey, (size_t)T->r_cap * 4 + 64);
  }
  need(T, T->dmask, (size_t)T->mask_cap * maskRowWords(nd) * 8 + 64);
  need(T, T->dfmt, (size_t)T->mask_cap + 64);
  need(T, T->dbox, dboxBytes(nd));
  need(T, T->ctr, sizeof(DevCounters));
  need(T, T->tile_rec, 64);
  need(T, T->cnt, (size_t)2 * nd * 4 + 64);
  need(T, T->off, (size_t)2 * (nd + 1) * 4 + 64);
  need(T, T->dbase, (size_t)2 * nd * 4 + 64);
  need(T, T->bcnt, (size_t)2 * nd * kQueueGrid * 4 + 64);
  need(T, T->bbase, (size_t)2 * nd * kQueueGrid * 4 + 64);
  need(T, T->wctr, (size_t)(nd + 3) * 4 + 64);
  need(T, T->wq, (size_t)3 * kWorkCounters * kWorkStride * 4);
  need(T, T->qa, (size_t)T->q_cap[0] * 4);
  need(T, T->qb, (size_t)T->q_cap[0] * 4);
  if (cfg.schedule == SPRAY_SCHEDULE_LEARNED) {
    need(T, T->dflag, ((size_t)T->q_cap[0] / 64 + 2) * 8);
    need(T, T->acnt, (size_t)2 * kAssignWaves * 4 + 64);
  }
  // shade records: one per accepted hit of a pass, far fewer than queue
  // entries (an overflow grows and re-renders like any other pool)
  T->sh_cap = std::min(T->q_cap[0], T->r_cap);
  need(T, T->sh_rec, (size_t)T->sh_cap * kShadeRec * 16);
  need(T, T->score, (size_t)nd * 8 + 64);
  need(T, T->order, (size_t)nd * 4 + 64);
  need(T, T->q[0], (size_t)T->q_cap[0] * sizeof(QEntry));
  need(T, T->q[1], (size_t)T->q_cap[1] * sizeof(QEntry));
  T->h_cnt.resize(2 * nd);
  T->h_order.resize(nd);
  need(T, T->image, (size_t)cfg.image_w * cfg.image_h * 16);
  hipDeviceProp_t prop;
  SPRAY_HIP(hipGetDeviceProperties(&prop, c->device));
  // resident blocks: enough waves per CU to hide traversal latency
  int block = 64 * std::max(1, c->batch_tile);
  int per_cu = std::max(1, 1536 / block);
  T->grid_persistent = prop.multiProcessorCount * per_cu;
  T->num_cu = prop.multiProcessorCount;
}

static TraceParams makeParams(Ctx* c) {
  TracerState* T = c->tracer;
  const spray_config& cfg = c->cfg;
  TraceParams P;
  std::memset(&P, 0, sizeof(P));
  P.cam = c->cam;
  P.image_w = cfg.image_w;
  P.spp = cfg.pixel_samples;
  {
    std::vector<int> ys, hs;
    tileList(cfg, &ys, &hs);
    P.tile_samples = (long long)cfg.image_w * hs[0] * cfg.pixel_samples;
  }
  P.ao_mode = cfg.ao_mode;
  P.ao_samples = cfg.ao_samples;
  P.bounces = cfg.bounces;
  P.rr_start = cfg.rr_start;
  P.ao_max_dist = cfg.ao_max_dist;
  P.shininess = cfg.shininess;
  for (int k = 0; k < 3; ++k) P.ks[k] = cfg.ks[k];
  P.specular = (cfg.ks[0] != 0.f || cfg.ks[1] != 0.f || cfg.ks[2] != 0.f) ? 1 : 0;
  P.inv_spp = 1.0 / (double)cfg.pixel_samples;
  P.inv_spp_f = (float)P.inv_spp;
  P.ndomains = c->pool.ndomains;
  int nl = cfg.ao_mode ? 0 : (int)c->scene.lights.size();
  if (nl > kMaxLights) throw StatusError(SPRAY_ERR_CAPACITY, "more than 8 lights");
  P.nlights = nl;
  for (int l = 0; l < nl; ++l) {
    const HostLight& L = c->scene.lights[l];
    DevLight& D = P.lights[l];
    D.type = L.type;
    for (int k = 0; k < 3; ++k) {
      D.pos[k] = L.pos[k];
      D.e1[k] = L.e1[k];
      D.e2[k] = L.e2[k];
      D.radiance[k] = L.radiance[k];
      D.normal[k] = L.normal[k];
    }
    D.area = L.area;
  }
  const DomainPool& pool = c->pool;
  P.desc = pool.d_desc.as<DomainDesc>();
  P.dbox = c->tracer->dbox.as<float4>();
  P.nodes = pool.nodes.as<BNode4>();
  P.tris = pool.tris.as<LTri>();
  P.vtx = pool.vtx.as<float>();
  P.face = pool.face.as<uint32_t>();
  P.rgb = pool.rgb.as<uint32_t>();
  P.nrm = pool.nrm.as<float>();
  P.fshade = pool.fshade.as<float4>();
  auto pool4 = [](DevBuf* b) {
    RayPool r;
    r.op = b[0].as<float4>();
    r.ds = b[1].as<float4>();
    r.wd = b[2].as<float4>();
    r.h = b[3].as<float4>();
    return r;
  };
  P.R = pool4(T->R);
  P.Pn = pool4(T->Pn);
  P.S = pool4(T->Sh);
  P.Sx.tmax = T->s_tmax.as<float>();
  P.Sx.occluded = T->s_occ.as<uint32_t>();
  P.r_cap = T->r_cap;
  P.s_cap = T->s_cap;
  P.p_cap = T->p_cap;
  P.vbuf = T->vbuf.as<unsigned long long>();
  P.image = T->image.as<float>();
  P.ctr = T->ctr.as<DevCounters>();
  P.tile_rec = T->tile_rec.as<int>();
  int nd = P.ndomains;
  P.cnt = T->cnt.as<uint32_t>();
  P.off = T->off.as<uint32_t>();
  P.dbase = T->dbase.as<uint32_t>();
  P.bcnt = T->bcnt.as<uint32_t>();
  P.bbase = T->bbase.as<uint32_t>();
  P.wctr = T->wctr.as<uint32_t>();
  P.wq = T->wq.as<uint32_t>();
  P.score = T->score.as<unsigned long long>();
  P.order = T->order.as<int>();
  P.q[0] = T->q[0].as<QEntry>();
  P.q[1] = T->q[1].as<QEntry>();
  P.ask = T->ask.as<float4>();
  P.ask_cnt = T->ask_cnt.as<uint32_t>();
  P.ask_regions = kQueueGrid;
  P.ask_region = T->ask.bytes >= (size_t)kQueueGrid * askRegion(T->r_cap) * 48 ? askRegion(T->r_cap) : 0u;
  P.rkey = T->rkey.as<uint32_t>();
  P.dmask = T->dmask.as<unsigned long long>();
  P.dfmt = T->dfmt.as<uint8_t>();
  P.mask_cap = T->mask_cap;
  P.q_cap[0] = T->q_cap[0];
  P.q_cap[1] = T->q_cap[1];
  P.refill = std::min(64, std::max(1, c->batch_chunk));
  P.sh_rec = T->sh_rec.as<float4>();
  P.sh_cap = T->sh_cap;
  P.qa = T->qa.as<uint32_t>();
  P.qb = T->qb.as<uint32_t>();
  P.dflag = T->dflag.as<unsigned long long>();
  P.acnt = T->acnt.as<uint32_t>();
  P.learned = (cfg.schedule == SPRAY_SCHEDULE_LEARNED) ? 1 : 0;
  P.log_on = T->logging ? 1 : 0;
  P.vlog = T->vlog.as<uint32_t>();
  P.vlog_cap = T->logging ? (uint32_t)std::min<size_t>(T->vlog.bytes / 12, 0xffffffffu) : 0u;
  P.log_buf = T->log.as<float>();
  P.log_cap = T->logging ? T->log_cap : 0ull;
  P.log_count = T->log_ctr.as<unsigned long long>();
  (void)nd;
  return P;
}

static size_t countShmem(int nd) { return countLds(nd, kBlock).total; }

// Domain boxes for the ray/domain slab tests, plus kCount's domain grid:
// header, cell offsets and the cells' domain ids (ascending).
static void domainBoxesUpload(Ctx* c, hipStream_t s) {
  TracerState* T = c->tracer;
  const int nd = c->pool.ndomains;
  const auto& D = c->pool.desc;
  float glo[3], ghi[3], h[3], inv[3];
  int r[3];
  std::vector<uint32_t> off;
  std::vector<uint16_t> ids;
  // grid over the union of the domain boxes, resolution from the mean box
  // extent per axis (one cell per domain on a lattice of domains)
  float avg[3];
  for (int k = 0; k < 3; ++k) {
    glo[k] = kInfF;
    ghi[k] = -kInfF;
    double sum = 0;
    for (int d = 0; d < nd; ++d) {
      glo[k] = std::min(glo[k], D[d].lo[k]);
      ghi[k] = std::max(ghi[k], D[d].hi[k]);
      sum += (double)D[d].hi[k] - D[d].lo[k];
    }
    avg[k] = (float)(sum / nd);
  }
  for (int k = 0; k < 3; ++k) {
    const float ext = ghi[k] - glo[k];
    r[k] = (ext > 0.f && avg[k] > 0.f) ? (int)std::lround(ext / avg[k]) : 1;
    r[k] = std::max(1, std::min(kGridMaxRes, r[k]));
  }
  for (;;) {
    while ((long long)r[0] * r[1] * r[2] > kGridCells) *std::max_element(r, r + 3) /= 2;
    for (int k = 0; k < 3; ++k) {
      const float ext = ghi[k] - glo[k];
      h[k] = ext > 0.f ? ext / (float)r[k] : 1.f;
      inv[k] = 1.f / h[k];
    }
    const int ncell = r[0] * r[1] * r[2];
    std::vector<uint32_t> cnt(ncell + 1, 0);
    auto range = [&](int d, int k, int* c0, int* c1) {
      *c0 = std::max(0, std::min(r[k] - 1, (int)std::floor((D[d].lo[k] - glo[k]) * inv[k])));
      *c1 = std::max(0, std::min(r[k] - 1, (int)std::ceil((D[d].hi[k] - glo[k]) * inv[k]) - 1));
      *c1 = std::max(*c0, *c1);
    };
    long long total = 0;
    for (int d = 0; d < nd; ++d) {
      int a[3], b[3];
      for (int k = 0; k < 3; ++k) range(d, k, &a[k], &b[k]);
      total += (long long)(b[0] - a[0] + 1) * (b[1] - a[1] + 1) * (b[2] - a[2] + 1);
    }
    if (total > (long long)kGridIdsPerDomain * nd && r[0] * r[1] * r[2] > 1) {
      *std::max_element(r, r + 3) /= 2;  // domains span too many cells: coarser grid
      continue;
    }
    for (int d = 0; d < nd; ++d) {
      int a[3], b[3];
      for (int k = 0; k < 3; ++k) range(d, k, &a[k], &b[k]);
      for (int z = a[2]; z <= b[2]; ++z)
        for (int y = a[1]; y <= b[1]; ++y)
          for (int x = a[0]; x <= b[0]; ++x) ++cnt[(z * r[1] + y) * r[0] + x];
    }
    off.assign(ncell + 1, 0);
    for (int i = 0; i < ncell; ++i) off[i + 1] = off[i] + cnt[i];
    ids.assign(off[ncell], 0);
    std::vector<uint32_t> cur(off.begin(), off.end() - 1);
    for (int d = 0; d < nd; ++d) {  // ascending domain ids within a cell
      int a[3], b[3];
      for (int k = 0; k < 3; ++k) range(d, k, &a[k], &b[k]);
      for (int z = a[2]; z <= b[2]; ++z)
        for (int y = a[1]; y <= b[1]; ++y)
          for (int x = a[0]; x <= b[0]; ++x) ids[cur[(z * r[1] + y) * r[0] + x]++] = (uint16_t)d;
    }
    break;
  }
  const size_t bytes = gridIdsBytes(nd) + ids.size() * 2;
  T->h_box.assign((bytes + 3) / 4, 0.f);
  float* box = T->h_box.data();
  for (int d = 0; d < nd; ++d)
    for (int k = 0; k < 3; ++k) {
      box[(size_t)d * 8 + k] = D[d].lo[k];
      box[(size_t)d * 8 + 4 + k] = D[d].hi[k];
    }
  float* gh = (float*)((char*)box + gridHeadBytes(nd));
  for (int k = 0; k < 3; ++k) {
    gh[k] = glo[k];
    gh[4 + k] = ghi[k];
    gh[8 + k] = h[k];
    gh[12 + k] = inv[k];
  }
  std::memcpy(&gh[3], &r[0], 4);
  std::memcpy(&gh[7], &r[1], 4);
  std::memcpy(&gh[11], &r[2], 4);
  std::memcpy((char*)box + gridOffBytes(nd), off.data(), off.size() * 4);
  std::memcpy((char*)box + gridIdsBytes(nd), ids.data(), ids.size() * 2);
  SPRAY_HIP(hipMemcpyAsync(T->dbox.p, T->h_box.data(), bytes, hipMemcpyHostToDevice, s));
}

static size_t scatterShmem(int nd) { return (size_t)nd * (8 + 4) * 4; }  // boxes, slots, bases

static void ensureOverflow(TracerState* T, size_t lanes) { need(T, T->ovf, lanes * kOverflow * 4); }

static size_t travShmem(int block) { return (size_t)kRing * block * 4; }

// Blocks of kTraverse with `block` threads that the GPU holds at once.
static int travResidentBlocks(Ctx* c, int block) {
  TracerState* T = c->tracer;
  if (T->trav_res_block != block) {
    int nb = 0;
    SPRAY_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, kTraverse, block, travShmem(block)));
    T->trav_res_blocks = std::max(1, nb) * std::max(1, T->num_cu);
    T->trav_res_block = block;
  }
  return T->trav_res_blocks;
}

static void passQueues(Ctx* c, const TraceParams& P) {
  hipStream_t s = c->stream;
  TracerState* T = c->tracer;
  int nd = P.ndomains;
  kPassBegin<<<1, 1, 0, s>>>(P);
  kClearQ<<<(nd + 255) / 256, 256, 0, s>>>(P);
  const bool list = (nd + 63) / 64 > kHitListWords;  // packed hit lists (kScatter agrees)
  const size_t sh = countShmem(nd);
  if (P.learned && list)
    kCount<true, true><<<kQueueGrid, kBlock, sh, s>>>(P);
  else if (P.learned)
    kCount<true, false><<<kQueueGrid, kBlock, sh, s>>>(P);
  else if (list)
    kCount<false, true><<<kQueueGrid, kBlock, sh, s>>>(P);
  else
    kCount<false, false><<<kQueueGrid, kBlock, sh, s>>>(P);
  // learned form: the predictor needs only kCount's ask records, so it runs
  // on the second stream beside the schedule and the scatter
  if (P.learned) {
    SPRAY_HIP(hipEventRecord(T->fork_ev, s));
    SPRAY_HIP(hipStreamWaitEvent(T->aux, T->fork_ev, 0));
    predictorInfer(c, P, T->aux);
    SPRAY_HIP(hipEventRecord(T->join_ev, T->aux));
  }
  kSchedule<<<1, 1024, 0, s>>>(P);
  kBlockBases<<<2 * nd, kBlock, 0, s>>>(P);
  if (list)
    kScatter<true><<<kQueueGrid, kBlock, scatterShmem(nd), s>>>(P);
  else
    kScatter<false><<<kQueueGrid, kBlock, scatterShmem(nd), s>>>(P);
  if (P.learned) {
    SPRAY_HIP(hipStreamWaitEvent(s, T->join_ev, 0));
    predictorAssign(c, P);
  }
}

// Shading of the pass's accepted hits (those not yet shaded), then the new
// shadow rays' any-hit test in their spawning domain (procShads2).
static void passShade(Ctx* c, const TraceParams& P) {
  TracerState* T = c->tracer;
  hipStream_t s = c->stream;
  kShade<<<1024, kShadeBlock, 0, s>>>(P);
  TraceParams Q = P;
  Q.ovf = T->ovf.as<uint32_t>();
  int block = 64 * std::max(1, c->batch_tile);
  Q.trav_static = travResidentBlocks(c, block);
  int ei = c->profile ? T->evBegin(s) : -1;
  kTraverse<<<T->grid_persistent, block, travShmem(block), s>>>(Q, -2, 0);
  if (ei >= 0) T->evEnd(ei, s);
}

static void passRetire(Ctx* c, const TraceParams& P) {
  hipStream_t s = c->stream;
  kRetire<<<1024, 256, 0, s>>>(P);
  if (P.log_on) kLogFinalize<<<1024, 256, 0, s>>>(P);
}

static void passFinish(Ctx* c, const TraceParams& P) {
  passShade(c, P);
  passRetire(c, P);
}

// One pass of the plain schedule: the host reads the per-domain counts and
// launches one trace kernel per domain in schedule order (the reference's
// sequential domain sweep, ooc_pcontext.h:127-160).
static bool passPlain(Ctx* c, const TraceParams& P) {
  TracerState* T = c->tracer;
  hipStream_t s = c->stream;
  int nd = P.ndomains;
  passQueues(c, P);
  DevCounters h;
  SPRAY_HIP(hipMemcpyAsync(&h, P.ctr, sizeof(h), hipMemcpyDeviceToHost, s));
  SPRAY_HIP(hipMemcpyAsync(T->h_cnt.data(), P.cnt, 2 * nd * 4, hipMemcpyDeviceToHost, s));
  SPRAY_HIP(hipMemcpyAsync(T->h_order.data(), P.order, nd * 4, hipMemcpyDeviceToHost, s));
  SPRAY_HIP(hipStreamSynchronize(s));
  if (!h.active) return false;
  if ((int)T->slog_meta.size() / 4 < T->slog_max) {
    unsigned long long entries = 0;
    for (int k = 0; k < 2 * nd; ++k) entries += T->h_cnt[k];
    T->slog_meta.insert(T->slog_meta.end(), {T->cur_tile, T->cur_window, T->cur_pass, (int)entries});
    const size_t at = T->slog_scores.size();
    T->slog_scores.resize(at + nd);
    SPRAY_HIP(hipMemcpy(T->slog_scores.data() + at, P.score, (size_t)nd * 8, hipMemcpyDeviceToHost));
    T->slog_order.insert(T->slog_order.end(), T->h_order.begin(), T->h_order.begin() + nd);
  }
  hipDeviceProp_t prop;
  SPRAY_HIP(hipGetDeviceProperties(&prop, c->device));
  const int max_grid = prop.multiProcessorCount * (1536 / kTravBlock);
  ensureOverflow(T, (size_t)max_grid * kTravBlock);
  TraceParams Q = P;
  Q.ovf = T->ovf.as<uint32_t>();
  Q.trav_static = travResidentBlocks(c, kTravBlock);
  std::vector<int> doms;  // domains with entries, in schedule order
  for (int k = 0; k < nd; ++k) {
    int d = T->h_order[k];
    if (T->h_cnt[d] + T->h_cnt[nd + d] > 0) doms.push_back(d);
  }
  // with a paged pool every domain is loaded in turn (the next one ahead, on
  // the copy stream), and its hits are shaded and its new shadow rays tested
  // while it is resident
  const bool paged = c->pool.paged;
  for (size_t j = 0; j < doms.size(); ++j) {
    const int d = doms[j];
    poolAcquire(c, d, s);
    if (j + 1 < doms.size()) poolPrefetch(c, doms[j + 1], d);
    uint32_t n = T->h_cnt[d] + T->h_cnt[nd + d];
    int grid = (int)std::min<uint32_t>(max_grid, (n + kTravBlock - 1) / kTravBlock);
    int ei = c->profile ? T->evBegin(s) : -1;
    kTraverse<<<grid, kTravBlock, travShmem(kTravBlock), s>>>(Q, d, 0);
    if (ei >= 0) T->evEnd(ei, s);
    if (paged) {
      passShade(c, P);
      kDomainDone<<<1, 1, 0, s>>>(P);
    }
    poolRelease(c, d, s);
  }
  if (paged)
    passRetire(c, P);
  else
    passFinish(c, P);
  return true;
}

// Batched pass over a paged pool.  The queues are laid out in schedule
// order, so consecutive scheduled domains own contiguous entry ranges: the
// pass traces them in chunks of half the cache, one persistent launch per
// chunk, while the next chunk's domains load on the copy stream.  Each
// chunk's hits are shaded and its new shadow rays tested while it is
// resident.  Chunks run forward and backward in alternate passes, so the
// last chunk of a pass is still resident for the first of the next.
static bool passPagedBatched(Ctx* c, const TraceParams& P) {
  TracerState* T = c->tracer;
  DomainPool& pool = c->pool;
  hipStream_t s = c->stream;
  const int nd = P.ndomains;
  passQueues(c, P);
  DevCounters h;
  SPRAY_HIP(hipMemcpyAsync(&h, P.ctr, sizeof(h), hipMemcpyDeviceToHost, s));
  SPRAY_HIP(hipMemcpyAsync(T->h_cnt.data(), P.cnt, 2 * nd * 4, hipMemcpyDeviceToHost, s));
  SPRAY_HIP(hipMemcpyAsync(T->h_order.data(), P.order, nd * 4, hipMemcpyDeviceToHost, s));
  SPRAY_HIP(hipStreamSynchronize(s));
  if (!h.active) return false;
  // entry offsets by schedule rank, and chunks of ranks holding `per` domains with entries
  std::vector<uint32_t> off0(nd + 1, 0), off1(nd + 1, 0);
  for (int k = 0; k < nd; ++k) {
    off0[k + 1] = off0[k] + T->h_cnt[T->h_order[k]];
    off1[k + 1] = off1[k] + T->h_cnt[nd + T->h_order[k]];
  }
  const int per = std::max(1, pool.capacity / 2);
  std::vector<std::pair<int, int>> chunks;  // [k0, k1) in schedule rank
  for (int k = 0, k0 = 0, n = 0; k < nd; ++k) {
    const int d = T->h_order[k];
    n += (T->h_cnt[d] + T->h_cnt[nd + d] > 0) ? 1 : 0;
    if (n == per || k == nd - 1) {
      if (n > 0) chunks.push_back({k0, k + 1});
      k0 = k + 1;
      n = 0;
    }
  }
  if (T->paged_passes++ & 1) std::reverse(chunks.begin(), chunks.end());
  auto chunkDomains = [&](const std::pair<int, int>& ch) {
    std::vector<int> v;
    for (int k = ch.first; k < ch.second; ++k) {
      const int d = T->h_order[k];
      if (T->h_cnt[d] + T->h_cnt[nd + d] > 0) v.push_back(d);
    }
    return v;
  };
  const int block = 64 * std::max(1, c->batch_tile);
  TraceParams Q = P;
  Q.ovf = T->ovf.as<uint32_t>();
  Q.trav_static = travResidentBlocks(c, block);
  for (size_t j = 0; j < chunks.size(); ++j) {
    const std::vector<int> cur = chunkDomains(chunks[j]);
    for (int d : cur) {
      poolAcquire(c, d, s);
      pool.pinned[d] = 1;
    }
    if (j + 1 < chunks.size())
      for (int d : chunkDomains(chunks[j + 1])) poolPrefetch(c, d, -1);
    const int k0 = chunks[j].first, k1 = chunks[j].second;
    Q.rng[0] = off0[k0];
    Q.rng[1] = off0[k1] - off0[k0];
    Q.rng[2] = off1[k0];
    Q.rng[3] = off1[k1] - off1[k0];
    int ei = c->profile ? T->evBegin(s) : -1;
    kTraverse<<<T->grid_persistent, block, travShmem(block), s>>>(Q, -3, 0);
    if (ei >= 0) T->evEnd(ei, s);
    passShade(c, P);
    kDomainDone<<<1, 1, 0, s>>>(P);
    for (int d : cur) {
      poolRelease(c, d, s);
      pool.pinned[d] = 0;
    }
  }
  passRetire(c, P);
  return true;
}

// One pass of the batched schedule: every queue of the pass in one launch
// (with the learned form the deferred entries come last in it, or, in small
// passes, in a second launch); sizes stay on the device so the pass can be
// captured in a graph.
static void passBatched(Ctx* c, const TraceParams& P) {
  TracerState* T = c->tracer;
  hipStream_t s = c->stream;
  int block = 64 * std::max(1, c->batch_tile);
  passQueues(c, P);
  TraceParams Q = P;
  Q.ovf = T->ovf.as<uint32_t>();
  Q.trav_static = travResidentBlocks(c, block);
  int ei = c->profile ? T->evBegin(s) : -1;
  kTraverse<<<T->grid_persistent, block, travShmem(block), s>>>(Q, -1, 0);
  if (ei >= 0) T->evEnd(ei, s);
  if (P.learned) {  // the deferred list, when phase 0 left it (small passes)
    ei = c->profile ? T->evBegin(s) : -1;
    kTraverse<<<T->grid_persistent, block, travShmem(block), s>>>(Q, -1, 1);
    if (ei >= 0) T->evEnd(ei, s);
  }
  passFinish(c, P);
}

static void tileBody(Ctx* c, const TraceParams& P, bool batched, int windows, int* passes_run) {
  hipStream_t s = c->stream;
  const long long S = c->tracer->S;
  int eye_grid = (int)std::min<long long>(65535, (S + 255) / 256);
  kTileStart<<<1, 1, 0, s>>>(P);
  kGenEyes<<<eye_grid, 256, 0, s>>>(P);
  for (int w = 0; w < windows; ++w) {
    c->tracer->cur_window = w;
    if (w > 0) {
      kWindowStart<<<1, 1, 0, s>>>(P);
      kRestartPending<<<1024, 256, 0, s>>>(P);
      kRestartDone<<<1, 1, 0, s>>>(P);
    }
    kVbufReset<<<eye_grid, 256, 0, s>>>(P);
    for (int p = 0; p <= kHistory; ++p) {
      c->tracer->cur_pass = p;
      if (batched && c->pool.paged) {
        if (!passPagedBatched(c, P)) break;
      } else if (batched) {
        passBatched(c, P);
      } else {
        if (!passPlain(c, P)) break;
      }
      if (passes_run) ++*passes_run;
    }
    kWindowEnd<<<1024, 256, 0, s>>>(P);
    kWindowAdvance<<<1, 1, 0, s>>>(P);
  }
}

void tracerTraceFrame(Ctx* c, float* rgba_out, spray_stats* st) {
  const spray_config& cfg = c->cfg;
  if (c->pool.paged && cfg.schedule == SPRAY_SCHEDULE_LEARNED)
    throw StatusError(SPRAY_ERR_STATE,
                      "the learned schedule needs the whole scene resident (no partial domain cache)");
  std::vector<int> ys, hs;
  tileList(cfg, &ys, &hs);
  const unsigned long long hits0 = c->pool.hits, misses0 = c->pool.misses,
                           bytes0 = c->pool.bytes_in;
  double grow = c->pool_headroom;
  for (unsigned long long regrows = 0;; ++regrows) {
    ensureState(c, grow);
    TracerState* T = c->tracer;
    hipStream_t s = c->stream;
    bool batched = cfg.schedule != SPRAY_SCHEDULE_PLAIN;
    int block = 64 * std::max(1, c->batch_tile);
    ensureOverflow(T, std::max<size_t>((size_t)T->grid_persistent * block, 4096));
    if (cfg.schedule == SPRAY_SCHEDULE_LEARNED) {
      if (!predictorReady(c))
        throw StatusError(SPRAY_ERR_STATE, "the learned schedule needs a trained or loaded predictor");
      predictorPrepare(c);
      if (!T->aux) {
        SPRAY_HIP(hipStreamCreateWithFlags(&T->aux, hipStreamNonBlocking));
        SPRAY_HIP(hipEventCreateWithFlags(&T->fork_ev, hipEventDisableTiming));
        SPRAY_HIP(hipEventCreateWithFlags(&T->join_ev, hipEventDisableTiming));
      }
    }
    TraceParams P = makeParams(c);
    P.ovf = T->ovf.as<uint32_t>();
    domainBoxesUpload(c, s);
    SPRAY_HIP(hipMemsetAsync(P.image, 0, (size_t)cfg.image_w * cfg.image_h * 16, s));
    SPRAY_HIP(hipMemsetAsync(P.ctr, 0, sizeof(DevCounters), s));
    int windows = (cfg.bounces + kHistory - 1) / kHistory;
    hipEvent_t e0, e1;
    SPRAY_HIP(hipEventCreate(&e0));
    SPRAY_HIP(hipEventCreate(&e1));
    SPRAY_HIP(hipEventRecord(e0, s));
    int key[6] = {(int)batched, windows, (int)P.learned, c->batch_chunk, c->batch_tile,
                  P.log_on + 2 * c->param_version};
    // profiling times each launch instead; a paged pool is driven from the host
    const bool use_graph = batched && !c->profile && !c->pool.paged;
    hipGraphExec_t exec = use_graph ? T->findGraph(key) : nullptr;
    if (use_graph && exec == nullptr) {
      hipGraph_t g;
      SPRAY_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
      tileBody(c, P, true, windows, nullptr);
      SPRAY_HIP(hipStreamEndCapture(s, &g));
      SPRAY_HIP(hipGraphInstantiate(&exec, g, nullptr, nullptr, 0));
      SPRAY_HIP(hipGraphDestroy(g));
      // graphs of an older configuration version (remeshed pool, new camera)
      // are never launched again
      std::vector<TracerState::KeyedGraph> keep;
      for (const TracerState::KeyedGraph& g : T->graphs)
        if ((g.key[5] >> 1) == (key[5] >> 1))
          keep.push_back(g);
        else
          T->retired.push_back(g.exec);
      T->graphs.swap(keep);
      T->trimRetired();
      TracerState::KeyedGraph kg;
      std::memcpy(kg.key, key, sizeof(key));
      kg.exec = exec;
      T->graphs.push_back(kg);
    }
    const int per_batch = batched ? tilesPerBatch(c, hs) : 1;
    for (size_t b0 = 0; b0 < ys.size(); b0 += per_batch) {
      int rows = 0;
      for (size_t ti = b0; ti < std::min(ys.size(), b0 + (size_t)per_batch); ++ti) rows += hs[ti];
      int rec[4] = {ys[b0], rows, 0, 0};
      T->cur_tile = (int)b0;
      SPRAY_HIP(hipMemcpyAsync(T->tile_rec.p, rec, sizeof(rec), hipMemcpyHostToDevice, s));
      if (use_graph)
        SPRAY_HIP(hipGraphLaunch(exec, s));
      else
        tileBody(c, P, batched, windows, nullptr);
    }
    SPRAY_HIP(hipEventRecord(e1, s));
    DevCounters h;
    SPRAY_HIP(hipMemcpyAsync(&h, P.ctr, sizeof(h), hipMemcpyDeviceToHost, s));
    SPRAY_HIP(hipStreamSynchronize(s));
    float ms = 0;
    SPRAY_HIP(hipEventElapsedTime(&ms, e0, e1));
    T->evCollect();
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (h.err & 4u) throw StatusError(SPRAY_ERR_CAPACITY, "traversal stack overflow");
    if (h.err & 1u) {  // a pool or queue bound was hit: grow and render again
      if (grow >= 8.0) throw StatusError(SPRAY_ERR_CAPACITY, "ray pools exceeded after growing 8x");
      grow = std::min(8.0, 2.0 * grow);
      continue;
    }
    if (rgba_out)
      SPRAY_HIP(hipMemcpy(rgba_out, P.image, (size_t)cfg.image_w * cfg.image_h * 16,
                          hipMemcpyDeviceToHost));
    st->useful_rays = h.useful;
    st->eye_rays = (unsigned long long)cfg.image_w * cfg.image_h * cfg.pixel_samples;
    st->radiance_visits = h.radiance_visits;
    st->shadow_visits = h.shadow_visits;
    st->shades = h.shades;
    st->final_hits = h.final_hits;
    st->passes = h.passes;
    st->windows = (unsigned long long)windows * ys.size();
    st->tiles = ys.size();
    st->queue_entries = h.queue_entries;
    st->deferred_entries = h.deferred;
    st->pruned_entries = h.pruned;
    st->trace_ms = ms;
    st->node_visits = h.node_visits;
    st->tri_tests = h.tri_tests;
    st->cache_hits = c->pool.hits - hits0;
    st->cache_misses = c->pool.misses - misses0;
    st->cache_bytes = c->pool.bytes_in - bytes0;
    st->regrows = regrows;
    return;
  }
}

void tracerSetLogging(Ctx* c, bool on, long long max_samples) {
  ensureState(c, 1.0);
  TracerState* T = c->tracer;
  c->param_version++;
  if (!on) {
    T->logging = false;
    return;
  }
  if (max_samples <= 0) max_samples = 50000000;
  need(T, T->log, (size_t)max_samples * 32);
  need(T, T->vlog, (size_t)T->q_cap[0] * 12);
  need(T, T->log_ctr, 64);
  SPRAY_HIP(hipMemset(T->log_ctr.p, 0, 8));
  T->log_cap = (unsigned long long)max_samples;
  T->logging = true;
}

long long tracerLogCount(const Ctx* c) {
  const TracerState* T = c->tracer;
  if (!T || !T->log_ctr.p) return 0;
  unsigned long long n = 0;
  if (hipMemcpy(&n, T->log_ctr.p, 8, hipMemcpyDeviceToHost) != hipSuccess) return 0;
  return (long long)std::min<unsigned long long>(n, T->log_cap);
}

const float* tracerLogBuffer(Ctx* c) {
  if (!c->tracer || !c->tracer->log.p) throw StatusError(SPRAY_ERR_STATE, "no logged samples");
  return c->tracer->log.as<float>();
}

void tracerKernelTime(Ctx* c, long long* launches, double* ms, int reset) {
  TracerState* T = c->tracer;
  *launches = T ? T->prof_launches : 0;
  *ms = T ? T->prof_ms : 0.0;
  if (T && reset) {
    T->prof_launches = 0;
    T->prof_ms = 0;
  }
}

void tracerScheduleLog(Ctx* c, int max_records) {
  ensureState(c, 1.0);
  TracerState* T = c->tracer;
  T->slog_max = std::max(0, max_records);
  T->slog_meta.clear();
  T->slog_scores.clear();
  T->slog_order.clear();
}

int tracerScheduleRead(Ctx* c, int* meta, unsigned long long* scores, int* order) {
  const TracerState* T = c->tracer;
  if (!T) return 0;
  const int n = (int)T->slog_meta.size() / 4;
  if (meta) std::copy(T->slog_meta.begin(), T->slog_meta.end(), meta);
  if (scores) std::copy(T->slog_scores.begin(), T->slog_scores.end(), scores);
  if (order) std::copy(T->slog_order.begin(), T->slog_order.end(), order);
  return n;
}

// Autotuner of the batched forms: times every (queue chunk, waves per
// block) pair on a tuning frame and keeps the fastest.  The tuning frame is
// the loaded scene and camera at half the samples per pixel (at least one),
// so the pair is not chosen on the frames it then renders.  Results are kept
// in a text cache (SPRAY_AUTOTUNE_CACHE, else $HOME/.cache/spray_amd/
// autotune.txt), one line per key: the device (name, compute units), the
// frame settings (image, samples, bounces, AO, Russian roulette, schedule,
// tiles in flight, BVH builder) and a hash of the scene (every domain's
// triangle count and bounds).  A key found in the cache is used without
// tuning.
namespace {
uint64_t fnv(uint64_t h, const void* p, size_t n) {
  const unsigned char* b = (const unsigned char*)p;
  for (size_t i = 0; i < n; ++i) h = (h ^ b[i]) * 1099511628211ull;
  return h;
}

std::string autotuneKey(Ctx* c) {
  hipDeviceProp_t prop;
  SPRAY_HIP(hipGetDeviceProperties(&prop, c->device));
  uint64_t h = 1469598103934665603ull;
  for (const DomainDesc& D : c->pool.desc) {
    h = fnv(h, &D.num_tris, sizeof(D.num_tris));
    h = fnv(h, D.lo, sizeof(D.lo));
    h = fnv(h, D.hi, sizeof(D.hi));
    h = fnv(h, &D.mtl, sizeof(D.mtl));
  }
  const spray_config& g = c->cfg;
  char buf[512];
  std::snprintf(buf, sizeof(buf), "%s|cu%d|%dx%d|spp%d|b%d|ao%d,%d,%.9g|rr%d|s%d|tif%d|bvh%d|scene%016llx",
                prop.name, prop.multiProcessorCount, g.image_w, g.image_h, g.pixel_samples, g.bounces,
                g.ao_mode, g.ao_samples, (double)g.ao_max_dist, g.rr_start, g.schedule, c->tiles_in_flight,
                c->bvh_builder, (unsigned long long)h);
  std::string key(buf);
  for (char& ch : key)
    if (ch == ' ' || ch == '\t' || ch == '\n') ch = '_';
  return key;
}

std::string autotuneCachePath() {
  if (const char* p = std::getenv("SPRAY_AUTOTUNE_CACHE")) return p;
  const char* home = std::getenv("HOME");
  return std::string(home ? home : "/tmp") + "/.cache/spray_amd/autotune.txt";
}

bool autotuneLookup(const std::string& path, const std::string& key, int* chunk, int* tile) {
  std::ifstream in(path);
  std::string k;
  int ch = 0, tl = 0;
  double ms = 0;
  bool found = false;
  while (in >> k >> ch >> tl >> ms)
    if (k == key && ch >= 1 && ch <= 64 && tl >= 1 && tl <= 4) {
      *chunk = ch;  // the last line of a key wins
      *tile = tl;
      found = true;
    }
  return found;
}

void autotuneStore(const std::string& path, const std::string& key, int chunk, int tile, double ms) {
  const size_t slash = path.find_last_of('/');
  if (slash != std::string::npos && slash > 0) {
    std::string dir = path.substr(0, slash);
    for (size_t i = 1; i <= dir.size(); ++i)
      if (i == dir.size() || dir[i] == '/') (void)mkdir(dir.substr(0, i).c_str(), 0755);
  }
  std::ofstream out(path, std::ios::app);
  if (out) out << key << ' ' << chunk << ' ' << tile << ' ' << ms << '\n';  // a cache: failures are not fatal
}
}  // namespace

void tracerAutotune(Ctx* c, int* chunk_out, int* tile_out, int* cached_out) {
  const int saved_sched = c->cfg.schedule, saved_spp = c->cfg.pixel_samples;
  if (c->cfg.schedule == SPRAY_SCHEDULE_PLAIN) c->cfg.schedule = SPRAY_SCHEDULE_BATCHED;
  const std::string key = autotuneKey(c), path = autotuneCachePath();
  int bc = c->batch_chunk, bt = c->batch_tile;
  if (autotuneLookup(path, key, &bc, &bt)) {
    c->batch_chunk = bc;
    c->batch_tile = bt;
    c->cfg.schedule = saved_sched;
    *chunk_out = bc;
    *tile_out = bt;
    *cached_out = 1;
    return;
  }
  // queue chunk = lanes a wave refills at once; tile = waves per block
  const int chunks[] = {8, 16, 32, 64};
  const int tiles[] = {1, 2, 4};
  double best = 1e30;
  c->cfg.pixel_samples = std::max(1, saved_spp / 2);
  c->param_version++;
  try {
    for (int ch : chunks)
      for (int tl : tiles) {
        c->batch_chunk = ch;
        c->batch_tile = tl;
        spray_stats st;
        std::memset(&st, 0, sizeof(st));
        tracerTraceFrame(c, nullptr, &st);  // warm (graph capture)
        tracerTraceFrame(c, nullptr, &st);
        if (st.trace_ms < best) {
          best = st.trace_ms;
          bc = ch;
          bt = tl;
        }
      }
  } catch (...) {
    c->cfg.pixel_samples = saved_spp;
    c->cfg.schedule = saved_sched;
    c->param_version++;
    throw;
  }
  c->cfg.pixel_samples = saved_spp;
  c->cfg.schedule = saved_sched;
  c->param_version++;
  c->batch_chunk = bc;
  c->batch_tile = bt;
  autotuneStore(path, key, bc, bt, best);
  *chunk_out = bc;
  *tile_out = bt;
  *cached_out = 0;
}

}  // namespace spray
