// Device-side state and helpers of the speculative tracer.
#pragma once

#include "dmath.h"
#include "kernels_common.h"
#include "spray_internal.h"
#include "traverse.h"

namespace spray {

constexpr int kMaxLights = 8;
constexpr int kTopDepths = kListSize;  // list positions that carry score weight
// a ray's hit domains as kTopDepths packed uint16 ids: the words of a mask of 256 domains
constexpr int kHitListWords = kTopDepths * 2 / 8;
static_assert(kTopDepths % 8 == 0 && kHitListWords == 4, "32-byte hit lists");
// words per hit-mask row: even above 256 domains, so that a row's packed list
// (two 16-byte stores) stays 16-byte aligned
__host__ __device__ inline int maskRowWords(int nd) {
  const int nw = (nd + 63) >> 6;
  return nw > kHitListWords ? (nw + 1) & ~1 : nw;
}

// Ray pools are SoA streams of float4 (16-byte loads per lane):
//   op = (org.xyz, pixid)  ds = (dir.xyz, samid)
//   wd = (w.xyz, depth | spawn_domain << 8)  h = history[4]
struct RayPool {
  float4 *op, *ds, *wd, *h;
};
struct ShadowExtra {
  float* tmax;
  uint32_t* occluded;
};

// Queue entry: ray index and the distance at which the ray enters the domain
struct QEntry {
  uint32_t ray;
  float tdom;
};

struct DevCounters {
  uint32_t r_count, s_count, p_count;  // pool fill levels (atomic append)
  uint32_t r_beg, r_end, s_beg, s_end;  // rays to enqueue / trace in this pass
  uint32_t r_init;                      // initial rays of the window (eyes / restarts)
  uint32_t active;                      // pass has work
  uint32_t err;                         // bit 0 pool overflow, 1 list overflow, 2 stack overflow
  uint32_t r_mark, s_mark;              // end of the last enqueued ranges
  uint32_t q_total[2];                  // entries in this pass's queues
  uint32_t vlog_count;                  // visit-log entries of this pass
  uint32_t log_live;                    // predictor logging on and its sample log not yet full
  uint32_t shade_count;                 // accepted hits of this pass (shade records)
  uint32_t sh_mark, s_local;            // shade records / spawned shadow rays already processed
                                        // (advanced per domain by the paged domain sweep)
  uint32_t n_phase[2];                  // learned form: radiance entries in phase A / B
  unsigned long long useful, radiance_visits, shadow_visits, shades, final_hits, queue_entries,
      pruned, deferred, passes, node_visits, tri_tests, log_count;
};

struct TraceParams {
  CameraDev cam;
  int image_w, spp, tile_y, tile_w, tile_h;  // tile_* come from the device tile record
  int tile_samples;  // samples of a full blocking tile (RNG seeds are tile-local)
  int ao_mode, ao_samples, bounces, rr_start, nlights;
  float ao_max_dist, shininess, ks[3];
  int specular;  // any ks component non-zero (the specular term is ks * pow(.))
  float inv_spp_f;
  double inv_spp;
  int ndomains;
  DevLight lights[kMaxLights];
  // pool
  const DomainDesc* desc;
  // per domain (lo, 0), (hi, 0); then the domain grid: header (4 float4),
  // cell offsets (uint32, at gridOffBytes) and cell domain ids (uint16, at
  // gridIdsBytes)
  const float4* dbox;
  const BNode4* nodes;
  const LTri* tris;
  const float* vtx;
  const uint32_t* face;
  const uint32_t* rgb;
  const float* nrm;
  const float4* fshade;  // per face: 3 x (vertex normal, vertex rgb bits)
  // tracer state
  RayPool R, Pn;  // radiance (window) and pending pools
  RayPool S;
  ShadowExtra Sx;
  uint32_t r_cap, s_cap, p_cap;
  unsigned long long* vbuf;  // 4 words per sample: t bits << 32 | parent bits
  float* image;              // rgba per pixel
  DevCounters* ctr;
  const int* tile_rec;       // [tile_y, tile_h, ray_depth(window offset)]
  // queues
  uint32_t* cnt;      // [2][nd] entries per domain (radiance, shadow)
  uint32_t* off;      // [2][nd + 1] queue offsets in schedule order
  uint32_t* dbase;    // [2][nd] queue offset of each domain (by id)
  uint32_t* bcnt;     // [2][nd][queue grid] entries per block (kCount)
  uint32_t* bbase;    // [2][nd][queue grid] queue base per block (kScatter)
  uint32_t* wctr;     // [nd + 3] work counters: per domain (plain), phase A / B, local shadows
  uint32_t* wq;       // [3][kWorkCounters] split counters of the persistent launches, kWorkStride apart
  unsigned long long* score;  // [nd]
  int* order;         // [nd] schedule
  QEntry* q[2];       // queues (radiance, shadow)
  // learned form: one ask record (3 float4) per radiance ray of the pass with
  // two or more candidate domains, written by kCount into its block's region
  // (block b's records at [b * ask_region, b * ask_region + ask_cnt[b])):
  // (origin, dir.x), (dir.y, dir.z, entry t of the nearest and of the second
  // candidate), (ray, nearest domain, second domain, candidates)
  float4* ask;
  uint32_t* ask_cnt;    // [ask_regions]
  uint32_t ask_regions;  // kCount's blocks
  uint32_t ask_region;   // records per block region (0: no ask buffer)
  // learned form, per radiance ray of the pass: the phase-A bound, the
  // ordered bits of its entry t (entries entering later are deferred), or
  // kNoBound
  uint32_t* rkey;
  uint32_t rng[4];    // range launches: radiance base, count, shadow base, count
  unsigned long long* dmask;  // [pass ray][maskRowWords(nd)]: domains hit (kCount -> kScatter)
  uint32_t mask_cap;          // pass rays the masks hold
  // more than 256 domains: a ray with at most kTopDepths hit domains keeps
  // their ids (uint16, 0xffff after the last) in the first 32 bytes of its
  // mask row instead of the mask, and dfmt[pass ray] = 1 says so
  uint8_t* dfmt;
  uint32_t q_cap[2];
  int refill;         // a wave refetches work once this many lanes are idle
  int trav_static;    // blocks of the trace kernel resident at once: their waves take a first range
  uint32_t* ovf;      // traversal stack overflow areas
  // shade records of the pass (accepted hits), SoA
  // shade records, kShadeRec float4 each: the ray's op, ds, wd (w: depth |
  // domain << 8) and history, then (t, u, v, prim); prim < 0 marks an
  // empty slot
  float4* sh_rec;
  uint32_t sh_cap;
  // learned form
  uint32_t* qa;       // learned form: radiance queue indices of phase A
  uint32_t* qb;       //   and of the deferred entries (traced last)
  unsigned long long* dflag;  // learned form: deferral bit of each radiance queue entry
  uint32_t* acnt;     // learned form: [2][kAssignWaves] span counts, then list bases
  int learned;
  // predictor logging: visit log of the pass (ray, domain, t) and samples
  int log_on;
  uint32_t* vlog;
  uint32_t vlog_cap;           // entries of 3 words
  float* log_buf;              // 8 floats per sample
  unsigned long long log_cap;  // samples
  unsigned long long* log_count;  // samples logged so far (kept across frames)
};

constexpr uint32_t kNoBound = 0xffffffffu;
constexpr int kAssignBlocks = 1024;  // blocks of the phase split (4 waves each)
constexpr int kAssignWaves = kAssignBlocks * 4;

struct Ctx;
// mlp.hip, learned form: the phase-A bound of every asking ray of the pass
// (after kCount; on stream s), then every radiance queue entry of the pass
// into phase A or the deferred list (after kScatter and the bounds)
void predictorInfer(Ctx* c, const TraceParams& P, hipStream_t s);
void predictorAssign(Ctx* c, const TraceParams& P);
bool predictorReady(Ctx* c);
void predictorPrepare(Ctx* c);

// ---- history buffer (VBuf, ooc/ooc_vbuf.cc:41-112) -------------------------
// Each of the 4 depth slots of a sample is one 64-bit word: the accepted t in
// the high half and the t of the parent slot in the low half.  A slot whose
// parent field differs from the t of the slot above it is stale and reads as
// infinity, which stands in for the reference's "reset deeper entries".
__device__ __forceinline__ float wT(unsigned long long w) { return __uint_as_float((uint32_t)(w >> 32)); }
__device__ __forceinline__ uint32_t wP(unsigned long long w) { return (uint32_t)w; }
__device__ __forceinline__ unsigned long long mkW(float t, uint32_t parent) {
  return ((unsigned long long)__float_as_uint(t) << 32) | parent;
}
// Relaxed device-scope loads read the L2 copy (no L1 invalidation).  Within
// one launch only stale-chain candidates can race with a prefix change, and
// their writes are invalid by construction, so no acquire ordering is needed;
// launches are ordered by the stream.
__device__ __forceinline__ unsigned long long ldW(const unsigned long long* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// VBuf::correct: history[0..n) equals the buffer's chain
__device__ __forceinline__ bool vbufCorrect(const unsigned long long* vb, int samid,
                                            const float h[4], int n) {
  const unsigned long long* w = vb + (size_t)samid * kHistory;
  // all words are read up front: independent loads, not a chain of
  // loads each waiting on the previous comparison
  unsigned long long x[kHistory];
#pragma unroll
  for (int i = 0; i < kHistory; ++i) x[i] = ldW(&w[i]);
  bool ok = true;
#pragma unroll
  for (int i = 0; i < kHistory; ++i)  // constant indices keep h[] in registers
    if (i < n)
      ok = ok && wT(x[i]) == h[i] && (i == 0 || wP(x[i]) == __float_as_uint(h[i > 0 ? i - 1 : 0]));
  return ok;
}

// Shading test: the ray's history prefix is the buffer's chain and the
// buffer's entry at `depth` is this hit (t with the parent's t).
__device__ __forceinline__ bool vbufIsFinal(const unsigned long long* vb, int samid,
                                            const float h[4], int depth, float t) {
  const unsigned long long* w = vb + (size_t)samid * kHistory;
  unsigned long long x[kHistory];
#pragma unroll
  for (int i = 0; i < kHistory; ++i) x[i] = ldW(&w[i]);
  bool ok = true;
#pragma unroll
  for (int i = 0; i < kHistory; ++i) {
    const uint32_t parent = i > 0 ? __float_as_uint(h[i > 0 ? i - 1 : 0]) : 0u;
    if (i < depth) ok = ok && wT(x[i]) == h[i] && (i == 0 || wP(x[i]) == parent);
    if (i == depth) ok = ok && x[i] == mkW(t, parent);
  }
  return ok;
}
__device__ __forceinline__ float pick4(const float h[4], int i) {
  return i == 0 ? h[0] : (i == 1 ? h[1] : (i == 2 ? h[2] : h[3]));
}

// VBuf::update for a ray at `depth` with hit distance t: true when the
// candidate (history[0..depth), t) became the buffer's lexicographic minimum.
__device__ __forceinline__ bool vbufUpdate(unsigned long long* vb, int samid, const float h[4],
                                           int depth, float t) {
  unsigned long long* w = vb + (size_t)samid * kHistory + depth;
  uint32_t parent = depth > 0 ? __float_as_uint(pick4(h, depth - 1)) : 0u;
  unsigned long long want = mkW(t, parent);
  unsigned long long old = ldW(w);
  while (true) {
    if (!vbufCorrect(vb, samid, h, depth)) return false;
    if (wP(old) == parent && !(t < wT(old))) return false;
    // on failure `old` receives the current value (lock-free: another
    // candidate succeeded)
    if (__hip_atomic_compare_exchange_strong(w, &old, want, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                             __HIP_MEMORY_SCOPE_AGENT))
      return true;
  }
}

}  // namespace spray
