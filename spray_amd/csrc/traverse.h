// Closest-hit and any-hit traversal of one domain's 4-wide BVH.
//
// Per-lane short stack: a 16-entry ring in LDS.  Each wave owns kRing x 64
// words (slot s of lane l of wave w at lds[(w * kRing + s) * 64 + l], so a
// wave's pushes hit 64 consecutive banks and the address needs no block
// size);
// when the ring is full its bottom entry spills to a per-lane overflow
// stack in HBM, which is popped only once the ring is empty.  The overflow
// depth bounds the BVH depth; the build rejects deeper trees.
#pragma once

#include "dmath.h"
#include "spray_internal.h"

namespace spray {

constexpr int kRing = 16;
constexpr int kOverflow = 112;
constexpr uint32_t kStackEmpty = 0xffffffffu;
// robust slab test (Ize, "Robust BVH Ray Traversal", JCGT 2013): 1 + 2*gamma(3)
constexpr float kSlabInflate = 1.0000003576278687f;
// a ray visits each node of a tree at most once; the bound only stops a
// corrupt (cyclic) tree from spinning forever
constexpr uint32_t kMaxNodeVisits = 1u << 22;

// The stack state is one packed word per lane: ring head (bits 0-3), ring
// fill (4-8), overflow fill (9-15), overflow error (16).  The ring pointer
// carries the LDS address space, so ring accesses stay ds_read/ds_write and
// are never merged with the overflow path into flat accesses.
typedef __attribute__((address_space(3))) uint32_t lds_u32;

struct LaneStack {
  lds_u32* lds;        // this block's ring base
  uint32_t* ovf;       // overflow areas of the grid (kOverflow words per thread)
  uint32_t w;
  __device__ void init(uint32_t* lds_base, uint32_t* ovf_base) {
    lds = (lds_u32*)lds_base;
    ovf = ovf_base;
    w = 0;
  }
  __device__ void reset() { w &= 1u << 16; }
  __device__ bool overflowed() const { return (w >> 16) & 1u; }
  __device__ uint32_t count() const { return (w >> 4) & 31u; }
  __device__ uint32_t* lane_ovf() const {  // only on the rare spill path
    return ovf + (blockIdx.x * (size_t)blockDim.x + threadIdx.x) * kOverflow;
  }
  __device__ lds_u32& slot(uint32_t s) const {
    return lds[(((threadIdx.x >> 6) * kRing + (s & (kRing - 1))) << 6) + (threadIdx.x & 63)];
  }
  __device__ void push(uint32_t x) {
    uint32_t head = w & 15u, count = (w >> 4) & 31u, spilled = (w >> 9) & 127u, err = w >> 16;
    if (count == kRing) {
      uint32_t bottom = slot(head);
      if (spilled < kOverflow)
        lane_ovf()[spilled++] = bottom;
      else
        err = 1;  // reported to the host; never silently dropped
      head = (head + 1) & (kRing - 1);
      --count;
    }
    slot(head + count) = x;
    ++count;
    w = head | (count << 4) | (spilled << 9) | (err << 16);
  }
  // Pushes the first n (<= 3) of x0, x1, x2 (x0 deepest).  With room for
  // three entries all three slots above the top are written (the unused
  // ones hold nothing live), so the common case has no branch; a nearly
  // full ring takes the spilling path.
  __device__ void push3(uint32_t x0, uint32_t x1, uint32_t x2, uint32_t n) {
    uint32_t head = w & 15u, count = (w >> 4) & 31u;
    if (count + 3u <= (uint32_t)kRing) {
      slot(head + count) = x0;
      slot(head + count + 1u) = x1;
      slot(head + count + 2u) = x2;
      w += n << 4;
    } else {
      if (n >= 1u) push(x0);
      if (n >= 2u) push(x1);
      if (n >= 3u) push(x2);
    }
  }
  __device__ uint32_t pop() {
    uint32_t head = w & 15u, count = (w >> 4) & 31u, spilled = (w >> 9) & 127u;
    if (count > 0) {
      w -= 1u << 4;
      return slot(head + count - 1u);
    }
    if (spilled > 0) {
      --spilled;
      w = (w & ~(127u << 9)) | (spilled << 9);
      return lane_ovf()[spilled];
    }
    return kStackEmpty;
  }
};

// Work counters: node steps per wave (a sum of ballots, kept in scalar
// registers) and triangle tests per lane.
struct TravCount {
  unsigned long long nodes = 0;
  uint32_t tris = 0;
};

// Tests the 4 child boxes of a node against [tnear, tfar]; returns a 4-bit
// hit mask and the entry distances.  kSafe: a ray parallel to a slab with
// its origin on one of the slab's planes gives 0 * inf = NaN there; that
// axis then bounds nothing (only rays with an infinite 1/dir component
// can produce it, so waves without such rays use the unchecked form).
template <bool kSafe>
__device__ __forceinline__ int nodeBoxes(const BNode4* __restrict__ nd, const float org[3],
                                         const float inv[3], float tnear, float tfar,
                                         float tmin_o[4], int child_o[4]) {
  const float4* q = (const float4*)nd;
  float4 lox = q[0], hix = q[1], loy = q[2], hiy = q[3], loz = q[4], hiz = q[5];
  int4 ch = ((const int4*)nd)[6];
  float lx[4] = {lox.x, lox.y, lox.z, lox.w}, hx[4] = {hix.x, hix.y, hix.z, hix.w};
  float ly[4] = {loy.x, loy.y, loy.z, loy.w}, hy[4] = {hiy.x, hiy.y, hiy.z, hiy.w};
  float lz[4] = {loz.x, loz.y, loz.z, loz.w}, hz[4] = {hiz.x, hiz.y, hiz.z, hiz.w};
  child_o[0] = ch.x;
  child_o[1] = ch.y;
  child_o[2] = ch.z;
  child_o[3] = ch.w;
  int mask = 0;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    float ax = (lx[i] - org[0]) * inv[0], bx = (hx[i] - org[0]) * inv[0];
    float ay = (ly[i] - org[1]) * inv[1], by = (hy[i] - org[1]) * inv[1];
    float az = (lz[i] - org[2]) * inv[2], bz = (hz[i] - org[2]) * inv[2];
    float lx0 = fminf(ax, bx), lx1 = fmaxf(ax, bx);
    float ly0 = fminf(ay, by), ly1 = fmaxf(ay, by);
    float lz0 = fminf(az, bz), lz1 = fmaxf(az, bz);
    if (kSafe) {
      if (ax != ax || bx != bx) lx0 = -kInfF, lx1 = kInfF;
      if (ay != ay || by != by) ly0 = -kInfF, ly1 = kInfF;
      if (az != az || bz != bz) lz0 = -kInfF, lz1 = kInfF;
    }
    float t0 = fmaxf(fmaxf(lx0, ly0), fmaxf(lz0, tnear));
    float t1 = fminf(fminf(lx1, ly1), fminf(lz1, tfar));
    t1 = t1 * kSlabInflate;
    tmin_o[i] = t0;
    if (t0 <= t1 && lx[i] <= hx[i]) mask |= 1 << i;
  }
  return mask;
}

// Resumable traversal of one ray (closest or any hit) in two wave-wide
// phases (Aila & Laine 2009, "while-while" with speculative traversal):
// in the node phase every lane descends inner nodes; a lane that reaches a
// leaf postpones it (one slot, `pend`) and keeps descending while other
// lanes of its wave have not found a leaf yet.  In the leaf phase every lane
// tests its postponed leaf and then the leaves at the top of its stack.
// Node refs and leaf refs share the stack: a leaf ref is the child word of
// its parent (bit 31, count << 24, first triangle).
//
// Closest hit: `tlim` is the best t so far (initially the ray's tfar, which
// is therefore also the reported t of a miss); an exact tie in t goes to the
// lower primitive id.  Any hit: `tlim` stays tfar and prim becomes 1.
struct TravState {
  float org[3], inv[3];
  float Sx, Sy, Sz;
  uint32_t kk;          // kx | ky << 2 | kz << 4 | any << 8
  float tnear, tlim, u, v;
  int prim;
  uint32_t cur, pend, guard;
  uint32_t node_base, tri_base;  // the domain's first node / leaf-ordered triangle
};

__device__ __forceinline__ bool isInnerRef(uint32_t r) { return (int32_t)r >= 0; }
__device__ __forceinline__ bool isLeafRef(uint32_t r) { return (int32_t)r < 0 && r != kStackEmpty; }

__device__ __forceinline__ void travInit(TravState& s, uint32_t node_base, uint32_t tri_base,
                                         const float org[3], const float dir[3], float tnear,
                                         float tfar, bool any, LaneStack& st) {
  s.org[0] = org[0];
  s.org[1] = org[1];
  s.org[2] = org[2];
  s.inv[0] = 1.0f / dir[0];
  s.inv[1] = 1.0f / dir[1];
  s.inv[2] = 1.0f / dir[2];
  RayPre rp = rayPre(dir);
  s.Sx = rp.Sx;
  s.Sy = rp.Sy;
  s.Sz = rp.Sz;
  s.kk = (uint32_t)rp.kx | ((uint32_t)rp.ky << 2) | ((uint32_t)rp.kz << 4) | (any ? 256u : 0u);
  s.tnear = tnear;
  s.tlim = tfar;
  s.u = s.v = 0.f;
  s.prim = -1;
  s.cur = 0;
  s.pend = kStackEmpty;
  s.guard = 0;
  s.node_base = node_base;
  s.tri_base = tri_base;
  st.reset();
}

__device__ __forceinline__ bool travDone(const TravState& s) {
  return s.cur == kStackEmpty && s.pend == kStackEmpty;
}

__device__ __forceinline__ void cswap(float& ka, int& ia, float& kb, int& ib) {
  bool sw = kb < ka;
  float tk = sw ? kb : ka;
  int ti = sw ? ib : ia;
  kb = sw ? ka : kb;
  ib = sw ? ia : ib;
  ka = tk;
  ia = ti;
}

// One inner node: the hit children (inner nodes and leaves) in order of
// entry distance; the nearest becomes `cur`, the others are pushed.
template <bool kSafe>
__device__ __forceinline__ void travNode(TravState& s, LaneStack& st,
                                         const BNode4* __restrict__ nodes) {
  if (++s.guard > kMaxNodeVisits) {
    st.w |= 1u << 16;
    s.cur = s.pend = kStackEmpty;
    return;
  }
  float tmin[4];
  int child[4];
  const BNode4* nd = &nodes[s.node_base + s.cur];
  int mask = nodeBoxes<kSafe>(nd, s.org, s.inv, s.tnear, s.tlim, tmin, child);
  float ks[4];
  int m = 0;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    bool ok = (mask >> i) & 1;
    ks[i] = ok ? fminf(tmin[i], 3.0e38f) : kInfF;
    m += ok ? 1 : 0;
  }
  float k0 = ks[0], k1 = ks[1], k2 = ks[2], k3 = ks[3];
  int i0 = child[0], i1 = child[1], i2 = child[2], i3 = child[3];
  cswap(k0, i0, k1, i1);
  cswap(k2, i2, k3, i3);
  cswap(k0, i0, k2, i2);
  cswap(k1, i1, k3, i3);
  cswap(k1, i1, k2, i2);
  if (m == 0) {
    s.cur = st.pop();
    return;
  }
  // farther children below nearer ones: i3 (if hit) deepest, i1 on top
  uint32_t x0 = (uint32_t)(m > 3 ? i3 : (m > 2 ? i2 : i1));
  uint32_t x1 = (uint32_t)(m > 3 ? i2 : i1);
  st.push3(x0, x1, (uint32_t)i1, (uint32_t)(m - 1));
  s.cur = (uint32_t)i0;
}

// Tests one triangle (index relative to the domain's leaf-ordered array);
// returns true when an any-hit query is answered.
__device__ __forceinline__ bool travTri(TravState& s, uint32_t tri, const LTri* __restrict__ tris) {
  const RayPre rp{(int)(s.kk & 3u), (int)((s.kk >> 2) & 3u), (int)((s.kk >> 4) & 3u), s.Sx, s.Sy,
                  s.Sz};
  const bool any = (s.kk >> 8) & 1u;
  const float4* tq = (const float4*)&tris[s.tri_base + tri];
  float4 a = tq[0], b = tq[1], c = tq[2];
  float v0[3] = {a.x, a.y, a.z}, v1[3] = {b.x, b.y, b.z}, v2[3] = {c.x, c.y, c.z};
  // closest: accept t <= tlim once a hit exists (ties go to the lower id)
  float ub = (any || s.prim < 0) ? s.tlim : __uint_as_float(__float_as_uint(s.tlim) + 1u);
  float t, u, v;
  if (triHit(rp, s.org, v0, v1, v2, s.tnear, ub, &t, &u, &v)) {
    if (any) {
      s.prim = 1;
      return true;
    }
    int p = __float_as_int(a.w);
    if (s.prim < 0 || t < s.tlim || p < s.prim) {
      s.tlim = t;
      s.u = u;
      s.v = v;
      s.prim = p;
    }
  }
  return false;
}

// Node phase over the wave: lanes with an inner `cur` descend until every
// busy lane holds a postponed leaf, none can descend further, or fewer than
// half of the wave's rays would take a step.
template <bool kSafe>
__device__ __forceinline__ void travNodeLoop(TravState& s, LaneStack& st, TravCount& cnt,
                                             bool active, const BNode4* __restrict__ nodes) {
  while (true) {
    const bool step = active && isInnerRef(s.cur);
    cnt.nodes += (unsigned long long)__popcll(__ballot(step));
    if (step) {
      travNode<kSafe>(s, st, nodes);
      if (isLeafRef(s.cur) && s.pend == kStackEmpty) {
        s.pend = s.cur;
        s.cur = st.pop();
      }
    }
    bool more = active && isInnerRef(s.cur);
    bool hungry = active && s.pend == kStackEmpty && s.cur != kStackEmpty;
    // leave when no lane lacks a leaf, or when fewer than half of the
    // wave's rays would step: the rest go to the leaf phase instead of
    // idling behind a few deep lanes (they resume in the next round).  On
    // C2, C3 and C5 half was 1% faster than a third and than two thirds.
    const int nmore = __popcll(__ballot(more));
    if (nmore == 0 || __ballot(hungry) == 0ull || nmore * 2 < __popcll(__ballot(active))) break;
  }
}

// NaN-checked box tests only in waves with an axis-parallel ray: two copies
// of the loop, so that the common one carries no NaN handling (whose merges
// would also cost a canonicalisation before every min/max).
__device__ __forceinline__ void travNodePhase(TravState& s, LaneStack& st, TravCount& cnt,
                                              bool active, const BNode4* __restrict__ nodes) {
  if (__ballot(active && hasInfInv(s.inv)) != 0ull)
    travNodeLoop<true>(s, st, cnt, active, nodes);
  else
    travNodeLoop<false>(s, st, cnt, active, nodes);
}

// Leaf phase over the wave, one triangle per lane and step: each lane
// walks a leaf cursor (the leaf word with its count decremented and its
// first triangle advanced as triangles are tested), taking its postponed
// leaf first and then consecutive leaves from the stack, so lanes with
// short leaves move on instead of waiting for the longest leaf of the wave.
__device__ __forceinline__ void travLeafPhase(TravState& s, LaneStack& st, TravCount& cnt,
                                              bool active, const LTri* __restrict__ tris) {
  uint32_t lf = 0;  // leaf cursor (bit 31 | remaining << 24 | next triangle), 0 = none
  while (true) {
    if (active && lf == 0) {
      if (s.pend != kStackEmpty) {
        lf = s.pend;
        s.pend = kStackEmpty;
      } else if (isLeafRef(s.cur)) {
        lf = s.cur;
        s.cur = st.pop();
      }
      if (((lf >> 24) & 0x7fu) == 0u) lf = 0;
    }
    const bool tri = lf != 0;
    if (__ballot(tri) == 0ull) break;
    if (tri) {
      cnt.tris += 1;
      if (travTri(s, lf & 0x00ffffffu, tris)) {
        s.cur = s.pend = kStackEmpty;
        lf = 0;
      } else {
        lf += 1u - (1u << 24);
        if (((lf >> 24) & 0x7fu) == 0u) lf = 0;
      }
    }
  }
}

}  // namespace spray
