// Closest-hit and any-hit traversal of one domain's 4-wide BVH.
//
// Per-lane short stack: a 16-entry ring in LDS (slot s of lane t at
// lds[s * blockDim.x + t], so a wave's pushes hit 64 consecutive banks);
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

struct LaneStack {
  uint32_t* lds;  // this block's ring base
  uint32_t* ovf;  // this lane's overflow area (kOverflow entries)
  int head, count, spilled;
  bool overflowed;
  __device__ void init(uint32_t* lds_base, uint32_t* ovf_lane) {
    lds = lds_base;
    ovf = ovf_lane;
    head = 0;
    count = 0;
    spilled = 0;
    overflowed = false;
  }
  __device__ void reset() {
    head = 0;
    count = 0;
    spilled = 0;
  }
  __device__ void push(uint32_t x) {
    if (count == kRing) {
      uint32_t bottom = lds[head * blockDim.x + threadIdx.x];
      if (spilled < kOverflow) {
        ovf[spilled++] = bottom;
      } else {
        overflowed = true;  // reported to the host; never silently dropped
      }
      head = (head + 1) & (kRing - 1);
      --count;
    }
    lds[((head + count) & (kRing - 1)) * blockDim.x + threadIdx.x] = x;
    ++count;
  }
  __device__ uint32_t pop() {
    if (count > 0) {
      --count;
      return lds[((head + count) & (kRing - 1)) * blockDim.x + threadIdx.x];
    }
    if (spilled > 0) return ovf[--spilled];
    return kStackEmpty;
  }
};

struct TravCount {
  uint32_t nodes = 0, tris = 0;
};

// Tests the 4 child boxes of a node against [tnear, tfar]; returns a 4-bit
// hit mask and the entry distances.
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
    // fminf/fmaxf drop a NaN operand (0 * inf on a slab plane): conservative
    float t0 = fmaxf(fmaxf(fminf(ax, bx), fminf(ay, by)), fmaxf(fminf(az, bz), tnear));
    float t1 = fminf(fminf(fmaxf(ax, bx), fmaxf(ay, by)), fminf(fmaxf(az, bz), tfar));
    t1 = t1 * kSlabInflate;
    tmin_o[i] = t0;
    if (t0 <= t1 && lx[i] <= hx[i]) mask |= 1 << i;
  }
  return mask;
}

// Closest hit in (tnear, tfar): on return prim >= 0 on a hit with t/u/v set;
// equal t resolves to the lower primitive id.
template <bool kCount>
__device__ __forceinline__ void closestHit(const BNode4* __restrict__ nodes,
                                           const LTri* __restrict__ tris, const float org[3],
                                           const float dir[3], float tnear, float tfar,
                                           uint32_t max_nodes, LaneStack& st, float* t_o,
                                           float* u_o, float* v_o, int* prim_o, TravCount* cnt) {
  RayPre rp = rayPre(dir);
  float inv[3] = {1.0f / dir[0], 1.0f / dir[1], 1.0f / dir[2]};
  float tbest = tfar, ub = 0.f, vb = 0.f;
  int pbest = -1;
  st.reset();
  uint32_t cur = 0;
  uint32_t guard = 0;
  while (cur != kStackEmpty) {
    // a ray visits each node of a tree at most once; more visits mean a corrupt tree
    if (++guard > max_nodes) {
      st.overflowed = true;
      break;
    }
    float tmin[4];
    int child[4];
    if (kCount) cnt->nodes++;
    int mask = nodeBoxes(&nodes[cur], org, inv, tnear, tbest, tmin, child);
    // leaves in slot order
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      if (!((mask >> i) & 1) || child[i] >= 0) continue;
      int first = child[i] & 0x00ffffff;
      int n = (child[i] >> 24) & 0x7f;
      for (int k = 0; k < n; ++k) {
        const float4* tq = (const float4*)&tris[first + k];
        float4 a = tq[0], b = tq[1], c = tq[2];
        float v0[3] = {a.x, a.y, a.z}, v1[3] = {b.x, b.y, b.z}, v2[3] = {c.x, c.y, c.z};
        int p = __float_as_int(a.w);
        float t, u, v;
        if (kCount) cnt->tris++;
        if (triHit(rp, org, v0, v1, v2, tnear, tfar, &t, &u, &v)) {
          if (pbest < 0 || t < tbest || (t == tbest && p < pbest)) {
            tbest = t;
            ub = u;
            vb = v;
            pbest = p;
          }
        }
      }
    }
    // inner children, nearest first
    int ids[4];
    float ts[4];
    int m = 0;
    float lim = tbest * kSlabInflate;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      if (((mask >> i) & 1) && child[i] >= 0 && tmin[i] <= lim) {
        ids[m] = child[i];
        ts[m] = tmin[i];
        ++m;
      }
    }
    // insertion sort by entry distance (m <= 4)
    for (int i = 1; i < m; ++i) {
      int id = ids[i];
      float t = ts[i];
      int j = i - 1;
      while (j >= 0 && ts[j] > t) {
        ids[j + 1] = ids[j];
        ts[j + 1] = ts[j];
        --j;
      }
      ids[j + 1] = id;
      ts[j + 1] = t;
    }
    if (m == 0) {
      cur = st.pop();
    } else {
      for (int i = m - 1; i >= 1; --i) st.push((uint32_t)ids[i]);
      cur = (uint32_t)ids[0];
    }
  }
  *t_o = pbest >= 0 ? tbest : tfar;
  *u_o = ub;
  *v_o = vb;
  *prim_o = pbest;
}

template <bool kCount>
__device__ __forceinline__ bool anyHit(const BNode4* __restrict__ nodes,
                                       const LTri* __restrict__ tris, const float org[3],
                                       const float dir[3], float tnear, float tfar,
                                       uint32_t max_nodes, LaneStack& st, TravCount* cnt) {
  RayPre rp = rayPre(dir);
  float inv[3] = {1.0f / dir[0], 1.0f / dir[1], 1.0f / dir[2]};
  st.reset();
  uint32_t cur = 0;
  uint32_t guard = 0;
  while (cur != kStackEmpty) {
    if (++guard > max_nodes) {
      st.overflowed = true;
      break;
    }
    float tmin[4];
    int child[4];
    if (kCount) cnt->nodes++;
    int mask = nodeBoxes(&nodes[cur], org, inv, tnear, tfar, tmin, child);
    for (int i = 0; i < 4; ++i) {
      if (!((mask >> i) & 1) || child[i] >= 0) continue;
      int first = child[i] & 0x00ffffff;
      int n = (child[i] >> 24) & 0x7f;
      for (int k = 0; k < n; ++k) {
        const float4* tq = (const float4*)&tris[first + k];
        float4 a = tq[0], b = tq[1], c = tq[2];
        float v0[3] = {a.x, a.y, a.z}, v1[3] = {b.x, b.y, b.z}, v2[3] = {c.x, c.y, c.z};
        float t, u, v;
        if (kCount) cnt->tris++;
        if (triHit(rp, org, v0, v1, v2, tnear, tfar, &t, &u, &v)) return true;
      }
    }
    uint32_t next = kStackEmpty;
    for (int i = 0; i < 4; ++i) {
      if (((mask >> i) & 1) && child[i] >= 0) {
        if (next == kStackEmpty)
          next = (uint32_t)child[i];
        else
          st.push((uint32_t)child[i]);
      }
    }
    cur = (next != kStackEmpty) ? next : st.pop();
  }
  return false;
}

}  // namespace spray

namespace spray {

// Resumable traversal of one ray (closest or any hit): one call of
// travStep() processes one node, so a wave can refill lanes whose ray has
// finished while the others keep traversing (Aila & Laine 2009,
// "persistent while-while" with per-lane replacement).
struct TravState {
  const BNode4* nodes;
  const LTri* tris;
  float org[3], inv[3];
  RayPre rp;
  float tnear, tfar, tlim;  // tlim: current best t (closest) or tfar (any)
  float u, v;
  int prim;
  uint32_t cur, guard, max_nodes;
  bool any;
};

__device__ __forceinline__ void travInit(TravState& s, const BNode4* nodes, const LTri* tris,
                                         uint32_t max_nodes, const float org[3],
                                         const float dir[3], float tnear, float tfar, bool any,
                                         LaneStack& st) {
  s.nodes = nodes;
  s.tris = tris;
  s.org[0] = org[0];
  s.org[1] = org[1];
  s.org[2] = org[2];
  s.inv[0] = 1.0f / dir[0];
  s.inv[1] = 1.0f / dir[1];
  s.inv[2] = 1.0f / dir[2];
  s.rp = rayPre(dir);
  s.tnear = tnear;
  s.tfar = tfar;
  s.tlim = tfar;
  s.u = s.v = 0.f;
  s.prim = -1;
  s.cur = 0;
  s.guard = 0;
  s.max_nodes = max_nodes;
  s.any = any;
  st.reset();
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

// Processes one node; returns true when the ray is finished (stack empty,
// any-hit found, or the visit bound exceeded -> st.overflowed).
__device__ __forceinline__ bool travStep(TravState& s, LaneStack& st, TravCount& cnt) {
  if (++s.guard > s.max_nodes) {
    st.overflowed = true;
    return true;
  }
  float tmin[4];
  int child[4];
  cnt.nodes++;
  int mask = nodeBoxes(&s.nodes[s.cur], s.org, s.inv, s.tnear, s.tlim, tmin, child);
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    if (!((mask >> i) & 1) || child[i] >= 0) continue;
    int first = child[i] & 0x00ffffff;
    int n = (child[i] >> 24) & 0x7f;
    for (int k = 0; k < n; ++k) {
      const float4* tq = (const float4*)&s.tris[first + k];
      float4 a = tq[0], b = tq[1], c = tq[2];
      float v0[3] = {a.x, a.y, a.z}, v1[3] = {b.x, b.y, b.z}, v2[3] = {c.x, c.y, c.z};
      float t, u, v;
      cnt.tris++;
      if (triHit(s.rp, s.org, v0, v1, v2, s.tnear, s.any ? s.tlim : s.tfar, &t, &u, &v)) {
        if (s.any) {
          s.prim = 1;
          return true;
        }
        int p = __float_as_int(a.w);
        if (s.prim < 0 || t < s.tlim || (t == s.tlim && p < s.prim)) {
          s.tlim = t;
          s.u = u;
          s.v = v;
          s.prim = p;
        }
      }
    }
  }
  // inner children, nearest first (sorting network on 4 keys; no local arrays)
  const float lim = s.tlim * kSlabInflate;
  float k0 = 3.4e38f, k1 = 3.4e38f, k2 = 3.4e38f, k3 = 3.4e38f;
  float ks[4];
  int m = 0;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    bool ok = ((mask >> i) & 1) && child[i] >= 0 && tmin[i] <= lim;
    ks[i] = ok ? fminf(tmin[i], 3.0e38f) : kInfF;
    m += ok ? 1 : 0;
  }
  k0 = ks[0];
  k1 = ks[1];
  k2 = ks[2];
  k3 = ks[3];
  int i0 = child[0], i1 = child[1], i2 = child[2], i3 = child[3];
  cswap(k0, i0, k1, i1);
  cswap(k2, i2, k3, i3);
  cswap(k0, i0, k2, i2);
  cswap(k1, i1, k3, i3);
  cswap(k1, i1, k2, i2);
  if (m == 0) {
    s.cur = st.pop();
    return s.cur == kStackEmpty;
  }
  if (m > 3) st.push((uint32_t)i3);
  if (m > 2) st.push((uint32_t)i2);
  if (m > 1) st.push((uint32_t)i1);
  s.cur = (uint32_t)i0;
  return false;
}

}  // namespace spray
