// Paged domain pool: an LRU cache of domain blocks in HBM over per-domain
// images in pinned host memory.
//
// Replaces the reference's LruCache (render/lru_cache.cc:41-171) and
// TriMeshBuffer::load (render/trimesh_buffer.cc:117-169): the reference
// reads a missing domain's PLY file and builds its Embree scene; here every
// domain is built once at load time and its image (BVH nodes, leaf-ordered
// triangles, shading records) is kept in pinned host memory, so a miss is
// one host-to-HBM copy into the evicted block.  Copies run on the pool's
// copy stream; the trace stream waits for a block's copy, and a copy into a
// block waits for the last kernels that used the block.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "ctx.h"

namespace spray {

void DomainPool::releasePaging() {
  for (hipEvent_t e : block_free) (void)hipEventDestroy(e);
  for (hipEvent_t e : block_ready) (void)hipEventDestroy(e);
  block_free.clear();
  block_ready.clear();
  if (h_image) (void)hipHostFree(h_image);
  if (h_desc_pinned) (void)hipHostFree(h_desc_pinned);
  if (copy_stream) (void)hipStreamDestroy(copy_stream);
  h_image = nullptr;
  h_desc_pinned = nullptr;
  copy_stream = nullptr;
  h_image_off.clear();
  h_nrm.clear();
  block_domain.clear();
  domain_block.clear();
  block_tick.clear();
  pending.clear();
  pinned.clear();
  paged = false;
  capacity = 0;
  tick = hits = misses = bytes_in = 0;
}

// Bytes of domain d's image: nodes, leaf-ordered triangles, shading records.
static size_t imageBytes(const DomainDesc& D) {
  return (size_t)D.num_nodes * sizeof(BNode4) + (size_t)D.num_tris * (sizeof(LTri) + 48);
}

// Points domain d's descriptor at block b (host copy and device copy, the
// latter ordered on stream s).
static void bindBlock(DomainPool& P, int d, int b, hipStream_t s) {
  DomainDesc& D = P.desc[d];
  D.node_off = (long long)b * P.block_nodes;
  D.tri_off = (long long)b * P.block_tris;
  D.shade_off = (long long)b * P.block_tris;
  P.h_desc_pinned[d] = D;
  SPRAY_HIP(hipMemcpyAsync(P.d_desc.as<DomainDesc>() + d, P.h_desc_pinned + d, sizeof(DomainDesc),
                           hipMemcpyHostToDevice, s));
}

// LruCache::load (lru_cache.cc:65-171): a hit promotes the domain's block to
// most recently used; a miss takes the next free block or evicts the least
// recently used one (never `keep`'s block) and copies the image in.
// Returns the block.  A domain loaded ahead of its access (pending) is not
// counted again when the access comes.
static int lruLoad(Ctx* c, int d, int keep) {
  DomainPool& P = c->pool;
  if (P.pending[d]) {
    P.pending[d] = 0;
    return P.domain_block[d];
  }
  int b = P.domain_block[d];
  if (b >= 0) {
    ++P.hits;
    P.block_tick[b] = ++P.tick;
    return b;
  }
  ++P.misses;
  int victim = -1;
  for (int i = 0; i < P.capacity; ++i) {
    if (P.block_domain[i] < 0) {  // not full: blocks fill in order
      victim = i;
      break;
    }
    if ((keep >= 0 && P.block_domain[i] == keep) || P.pinned[P.block_domain[i]]) continue;
    if (victim < 0 || P.block_tick[i] < P.block_tick[victim]) victim = i;
  }
  if (victim < 0) throw StatusError(SPRAY_ERR_STATE, "no evictable cache block");
  if (P.block_domain[victim] >= 0) P.domain_block[P.block_domain[victim]] = -1;
  P.block_domain[victim] = d;
  P.domain_block[d] = victim;
  P.block_tick[victim] = ++P.tick;
  // the copy waits for the block's last users, then the descriptor follows
  const DomainDesc& D = P.desc[d];
  const char* src = (const char*)P.h_image + P.h_image_off[d];
  hipStream_t cs = P.copy_stream;
  SPRAY_HIP(hipStreamWaitEvent(cs, P.block_free[victim], 0));
  const size_t nb = (size_t)D.num_nodes * sizeof(BNode4), tb = (size_t)D.num_tris * sizeof(LTri);
  SPRAY_HIP(hipMemcpyAsync(P.nodes.as<BNode4>() + (size_t)victim * P.block_nodes, src, nb,
                           hipMemcpyHostToDevice, cs));
  SPRAY_HIP(hipMemcpyAsync(P.tris.as<LTri>() + (size_t)victim * P.block_tris, src + nb, tb,
                           hipMemcpyHostToDevice, cs));
  SPRAY_HIP(hipMemcpyAsync(P.fshade.as<float4>() + (size_t)victim * P.block_tris * 3, src + nb + tb,
                           (size_t)D.num_tris * 48, hipMemcpyHostToDevice, cs));
  bindBlock(P, d, victim, cs);
  SPRAY_HIP(hipEventRecord(P.block_ready[victim], cs));
  P.bytes_in += imageBytes(D);
  return victim;
}

void poolAcquire(Ctx* c, int d, hipStream_t s) {
  DomainPool& P = c->pool;
  if (!P.paged) return;
  const int b = lruLoad(c, d, -1);
  SPRAY_HIP(hipStreamWaitEvent(s, P.block_ready[b], 0));
}

void poolPrefetch(Ctx* c, int d, int keep) {
  DomainPool& P = c->pool;
  if (!P.paged || P.capacity < 2 || P.pending[d]) return;
  if (P.domain_block[d] >= 0) return;  // resident: its access is counted when it comes
  lruLoad(c, d, keep);
  P.pending[d] = 1;
}

void poolRelease(Ctx* c, int d, hipStream_t s) {
  DomainPool& P = c->pool;
  if (!P.paged) return;
  SPRAY_HIP(hipEventRecord(P.block_free[P.domain_block[d]], s));
}

// Sets up the cache after the build: `resident` lists the (domain, block)
// pairs the build left in HBM, least recently used first.
void poolInitPaging(Ctx* c, const std::vector<std::pair<int, int>>& resident) {
  DomainPool& P = c->pool;
  const int nd = P.ndomains;
  P.block_domain.assign(P.capacity, -1);
  P.domain_block.assign(nd, -1);
  P.block_tick.assign(P.capacity, 0);
  P.pending.assign(nd, 0);
  P.pinned.assign(nd, 0);
  P.block_free.resize(P.capacity);
  P.block_ready.resize(P.capacity);
  for (int b = 0; b < P.capacity; ++b) {
    SPRAY_HIP(hipEventCreateWithFlags(&P.block_free[b], hipEventDisableTiming));
    SPRAY_HIP(hipEventCreateWithFlags(&P.block_ready[b], hipEventDisableTiming));
  }
  for (size_t i = 0; i < resident.size(); ++i) {
    const int d = resident[i].first, b = resident[i].second;
    P.block_domain[b] = d;
    P.domain_block[d] = b;
    P.block_tick[b] = ++P.tick;
  }
  if (!P.h_desc_pinned)
    SPRAY_HIP(hipHostMalloc((void**)&P.h_desc_pinned, sizeof(DomainDesc) * nd, hipHostMallocDefault));
  std::copy(P.desc.begin(), P.desc.end(), P.h_desc_pinned);
  P.hits = P.misses = P.bytes_in = 0;
}

}  // namespace spray
