// spray_ctx: everything one rank's tracer owns on the device.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdlib>
#include <string>
#include <vector>

#include "dmath.h"
#include "spray_internal.h"

struct spray_ctx;

namespace spray {

#define SPRAY_HIP(call)                                                              \
  do {                                                                               \
    hipError_t e_ = (call);                                                          \
    if (e_ != hipSuccess) {                                                          \
      throw DeviceError(std::string(#call) + ": " + hipGetErrorString(e_));          \
    }                                                                                \
  } while (0)

struct DeviceError {
  std::string msg;
  explicit DeviceError(std::string m) : msg(std::move(m)) {}
};
struct StatusError {
  int code;
  std::string msg;
  StatusError(int c, std::string m) : code(c), msg(std::move(m)) {}
};

// A growable device buffer (grows only; contents not preserved on growth).
struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  void reserve(size_t b) {
    if (b <= bytes) return;
    if (p) SPRAY_HIP(hipFree(p));
    p = nullptr;
    bytes = 0;
    if (hipMalloc(&p, b) != hipSuccess) {
      p = nullptr;
      (void)hipGetLastError();
      throw StatusError(SPRAY_ERR_OOM, "device allocation of " + std::to_string(b) + " bytes failed");
    }
    bytes = b;
  }
  template <typename T>
  T* as() const {
    return (T*)p;
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
  }
};

// HBM-resident domain pool (replaces the domain cache + Embree scenes)
struct DomainPool {
  int ndomains = 0;
  long long total_tris = 0, total_verts = 0, total_nodes = 0;
  std::vector<DomainDesc> desc;  // host copy
  std::vector<long long> nverts; // vertices per domain as loaded
  DevBuf d_desc;                 // DomainDesc[ndomains]
  DevBuf vtx, face, rgb, nrm;    // float3 / uint3 / uint32 / float3
  DevBuf nodes;                  // BNode4
  DevBuf tris;                   // LTri (leaf order)
  DevBuf fshade;                 // float4[3] per face: vertex normal xyz + vertex rgb bits
  std::vector<int> bvh_depth;
  double build_ms = 0;
  // LRU paging (lru_cache.cc:41-171): with a cache smaller than the scene,
  // HBM holds `capacity` blocks of block_nodes nodes and block_tris
  // triangles + shading records, and every domain's image (nodes, leaf-ordered
  // triangles, shading records) stays in pinned host memory.
  bool paged = false;
  int capacity = 0;
  long long block_nodes = 0, block_tris = 0;
  void* h_image = nullptr;                // pinned: per domain nodes, tris, shading records
  std::vector<long long> h_image_off;     // byte offset of each domain's image
  std::vector<std::vector<float>> h_nrm;  // per-domain vertex normals (query API)
  DomainDesc* h_desc_pinned = nullptr;    // staging for descriptor uploads
  std::vector<int> block_domain, domain_block;
  std::vector<unsigned long long> block_tick;
  std::vector<hipEvent_t> block_free, block_ready;
  std::vector<char> pending;              // prefetched, access not yet made
  std::vector<char> pinned;               // in use by the chunk being traced: never evicted
  unsigned long long tick = 0, hits = 0, misses = 0, bytes_in = 0;
  hipStream_t copy_stream = nullptr;
  void releasePaging();
  void release() {
    d_desc.release();
    vtx.release();
    face.release();
    rgb.release();
    nrm.release();
    nodes.release();
    tris.release();
    fshade.release();
    releasePaging();
    desc.clear();
    ndomains = 0;
  }
};

// Where buildPool takes each domain's raw mesh from: host meshes, PLY
// bodies decoded on the GPU, or the GPU marching-cubes generator.
struct Ctx;
struct MeshSource {
  virtual ~MeshSource() = default;
  virtual long long numVerts(int d) const = 0;
  virtual long long numFaces(int d) const = 0;
  // Issues, on stream s, the writes of domain d's vertices (3 floats each),
  // faces (3 indices each, every index < numVerts(d)) and packed colours.
  virtual void fill(Ctx* c, int d, float* v, uint32_t* f, uint32_t* rgb, hipStream_t s) = 0;
  // After the fills issued so far: throws if one of them found bad input.
  virtual void check(Ctx* c) { (void)c; }
  // fill() of different domains may run concurrently on several streams,
  // with no host wait inside.
  virtual bool pipelined() const { return false; }
  // The scene is rebuilt from this kind of source again and again
  // (remeshing): the build keeps its scratch for the next build.
  virtual bool rebuilds() const { return false; }
};

// Pipelines that side-by-side work should use: at most `want`, and no more
// than the hardware queues HIP gives a process (GPU_MAX_HW_QUEUES, 4 by
// default).  Pipeline 0 runs on the context's stream, the others on a
// StreamSet of pipelines - 1 streams.  A stream beyond the queues shares a
// queue with another and runs after it: four pipelines on side streams
// alone left the context's queue idle and put two pipelines on one queue,
// which set the length of the whole build.
inline int pipelineCount(int want) {
  int hw = 4;
  if (const char* e = std::getenv("GPU_MAX_HW_QUEUES")) {
    const int v = std::atoi(e);
    if (v > 0) hw = v;
  }
  return std::max(1, std::min(want, hw));
}

// Non-blocking streams that start after the work issued so far on `after`;
// join() makes `after` wait for everything issued on them, and the
// destructor waits for them and destroys them.
struct StreamSet {
  std::vector<hipStream_t> st;
  std::vector<hipEvent_t> ev;
  StreamSet(int k, hipStream_t after) {
    if (k <= 0) return;
    st.assign(k, nullptr);
    ev.assign(k, nullptr);
    try {
      for (int i = 0; i < k; ++i) {
        SPRAY_HIP(hipStreamCreateWithFlags(&st[i], hipStreamNonBlocking));
        SPRAY_HIP(hipEventCreateWithFlags(&ev[i], hipEventDisableTiming));
      }
      SPRAY_HIP(hipEventRecord(ev[0], after));
      for (int i = 0; i < k; ++i) SPRAY_HIP(hipStreamWaitEvent(st[i], ev[0], 0));
    } catch (...) {
      release();
      throw;
    }
  }
  StreamSet(const StreamSet&) = delete;
  StreamSet& operator=(const StreamSet&) = delete;
  int size() const { return (int)st.size(); }
  hipStream_t operator[](int i) const { return st[i]; }
  void join(hipStream_t after) {
    for (size_t i = 0; i < st.size(); ++i) {
      SPRAY_HIP(hipEventRecord(ev[i], st[i]));
      SPRAY_HIP(hipStreamWaitEvent(after, ev[i], 0));
    }
  }
  void release() {
    for (hipStream_t x : st)
      if (x) {
        (void)hipStreamSynchronize(x);
        (void)hipStreamDestroy(x);
      }
    for (hipEvent_t e : ev)
      if (e) (void)hipEventDestroy(e);
    st.clear();
    ev.clear();
  }
  ~StreamSet() { release(); }
};

struct HostMeshSource : MeshSource {
  std::vector<HostMesh>& m;
  explicit HostMeshSource(std::vector<HostMesh>& meshes) : m(meshes) {}
  long long numVerts(int d) const override { return (long long)m[d].v.size() / 3; }
  long long numFaces(int d) const override { return (long long)m[d].f.size() / 3; }
  void fill(Ctx* c, int d, float* v, uint32_t* f, uint32_t* rgb, hipStream_t s) override;
};

struct TracerState;  // tracer.hip
struct Predictor;    // mlp.hip

// Smallest sample budget of a batch of blocking tiles (the budget grows with
// the HBM left after the scene is loaded).
constexpr long long kBatchSamples = 8ll << 20;

struct BuildCache;  // bvh.hip: build scratch kept between remeshes

struct Ctx {
  spray_config cfg;
  int device = 0;
  hipStream_t stream = nullptr;
  HostScene scene;
  bool loaded = false;
  DomainPool pool;
  CameraDev cam;
  bool camera_set = false;
  TracerState* tracer = nullptr;
  Predictor* pred = nullptr;
  int batch_chunk = 16;  // refill threshold of the trace kernel (idle lanes per wave)
  int tiles_in_flight = 0;  // blocking tiles traced together (batched forms); 0 = automatic
  long long batch_samples = 0;  // automatic batches: samples in flight (set from free HBM)
  int bvh_builder = 0;     // 0 = Karras LBVH, 1 = PLOC (next scene load)
  int cache_size = -1;     // domains resident in HBM (next scene load); < 0 or >= ndomains: all
  int batch_tile = 4;      // waves per block of the persistent kernel (block = 64 * tile)
  int param_version = 0;   // bumped when camera / config change (graph re-capture)
  bool profile = false;    // time every trace-kernel launch with hipEvents
  double pool_headroom = 1.0;  // factor on the initial ray-pool estimate (test hook)
  std::string err;
  // scratch for spray_trace_rays / spray_domain_list
  DevBuf q_in, q_out;
  // a scene made by the GPU generator (spray_scenegen_load) can be remeshed;
  // the generator's per-domain buffers stay allocated between remeshes
  bool generated = false;
  DevBuf gen_store;
  BuildCache* build_cache = nullptr;
  spray_scenegen_params gen{};
  std::string gen_name, gen_extra;
};

// bvh.hip: builds the pool from `src` (the domains of c->scene); with
// keep_buffers the pool's device buffers are reused where large enough
void buildPool(Ctx* c, MeshSource& src, bool keep_buffers = false);
void buildCacheRelease(Ctx* c);
// mesh_gpu.hip: scene loading from PLY files (binary bodies decoded on the
// GPU), the GPU marching-cubes generator, mesh readback and the scene writer
void loadScenePly(Ctx* c, HostScene& scene);
void generateScene(Ctx* c, const spray_scenegen_params& p, const std::string& name,
                   const std::string& extra, bool remesh);
void domainMesh(Ctx* c, int d, float* v, uint32_t* f, uint32_t* rgb, long long* nv_out,
                long long* nf_out);
void writeScene(Ctx* c, const std::string& dir, const std::string& name);
void plyInfo(Ctx* c, const std::string& path, long long* nv, long long* nf, float lo[3],
             float hi[3]);
// paged pool (pool_cache.cc): make domain d resident and have stream s wait
// for it; start loading d ahead (never evicting `keep`); record that the
// kernels on s that use d's block have been issued; cache state after a build
void poolAcquire(Ctx* c, int d, hipStream_t s);
void poolPrefetch(Ctx* c, int d, int keep);
void poolRelease(Ctx* c, int d, hipStream_t s);
void poolInitPaging(Ctx* c, const std::vector<std::pair<int, int>>& resident);
// spray_trace_rays_device (bvh.hip): scratch set-up per device, and the
// query on caller-owned device arrays using that scratch
void scratchInit(int device);
size_t scratchBytes();
void traceRaysNoAlloc(Ctx* c, const float* d_org, const float* d_dir, const float* d_tnear,
                      const float* d_tfar, size_t n, int mode, int domain, float* d_t, float* d_u,
                      float* d_v, int* d_prim);
void traceRaysDevice(Ctx* c, const float* d_org, const float* d_dir, const float* d_tnear,
                     const float* d_tfar, size_t n, int mode, int domain, float* d_t, float* d_u,
                     float* d_v, int* d_prim);
void domainListDevice(Ctx* c, const float* d_org, const float* d_dir, size_t n, int maxk,
                      int* d_counts, int* d_ids, float* d_ts);
// tracer.hip
void tracerTraceFrame(Ctx* c, float* rgba_out, spray_stats* st);
void tracerRelease(Ctx* c);
void tracerAutotune(Ctx* c, int* chunk, int* tile, int* cached);
void tracerKernelTime(Ctx* c, long long* launches, double* ms, int reset);
void tracerScheduleLog(Ctx* c, int max_records);
int tracerScheduleRead(Ctx* c, int* meta, unsigned long long* scores, int* order);
void tracerSetLogging(Ctx* c, bool on, long long max_samples);
long long tracerLogCount(const Ctx* c);
const float* tracerLogBuffer(Ctx* c);
// mlp.hip
void predictorRelease(Ctx* c);

}  // namespace spray

struct spray_ctx {
  spray::Ctx c;
};
