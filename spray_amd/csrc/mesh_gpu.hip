// Scene input and output on the GPU: binary PLY bodies decoded in HBM, the
// marching-cubes scene generator as kernels (a generated scene can be
// remeshed in place every frame), mesh readback, the scene writer and the
// PLY info query.
//
// Replaces the reference's PlyLoader::load (io/ply_loader.cc:179-324) as
// called by TriMeshBuffer::load (render/trimesh_buffer.cc:117-169), and the
// ply_header_reader app (apps/ply_header_reader.cc:36-108: a PLY file's
// vertex and face counts and the bounds of its vertices).
//
// Binary little-endian bodies are read from the file straight into pinned
// memory, copied to HBM as they are and decoded there: each block stages
// 256 records with aligned dword loads in LDS and every lane takes one
// record out of LDS (vertex: float x y z [uchar r g b]; face: list count +
// 3 indices of 1, 2 or 4 bytes, unsigned little-endian as the host parser
// reads them).  ASCII files stay on the host parser.  Errors are the host
// parser's: a face whose count is not 3 or whose index is out of range
// (the first such face of the file decides the message).
//
// The generator reproduces scenegen.cc's CPU meshes in the same vertex and
// face order: field values in double at the grid points, one vertex per
// sign-changing grid edge numbered in (axis, k, j, i) edge order (a cut
// bitmask, per-block counts of 256 edges scanned, a popcount rank inside the
// block), then the case table's triangles per cell numbered in (k, j, i)
// cell order (per-block counts scanned, an LDS scan inside the block).  Every double operation is
// the CPU's, in the CPU's order (no contraction), so meshes agree
// bit-for-bit wherever the device and host sin agree in the last place.
#include <hip/hip_runtime.h>

#include <fcntl.h>
#include <omp.h>
#include <unistd.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "ctx.h"
#include "kernels_common.h"

namespace spray {

namespace {

int gridOf(long long n, int block, long long cap = 65535) {
  long long g = (n + block - 1) / block;
  return (int)std::max<long long>(1, std::min<long long>(g, cap));
}

// Upper bound on kMcVerts' grid.  The environment variable
// SPRAY_MC_VERTS_BLOCKS (a positive block count) lowers it, so a test can run
// the path where each wave takes several groups of words; unset, only the
// launch's own cap applies.
long long mcVertsCap(long long cap) {
  const char* e = std::getenv("SPRAY_MC_VERTS_BLOCKS");
  const long long v = e ? std::atoll(e) : 0;
  return v >= 1 ? std::min(v, cap) : cap;
}

// ---------------------------------------------------------------------------
// Bounds: per-block min/max of float points, then one block over the partials
constexpr int kBlk = 256;
constexpr int kBoundBlocks = 1024;

__device__ void blockBounds(float lo[3], float hi[3], float* part) {
  __shared__ float red[6][kBlk / 64];
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  for (int c = 0; c < 3; ++c) {
    float a = waveMin(lo[c]), b = waveMax(hi[c]);
    if (l == 0) {
      red[c][w] = a;
      red[3 + c][w] = b;
    }
  }
  __syncthreads();
  if (threadIdx.x < 6) {
    float v = red[threadIdx.x][0];
    for (int i = 1; i < (int)(blockDim.x >> 6); ++i)
      v = threadIdx.x < 3 ? fminf(v, red[threadIdx.x][i]) : fmaxf(v, red[threadIdx.x][i]);
    part[blockIdx.x * 6 + threadIdx.x] = v;
  }
}

__global__ void __launch_bounds__(kBlk) kBoundsFinal(const float* part, int nparts, float* out) {
  float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (int i = threadIdx.x; i < nparts; i += blockDim.x)
    for (int c = 0; c < 3; ++c) {
      lo[c] = fminf(lo[c], part[i * 6 + c]);
      hi[c] = fmaxf(hi[c], part[i * 6 + 3 + c]);
    }
  blockBounds(lo, hi, out);  // a single block: writes out[0..5]
}

__global__ void __launch_bounds__(kBlk) kPointBounds(const float* v, long long nv, float* part) {
  float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < nv;
       i += (long long)gridDim.x * blockDim.x)
    for (int c = 0; c < 3; ++c) {
      lo[c] = fminf(lo[c], v[i * 3 + c]);
      hi[c] = fmaxf(hi[c], v[i * 3 + c]);
    }
  blockBounds(lo, hi, part);
}

// ---------------------------------------------------------------------------
// PLY body decode
constexpr int kMaxStride = 16;  // face: 4-byte count + 3 four-byte indices

struct Xform {
  float m[16];
};

// Stages records [r0, r0 + blockDim.x) of `stride` bytes in LDS; the body
// buffer is dword aligned and padded past its end.  Returns the first
// record's address in LDS.
__device__ __forceinline__ const uint8_t* stageRecords(const uint8_t* body, long long r0,
                                                       long long n, int stride, uint32_t* lds) {
  const long long b0 = r0 * stride, b1 = min(n, r0 + (long long)blockDim.x) * stride;
  const long long w0 = b0 >> 2, w1 = (b1 + 3) >> 2;
  const uint32_t* words = (const uint32_t*)body;
  for (long long w = w0 + threadIdx.x; w < w1; w += blockDim.x) lds[w - w0] = words[w];
  __syncthreads();
  return (const uint8_t*)lds + (b0 - (w0 << 2));
}

__device__ __forceinline__ uint32_t leBytes(const uint8_t* p, int n) {
  uint32_t v = 0;
  for (int i = 0; i < n; ++i) v |= (uint32_t)p[i] << (8 * i);
  return v;
}

__global__ void __launch_bounds__(kBlk) kPlyVerts(const uint8_t* __restrict__ body, long long nv,
                                                  int stride, int has_rgb, Xform xf, int apply,
                                                  float* __restrict__ v,
                                                  uint32_t* __restrict__ rgb) {
  __shared__ uint32_t lds[(kBlk * kMaxStride + 8) / 4];
  for (long long r0 = blockIdx.x * (long long)kBlk; r0 < nv; r0 += (long long)gridDim.x * kBlk) {
    const uint8_t* rec = stageRecords(body, r0, nv, stride, lds);
    const long long i = r0 + threadIdx.x;
    if (i < nv) {
      const uint8_t* p = rec + threadIdx.x * stride;
      float q[3];
      for (int k = 0; k < 3; ++k) q[k] = __uint_as_float(leBytes(p + 4 * k, 4));
      if (apply) {  // trimesh_buffer.cc:139-156, glm mat4 * vec4(p, 1)
        float o[3];
        for (int r = 0; r < 3; ++r) {
          float a = xf.m[r] * q[0] + xf.m[4 + r] * q[1];
          float b = xf.m[8 + r] * q[2] + xf.m[12 + r] * 1.0f;
          o[r] = a + b;
        }
        for (int k = 0; k < 3; ++k) q[k] = o[k];
      }
      for (int k = 0; k < 3; ++k) v[i * 3 + k] = q[k];
      rgb[i] = has_rgb ? ((uint32_t)p[12] << 16) | ((uint32_t)p[13] << 8) | p[14] : 0u;
    }
    __syncthreads();  // the next stage overwrites lds
  }
}

// err: min over bad faces of (face << 1 | kind), kind 0 = not a triangle,
// 1 = index out of range.  Bad faces get index 0 so later kernels stay in
// bounds.
__global__ void __launch_bounds__(kBlk) kPlyFaces(const uint8_t* __restrict__ body, long long nf,
                                                  int lb, int ib, long long nv,
                                                  uint32_t* __restrict__ f,
                                                  unsigned long long* err) {
  __shared__ uint32_t lds[(kBlk * kMaxStride + 8) / 4];
  const int stride = lb + 3 * ib;
  for (long long r0 = blockIdx.x * (long long)kBlk; r0 < nf; r0 += (long long)gridDim.x * kBlk) {
    const uint8_t* rec = stageRecords(body, r0, nf, stride, lds);
    const long long i = r0 + threadIdx.x;
    if (i < nf) {
      const uint8_t* p = rec + threadIdx.x * stride;
      const uint32_t cnt = leBytes(p, lb);
      uint32_t idx[3];
      for (int k = 0; k < 3; ++k) idx[k] = leBytes(p + lb + k * ib, ib);
      int bad = -1;
      if (cnt != 3)
        bad = 0;
      else if (idx[0] >= nv || idx[1] >= nv || idx[2] >= nv)
        bad = 1;
      for (int k = 0; k < 3; ++k) f[i * 3 + k] = bad < 0 ? idx[k] : 0u;
      if (bad >= 0) atomicMin(err, ((unsigned long long)i << 1) | (unsigned long long)bad);
    }
    __syncthreads();
  }
}

size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

// Reads n bytes at `off` of `path` into dst with parallel preads.
bool readRange(const std::string& path, long long off, size_t n, char* dst, int nthreads) {
  int fd = open(path.c_str(), O_RDONLY);
  if (fd < 0) return false;
  const size_t chunk = (size_t)8 << 20;
  const long long nchunks = (long long)((n + chunk - 1) / chunk);
  int failed = 0;
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads)
  for (long long k = 0; k < nchunks; ++k) {
    const size_t o = (size_t)k * chunk, len = std::min(chunk, n - o);
    size_t got = 0;
    while (got < len) {
      ssize_t r = pread(fd, dst + o + got, len - got, (off_t)(off + (long long)(o + got)));
      if (r <= 0) {
#pragma omp atomic write
        failed = 1;
        break;
      }
      got += (size_t)r;
    }
  }
  close(fd);
  return !failed;
}

// Host checks of a binary body against the file size (the host parser's
// "truncated" errors, first element first).
int checkBody(const std::string& file, const PlyLayout& L, std::string* err) {
  const long long vb = L.nv * L.vertexStride(), fb = L.nf * L.faceStride();
  const long long first = L.faces_first ? fb : vb;
  if (L.file_bytes < L.body + first) {
    *err = file + (L.faces_first ? ": truncated faces" : ": truncated vertices");
    return SPRAY_ERR_FORMAT;
  }
  if (L.file_bytes < L.body + vb + fb) {
    *err = file + (L.faces_first ? ": truncated vertices" : ": truncated faces");
    return SPRAY_ERR_FORMAT;
  }
  return SPRAY_OK;
}

// Pinned double buffer + device staging for PLY bodies.
struct BodyStager {
  void* pinned[2] = {nullptr, nullptr};
  hipEvent_t ev[2] = {nullptr, nullptr};
  size_t pin_bytes = 0;
  DevBuf dev;
  int slot = 0;
  hipStream_t last = nullptr;
  explicit BodyStager(size_t max_body) : pin_bytes(max_body) {
    for (int k = 0; k < 2; ++k) {
      SPRAY_HIP(hipHostMalloc(&pinned[k], std::max<size_t>(pin_bytes, 64), hipHostMallocDefault));
      SPRAY_HIP(hipEventCreateWithFlags(&ev[k], hipEventDisableTiming));
    }
    dev.reserve(max_body + 512 + 64);
  }
  ~BodyStager() {
    if (last) (void)hipStreamSynchronize(last);
    for (int k = 0; k < 2; ++k) {
      if (pinned[k]) (void)hipHostFree(pinned[k]);
      if (ev[k]) (void)hipEventDestroy(ev[k]);
    }
    dev.release();
  }
  // Reads the body of `file` and copies its vertex and face element to the
  // device (each at a 256-byte aligned address); decodes them on s.
  void decode(const std::string& file, const PlyLayout& L, const float* xf, bool apply,
              float* v, uint32_t* f, uint32_t* rgb, unsigned long long* err, hipStream_t s,
              int nthreads) {
    last = s;
    const size_t vb = (size_t)L.nv * L.vertexStride(), fb = (size_t)L.nf * L.faceStride();
    char* host = (char*)pinned[slot];
    SPRAY_HIP(hipEventSynchronize(ev[slot]));  // the last copy out of this slot is done
    if (!readRange(file, L.body, vb + fb, host, nthreads))
      throw StatusError(SPRAY_ERR_IO, "unable to read " + file);
    char* dv = dev.as<char>();
    char* df = dv + align256(vb);
    const char* hv = host + (L.faces_first ? fb : 0);
    const char* hf = host + (L.faces_first ? 0 : vb);
    SPRAY_HIP(hipMemcpyAsync(dv, hv, vb, hipMemcpyHostToDevice, s));
    SPRAY_HIP(hipMemcpyAsync(df, hf, fb, hipMemcpyHostToDevice, s));
    SPRAY_HIP(hipEventRecord(ev[slot], s));
    slot ^= 1;
    Xform X;
    std::memcpy(X.m, xf, sizeof(X.m));
    kPlyVerts<<<gridOf(L.nv, kBlk), kBlk, 0, s>>>((const uint8_t*)dv, L.nv, L.vertexStride(),
                                                   L.rgb ? 1 : 0, X, apply ? 1 : 0, v, rgb);
    kPlyFaces<<<gridOf(L.nf, kBlk), kBlk, 0, s>>>((const uint8_t*)df, L.nf, L.list_bytes,
                                                   L.index_bytes, L.nv, f, err);
    SPRAY_HIP(hipGetLastError());
  }
};

std::string faceError(const std::string& file, unsigned long long key) {
  return file + ((key & 1ull) ? ": face index out of range" : ": non-triangle face");
}

// Domains from PLY files: binary bodies decoded on the GPU, ASCII files
// parsed on the host (with the domain transform applied there).
struct PlySource : MeshSource {
  const HostScene& scene;
  std::vector<PlyLayout> L;
  std::vector<HostMesh> host;
  std::vector<int> codes;
  std::vector<std::string> errs;
  std::vector<char> filled;
  BodyStager* stager = nullptr;
  DevBuf errbuf;
  int nthreads;
  explicit PlySource(Ctx* c) : scene(c->scene) {
    const int nd = (int)scene.domains.size();
    L.resize(nd);
    host.resize(nd);
    codes.assign(nd, SPRAY_OK);
    errs.resize(nd);
    filled.assign(nd, 0);
    nthreads = std::max(1, c->cfg.nthreads);
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads)
    for (int d = 0; d < nd; ++d) {
      const HostDomain& hd = scene.domains[d];
      codes[d] = readPlyHeader(hd.filename, &L[d], &errs[d]);
      if (codes[d] != SPRAY_OK) continue;
      if (L[d].binary) {
        codes[d] = checkBody(hd.filename, L[d], &errs[d]);
        continue;
      }
      codes[d] = readPly(hd.filename, &host[d], &errs[d]);
      if (codes[d] != SPRAY_OK || transformIsIdentity(hd.xf)) continue;
      // trimesh_buffer.cc:139-156: transform vertices when not identity
      size_t nv = host[d].v.size() / 3;
      for (size_t i = 0; i < nv; ++i) {
        float p[3] = {host[d].v[i * 3], host[d].v[i * 3 + 1], host[d].v[i * 3 + 2]}, o[3];
        transformPoint(hd.xf, p, o);
        std::memcpy(&host[d].v[i * 3], o, 12);
      }
    }
    for (int d = 0; d < nd; ++d)
      if (codes[d] != SPRAY_OK) throw StatusError(codes[d], errs[d]);
    size_t max_body = 0;
    for (int d = 0; d < nd; ++d)
      if (L[d].binary)
        max_body = std::max(max_body, (size_t)(L[d].nv * L[d].vertexStride() +
                                               L[d].nf * L[d].faceStride()));
    errbuf.reserve((size_t)nd * 8 + 64);
    SPRAY_HIP(hipMemsetAsync(errbuf.p, 0xff, (size_t)nd * 8, c->stream));
    if (max_body) stager = new BodyStager(max_body);
  }
  ~PlySource() override {
    delete stager;
    errbuf.release();
  }
  long long numVerts(int d) const override {
    return L[d].binary ? L[d].nv : (long long)host[d].v.size() / 3;
  }
  long long numFaces(int d) const override {
    return L[d].binary ? L[d].nf : (long long)host[d].f.size() / 3;
  }
  void fill(Ctx* c, int d, float* v, uint32_t* f, uint32_t* rgb, hipStream_t s) override {
    (void)c;
    filled[d] = 1;
    if (!L[d].binary) {
      HostMeshSource hs(host);
      hs.fill(c, d, v, f, rgb, s);
      return;
    }
    const HostDomain& hd = scene.domains[d];
    stager->decode(hd.filename, L[d], hd.xf, !transformIsIdentity(hd.xf), v, f, rgb,
                   errbuf.as<unsigned long long>() + d, s, nthreads);
  }
  void check(Ctx* c) override {
    SPRAY_HIP(hipStreamSynchronize(c->stream));
    const int nd = (int)L.size();
    std::vector<unsigned long long> e(nd);
    SPRAY_HIP(hipMemcpy(e.data(), errbuf.p, (size_t)nd * 8, hipMemcpyDeviceToHost));
    for (int d = 0; d < nd; ++d)
      if (filled[d] && e[d] != ~0ull)
        throw StatusError(SPRAY_ERR_FORMAT, faceError(scene.domains[d].filename, e[d]));
  }
};

// ---------------------------------------------------------------------------
// Marching cubes (scenegen.cc's mcGenerate on the GPU)
// x / d for 32-bit x as (umulhi(x, m) + x) >> s, s = ceil(log2 d) and
// m = floor(2^32 (2^s - d) / d) + 1 (round-up multiplier, Granlund and
// Montgomery 1994): point, edge and cell numbers are split into (i, j, k)
// without integer division.
struct McDiv {
  uint32_t m, s;
};

McDiv mcDiv(uint32_t d) {
  uint32_t s = 0;
  while ((1ull << s) < d) ++s;
  return McDiv{(uint32_t)(((1ull << 32) * ((1ull << s) - d)) / d + 1), s};
}

__device__ __forceinline__ uint32_t mcDivide(uint32_t x, const McDiv& v) {
  return (uint32_t)(((unsigned long long)__umulhi(x, v.m) + x) >> v.s);
}

struct McGrid {
  int n, N;              // cells and grid points per axis
  long long E;           // grid edges per axis: n * N * N
  McDiv dn, dN;          // division by n and by N
  double ox, oy, oz, h;  // domain origin and cell size
  double a, b, phase;
};

struct McCorners {  // per local edge: corner offset and axis
  int dx[12], dy[12], dz[12], axis[12];
};

// position of edge (axis, i, j, k) in the axis' edge numbering
__device__ __forceinline__ long long mcEdgeIndex(const McGrid& g, int axis, int i, int j, int k) {
  if (axis == 0) return ((long long)k * g.N + j) * g.n + i;
  if (axis == 1) return ((long long)k * g.n + j) * g.N + i;
  return ((long long)k * g.N + j) * g.N + i;
}

__device__ __forceinline__ void mcEdgeDecode(const McGrid& g, long long e, int* axis, int* i,
                                             int* j, int* k) {
  const int ax = e < g.E ? 0 : (e < 2 * g.E ? 1 : 2);
  const uint32_t r = (uint32_t)(e - ax * g.E);
  const uint32_t ni = ax == 0 ? g.n : g.N, nj = ax == 1 ? g.n : g.N;
  const uint32_t q = mcDivide(r, ax == 0 ? g.dn : g.dN);
  const uint32_t kk = mcDivide(q, ax == 1 ? g.dn : g.dN);
  *axis = ax;
  *i = (int)(r - q * ni);
  *j = (int)(q - kk * nj);
  *k = (int)kk;
}

// Kernels that evaluate the field take this many bytes of dynamic LDS for
// its separable factors.
size_t mcTabBytes(const McGrid& g) { return (size_t)3 * g.N * sizeof(double); }

// The field's separable factors sin(a x + phase), sin(a y) and sin(a z) at
// every grid coordinate of the domain, into LDS (3 N doubles), the
// coordinates formed as in mcFieldAt.
__device__ void mcAxisTables(const McGrid& g, double* tab) {
  for (int q = threadIdx.x; q < 3 * g.N; q += blockDim.x) {
    const int ax = q / g.N, i = q - ax * g.N;
    const double c = (ax == 0 ? g.ox : (ax == 1 ? g.oy : g.oz)) + (double)i * g.h;
    tab[q] = ax == 0 ? sinD(g.a * c + g.phase) : sinD(g.a * c);
  }
  __syncthreads();
}

// Field value at grid point (i, j, k), in the CPU generator's operations
// and order (the file is built without contraction):
// sin(a x + phase) sin(a y) sin(a z) + 0.3 sin(b (x + y + z) + phase), the
// first three factors read from the axis tables.
__device__ __forceinline__ double mcFieldAt(const McGrid& g, const double* tab, int i, int j,
                                            int k) {
  const double x = g.ox + (double)i * g.h, y = g.oy + (double)j * g.h, z = g.oz + (double)k * g.h;
  return tab[i] * tab[g.N + j] * tab[2 * g.N + k] + 0.3 * sinD(g.b * (x + y + z) + g.phase);
}

// Field signs as bytes (1: inside, f < 0): the cut and case tests read one
// byte per grid point; field values are evaluated again only at the ends
// of cut edges.
__global__ void __launch_bounds__(kBlk) kMcSign(McGrid g, uint8_t* __restrict__ sgn) {
  extern __shared__ double tab[];
  mcAxisTables(g, tab);
  const long long np = (long long)g.N * g.N * g.N;
  for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < np;
       t += (long long)gridDim.x * blockDim.x) {
    const uint32_t q = mcDivide((uint32_t)t, g.dN), k = mcDivide(q, g.dN);
    const int i = (int)((uint32_t)t - q * g.N), j = (int)(q - k * g.N);
    sgn[t] = mcFieldAt(g, tab, i, j, k) < 0.0 ? 1 : 0;
  }
}

__device__ __forceinline__ int mcS(const uint8_t* sgn, const McGrid& g, int i, int j, int k) {
  return sgn[((long long)k * g.N + j) * g.N + i];
}

__device__ __forceinline__ void mcVertex(const McGrid& g, int axis, int i, int j, int k,
                                         double f0, double f1, float out[3]) {
  const double t = (0.0 - f0) / (f1 - f0);
  double pos[3] = {g.ox + (double)i * g.h, g.oy + (double)j * g.h, g.oz + (double)k * g.h};
  pos[axis] += t * g.h;
  for (int c = 0; c < 3; ++c) out[c] = (float)pos[c];
}

// Edge cuts as bits: bit e of the mask is set where the field changes sign
// along edge e; edges go in blocks of 256 (four 64-bit words), bcnt[blk] =
// cuts in the block (bcnt[nblk] = 0, for the exclusive scan of the block
// counts).
__global__ void __launch_bounds__(kBlk) kMcEdgeBits(McGrid g, const uint8_t* __restrict__ sgn,
                                                    unsigned long long* __restrict__ bits,
                                                    uint32_t* __restrict__ bcnt) {
  __shared__ uint32_t wsum[kBlk / 64];
  const long long ne = 3 * g.E, nblk = (ne + kBlk - 1) / kBlk;
  const int w = threadIdx.x >> 6;
  for (long long blk = blockIdx.x; blk < nblk; blk += gridDim.x) {
    const long long e = blk * kBlk + threadIdx.x;
    bool cut = false;
    if (e < ne) {
      int axis, i, j, k;
      mcEdgeDecode(g, e, &axis, &i, &j, &k);
      cut = mcS(sgn, g, i, j, k) != mcS(sgn, g, i + (axis == 0), j + (axis == 1), k + (axis == 2));
    }
    const unsigned long long m = __ballot(cut);
    if ((threadIdx.x & 63) == 0) {
      bits[blk * (kBlk / 64) + w] = m;
      wsum[w] = (uint32_t)__popcll(m);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      uint32_t t = 0;
      for (int q = 0; q < kBlk / 64; ++q) t += wsum[q];
      bcnt[blk] = t;
    }
    __syncthreads();
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) bcnt[nblk] = 0u;
}

// Vertex id of a cut edge: its block's offset plus the cuts before it in
// the block.
__device__ __forceinline__ uint32_t mcEdgeId(const unsigned long long* __restrict__ bits,
                                             const uint32_t* __restrict__ eoff, long long e) {
  const long long blk = e / kBlk;
  const int w = (int)((e >> 6) & (kBlk / 64 - 1)), b = (int)(e & 63);
  uint32_t id = eoff[blk];
  for (int q = 0; q < w; ++q) id += (uint32_t)__popcll(bits[blk * (kBlk / 64) + q]);
  return id + (uint32_t)__popcll(bits[blk * (kBlk / 64) + w] & ((1ull << b) - 1ull));
}

// LDS written by some lanes and read by others of the same wave
__device__ __forceinline__ void waveSync() {
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
}

// Position of the k-th set bit of m (k < popcount(m)).
__device__ __forceinline__ int mcNthBit(unsigned long long m, uint32_t k) {
  int pos = 0;
  for (int wd = 32; wd; wd >>= 1) {
    const unsigned long long low = m & ((1ull << wd) - 1ull);
    const uint32_t c = (uint32_t)__popcll(low);
    if (k >= c) {
      k -= c;
      m >>= wd;
      pos += wd;
    } else {
      m = low;
    }
  }
  return pos;
}

// Vertices of the cut edges.  Each wave takes 64 words of the cut mask
// (16 blocks of 256 edges), scans their cut counts, and gives each lane one
// cut (its word found by a binary search of the scan, its bit by a
// popcount search), so the field is evaluated for 64 cuts at a time and the
// vertex stores are contiguous.  With v == nullptr only the per-block
// bounds of the vertices are produced (part, one slot per block).
__global__ void __launch_bounds__(kBlk) kMcVerts(McGrid g,
                                                 const unsigned long long* __restrict__ bits,
                                                 const uint32_t* __restrict__ eoff, uint32_t color,
                                                 float* __restrict__ v, uint32_t* __restrict__ rgb,
                                                 float* part) {
  static_assert(kBlk % 256 == 0, "a wave's 64 words are whole blocks of edges");
  __shared__ uint32_t scan[kBlk];
  __shared__ unsigned long long word[kBlk];
  extern __shared__ double tab[];
  mcAxisTables(g, tab);
  float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
  const long long nwords = ((3 * g.E + kBlk - 1) / kBlk) * (kBlk / 64);
  const long long ngroups = (nwords + 63) / 64;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  uint32_t* ex = scan + w * 64;
  unsigned long long* wd = word + w * 64;
  for (long long gi = (long long)blockIdx.x * (kBlk / 64) + w; gi < ngroups;
       gi += (long long)gridDim.x * (kBlk / 64)) {
    const long long wi = gi * 64 + lane;
    const unsigned long long m = wi < nwords ? bits[wi] : 0ull;
    const uint32_t c = (uint32_t)__popcll(m);
    uint32_t incl = c;
    for (int o = 1; o < 64; o <<= 1) {
      const uint32_t x = __shfl_up(incl, o, 64);
      if (lane >= o) incl += x;
    }
    const uint32_t total = (uint32_t)__shfl(incl, 63, 64);
    if (total == 0) continue;  // uniform over the wave
    waveSync();  // the previous group's reads of ex and wd are done
    ex[lane] = incl - c;
    wd[lane] = m;
    waveSync();
    const uint32_t base = eoff[gi * 64 / (kBlk / 64)];  // first block of the group
    for (uint32_t r = lane; r < total; r += 64) {
      int L = 0;  // last word whose scan start is <= r: it holds cut r
      for (int step = 32; step; step >>= 1)
        if (ex[L + step] <= r) L += step;
      const int b = mcNthBit(wd[L], r - ex[L]);
      const long long e = (gi * 64 + L) * 64 + b;
      int axis, i, j, k;
      mcEdgeDecode(g, e, &axis, &i, &j, &k);
      const double f0 = mcFieldAt(g, tab, i, j, k);
      const double f1 = mcFieldAt(g, tab, i + (axis == 0), j + (axis == 1), k + (axis == 2));
      float p[3];
      mcVertex(g, axis, i, j, k, f0, f1, p);
      if (v) {
        const uint32_t id = base + r;
        for (int q = 0; q < 3; ++q) v[(size_t)id * 3 + q] = p[q];
        rgb[id] = color;
      } else {
        for (int q = 0; q < 3; ++q) {
          lo[q] = fminf(lo[q], p[q]);
          hi[q] = fmaxf(hi[q], p[q]);
        }
      }
    }
  }
  if (part) blockBounds(lo, hi, part);
}

__device__ __forceinline__ int mcCase(const McGrid& g, const uint8_t* sgn, int i, int j, int k) {
  int cs = 0;
  for (int c = 0; c < 8; ++c) cs |= mcS(sgn, g, i + (c & 1), j + ((c >> 1) & 1), k + ((c >> 2) & 1)) << c;
  return cs;
}

__device__ __forceinline__ void mcCellDecode(const McGrid& g, long long t, int* i, int* j, int* k) {
  const uint32_t q = mcDivide((uint32_t)t, g.dn), kk = mcDivide(q, g.dn);
  *i = (int)((uint32_t)t - q * g.n);
  *j = (int)(q - kk * g.n);
  *k = (int)kk;
}

// Triangles per block of 256 cells (bcnt[nblk] = 0).
__global__ void __launch_bounds__(kBlk) kMcCellCount(McGrid g, const uint8_t* __restrict__ sgn,
                                                     const uint8_t* __restrict__ ntri,
                                                     uint32_t* __restrict__ bcnt) {
  __shared__ uint32_t wsum[kBlk / 64];
  const long long nc = (long long)g.n * g.n * g.n, nblk = (nc + kBlk - 1) / kBlk;
  for (long long blk = blockIdx.x; blk < nblk; blk += gridDim.x) {
    const long long t = blk * kBlk + threadIdx.x;
    uint32_t n = 0;
    if (t < nc) {
      int i, j, k;
      mcCellDecode(g, t, &i, &j, &k);
      n = ntri[mcCase(g, sgn, i, j, k)];
    }
    n = waveSumU(n);
    if ((threadIdx.x & 63) == 0) wsum[threadIdx.x >> 6] = n;
    __syncthreads();
    if (threadIdx.x == 0) {
      uint32_t s = 0;
      for (int q = 0; q < kBlk / 64; ++q) s += wsum[q];
      bcnt[blk] = s;
    }
    __syncthreads();
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) bcnt[nblk] = 0u;
}

// Grid cap of kMcSign and kMcCells: eight blocks per CU, each stages its
// tables into LDS once.
constexpr long long kMcCellBlocks = 2048;

// Triangles of every cell, in cell order: the block's offset plus an
// exclusive scan of the per-cell counts inside the block.  The vertices of a
// wave's triangles are then spread over its lanes, one vertex per lane, its
// cell found by a binary search of the wave's scan, so the edge-id loads of
// all of them are in flight together and the index stores are contiguous.
__global__ void __launch_bounds__(kBlk) kMcCells(McGrid g, const uint8_t* __restrict__ sgn,
                                                 const int8_t* __restrict__ table,
                                                 const uint8_t* __restrict__ ntri, McCorners ec,
                                                 const unsigned long long* __restrict__ bits,
                                                 const uint32_t* __restrict__ eoff,
                                                 const uint32_t* __restrict__ coff,
                                                 uint32_t* __restrict__ out) {
  static_assert(kBlk == 256, "one case per thread when the tables are staged");
  __shared__ int8_t tab[256 * 24];  // the table's rows without their terminator
  __shared__ uint8_t cnt[256];
  __shared__ int4 corner[12];  // dx, dy, dz, axis of each local edge
  __shared__ uint32_t wsum[kBlk / 64];
  __shared__ uint32_t excl[kBlk];
  __shared__ int ci[kBlk], cj[kBlk], ck[kBlk];
  __shared__ uint8_t ccase[kBlk];
  for (int x = threadIdx.x; x < 256 * 24; x += kBlk) tab[x] = table[(x / 24) * 25 + x % 24];
  cnt[threadIdx.x] = ntri[threadIdx.x];
  if (threadIdx.x < 12) {
    const int e = threadIdx.x;
    corner[e] = make_int4(ec.dx[e], ec.dy[e], ec.dz[e], ec.axis[e]);
  }
  __syncthreads();
  const long long nc = (long long)g.n * g.n * g.n, nblk = (nc + kBlk - 1) / kBlk;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const uint32_t* ex = excl + w * 64;
  for (long long blk = blockIdx.x; blk < nblk; blk += gridDim.x) {
    const long long t = blk * kBlk + threadIdx.x;
    int i = 0, j = 0, k = 0, cs = 0;
    uint32_t n = 0;
    if (t < nc) {
      mcCellDecode(g, t, &i, &j, &k);
      cs = mcCase(g, sgn, i, j, k);
      n = cnt[cs];
    }
    uint32_t incl = n;
    for (int o = 1; o < 64; o <<= 1) {
      const uint32_t x = __shfl_up(incl, o, 64);
      if (lane >= o) incl += x;
    }
    if (lane == 63) wsum[w] = incl;
    excl[threadIdx.x] = incl - n;
    ci[threadIdx.x] = i;
    cj[threadIdx.x] = j;
    ck[threadIdx.x] = k;
    ccase[threadIdx.x] = (uint8_t)cs;
    __syncthreads();
    uint32_t base = coff[blk];
    for (int q = 0; q < w; ++q) base += wsum[q];
    const uint32_t nv = 3u * (uint32_t)__shfl(incl, 63, 64);
    for (uint32_t vtx = lane; vtx < nv; vtx += 64) {
      const uint32_t tri = vtx / 3u, r = vtx - 3u * tri;
      // last lane whose scan start is <= tri: it owns triangle tri (n > 0)
      int L = 0;
      for (int step = 32; step; step >>= 1)
        if (ex[L + step] <= tri) L += step;
      const int c = w * 64 + L;
      const int e = tab[ccase[c] * 24 + 3 * (int)(tri - ex[L]) + (int)r];
      const int4 cr = corner[e];
      const long long idx =
          cr.w * g.E + mcEdgeIndex(g, cr.w, ci[c] + cr.x, cj[c] + cr.y, ck[c] + cr.z);
      out[(size_t)base * 3 + vtx] = mcEdgeId(bits, eoff, idx);
    }
    __syncthreads();  // the LDS rows are rewritten by the next block of cells
  }
}

__global__ void kMcTotals(const uint32_t* vtotal, const uint32_t* ftotal, uint32_t* out) {
  out[0] = *vtotal;
  out[1] = *ftotal;
}

// Per domain, the count pass leaves what the emission needs: sign bytes,
// cut bits and the scanned block offsets of edges and cells.  When those
// buffers for every domain would exceed kMcKeepBudget (or a quarter of the
// free HBM), they are kept for none and the emission recomputes them per
// domain into scratch.
constexpr size_t kMcKeepBudget = (size_t)16 << 30;
constexpr int kCountStreams = 4;  // streams of the count pass over the domains

struct McSource : MeshSource {
  struct Parts {  // byte offsets of one domain's buffers
    size_t sgn, bits, eoff, coff;
  };
  spray_scenegen_params p;
  int nd;
  bool keep = false;
  std::vector<McGrid> grid;
  std::vector<Parts> parts;
  std::vector<long long> nv, nf;
  std::vector<float> bounds;  // 6 per domain: lo xyz, hi xyz
  DevBuf& store;  // the context's generator store (kept between remeshes)
  DevBuf table, ntri, part, dev_out;
  McCorners ec;
  static long long edgeBlocks(const McGrid& g) { return (3 * g.E + kBlk - 1) / kBlk; }
  static long long cellBlocks(const McGrid& g) {
    return ((long long)g.n * g.n * g.n + kBlk - 1) / kBlk;
  }
  static Parts sizes(const McGrid& g) {  // bytes of each buffer, 256-aligned
    auto al = [](size_t b) { return (b + 255) & ~(size_t)255; };
    return Parts{al((size_t)g.N * g.N * g.N), al((size_t)edgeBlocks(g) * (kBlk / 64) * 8),
                 al((size_t)(edgeBlocks(g) + 1) * 4), al((size_t)(cellBlocks(g) + 1) * 4)};
  }
  // sign bytes, cut bits and scanned block offsets of domain d; with bnd,
  // also its vertex bounds (partials in bpart), and with tot its vertex and
  // face totals.  sums: the scans' tile sums (scanSumsWords of the larger
  // scan), or null for the calling thread's own buffer.
  void count(const McGrid& g, const Parts& at, float* bnd, uint32_t* tot, hipStream_t s,
             float* bpart = nullptr, uint32_t* sums = nullptr) {
    uint8_t* sgn = store.as<uint8_t>() + at.sgn;
    unsigned long long* bits = (unsigned long long*)(store.as<char>() + at.bits);
    uint32_t* eoff = (uint32_t*)(store.as<char>() + at.eoff);
    uint32_t* coff = (uint32_t*)(store.as<char>() + at.coff);
    const long long np = (long long)g.N * g.N * g.N, ne = 3 * g.E, nc = (long long)g.n * g.n * g.n;
    const long long eb = edgeBlocks(g), cb = cellBlocks(g);
    kMcSign<<<gridOf(np, kBlk, kMcCellBlocks), kBlk, mcTabBytes(g), s>>>(g, sgn);
    kMcEdgeBits<<<gridOf(ne, kBlk), kBlk, 0, s>>>(g, sgn, bits, eoff);
    auto scan = [&](uint32_t* x, long long n) {
      if (sums)
        exclusiveScan(x, x, n, s, sums);
      else
        exclusiveScan(x, x, n, s);
    };
    scan(eoff, eb + 1);
    if (bnd) {
      float* bp = bpart ? bpart : part.as<float>();
      const int vgrid = gridOf(eb * (kBlk / 64), kBlk, mcVertsCap(kBoundBlocks));
      kMcVerts<<<vgrid, kBlk, mcTabBytes(g), s>>>(g, bits, eoff, 0u, nullptr, nullptr, bp);
      kBoundsFinal<<<1, kBlk, 0, s>>>(bp, vgrid, bnd);
    }
    kMcCellCount<<<gridOf(nc, kBlk), kBlk, 0, s>>>(g, sgn, ntri.as<uint8_t>(), coff);
    scan(coff, cb + 1);
    if (tot) kMcTotals<<<1, 1, 0, s>>>(eoff + eb, coff + cb, tot);
  }
  McSource(Ctx* c, const spray_scenegen_params& prm) : p(prm), store(c->gen_store) {
    nd = p.nx * p.ny * p.nz;
    grid.resize(nd);
    parts.resize(nd);
    Parts mx{0, 0, 0, 0};
    size_t total = 0;
    for (int d = 0; d < nd; ++d) {
      double a, b;
      int n;
      mcDomainParams(p, d, &a, &b, &n);
      McGrid& g = grid[d];
      g.n = n;
      g.N = n + 1;
      g.E = (long long)n * g.N * g.N;
      g.dn = mcDiv((uint32_t)n);
      g.dN = mcDiv((uint32_t)g.N);
      g.ox = (double)(d % p.nx);
      g.oy = (double)((d / p.nx) % p.ny);
      g.oz = (double)(d / (p.nx * p.ny));
      g.h = 1.0 / n;
      g.a = a;
      g.b = b;
      g.phase = p.phase;
      if (3 * g.E + 1 >= (1ll << 32) || (long long)n * n * n + 1 >= (1ll << 32))
        throw StatusError(SPRAY_ERR_CAPACITY, "scenegen: grid too fine for 32-bit numbering");
      const Parts z = sizes(g);
      mx = Parts{std::max(mx.sgn, z.sgn), std::max(mx.bits, z.bits), std::max(mx.eoff, z.eoff),
                 std::max(mx.coff, z.coff)};
      total += z.sgn + z.bits + z.eoff + z.coff;
    }
    size_t free_b = 0, total_b = 0;
    SPRAY_HIP(hipMemGetInfo(&free_b, &total_b));
    keep = total <= std::min(kMcKeepBudget, free_b / 4);
    if (keep) {
      size_t at = 0;
      for (int d = 0; d < nd; ++d) {
        const Parts z = sizes(grid[d]);
        parts[d] = Parts{at, at + z.sgn, at + z.sgn + z.bits, at + z.sgn + z.bits + z.eoff};
        at += z.sgn + z.bits + z.eoff + z.coff;
      }
      store.reserve(at + 64);
    } else {  // one scratch set, sized for the largest domain
      const Parts one{0, mx.sgn, mx.sgn + mx.bits, mx.sgn + mx.bits + mx.eoff};
      for (int d = 0; d < nd; ++d) parts[d] = one;
      store.reserve(mx.sgn + mx.bits + mx.eoff + mx.coff + 64);
    }
    const int8_t* T = mcTriTable();
    uint8_t counts[256];
    for (int cs = 0; cs < 256; ++cs) {
      int q = 0;
      while (T[cs * 25 + q] >= 0) q += 3;
      counts[cs] = (uint8_t)(q / 3);
    }
    for (int e = 0; e < 12; ++e) mcEdgeCorner(e, &ec.dx[e], &ec.dy[e], &ec.dz[e], &ec.axis[e]);
    table.reserve(256 * 25);
    ntri.reserve(256);
    part.reserve((size_t)kBoundBlocks * 6 * 4 + 64);
    dev_out.reserve((size_t)nd * 8 * 4 + 64);
    hipStream_t s = c->stream;
    SPRAY_HIP(hipMemcpyAsync(table.p, T, 256 * 25, hipMemcpyHostToDevice, s));
    SPRAY_HIP(hipMemcpyAsync(ntri.p, counts, 256, hipMemcpyHostToDevice, s));
    // count pass: vertex and face totals and vertex bounds of every domain;
    // with every domain's buffers kept, the domains run on up to
    // kCountStreams streams (the context's among them), each with its own
    // bound partials and scan tile sums
    uint32_t* tot = dev_out.as<uint32_t>();
    float* bnd = (float*)(tot + 2 * (size_t)nd);
    if (keep && nd > 1) {
      const int K = std::min(pipelineCount(kCountStreams), nd);
      long long scan_max = 0;
      for (int d = 0; d < nd; ++d)
        scan_max = std::max(scan_max, std::max(edgeBlocks(grid[d]), cellBlocks(grid[d])) + 1);
      const size_t part_b = align256((size_t)kBoundBlocks * 6 * 4);
      const size_t sums_b = align256((size_t)scanSumsWords(scan_max) * 4);
      struct Per {  // per stream: bound partials, tile sums
        DevBuf b;
        ~Per() { b.release(); }
      } pb;
      DevBuf& per = pb.b;
      per.reserve((part_b + sums_b) * K + 64);
      {
        StreamSet st(K - 1, s);  // stream 0 is s itself
        for (int d = 0; d < nd; ++d) {
          char* base = per.as<char>() + (part_b + sums_b) * (d % K);
          count(grid[d], parts[d], bnd + 6 * d, tot + 2 * d, d % K ? st[d % K - 1] : s,
                (float*)base, (uint32_t*)(base + part_b));
        }
        st.join(s);
      }
    } else {
      for (int d = 0; d < nd; ++d) count(grid[d], parts[d], bnd + 6 * d, tot + 2 * d, s);
    }
    SPRAY_HIP(hipGetLastError());
    std::vector<uint32_t> h((size_t)nd * 2);
    bounds.resize((size_t)nd * 6);
    SPRAY_HIP(hipMemcpyAsync(h.data(), tot, (size_t)nd * 8, hipMemcpyDeviceToHost, s));
    SPRAY_HIP(hipMemcpyAsync(bounds.data(), bnd, (size_t)nd * 24, hipMemcpyDeviceToHost, s));
    SPRAY_HIP(hipStreamSynchronize(s));
    nv.resize(nd);
    nf.resize(nd);
    for (int d = 0; d < nd; ++d) {
      nv[d] = h[2 * d];
      nf[d] = h[2 * d + 1];
    }
  }
  ~McSource() override {
    for (DevBuf* b : {&table, &ntri, &part, &dev_out}) b->release();
  }
  long long numVerts(int d) const override { return nv[d]; }
  long long numFaces(int d) const override { return nf[d]; }
  // with every domain's count buffers kept, a fill reads only its own
  bool pipelined() const override { return keep; }
  bool rebuilds() const override { return true; }
  void fill(Ctx* c, int d, float* v, uint32_t* f, uint32_t* rgb, hipStream_t s) override {
    (void)c;
    const McGrid& g = grid[d];
    const Parts& at = parts[d];
    if (!keep) count(g, at, nullptr, nullptr, s);
    const uint8_t* sgn = store.as<uint8_t>() + at.sgn;
    const unsigned long long* bits = (const unsigned long long*)(store.as<char>() + at.bits);
    const uint32_t* eoff = (const uint32_t*)(store.as<char>() + at.eoff);
    const uint32_t* coff = (const uint32_t*)(store.as<char>() + at.coff);
    const long long ne = 3 * g.E, nc = (long long)g.n * g.n * g.n;
    kMcVerts<<<gridOf(edgeBlocks(g) * (kBlk / 64), kBlk, mcVertsCap(65535)), kBlk, mcTabBytes(g), s>>>(
        g, bits, eoff, p.rgb, v, rgb, nullptr);
    kMcCells<<<gridOf(nc, kBlk, kMcCellBlocks), kBlk, 0, s>>>(g, sgn, table.as<int8_t>(),
                                                             ntri.as<uint8_t>(), ec, bits, eoff,
                                                             coff, f);
    SPRAY_HIP(hipGetLastError());
  }
};

long long domainVerts(const DomainPool& P, int d) { return P.nverts[d]; }

}  // namespace

void loadScenePly(Ctx* c, HostScene& scene) {
  c->scene = std::move(scene);
  c->generated = false;
  c->gen_store.release();
  buildCacheRelease(c);
  PlySource src(c);
  tracerRelease(c);
  buildPool(c, src);
}

void generateScene(Ctx* c, const spray_scenegen_params& p, const std::string& name,
                   const std::string& extra, bool remesh) {
  const int nd = p.nx * p.ny * p.nz;
  if (p.nx <= 0 || p.ny <= 0 || p.nz <= 0 || (p.res <= 0 && !p.per_domain_random))
    throw StatusError(SPRAY_ERR_INVALID, "scenegen: empty domain grid or resolution");
  if (nd > kMaxDomains) throw StatusError(SPRAY_ERR_INVALID, "scenegen: too many domains");
  McSource src(c, p);
  // the .spray description spray_scenegen_write would write for this scene
  std::string text = extra + "\n";
  const float white[3] = {1.f, 1.f, 1.f};
  for (int d = 0; d < nd; ++d) {
    if (src.nf[d] == 0)
      throw StatusError(SPRAY_ERR_INVALID, "scenegen: a domain has no surface (empty mesh)");
    char fname[64];
    std::snprintf(fname, sizeof(fname), "%s_d%04d.ply", name.c_str(), d);
    text += sceneDomainBlock(d, fname, kDiffuse, white, 1.f, 1.f, &src.bounds[6 * d],
                             &src.bounds[6 * d + 3], src.nv[d], src.nf[d]);
  }
  HostScene scene;
  std::string err;
  int rc = parseSceneText(text, name + ".spray", "", &scene, &err);
  if (rc != SPRAY_OK) throw StatusError(rc, err);
  c->scene = std::move(scene);
  if (remesh) {
    c->param_version++;  // pool addresses change: re-capture the pass graphs
  } else {
    tracerRelease(c);
  }
  buildPool(c, src, remesh);
  c->generated = true;
  c->gen = p;
  c->gen_name = name;
  c->gen_extra = extra;
}

void domainMesh(Ctx* c, int d, float* v, uint32_t* f, uint32_t* rgb, long long* nv_out,
                long long* nf_out) {
  const DomainPool& P = c->pool;
  const DomainDesc& D = P.desc[d];
  const long long nv = domainVerts(P, d);
  if (nv_out) *nv_out = nv;
  if (nf_out) *nf_out = D.num_tris;
  if (!v && !f && !rgb) return;
  if (P.paged) throw StatusError(SPRAY_ERR_STATE, "a paged pool keeps no raw meshes");
  SPRAY_HIP(hipStreamSynchronize(c->stream));
  if (v)
    SPRAY_HIP(hipMemcpy(v, P.vtx.as<float>() + D.vtx_off * 3, (size_t)nv * 12, hipMemcpyDeviceToHost));
  if (f)
    SPRAY_HIP(hipMemcpy(f, P.face.as<uint32_t>() + D.face_off * 3, (size_t)D.num_tris * 12,
                        hipMemcpyDeviceToHost));
  if (rgb)
    SPRAY_HIP(hipMemcpy(rgb, P.rgb.as<uint32_t>() + D.vtx_off, (size_t)nv * 4, hipMemcpyDeviceToHost));
}

void writeScene(Ctx* c, const std::string& dir, const std::string& name) {
  const DomainPool& P = c->pool;
  if (P.paged) throw StatusError(SPRAY_ERR_STATE, "a paged pool keeps no raw meshes");
  const int nd = P.ndomains;
  std::vector<std::string> blocks(nd);
  std::vector<int> failed(nd, 0);
  long long total = 0;
  for (int d = 0; d < nd; ++d) total += P.desc[d].num_tris;
  SPRAY_HIP(hipStreamSynchronize(c->stream));
  const int nthreads = std::max(1, std::min(c->cfg.nthreads, 16));
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads)
  for (int d = 0; d < nd; ++d) {
    if (hipSetDevice(c->device) != hipSuccess) {
      failed[d] = 2;
      continue;
    }
    const HostDomain& hd = c->scene.domains[d];
    const long long nv = domainVerts(P, d), nf = P.desc[d].num_tris;
    HostMesh m;
    m.v.resize((size_t)nv * 3);
    m.f.resize((size_t)nf * 3);
    m.rgb.resize((size_t)nv);
    const DomainDesc& D = P.desc[d];
    if (hipMemcpy(m.v.data(), P.vtx.as<float>() + D.vtx_off * 3, (size_t)nv * 12,
                  hipMemcpyDeviceToHost) != hipSuccess ||
        hipMemcpy(m.f.data(), P.face.as<uint32_t>() + D.face_off * 3, (size_t)nf * 12,
                  hipMemcpyDeviceToHost) != hipSuccess ||
        hipMemcpy(m.rgb.data(), P.rgb.as<uint32_t>() + D.vtx_off, (size_t)nv * 4,
                  hipMemcpyDeviceToHost) != hipSuccess) {
      failed[d] = 2;
      continue;
    }
    char fname[64];
    std::snprintf(fname, sizeof(fname), "%s_d%04d.ply", name.c_str(), d);
    if (!writePlyBinary(dir + "/" + fname, m.v.data(), (size_t)nv, m.f.data(), (size_t)nf,
                        m.rgb.data())) {
      failed[d] = 1;
      continue;
    }
    blocks[d] = sceneDomainBlock(d, fname, hd.mtl, hd.mtl_rgb, hd.eta_a, hd.eta_b, hd.lo, hd.hi, nv,
                                 nf);
  }
  for (int d = 0; d < nd; ++d) {
    if (failed[d] == 2) throw DeviceError("mesh readback failed");
    if (failed[d]) throw StatusError(SPRAY_ERR_IO, "cannot write PLY into " + dir);
  }
  const std::string spath = dir + "/" + name + ".spray";
  FILE* fp = std::fopen(spath.c_str(), "w");
  if (!fp) throw StatusError(SPRAY_ERR_IO, "cannot write " + spath);
  std::fprintf(fp, "# spray_amd scene: %d domains, %lld triangles\n", nd, total);
  for (const HostLight& L : c->scene.lights) std::fprintf(fp, "%s\n", sceneLightLine(L).c_str());
  std::fprintf(fp, "\n");
  for (auto& b : blocks) std::fputs(b.c_str(), fp);
  if (std::fclose(fp) != 0) throw StatusError(SPRAY_ERR_IO, "cannot write " + spath);
}

void plyInfo(Ctx* c, const std::string& path, long long* nv, long long* nf, float lo[3],
             float hi[3]) {
  PlyLayout L;
  std::string err;
  int rc = readPlyHeader(path, &L, &err);
  if (rc != SPRAY_OK) throw StatusError(rc, err);
  HostMesh host;
  if (L.binary) {
    rc = checkBody(path, L, &err);
  } else {
    rc = readPly(path, &host, &err);  // ASCII bodies are parsed on the host
  }
  if (rc != SPRAY_OK) throw StatusError(rc, err);
  hipStream_t s = c->stream;
  struct Bufs {
    DevBuf v, f, rgb, aux;
    ~Bufs() {
      for (DevBuf* b : {&v, &f, &rgb, &aux}) b->release();
    }
  } B;
  // aux: block partials, the final bounds, the face error key
  B.aux.reserve((size_t)kBoundBlocks * 24 + 64);
  float* part = B.aux.as<float>();
  float* out = part + (size_t)kBoundBlocks * 6;
  unsigned long long* ekey = (unsigned long long*)(out + 8);
  B.v.reserve((size_t)L.nv * 12 + 64);
  if (L.binary) {
    B.f.reserve((size_t)L.nf * 12 + 64);
    B.rgb.reserve((size_t)L.nv * 4 + 64);
    BodyStager st((size_t)(L.nv * L.vertexStride() + L.nf * L.faceStride()));
    float id[16];
    for (int i = 0; i < 16; ++i) id[i] = (i % 5 == 0) ? 1.f : 0.f;
    SPRAY_HIP(hipMemsetAsync(ekey, 0xff, 8, s));
    st.decode(path, L, id, false, B.v.as<float>(), B.f.as<uint32_t>(), B.rgb.as<uint32_t>(), ekey,
              s, std::max(1, c->cfg.nthreads));
    unsigned long long e = ~0ull;
    SPRAY_HIP(hipMemcpyAsync(&e, ekey, 8, hipMemcpyDeviceToHost, s));
    SPRAY_HIP(hipStreamSynchronize(s));
    if (e != ~0ull) throw StatusError(SPRAY_ERR_FORMAT, faceError(path, e));
  } else {
    SPRAY_HIP(hipMemcpyAsync(B.v.p, host.v.data(), host.v.size() * 4, hipMemcpyHostToDevice, s));
  }
  // Aabb::merge over every vertex (apps/ply_header_reader.cc:84-88)
  const int nb = gridOf(L.nv, kBlk, kBoundBlocks);
  kPointBounds<<<nb, kBlk, 0, s>>>(B.v.as<float>(), L.nv, part);
  kBoundsFinal<<<1, kBlk, 0, s>>>(part, nb, out);
  SPRAY_HIP(hipGetLastError());
  float b[6];
  SPRAY_HIP(hipMemcpyAsync(b, out, 24, hipMemcpyDeviceToHost, s));
  SPRAY_HIP(hipStreamSynchronize(s));
  *nv = L.nv;
  *nf = L.nf;
  for (int k = 0; k < 3; ++k) {
    lo[k] = b[k];
    hi[k] = b[3 + k];
  }
}

}  // namespace spray
