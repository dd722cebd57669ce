// Synthetic scene generator: per-domain marching-cubes isosurfaces of
// f(x) = sin(a(x+phase))... written as SpRay domains (binary little-endian
// PLY with float xyz + uchar rgb and `list uchar int` faces, plus a .spray
// file with exact bound/vertex/face lines, the layout the reference's
// io/scene_loader.cc:315-358 and io/ply_loader.cc:114-324 read).
//
// The marching-cubes triangle table is derived at start-up rather than
// tabulated: on each cube face the sign-change points are joined by directed
// segments (exit crossing -> entry crossing, walking the face boundary
// counter-clockwise seen from outside; on an ambiguous face each inside
// corner is cut off separately).  Adjacent cubes make the same choice on a
// shared face, so chaining the segments into loops and fanning each loop
// gives a watertight, consistently oriented surface.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "dmath.h"
#include "spray_internal.h"

namespace spray {
namespace {

struct McTable {
  // up to 8 triangles per case, as 3 local edge ids each; -1 terminated
  int8_t tri[256][25];
  McTable() {
    // corners: bit0 = x, bit1 = y, bit2 = z
    // edges: (corner with 0 on `axis`, axis); id = axis*4 + k
    int edge_of[8][8];
    for (int a = 0; a < 8; ++a)
      for (int b = 0; b < 8; ++b) edge_of[a][b] = -1;
    for (int axis = 0; axis < 3; ++axis) {
      int k = 0;
      for (int c = 0; c < 8; ++c) {
        if (c & (1 << axis)) continue;
        int d = c | (1 << axis);
        edge_of[c][d] = edge_of[d][c] = axis * 4 + k;
        ++k;
      }
    }
    // faces: counter-clockwise corner order seen from outside
    static const int face[6][4] = {
        {0, 4, 6, 2},  // x = 0, normal -x
        {1, 3, 7, 5},  // x = 1, normal +x
        {0, 1, 5, 4},  // y = 0, normal -y
        {2, 6, 7, 3},  // y = 1, normal +y
        {0, 2, 3, 1},  // z = 0, normal -z
        {4, 5, 7, 6},  // z = 1, normal +z
    };
    for (int cs = 0; cs < 256; ++cs) {
      int next[12];  // segment start edge -> end edge
      for (int e = 0; e < 12; ++e) next[e] = -1;
      for (int fi = 0; fi < 6; ++fi) {
        int cr_edge[4], cr_exit[4], ncr = 0;
        for (int i = 0; i < 4; ++i) {
          int a = face[fi][i], b = face[fi][(i + 1) & 3];
          bool ia = (cs >> a) & 1, ib = (cs >> b) & 1;
          if (ia != ib) {
            cr_edge[ncr] = edge_of[a][b];
            cr_exit[ncr] = ia ? 1 : 0;  // inside -> outside along the walk
            ++ncr;
          }
        }
        for (int i = 0; i < ncr; ++i) {
          if (!cr_exit[i]) continue;
          int prev = (i + ncr - 1) % ncr;  // the entry that opened this inside arc
          next[cr_edge[i]] = cr_edge[prev];
        }
      }
      int ntri = 0;
      bool used[12] = {false};
      for (int e0 = 0; e0 < 12; ++e0) {
        if (next[e0] < 0 || used[e0]) continue;
        int loop[12], nl = 0;
        int e = e0;
        while (!used[e]) {
          used[e] = true;
          loop[nl++] = e;
          e = next[e];
        }
        for (int i = 1; i + 1 < nl; ++i) {
          tri[cs][ntri * 3 + 0] = (int8_t)loop[0];
          tri[cs][ntri * 3 + 1] = (int8_t)loop[i + 1];
          tri[cs][ntri * 3 + 2] = (int8_t)loop[i];
          ++ntri;
        }
      }
      tri[cs][ntri * 3] = -1;
    }
  }
};

const McTable& table() {
  static McTable t;
  return t;
}

}  // namespace

const int8_t* mcTriTable() { return &table().tri[0][0]; }

// local edge id -> (corner with 0 on axis, axis)
void mcEdgeCorner(int e, int* dx, int* dy, int* dz, int* axis) {
  *axis = e / 4;
  int k = e % 4;
  int c = 0, m = 0;
  for (int cc = 0; cc < 8; ++cc) {
    if (cc & (1 << *axis)) continue;
    if (m == k) {
      c = cc;
      break;
    }
    ++m;
  }
  *dx = c & 1;
  *dy = (c >> 1) & 1;
  *dz = (c >> 2) & 1;
}

namespace {

uint64_t splitmix64(uint64_t& x) {
  uint64_t z = (x += 0x9e3779b97f4a7c15ull);
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  return z ^ (z >> 31);
}
double uniform01(uint64_t& s) { return (double)(splitmix64(s) >> 11) * (1.0 / 9007199254740992.0); }

}  // namespace

void mcDomainParams(const spray_scenegen_params& p, int d, double* a, double* b, int* res) {
  *a = p.a;
  *b = p.b;
  *res = p.res;
  if (!p.per_domain_random) return;
  // log-uniform triangle budget per domain, rescaled to the requested total
  int nd = p.nx * p.ny * p.nz;
  double total = 0, mine = 0;
  for (int i = 0; i < nd; ++i) {
    uint64_t s = (p.seed ^ (uint64_t)i);
    double u = uniform01(s);
    double T = std::exp(std::log(1e4) + u * (std::log(2e6) - std::log(1e4)));
    total += T;
    if (i == d) mine = T;
  }
  if (p.tri_total > 0) mine *= p.tri_total / total;
  uint64_t s = (p.seed ^ (uint64_t)d);
  (void)uniform01(s);
  *a = 4.0 + 4.0 * uniform01(s);
  *b = 12.0 + 8.0 * uniform01(s);
  int r = (int)std::lround(std::sqrt(mine / 19.5));
  *res = r < 4 ? 4 : r;
}

void mcGenerate(const spray_scenegen_params& p, int d, HostMesh* m) {
  const McTable& T = table();
  int di = d % p.nx, dj = (d / p.nx) % p.ny, dk = d / (p.nx * p.ny);
  double a, b;
  int n;
  mcDomainParams(p, d, &a, &b, &n);
  const int N = n + 1;
  const double h = 1.0 / n;
  auto X = [&](int i, double o) { return o + (double)i * h; };
  std::vector<double> f((size_t)N * N * N);
  for (int k = 0; k < N; ++k)
    for (int j = 0; j < N; ++j)
      for (int i = 0; i < N; ++i) {
        double x = X(i, di), y = X(j, dj), z = X(k, dk);
        f[((size_t)k * N + j) * N + i] =
            sinD(a * x + p.phase) * sinD(a * y) * sinD(a * z) + 0.3 * sinD(b * (x + y + z) + p.phase);
      }
  auto F = [&](int i, int j, int k) { return f[((size_t)k * N + j) * N + i]; };
  // edges along axis: count of edges per axis = n * N * N
  const size_t E = (size_t)n * N * N;
  auto eidx = [&](int axis, int i, int j, int k) -> size_t {
    // the axis coordinate ranges over n, the others over N
    if (axis == 0) return ((size_t)k * N + j) * n + i;
    if (axis == 1) return ((size_t)k * n + j) * N + i;
    return ((size_t)k * N + j) * N + i;
  };
  std::vector<int32_t> vid(3 * E, -1);
  m->v.clear();
  for (int axis = 0; axis < 3; ++axis) {
    int ni = axis == 0 ? n : N, nj = axis == 1 ? n : N, nk = axis == 2 ? n : N;
    for (int k = 0; k < nk; ++k)
      for (int j = 0; j < nj; ++j)
        for (int i = 0; i < ni; ++i) {
          int i1 = i + (axis == 0), j1 = j + (axis == 1), k1 = k + (axis == 2);
          double f0 = F(i, j, k), f1 = F(i1, j1, k1);
          bool in0 = f0 < 0.0, in1 = f1 < 0.0;
          if (in0 == in1) continue;
          double t = (0.0 - f0) / (f1 - f0);
          double pos[3] = {X(i, di), X(j, dj), X(k, dk)};
          pos[axis] += t * h;
          vid[axis * E + eidx(axis, i, j, k)] = (int32_t)(m->v.size() / 3);
          for (int c = 0; c < 3; ++c) m->v.push_back((float)pos[c]);
        }
  }
  m->f.clear();
  int ex[12], ey[12], ez[12], eax[12];
  for (int e = 0; e < 12; ++e) mcEdgeCorner(e, &ex[e], &ey[e], &ez[e], &eax[e]);
  for (int k = 0; k < n; ++k)
    for (int j = 0; j < n; ++j)
      for (int i = 0; i < n; ++i) {
        int cs = 0;
        for (int c = 0; c < 8; ++c)
          if (F(i + (c & 1), j + ((c >> 1) & 1), k + ((c >> 2) & 1)) < 0.0) cs |= 1 << c;
        const int8_t* tr = T.tri[cs];
        for (int q = 0; tr[q] >= 0; q += 3) {
          for (int r = 0; r < 3; ++r) {
            int e = tr[q + r];
            int32_t id = vid[eax[e] * E + eidx(eax[e], i + ex[e], j + ey[e], k + ez[e])];
            m->f.push_back((uint32_t)id);
          }
        }
      }
  m->rgb.assign(m->v.size() / 3, p.rgb);
}


int scenegenWrite(const spray_scenegen_params& p, const char* dir, const char* name,
                  const char* extra_lines, int64_t* tris_out, std::string* err) {
  int nd = p.nx * p.ny * p.nz;
  if (nd <= 0 || (p.res <= 0 && !p.per_domain_random)) {
    *err = "scenegen: empty domain grid or resolution";
    return SPRAY_ERR_INVALID;
  }
  std::vector<std::string> blocks(nd);
  std::vector<int64_t> tris(nd, 0);
  std::vector<int> failed(nd, 0);
#pragma omp parallel for schedule(dynamic, 1)
  for (int d = 0; d < nd; ++d) {
    HostMesh m;
    mcGenerate(p, d, &m);
    size_t nv = m.v.size() / 3, nf = m.f.size() / 3;
    tris[d] = (int64_t)nf;
    char fname[64];
    std::snprintf(fname, sizeof(fname), "%s_d%04d.ply", name, d);
    std::string path = std::string(dir) + "/" + fname;
    if (nf == 0 || !writePlyBinary(path, m.v.data(), nv, m.f.data(), nf, m.rgb.data())) {
      failed[d] = nf == 0 ? 2 : 1;
      continue;
    }
    float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (size_t i = 0; i < nv; ++i)
      for (int c = 0; c < 3; ++c) {
        lo[c] = std::fmin(lo[c], m.v[i * 3 + c]);
        hi[c] = std::fmax(hi[c], m.v[i * 3 + c]);
      }
    const float white[3] = {1.f, 1.f, 1.f};
    blocks[d] = sceneDomainBlock(d, fname, kDiffuse, white, 1.f, 1.f, lo, hi, (long long)nv,
                                 (long long)nf);
  }
  int64_t total = 0;
  for (int d = 0; d < nd; ++d) {
    if (failed[d]) {
      *err = failed[d] == 2 ? "scenegen: a domain has no surface (empty mesh)"
                            : "scenegen: cannot write PLY into " + std::string(dir);
      return failed[d] == 2 ? SPRAY_ERR_INVALID : SPRAY_ERR_IO;
    }
    total += tris[d];
  }
  std::string spath = std::string(dir) + "/" + name + ".spray";
  FILE* fp = std::fopen(spath.c_str(), "w");
  if (!fp) {
    *err = "scenegen: cannot write " + spath;
    return SPRAY_ERR_IO;
  }
  std::fprintf(fp, "# spray_amd synthetic scene: %d domains, %lld triangles\n", nd,
               (long long)total);
  if (extra_lines && extra_lines[0]) std::fprintf(fp, "%s\n", extra_lines);
  std::fprintf(fp, "\n");
  for (auto& b : blocks) std::fputs(b.c_str(), fp);
  std::fclose(fp);
  if (tris_out) *tris_out = total;
  return SPRAY_OK;
}

}  // namespace spray
