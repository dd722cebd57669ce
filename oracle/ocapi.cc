// Oracle (test infrastructure only): C ABI used by tests/ through ctypes
// and by bench.py's cpu_baseline leg.  Nothing in the product links it.
#include <omp.h>

#include <algorithm>
#include <cstring>
#include <vector>
#include <memory>
#include <stdexcept>
#include <string>

#include "oscene.h"
#include "otracer.h"

using namespace orc;

namespace {
thread_local std::string g_err;
int setErr(const std::exception& e) {
  g_err = e.what();
  return -1;
}
}  // namespace

extern "C" {

const char* orc_last_error() { return g_err.c_str(); }

// ---- scenes ---------------------------------------------------------------
void* orc_scene_new() { return new Scene(); }
void orc_scene_free(void* h) { delete (Scene*)h; }

// Loads a .spray file and every PLY it names (infinite-cache warm-up,
// render/scene.inl:85-93), computes normals and builds the oracle BVHs.
int orc_scene_load(void* h, const char* spray_file, const char* ply_path, int nthreads) {
  try {
    Scene* s = (Scene*)h;
    loadSceneFile(spray_file, ply_path ? ply_path : "", s);
    mergeBounds(s);
    int n = (int)s->domains.size();
    std::string err;
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads)
    for (int i = 0; i < n; ++i) {
      try {
        Domain& d = s->domains[i];
        size_t nv, nf;
        loadPly(d.filename, &d.mesh, &nv, &nf);
        finishMesh(&d);
        buildBvh(&d.mesh);
      } catch (const std::exception& e) {
#pragma omp critical(orc_load_err)
        err = e.what();
      }
    }
    if (!err.empty()) throw std::runtime_error(err);
    return 0;
  } catch (const std::exception& e) {
    return setErr(e);
  }
}

// Parses only the .spray file (no PLY reads): header-level known answers.
int orc_scene_parse(void* h, const char* spray_file, const char* ply_path) {
  try {
    Scene* s = (Scene*)h;
    loadSceneFile(spray_file, ply_path ? ply_path : "", s);
    mergeBounds(s);
    return 0;
  } catch (const std::exception& e) {
    return setErr(e);
  }
}

int orc_ply_read(const char* file, int64_t* nv, int64_t* nf, int* has_color, float* lo,
                 float* hi) {
  try {
    Mesh m;
    size_t a, b;
    loadPly(file, &m, &a, &b);
    *nv = (int64_t)a;
    *nf = (int64_t)b;
    *has_color = m.has_color ? 1 : 0;
    for (int k = 0; k < 3; ++k) {
      lo[k] = kInf;
      hi[k] = -kInf;
    }
    for (size_t i = 0; i < a; ++i)
      for (int k = 0; k < 3; ++k) {
        lo[k] = std::min(lo[k], m.v[i * 3 + k]);
        hi[k] = std::max(hi[k], m.v[i * 3 + k]);
      }
    return 0;
  } catch (const std::exception& e) {
    return setErr(e);
  }
}

int orc_scene_num_domains(void* h) { return (int)((Scene*)h)->domains.size(); }
int orc_scene_num_lights(void* h) { return (int)((Scene*)h)->lights.size(); }
void orc_scene_bound(void* h, float* lo, float* hi) {
  Scene* s = (Scene*)h;
  for (int k = 0; k < 3; ++k) {
    lo[k] = s->lo[k];
    hi[k] = s->hi[k];
  }
}
int orc_scene_domain(void* h, int i, int64_t* nv, int64_t* nf, float* lo, float* hi,
                     int* mtl) {
  Scene* s = (Scene*)h;
  if (i < 0 || i >= (int)s->domains.size()) return -1;
  Domain& d = s->domains[i];
  *nv = (int64_t)d.num_vertices;
  *nf = (int64_t)d.num_faces;
  for (int k = 0; k < 3; ++k) {
    lo[k] = d.lo[k];
    hi[k] = d.hi[k];
  }
  *mtl = d.mtl.type;
  return 0;
}
int64_t orc_mesh_sizes(void* h, int i, int64_t* nf) {
  Scene* s = (Scene*)h;
  *nf = (int64_t)(s->domains[i].mesh.f.size() / 3);
  return (int64_t)(s->domains[i].mesh.v.size() / 3);
}
void orc_mesh_copy(void* h, int i, float* v, uint32_t* f, uint32_t* rgb, float* n) {
  const Mesh& m = ((Scene*)h)->domains[i].mesh;
  if (v) std::memcpy(v, m.v.data(), m.v.size() * 4);
  if (f) std::memcpy(f, m.f.data(), m.f.size() * 4);
  if (rgb) std::memcpy(rgb, m.rgb.data(), m.rgb.size() * 4);
  if (n) std::memcpy(n, m.n.data(), m.n.size() * 4);
}

// ---- geometry queries ------------------------------------------------------
int orc_closest_hit(void* h, int dom, int64_t n, const float* org, const float* dir,
                    float tnear, float tfar, int brute, float* t, float* u, float* v,
                    int* prim) {
  Scene* s = (Scene*)h;
  if (dom < 0 || dom >= (int)s->domains.size()) return -1;
  const Mesh& m = s->domains[dom].mesh;
#pragma omp parallel for schedule(static, 1024)
  for (int64_t i = 0; i < n; ++i) {
    Hit r = brute ? closestHitBrute(m, &org[i * 3], &dir[i * 3], tnear, tfar)
                  : closestHit(m, &org[i * 3], &dir[i * 3], tnear, tfar);
    t[i] = r.t;
    u[i] = r.u;
    v[i] = r.v;
    prim[i] = r.prim;
  }
  return 0;
}
int orc_any_hit(void* h, int dom, int64_t n, const float* org, const float* dir, float tnear,
                const float* tfar, int* out) {
  Scene* s = (Scene*)h;
  if (dom < 0 || dom >= (int)s->domains.size()) return -1;
  const Mesh& m = s->domains[dom].mesh;
#pragma omp parallel for schedule(static, 1024)
  for (int64_t i = 0; i < n; ++i) out[i] = anyHit(m, &org[i * 3], &dir[i * 3], tnear, tfar[i]) ? 1 : 0;
  return 0;
}
// rays.h:57-93 + wbvh_embree.cc:126-148: sorted (t, id) domain list per ray
int orc_domain_list(void* h, int64_t n, const float* org, const float* dir, int maxk,
                    int* count, int* ids, float* ts) {
  Scene* s = (Scene*)h;
  int nd = (int)s->domains.size();
  for (int64_t i = 0; i < n; ++i) {
    struct DH {
      int id;
      float t;
    };
    std::vector<DH> hits;
    for (int d = 0; d < nd; ++d) {
      float tmin;
      if (intersectAabb(s->domains[d].lo, s->domains[d].hi, &org[i * 3], &dir[i * 3], kRayEps,
                        kInf, &tmin))
        hits.push_back({d, tmin});
    }
    std::sort(hits.begin(), hits.end(), [](const DH& a, const DH& b) {
      return (a.t < b.t) || ((a.t == b.t) && (a.id < b.id));
    });
    count[i] = (int)hits.size();
    for (int k = 0; k < (int)hits.size() && k < maxk; ++k) {
      ids[i * maxk + k] = hits[k].id;
      ts[i * maxk + k] = hits[k].t;
    }
  }
  return 0;
}
int orc_interpolate(void* h, int dom, int64_t n, const int* prim, const float* u,
                    const float* v, uint32_t* color, float* Ns) {
  Scene* s = (Scene*)h;
  const Mesh& m = s->domains[dom].mesh;
  for (int64_t i = 0; i < n; ++i) interpolate(m, prim[i], u[i], v[i], &color[i], &Ns[i * 3]);
  return 0;
}
int orc_tri_intersect(int64_t n, const float* org, const float* dir, const float* tri,
                      float tnear, float tfar, float* t, float* u, float* v, int* hit) {
  for (int64_t i = 0; i < n; ++i) {
    RayPre rp = rayPrecompute(&dir[i * 3]);
    hit[i] = triIntersect(rp, &org[i * 3], &tri[i * 9], &tri[i * 9 + 3], &tri[i * 9 + 6], tnear,
                          tfar, &t[i], &u[i], &v[i])
                 ? 1
                 : 0;
  }
  return 0;
}

// ---- math known answers -------------------------------------------------------
uint32_t orc_rng_seed(int id) { return rngInit(id).s; }
uint32_t orc_murmur_mix(uint32_t h, uint32_t k) { return murmurMix(h, k); }
uint32_t orc_murmur_fmix(uint32_t h) { return murmurFinalize(h); }
void orc_rng_floats(int id, int n, float* out) {
  Rng r = rngInit(id);
  for (int i = 0; i < n; ++i) out[i] = rngGet1D(r);
}
void orc_sincos(int64_t n, const float* x, float* s, float* c) {
  for (int64_t i = 0; i < n; ++i) sincosRef(x[i], &s[i], &c[i]);
}
void orc_concentric(int64_t n, const float* u1, const float* u2, float* dx, float* dy) {
  for (int64_t i = 0; i < n; ++i) concentricDisk(u1[i], u2[i], &dx[i], &dy[i]);
}
void orc_cosine_sample(int64_t n, const float* u1, const float* u2, const float* N, float* wi,
                       float* pdf) {
  for (int64_t i = 0; i < n; ++i) {
    V3 w;
    cosineHemisphereSample(u1[i], u2[i], v3(N[i * 3], N[i * 3 + 1], N[i * 3 + 2]), &w, &pdf[i]);
    wi[i * 3] = w.x;
    wi[i * 3 + 1] = w.y;
    wi[i * 3 + 2] = w.z;
  }
}
void orc_camera_dirs(const float* pos, const float* lookat, const float* up, float fov, int w,
                     int h, int64_t n, const float* xs, const float* ys, float* dirs) {
  Camera cam;
  cam.init(v3(pos[0], pos[1], pos[2]), v3(lookat[0], lookat[1], lookat[2]),
           v3(up[0], up[1], up[2]), fov, w, h);
  for (int64_t i = 0; i < n; ++i) {
    V3 d = cam.direction(xs[i], ys[i]);
    dirs[i * 3] = d.x;
    dirs[i * 3 + 1] = d.y;
    dirs[i * 3 + 2] = d.z;
  }
}

// ---- rendering --------------------------------------------------------------
int orc_num_tiles(const TraceConfig* cfg) {
  try {
    return numTiles(*cfg);
  } catch (const std::exception& e) {
    return setErr(e);
  }
}
int orc_render(void* h, const TraceConfig* cfg, float* rgba, TraceStats* st) {
  try {
    trace(*(Scene*)h, *cfg, rgba, st);
    return 0;
  } catch (const std::exception& e) {
    return setErr(e);
  }
}
// Renders like orc_render and returns the DomainStats schedule of the first
// max_records passes (ooc_domain_stats.cc:60-111): meta[4] = {tile, window,
// pass in window, entries queued}, scores[ndomains], order[ndomains].
int orc_render_schedule(void* h, const TraceConfig* cfg, float* rgba, TraceStats* st, int max_records,
                        int* nrec, int* meta, long long* scores, int* order) {
  try {
    SchedLog L;
    L.max_records = max_records;
    trace(*(Scene*)h, *cfg, rgba, st, &L);
    const int n = L.records(), nd = L.ndomains;
    *nrec = n;
    std::copy(L.meta.begin(), L.meta.end(), meta);
    for (size_t i = 0; i < (size_t)n * nd; ++i) scores[i] = (long long)L.scores[i];
    std::copy(L.order.begin(), L.order.end(), order);
    return 0;
  } catch (const std::exception& e) {
    return setErr(e);
  }
}
int orc_sizeof_config() { return (int)sizeof(TraceConfig); }
int orc_sizeof_stats() { return (int)sizeof(TraceStats); }

}  // extern "C"
