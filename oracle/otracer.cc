// Oracle (test infrastructure only): restatement of SpRay's per-rank
// speculative image-scheduled tracer (src/ooc) on the CPU with OpenMP, plus
// an order-independent per-sample tracer used to cross-check it.
#include "otracer.h"

#include <omp.h>

#include <algorithm>
#include <chrono>
#include <deque>
#include <numeric>
#include <stdexcept>
#include <vector>

namespace orc {

// ---------------------------------------------------------------------------
// ooc/ooc_ray.h:28-40 (Ray), :42-46 (RayData).  `tmax` is new: the far
// bound of a shadow ray (the reference's is always infinity, rays.h:398).
struct Ray {
  float org[3];
  int pixid;
  float dir[3];
  int samid;
  float w[3];
  int depth;
  float history[kHistory];
  float tdom;
  int committed;
  int light;
  int occluded;
  float tmax;
};
struct RayData {
  Ray* ray;
  float tdom;
  int dom_depth;
};

// deps/pbrt/memory.h:67-158 semantics: bump allocation with stable
// addresses, Reset() recycles the blocks.
class Arena {
 public:
  Ray* alloc() {
    if (blk_ == blocks_.size()) blocks_.emplace_back(kBlock);
    Ray* r = &blocks_[blk_][pos_];
    if (++pos_ == kBlock) {
      ++blk_;
      pos_ = 0;
    }
    return r;
  }
  void reset() {
    blk_ = 0;
    pos_ = 0;
  }

 private:
  static constexpr size_t kBlock = 4096;
  std::vector<std::vector<Ray>> blocks_;
  size_t blk_ = 0, pos_ = 0;
};

// ---------------------------------------------------------------------------
// Scene context shared by both tracers
struct Ctx {
  const Scene* s;
  TraceConfig cfg;
  Camera cam;
  int ndomains;
  int nlights;     // AO: ao_samples; PT: scene lights (ooc_tracer.inl:54-60)
  int tile_w = 0;  // largest blocking tile
  int tile_h = 0;
  SchedLog* log = nullptr;
};

struct Isect {
  float t, u, v;
  int prim;
  uint32_t color;
  float Ns[3];
};

// rays.h:57-93 (DomainList) + wbvh_embree.cc:126-148 (cbIntersect1):
// all domains whose AABB the ray [0.001, inf) enters, sorted by (t, id).
struct DomHit {
  int id;
  float t;
};
static void domainList(const Ctx& C, const float org[3], const float dir[3],
                       std::vector<DomHit>* out) {
  out->clear();
  for (int i = 0; i < C.ndomains; ++i) {
    const Domain& d = C.s->domains[i];
    float tmin;
    if (intersectAabb(d.lo, d.hi, org, dir, kRayEps, kInf, &tmin)) out->push_back({i, tmin});
  }
  std::sort(out->begin(), out->end(), [](const DomHit& a, const DomHit& b) {
    return (a.t < b.t) || ((a.t == b.t) && (a.id < b.id));
  });
}

// scene.inl:189-199 + trimesh_buffer.cc:327-360
static bool intersectDomain(const Ctx& C, int id, const float org[3], const float dir[3],
                            Isect* is) {
  const Mesh& m = C.s->domains[id].mesh;
  Hit h = closestHit(m, org, dir, kRayEps, kInf);
  if (h.prim < 0) return false;
  is->t = h.t;
  is->u = h.u;
  is->v = h.v;
  is->prim = h.prim;
  interpolate(m, h.prim, h.u, h.v, &is->color, is->Ns);
  return true;
}
// scene.inl:201-209 with the far bound as a parameter (deviation Q3)
static bool occludedDomain(const Ctx& C, int id, const float org[3], const float dir[3],
                           float tmax) {
  return anyHit(C.s->domains[id].mesh, org, dir, kRayEps, tmax);
}

// ooc_ray.h:54-83 (makeShadow) and :85-111 (makeRay).  New rays start with
// history[depth] = inf: the reference leaves that slot uninitialised while
// filterRqs (ooc_tcontext.inl:146) compares against it.
static void makeShadow(const Ray& ray, int light, V3 pos, V3 dir, V3 w, float t, float tmax,
                       Ray* sh) {
  sh->org[0] = pos.x;
  sh->org[1] = pos.y;
  sh->org[2] = pos.z;
  sh->pixid = ray.pixid;
  sh->dir[0] = dir.x;
  sh->dir[1] = dir.y;
  sh->dir[2] = dir.z;
  sh->samid = ray.samid;
  sh->w[0] = w.x;
  sh->w[1] = w.y;
  sh->w[2] = w.z;
  sh->depth = ray.depth + 1;
  for (int d = 0; d < ray.depth; ++d) sh->history[d] = ray.history[d];
  sh->history[ray.depth] = t;
  for (int d = ray.depth + 1; d < kHistory; ++d) sh->history[d] = kInf;
  sh->committed = 0;
  sh->light = light;
  sh->occluded = 0;
  sh->tdom = 0;
  sh->tmax = tmax;
}
static void makeRay(const Ray& in, V3 pos, V3 dir, V3 w, float t, Ray* out, int depth) {
  out->org[0] = pos.x;
  out->org[1] = pos.y;
  out->org[2] = pos.z;
  out->pixid = in.pixid;
  out->dir[0] = dir.x;
  out->dir[1] = dir.y;
  out->dir[2] = dir.z;
  out->samid = in.samid;
  out->w[0] = w.x;
  out->w[1] = w.y;
  out->w[2] = w.z;
  out->depth = depth;
  for (int d = 0; d < in.depth; ++d) out->history[d] = in.history[d];
  out->history[in.depth] = t;
  for (int d = in.depth + 1; d < kHistory; ++d) out->history[d] = kInf;
  out->committed = 0;
  out->light = 0;
  out->occluded = 0;
  out->tdom = 0;
  out->tmax = kInf;
}

// render/reflection.h:126-131 (Reflect), :133-157 (FrDielectric),
// :159-173 (Refract)
static V3 reflectDir(V3 wo, V3 n) { return (-wo + ((2.0f * dot(wo, n)) * n)); }
static float frDielectric(float cosI, float etaI, float etaT, V3 wo, V3 n, V3* wt) {
  float sin2I = std::max(0.0f, 1.0f - (cosI * cosI));
  float eta = etaI / etaT;
  float sin2T = eta * eta * sin2I;
  if (sin2T >= 1.0f) return 1.0f;
  float cosT = std::sqrt(1.0f - sin2T);
  float Rparl = ((etaT * cosI) - (etaI * cosT)) / ((etaT * cosI) + (etaI * cosT));
  float Rperp = ((etaI * cosI) - (etaT * cosT)) / ((etaI * cosI) + (etaT * cosT));
  float fr = (Rparl * Rparl + Rperp * Rperp) / 2.0f;
  *wt = (eta * -wo) + (n * (eta * cosI - cosT));
  return fr;
}
static bool refractDir(float cosI, float etaI, float etaT, V3 wo, V3 n, V3* wt) {
  float sin2I = std::max(0.0f, 1.0f - (cosI * cosI));
  float eta = etaI / etaT;
  float sin2T = eta * eta * sin2I;
  if (sin2T >= 1.0f) return false;
  float cosT = std::sqrt(1.0f - sin2T);
  *wt = (eta * -wo) + (n * (eta * cosI - cosT));
  return true;
}

// reflection.h:202-214
static V3 blinnPhong(float costheta, V3 kd, V3 ks, float shininess, V3 li, V3 wi, V3 n, V3 wo) {
  V3 half = normalize(wi + wo);
  float ndh = clampf(dot(n, half), 0.0f, 1.0f);
  float p = (float)std::pow((double)ndh, (double)shininess);
  V3 cs = ks * p;
  V3 cd = kd * costheta;
  return li * (cd + cs);
}

// Light sampling: light.h:41-89 and the new rect light.  Returns radiance,
// sets wi, pdf and the shadow far bound.
static V3 sampleLight(const Light& L, Rng& rng, V3 pos, V3 nff, V3* wi, float* pdf,
                      float* tmax) {
  if (L.type == kPointLight) {
    *wi = normalize(L.pos - pos);
    *pdf = 1.0f;
    *tmax = kInf;
    return L.radiance;
  }
  float u1 = rngGet1D(rng);
  float u2 = rngGet1D(rng);
  if (L.type == kHemisphereLight) {
    cosineHemisphereSample(u1, u2, nff, wi, pdf);
    *tmax = kInf;
    return L.radiance;
  }
  // rect: uniform point on the parallelogram, solid-angle pdf
  V3 p = L.pos + (u1 * L.e1) + (u2 * L.e2);
  V3 d = p - pos;
  float dist2 = dot(d, d);
  float dist = std::sqrt(dist2);
  *wi = d * (1.0f / dist);
  float cosl = -dot(L.normal, *wi);
  *tmax = dist;
  if (!(cosl > 0.0f)) {
    *pdf = 0.0f;
    return L.radiance;
  }
  *pdf = dist2 / (L.area * cosl);
  return L.radiance;
}

// ooc_shader_ao.h:81-197 and ooc_shader_pt.h:93-227 restated as one
// function.  `shadow(Ray)` and `bounce(Ray, bool pending)` receive the
// spawned rays.
template <typename ShadowFn, typename BounceFn>
static void shade(const Ctx& C, int domain_id, const Ray& rayin, const Isect& is, int ray_depth,
                  ShadowFn shadow, BounceFn bounce) {
  const TraceConfig& cfg = C.cfg;
  V3 dir = v3(rayin.dir[0], rayin.dir[1], rayin.dir[2]);
  V3 pos = v3(rayin.dir[0] * is.t + rayin.org[0], rayin.dir[1] * is.t + rayin.org[1],
              rayin.dir[2] * is.t + rayin.org[2]);
  V3 surf = unpackColor(is.color);
  V3 normal = v3(is.Ns[0], is.Ns[1], is.Ns[2]);
  V3 wo = -dir;
  V3 Lin = v3(rayin.w[0], rayin.w[1], rayin.w[2]);
  float cos_theta_i = dot(wo, normal);
  bool entering = (cos_theta_i > 0.0f);
  V3 nff = entering ? normal : -normal;
  nff = normalize(nff);
  const Material& mtl = C.s->domains[domain_id].mtl;
  bool delta = (mtl.type != kDiffuse);
  int next_virtual_depth = rayin.depth + 1;
  int next_actual_depth = ray_depth + next_virtual_depth;
  V3 wi, Lr;
  float pdf, costheta;

  auto genR2 = [&](V3 d, V3 w) {
    Ray r2;
    bool pending = (next_virtual_depth == kHistory);
    makeRay(rayin, pos, d, w, is.t, &r2, pending ? 0 : next_virtual_depth);
    bounce(r2, pending);
  };

  if (cfg.ao_mode) {
    const float ao_weight = 1.0f / (float)cfg.ao_samples;
    for (int l = 0; l < cfg.ao_samples; ++l) {
      Rng rng = rngInit(rayin.pixid * (l + 1));
      float u1 = rngGet1D(rng), u2 = rngGet1D(rng);
      cosineHemisphereSample(u1, u2, nff, &wi, &pdf);
      costheta = clampf(dot(nff, wi), 0.0f, 1.0f);
      Lr = (Lin * surf) * (((kOneOverPi * costheta) * ao_weight) / pdf);
      if (hasPositive(Lr)) {
        Ray sh;
        makeShadow(rayin, l, pos, wi, Lr, is.t, cfg.ao_max_dist, &sh);
        shadow(sh);
      }
    }
  } else if (!delta) {
    Rng ls = rngInit(rayin.samid * next_actual_depth);
    V3 ks = v3(cfg.ks[0], cfg.ks[1], cfg.ks[2]);
    for (int l = 0; l < C.nlights; ++l) {
      const Light& L = C.s->lights[l];
      int nsamp = L.isArea() ? cfg.ao_samples : 1;
      for (int s = 0; s < nsamp; ++s) {
        float tmax;
        V3 Li = sampleLight(L, ls, pos, nff, &wi, &pdf, &tmax);
        if (pdf > 0.0f) {
          costheta = clampf(dot(nff, wi), 0.0f, 1.0f);
          float scale = L.isArea() ? (1.0f / (pdf * (float)nsamp)) : (1.0f / pdf);
          Lr = (Lin * blinnPhong(costheta, surf, ks, cfg.shininess, Li, wi, nff, wo)) * scale;
          if (hasPositive(Lr)) {
            Ray sh;
            makeShadow(rayin, l, pos, wi, Lr, is.t, tmax, &sh);
            shadow(sh);
          }
        }
      }
    }
  }

  if (next_actual_depth < cfg.bounces) {
    wo = normalize(wo);
    if (delta) {
      if (cos_theta_i != 0.0f) {
        cos_theta_i = clampf(cos_theta_i, -1.0f, 1.0f);
        float abs_cos = std::fabs(cos_theta_i);
        if (!entering) cos_theta_i = abs_cos;
        bool refl = false, trans = false;
        float fr = 1.0f;
        V3 wt = v3(0, 0, 0);
        if (mtl.type == kMirror) {
          refl = true;
          fr = 1.0f;
        } else if (mtl.type == kGlass) {
          float etaI = entering ? mtl.eta_a : mtl.eta_b;
          float etaT = entering ? mtl.eta_b : mtl.eta_a;
          fr = frDielectric(cos_theta_i, etaI, etaT, wo, nff, &wt);
          refl = (fr != 0.0f);
          trans = (fr != 1.0f);
        } else {  // transmission
          float etaI = entering ? mtl.eta_a : mtl.eta_b;
          float etaT = entering ? mtl.eta_b : mtl.eta_a;
          if (refractDir(cos_theta_i, etaI, etaT, wo, nff, &wt)) {
            trans = true;
            fr = 0.0f;
          } else {
            refl = true;
            fr = 1.0f;
          }
        }
        if (refl) {
          wi = normalize(reflectDir(wo, nff));
          Lr = Lin * (fr / abs_cos);
          if (hasPositive(Lr)) genR2(wi, Lr);
        }
        if (trans) {
          // reference aborts when both are set (ooc_shader_pt.h:197); both are spawned here
          wi = normalize(wt);
          Lr = Lin * ((1.0f - fr) / abs_cos);
          if (hasPositive(Lr)) genR2(wi, Lr);
        }
      }
    } else {
      Rng rng = rngInit(rayin.samid * next_actual_depth);
      float u1 = rngGet1D(rng), u2 = rngGet1D(rng);
      cosineHemisphereSample(u1, u2, nff, &wi, &pdf);
      costheta = clampf(dot(nff, wi), 0.0f, 1.0f);
      Lr = (((Lin * surf) * kOneOverPi) * costheta) / pdf;
      if (hasPositive(Lr)) {
        if (cfg.rr_start > 0 && next_actual_depth >= cfg.rr_start) {
          // Russian roulette (new): survive with p = min(1, max(Lr)), the
          // draw continues the bounce sampler's stream
          float p = std::max(Lr.x, std::max(Lr.y, Lr.z));
          if (p > 1.0f) p = 1.0f;
          float u = rngGet1D(rng);
          if (u < p) genR2(wi, Lr / p);
        } else {
          genR2(wi, Lr);
        }
      }
    }
  }
}

// ---------------------------------------------------------------------------
// render/tile.cc:317-391 (ImageScheduleTileList::init), single rank
static void tiles(const TraceConfig& cfg, std::vector<int>* ys, std::vector<int>* hs) {
  int64_t num_samples = (int64_t)cfg.image_w * cfg.image_h * cfg.pixel_samples;
  int64_t est = (num_samples + cfg.max_samples_per_rank - 1) / cfg.max_samples_per_rank;
  int ntiles = (int)est;
  int tile_h = cfg.image_h / ntiles;
  if (tile_h <= 0) throw std::runtime_error("blocking tile height is zero");
  ys->clear();
  hs->clear();
  for (int y = 0; y < cfg.image_h; y += tile_h) {
    ys->push_back(y);
    hs->push_back(std::min(tile_h, cfg.image_h - y));
  }
}
int numTiles(const TraceConfig& cfg) {
  std::vector<int> ys, hs;
  tiles(cfg, &ys, &hs);
  return (int)ys.size();
}
void tileRect(const TraceConfig& cfg, int i, int* x, int* y, int* w, int* h) {
  std::vector<int> ys, hs;
  tiles(cfg, &ys, &hs);
  *x = 0;
  *y = ys.at(i);
  *w = cfg.image_w;
  *h = hs.at(i);
}

// ooc_tracer.inl:84-172 (genSingleEyes / genMultiEyes) for one eye ray
static void genEye(const Ctx& C, int tile_y, int tile_w, int bufid, Ray* ray) {
  int spp = C.cfg.pixel_samples;
  int pix_local = bufid / spp;
  int s = bufid % spp;
  (void)s;
  int y0 = pix_local / tile_w, x0 = pix_local % tile_w;
  int x = x0, y = tile_y + y0;
  ray->org[0] = C.cam.pos.x;
  ray->org[1] = C.cam.pos.y;
  ray->org[2] = C.cam.pos.z;
  ray->pixid = y * C.cfg.image_w + x;
  V3 d;
  if (spp > 1) {
    Rng rng = rngInit(bufid);
    float fx = (float)x + rngGet1D(rng);
    float fy = (float)y + rngGet1D(rng);
    d = C.cam.direction(fx, fy);
  } else {
    d = C.cam.direction((float)x, (float)y);
  }
  ray->dir[0] = d.x;
  ray->dir[1] = d.y;
  ray->dir[2] = d.z;
  ray->samid = bufid;
  ray->w[0] = ray->w[1] = ray->w[2] = 1.f;
  ray->depth = 0;
  ray->history[0] = kInf;
  for (int k = 1; k < kHistory; ++k) ray->history[k] = kInf;
  ray->committed = 0;
  ray->light = 0;
  ray->occluded = 0;
  ray->tdom = 0;
  ray->tmax = kInf;
}

// ---------------------------------------------------------------------------
// ooc/ooc_domain_stats.{h,cc}.  Depths >= 16 share one overflow slot whose
// weight max(0, 16 - d) is zero (deviation Q2: the reference indexes past
// its 16 slots).
class DomainStats {
 public:
  static constexpr int kSlots = kDomainListSize + 1;
  void resize(int n) {
    n_ = n;
    stats_.assign((size_t)n * kSlots, 0);
  }
  void reset() { std::fill(stats_.begin(), stats_.end(), 0); }
  static int slot(int depth) { return depth < kDomainListSize ? depth : kDomainListSize; }
  void increment(int id, int depth) { ++stats_[(size_t)id * kSlots + slot(depth)]; }
  // ooc_domain_stats.h:73-77 CHECK_GT(s, 0): counted, reported by trace()
  void decrement(int id, int depth) {
    int64_t& s = stats_[(size_t)id * kSlots + slot(depth)];
    if (s <= 0) ++violations;
    --s;
  }
  uint64_t violations = 0;
  void add(int id, const DomainStats& o) {
    for (int d = 0; d < kSlots; ++d) stats_[(size_t)id * kSlots + d] += o.stats_[(size_t)id * kSlots + d];
  }
  int64_t total() const {
    int64_t s = 0;
    for (int64_t v : stats_) s += v;
    return s;
  }
  // evalScores + sort descending (ties by id: deviation Q12) + traversal order
  void schedule(std::vector<int>* order, std::vector<int64_t>* scores = nullptr) const {
    std::vector<int64_t> score(n_, 0);
    for (int id = 0; id < n_; ++id) {
      int64_t sc = 0;
      for (int d = 0; d < kSlots; ++d) sc += stats_[(size_t)id * kSlots + d] * std::max(0, kDomainListSize - d);
      score[id] = sc;
    }
    if (scores) *scores = score;
    order->resize(n_);
    std::iota(order->begin(), order->end(), 0);
    std::stable_sort(order->begin(), order->end(),
                     [&](int a, int b) { return score[a] > score[b]; });
  }

 private:
  int n_ = 0;
  std::vector<int64_t> stats_;
};

// ooc/ooc_vbuf.{h,cc}
class VBuf {
 public:
  void resize(size_t nsamples) { t_.assign(nsamples * kHistory, kInf); }
  void reset() { std::fill(t_.begin(), t_.end(), kInf); }
  bool correct(const Ray& r) const { return correctN(r, r.depth); }
  bool correctN(const Ray& r, int n) const {
    size_t off = (size_t)r.samid * kHistory;
    for (int i = 0; i < n; ++i)
      if (r.history[i] != t_[off + i]) return false;
    return true;
  }
  bool update(float t, Ray* r) {
    const int depth = r->depth;
    int pos = depth;
    size_t off = (size_t)r->samid * kHistory;
    for (int i = 0; i < depth; ++i) {
      float lt = r->history[i], gt = t_[off + i];
      if (lt < gt) {
        pos = i;
        break;
      } else if (lt > gt) {
        pos = -1;
        break;
      }
    }
    bool updated;
    if (pos == -1) {
      updated = false;
    } else if (pos < depth) {
      updated = true;
      r->history[depth] = t;
      for (int i = pos; i <= depth; ++i) t_[off + i] = r->history[i];
    } else {
      if (t < t_[off + depth]) {
        updated = true;
        t_[off + depth] = t;
        r->history[depth] = t;
      } else {
        updated = false;
      }
    }
    if (updated)
      for (int i = depth + 1; i < kHistory; ++i) t_[off + i] = kInf;
    return updated;
  }
  uint64_t countFinite() const {
    uint64_t c = 0;
    for (float v : t_) c += (v != kInf);
    return c;
  }

 private:
  std::vector<float> t_;
};

// ooc/ooc_tcontext.{h,inl}
struct TContext {
  const Ctx* C;
  VBuf vbuf;
  float* image;
  Arena mem0, mem1;
  Arena* mem_in = &mem0;
  Arena* mem_out = &mem1;
  std::vector<std::deque<RayData>> rqs, sqs0, sqs1;
  std::vector<std::deque<RayData>>* sqs_in = &sqs0;
  std::vector<std::deque<RayData>>* sqs_out = &sqs1;
  std::deque<Ray*> sq2, rq2, pending_q, frq, fsq_in, fsq_out, cr0, cr1;
  std::deque<Ray*>* commit_q = &cr0;
  std::deque<Ray*>* retire_q = &cr1;
  DomainStats rstats;
  std::vector<DomHit> dl;
  std::vector<Ray*> census;  // non-pending rays spawned in this window
  TraceStats st;

  void resize(const Ctx* c, float* img, size_t vbuf_samples) {
    C = c;
    image = img;
    vbuf.resize(vbuf_samples);
    rqs.assign(C->ndomains, {});
    sqs0.assign(C->ndomains, {});
    sqs1.assign(C->ndomains, {});
    rstats.resize(C->ndomains);
  }
  bool anyEmpty(const std::vector<std::deque<RayData>>& q) const {
    for (auto& d : q)
      if (!d.empty()) return false;
    return true;
  }
  // ooc_isector.h:126-145 (isectAll)
  void isectAll(Ray* r) {
    domainList(*C, r->org, r->dir, &dl);
    for (size_t i = 0; i < dl.size(); ++i) {
      rqs[dl[i].id].push_back(RayData{r, dl[i].t, (int)i});
      rstats.increment(dl[i].id, (int)i);
    }
  }
  // ooc_isector.h:147-174 (isectWithoutCurrentDomain)
  bool isectExcept(int exclude, Ray* r) {
    domainList(*C, r->org, r->dir, &dl);
    int n = 0;
    for (size_t i = 0; i < dl.size(); ++i) {
      if (dl[i].id == exclude) continue;
      ++n;
      (*sqs_out)[dl[i].id].push_back(RayData{r, dl[i].t, (int)i});
      rstats.increment(dl[i].id, (int)i);
    }
    return n > 0;
  }
  // ooc_tcontext.h:206-216; with the fix of deviation Q1 the parent's whole
  // history is validated before the history buffer is reset.
  void procPendingQ() {
    while (!pending_q.empty()) {
      Ray* r = pending_q.front();
      pending_q.pop_front();
      bool ok = C->cfg.ref_pending_bug ? vbuf.correct(*r) : vbuf.correctN(*r, kHistory);
      if (ok) {
        r->depth = 0;
        r->history[0] = kInf;
        for (int k = 1; k < kHistory; ++k) r->history[k] = kInf;
        ++st.useful_rays;
        isectAll(r);
      }
    }
  }
  // ooc_tcontext.inl:137-170
  void filterQs(int id) {
    auto& rq = rqs[id];
    while (!rq.empty()) {
      RayData d = rq.front();
      rq.pop_front();
      rstats.decrement(id, d.dom_depth);
      if (d.tdom <= d.ray->history[d.ray->depth] && vbuf.correct(*d.ray)) frq.push_back(d.ray);
    }
    filterSqs(id, *sqs_in, &fsq_in);
    filterSqs(id, *sqs_out, &fsq_out);
  }
  void filterSqs(int id, std::vector<std::deque<RayData>>& sqs, std::deque<Ray*>* fsq) {
    auto& sq = sqs[id];
    while (!sq.empty()) {
      RayData d = sq.front();
      sq.pop_front();
      rstats.decrement(id, d.dom_depth);
      if (!d.ray->occluded && vbuf.correct(*d.ray)) fsq->push_back(d.ray);
    }
  }
  bool allFilterQsEmpty() const { return frq.empty() && fsq_in.empty() && fsq_out.empty(); }
  // ooc_tcontext.inl:28-100
  void procFilterQs(int id, int ray_depth) {
    while (!frq.empty()) {
      Ray* r = frq.front();
      frq.pop_front();
      Isect is;
      ++st.radiance_visits;
      if (intersectDomain(*C, id, r->org, r->dir, &is)) {
        if (vbuf.update(is.t, r)) {
          ++st.shades;
          shade(
              *C, id, *r, is, ray_depth,
              [&](const Ray& sh) {
                Ray* p = mem_out->alloc();
                *p = sh;
                census.push_back(p);
                sq2.push_back(p);
              },
              [&](const Ray& b, bool pending) {
                Ray* p = mem_out->alloc();
                *p = b;
                if (pending) {
                  pending_q.push_back(p);
                } else {
                  census.push_back(p);
                  rq2.push_back(p);
                }
              });
          // procShads2
          while (!sq2.empty()) {
            Ray* s = sq2.front();
            sq2.pop_front();
            ++st.shadow_visits;
            if (!occludedDomain(*C, id, s->org, s->dir, s->tmax)) {
              if (!isectExcept(id, s)) {
                s->committed = 1;
                commit_q->push_back(s);
              }
            }
          }
          // procRads2
          while (!rq2.empty()) {
            Ray* b = rq2.front();
            rq2.pop_front();
            isectAll(b);
          }
        }
      }
    }
    procShads(id, &fsq_in, retire_q);
    procShads(id, &fsq_out, commit_q);
  }
  void procShads(int id, std::deque<Ray*>* qin, std::deque<Ray*>* qout) {
    while (!qin->empty()) {
      Ray* r = qin->front();
      qin->pop_front();
      ++st.shadow_visits;
      bool occ = occludedDomain(*C, id, r->org, r->dir, r->tmax);
      if (occ) r->occluded = 1;
      if (!occ && !r->committed) {
        r->committed = 1;
        qout->push_back(r);
      }
    }
  }
  // ooc_tcontext.inl:123-135
  void retire() {
    double scale = 1.0 / (double)C->cfg.pixel_samples;
    while (!retire_q->empty()) {
      Ray* r = retire_q->front();
      retire_q->pop_front();
      if (!r->occluded && vbuf.correct(*r)) {
        float* px = &image[(size_t)r->pixid * 4];
        for (int k = 0; k < 3; ++k) px[k] = (float)((double)px[k] + scale * (double)r->w[k]);
      }
    }
  }
  void swapQs() {
    std::swap(commit_q, retire_q);
    std::swap(sqs_in, sqs_out);
  }
  // end-of-window census of useful rays (non-stale spawned rays)
  void censusWindow() {
    for (Ray* r : census)
      if (vbuf.correct(*r)) ++st.useful_rays;
    census.clear();
    st.final_hits += vbuf.countFinite();
  }
};

// ooc/ooc_pcontext.h:91-181 (isectPrims), run by every thread of the team
static void isectPrims(const Ctx& C, std::vector<TContext>& tcs, DomainStats& merged,
                       std::vector<int>& order, std::vector<int>& votes, int tid, int tile) {
  TContext& tc = tcs[tid];
  const int nt = (int)tcs.size();
  auto vote = [&](int slot, bool v) {
    votes[slot * nt + tid] = v ? 1 : 0;
#pragma omp barrier
    int s = 0;
    for (int i = 0; i < nt; ++i) s += votes[slot * nt + i];
    return s;
  };
  int ray_depth = 0;
  int window = 0;
  while (true) {
    tc.procPendingQ();
    tc.vbuf.reset();
    int pass = 0;
    while (true) {
      bool sqs_in_empty = tc.anyEmpty(*tc.sqs_in);
      bool rqs_empty = tc.anyEmpty(tc.rqs);
      bool retire_empty = tc.retire_q->empty();
      if (vote(0, sqs_in_empty && rqs_empty && retire_empty) == nt) break;
      if (tid == 0) tc.st.passes++;
      if (vote(1, sqs_in_empty && rqs_empty) < nt) {
#pragma omp single
        merged.reset();
#pragma omp critical(orc_merge_stats)
        for (int id = 0; id < C.ndomains; ++id) merged.add(id, tc.rstats);
#pragma omp barrier
#pragma omp single
        {
          SchedLog* L = C.log;
          if (L && L->records() < L->max_records) {
            std::vector<int64_t> sc;
            merged.schedule(&order, &sc);
            L->meta.insert(L->meta.end(), {tile, window, pass, (int)merged.total()});
            L->scores.insert(L->scores.end(), sc.begin(), sc.end());
            L->order.insert(L->order.end(), order.begin(), order.end());
          } else {
            merged.schedule(&order);
          }
        }
        for (int i = 0; i < C.ndomains; ++i) {
          int id = order[i];
          tc.filterQs(id);
          if (vote(2, tc.allFilterQsEmpty()) < nt) tc.procFilterQs(id, ray_depth);
#pragma omp barrier
        }
      }
#pragma omp critical(orc_retire)
      tc.retire();
      tc.swapQs();
      ++pass;
    }
    tc.censusWindow();
    tc.mem_in->reset();
    std::swap(tc.mem_in, tc.mem_out);
    ray_depth += kHistory;
    ++window;
    if (tid == 0) tc.st.windows++;
    if (vote(3, tc.pending_q.empty()) == nt) break;
  }
}

static void traceSpeculative(const Ctx& C, float* rgba, TraceStats* out) {
  std::vector<int> ys, hs;
  tiles(C.cfg, &ys, &hs);
  int nt = C.cfg.nthreads;
  std::vector<TContext> tcs(nt);
  size_t vbuf_samples = (size_t)C.tile_w * C.tile_h * C.cfg.pixel_samples;
  for (auto& tc : tcs) tc.resize(&C, rgba, vbuf_samples);
  DomainStats merged;
  merged.resize(C.ndomains);
  std::vector<int> order(C.ndomains);
  std::vector<int> votes(4 * nt);
  std::vector<Ray> eyes;
  omp_set_dynamic(0);
  int t0 = std::max(0, C.cfg.tile_begin), t1 = std::min((int)ys.size(), C.cfg.tile_end);
  uint64_t eye_total = 0;
#pragma omp parallel num_threads(nt)
  {
    int tid = omp_get_thread_num();
    TContext& tc = tcs[tid];
    for (int ti = t0; ti < t1; ++ti) {
#pragma omp barrier
      int tile_h = hs[ti], tile_w = C.cfg.image_w;
      size_t neyes = (size_t)tile_w * tile_h * C.cfg.pixel_samples;
#pragma omp single
      {
        eyes.resize(neyes);
        eye_total += neyes;
      }
      tc.mem_in->reset();
      tc.mem_out->reset();
#pragma omp for schedule(static, 1)
      for (long i = 0; i < (long)neyes; ++i) genEye(C, ys[ti], tile_w, (int)i, &eyes[i]);
      tc.rstats.reset();
#pragma omp for schedule(static, 1)
      for (long i = 0; i < (long)neyes; ++i) tc.isectAll(&eyes[i]);
      isectPrims(C, tcs, merged, order, votes, tid, ti);
    }
  }
  for (auto& tc : tcs)
    if (tc.rstats.violations) throw std::runtime_error("DomainStats count went negative");
  for (auto& tc : tcs) {
    out->useful_rays += tc.st.useful_rays;
    out->radiance_visits += tc.st.radiance_visits;
    out->shadow_visits += tc.st.shadow_visits;
    out->shades += tc.st.shades;
    out->final_hits += tc.st.final_hits;
    out->passes += tc.st.passes;
    out->windows += tc.st.windows;
  }
  out->eye_rays = eye_total;
  out->useful_rays += eye_total;
  out->tiles = (uint64_t)(t1 - t0);
}

// ---------------------------------------------------------------------------
// Order-independent tracer: per sample, the closest hit over all domains at
// each depth, shading exactly the hits the speculative tracer keeps.
static void traceCanonical(const Ctx& C, float* rgba, TraceStats* out) {
  std::vector<int> ys, hs;
  tiles(C.cfg, &ys, &hs);
  int t0 = std::max(0, C.cfg.tile_begin), t1 = std::min((int)ys.size(), C.cfg.tile_end);
  uint64_t useful = 0, eyes_n = 0, rvis = 0, svis = 0, hits = 0;
  const double scale = 1.0 / (double)C.cfg.pixel_samples;
  for (int ti = t0; ti < t1; ++ti) {
    int tile_w = C.cfg.image_w, tile_h = hs[ti];
    long neyes = (long)tile_w * tile_h * C.cfg.pixel_samples;
    eyes_n += neyes;
#pragma omp parallel for schedule(dynamic, 256) num_threads(C.cfg.nthreads) \
    reduction(+ : useful, rvis, svis, hits)
    for (long i = 0; i < neyes; ++i) {
      std::vector<DomHit> dl;
      Ray eye;
      genEye(C, ys[ti], tile_w, (int)i, &eye);
      std::vector<Ray> cur{eye};
      int ray_depth = 0;  // actual depth of window start
      while (!cur.empty()) {
        // closest hit among the current candidates (children of one hit)
        int best_k = -1, best_dom = -1;
        Isect best{};
        for (size_t k = 0; k < cur.size(); ++k) {
          Ray& r = cur[k];
          domainList(C, r.org, r.dir, &dl);
          for (auto& dh : dl) {
            if (best_k >= 0 && dh.t > best.t) break;
            Isect is;
            ++rvis;
            if (intersectDomain(C, dh.id, r.org, r.dir, &is)) {
              if (best_k < 0 || is.t < best.t) {
                best = is;
                best_k = (int)k;
                best_dom = dh.id;
              }
            }
          }
        }
        if (best_k < 0) break;
        ++hits;
        Ray r = cur[best_k];
        r.history[r.depth] = best.t;
        std::vector<Ray> next;
        bool next_pending = false;
        shade(
            C, best_dom, r, best, ray_depth,
            [&](const Ray& sh) {
              ++useful;
              bool occ = occludedDomain(C, best_dom, sh.org, sh.dir, sh.tmax);
              ++svis;
              if (!occ) {
                domainList(C, sh.org, sh.dir, &dl);
                for (auto& dh : dl) {
                  if (dh.id == best_dom) continue;
                  ++svis;
                  if (occludedDomain(C, dh.id, sh.org, sh.dir, sh.tmax)) {
                    occ = true;
                    break;
                  }
                }
              }
              if (!occ) {
                float* px = &rgba[(size_t)sh.pixid * 4];
                double add[3] = {scale * (double)sh.w[0], scale * (double)sh.w[1],
                                 scale * (double)sh.w[2]};
#pragma omp atomic
                px[0] += (float)add[0];
#pragma omp atomic
                px[1] += (float)add[1];
#pragma omp atomic
                px[2] += (float)add[2];
              }
            },
            [&](const Ray& b, bool pending) {
              ++useful;
              next.push_back(b);
              next_pending = pending;
            });
        if (next_pending) {
          ray_depth += kHistory;
          for (auto& b : next) {
            b.depth = 0;
            for (int k = 0; k < kHistory; ++k) b.history[k] = kInf;
          }
        }
        cur.swap(next);
      }
    }
  }
  out->useful_rays += useful + eyes_n;
  out->eye_rays += eyes_n;
  out->radiance_visits += rvis;
  out->shadow_visits += svis;
  out->final_hits += hits;
  out->shades += hits;
  out->tiles += (uint64_t)(t1 - t0);
}

void trace(const Scene& s, const TraceConfig& cfg, float* rgba, TraceStats* st, SchedLog* log) {
  Ctx C;
  C.s = &s;
  C.log = log;
  if (log) log->ndomains = (int)s.domains.size();
  C.cfg = cfg;
  C.ndomains = (int)s.domains.size();
  if (C.ndomains <= 0) throw std::runtime_error("empty scene");
  if (cfg.pixel_samples <= 0 || cfg.bounces <= 0 || cfg.image_w <= 0 || cfg.image_h <= 0)
    throw std::runtime_error("invalid configuration");
  if (cfg.nthreads <= 0) throw std::runtime_error("nthreads must be positive");
  if (cfg.ao_mode) {
    C.nlights = cfg.ao_samples;
  } else {  // a scene without lights renders black (ooc_shader_pt.h:114 loops over none)
    C.nlights = (int)s.lights.size();
  }
  C.cam.init(v3(cfg.cam_pos[0], cfg.cam_pos[1], cfg.cam_pos[2]),
             v3(cfg.cam_lookat[0], cfg.cam_lookat[1], cfg.cam_lookat[2]),
             v3(cfg.cam_up[0], cfg.cam_up[1], cfg.cam_up[2]), cfg.fov, cfg.image_w,
             cfg.image_h);
  std::vector<int> ys, hs;
  tiles(cfg, &ys, &hs);
  C.tile_w = cfg.image_w;
  C.tile_h = *std::max_element(hs.begin(), hs.end());  // deviation Q10
  auto t_start = std::chrono::steady_clock::now();
  if (cfg.canonical)
    traceCanonical(C, rgba, st);
  else
    traceSpeculative(C, rgba, st);
  st->seconds +=
      std::chrono::duration<double>(std::chrono::steady_clock::now() - t_start).count();
}

}  // namespace orc
