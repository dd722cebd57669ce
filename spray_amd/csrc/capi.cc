// C ABI of the spray_amd library (declared in include/spray_gpu.h).
#include <hip/hip_runtime.h>
#include <omp.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <exception>
#include <string>

#include "ctx.h"

using namespace spray;

namespace {

thread_local std::string g_create_err;
constexpr long long kTraceRaysDeviceMax = 2147483647LL;

template <typename F>
int guarded(spray_ctx* ctx, F&& f) {
  try {
    f();
    if (ctx) ctx->c.err.clear();
    return SPRAY_OK;
  } catch (const StatusError& e) {
    if (ctx) ctx->c.err = e.msg;
    return e.code;
  } catch (const DeviceError& e) {
    if (ctx) ctx->c.err = e.msg;
    return SPRAY_ERR_DEVICE;
  } catch (const std::bad_alloc&) {
    if (ctx) ctx->c.err = "host allocation failed";
    return SPRAY_ERR_OOM;
  } catch (const std::exception& e) {
    if (ctx) ctx->c.err = e.what();
    return SPRAY_ERR_INVALID;
  }
}

void validateConfig(const spray_config& c) {
  if (c.image_w <= 0 || c.image_h <= 0) throw StatusError(SPRAY_ERR_INVALID, "image size must be positive");
  if (c.pixel_samples <= 0) throw StatusError(SPRAY_ERR_INVALID, "pixel_samples must be positive");
  if (c.bounces <= 0) throw StatusError(SPRAY_ERR_INVALID, "bounces must be positive");
  if (c.ao_samples <= 0) throw StatusError(SPRAY_ERR_INVALID, "ao_samples must be positive");
  if (c.max_samples_per_rank <= 0) throw StatusError(SPRAY_ERR_INVALID, "max_samples_per_rank must be positive");
  if (c.schedule < SPRAY_SCHEDULE_PLAIN || c.schedule > SPRAY_SCHEDULE_LEARNED)
    throw StatusError(SPRAY_ERR_INVALID, "unknown schedule");
  if (!(c.fov > 0.f && c.fov < 180.f)) throw StatusError(SPRAY_ERR_INVALID, "fov must be in (0, 180)");
  long long samples = (long long)c.image_w * c.image_h * c.pixel_samples;
  long long ntiles = (samples + c.max_samples_per_rank - 1) / c.max_samples_per_rank;
  if (c.image_h / ntiles <= 0)
    throw StatusError(SPRAY_ERR_INVALID, "max_samples_per_rank too small for one image row");
}

// Rays the reference would hand to Embree with an undefined result (a NaN
// or all-zero direction) are rejected with the first offending index.
void checkRays(const float* dx, const float* dy, const float* dz, size_t stride, size_t n) {
  for (size_t i = 0; i < n; ++i) {
    const float x = dx[i * stride], y = dy[i * stride], z = dz[i * stride];
    if (std::isnan(x) || std::isnan(y) || std::isnan(z) || (x == 0.f && y == 0.f && z == 0.f))
      throw StatusError(SPRAY_ERR_INVALID,
                        "ray " + std::to_string(i) + " has a NaN or zero direction");
  }
}

// render/camera.h:128-166 (Camera::init) on the host
void cameraInit(Ctx* c, const float pos[3], const float look[3], const float up[3], float vfov) {
  const spray_config& cfg = c->cfg;
  float aspect = (float)cfg.image_w / (float)cfg.image_h;
  float theta = (float)((double)vfov * kPiD / 180.0f);
  float half_h = (float)std::tan((double)(theta / 2.0f));
  float half_w = aspect * half_h;
  f3 cp = mk(pos[0], pos[1], pos[2]);
  f3 w = sub(cp, mk(look[0], look[1], look[2]));
  f3 u = cross(mk(up[0], up[1], up[2]), w);
  f3 v = cross(w, u);
  w = normalize(w);
  u = normalize(u);
  v = normalize(v);
  f3 center = sub(cp, w);
  c->cam.lowerleft = sub(sub(center, scl(half_w, u)), scl(half_h, v));
  c->cam.wvec = scl(2.0f * half_w, u);
  c->cam.hvec = scl(2.0f * half_h, v);
  c->cam.pos = cp;
  c->cam.img_w = (float)cfg.image_w;
  c->cam.img_h = (float)cfg.image_h;
  c->camera_set = true;
}

// spray_renderer.inl:86-105 (initCamera)
void defaultCamera(Ctx* c) {
  if (c->cfg.has_camera) {
    cameraInit(c, c->cfg.camera_pos, c->cfg.camera_lookat, c->cfg.camera_up, c->cfg.fov);
    return;
  }
  const HostScene& s = c->scene;
  float ext[3], look[3];
  for (int k = 0; k < 3; ++k) {
    ext[k] = s.hi[k] - s.lo[k];
    look[k] = (float)(0.5 * (double)(s.lo[k] + s.hi[k]));
  }
  float len = std::sqrt(ext[0] * ext[0] + ext[1] * ext[1] + ext[2] * ext[2]);
  float z = len * 0.5f * 1.0f;
  float pos[3] = {look[0] + len * 0.5f * 0.0f, look[1] + len * 0.5f * 0.0f, look[2] + z};
  float up[3] = {0.f, 1.f, 0.f};
  cameraInit(c, pos, look, up, c->cfg.fov);
}

void needScene(Ctx* c) {
  if (!c->loaded) throw StatusError(SPRAY_ERR_STATE, "no scene loaded");
}

}  // namespace

extern "C" {

void spray_config_defaults(spray_config* cfg) {
  std::memset(cfg, 0, sizeof(*cfg));
  cfg->image_w = 400;
  cfg->image_h = 400;
  cfg->has_camera = 0;
  cfg->camera_up[1] = 1.f;
  cfg->fov = 90.f;
  cfg->pixel_samples = 4;
  cfg->ao_samples = 8;
  cfg->ao_mode = 0;
  cfg->bounces = 1;
  cfg->ks[0] = cfg->ks[1] = cfg->ks[2] = 0.f;
  cfg->shininess = 1.f;
  cfg->max_samples_per_rank = 1024 * 1024;
  cfg->nthreads = 1;
  cfg->ao_max_dist = INFINITY;
  cfg->rr_start = 0;
  cfg->schedule = SPRAY_SCHEDULE_PLAIN;
  cfg->device = 0;
}

int spray_create(const spray_config* cfg, spray_ctx** out) {
  g_create_err.clear();
  if (!cfg || !out) {
    g_create_err = "null argument";
    return SPRAY_ERR_INVALID;
  }
  *out = nullptr;
  spray_ctx* ctx = new spray_ctx();
  int rc = guarded(ctx, [&] {
    validateConfig(*cfg);
    ctx->c.cfg = *cfg;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) {
      (void)hipGetLastError();
      throw StatusError(SPRAY_ERR_DEVICE, "no HIP device available");
    }
    if (cfg->device < 0 || cfg->device >= ndev) throw StatusError(SPRAY_ERR_INVALID, "bad device ordinal");
    ctx->c.device = cfg->device;
    SPRAY_HIP(hipSetDevice(cfg->device));
    SPRAY_HIP(hipStreamCreateWithFlags(&ctx->c.stream, hipStreamNonBlocking));
    scratchInit(cfg->device);  // this device's copy of the no-allocation entry point's scratch
  });
  if (rc != SPRAY_OK) {
    g_create_err = ctx->c.err;
    delete ctx;
    return rc;
  }
  *out = ctx;
  return SPRAY_OK;
}

void spray_destroy(spray_ctx* ctx) {
  if (!ctx) return;
  (void)hipSetDevice(ctx->c.device);
  if (ctx->c.stream) (void)hipStreamSynchronize(ctx->c.stream);
  tracerRelease(&ctx->c);
  predictorRelease(&ctx->c);
  ctx->c.pool.release();
  ctx->c.q_in.release();
  ctx->c.q_out.release();
  ctx->c.gen_store.release();
  buildCacheRelease(&ctx->c);
  if (ctx->c.stream) (void)hipStreamDestroy(ctx->c.stream);
  delete ctx;
}

const char* spray_last_error(const spray_ctx* ctx) {
  return ctx ? ctx->c.err.c_str() : g_create_err.c_str();
}

int spray_set_config(spray_ctx* ctx, const spray_config* cfg) {
  if (!ctx || !cfg) return SPRAY_ERR_INVALID;
  return guarded(ctx, [&] {
    validateConfig(*cfg);
    if (cfg->device != ctx->c.device) throw StatusError(SPRAY_ERR_INVALID, "device cannot change");
    ctx->c.cfg = *cfg;
    ctx->c.param_version++;
    if (ctx->c.loaded) defaultCamera(&ctx->c);
  });
}

int spray_load_scene(spray_ctx* ctx, const char* spray_file, const char* ply_path) {
  if (!ctx || !spray_file) return SPRAY_ERR_INVALID;
  return guarded(ctx, [&] {
    Ctx* c = &ctx->c;
    SPRAY_HIP(hipSetDevice(c->device));
    c->loaded = false;
    HostScene scene;
    std::string err;
    int rc = parseSceneFile(spray_file, ply_path ? ply_path : "", &scene, &err);
    if (rc != SPRAY_OK) throw StatusError(rc, err);
    loadScenePly(c, scene);
    c->loaded = true;
    defaultCamera(c);
  });
}

int spray_scenegen_load(spray_ctx* ctx, const spray_scenegen_params* p, const char* name,
                        const char* extra_lines) {
  if (!ctx || !p || !name) return SPRAY_ERR_INVALID;
  return guarded(ctx, [&] {
    Ctx* c = &ctx->c;
    SPRAY_HIP(hipSetDevice(c->device));
    c->loaded = false;
    c->generated = false;
    generateScene(c, *p, name, extra_lines ? extra_lines : "", false);
    c->loaded = true;
    defaultCamera(c);
  });
}

int spray_remesh(spray_ctx* ctx, double phase) {
  if (!ctx) return SPRAY_ERR_INVALID;
  return guarded(ctx, [&] {
    Ctx* c = &ctx->c;
    if (!c->loaded || !c->generated)
      throw StatusError(SPRAY_ERR_STATE, "remeshing needs a scene made by spray_scenegen_load");
    SPRAY_HIP(hipSetDevice(c->device));
    spray_scenegen_params p = c->gen;
    p.phase = phase;
    c->loaded = false;
    generateScene(c, p, c->gen_name, c->gen_extra, true);
    c->loaded = true;
  });
}

int spray_domain_mesh(spray_ctx* ctx, int id, float* v, unsigned int* f, unsigned int* rgb,
                      long long* nverts, long long* nfaces) {
  if (!ctx) return SPRAY_ERR_INVALID;
  return guarded(ctx, [&] {
    needScene(&ctx->c);
    SPRAY_HIP(hipSetDevice(ctx->c.device));
    if (id < 0 || id >= ctx->c.pool.ndomains) throw StatusError(SPRAY_ERR_INVALID, "bad domain id");
    domainMesh(&ctx->c, id, v, f, rgb, nverts, nfaces);
  });
}

int spray_write_scene(spray_ctx* ctx, const char* dir, const char* name) {
  if (!ctx || !dir || !name) return SPRAY_ERR_INVALID;
  return guarded(ctx, [&] {
    needScene(&ctx->c);
    SPRAY_HIP(hipSetDevice(ctx->c.device));
    writeScene(&ctx->c, dir, name);
  });
}

int spray_ply_info(spray_ctx* ctx, const char* path, long long* nverts, long long* nfaces,
                   float lo[3], float hi[3]) {
  if (!ctx || !path || !nverts || !nfaces || !lo || !hi) return SPRAY_ERR_INVALID;
  return guarded(ctx, [&] {
    SPRAY_HIP(hipSetDevice(ctx->c.device));
    plyInfo(&ctx->c, path, nverts, nfaces, lo, hi);
  });
}

int spray_num_domains(const spray_ctx* ctx) {
  if (!ctx || !ctx->c.loaded) return 0;
  return ctx->c.pool.ndomains;
}

int spray_domain_info(const spray_ctx* ctx, int id, long long* nverts, long long* nfaces,
                      float lo[3], float hi[3]) {
  if (!ctx || !ctx->c.loaded) return SPRAY_ERR_STATE;
  if (id < 0 || id >= ctx->c.pool.ndomains) return SPRAY_ERR_INVALID;
  const HostDomain& d = ctx->c.scene.domains[id];
  if (nverts) *nverts = d.num_vertices;
  if (nfaces) *nfaces = d.num_faces;
  for (int k = 0; k < 3; ++k) {
    if (lo) lo[k] = d.lo[k];
    if (hi) hi[k] = d.hi[k];
  }
  return SPRAY_OK;
}

int spray_domain_normals(spray_ctx* ctx, int id, float* out) {
  if (!ctx || !out) return SPRAY_ERR_INVALID;
  return guarded(ctx, [&] {
    needScene(&ctx->c);
    const DomainPool& P = ctx->c.pool;
    if (id < 0 || id >= P.ndomains) throw StatusError(SPRAY_ERR_INVALID, "bad domain id");
    if (P.paged) {  // normals of a paged pool stay on the host
      std::copy(P.h_nrm[id].begin(), P.h_nrm[id].end(), out);
      return;
    }
    SPRAY_HIP(hipMemcpy(out, P.nrm.as<float>() + P.desc[id].vtx_off * 3, (size_t)P.nverts[id] * 12,
                        hipMemcpyDeviceToHost));
  });
}

int spray_set_camera(spray_ctx* ctx, const float pos[3], const float lookat[3], const float up[3],
                     float vfov_deg) {
  if (!ctx || !pos || !lookat || !up) return SPRAY_ERR_INVALID;
  return guarded(ctx, [&] {
    if (!(vfov_deg > 0.f && vfov_deg < 180.f)) throw StatusError(SPRAY_ERR_INVALID, "bad fov");
    ctx->c.cfg.has_camera = 1;
    for (int k = 0; k < 3; ++k) {
      ctx->c.cfg.camera_pos[k] = pos[k];
      ctx->c.cfg.camera_lookat[k] = lookat[k];
      ctx->c.cfg.camera_up[k] = up[k];
    }
    ctx->c.cfg.fov = vfov_deg;
    ctx->c.param_version++;
    cameraInit(&ctx->c, pos, lookat, up, vfov_deg);
  });
}

int spray_trace_rays(spray_ctx* ctx, const spray_rays_soa* rays, size_t n, spray_hits_soa* hits,
                     int mode, int domain) {
  if (!ctx || !rays || !hits) return SPRAY_ERR_INVALID;
  return guarded(ctx, [&] {
    Ctx* c = &ctx->c;
    needScene(c);
    if (domain < 0 || domain >= c->pool.ndomains) throw StatusError(SPRAY_ERR_INVALID, "bad domain id");
    if (mode != SPRAY_QUERY_CLOSEST && mode != SPRAY_QUERY_ANY)
      throw StatusError(SPRAY_ERR_INVALID, "bad query mode");
    if (!hits->prim || (mode == SPRAY_QUERY_CLOSEST && (!hits->t || !hits->u || !hits->v)))
      throw StatusError(SPRAY_ERR_INVALID, "missing output arrays");
    if (n == 0) return;
    checkRays(rays->dx, rays->dy, rays->dz, 1, n);
    SPRAY_HIP(hipSetDevice(c->device));
    c->q_in.reserve(n * 48 + 256);  // 8 input + 4 output floats per ray
    float* d_org = c->q_in.as<float>();
    float* d_dir = d_org + n * 3;
    float* d_tn = d_dir + n * 3;
    float* d_tf = d_tn + n;
    std::vector<float> tmp(n * 3);
    for (size_t i = 0; i < n; ++i) {
      tmp[i * 3] = rays->ox[i];
      tmp[i * 3 + 1] = rays->oy[i];
      tmp[i * 3 + 2] = rays->oz[i];
    }
    SPRAY_HIP(hipMemcpyAsync(d_org, tmp.data(), n * 12, hipMemcpyHostToDevice, c->stream));
    SPRAY_HIP(hipStreamSynchronize(c->stream));
    for (size_t i = 0; i < n; ++i) {
      tmp[i * 3] = rays->dx[i];
      tmp[i * 3 + 1] = rays->dy[i];
      tmp[i * 3 + 2] = rays->dz[i];
    }
    SPRAY_HIP(hipMemcpyAsync(d_dir, tmp.data(), n * 12, hipMemcpyHostToDevice, c->stream));
    SPRAY_HIP(hipMemcpyAsync(d_tn, rays->tnear, n * 4, hipMemcpyHostToDevice, c->stream));
    SPRAY_HIP(hipMemcpyAsync(d_tf, rays->tfar, n * 4, hipMemcpyHostToDevice, c->stream));
    // outputs live after the inputs
    float* d_t = d_tf + n;
    float* d_u = d_t + n;
    float* d_v = d_u + n;
    int* d_p = (int*)(d_v + n);
    poolAcquire(c, domain, c->stream);  // paged pool: the domain's block first
    traceRaysDevice(c, d_org, d_dir, d_tn, d_tf, n, mode, domain, d_t, d_u, d_v, d_p);
    poolRelease(c, domain, c->stream);
    if (mode == SPRAY_QUERY_CLOSEST) {
      SPRAY_HIP(hipMemcpy(hits->t, d_t, n * 4, hipMemcpyDeviceToHost));
      SPRAY_HIP(hipMemcpy(hits->u, d_u, n * 4, hipMemcpyDeviceToHost));
      SPRAY_HIP(hipMemcpy(hits->v, d_v, n * 4, hipMemcpyDeviceToHost));
    }
    SPRAY_HIP(hipMemcpy(hits->prim, d_p, n * 4, hipMemcpyDeviceToHost));
  });
}

int spray_trace_rays_device(spray_ctx* ctx, const float* org, const float* dir, const float* tnear,
                            const float* tfar, size_t n, float* t, float* u, float* v, int* prim,
                            int mode, int domain) {
  if (!ctx) return SPRAY_ERR_INVALID;
  return guarded(ctx, [&] {
    Ctx* c = &ctx->c;
    needScene(c);
    if (domain < 0 || domain >= c->pool.ndomains) throw StatusError(SPRAY_ERR_INVALID, "bad domain id");
    if (mode != SPRAY_QUERY_CLOSEST && mode != SPRAY_QUERY_ANY)
      throw StatusError(SPRAY_ERR_INVALID, "bad query mode");
    if (c->pool.paged)
      throw StatusError(SPRAY_ERR_STATE, "spray_trace_rays_device needs the whole scene resident");
    if (n > (size_t)kTraceRaysDeviceMax)
      throw StatusError(SPRAY_ERR_CAPACITY, "spray_trace_rays_device takes at most 2^31 - 1 rays per call");
    if (n == 0) return;
    if (!org || !dir || !tnear || !tfar || !prim || (mode == SPRAY_QUERY_CLOSEST && (!t || !u || !v)))
      throw StatusError(SPRAY_ERR_INVALID, "missing device arrays");
    SPRAY_HIP(hipSetDevice(c->device));
    traceRaysNoAlloc(c, org, dir, tnear, tfar, n, mode, domain, t, u, v, prim);
  });
}

int spray_domain_list(spray_ctx* ctx, const float* org, const float* dir, size_t n, int maxk,
                      int* counts, int* ids, float* ts) {
  if (!ctx || !org || !dir || !counts || !ids || !ts || maxk <= 0) return SPRAY_ERR_INVALID;
  return guarded(ctx, [&] {
    Ctx* c = &ctx->c;
    needScene(c);
    if (c->pool.ndomains > 1024) throw StatusError(SPRAY_ERR_CAPACITY, "domain list query supports <= 1024 domains");
    if (n == 0) return;
    checkRays(dir, dir + 1, dir + 2, 3, n);
    SPRAY_HIP(hipSetDevice(c->device));
    size_t bytes = n * 24 + n * 4 + n * (size_t)maxk * 8 + 256;
    c->q_in.reserve(bytes);
    float* d_org = c->q_in.as<float>();
    float* d_dir = d_org + n * 3;
    int* d_cnt = (int*)(d_dir + n * 3);
    int* d_ids = d_cnt + n;
    float* d_ts = (float*)(d_ids + n * maxk);
    SPRAY_HIP(hipMemcpyAsync(d_org, org, n * 12, hipMemcpyHostToDevice, c->stream));
    SPRAY_HIP(hipMemcpyAsync(d_dir, dir, n * 12, hipMemcpyHostToDevice, c->stream));
    domainListDevice(c, d_org, d_dir, n, maxk, d_cnt, d_ids, d_ts);
    SPRAY_HIP(hipMemcpy(counts, d_cnt, n * 4, hipMemcpyDeviceToHost));
    SPRAY_HIP(hipMemcpy(ids, d_ids, n * maxk * 4, hipMemcpyDeviceToHost));
    SPRAY_HIP(hipMemcpy(ts, d_ts, n * maxk * 4, hipMemcpyDeviceToHost));
  });
}

int spray_trace_frame(spray_ctx* ctx, float* rgba_out, spray_stats* out) {
  if (!ctx) return SPRAY_ERR_INVALID;
  return guarded(ctx, [&] {
    needScene(&ctx->c);
    SPRAY_HIP(hipSetDevice(ctx->c.device));
    spray_stats st;
    std::memset(&st, 0, sizeof(st));
    tracerTraceFrame(&ctx->c, rgba_out, &st);
    st.build_ms = ctx->c.pool.build_ms;
    if (out) *out = st;
  });
}

int spray_set_schedule(spray_ctx* ctx, int schedule) {
  if (!ctx) return SPRAY_ERR_INVALID;
  return guarded(ctx, [&] {
    if (schedule < SPRAY_SCHEDULE_PLAIN || schedule > SPRAY_SCHEDULE_LEARNED)
      throw StatusError(SPRAY_ERR_INVALID, "unknown schedule");
    ctx->c.cfg.schedule = schedule;
    ctx->c.param_version++;
  });
}

int spray_autotune(spray_ctx* ctx, int* chunk_out, int* tile_out) {
  if (!ctx) return SPRAY_ERR_INVALID;
  return guarded(ctx, [&] {
    needScene(&ctx->c);
    SPRAY_HIP(hipSetDevice(ctx->c.device));
    int ch = 0, tl = 0, cached = 0;
    tracerAutotune(&ctx->c, &ch, &tl, &cached);
    if (chunk_out) *chunk_out = ch;
    if (tile_out) *tile_out = tl;
  });
}

int spray_set_batch_params(spray_ctx* ctx, int chunk, int tile) {
  if (!ctx) return SPRAY_ERR_INVALID;
  return guarded(ctx, [&] {
    // chunk: idle lanes a wave gathers before it fetches queue entries for
    // them at once (1..64); tile: waves per workgroup of the trace kernel
    if (chunk < 1 || chunk > 64) throw StatusError(SPRAY_ERR_INVALID, "chunk must be in [1, 64]");
    if (tile < 1 || tile > 4) throw StatusError(SPRAY_ERR_INVALID, "tile must be in [1, 4]");
    ctx->c.batch_chunk = chunk;
    ctx->c.batch_tile = tile;
  });
}

int spray_set_tiles_in_flight(spray_ctx* ctx, int tiles) {
  if (!ctx) return SPRAY_ERR_INVALID;
  return guarded(ctx, [&] {
    if (tiles < 0 || tiles > 4096) throw StatusError(SPRAY_ERR_INVALID, "tiles must be in [0, 4096]");
    ctx->c.tiles_in_flight = tiles;
  });
}

int spray_set_cache_size(spray_ctx* ctx, int domains) {
  if (!ctx) return SPRAY_ERR_INVALID;
  return guarded(ctx, [&] {
    // render/config.cc:200-202 takes any value; lru_cache.cc:44 rejects 0
    if (domains == 0) throw StatusError(SPRAY_ERR_INVALID, "cache size must not be 0");
    ctx->c.cache_size = domains;
  });
}

int spray_set_pool_headroom(spray_ctx* ctx, double factor) {
  if (!ctx) return SPRAY_ERR_INVALID;
  return guarded(ctx, [&] {
    if (!(factor > 0.0 && factor <= 8.0)) throw StatusError(SPRAY_ERR_INVALID, "headroom must be in (0, 8]");
    ctx->c.pool_headroom = factor;
  });
}

int spray_cache_state(spray_ctx* ctx, int* capacity, int* lru_domains,
                      unsigned long long counters[3]) {
  if (!ctx || !capacity) return SPRAY_ERR_INVALID;
  return guarded(ctx, [&] {
    needScene(&ctx->c);
    const DomainPool& P = ctx->c.pool;
    if (!P.paged) {
      *capacity = P.ndomains;
      if (lru_domains)
        for (int d = 0; d < P.ndomains; ++d) lru_domains[d] = d;
      if (counters) counters[0] = counters[1] = counters[2] = 0;
      return;
    }
    *capacity = P.capacity;
    if (lru_domains) {
      std::vector<std::pair<unsigned long long, int>> order;
      for (int b = 0; b < P.capacity; ++b)
        if (P.block_domain[b] >= 0) order.push_back({P.block_tick[b], P.block_domain[b]});
      std::sort(order.begin(), order.end());
      for (int i = 0; i < P.capacity; ++i) lru_domains[i] = i < (int)order.size() ? order[i].second : -1;
    }
    if (counters) {
      counters[0] = P.hits;
      counters[1] = P.misses;
      counters[2] = P.bytes_in;
    }
  });
}

int spray_set_bvh_builder(spray_ctx* ctx, int builder) {
  if (!ctx) return SPRAY_ERR_INVALID;
  return guarded(ctx, [&] {
    if (builder < 0 || builder > 1) throw StatusError(SPRAY_ERR_INVALID, "builder must be 0 or 1");
    ctx->c.bvh_builder = builder;
  });
}

int spray_profile_kernels(spray_ctx* ctx, int enable) {
  if (!ctx) return SPRAY_ERR_INVALID;
  ctx->c.profile = enable != 0;
  return SPRAY_OK;
}

int spray_schedule_log(spray_ctx* ctx, int max_records) {
  if (!ctx || max_records < 0) return SPRAY_ERR_INVALID;
  return guarded(ctx, [&] {
    if (!ctx->c.loaded) throw StatusError(SPRAY_ERR_STATE, "no scene loaded");
    tracerScheduleLog(&ctx->c, max_records);
  });
}

int spray_schedule_read(spray_ctx* ctx, int* nrec, int* meta, unsigned long long* scores, int* order) {
  if (!ctx || !nrec) return SPRAY_ERR_INVALID;
  *nrec = tracerScheduleRead(&ctx->c, meta, scores, order);
  return SPRAY_OK;
}

int spray_kernel_time(spray_ctx* ctx, long long* launches, double* total_ms, int reset) {
  if (!ctx || !launches || !total_ms) return SPRAY_ERR_INVALID;
  tracerKernelTime(&ctx->c, launches, total_ms, reset);
  return SPRAY_OK;
}

int spray_scenegen_domain(const spray_scenegen_params* p, int domain, float* v, unsigned int* f,
                          unsigned int* rgb, long long* nverts, long long* nfaces) {
  if (!p || domain < 0 || domain >= p->nx * p->ny * p->nz || p->nx <= 0 || p->ny <= 0 || p->nz <= 0)
    return SPRAY_ERR_INVALID;
  if (!p->per_domain_random && p->res <= 0) return SPRAY_ERR_INVALID;
  try {
    HostMesh m;
    mcGenerate(*p, domain, &m);
    if (nverts) *nverts = (long long)m.v.size() / 3;
    if (nfaces) *nfaces = (long long)m.f.size() / 3;
    if (v) std::copy(m.v.begin(), m.v.end(), v);
    if (f) std::copy(m.f.begin(), m.f.end(), f);
    if (rgb) std::copy(m.rgb.begin(), m.rgb.end(), rgb);
    return SPRAY_OK;
  } catch (const std::bad_alloc&) {
    return SPRAY_ERR_OOM;
  }
}

int spray_scenegen_write(const spray_scenegen_params* p, const char* dir, const char* name,
                         const char* extra_lines, long long* triangles_out) {
  if (!p || !dir || !name) return SPRAY_ERR_INVALID;
  std::string err;
  int64_t tris = 0;
  int rc = scenegenWrite(*p, dir, name, extra_lines, &tris, &err);
  if (rc != SPRAY_OK) {
    g_create_err = err;
    return rc;
  }
  if (triangles_out) *triangles_out = tris;
  return SPRAY_OK;
}

// HdrImage::writePpm (display/image.h:183-204): ASCII P3, maximum value
// 1023, rows written from the top (y = h-1) down, each channel
// clamp(v * 1023, 0, 1023) truncated to an integer.
int spray_write_ppm(const float* rgba, int w, int h, const char* path) {
  g_create_err.clear();
  if (!rgba || !path || w <= 0 || h <= 0) {
    g_create_err = "invalid image or path";
    return SPRAY_ERR_INVALID;
  }
  FILE* f = std::fopen(path, "w");
  if (!f) {
    g_create_err = std::string("unable to open ") + path;
    return SPRAY_ERR_IO;
  }
  std::fprintf(f, "P3\n%d %d\n1023\n", w, h);
  for (int y = h - 1; y >= 0; --y) {
    for (int x = 0; x < w; ++x) {
      const float* px = rgba + 4 * ((size_t)y * w + x);
      unsigned c[3];
      for (int k = 0; k < 3; ++k) {
        float v = px[k] * 1023.f;
        v = std::fmin(std::fmax(v, 0.0f), 1023.0f);  // NaN becomes 0
        c[k] = (unsigned)v;
      }
      std::fprintf(f, "%u %u %u\n", c[0], c[1], c[2]);
    }
  }
  if (std::fclose(f) != 0) {
    g_create_err = std::string("write failed: ") + path;
    return SPRAY_ERR_IO;
  }
  return SPRAY_OK;
}

}  // extern "C"
