/*
 * spray_gpu.h - C ABI of the MI355X-native SpRay per-rank tracing path.
 *
 * One spray_ctx stands in for one SpRay rank's ooc tracer: it owns the
 * scene (an HBM-resident domain pool with one LBVH per domain), the ray
 * queues, the history buffer and the image.  Every entry point returns
 * SPRAY_OK or an SPRAY_ERR_* code; the message of the last failure is read
 * with spray_last_error().  Nothing aborts across this boundary (the
 * reference's glog CHECK/LOG(FATAL) paths become status codes).
 *
 * Plain types only: no torch or HIP types cross the boundary.  Host buffers
 * passed in are not retained after a call returns.  A context is used from
 * one host thread (the reference runs MPI_THREAD_FUNNELED,
 * apps/spray_ooc.cc:33-45).
 */
#ifndef SPRAY_GPU_H_
#define SPRAY_GPU_H_

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SPRAY_OK 0
#define SPRAY_ERR_INVALID 1  /* invalid argument or configuration */
#define SPRAY_ERR_IO 2       /* file cannot be opened or written */
#define SPRAY_ERR_FORMAT 3   /* malformed .spray or PLY content */
#define SPRAY_ERR_OOM 4      /* device or host allocation failed */
#define SPRAY_ERR_CAPACITY 5 /* a queue or pool bound was exceeded */
#define SPRAY_ERR_DEVICE 6   /* HIP runtime error or no device */
#define SPRAY_ERR_STATE 7    /* call order (e.g. trace before load) */

#define SPRAY_SCHEDULE_PLAIN 0   /* one launch per scheduled domain (reference order) */
#define SPRAY_SCHEDULE_BATCHED 1 /* persistent kernel over all queues, hipGraph passes */
#define SPRAY_SCHEDULE_LEARNED 2 /* batched + MLP-ordered speculative visits */

#define SPRAY_QUERY_CLOSEST 0 /* rtcIntersect: closest hit */
#define SPRAY_QUERY_ANY 1     /* rtcOccluded: any hit */

typedef struct spray_ctx spray_ctx;

/* Mirrors the ooc-path fields of spray::Config (render/config.h:34-95,
 * defaults render/config.cc:33-73) plus the new knobs marked "new". */
typedef struct spray_config {
  int image_w, image_h;
  int has_camera; /* 0: default framing from the scene bound (spray_renderer.inl:86-105) */
  float camera_pos[3], camera_lookat[3], camera_up[3];
  float fov;            /* vertical, degrees */
  int pixel_samples;    /* --pixel-samples */
  int ao_samples;       /* --ao-samples (also area-light samples in PT) */
  int ao_mode;          /* --ao-mode */
  int bounces;          /* --bounces */
  float ks[3];          /* --blinn ks */
  float shininess;      /* --blinn shininess */
  long long max_samples_per_rank; /* --max-samples-per-rank (blocking tile cap) */
  int nthreads;         /* host threads used while loading PLY files */
  float ao_max_dist;    /* new: AO ray length (reference: unbounded = INFINITY) */
  int rr_start;         /* new: Russian roulette from this bounce on (0 = off) */
  int schedule;         /* new: SPRAY_SCHEDULE_* */
  int device;           /* HIP device ordinal */
} spray_config;

/* Fills cfg with the reference defaults (render/config.cc:33-73), except
 * that unset ks/shininess become 0/1 (the reference leaves them undefined). */
void spray_config_defaults(spray_config* cfg);

typedef struct spray_stats {
  unsigned long long useful_rays;     /* eye + non-stale bounce + non-stale shadow rays */
  unsigned long long eye_rays;
  unsigned long long radiance_visits; /* (ray, domain) closest-hit traversals */
  unsigned long long shadow_visits;   /* (ray, domain) any-hit traversals */
  unsigned long long shades;          /* accepted history updates (shader calls) */
  unsigned long long final_hits;      /* hits on the final history chains */
  unsigned long long passes;
  unsigned long long windows;
  unsigned long long tiles;
  unsigned long long queue_entries;   /* (ray, domain) pairs enqueued */
  unsigned long long deferred_entries;/* learned: pairs sent to the fallback queue */
  unsigned long long pruned_entries;  /* pairs dropped by the tdom / history filter */
  unsigned long long node_visits;     /* BVH nodes fetched by all traversals */
  unsigned long long tri_tests;       /* triangles tested by all traversals */
  double trace_ms;                    /* device time of the frame (hipEvents) */
  double build_ms;                    /* domain-pool build time of the last load */
  unsigned long long cache_hits;      /* domain loads served by the HBM cache (paged pool) */
  unsigned long long cache_misses;    /* domain loads copied from host memory */
  unsigned long long cache_bytes;     /* bytes copied host -> HBM by those loads */
  unsigned long long regrows;         /* frame re-renders after a ray pool or queue overflowed */
} spray_stats;

typedef struct spray_rays_soa { /* n entries each; host memory */
  const float *ox, *oy, *oz;
  const float *dx, *dy, *dz;
  const float *tnear, *tfar;
} spray_rays_soa;

typedef struct spray_hits_soa { /* n entries each; host memory */
  float *t, *u, *v; /* closest: hit t (tfar on a miss) and barycentrics */
  int *prim;        /* closest: primitive id, -1 on a miss; any: 1 if occluded else 0 */
} spray_hits_soa;

/* Scene generator parameters (synthetic marching-cubes domains). */
typedef struct spray_scenegen_params {
  int nx, ny, nz;        /* domain grid; domain (i,j,k) spans [i,i+1]x[j,j+1]x[k,k+1] */
  int res;               /* cells per domain axis */
  double a, b, phase;    /* f = sin(a x + phase) sin(a y) sin(a z) + 0.3 sin(b (x+y+z) + phase) */
  int per_domain_random; /* 1: per-domain a, b and log-uniform triangle budget from seed */
  unsigned long long seed;
  double tri_total;      /* per_domain_random: rescale budgets to this total */
  unsigned int rgb;      /* packed vertex colour r<<16|g<<8|b */
} spray_scenegen_params;

/* ---- context lifecycle (SprayRenderer/Tracer init, spray_renderer.inl:28-84) ---- */
int spray_create(const spray_config* cfg, spray_ctx** out);
void spray_destroy(spray_ctx* ctx);
const char* spray_last_error(const spray_ctx* ctx); /* ctx may be NULL (create errors) */
int spray_set_config(spray_ctx* ctx, const spray_config* cfg);

/* Scene::init (render/scene.inl:30-100): parses the .spray file
 * (io/scene_loader.cc:315-358), loads every PLY (io/ply_loader.cc:179-204)
 * into the HBM domain pool (the infinite-cache warm-up, scene.inl:85-93),
 * computes vertex normals (trimesh_buffer.cc:267-326) and builds one LBVH
 * per domain (replacing rtcCommit, trimesh_buffer.cc:213-224). */
int spray_load_scene(spray_ctx* ctx, const char* spray_file, const char* ply_path);
int spray_num_domains(const spray_ctx* ctx);
/* Domain record: counts and world AABB (render/domain.h:32-44). */
int spray_domain_info(const spray_ctx* ctx, int id, long long* nverts, long long* nfaces,
                      float lo[3], float hi[3]);
/* Copies a domain's vertex normals (3 floats per vertex, as many vertices
 * as spray_domain_mesh reports) computed on the GPU. */
int spray_domain_normals(spray_ctx* ctx, int id, float* out);

/* Camera::init (render/camera.h:128-166). */
int spray_set_camera(spray_ctx* ctx, const float pos[3], const float lookat[3],
                     const float up[3], float vfov_deg);

/* Tracer::trace (ooc/ooc_tracer.inl:185-234) + HdrImage accumulation: renders
 * one frame and copies W*H*4 floats (rgba, row-major, y up) into rgba_out
 * (may be NULL). */
int spray_trace_frame(spray_ctx* ctx, float* rgba_out, spray_stats* out);

/* Scene::intersect / Scene::occluded for a batch of rays against one domain
 * (rtcIntersect / rtcOccluded, render/scene.inl:189-209).  The parity entry
 * point: mode is SPRAY_QUERY_CLOSEST or SPRAY_QUERY_ANY. */
int spray_trace_rays(spray_ctx* ctx, const spray_rays_soa* rays, size_t n,
                     spray_hits_soa* hits, int mode, int domain);

/* No counterpart in the reference: spray_trace_rays for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  Same query and results, on caller-owned
 * DEVICE arrays: org[n*3], dir[n*3] (xyz per ray), tnear[n], tfar[n] in;
 * prim[n] out, and t/u/v[n] out for SPRAY_QUERY_CLOSEST (may be NULL for
 * SPRAY_QUERY_ANY).  It calls no hipMalloc/hipFree and no allocating
 * library routine: the traversal's scratch (overflow stacks of at most
 * 65,536 lanes plus an error word, 29.4 MB) is a static device array of the
 * library, one per GPU, held until the process ends.  Every call on a GPU
 * shares that GPU's array, so calls that overlap on one GPU race; the call
 * returns after its work on the context's stream has finished.  At most
 * 2^31 - 1 rays per call; the scene must be wholly resident (no partial
 * domain cache). */
int spray_trace_rays_device(spray_ctx* ctx, const float* org, const float* dir, const float* tnear,
                            const float* tfar, size_t n, float* t, float* u, float* v, int* prim,
                            int mode, int domain);

/* Isector::isectDomains (ooc/ooc_isector.h:116-124 -> WbvhEmbree,
 * render/wbvh_embree.cc:126-148): per-ray domain list sorted by (t, id).
 * counts[n]; ids/ts[n*maxk] (entries beyond maxk are dropped). */
int spray_domain_list(spray_ctx* ctx, const float* org, const float* dir, size_t n, int maxk,
                      int* counts, int* ids, float* ts);

/* Selects the speculative schedule (new). */
int spray_set_schedule(spray_ctx* ctx, int schedule);

/* Batched form: autotunes the queue chunk (how many lanes of a wave must be
 * idle before the wave fetches that many queue entries with one atomic,
 * 1..64) and the wavefront tile (waves per workgroup of the persistent trace
 * kernel, 1..4) on a tuning frame (the loaded scene and current camera at
 * half the samples per pixel); the choice is kept in the context and in a
 * cache file (SPRAY_AUTOTUNE_CACHE, else ~/.cache/spray_amd/autotune.txt)
 * keyed by the device, the frame settings and the scene, so a later call
 * with the same key tunes nothing. */
int spray_autotune(spray_ctx* ctx, int* chunk_out, int* tile_out);
int spray_set_batch_params(spray_ctx* ctx, int chunk, int tile);

/* Batched forms (new): how many consecutive blocking tiles (tile.cc:317-391)
 * share every pass; each keeps its own sample numbering, so the image does
 * not depend on it.  0 (default) groups tiles up to about 8M samples. */
int spray_set_tiles_in_flight(spray_ctx* ctx, int tiles);

/* BVH builder used by the next spray_load_scene: 0 = Karras LBVH (default),
 * 1 = PLOC clustering over Morton-ordered triangles.  Stands in for the
 * Embree scene build-quality flags of the per-domain scenes
 * (render/trimesh_buffer.cc:112, RTC_SCENE_DYNAMIC); hits do not depend on
 * the choice, only traversal cost does. */
int spray_set_bvh_builder(spray_ctx* ctx, int builder);

/* Domain cache size for the next spray_load_scene, in domains
 * (render/config.cc:200-202 --cache-size; lru_cache.cc:41-171): < 0 or
 * >= the domain count keeps every domain in HBM (the infinite cache); a
 * smaller value keeps that many domains in HBM, least recently used evicted,
 * every domain's image in pinned host memory.  With a partial cache the
 * plain schedule sweeps the domains one by one (the reference's domain
 * loop) and the batched schedule traces chunks of consecutive scheduled
 * domains that fit in half the cache while the next chunk loads; the learned
 * schedule needs the whole scene resident. */
int spray_set_cache_size(spray_ctx* ctx, int domains);

/* Ray-pool headroom (new; test hook): the factor on the tracer's initial
 * estimate of its ray pools and queues (default 1).  A frame that overflows
 * them is rendered again with doubled capacities, up to 8 times the
 * estimate, and spray_stats.regrows counts the re-renders; a smaller factor
 * exercises that path.  The reference's per-thread arenas and std::queues
 * grow without bound instead (ooc/ooc_tcontext.h:74-115). */
int spray_set_pool_headroom(spray_ctx* ctx, double factor);

/* Domain cache state (render/lru_cache.h:50-57, getCacheSize and
 * getBlocks): *capacity = domains the cache holds; lru_domains (capacity
 * entries) = the resident domains from least to most recently used, -1
 * past the last; counters = hits, misses and bytes copied into HBM since the
 * scene was loaded.  A scene held whole reports every domain resident and
 * zero counters. */
int spray_cache_state(spray_ctx* ctx, int* capacity, int* lru_domains,
                      unsigned long long counters[3]);

/* Kernel timing (new): when enabled, every trace-kernel launch (the
 * per-domain kernel of the plain form, the persistent kernel of the batched
 * forms) is bracketed by hipEvents on the context's stream (the batched
 * form then runs without its graph).  spray_kernel_time returns the launch
 * count and summed device time since the last reset. */
int spray_profile_kernels(spray_ctx* ctx, int enable);
int spray_kernel_time(spray_ctx* ctx, long long* launches, double* total_ms, int reset);

/* Schedule test hook (DomainStats::schedule, ooc/ooc_domain_stats.cc:60-111,
 * as PContext::isectPrims uses it, ooc/ooc_pcontext.h:127-131): with
 * max_records > 0 the plain schedule records its first max_records passes
 * from the next frame on (0 stops; each call clears the record).
 * spray_schedule_read returns the record count and, per record i,
 * meta[4i..4i+3] = {blocking tile, window, pass in the window, queue entries},
 * scores[i*ndomains + d] = sum over the pass's entries of domain d of
 * max(0, 16 - position in the ray's domain list), and order[i*ndomains ..]
 * = the domains by descending score, ties by id.  Arrays may be NULL. */
int spray_schedule_log(spray_ctx* ctx, int max_records);
int spray_schedule_read(spray_ctx* ctx, int* nrec, int* meta, unsigned long long* scores, int* order);

/* Learned form (new): domain-hit predictor.  Logging collects
 * (origin, direction, domain, hit) samples from the frames traced while it
 * is on (any schedule; the benchmark logs plain frames, and batched or
 * learned frames for online fine-tuning) until max_samples are held;
 * training runs the MLP's own forward/backward/Adam kernels on the GPU. */
int spray_predictor_log(spray_ctx* ctx, int enable, long long max_samples);
long long spray_predictor_num_samples(const spray_ctx* ctx);
int spray_predictor_train(spray_ctx* ctx, int epochs, int batch, float lr, float* loss_out);
int spray_predictor_eval(spray_ctx* ctx, float* loss_out, float* accuracy_out);
int spray_predictor_save(spray_ctx* ctx, const char* path);
int spray_predictor_load(spray_ctx* ctx, const char* path);

/* Predictor test hooks (new): one forward/backward of the MLP kernels on
 * host samples (8 floats each: origin, direction, domain, label) with host
 * parameters (flat fp32: W1[128][22], b1[128], W2[128][128], b2[128],
 * w3[128], b3, emb[ndomains][16]); returns the summed BCE loss and the
 * gradient of the mean BCE (n a multiple of 32).  spray_mlp_logits runs the
 * forward only, with the learned schedule's inference arithmetic (params
 * may be NULL: current weights). */
int spray_mlp_num_params(spray_ctx* ctx);
int spray_mlp_grad(spray_ctx* ctx, const float* samples, int n, const float* params,
                   float* loss_out, float* grad_out);
int spray_mlp_logits(spray_ctx* ctx, const float* samples, int n, const float* params,
                     float* logits_out);

/* Synthetic scene writer: domains as PLY files plus a .spray file
 * (dir/name.spray); extra_lines (e.g. light lines) go to the header. */
int spray_scenegen_write(const spray_scenegen_params* p, const char* dir, const char* name,
                         const char* extra_lines, long long* triangles_out);

/* One domain of that scene as the CPU generator builds it (host only): the
 * mesh spray_scenegen_write writes to name_dNNNN.ply, in the same vertex,
 * face and colour order.  With v, f and rgb NULL only the counts are
 * returned; otherwise the arrays hold nverts*3, nfaces*3 and nverts
 * entries. */
int spray_scenegen_domain(const spray_scenegen_params* p, int domain, float* v, unsigned int* f,
                          unsigned int* rgb, long long* nverts, long long* nfaces);

/* ---- on-disk formats and remeshing on the GPU (new) ----
 * spray_load_scene decodes binary little-endian PLY bodies on the GPU (the
 * file bytes go to HBM unparsed; io/ply_loader.cc:179-324 semantics and
 * errors); ASCII PLY stays on the host parser. */

/* The scene spray_scenegen_write + spray_load_scene would give (the same
 * meshes, domain bounds and .spray description, domain files named
 * name_dNNNN.ply), generated by marching-cubes kernels straight into the
 * HBM domain pool with no files.  Stands in for the reference's scene input
 * path (Scene::init, render/scene.inl:30-100) for in-situ data. */
int spray_scenegen_load(spray_ctx* ctx, const spray_scenegen_params* p, const char* name,
                        const char* extra_lines);
/* Remeshes a scene made by spray_scenegen_load at a new field phase (a
 * time-varying field remeshed every frame): meshes, domain bounds, normals
 * and BVHs are rebuilt on the GPU; camera, schedule, batch parameters and
 * predictor stay. */
int spray_remesh(spray_ctx* ctx, double phase);
/* Domain id's mesh as held in HBM (TriMeshBuffer's vertex, face and colour
 * buffers, render/trimesh_buffer.cc:52-110).  *nverts / *nfaces (when not
 * NULL) receive the loaded counts (the scene file's vertex/face lines are
 * not checked against the PLY); with v, f and rgb all NULL nothing else
 * happens, else v gets nverts*3 floats, f nfaces*3 indices and rgb nverts
 * packed colours (r<<16|g<<8|b), each pointer may be NULL.  Copying needs
 * the whole scene resident (no partial domain cache). */
int spray_domain_mesh(spray_ctx* ctx, int id, float* v, unsigned int* f, unsigned int* rgb,
                      long long* nverts, long long* nfaces);
/* Writes the loaded scene (remeshed frames included) as binary PLY files
 * dir/name_dNNNN.ply plus dir/name.spray (lights, materials, world bounds,
 * counts; vertices in world space), readable by spray_load_scene and the
 * reference's loaders (io/scene_loader.cc:315-358, io/ply_loader.cc). */
int spray_write_scene(spray_ctx* ctx, const char* dir, const char* name);
/* ply_header_reader (apps/ply_header_reader.cc:36-108): vertex and face
 * counts of a PLY file and the bounds of its vertices (binary bodies decoded
 * and reduced on the GPU). */
int spray_ply_info(spray_ctx* ctx, const char* path, long long* nverts, long long* nfaces,
                   float lo[3], float hi[3]);

/* HdrImage::writePpm (display/image.h:183-204): writes a W*H*4 rgba float
 * image (as returned by spray_trace_frame) as an ASCII P3 file with maximum
 * value 1023, top row first.  Host only. */
int spray_write_ppm(const float* rgba, int w, int h, const char* path);

#ifdef __cplusplus
}
#endif

#endif /* SPRAY_GPU_H_ */
