// Oracle (test infrastructure only): restatement of SpRay's ooc::Tracer.
#pragma once

#include <cstdint>
#include <vector>

#include "oscene.h"

namespace orc {

// render/config.h fields used by the ooc path, plus the documented new knobs.
struct TraceConfig {
  int image_w = 400, image_h = 400;
  float cam_pos[3] = {0, 0, 0}, cam_lookat[3] = {0, 0, -1}, cam_up[3] = {0, 1, 0};
  float fov = 90.f;
  int pixel_samples = 4;
  int ao_samples = 8;
  int ao_mode = 0;
  int bounces = 1;
  float ks[3] = {0, 0, 0};
  float shininess = 1.f;
  int64_t max_samples_per_rank = 1024 * 1024;
  int nthreads = 1;
  float ao_max_dist = kInf;  // reference: unbounded AO rays
  int rr_start = 0;          // 0: no Russian roulette (reference)
  int ref_pending_bug = 0;   // 1: restart pending rays without validation (reference)
  int canonical = 0;         // 1: order-independent per-sample tracer
  int tile_begin = 0;        // tiles [tile_begin, tile_end) are traced
  int tile_end = 1 << 30;
};

struct TraceStats {
  uint64_t useful_rays = 0;
  uint64_t eye_rays = 0;
  uint64_t radiance_visits = 0;  // (ray, domain) closest-hit traversals
  uint64_t shadow_visits = 0;    // (ray, domain) any-hit traversals
  uint64_t shades = 0;           // accepted history updates (shader calls)
  uint64_t final_hits = 0;       // hits on final history chains
  uint64_t passes = 0;
  uint64_t windows = 0;
  uint64_t tiles = 0;
  double seconds = 0;
};

// DomainStats schedule of every pass of the speculative tracer (test hook):
// per record meta = {tile, window, pass in window, entries queued}, the
// score of every domain and the visit order.
struct SchedLog {
  int max_records = 0;
  int ndomains = 0;
  std::vector<int> meta;        // 4 per record
  std::vector<int64_t> scores;  // ndomains per record
  std::vector<int> order;       // ndomains per record
  int records() const { return (int)meta.size() / 4; }
};

// Renders into rgba (image_w*image_h*4 floats, cleared by the caller).
// `log` (may be null) receives the schedule of the first passes.
void trace(const Scene& s, const TraceConfig& cfg, float* rgba, TraceStats* st,
           SchedLog* log = nullptr);

// Number of blocking tiles (render/tile.cc:317-391) for this configuration.
int numTiles(const TraceConfig& cfg);
void tileRect(const TraceConfig& cfg, int i, int* x, int* y, int* w, int* h);

}  // namespace orc
