// Oracle (test infrastructure only): CPU restatement of SpRay's per-rank
// tracing arithmetic.  Only tests/, __graft_entry__.smoke() and bench.py's
// cpu_baseline leg may load the oracle; the product never does.
//
// Every function states the reference file:line it restates.  All float
// arithmetic is written in the reference's operation order and the oracle
// is compiled with -ffp-contract=off, so results are reproducible bit for
// bit by any IEEE-754 single-precision implementation that follows the
// same order.
#pragma once

#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>

namespace orc {

constexpr float kInf = std::numeric_limits<float>::infinity();
// render/spray.h:46 (SPRAY_RAY_EPSILON)
constexpr float kRayEps = 0.001f;
// render/spray.h:48-53
constexpr double kPi = 3.14159265358979323846;  // M_PI
constexpr float kOneOverPi = 0.3183098861837907f;
constexpr double k1Over255 = 0.00392156862745098;
// src/CMakeLists.txt:32-33 (SPRAY_RAY_DOMAIN_LIST_SIZE, SPRAY_SPECU_HISTORY_SIZE)
constexpr int kDomainListSize = 16;
constexpr int kHistory = 4;

struct V3 {
  float x, y, z;
};
inline V3 v3(float x, float y, float z) { return V3{x, y, z}; }
inline V3 operator+(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V3 operator-(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 operator-(V3 a) { return {-a.x, -a.y, -a.z}; }
inline V3 operator*(V3 a, V3 b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
inline V3 operator*(V3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
inline V3 operator*(float s, V3 a) { return {s * a.x, s * a.y, s * a.z}; }
inline V3 operator/(V3 a, float s) { return {a.x / s, a.y / s, a.z / s}; }
inline float get(V3 a, int i) { return i == 0 ? a.x : (i == 1 ? a.y : a.z); }

// glm::dot: (a.x*b.x + a.y*b.y) + a.z*b.z
inline float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
// glm::cross
inline V3 cross(V3 a, V3 b) {
  return {a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y};
}
// glm::normalize = v * (1 / sqrt(dot(v, v)))
inline V3 normalize(V3 v) {
  float s = 1.0f / std::sqrt(dot(v, v));
  return v * s;
}
// glm::clamp = min(max(x, lo), hi) with glm's comparison forms
inline float clampf(float x, float lo, float hi) {
  float m = (x < lo) ? lo : x;
  return (hi < m) ? hi : m;
}
// utils/math.h:76-78
inline bool hasPositive(V3 v) { return v.x > 0.0f || v.y > 0.0f || v.z > 0.0f; }

// ---------------------------------------------------------------------------
// RNG: deps/embree/random_sampler.h:34-114 (MurmurHash3 mix/finalize + LCG)
inline uint32_t murmurMix(uint32_t hash, uint32_t k) {
  k *= 0xcc9e2d51u;
  k = (k << 15) | (k >> 17);
  k *= 0x1b873593u;
  hash ^= k;
  hash = ((hash << 13) | (hash >> 19)) * 5u + 0xe6546b64u;
  return hash;
}
inline uint32_t murmurFinalize(uint32_t h) {
  h ^= h >> 16;
  h *= 0x85ebca6bu;
  h ^= h >> 13;
  h *= 0xc2b2ae35u;
  h ^= h >> 16;
  return h;
}
struct Rng {
  uint32_t s;
};
// RandomSampler_init(self, id) :69-75
inline Rng rngInit(int id) { return Rng{murmurFinalize(murmurMix(0u, (uint32_t)id))}; }
// RandomSampler_get1D :92-106: (int)(s>>1) * 2^-31
inline float rngGet1D(Rng& r) {
  r.s = r.s * 1664525u + 1013904223u;
  return (float)(int)(r.s >> 1) * 4.656612873077392578125e-10f;
}

// ---------------------------------------------------------------------------
// sin/cos of a float argument.  The reference calls glm::cos/glm::sin
// (libm cosf/sinf) in ConcentricDiskSampling (render/sampler.h:84-85).  Here
// they are evaluated in double precision (Cody-Waite reduction by pi/2 and
// Taylor polynomials through degree 17/16 on |r| <= pi/4, absolute error
// below 1e-16) and rounded once to float, which is the correctly rounded
// result except when the exact value lies within ~1e-16 of a float rounding
// boundary.  The product's kernels evaluate the same double-precision
// sequence, so both sides agree bit for bit.
inline void sincosRef(float xf, float* s_out, float* c_out) {
  const double x = (double)xf;
  const double kTwoOverPi = 0.63661977236758134308;
  const double kPio2Hi = 1.57079632679489655800e+00;  // pi/2 high part
  const double kPio2Lo = 6.12323399573676603587e-17;  // pi/2 - kPio2Hi
  double kd = std::nearbyint(x * kTwoOverPi);
  double r = (x - kd * kPio2Hi) - kd * kPio2Lo;
  long k = (long)kd;
  double r2 = r * r;
  // sin(r) = r - r^3/3! + ... + r^17/17!
  double ps = 1.0 / 355687428096000.0;              // 1/17!
  ps = ps * r2 - 1.0 / 1307674368000.0;             // 1/15!
  ps = ps * r2 + 1.0 / 6227020800.0;                // 1/13!
  ps = ps * r2 - 1.0 / 39916800.0;                  // 1/11!
  ps = ps * r2 + 1.0 / 362880.0;                    // 1/9!
  ps = ps * r2 - 1.0 / 5040.0;                      // 1/7!
  ps = ps * r2 + 1.0 / 120.0;                       // 1/5!
  ps = ps * r2 - 1.0 / 6.0;                         // 1/3!
  double sn = r + r * r2 * ps;
  // cos(r) = 1 - r^2/2! + ... + r^16/16!
  double pc = 1.0 / 20922789888000.0;               // 1/16!
  pc = pc * r2 - 1.0 / 87178291200.0;               // 1/14!
  pc = pc * r2 + 1.0 / 479001600.0;                 // 1/12!
  pc = pc * r2 - 1.0 / 3628800.0;                   // 1/10!
  pc = pc * r2 + 1.0 / 40320.0;                     // 1/8!
  pc = pc * r2 - 1.0 / 720.0;                       // 1/6!
  pc = pc * r2 + 1.0 / 24.0;                        // 1/4!
  pc = pc * r2 - 0.5;                               // 1/2!
  double cs = 1.0 + r2 * pc;
  double s, c;
  switch (k & 3) {
    case 0: s = sn; c = cs; break;
    case 1: s = cs; c = -sn; break;
    case 2: s = -sn; c = -cs; break;
    default: s = -cs; c = sn; break;
  }
  *s_out = (float)s;
  *c_out = (float)c;
}

// ---------------------------------------------------------------------------
// Sampling: render/sampler.h:49-92 (ConcentricDiskSampling),
// :101-110 (localToWorld), :114-128 (cosineHemisphereSample/Pdf),
// render/sampler.cc:54-60 (getCosineHemisphereSample)
inline void concentricDisk(float u1, float u2, float* dx, float* dy) {
  float r, theta;
  float sx = 2 * u1 - 1;
  float sy = 2 * u2 - 1;
  if (sx == 0.0f && sy == 0.0f) {
    *dx = 0.0f;
    *dy = 0.0f;
    return;
  }
  if (sx >= -sy) {
    if (sx > sy) {
      r = sx;
      if (sy > 0.0f)
        theta = sy / r;
      else
        theta = 8.0f + sy / r;
    } else {
      r = sy;
      theta = 2.0f - sx / r;
    }
  } else {
    if (sx <= sy) {
      r = -sx;
      theta = 4.0f - sy / r;
    } else {
      r = -sy;
      theta = 6.0f + sx / r;
    }
  }
  // theta *= SPRAY_PI / 4.f  (double product, rounded to float)
  theta = (float)((double)theta * (kPi / 4.0));
  float s, c;
  sincosRef(theta, &s, &c);
  *dx = r * c;
  *dy = r * s;
}

inline V3 localToWorld(V3 N, V3 v) {
  V3 dx0 = v3(0, N.z, -N.y);
  V3 dx1 = v3(-N.z, 0, N.x);
  V3 dx = normalize(dot(dx0, dx0) > dot(dx1, dx1) ? dx0 : dx1);
  V3 dy = normalize(cross(N, dx));
  // glm mat3(dx, dy, N) * v
  return {dx.x * v.x + dy.x * v.y + N.x * v.z, dx.y * v.x + dy.y * v.y + N.y * v.z,
          dx.z * v.x + dy.z * v.y + N.z * v.z};
}

// getCosineHemisphereSample(u1, u2, N, &wi, &pdf)
inline void cosineHemisphereSample(float u1, float u2, V3 N, V3* wi, float* pdf) {
  V3 v;
  concentricDisk(u1, u2, &v.x, &v.y);
  float z2 = 1.f - v.x * v.x - v.y * v.y;
  v.z = std::sqrt(z2 < 0.f ? 0.f : z2);  // glm::max(0.f, z2) = (0 < z2) ? z2 : 0
  v = normalize(v);
  *wi = normalize(localToWorld(N, v));
  *pdf = v.z * kOneOverPi;
}

// utils/util.h:69-74 unpack(uint32, vec3): channel * SPRAY_1_OVER_255 (double)
inline V3 unpackColor(uint32_t c) {
  return {(float)((double)((c >> 16) & 0xff) * k1Over255),
          (float)((double)((c >> 8) & 0xff) * k1Over255), (float)((double)(c & 0xff) * k1Over255)};
}

// ---------------------------------------------------------------------------
// Camera: render/camera.h:128-166 (init), :177-209 (getDirection/generateRay)
struct Camera {
  V3 pos, lookat, up;
  V3 lowerleft, wvec, hvec;
  float img_w, img_h;
  void init(V3 campos, V3 look, V3 upv, float vfov, int w, int h) {
    float aspect = (float)w / (float)h;
    float theta = (float)((double)vfov * kPi / 180.0f);
    float half_h = (float)std::tan((double)(theta / 2.0f));
    float half_w = aspect * half_h;
    V3 ww = campos - look;
    V3 uu = cross(upv, ww);
    V3 vv = cross(ww, uu);
    ww = normalize(ww);
    uu = normalize(uu);
    vv = normalize(vv);
    V3 center = campos - ww;
    lowerleft = center - (half_w * uu) - (half_h * vv);
    wvec = (2.0f * half_w) * uu;
    hvec = (2.0f * half_h) * vv;
    pos = campos;
    lookat = look;
    up = upv;
    img_w = (float)w;
    img_h = (float)h;
  }
  V3 direction(float x, float y) const {
    float u = x / img_w;
    float v = y / img_h;
    V3 d = lowerleft + (u * wvec) + (v * hvec) - pos;
    return normalize(d);
  }
};

// ---------------------------------------------------------------------------
// Ray/box slab test: render/aabb.h:139-169 (intersectAabb with float[3]).
// The reference's slab distances are (plane - org) / dir; a ray parallel to
// a slab (dir 0, 1/dir infinite) whose origin lies on one of its planes gives
// 0 * inf = NaN there, which the reference's comparisons then mishandle
// (deviation Q13: it drops the domain).  Here such a ray is inside the slab
// when its origin is on the lower plane and outside when on the upper one
// (unless the slab is flat), so a ray running along a face shared by two
// domains belongs to exactly one of them.  Without NaNs the result equals
// the reference's (its comparison chain computes the same max of entries
// and min of exits).
inline bool intersectAabb(const float lo[3], const float hi[3], const float org[3],
                          const float dir[3], float t0, float t1, float* tmin_o) {
  float tmin = -kInf, tmax = kInf;
  for (int k = 0; k < 3; ++k) {
    float inv = 1.0f / dir[k];
    float a = ((inv < 0.0f ? hi[k] : lo[k]) - org[k]) * inv;
    float b = ((inv < 0.0f ? lo[k] : hi[k]) - org[k]) * inv;
    if (a != a || b != b) {  // parallel ray on a slab plane
      if (org[k] == hi[k] && lo[k] < hi[k]) return false;
      continue;
    }
    if (a > tmin) tmin = a;
    if (b < tmax) tmax = b;
  }
  if (tmin > tmax) return false;
  *tmin_o = tmin;
  return (tmin < t1) && (tmax > t0);
}

}  // namespace orc
