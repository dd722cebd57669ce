// Device math for the tracing path: vectors, the counter-based RNG,
// cosine-hemisphere sampling, the pinhole camera, the domain slab test and
// the watertight triangle test.  Every expression keeps the reference's
// operation order and the library is built with -ffp-contract=off, so a
// kernel reproduces the reference arithmetic bit for bit.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define SPRAY_HD __host__ __device__ __forceinline__

namespace spray {

constexpr float kInfF = __builtin_huge_valf();
constexpr float kRayEps = 0.001f;                  // render/spray.h:46
constexpr double kPiD = 3.14159265358979323846;    // M_PI
constexpr float kOneOverPi = 0.3183098861837907f;  // render/spray.h:50
constexpr double k1Over255 = 0.00392156862745098;  // render/spray.h:54

struct f3 {
  float x, y, z;
};
SPRAY_HD f3 mk(float x, float y, float z) { return f3{x, y, z}; }
SPRAY_HD f3 add(f3 a, f3 b) { return f3{a.x + b.x, a.y + b.y, a.z + b.z}; }
SPRAY_HD f3 sub(f3 a, f3 b) { return f3{a.x - b.x, a.y - b.y, a.z - b.z}; }
SPRAY_HD f3 neg(f3 a) { return f3{-a.x, -a.y, -a.z}; }
SPRAY_HD f3 mul(f3 a, f3 b) { return f3{a.x * b.x, a.y * b.y, a.z * b.z}; }
SPRAY_HD f3 scl(f3 a, float s) { return f3{a.x * s, a.y * s, a.z * s}; }
SPRAY_HD f3 scl(float s, f3 a) { return f3{s * a.x, s * a.y, s * a.z}; }
SPRAY_HD f3 dvs(f3 a, float s) { return f3{a.x / s, a.y / s, a.z / s}; }
SPRAY_HD float dot(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
SPRAY_HD f3 cross(f3 a, f3 b) {
  return f3{a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y};
}
SPRAY_HD f3 normalize(f3 v) {  // glm::normalize: v * (1 / sqrt(dot(v, v)))
  float s = 1.0f / sqrtf(dot(v, v));
  return scl(v, s);
}
SPRAY_HD float clamp01(float x) {  // glm::clamp(x, 0, 1)
  float m = (x < 0.0f) ? 0.0f : x;
  return (1.0f < m) ? 1.0f : m;
}
SPRAY_HD bool hasPositive(f3 v) { return v.x > 0.0f || v.y > 0.0f || v.z > 0.0f; }

// deps/embree/random_sampler.h:34-114
SPRAY_HD uint32_t rngInit(int id) {
  uint32_t k = (uint32_t)id;
  k *= 0xcc9e2d51u;
  k = (k << 15) | (k >> 17);
  k *= 0x1b873593u;
  uint32_t h = 0u ^ k;
  h = ((h << 13) | (h >> 19)) * 5u + 0xe6546b64u;
  h ^= h >> 16;
  h *= 0x85ebca6bu;
  h ^= h >> 13;
  h *= 0xc2b2ae35u;
  h ^= h >> 16;
  return h;
}
SPRAY_HD float rng1D(uint32_t& s) {
  s = s * 1664525u + 1013904223u;
  return (float)(int)(s >> 1) * 4.656612873077392578125e-10f;
}

// sin and cos of a float angle evaluated in double (Cody-Waite reduction by
// pi/2, Taylor series through degree 17/16 on |r| <= pi/4) and rounded once.
SPRAY_HD void sincosF(float xf, float* s_out, float* c_out) {
  const double x = (double)xf;
  double kd = rint(x * 0.63661977236758134308);
  double r = (x - kd * 1.57079632679489655800e+00) - kd * 6.12323399573676603587e-17;
  long long k = (long long)kd;
  double r2 = r * r;
  double ps = 1.0 / 355687428096000.0;
  ps = ps * r2 - 1.0 / 1307674368000.0;
  ps = ps * r2 + 1.0 / 6227020800.0;
  ps = ps * r2 - 1.0 / 39916800.0;
  ps = ps * r2 + 1.0 / 362880.0;
  ps = ps * r2 - 1.0 / 5040.0;
  ps = ps * r2 + 1.0 / 120.0;
  ps = ps * r2 - 1.0 / 6.0;
  double sn = r + r * r2 * ps;
  double pc = 1.0 / 20922789888000.0;
  pc = pc * r2 - 1.0 / 87178291200.0;
  pc = pc * r2 + 1.0 / 479001600.0;
  pc = pc * r2 - 1.0 / 3628800.0;
  pc = pc * r2 + 1.0 / 40320.0;
  pc = pc * r2 - 1.0 / 720.0;
  pc = pc * r2 + 1.0 / 24.0;
  pc = pc * r2 - 0.5;
  double cs = 1.0 + r2 * pc;
  double s, c;
  switch ((int)(k & 3)) {
    case 0: s = sn; c = cs; break;
    case 1: s = cs; c = -sn; break;
    case 2: s = -sn; c = -cs; break;
    default: s = -cs; c = sn; break;
  }
  *s_out = (float)s;
  *c_out = (float)c;
}

// sin of a double, with the same double operations wherever it runs (host
// or GPU, both built without contraction): Cody-Waite reduction by pi/2
// with a three-part constant (pi/2 rounded to 33 bits, the next 33 bits,
// the rest), then the Taylor series of sin (through degree 23) or cos
// (through degree 22) on |r| <= pi/4, chosen by the quadrant.  Within a
// few ulp of the true sine for |x| < 1e5.  The marching-cubes generators
// evaluate their field with it on the CPU and on the GPU, so a grid point's
// field value, and with it the mesh, is the same bit for bit on both.
SPRAY_HD double sinD(double x) {
  const double kd = rint(x * 0.63661977236758134308);
  double r = x - kd * 1.5707963267341256;
  r = r - kd * 6.077100506303966e-11;
  r = r - kd * 2.0222662487959506e-21;
  const double r2 = r * r;
  const long long k = (long long)kd;
  double p;
  if ((k & 1) == 0) {  // sin r = r (1 - r2/3! + r2^2/5! - ...)
    p = -1.0 / 25852016738884976640000.0;
    p = p * r2 + 1.0 / 51090942171709440000.0;
    p = p * r2 - 1.0 / 121645100408832000.0;
    p = p * r2 + 1.0 / 355687428096000.0;
    p = p * r2 - 1.0 / 1307674368000.0;
    p = p * r2 + 1.0 / 6227020800.0;
    p = p * r2 - 1.0 / 39916800.0;
    p = p * r2 + 1.0 / 362880.0;
    p = p * r2 - 1.0 / 5040.0;
    p = p * r2 + 1.0 / 120.0;
    p = p * r2 - 1.0 / 6.0;
    p = r + r * r2 * p;
  } else {  // cos r = 1 - r2/2! + r2^2/4! - ...
    p = 1.0 / 1124000727777607680000.0;
    p = p * r2 - 1.0 / 2432902008176640000.0;
    p = p * r2 + 1.0 / 6402373705728000.0;
    p = p * r2 - 1.0 / 20922789888000.0;
    p = p * r2 + 1.0 / 87178291200.0;
    p = p * r2 - 1.0 / 479001600.0;
    p = p * r2 + 1.0 / 3628800.0;
    p = p * r2 - 1.0 / 40320.0;
    p = p * r2 + 1.0 / 720.0;
    p = p * r2 - 1.0 / 24.0;
    p = p * r2 + 0.5;
    p = 1.0 - r2 * p;
  }
  return (k & 2) ? -p : p;
}

// render/sampler.h:49-92 ConcentricDiskSampling
SPRAY_HD void concentricDisk(float u1, float u2, float* dx, float* dy) {
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
      theta = (sy > 0.0f) ? sy / r : 8.0f + sy / r;
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
  theta = (float)((double)theta * (kPiD / 4.0));
  float s, c;
  sincosF(theta, &s, &c);
  *dx = r * c;
  *dy = r * s;
}

// render/sampler.h:101-110 localToWorld; render/sampler.cc:54-60
SPRAY_HD void cosineSample(float u1, float u2, f3 N, f3* wi, float* pdf) {
  f3 v;
  concentricDisk(u1, u2, &v.x, &v.y);
  float z2 = 1.f - v.x * v.x - v.y * v.y;
  v.z = sqrtf(z2 < 0.f ? 0.f : z2);
  v = normalize(v);
  f3 dx0 = mk(0, N.z, -N.y);
  f3 dx1 = mk(-N.z, 0, N.x);
  f3 dx = normalize(dot(dx0, dx0) > dot(dx1, dx1) ? dx0 : dx1);
  f3 dy = normalize(cross(N, dx));
  f3 w = mk(dx.x * v.x + dy.x * v.y + N.x * v.z, dx.y * v.x + dy.y * v.y + N.y * v.z,
            dx.z * v.x + dy.z * v.y + N.z * v.z);
  *wi = normalize(w);
  *pdf = v.z * kOneOverPi;
}

// utils/util.h:69-74
SPRAY_HD f3 unpackColor(uint32_t c) {
  return mk((float)((double)((c >> 16) & 0xff) * k1Over255),
            (float)((double)((c >> 8) & 0xff) * k1Over255),
            (float)((double)(c & 0xff) * k1Over255));
}

struct CameraDev {
  f3 pos, lowerleft, wvec, hvec;
  float img_w, img_h;
};
// render/camera.h:177-209
SPRAY_HD f3 cameraDir(const CameraDev& c, float x, float y) {
  float u = x / c.img_w;
  float v = y / c.img_h;
  f3 d = sub(add(add(c.lowerleft, scl(u, c.wvec)), scl(v, c.hvec)), c.pos);
  return normalize(d);
}

// render/aabb.h:139-169, returns tmin through *tmin_o
// Domain AABB slab test (render/aabb.h:139-169), entry distance in *tmin_o.
// A ray parallel to a slab whose origin lies on one of its planes gives
// 0 * inf = NaN there: the ray is inside the slab when its origin is on the
// lower plane and outside when on the upper one (unless the slab is flat),
// so a ray along a face shared by two domains belongs to exactly one of
// them.  Only rays with a zero (or denormal) direction component can
// produce NaN: safe = false (uniform over a wave) skips the check.
SPRAY_HD bool slabDomainS(const float lo[3], const float hi[3], const float org[3],
                          const float inv[3], bool safe, float* tmin_o) {
  float tmin = -kInfF, tmax = kInfF;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    float a = ((inv[k] < 0.0f ? hi[k] : lo[k]) - org[k]) * inv[k];
    float b = ((inv[k] < 0.0f ? lo[k] : hi[k]) - org[k]) * inv[k];
    if (safe && (a != a || b != b)) {
      if (org[k] == hi[k] && lo[k] < hi[k]) return false;
      continue;
    }
    if (a > tmin) tmin = a;
    if (b < tmax) tmax = b;
  }
  *tmin_o = tmin;
  return tmin <= tmax && (tmin < kInfF) && (tmax > kRayEps);
}

// Branch-free form of the same test against a box held as two float4 (lo,
// hi).  For finite 1/dir, min/max of the two plane distances equal the
// reference's sign-selected ones bit for bit (they are the same products).
// kSafe (uniform over a wave) adds the NaN rule of slabDomainS.
SPRAY_HD bool slabBox4(const float4& lo, const float4& hi, const float org[3], const float inv[3],
                       bool kSafe, float* tmin_o) {
  const float ax = (lo.x - org[0]) * inv[0], bx = (hi.x - org[0]) * inv[0];
  const float ay = (lo.y - org[1]) * inv[1], by = (hi.y - org[1]) * inv[1];
  const float az = (lo.z - org[2]) * inv[2], bz = (hi.z - org[2]) * inv[2];
  float nx = fminf(ax, bx), fx = fmaxf(ax, bx);
  float ny = fminf(ay, by), fy = fmaxf(ay, by);
  float nz = fminf(az, bz), fz = fmaxf(az, bz);
  bool out = false;
  if (kSafe) {
    const bool qx = ax != ax || bx != bx, qy = ay != ay || by != by, qz = az != az || bz != bz;
    out = (qx && org[0] == hi.x && lo.x < hi.x) || (qy && org[1] == hi.y && lo.y < hi.y) ||
          (qz && org[2] == hi.z && lo.z < hi.z);
    nx = qx ? -kInfF : nx;
    fx = qx ? kInfF : fx;
    ny = qy ? -kInfF : ny;
    fy = qy ? kInfF : fy;
    nz = qz ? -kInfF : nz;
    fz = qz ? kInfF : fz;
  }
  const float tmin = fmaxf(fmaxf(nx, ny), nz), tmax = fminf(fminf(fx, fy), fz);
  *tmin_o = tmin;
  return !out && tmin <= tmax && tmin < kInfF && tmax > kRayEps;
}

// Closed variant for culling boxes that contain other boxes: no half-open
// rejection, NaN axes unbounded.  A box that contains B passes whenever
// slabBox4 accepts B (the slab arithmetic is monotone in the planes).
SPRAY_HD bool slabBoxClosed(const float4& lo, const float4& hi, const float org[3],
                            const float inv[3], bool kSafe) {
  const float ax = (lo.x - org[0]) * inv[0], bx = (hi.x - org[0]) * inv[0];
  const float ay = (lo.y - org[1]) * inv[1], by = (hi.y - org[1]) * inv[1];
  const float az = (lo.z - org[2]) * inv[2], bz = (hi.z - org[2]) * inv[2];
  float nx = fminf(ax, bx), fx = fmaxf(ax, bx);
  float ny = fminf(ay, by), fy = fmaxf(ay, by);
  float nz = fminf(az, bz), fz = fmaxf(az, bz);
  if (kSafe) {
    const bool qx = ax != ax || bx != bx, qy = ay != ay || by != by, qz = az != az || bz != bz;
    nx = qx ? -kInfF : nx;
    fx = qx ? kInfF : fx;
    ny = qy ? -kInfF : ny;
    fy = qy ? kInfF : fy;
    nz = qz ? -kInfF : nz;
    fz = qz ? kInfF : fz;
  }
  const float tmin = fmaxf(fmaxf(nx, ny), nz), tmax = fminf(fminf(fx, fy), fz);
  return tmin <= tmax && tmin < kInfF && tmax > kRayEps;
}

// True when 1/dir has an infinite component (dir component 0 or denormal).
SPRAY_HD bool hasInfInv(const float inv[3]) {
  return fabsf(inv[0]) == kInfF || fabsf(inv[1]) == kInfF || fabsf(inv[2]) == kInfF;
}

SPRAY_HD bool slabDomain(const float lo[3], const float hi[3], const float org[3],
                         const float dir[3], float* tmin_o) {
  float inv[3] = {1.0f / dir[0], 1.0f / dir[1], 1.0f / dir[2]};
  return slabDomainS(lo, hi, org, inv, true, tmin_o);
}

// Watertight ray/triangle test (Woop, Benthin, Wald, JCGT 2(1), 2013).
struct RayPre {
  int kx, ky, kz;
  float Sx, Sy, Sz;
};
SPRAY_HD RayPre rayPre(const float dir[3]) {
  RayPre p;
  float ax = fabsf(dir[0]), ay = fabsf(dir[1]), az = fabsf(dir[2]);
  int kz = (ax > ay) ? ((ax > az) ? 0 : 2) : ((ay > az) ? 1 : 2);
  int kx = (kz == 2) ? 0 : kz + 1;
  int ky = (kx == 2) ? 0 : kx + 1;
  float dz = dir[kz];
  if (dz < 0.0f) {
    int t = kx;
    kx = ky;
    ky = t;
  }
  p.kx = kx;
  p.ky = ky;
  p.kz = kz;
  p.Sx = dir[kx] / dz;
  p.Sy = dir[ky] / dz;
  p.Sz = 1.0f / dz;
  return p;
}
SPRAY_HD float pick(const float* a, int k) { return k == 0 ? a[0] : (k == 1 ? a[1] : a[2]); }

// Returns true for tnear < t < tfar and writes t, u (weight of v1), v (weight of v2).
SPRAY_HD bool triHit(const RayPre& rp, const float org[3], const float* v0, const float* v1,
                     const float* v2, float tnear, float tfar, float* t_o, float* u_o,
                     float* v_o) {
  float A[3] = {v0[0] - org[0], v0[1] - org[1], v0[2] - org[2]};
  float B[3] = {v1[0] - org[0], v1[1] - org[1], v1[2] - org[2]};
  float C[3] = {v2[0] - org[0], v2[1] - org[1], v2[2] - org[2]};
  float Akz = pick(A, rp.kz), Bkz = pick(B, rp.kz), Ckz = pick(C, rp.kz);
  float Ax = pick(A, rp.kx) - rp.Sx * Akz;
  float Ay = pick(A, rp.ky) - rp.Sy * Akz;
  float Bx = pick(B, rp.kx) - rp.Sx * Bkz;
  float By = pick(B, rp.ky) - rp.Sy * Bkz;
  float Cx = pick(C, rp.kx) - rp.Sx * Ckz;
  float Cy = pick(C, rp.ky) - rp.Sy * Ckz;
  float U = Cx * By - Cy * Bx;
  float V = Ax * Cy - Ay * Cx;
  float W = Bx * Ay - By * Ax;
  if (U == 0.0f || V == 0.0f || W == 0.0f) {
    U = (float)((double)Cx * (double)By - (double)Cy * (double)Bx);
    V = (float)((double)Ax * (double)Cy - (double)Ay * (double)Cx);
    W = (float)((double)Bx * (double)Ay - (double)By * (double)Ax);
  }
  if ((U < 0.0f || V < 0.0f || W < 0.0f) && (U > 0.0f || V > 0.0f || W > 0.0f)) return false;
  float det = U + V + W;
  if (det == 0.0f) return false;
  float Az = rp.Sz * Akz;
  float Bz = rp.Sz * Bkz;
  float Cz = rp.Sz * Ckz;
  float T = U * Az + V * Bz + W * Cz;
  float rcp = 1.0f / det;
  float t = T * rcp;
  if (!(t > tnear && t < tfar)) return false;
  *t_o = t;
  *u_o = V * rcp;
  *v_o = W * rcp;
  return true;
}

}  // namespace spray
