// Oracle (test infrastructure only): scene-file, PLY, normals, BVH and
// triangle-test restatement.
#include "oscene.h"

#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <stdexcept>

namespace orc {

namespace {

void fail(const std::string& msg) { throw std::runtime_error(msg); }

void identity(float m[16]) {
  for (int i = 0; i < 16; ++i) m[i] = (i % 5 == 0) ? 1.f : 0.f;
}

// glm column access: m[c*4 + r]
// glm::scale(m, v): columns 0..2 scaled (glm/ext/matrix_transform.inl)
void mscale(float m[16], float x, float y, float z) {
  float s[3] = {x, y, z};
  for (int c = 0; c < 3; ++c)
    for (int r = 0; r < 4; ++r) m[c * 4 + r] = m[c * 4 + r] * s[c];
}
// glm::translate(m, v): m[3] = m[0]*v0 + m[1]*v1 + m[2]*v2 + m[3]
void mtranslate(float m[16], float x, float y, float z) {
  for (int r = 0; r < 4; ++r)
    m[12 + r] = m[0 + r] * x + m[4 + r] * y + m[8 + r] * z + m[12 + r];
}
// glm::rotate(m, angle, axis)
void mrotate(float m[16], float angle, V3 axis_in) {
  float c = std::cos(angle), s = std::sin(angle);
  V3 axis = normalize(axis_in);
  V3 temp = (1.f - c) * axis;
  float R[3][3];
  R[0][0] = c + temp.x * axis.x;
  R[0][1] = temp.x * axis.y + s * axis.z;
  R[0][2] = temp.x * axis.z - s * axis.y;
  R[1][0] = temp.y * axis.x - s * axis.z;
  R[1][1] = c + temp.y * axis.y;
  R[1][2] = temp.y * axis.z + s * axis.x;
  R[2][0] = temp.z * axis.x + s * axis.y;
  R[2][1] = temp.z * axis.y - s * axis.x;
  R[2][2] = c + temp.z * axis.z;
  float out[16];
  for (int j = 0; j < 3; ++j)
    for (int r = 0; r < 4; ++r)
      out[j * 4 + r] = m[0 + r] * R[j][0] + m[4 + r] * R[j][1] + m[8 + r] * R[j][2];
  for (int r = 0; r < 4; ++r) out[12 + r] = m[12 + r];
  std::memcpy(m, out, sizeof(out));
}
// glm mat4 * vec4 (pairwise: (m0*x + m1*y) + (m2*z + m3*w))
void mxform(const float m[16], const float p[3], float out[3]) {
  for (int r = 0; r < 3; ++r) {
    float a0 = m[0 + r] * p[0] + m[4 + r] * p[1];
    float a1 = m[8 + r] * p[2] + m[12 + r] * 1.0f;
    out[r] = a0 + a1;
  }
}
bool isIdentity(const float m[16]) {
  for (int i = 0; i < 16; ++i)
    if (m[i] != ((i % 5 == 0) ? 1.f : 0.f)) return false;
  return true;
}

std::vector<std::string> tokenize(const std::string& line) {
  // strtok(line, " ") semantics: split on spaces only, skip empties
  std::vector<std::string> out;
  std::string cur;
  for (char ch : line) {
    if (ch == ' ') {
      if (!cur.empty()) out.push_back(cur);
      cur.clear();
    } else {
      cur.push_back(ch);
    }
  }
  if (!cur.empty()) out.push_back(cur);
  return out;
}

float f(const std::string& s) { return (float)atof(s.c_str()); }

}  // namespace

void loadSceneFile(const std::string& spray_file, const std::string& ply_path, Scene* s) {
  std::ifstream in(spray_file);
  if (!in.is_open()) fail("unable to open input file " + spray_file);
  s->domains.clear();
  s->lights.clear();
  std::string line;
  Domain* cur = nullptr;
  // the reference allocates the domain vector up front (countAndAllocate),
  // so references stay valid; a deque-like growth via reserve is enough here
  std::vector<std::string> lines;
  while (std::getline(in, line)) lines.push_back(line);
  size_t ndom = 0;
  for (auto& l : lines) {
    auto t = tokenize(l);
    if (!t.empty() && t[0] == "domain") ++ndom;
  }
  if (ndom == 0) fail("no domain in scene file");
  s->domains.reserve(ndom);
  for (auto& l : lines) {
    auto t = tokenize(l);
    if (t.empty()) continue;
    const std::string& tag = t[0];
    if (tag[0] == '#') continue;
    if (tag == "domain") {
      s->domains.emplace_back();
      cur = &s->domains.back();
      cur->id = (int)s->domains.size() - 1;
      identity(cur->xf);
      continue;
    }
    if (tag == "light") {
      if (t.size() < 2) fail("bad light line");
      Light L;
      if (t[1] == "point") {
        if (t.size() != 8) fail("light point needs 6 values");
        L.type = kPointLight;
        L.pos = v3(f(t[2]), f(t[3]), f(t[4]));
        L.radiance = v3(f(t[5]), f(t[6]), f(t[7]));
      } else if (t[1] == "diffuse") {
        if (t.size() != 5) fail("light diffuse needs 3 values");
        L.type = kHemisphereLight;
        L.radiance = v3(f(t[2]), f(t[3]), f(t[4]));
      } else if (t[1] == "rect") {
        if (t.size() != 14) fail("light rect needs 12 values");
        L.type = kRectLight;
        L.pos = v3(f(t[2]), f(t[3]), f(t[4]));
        L.e1 = v3(f(t[5]), f(t[6]), f(t[7]));
        L.e2 = v3(f(t[8]), f(t[9]), f(t[10]));
        L.radiance = v3(f(t[11]), f(t[12]), f(t[13]));
        V3 c = cross(L.e1, L.e2);
        L.area = std::sqrt(dot(c, c));
        L.normal = normalize(c);
      } else {
        fail("unknown light source " + t[1]);
      }
      s->lights.push_back(L);
      continue;
    }
    static const char* known[] = {"file", "mtl", "bound", "scale", "rotate",
                                  "translate", "face", "vertex"};
    bool ok = false;
    for (auto k : known) ok = ok || (tag == k);
    if (!ok) fail("unknown tag name " + tag);
    if (!cur) fail("tag before first domain: " + tag);
    Domain& d = *cur;
    if (tag == "file") {
      if (t.size() != 2) fail("file needs 1 value");
      d.filename = ply_path.empty() ? t[1] : ply_path + "/" + t[1];
    } else if (tag == "mtl") {
      if (t.size() < 2) fail("bad mtl");
      if (t[1] == "diffuse" || t[1] == "mirror") {
        if (t.size() != 5) fail("mtl needs 3 values");
        d.mtl.type = (t[1] == "diffuse") ? kDiffuse : kMirror;
        d.mtl.rgb[0] = f(t[2]);
        d.mtl.rgb[1] = f(t[3]);
        d.mtl.rgb[2] = f(t[4]);
      } else if (t[1] == "glass" || t[1] == "transmission") {
        if (t.size() != 4) fail("mtl needs 2 values");
        d.mtl.type = (t[1] == "glass") ? kGlass : kTransmission;
        d.mtl.eta_a = f(t[2]);
        d.mtl.eta_b = f(t[3]);
      } else {
        fail("unknown material type " + t[1]);
      }
    } else if (tag == "bound") {
      if (t.size() != 7) fail("bound needs 6 values");
      for (int i = 0; i < 3; ++i) {
        d.obj_lo[i] = f(t[1 + i]);
        d.obj_hi[i] = f(t[4 + i]);
      }
      d.has_bound = true;
    } else if (tag == "scale") {
      if (t.size() != 4) fail("scale needs 3 values");
      mscale(d.xf, f(t[1]), f(t[2]), f(t[3]));
    } else if (tag == "rotate") {
      if (t.size() != 3) fail("rotate needs 2 values");
      V3 axis;
      if (t[1] == "x")
        axis = v3(1, 0, 0);
      else if (t[1] == "y")
        axis = v3(0, 1, 0);
      else if (t[1] == "z")
        axis = v3(0, 0, 1);
      else
        fail("invalid axis name " + t[1]);
      // glm::radians(atof(...)) in double, then (float)
      mrotate(d.xf, (float)(atof(t[2].c_str()) * 0.01745329251994329576923690768489), axis);
    } else if (tag == "translate") {
      if (t.size() != 4) fail("translate needs 3 values");
      mtranslate(d.xf, f(t[1]), f(t[2]), f(t[3]));
    } else if (tag == "face") {
      if (t.size() != 2) fail("face needs 1 value");
      d.num_faces = std::stoul(t[1]);
    } else if (tag == "vertex") {
      if (t.size() != 2) fail("vertex needs 1 value");
      d.num_vertices = std::stoul(t[1]);
    }
  }
  // scene_loader.cc:351-357: world AABB = transform * object bound corners
  for (auto& d : s->domains) {
    mxform(d.xf, d.obj_lo, d.lo);
    mxform(d.xf, d.obj_hi, d.hi);
  }
}

void mergeBounds(Scene* s) {
  for (int i = 0; i < 3; ++i) {
    s->lo[i] = kInf;
    s->hi[i] = -kInf;
  }
  for (auto& d : s->domains) {
    // scene.inl:314-317: bound, vertex and face counts are required
    bool valid = d.has_bound && d.lo[0] <= d.hi[0] && d.lo[1] <= d.hi[1] && d.lo[2] <= d.hi[2];
    if (!valid) fail("domain " + std::to_string(d.id) + ": invalid bound");
    if (d.num_vertices == 0 || d.num_faces == 0)
      fail("domain " + std::to_string(d.id) + ": missing vertex/face count");
    for (int i = 0; i < 3; ++i) {
      s->lo[i] = std::min(s->lo[i], d.lo[i]);
      s->hi[i] = std::max(s->hi[i], d.hi[i]);
    }
  }
}

namespace {

enum { kAscii = 1, kLE = 2, kBE = 3 };
int dtypeSize(const std::string& t) {
  if (t == "char" || t == "uchar" || t == "int8" || t == "uint8") return 1;
  if (t == "short" || t == "ushort" || t == "int16" || t == "uint16") return 2;
  if (t == "int" || t == "uint" || t == "float" || t == "int32" || t == "uint32" ||
      t == "float32")
    return 4;
  if (t == "double" || t == "float64") return 8;
  fail("unknown PLY data type " + t);
  return 0;
}

}  // namespace

void loadPly(const std::string& filename, Mesh* m, size_t* nv_out, size_t* nf_out) {
  std::ifstream in(filename, std::ios::in | std::ios::binary);
  if (!in.is_open()) fail("unable to open " + filename);
  std::string line;
  std::getline(in, line);
  if (line != "ply") fail("unknown file type");
  int format = 0;
  size_t nv = 0, nf = 0;
  int elem = -1;  // 0 vertex, 1 face
  int prop_state = 0;
  bool has_color = false;
  int list_size_bytes = 1, index_bytes = 4;
  bool end_header = false;
  int nelems = 0;
  std::vector<int> order;
  while (std::getline(in, line)) {
    std::istringstream ss(line);
    std::string word;
    ss >> word;
    if (word == "format") {
      ss >> word;
      if (word == "ascii")
        format = kAscii;
      else if (word == "binary_little_endian")
        format = kLE;
      else if (word == "binary_big_endian")
        fail("big-endian PLY is not supported");
      else
        fail("unknown PLY format " + word);
    } else if (word == "element") {
      if (!format) fail("element before format");
      if (nelems >= 2) fail("too many PLY elements");
      std::string name;
      size_t n = 0;
      ss >> name >> n;
      if (name == "vertex") {
        elem = 0;
        nv = n;
        prop_state = 0;
      } else if (name == "face") {
        elem = 1;
        nf = n;
      } else {
        fail("unknown element: " + name);
      }
      order.push_back(elem);
      ++nelems;
    } else if (word == "property") {
      if (!format) fail("property before format");
      if (elem == 0) {
        std::string dtype, name;
        ss >> dtype >> name;
        if (name == "x" || name == "y" || name == "z") {
          if (dtype != "float" && dtype != "float32") fail("vertex coordinates must be float");
          int want = (name == "x") ? 0 : (name == "y") ? 1 : 2;
          if (prop_state != want) fail("vertex properties out of order");
          prop_state = want + 1;
        } else if (name == "red" || name == "green" || name == "blue") {
          if (dtype != "uchar" && dtype != "uint8") fail("vertex colors must be uchar");
          int want = (name == "red") ? 3 : (name == "green") ? 4 : 5;
          if (prop_state != want) fail("color properties out of order");
          prop_state = want + 1;
          has_color = true;
        } else {
          fail("unknown property " + name);
        }
      } else if (elem == 1) {
        std::string w, lt, it, name;
        ss >> w >> lt >> it >> name;
        if (w != "list") fail("face property must be a list");
        list_size_bytes = dtypeSize(lt);
        index_bytes = dtypeSize(it);
        if (name != "vertex_index" && name != "vertex_indices") fail("unknown property " + name);
      } else {
        fail("property before element");
      }
    } else if (word == "end_header") {
      end_header = true;
      break;
    }
  }
  if (!end_header || !format || nv == 0 || nf == 0) fail("malformed PLY header");
  if (has_color && prop_state != 6) fail("incomplete color properties");
  m->v.assign(nv * 3, 0.f);
  m->f.assign(nf * 3, 0u);
  m->rgb.assign(nv, 0u);
  m->has_color = has_color;
  for (int e : order) {
    if (e == 0) {
      for (size_t n = 0; n < nv; ++n) {
        float x = 0, y = 0, z = 0;
        uint32_t r = 0, g = 0, b = 0;
        if (format == kLE) {
          in.read((char*)&x, 4);
          in.read((char*)&y, 4);
          in.read((char*)&z, 4);
          if (has_color) {
            in.read((char*)&r, 1);
            in.read((char*)&g, 1);
            in.read((char*)&b, 1);
          }
        } else {
          std::getline(in, line);
          std::istringstream ss(line);
          ss >> x >> y >> z;
          if (has_color) ss >> r >> g >> b;
        }
        if (!in) fail("truncated PLY vertex data");
        m->v[n * 3] = x;
        m->v[n * 3 + 1] = y;
        m->v[n * 3 + 2] = z;
        if (has_color) m->rgb[n] = (r << 16) | (g << 8) | b;
      }
    } else {
      for (size_t n = 0; n < nf; ++n) {
        int cnt = 0;
        uint32_t a = 0, b = 0, c = 0;
        if (format == kLE) {
          in.read((char*)&cnt, list_size_bytes);
          in.read((char*)&a, index_bytes);
          in.read((char*)&b, index_bytes);
          in.read((char*)&c, index_bytes);
        } else {
          std::getline(in, line);
          std::istringstream ss(line);
          ss >> cnt >> a >> b >> c;
        }
        if (!in) fail("truncated PLY face data");
        if (cnt != 3) fail("non-triangle face");
        if (a >= nv || b >= nv || c >= nv) fail("face index out of range");
        m->f[n * 3] = a;
        m->f[n * 3 + 1] = b;
        m->f[n * 3 + 2] = c;
      }
    }
  }
  *nv_out = nv;
  *nf_out = nf;
}

void computeNormals(Mesh* m) {
  size_t nv = m->v.size() / 3, nf = m->f.size() / 3;
  m->n.assign(nv * 3, 0.f);
  const float* V = m->v.data();
  float* N = m->n.data();
  for (size_t i = 0; i < nf; ++i) {
    size_t a = m->f[i * 3] * 3, b = m->f[i * 3 + 1] * 3, c = m->f[i * 3 + 2] * 3;
    V3 v0 = v3(V[a], V[a + 1], V[a + 2]);
    V3 v1 = v3(V[b], V[b + 1], V[b + 2]);
    V3 v2 = v3(V[c], V[c + 1], V[c + 2]);
    V3 u = v1 - v0, w = v2 - v0;
    float nx = u.y * w.z - u.z * w.y;
    float ny = u.z * w.x - u.x * w.z;
    float nz = u.x * w.y - u.y * w.x;
    N[a] += nx;
    N[a + 1] += ny;
    N[a + 2] += nz;
    N[b] += nx;
    N[b + 1] += ny;
    N[b + 2] += nz;
    N[c] += nx;
    N[c + 1] += ny;
    N[c + 2] += nz;
  }
}

void finishMesh(Domain* d) {
  if (!isIdentity(d->xf)) {
    size_t nv = d->mesh.v.size() / 3;
    for (size_t i = 0; i < nv; ++i) {
      float p[3] = {d->mesh.v[i * 3], d->mesh.v[i * 3 + 1], d->mesh.v[i * 3 + 2]}, o[3];
      mxform(d->xf, p, o);
      d->mesh.v[i * 3] = o[0];
      d->mesh.v[i * 3 + 1] = o[1];
      d->mesh.v[i * 3 + 2] = o[2];
    }
  }
  computeNormals(&d->mesh);
}

void interpolate(const Mesh& m, int prim, float u, float v, uint32_t* color, float Ns[3]) {
  uint32_t i0 = m.f[prim * 3], i1 = m.f[prim * 3 + 1], i2 = m.f[prim * 3 + 2];
  uint32_t c[3] = {m.rgb[i0], m.rgb[i1], m.rgb[i2]};
  uint32_t rgb[9];
  for (int k = 0; k < 3; ++k) {
    rgb[k * 3] = (c[k] >> 16) & 0xff;
    rgb[k * 3 + 1] = (c[k] >> 8) & 0xff;
    rgb[k * 3 + 2] = c[k] & 0xff;
  }
  float w = 1.f - u - v;
  uint32_t r = (uint32_t)((float)rgb[0] * w + (float)rgb[3] * u + (float)rgb[6] * v);
  uint32_t g = (uint32_t)((float)rgb[1] * w + (float)rgb[4] * u + (float)rgb[7] * v);
  uint32_t b = (uint32_t)((float)rgb[2] * w + (float)rgb[5] * u + (float)rgb[8] * v);
  *color = (r << 16) | (g << 8) | b;
  const float* n0 = &m.n[i0 * 3];
  const float* n1 = &m.n[i1 * 3];
  const float* n2 = &m.n[i2 * 3];
  for (int k = 0; k < 3; ++k) Ns[k] = (n0[k] * w) + (n1[k] * u) + (n2[k] * v);
}

// ---------------------------------------------------------------------------
// Watertight triangle test
RayPre rayPrecompute(const float dir[3]) {
  RayPre p;
  float ax = std::fabs(dir[0]), ay = std::fabs(dir[1]), az = std::fabs(dir[2]);
  int kz = (ax > ay) ? ((ax > az) ? 0 : 2) : ((ay > az) ? 1 : 2);
  int kx = kz + 1;
  if (kx == 3) kx = 0;
  int ky = kx + 1;
  if (ky == 3) ky = 0;
  if (dir[kz] < 0.0f) std::swap(kx, ky);
  p.kx = kx;
  p.ky = ky;
  p.kz = kz;
  p.Sx = dir[kx] / dir[kz];
  p.Sy = dir[ky] / dir[kz];
  p.Sz = 1.0f / dir[kz];
  return p;
}

bool triIntersect(const RayPre& rp, const float org[3], const float* v0, const float* v1,
                  const float* v2, float tnear, float tfar, float* t_o, float* u_o,
                  float* v_o) {
  const int kx = rp.kx, ky = rp.ky, kz = rp.kz;
  float A[3] = {v0[0] - org[0], v0[1] - org[1], v0[2] - org[2]};
  float B[3] = {v1[0] - org[0], v1[1] - org[1], v1[2] - org[2]};
  float C[3] = {v2[0] - org[0], v2[1] - org[1], v2[2] - org[2]};
  float Ax = A[kx] - rp.Sx * A[kz];
  float Ay = A[ky] - rp.Sy * A[kz];
  float Bx = B[kx] - rp.Sx * B[kz];
  float By = B[ky] - rp.Sy * B[kz];
  float Cx = C[kx] - rp.Sx * C[kz];
  float Cy = C[ky] - rp.Sy * C[kz];
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
  float Az = rp.Sz * A[kz];
  float Bz = rp.Sz * B[kz];
  float Cz = rp.Sz * C[kz];
  float T = U * Az + V * Bz + W * Cz;
  float rcp = 1.0f / det;
  float t = T * rcp;
  if (!(t > tnear && t < tfar)) return false;
  *t_o = t;
  *u_o = V * rcp;
  *v_o = W * rcp;
  return true;
}

// ---------------------------------------------------------------------------
// Binned-SAH BVH (the oracle's own; the GPU uses its own LBVH)
namespace {

struct Box {
  float lo[3], hi[3];
  void reset() {
    for (int i = 0; i < 3; ++i) {
      lo[i] = kInf;
      hi[i] = -kInf;
    }
  }
  void grow(const float p[3]) {
    for (int i = 0; i < 3; ++i) {
      lo[i] = std::min(lo[i], p[i]);
      hi[i] = std::max(hi[i], p[i]);
    }
  }
  void grow(const Box& b) {
    for (int i = 0; i < 3; ++i) {
      lo[i] = std::min(lo[i], b.lo[i]);
      hi[i] = std::max(hi[i], b.hi[i]);
    }
  }
  float area() const {
    float d[3];
    for (int i = 0; i < 3; ++i) d[i] = std::max(0.f, hi[i] - lo[i]);
    return 2.f * (d[0] * d[1] + d[1] * d[2] + d[2] * d[0]);
  }
};

struct Builder {
  std::vector<Box> tb;
  std::vector<float> cen;
  std::vector<BvhNode>* nodes;
  std::vector<uint32_t>* prims;

  // fills the pre-allocated node `slot` with the subtree over prims[first, first+count)
  void build(int slot, int first, int count) {
    Box b, cb;
    b.reset();
    cb.reset();
    for (int i = first; i < first + count; ++i) {
      uint32_t p = (*prims)[i];
      b.grow(tb[p]);
      cb.grow(&cen[p * 3]);
    }
    auto fillBox = [&]() {
      BvhNode& n = (*nodes)[slot];
      for (int k = 0; k < 3; ++k) {
        n.lo[k] = b.lo[k];
        n.hi[k] = b.hi[k];
      }
    };
    if (count <= 4) {
      fillBox();
      BvhNode& n = (*nodes)[slot];
      n.left = -1;
      n.first = first;
      n.count = count;
      return;
    }
    const int kBins = 16;
    float best_cost = kInf;
    int best_axis = -1, best_split = -1;
    for (int axis = 0; axis < 3; ++axis) {
      float lo = cb.lo[axis], hi = cb.hi[axis];
      if (!(hi > lo)) continue;
      Box bins[kBins];
      int cnt[kBins] = {0};
      for (int k = 0; k < kBins; ++k) bins[k].reset();
      float scale = kBins / (hi - lo);
      for (int i = first; i < first + count; ++i) {
        uint32_t p = (*prims)[i];
        int bi = std::min(kBins - 1, (int)((cen[p * 3 + axis] - lo) * scale));
        cnt[bi]++;
        bins[bi].grow(tb[p]);
      }
      float rarea[kBins];
      int rcnt[kBins];
      Box acc;
      acc.reset();
      int c = 0;
      for (int k = kBins - 1; k > 0; --k) {
        acc.grow(bins[k]);
        c += cnt[k];
        rarea[k] = acc.area();
        rcnt[k] = c;
      }
      acc.reset();
      c = 0;
      for (int k = 0; k < kBins - 1; ++k) {
        acc.grow(bins[k]);
        c += cnt[k];
        if (c == 0 || rcnt[k + 1] == 0) continue;
        float cost = acc.area() * c + rarea[k + 1] * rcnt[k + 1];
        if (cost < best_cost) {
          best_cost = cost;
          best_axis = axis;
          best_split = k;
        }
      }
    }
    int mid;
    if (best_axis < 0) {
      mid = first + count / 2;  // all centroids equal: split in the middle
    } else {
      float lo = cb.lo[best_axis], hi = cb.hi[best_axis];
      float scale = kBins / (hi - lo);
      auto it = std::partition(prims->begin() + first, prims->begin() + first + count,
                               [&](uint32_t p) {
                                 int bi = std::min(
                                     kBins - 1, (int)((cen[p * 3 + best_axis] - lo) * scale));
                                 return bi <= best_split;
                               });
      mid = (int)(it - prims->begin());
      if (mid == first || mid == first + count) mid = first + count / 2;
    }
    int L = (int)nodes->size();  // children adjacent: left at L, right at L+1
    nodes->push_back(BvhNode());
    nodes->push_back(BvhNode());
    build(L, first, mid - first);
    build(L + 1, mid, first + count - mid);
    fillBox();
    BvhNode& n = (*nodes)[slot];
    n.left = L;
    n.first = 0;
    n.count = 0;
  }
};

inline bool boxHit(const BvhNode& n, const float org[3], const float inv[3], float tnear,
                   float tfar, float* tentry) {
  float t0 = tnear, t1 = tfar;
  for (int k = 0; k < 3; ++k) {
    float a = (n.lo[k] - org[k]) * inv[k];
    float b = (n.hi[k] - org[k]) * inv[k];
    // a ray parallel to the slab with its origin on a slab plane gives
    // 0 * inf = NaN: inside the slab, no bound on this axis
    if (a != a || b != b) continue;
    t0 = std::fmax(t0, std::fmin(a, b));
    t1 = std::fmin(t1, std::fmax(a, b));
  }
  // robust slab test (Ize 2013): widen the far bound by 1 + 2*gamma(3)
  t1 = t1 * 1.0000003576278687f;
  *tentry = t0;
  return t0 <= t1;
}

}  // namespace

void buildBvh(Mesh* m) {
  size_t nf = m->f.size() / 3;
  m->nodes.clear();
  m->prims.resize(nf);
  for (size_t i = 0; i < nf; ++i) m->prims[i] = (uint32_t)i;
  if (nf == 0) return;
  Builder b;
  b.tb.resize(nf);
  b.cen.resize(nf * 3);
  for (size_t i = 0; i < nf; ++i) {
    b.tb[i].reset();
    for (int k = 0; k < 3; ++k) b.tb[i].grow(&m->v[m->f[i * 3 + k] * 3]);
    for (int k = 0; k < 3; ++k) b.cen[i * 3 + k] = 0.5f * (b.tb[i].lo[k] + b.tb[i].hi[k]);
  }
  b.nodes = &m->nodes;
  b.prims = &m->prims;
  m->nodes.reserve(nf);
  m->nodes.push_back(BvhNode());
  b.build(0, 0, (int)nf);
}

Hit closestHit(const Mesh& m, const float org[3], const float dir[3], float tnear, float tfar,
               uint64_t* node_visits, uint64_t* tri_tests) {
  Hit h{tfar, 0, 0, -1};
  if (m.nodes.empty()) return h;
  RayPre rp = rayPrecompute(dir);
  float inv[3] = {1.0f / dir[0], 1.0f / dir[1], 1.0f / dir[2]};
  int stack[128];
  int sp = 0;
  stack[sp++] = 0;
  uint64_t nv = 0, nt = 0;
  while (sp) {
    int ni = stack[--sp];
    const BvhNode& n = m.nodes[ni];
    float te;
    ++nv;
    if (!boxHit(n, org, inv, tnear, h.t, &te)) continue;
    if (n.left < 0) {
      for (int i = n.first; i < n.first + n.count; ++i) {
        uint32_t p = m.prims[i];
        const float* v0 = &m.v[m.f[p * 3] * 3];
        const float* v1 = &m.v[m.f[p * 3 + 1] * 3];
        const float* v2 = &m.v[m.f[p * 3 + 2] * 3];
        float t, u, v;
        ++nt;
        // t == current best is accepted only for a lower primID (tie rule)
        if (triIntersect(rp, org, v0, v1, v2, tnear, tfar, &t, &u, &v)) {
          if (h.prim < 0 || t < h.t || (t == h.t && (int)p < h.prim)) {
            h.t = t;
            h.u = u;
            h.v = v;
            h.prim = (int)p;
          }
        }
      }
    } else {
      const BvhNode& a = m.nodes[n.left];
      const BvhNode& b = m.nodes[n.left + 1];
      float ta, tb;
      bool ha = boxHit(a, org, inv, tnear, h.t, &ta);
      bool hb = boxHit(b, org, inv, tnear, h.t, &tb);
      if (ha && hb) {
        if (ta <= tb) {
          stack[sp++] = n.left + 1;
          stack[sp++] = n.left;
        } else {
          stack[sp++] = n.left;
          stack[sp++] = n.left + 1;
        }
      } else if (ha) {
        stack[sp++] = n.left;
      } else if (hb) {
        stack[sp++] = n.left + 1;
      }
    }
  }
  if (node_visits) *node_visits += nv;
  if (tri_tests) *tri_tests += nt;
  if (h.prim < 0) h.t = tfar;
  return h;
}

Hit closestHitBrute(const Mesh& m, const float org[3], const float dir[3], float tnear,
                    float tfar) {
  Hit h{tfar, 0, 0, -1};
  RayPre rp = rayPrecompute(dir);
  size_t nf = m.f.size() / 3;
  for (size_t p = 0; p < nf; ++p) {
    float t, u, v;
    if (triIntersect(rp, org, &m.v[m.f[p * 3] * 3], &m.v[m.f[p * 3 + 1] * 3],
                     &m.v[m.f[p * 3 + 2] * 3], tnear, tfar, &t, &u, &v)) {
      if (h.prim < 0 || t < h.t) {  // ascending primID: ties keep the first
        h.t = t;
        h.u = u;
        h.v = v;
        h.prim = (int)p;
      }
    }
  }
  return h;
}

bool anyHit(const Mesh& m, const float org[3], const float dir[3], float tnear, float tfar) {
  if (m.nodes.empty()) return false;
  RayPre rp = rayPrecompute(dir);
  float inv[3] = {1.0f / dir[0], 1.0f / dir[1], 1.0f / dir[2]};
  int stack[128];
  int sp = 0;
  stack[sp++] = 0;
  while (sp) {
    const BvhNode& n = m.nodes[stack[--sp]];
    float te;
    if (!boxHit(n, org, inv, tnear, tfar, &te)) continue;
    if (n.left < 0) {
      for (int i = n.first; i < n.first + n.count; ++i) {
        uint32_t p = m.prims[i];
        float t, u, v;
        if (triIntersect(rp, org, &m.v[m.f[p * 3] * 3], &m.v[m.f[p * 3 + 1] * 3],
                         &m.v[m.f[p * 3 + 2] * 3], tnear, tfar, &t, &u, &v))
          return true;
      }
    } else {
      stack[sp++] = n.left;
      stack[sp++] = n.left + 1;
    }
  }
  return false;
}

}  // namespace orc
