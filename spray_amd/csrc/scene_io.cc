// Host-side readers for SpRay's scene description (.spray) and PLY meshes.
// Accepted subset and checks follow io/scene_loader.cc:42-358 and
// io/ply_loader.cc:114-438; every failure becomes a status code instead of
// a glog CHECK abort.  New scene token: `light rect cx cy cz e1x e1y e1z
// e2x e2y e2z r g b` (a one-sided parallelogram emitter facing
// cross(e1, e2)).
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>

#include "spray_internal.h"

namespace spray {
namespace {

std::vector<std::string> split(const std::string& line) {
  // strtok(line, " ") (scene_loader.cc:333-342): spaces only
  std::vector<std::string> out;
  size_t i = 0;
  while (i < line.size()) {
    while (i < line.size() && line[i] == ' ') ++i;
    size_t j = i;
    while (j < line.size() && line[j] != ' ') ++j;
    if (j > i) out.emplace_back(line.substr(i, j - i));
    i = j;
  }
  return out;
}

float num(const std::string& s) { return (float)atof(s.c_str()); }

void setIdentity(float m[16]) {
  for (int i = 0; i < 16; ++i) m[i] = (i % 5 == 0) ? 1.f : 0.f;
}

// glm::translate / glm::scale / glm::rotate on a column-major mat4
void applyTranslate(float m[16], float x, float y, float z) {
  for (int r = 0; r < 4; ++r) m[12 + r] = m[r] * x + m[4 + r] * y + m[8 + r] * z + m[12 + r];
}
void applyScale(float m[16], float x, float y, float z) {
  const float s[3] = {x, y, z};
  for (int c = 0; c < 3; ++c)
    for (int r = 0; r < 4; ++r) m[c * 4 + r] = m[c * 4 + r] * s[c];
}
void applyRotate(float m[16], float angle, const float axis_in[3]) {
  float c = std::cos(angle), s = std::sin(angle);
  float l2 = axis_in[0] * axis_in[0] + axis_in[1] * axis_in[1] + axis_in[2] * axis_in[2];
  float inv = 1.0f / std::sqrt(l2);
  float a[3] = {axis_in[0] * inv, axis_in[1] * inv, axis_in[2] * inv};
  float t[3] = {(1.f - c) * a[0], (1.f - c) * a[1], (1.f - c) * a[2]};
  float R[3][3] = {{c + t[0] * a[0], t[0] * a[1] + s * a[2], t[0] * a[2] - s * a[1]},
                   {t[1] * a[0] - s * a[2], c + t[1] * a[1], t[1] * a[2] + s * a[0]},
                   {t[2] * a[0] + s * a[1], t[2] * a[1] - s * a[0], c + t[2] * a[2]}};
  float out[16];
  for (int j = 0; j < 3; ++j)
    for (int r = 0; r < 4; ++r)
      out[j * 4 + r] = m[r] * R[j][0] + m[4 + r] * R[j][1] + m[8 + r] * R[j][2];
  for (int r = 0; r < 4; ++r) out[12 + r] = m[12 + r];
  std::memcpy(m, out, sizeof(out));
}

}  // namespace

bool transformIsIdentity(const float xf[16]) {
  for (int i = 0; i < 16; ++i)
    if (xf[i] != ((i % 5 == 0) ? 1.f : 0.f)) return false;
  return true;
}

// glm mat4 * vec4(p, 1): (m0*x + m1*y) + (m2*z + m3*1)
void transformPoint(const float xf[16], const float p[3], float out[3]) {
  for (int r = 0; r < 3; ++r) {
    float a = xf[r] * p[0] + xf[4 + r] * p[1];
    float b = xf[8 + r] * p[2] + xf[12 + r] * 1.0f;
    out[r] = a + b;
  }
}

namespace {

int parseSceneStream(std::istream& in, const std::string& spray_file, const std::string& ply_path,
                     HostScene* s, std::string* err) {
  s->domains.clear();
  s->lights.clear();
  auto bad = [&](const std::string& m) {
    *err = spray_file + ": " + m;
    return SPRAY_ERR_FORMAT;
  };
  std::string line;
  int lineno = 0;
  while (std::getline(in, line)) {
    ++lineno;
    std::vector<std::string> t = split(line);
    if (t.empty() || t[0][0] == '#') continue;
    const std::string& tag = t[0];
    std::string where = " (line " + std::to_string(lineno) + ")";
    if (tag == "domain") {
      HostDomain d;
      d.id = (int)s->domains.size();
      setIdentity(d.xf);
      s->domains.push_back(d);
      continue;
    }
    if (tag == "light") {
      if (t.size() < 2) return bad("light without type" + where);
      HostLight L;
      if (t[1] == "point") {
        if (t.size() != 8) return bad("light point takes 6 numbers" + where);
        L.type = kPointLight;
        for (int k = 0; k < 3; ++k) {
          L.pos[k] = num(t[2 + k]);
          L.radiance[k] = num(t[5 + k]);
        }
      } else if (t[1] == "diffuse") {
        if (t.size() != 5) return bad("light diffuse takes 3 numbers" + where);
        L.type = kHemisphereLight;
        for (int k = 0; k < 3; ++k) L.radiance[k] = num(t[2 + k]);
      } else if (t[1] == "rect") {
        if (t.size() != 14) return bad("light rect takes 12 numbers" + where);
        L.type = kRectLight;
        for (int k = 0; k < 3; ++k) {
          L.pos[k] = num(t[2 + k]);
          L.e1[k] = num(t[5 + k]);
          L.e2[k] = num(t[8 + k]);
          L.radiance[k] = num(t[11 + k]);
        }
        float c[3] = {L.e1[1] * L.e2[2] - L.e2[1] * L.e1[2], L.e1[2] * L.e2[0] - L.e2[2] * L.e1[0],
                      L.e1[0] * L.e2[1] - L.e2[0] * L.e1[1]};
        float l2 = c[0] * c[0] + c[1] * c[1] + c[2] * c[2];
        L.area = std::sqrt(l2);
        if (!(L.area > 0.f)) return bad("degenerate rect light" + where);
        float inv = 1.0f / std::sqrt(l2);
        for (int k = 0; k < 3; ++k) L.normal[k] = c[k] * inv;
      } else {
        return bad("unknown light source " + t[1] + where);
      }
      s->lights.push_back(L);
      continue;
    }
    if (tag != "file" && tag != "mtl" && tag != "bound" && tag != "scale" && tag != "rotate" &&
        tag != "translate" && tag != "face" && tag != "vertex")
      return bad("unknown tag name " + tag + where);
    if (s->domains.empty()) return bad("'" + tag + "' before the first 'domain'" + where);
    HostDomain& d = s->domains.back();
    if (tag == "file") {
      if (t.size() != 2) return bad("file takes 1 name" + where);
      d.filename = ply_path.empty() ? t[1] : ply_path + "/" + t[1];
    } else if (tag == "mtl") {
      if (t.size() < 2) return bad("mtl without type" + where);
      if (t[1] == "diffuse" || t[1] == "mirror") {
        if (t.size() != 5) return bad("mtl " + t[1] + " takes 3 numbers" + where);
        d.mtl = t[1] == "diffuse" ? kDiffuse : kMirror;
        for (int k = 0; k < 3; ++k) d.mtl_rgb[k] = num(t[2 + k]);
      } else if (t[1] == "glass" || t[1] == "transmission") {
        if (t.size() != 4) return bad("mtl " + t[1] + " takes 2 numbers" + where);
        d.mtl = t[1] == "glass" ? kGlass : kTransmission;
        d.eta_a = num(t[2]);
        d.eta_b = num(t[3]);
      } else {
        return bad("unknown material type " + t[1] + where);
      }
    } else if (tag == "bound") {
      if (t.size() != 7) return bad("bound takes 6 numbers" + where);
      for (int k = 0; k < 3; ++k) {
        d.obj_lo[k] = num(t[1 + k]);
        d.obj_hi[k] = num(t[4 + k]);
      }
      d.has_bound = true;
    } else if (tag == "scale") {
      if (t.size() != 4) return bad("scale takes 3 numbers" + where);
      applyScale(d.xf, num(t[1]), num(t[2]), num(t[3]));
    } else if (tag == "rotate") {
      if (t.size() != 3) return bad("rotate takes an axis and an angle" + where);
      float axis[3] = {0, 0, 0};
      if (t[1] == "x")
        axis[0] = 1;
      else if (t[1] == "y")
        axis[1] = 1;
      else if (t[1] == "z")
        axis[2] = 1;
      else
        return bad("invalid axis name " + t[1] + where);
      applyRotate(d.xf, (float)(atof(t[2].c_str()) * 0.01745329251994329576923690768489), axis);
    } else if (tag == "translate") {
      if (t.size() != 4) return bad("translate takes 3 numbers" + where);
      applyTranslate(d.xf, num(t[1]), num(t[2]), num(t[3]));
    } else if (tag == "face" || tag == "vertex") {
      if (t.size() != 2) return bad(tag + " takes 1 count" + where);
      char* endp = nullptr;
      long long v = std::strtoll(t[1].c_str(), &endp, 10);
      if (!endp || *endp || v < 0) return bad("bad count " + t[1] + where);
      (tag == "face" ? d.num_faces : d.num_vertices) = v;
    }
  }
  if (s->domains.empty()) return bad("no domain");
  if ((int)s->domains.size() > kMaxDomains) return bad("too many domains");
  for (int k = 0; k < 3; ++k) {
    s->lo[k] = INFINITY;
    s->hi[k] = -INFINITY;
  }
  for (auto& d : s->domains) {
    // world AABB = transform * object bound corners (scene_loader.cc:351-357)
    transformPoint(d.xf, d.obj_lo, d.lo);
    transformPoint(d.xf, d.obj_hi, d.hi);
    // mergeDomainBounds requirements (render/scene.inl:314-317)
    if (!d.has_bound || !(d.lo[0] <= d.hi[0] && d.lo[1] <= d.hi[1] && d.lo[2] <= d.hi[2]))
      return bad("domain " + std::to_string(d.id) + " has an invalid bound");
    if (d.num_vertices <= 0 || d.num_faces <= 0)
      return bad("domain " + std::to_string(d.id) + " lacks vertex/face counts");
    if (d.filename.empty()) return bad("domain " + std::to_string(d.id) + " has no file");
    for (int k = 0; k < 3; ++k) {
      s->lo[k] = std::fmin(s->lo[k], d.lo[k]);
      s->hi[k] = std::fmax(s->hi[k], d.hi[k]);
    }
  }
  return SPRAY_OK;
}

}  // namespace

int parseSceneFile(const std::string& spray_file, const std::string& ply_path, HostScene* s,
                   std::string* err) {
  std::ifstream in(spray_file);
  if (!in.is_open()) {
    *err = "unable to open input file " + spray_file;
    return SPRAY_ERR_IO;
  }
  return parseSceneStream(in, spray_file, ply_path, s, err);
}

int parseSceneText(const std::string& text, const std::string& name, const std::string& ply_path,
                   HostScene* s, std::string* err) {
  std::istringstream in(text);
  return parseSceneStream(in, name, ply_path, s, err);
}

namespace {

// Parses the header; `in` is left at the first element's data.
int parsePlyHeader(std::istream& in, const std::string& filename, PlyLayout* L, std::string* err) {
  auto bad = [&](const std::string& msg) {
    *err = filename + ": " + msg;
    return SPRAY_ERR_FORMAT;
  };
  std::string line;
  if (!std::getline(in, line) || line != "ply") return bad("unknown file type");
  enum { kNone, kAscii, kLittle } fmt = kNone;
  long long nv = 0, nf = 0;
  int cur = -1, xyz = 0, rgb = 0;
  int list_bytes = 0, index_bytes = 0;
  std::vector<int> order;
  bool ended = false;
  auto size_of = [](const std::string& t) -> int {
    if (t == "char" || t == "uchar" || t == "int8" || t == "uint8") return 1;
    if (t == "short" || t == "ushort" || t == "int16" || t == "uint16") return 2;
    if (t == "int" || t == "uint" || t == "int32" || t == "uint32") return 4;
    return 0;
  };
  while (std::getline(in, line)) {
    std::istringstream ss(line);
    std::string w;
    ss >> w;
    if (w == "format") {
      ss >> w;
      if (w == "ascii")
        fmt = kAscii;
      else if (w == "binary_little_endian")
        fmt = kLittle;
      else if (w == "binary_big_endian")
        return bad("big-endian PLY is not supported");  // ply_loader.cc:148
      else
        return bad("unknown format " + w);
    } else if (w == "element") {
      if (fmt == kNone) return bad("element before format");
      if (order.size() >= 2) return bad("more than two elements");
      std::string name;
      long long n = -1;
      ss >> name >> n;
      if (n <= 0) return bad("element count must be positive");
      if (name == "vertex") {
        cur = 0;
        nv = n;
      } else if (name == "face") {
        cur = 1;
        nf = n;
      } else {
        return bad("unknown element: " + name);
      }
      order.push_back(cur);
    } else if (w == "property") {
      if (cur == 0) {
        std::string type, name;
        ss >> type >> name;
        static const char* seq[6] = {"x", "y", "z", "red", "green", "blue"};
        int pos = xyz + rgb;
        if (pos >= 6 || name != seq[pos]) return bad("unexpected vertex property " + name);
        if (pos < 3) {
          if (type != "float" && type != "float32") return bad("coordinates must be float");
          ++xyz;
        } else {
          if (type != "uchar" && type != "uint8") return bad("colours must be uchar");
          ++rgb;
        }
      } else if (cur == 1) {
        std::string list, lt, it, name;
        ss >> list >> lt >> it >> name;
        if (list != "list") return bad("face property must be a list");
        list_bytes = size_of(lt);
        index_bytes = size_of(it);
        if (!list_bytes || !index_bytes) return bad("unsupported list types");
        if (name != "vertex_index" && name != "vertex_indices")
          return bad("unknown property " + name);
      } else {
        return bad("property before element");
      }
    } else if (w == "end_header") {
      ended = true;
      break;
    }
  }
  if (!ended || fmt == kNone || nv <= 0 || nf <= 0) return bad("incomplete header");
  if (xyz != 3 || (rgb != 0 && rgb != 3)) return bad("vertex needs x y z [red green blue]");
  if (!list_bytes) return bad("face element has no index list");
  L->binary = fmt == kLittle;
  L->nv = nv;
  L->nf = nf;
  L->rgb = rgb != 0;
  L->list_bytes = list_bytes;
  L->index_bytes = index_bytes;
  L->faces_first = order[0] == 1;
  L->body = (long long)in.tellg();
  return SPRAY_OK;
}

}  // namespace

int readPlyHeader(const std::string& filename, PlyLayout* L, std::string* err) {
  std::ifstream in(filename, std::ios::in | std::ios::binary);
  if (!in.is_open()) {
    *err = "unable to open " + filename;
    return SPRAY_ERR_IO;
  }
  int rc = parsePlyHeader(in, filename, L, err);
  if (rc != SPRAY_OK) return rc;
  in.seekg(0, std::ios::end);
  L->file_bytes = (long long)in.tellg();
  return SPRAY_OK;
}

int readPly(const std::string& filename, HostMesh* m, std::string* err) {
  std::ifstream in(filename, std::ios::in | std::ios::binary);
  if (!in.is_open()) {
    *err = "unable to open " + filename;
    return SPRAY_ERR_IO;
  }
  auto bad = [&](const std::string& msg) {
    *err = filename + ": " + msg;
    return SPRAY_ERR_FORMAT;
  };
  PlyLayout L;
  int rc = parsePlyHeader(in, filename, &L, err);
  if (rc != SPRAY_OK) return rc;
  const long long nv = L.nv, nf = L.nf;
  const bool rgb = L.rgb;
  const int list_bytes = L.list_bytes, index_bytes = L.index_bytes;
  enum { kAscii, kLittle } fmt = L.binary ? kLittle : kAscii;
  const int order[2] = {L.faces_first ? 1 : 0, L.faces_first ? 0 : 1};
  std::string line;
  m->v.resize((size_t)nv * 3);
  m->f.resize((size_t)nf * 3);
  m->rgb.assign((size_t)nv, 0u);
  for (int e : order) {
    if (e == 0) {
      if (fmt == kLittle) {
        size_t stride = 12 + (rgb ? 3 : 0);
        std::vector<char> buf((size_t)nv * stride);
        if (!in.read(buf.data(), (std::streamsize)buf.size())) return bad("truncated vertices");
        for (long long i = 0; i < nv; ++i) {
          std::memcpy(&m->v[i * 3], &buf[i * stride], 12);
          if (rgb) {
            const unsigned char* c = (const unsigned char*)&buf[i * stride + 12];
            m->rgb[i] = ((uint32_t)c[0] << 16) | ((uint32_t)c[1] << 8) | c[2];
          }
        }
      } else {
        for (long long i = 0; i < nv; ++i) {
          if (!std::getline(in, line)) return bad("truncated vertices");
          std::istringstream ss(line);
          float x, y, z;
          ss >> x >> y >> z;
          m->v[i * 3] = x;
          m->v[i * 3 + 1] = y;
          m->v[i * 3 + 2] = z;
          if (rgb) {
            uint32_t r, g, b;
            ss >> r >> g >> b;
            m->rgb[i] = (r << 16) | (g << 8) | b;
          }
          if (!ss) return bad("malformed vertex line");
        }
      }
    } else {
      for (long long i = 0; i < nf; ++i) {
        long long cnt = 0;
        uint32_t idx[3] = {0, 0, 0};
        if (fmt == kLittle) {
          uint32_t c = 0;
          if (!in.read((char*)&c, list_bytes)) return bad("truncated faces");
          cnt = c;
          if (cnt == 3)
            for (int k = 0; k < 3; ++k)
              if (!in.read((char*)&idx[k], index_bytes)) return bad("truncated faces");
        } else {
          if (!std::getline(in, line)) return bad("truncated faces");
          std::istringstream ss(line);
          ss >> cnt >> idx[0] >> idx[1] >> idx[2];
          if (!ss) return bad("malformed face line");
        }
        if (cnt != 3) return bad("non-triangle face");  // ply_loader.cc:289
        for (int k = 0; k < 3; ++k) {
          if ((long long)idx[k] >= nv) return bad("face index out of range");
          m->f[i * 3 + k] = idx[k];
        }
      }
    }
  }
  return SPRAY_OK;
}

bool writePlyBinary(const std::string& path, const float* v, size_t nv, const uint32_t* f,
                    size_t nf, const uint32_t* rgb) {
  FILE* fp = std::fopen(path.c_str(), "wb");
  if (!fp) return false;
  std::fprintf(fp,
               "ply\nformat binary_little_endian 1.0\ncomment spray_amd marching cubes\n"
               "element vertex %zu\nproperty float x\nproperty float y\nproperty float z\n"
               "property uchar red\nproperty uchar green\nproperty uchar blue\n"
               "element face %zu\nproperty list uchar int vertex_indices\nend_header\n",
               nv, nf);
  std::vector<uint8_t> buf(nv * 15);
  for (size_t i = 0; i < nv; ++i) {
    std::memcpy(&buf[i * 15], &v[i * 3], 12);
    uint32_t c = rgb[i];
    buf[i * 15 + 12] = (c >> 16) & 0xff;
    buf[i * 15 + 13] = (c >> 8) & 0xff;
    buf[i * 15 + 14] = c & 0xff;
  }
  bool ok = std::fwrite(buf.data(), 1, buf.size(), fp) == buf.size();
  std::vector<uint8_t> fb(nf * 13);
  for (size_t i = 0; i < nf; ++i) {
    fb[i * 13] = 3;
    std::memcpy(&fb[i * 13 + 1], &f[i * 3], 12);
  }
  ok = std::fwrite(fb.data(), 1, fb.size(), fp) == fb.size() && ok;
  return std::fclose(fp) == 0 && ok;
}

// %.9g round-trips every float, so a written scene reads back bit-exactly.
std::string sceneLightLine(const HostLight& L) {
  char b[512];
  if (L.type == kPointLight)
    std::snprintf(b, sizeof(b), "light point %.9g %.9g %.9g %.9g %.9g %.9g", L.pos[0], L.pos[1],
                  L.pos[2], L.radiance[0], L.radiance[1], L.radiance[2]);
  else if (L.type == kHemisphereLight)
    std::snprintf(b, sizeof(b), "light diffuse %.9g %.9g %.9g", L.radiance[0], L.radiance[1],
                  L.radiance[2]);
  else
    std::snprintf(b, sizeof(b),
                  "light rect %.9g %.9g %.9g %.9g %.9g %.9g %.9g %.9g %.9g %.9g %.9g %.9g",
                  L.pos[0], L.pos[1], L.pos[2], L.e1[0], L.e1[1], L.e1[2], L.e2[0], L.e2[1],
                  L.e2[2], L.radiance[0], L.radiance[1], L.radiance[2]);
  return b;
}

std::string sceneDomainBlock(int id, const std::string& file, int mtl, const float mtl_rgb[3],
                             float eta_a, float eta_b, const float lo[3], const float hi[3],
                             long long nv, long long nf) {
  char m[160], b[768];
  if (mtl == kDiffuse || mtl == kMirror)
    std::snprintf(m, sizeof(m), "mtl %s %.9g %.9g %.9g", mtl == kDiffuse ? "diffuse" : "mirror",
                  mtl_rgb[0], mtl_rgb[1], mtl_rgb[2]);
  else
    std::snprintf(m, sizeof(m), "mtl %s %.9g %.9g", mtl == kGlass ? "glass" : "transmission",
                  eta_a, eta_b);
  std::snprintf(b, sizeof(b),
                "# domain %d\ndomain\nfile %s\n%s\nbound %.9g %.9g %.9g %.9g %.9g %.9g\n"
                "vertex %lld\nface %lld\n\n",
                id, file.c_str(), m, lo[0], lo[1], lo[2], hi[0], hi[1], hi[2], nv, nf);
  return b;
}

}  // namespace spray
