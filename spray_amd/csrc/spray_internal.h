// Internal types of the spray_amd library (host side and device layouts).
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/spray_gpu.h"

namespace spray {

constexpr int kHistory = 4;         // SPRAY_SPECU_HISTORY_SIZE (src/CMakeLists.txt:33)
constexpr int kListSize = 16;       // SPRAY_RAY_DOMAIN_LIST_SIZE (src/CMakeLists.txt:32)
constexpr int kStatSlots = kListSize + 1;  // depths >= 16 share one zero-weight slot
constexpr int kMaxDomains = 1024;  // kCount/kScatter stage 48 bytes per domain in LDS
constexpr int kLeafMax = 2;         // triangles per BVH leaf

enum MaterialType { kDiffuse = 0, kMirror = 1, kGlass = 2, kTransmission = 3 };
enum LightType { kPointLight = 0, kHemisphereLight = 1, kRectLight = 2 };

struct HostMesh {
  std::vector<float> v;       // xyz per vertex
  std::vector<uint32_t> f;    // 3 per face
  std::vector<uint32_t> rgb;  // packed per vertex
};

struct HostDomain {
  int id = 0;
  long long num_vertices = 0, num_faces = 0;
  std::string filename;
  float obj_lo[3] = {0, 0, 0}, obj_hi[3] = {0, 0, 0};
  bool has_bound = false;  // a domain without a bound line is invalid (render/scene.inl:314)
  float lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
  float xf[16];
  int mtl = kDiffuse;
  float mtl_rgb[3] = {1, 1, 1};
  float eta_a = 1, eta_b = 1;
};

struct HostLight {
  int type = kPointLight;
  float pos[3] = {0, 0, 0};
  float e1[3] = {0, 0, 0}, e2[3] = {0, 0, 0};
  float radiance[3] = {0, 0, 0};
  float normal[3] = {0, 0, 0};
  float area = 0;
};

struct HostScene {
  std::vector<HostDomain> domains;
  std::vector<HostLight> lights;
  float lo[3], hi[3];
};

// Layout of a PLY file as its header describes it (io/ply_loader.cc:114-178).
struct PlyLayout {
  bool binary = false;      // binary_little_endian; else ascii
  long long nv = 0, nf = 0;
  bool rgb = false;         // vertex has red green blue (uchar) after x y z (float)
  int list_bytes = 0;       // face list count type size
  int index_bytes = 0;      // face list index type size
  bool faces_first = false; // the face element precedes the vertex element
  long long body = 0;       // byte offset of the first element's data
  long long file_bytes = 0;
  int vertexStride() const { return 12 + (rgb ? 3 : 0); }
  int faceStride() const { return list_bytes + 3 * index_bytes; }
};

// scene_io.cc
int parseSceneFile(const std::string& spray_file, const std::string& ply_path, HostScene* s,
                   std::string* err);
// the same grammar over text held in memory (`name` labels errors)
int parseSceneText(const std::string& text, const std::string& name, const std::string& ply_path,
                   HostScene* s, std::string* err);
int readPlyHeader(const std::string& filename, PlyLayout* L, std::string* err);
int readPly(const std::string& filename, HostMesh* m, std::string* err);
bool transformIsIdentity(const float xf[16]);
void transformPoint(const float xf[16], const float p[3], float out[3]);
// binary little-endian PLY: float x y z, uchar red green blue, list uchar int
bool writePlyBinary(const std::string& path, const float* v, size_t nv, const uint32_t* f,
                    size_t nf, const uint32_t* rgb);
// .spray lines of a light / of a domain with identity transform
std::string sceneLightLine(const HostLight& L);
std::string sceneDomainBlock(int id, const std::string& file, int mtl, const float mtl_rgb[3],
                             float eta_a, float eta_b, const float lo[3], const float hi[3],
                             long long nv, long long nf);

// scenegen.cc
void mcDomainParams(const spray_scenegen_params& p, int d, double* a, double* b, int* res);
void mcGenerate(const spray_scenegen_params& p, int d, HostMesh* m);
// marching-cubes case table: up to 8 triangles (3 local edge ids each) per
// case, -1 terminated; local edge e = (corner with 0 on axis e/4, axis)
const int8_t* mcTriTable();  // [256][25]
void mcEdgeCorner(int e, int* dx, int* dy, int* dz, int* axis);
int scenegenWrite(const spray_scenegen_params& p, const char* dir, const char* name,
                  const char* extra_lines, int64_t* tris_out, std::string* err);

// ---------------------------------------------------------------------------
// Device layouts (shared by host and kernels)

// One 4-wide BVH node: 128 bytes = one HBM/L2 line.  child[i] >= 0 is an
// inner node index (relative to the domain's node base); a negative value
// is a leaf: bit 31 set, bits 24..30 the triangle count, bits 0..23 the
// first triangle in leaf order.  Unused slots have an empty box.
struct alignas(128) BNode4 {
  float lox[4], hix[4], loy[4], hiy[4], loz[4], hiz[4];
  int32_t child[4];
  int32_t pad[4];
};
static_assert(sizeof(BNode4) == 128, "BNode4 must be 128 bytes");

// Triangle in leaf order: 48 bytes; v0.w holds the primitive id.
struct alignas(16) LTri {
  float v0[4], v1[4], v2[4];
};

struct DomainDesc {
  long long node_off;  // into the pool's node array
  long long tri_off;   // into the pool's leaf-ordered triangle array
  long long vtx_off;   // into vertex / normal / colour arrays
  long long face_off;  // into the face array (3 uint32 per face)
  long long shade_off; // into the per-face shading records (the domain's block when paged)
  int num_tris;
  int num_nodes;
  float lo[3], hi[3];
  int mtl;
  float eta_a, eta_b;
};

struct DevLight {
  int type;
  float pos[3], e1[3], e2[3], radiance[3], normal[3];
  float area;
};

}  // namespace spray
