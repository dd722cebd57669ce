// Oracle (test infrastructure only): scene, PLY and BVH restatement.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "omath.h"

namespace orc {

// render/reflection.h:219-364 (Diffuse/Mirror/Glass/Transmission)
enum MaterialType { kDiffuse = 0, kMirror = 1, kGlass = 2, kTransmission = 3 };
struct Material {
  int type = kDiffuse;
  float rgb[3] = {1, 1, 1};
  float eta_a = 1, eta_b = 1;
};

// render/light.h:41-89 (point, diffuse hemisphere) plus the rectangular
// area light that the reference does not have (scene token "light rect").
enum LightType { kPointLight = 0, kHemisphereLight = 1, kRectLight = 2 };
struct Light {
  int type = kPointLight;
  V3 pos{0, 0, 0};  // point position or rect corner
  V3 e1{0, 0, 0}, e2{0, 0, 0};
  V3 radiance{0, 0, 0};
  V3 normal{0, 0, 0};  // rect: normalize(cross(e1, e2))
  float area = 0;      // rect: |cross(e1, e2)|
  bool isArea() const { return type != kPointLight; }
};

struct BvhNode {
  float lo[3], hi[3];
  int left;   // internal: index of left child (right = left + 1); leaf: -1
  int first;  // leaf: first index into prim order
  int count;  // leaf: number of prims (0 for internal)
};

struct Mesh {
  std::vector<float> v;       // xyz per vertex
  std::vector<uint32_t> f;    // 3 indices per face
  std::vector<uint32_t> rgb;  // packed r<<16|g<<8|b per vertex
  std::vector<float> n;       // area-weighted vertex normals (unnormalised)
  bool has_color = false;
  // BVH (binned SAH, built by the oracle)
  std::vector<BvhNode> nodes;
  std::vector<uint32_t> prims;
};

// render/domain.h:32-44
struct Domain {
  int id = 0;
  size_t num_vertices = 0, num_faces = 0;
  std::string filename;
  float obj_lo[3] = {0, 0, 0}, obj_hi[3] = {0, 0, 0};
  float lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};  // world AABB
  float xf[16];                                 // column-major glm::mat4
  bool has_bound = false;
  Material mtl;
  Mesh mesh;
};

struct Scene {
  std::vector<Domain> domains;
  std::vector<Light> lights;
  float lo[3], hi[3];
};

// io/scene_loader.cc:42-358; throws std::runtime_error on malformed input.
void loadSceneFile(const std::string& spray_file, const std::string& ply_path, Scene* s);
// io/ply_loader.cc:114-324
void loadPly(const std::string& filename, Mesh* m, size_t* nv, size_t* nf);
// render/scene.inl:295-349 (mergeDomainBounds) + world bound.
void mergeBounds(Scene* s);
// render/trimesh_buffer.cc:117-169 (load: transform) and :267-326 (computeNormals)
void finishMesh(Domain* d);
void computeNormals(Mesh* m);
// Build the oracle's own BVH (binned SAH) over a mesh.
void buildBvh(Mesh* m);

struct Hit {
  float t, u, v;
  int prim;
};
// Closest hit over (tnear, tfar) using the BVH; ties on t go to the smaller
// primID.  Returns prim = -1 on a miss.
Hit closestHit(const Mesh& m, const float org[3], const float dir[3], float tnear, float tfar,
               uint64_t* node_visits = nullptr, uint64_t* tri_tests = nullptr);
// Brute force over all triangles (same triangle test, same tie rule).
Hit closestHitBrute(const Mesh& m, const float org[3], const float dir[3], float tnear,
                    float tfar);
bool anyHit(const Mesh& m, const float org[3], const float dir[3], float tnear, float tfar);

// Watertight ray/triangle test (Woop, Benthin, Wald, JCGT 2(1), 2013).
struct RayPre {
  int kx, ky, kz;
  float Sx, Sy, Sz;
};
RayPre rayPrecompute(const float dir[3]);
bool triIntersect(const RayPre& rp, const float org[3], const float* v0, const float* v1,
                  const float* v2, float tnear, float tfar, float* t, float* u, float* v);

// render/trimesh_buffer.cc:327-360 (updateIntersection)
void interpolate(const Mesh& m, int prim, float u, float v, uint32_t* color, float Ns[3]);

}  // namespace orc
