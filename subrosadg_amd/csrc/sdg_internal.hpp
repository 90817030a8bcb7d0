// sdg_internal.hpp -- declarations shared by the host translation units of libsdg_hip.so.
#ifndef SDG_INTERNAL_HPP_
#define SDG_INTERNAL_HPP_

#include <cstddef>
#include <map>
#include <string>
#include <vector>

struct sdg_mesh;

namespace sdg {
struct Mesh;
struct MixedMesh;
void setGlobalError(const std::string& msg);
const char* globalError();
// the single-type mesh of a handle (nullptr for a mesh with both triangles and quadrangles)
const Mesh* meshOf(const sdg_mesh* m);
// the triangle/quadrangle description of a split-grid mesh (nullptr for other meshes)
const MixedMesh* mixedOf(const sdg_mesh* m);
// zstd raw files in the reference's layout (sdg_raw.cpp): header + one frame of the payload
int writeRawFile(const char* path, const void* data, size_t bytes);
int readRawFile(const char* path, std::vector<char>& out);
// gmsh MSH 4.1 ASCII file -> single-type affine mesh; boundary types by physical tag (sdg_gmsh.cpp)
bool readGmshMesh(const char* path, int order, const std::map<int, int>& bc_of_physical, Mesh& mesh,
                  std::string& error);
}  // namespace sdg

#endif  // SDG_INTERNAL_HPP_
