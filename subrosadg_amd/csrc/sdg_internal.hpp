// sdg_internal.hpp -- declarations shared by the host translation units of libsdg_hip.so.
#ifndef SDG_INTERNAL_HPP_
#define SDG_INTERNAL_HPP_

#include <string>

struct sdg_mesh;

namespace sdg {
struct Mesh;
void setGlobalError(const std::string& msg);
const char* globalError();
const Mesh* meshOf(const sdg_mesh* m);
}  // namespace sdg

#endif  // SDG_INTERNAL_HPP_
