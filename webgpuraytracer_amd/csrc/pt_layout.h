// pt_layout.h — the device layout of a scene, built on the host from the
// tracer.js arrays: child-pair and 4-wide BVH nodes, the root table, and the
// 48 B triangle geometry records.  Shared by the C ABI (which uploads it) and
// the CPU tests of the traversal code.
#pragma once
#include <map>
#include <string>
#include <vector>

#include "pt_abi.h"
#include "dev/geometry.h"

namespace pt {

struct HostLayout {
    std::vector<CNode> tn, bn;        // child-pair nodes: TLAS, every BLAS
    std::vector<WNode> wtn, wbn;      // 4-wide nodes: TLAS, every BLAS
    std::vector<Root> roots;          // slot 0: TLAS; then one per distinct BLAS
    std::vector<uint32_t> obj_root;   // per object: root slot of its BLAS
    std::vector<uint32_t> mesh_root;  // per list mesh: root slot
    std::map<uint32_t, uint32_t> slots;  // BLAS node_offset -> root slot
    uint32_t tlas_depth = 0, blas_depth = 0;  // child-pair tree depths
    uint32_t tlas_wneed = 0, blas_wneed = 0;  // wide traversal stack needs
};

// Builds every tree of the scene.  Returns PT_OK, or PT_ERR_LAYOUT with a
// message in err (indices out of range, cycles, oversize trees).
// skip_tlas: the top-level tree is built elsewhere (the device build);
// tn and wtn stay empty and roots[0] is a placeholder.
int build_layout(const pt_scene_desc* d, HostLayout& L, std::string& err, bool skip_tlas = false);

// 3 float4 per triangle: v0, v1 - v0, v2 - v0.
void build_tri_geo(const pt_scene_desc* d, std::vector<float4>& geo);

// Stack entries per lane for a TLAS and a BLAS need (tree depths for the
// child-pair traversal, wide needs for the wide one) plus slack.
inline uint32_t stack_need(uint32_t tlas, uint32_t blas) { return tlas + blas + 2; }

}  // namespace pt
