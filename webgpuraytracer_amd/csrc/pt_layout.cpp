// pt_layout.cpp — host-side construction of the device scene layout.
#include "pt_layout.h"

#include <algorithm>
#include <cstdarg>
#include <cstdio>
#include <cstring>

namespace pt {
namespace {

uint32_t f2u_host(float f) {
    uint32_t u;
    std::memcpy(&u, &f, 4);
    return u;
}

// Converts one reference BVH (32 B nodes, left = i+1) rooted at `root` into
// child-pair nodes appended to `out`.  For a BLAS the right child is
// node_offset + next (shader.wgsl:530); for the TLAS it is next
// (shader.wgsl:740).  Returns PT_OK and fills *r.
struct TreeConv {
    const pt_bvh_node* nodes;
    uint64_t n_nodes;
    bool blas;
    uint32_t node_offset;
    std::vector<CNode>* out;
    uint32_t max_depth = 0;
    std::string err;

    bool leaf_ref(uint32_t i, uint32_t* ref) {
        const pt_bvh_node& n = nodes[i];
        if (n.count <= 32 && n.next < (1u << 25)) {
            *ref = kLeafBit | ((n.count - 1u) << 25) | n.next;
        } else {
            if (i >= (1u << 30)) { err = "leaf index too large"; return false; }
            *ref = kLeafBit | kIndirectBit | i;
        }
        return true;
    }
    // Returns the reference for node i (recursively converting its subtree).
    bool convert(uint32_t i, uint32_t depth, uint32_t* ref) {
        if (i >= n_nodes) { err = "BVH child index out of range"; return false; }
        if (depth > 256) { err = "BVH deeper than 256 levels (cycle?)"; return false; }
        max_depth = std::max(max_depth, depth);
        const pt_bvh_node& n = nodes[i];
        if (n.count > 0) return leaf_ref(i, ref);
        const uint64_t left = uint64_t(i) + 1u;
        const uint64_t right = blas ? uint64_t(node_offset) + n.next : uint64_t(n.next);
        if (left >= n_nodes || right >= n_nodes) { err = "BVH child index out of range"; return false; }
        const uint32_t k = uint32_t(out->size());
        if (k >= 0x80000000u) { err = "too many BVH nodes"; return false; }
        out->push_back(CNode{});
        uint32_t lref, rref;
        if (!convert(uint32_t(left), depth + 1, &lref)) return false;
        if (!convert(uint32_t(right), depth + 1, &rref)) return false;
        const pt_bvh_node& L = nodes[left];
        const pt_bvh_node& R = nodes[right];
        CNode& c = (*out)[k];
        c.a = make_float4(L.aabb_min[0], L.aabb_min[1], L.aabb_min[2], L.aabb_max[0]);
        c.b = make_float4(L.aabb_max[1], L.aabb_max[2], R.aabb_min[0], R.aabb_min[1]);
        c.c = make_float4(R.aabb_min[2], R.aabb_max[0], R.aabb_max[1], R.aabb_max[2]);
        c.d = make_uint4(lref, rref, 0u, 0u);
        *ref = k;
        return true;
    }
    bool root(uint32_t i, Root* r) {
        if (i >= n_nodes) { err = "BVH root index out of range"; return false; }
        uint32_t ref;
        if (!convert(i, 0, &ref)) return false;
        const pt_bvh_node& n = nodes[i];
        for (int k = 0; k < 3; ++k) { r->mn[k] = n.aabb_min[k]; r->mx[k] = n.aabb_max[k]; }
        r->ref = ref;
        r->wref = 0;
        return true;
    }
};

// Collapses a child-pair tree into 4-wide nodes: starting from a node's two
// children, the inner child with the largest box surface is replaced by its
// own two children until four are gathered or only leaves remain.  Leaf refs
// are kept as they are.  `need` is the stack depth the wide traversal can
// reach below a node: it pushes at most (children - 1) refs per node on the
// path from the root.
struct WideConv {
    const std::vector<CNode>& bin;
    std::vector<WNode>* out;
    struct Child {
        float mn[3], mx[3];
        uint32_t ref;
    };
    static float area(const Child& c) {
        const float x = c.mx[0] - c.mn[0], y = c.mx[1] - c.mn[1], z = c.mx[2] - c.mn[2];
        return x * y + y * z + z * x;
    }
    void split(uint32_t k, Child& l, Child& r) const {
        const CNode& n = bin[k];
        l = Child{{n.a.x, n.a.y, n.a.z}, {n.a.w, n.b.x, n.b.y}, n.d.x};
        r = Child{{n.b.z, n.b.w, n.c.x}, {n.c.y, n.c.z, n.c.w}, n.d.y};
    }
    bool convert(uint32_t bref, uint32_t* wref, uint32_t* need) {
        if (bref & kLeafBit) {
            *wref = bref;
            *need = 0;
            return true;
        }
        Child ch[4];
        int n = 2;
        split(bref, ch[0], ch[1]);
        while (n < 4) {
            int best = -1;
            float best_area = -1.0f;
            for (int i = 0; i < n; ++i)
                if (!(ch[i].ref & kLeafBit) && area(ch[i]) > best_area) {
                    best = i;
                    best_area = area(ch[i]);
                }
            if (best < 0) break;
            Child a, b;
            split(ch[best].ref, a, b);
            for (int j = n; j > best + 1; --j) ch[j] = ch[j - 1];
            ch[best] = a;
            ch[best + 1] = b;
            ++n;
        }
        const size_t k = out->size();
        if (k >= 0x80000000u) return false;
        out->push_back(WNode{});
        uint32_t refs[4] = {0, 0, 0, 0}, sub = 0;
        for (int i = 0; i < n; ++i) {
            uint32_t nd;
            if (!convert(ch[i].ref, &refs[i], &nd)) return false;
            sub = std::max(sub, nd);
        }
        float v[6][4];
        for (int i = 0; i < 4; ++i)
            for (int a = 0; a < 3; ++a) {
                // unused slots: an empty box at +inf (never accepted; also masked by the count)
                v[a][i] = i < n ? ch[i].mn[a] : __builtin_inff();
                v[3 + a][i] = i < n ? ch[i].mx[a] : __builtin_inff();
            }
        WNode& w = (*out)[k];
        w.lox = make_float4(v[0][0], v[0][1], v[0][2], v[0][3]);
        w.loy = make_float4(v[1][0], v[1][1], v[1][2], v[1][3]);
        w.loz = make_float4(v[2][0], v[2][1], v[2][2], v[2][3]);
        w.hix = make_float4(v[3][0], v[3][1], v[3][2], v[3][3]);
        w.hiy = make_float4(v[4][0], v[4][1], v[4][2], v[4][3]);
        w.hiz = make_float4(v[5][0], v[5][1], v[5][2], v[5][3]);
        w.ref = make_uint4(refs[0], refs[1], refs[2], refs[3]);
        w.meta = make_uint4(uint32_t(n), 0u, 0u, 0u);
        *wref = uint32_t(k);
        *need = uint32_t(n - 1) + sub;
        return true;
    }
};

bool blas_ranges_ok(const pt_bvh_node* nodes, uint64_t n_nodes, uint32_t node_offset, uint32_t tri_offset,
                    uint64_t n_tris) {
    std::vector<uint32_t> st{node_offset};
    uint64_t visited = 0;
    while (!st.empty()) {
        const uint32_t i = st.back();
        st.pop_back();
        if (i >= n_nodes || ++visited > n_nodes) return false;
        const pt_bvh_node& n = nodes[i];
        if (n.count > 0) {
            if (uint64_t(tri_offset) + n.next + n.count > n_tris) return false;
        } else {
            st.push_back(i + 1);
            st.push_back(node_offset + n.next);
        }
    }
    return true;
}

}  // namespace

int build_layout(const pt_scene_desc* d, HostLayout& L, std::string& err, bool skip_tlas) {
    auto fail = [&](const char* fmt, ...) {
        char buf[512];
        va_list ap;
        va_start(ap, fmt);
        vsnprintf(buf, sizeof(buf), fmt, ap);
        va_end(ap);
        err = buf;
        return int(PT_ERR_LAYOUT);
    };
    std::vector<CNode>& tn = L.tn;
    std::vector<CNode>& bn = L.bn;
    std::vector<Root>& roots = L.roots;
    std::vector<uint32_t>& obj_root = L.obj_root;
    std::vector<uint32_t>& mesh_root = L.mesh_root;
    std::map<uint32_t, uint32_t>& slots = L.slots;
    uint32_t* tlas_depth = &L.tlas_depth;
    uint32_t* blas_depth = &L.blas_depth;
    const uint32_t any_obj = PT_HAS_SPHERES | PT_HAS_CUBES | PT_HAS_CYLINDERS | PT_HAS_TORI | PT_HAS_MESHES;
    const pt_bvh_node* tl = static_cast<const pt_bvh_node*>(d->tlas.ptr);
    const pt_bvh_node* bl = static_cast<const pt_bvh_node*>(d->bvh.ptr);
    const pt_object* ob = static_cast<const pt_object*>(d->objects.ptr);
    roots.assign(1, Root{});
    *tlas_depth = 0;
    *blas_depth = 0;
    if ((d->flags & any_obj) && !skip_tlas) {
        TreeConv t{tl, d->tlas.count, false, 0, &tn};
        if (!t.root(0, &roots[0])) return fail("tlas: %s", t.err.c_str());
        *tlas_depth = t.max_depth;
        WideConv w{tn, &L.wtn};
        if (!w.convert(roots[0].ref, &roots[0].wref, &L.tlas_wneed))
            return fail("tlas: too many wide nodes");
        // every TLAS leaf range must index real objects
        for (uint64_t i = 0; i < d->tlas.count; ++i)
            if (tl[i].count > 0 && uint64_t(tl[i].next) + tl[i].count > d->objects.count)
                return fail("tlas node %llu: object range out of bounds", (unsigned long long)i);
    }
    obj_root.assign(d->objects.count, 0u);
    auto slot_for = [&](uint32_t node_offset, uint32_t tri_offset, uint32_t* slot) -> int {
        if (!blas_ranges_ok(bl, d->bvh.count, node_offset, tri_offset, d->triangles.count))
            return fail("BLAS at node %u / triangle offset %u indexes outside the arrays", node_offset,
                           tri_offset);
        auto it = slots.find(node_offset);
        if (it != slots.end()) { *slot = it->second; return PT_OK; }
        TreeConv b{bl, d->bvh.count, true, node_offset, &bn};
        Root r;
        if (!b.root(node_offset, &r)) return fail("bvh: %s", b.err.c_str());
        *blas_depth = std::max(*blas_depth, b.max_depth);
        WideConv w{bn, &L.wbn};
        uint32_t wneed = 0;
        if (!w.convert(r.ref, &r.wref, &wneed)) return fail("bvh: too many wide nodes");
        L.blas_wneed = std::max(L.blas_wneed, wneed);
        *slot = uint32_t(roots.size());
        roots.push_back(r);
        slots[node_offset] = *slot;
        return PT_OK;
    };
    if (d->flags & PT_HAS_MESHES) {
        for (uint64_t i = 0; i < d->objects.count; ++i) {
            if (ob[i].object_type != PT_OBJ_MESH) continue;
            uint32_t s;
            int rc = slot_for(f2u_host(ob[i].param0), f2u_host(ob[i].param1), &s);
            if (rc) return rc;
            obj_root[i] = s;
        }
    }
    mesh_root.assign(d->meshes.count, 0u);
    if (d->flags & PT_HAS_LIST_MESHES) {
        const pt_mesh_instance* mi = static_cast<const pt_mesh_instance*>(d->meshes.ptr);
        for (uint64_t i = 0; i < d->meshes.count; ++i) {
            uint32_t s;
            int rc = slot_for(mi[i].node_offset, mi[i].tri_offset, &s);
            if (rc) return rc;
            mesh_root[i] = s;
        }
    }
    return PT_OK;
}


void build_tri_geo(const pt_scene_desc* d, std::vector<float4>& geo) {
    const pt_triangle* tr = static_cast<const pt_triangle*>(d->triangles.ptr);
    geo.resize(size_t(d->triangles.count) * 3);
    for (uint64_t i = 0; i < d->triangles.count; ++i) {
        const pt_triangle& t = tr[i];
        // w of the first record: the triangle's index (bits), the prim id of a hit
        float idx;
        const uint32_t u = uint32_t(i);
        std::memcpy(&idx, &u, 4);
        geo[3 * i] = make_float4(t.v0[0], t.v0[1], t.v0[2], idx);
        geo[3 * i + 1] = make_float4(t.v1[0] - t.v0[0], t.v1[1] - t.v0[1], t.v1[2] - t.v0[2], 0.f);
        geo[3 * i + 2] = make_float4(t.v2[0] - t.v0[0], t.v2[1] - t.v0[1], t.v2[2] - t.v0[2], 0.f);
    }
}

}  // namespace pt
