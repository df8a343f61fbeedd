// pt_build.hip — bottom-level BVHs built on the device (PT_OPT_BLAS_BUILD =
// PT_BLAS_BUILD_DEVICE): a linear BVH over the triangles' Morton codes,
// collapsed into the 4-wide nodes the wide traversal reads.
//
// Per BLAS of n triangles (global indices [t0, t0 + n)):
//   1. centroid bounds (block partials + one reducing block);
//   2. 30-bit Morton codes of the centroids, sorted with their indices
//      (rocPRIM radix sort);
//   3. the binary radix tree over the sorted codes, one internal node per
//      thread (Karras, "Maximizing parallelism in the construction of BVHs,
//      octrees, and k-d trees", HPG 2012), ties in codes broken by index;
//   4. node boxes bottom-up: each leaf thread climbs while it is the second
//      child to arrive (an acquire-release counter per node);
//   5. the 4-wide collapse, one launch per level: a wide node takes its
//      binary node's two children and repeatedly opens the child with the
//      largest surface area until it has four; a subtree of at most
//      kLeafMax triangles becomes a leaf (its triangles are contiguous in
//      the sorted order); the accumulated (children - 1) along each path
//      gives the traversal stack the tree needs;
//   6. the triangle records permuted into the sorted order (a separate copy:
//      the reference-order traversal keeps the uploaded order), each keeping
//      its original index in the first record's w.
// Node boxes come from the exact vertices (pt_triangle), widened by a few
// ulps so that a hit the rounded edge vectors produce is never culled.
#include <rocprim/device/device_radix_sort.hpp>

#include "pt_internal.h"

namespace pt {
namespace {

constexpr uint32_t kBlk = 256;
constexpr uint32_t kLeafMax = 4;

struct Box {
    float lo[3], hi[3];
};

__device__ __forceinline__ float widen_lo(float v) { return v - (fabsf(v) * 1.2e-7f + 1e-30f); }
__device__ __forceinline__ float widen_hi(float v) { return v + (fabsf(v) * 1.2e-7f + 1e-30f); }

__device__ __forceinline__ Box tri_box(const pt_triangle& t) {
    Box b;
    for (int k = 0; k < 3; ++k) {
        b.lo[k] = widen_lo(fminf(fminf(t.v0[k], t.v1[k]), t.v2[k]));
        b.hi[k] = widen_hi(fmaxf(fmaxf(t.v0[k], t.v1[k]), t.v2[k]));
    }
    return b;
}

__device__ __forceinline__ float centroid(const pt_triangle& t, int k) { return (t.v0[k] + t.v1[k] + t.v2[k]) * (1.0f / 3.0f); }

// The primitives a tree is built over: the triangles [t0, t0 + n) of the
// scene (a BLAS), or precomputed boxes (the TLAS: one per object).
struct TriSrc {
    const pt_triangle* tris;
    uint32_t t0;
    __device__ Box box(uint32_t i) const { return tri_box(tris[t0 + i]); }
    __device__ float center(uint32_t i, int k) const { return centroid(tris[t0 + i], k); }
};
struct BoxSrc {
    const Box* boxes;
    __device__ Box box(uint32_t i) const { return boxes[i]; }
    __device__ float center(uint32_t i, int k) const { return 0.5f * (boxes[i].lo[k] + boxes[i].hi[k]); }
};

// 1. per-block centroid bounds -> part[block][6]
template <class Src>
__global__ void __launch_bounds__(kBlk) cbounds_k(Src src, uint32_t n, float* __restrict__ part) {
    __shared__ float red[6][kBlk];
    const uint32_t i = blockIdx.x * kBlk + threadIdx.x;
    float v[6] = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY};
    if (i < n)
        for (int k = 0; k < 3; ++k) v[k] = v[3 + k] = src.center(i, k);
    for (int k = 0; k < 6; ++k) red[k][threadIdx.x] = v[k];
    __syncthreads();
    for (uint32_t s = kBlk / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s)
            for (int k = 0; k < 6; ++k)
                red[k][threadIdx.x] = k < 3 ? fminf(red[k][threadIdx.x], red[k][threadIdx.x + s])
                                            : fmaxf(red[k][threadIdx.x], red[k][threadIdx.x + s]);
        __syncthreads();
    }
    if (threadIdx.x < 6) part[blockIdx.x * 6 + threadIdx.x] = red[threadIdx.x][0];
}

__global__ void __launch_bounds__(kBlk) cbounds_fin_k(const float* __restrict__ part, uint32_t blocks,
                                                      float* __restrict__ out) {
    __shared__ float red[6][kBlk];
    float v[6] = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY};
    for (uint32_t b = threadIdx.x; b < blocks; b += kBlk)
        for (int k = 0; k < 6; ++k) v[k] = k < 3 ? fminf(v[k], part[b * 6 + k]) : fmaxf(v[k], part[b * 6 + k]);
    for (int k = 0; k < 6; ++k) red[k][threadIdx.x] = v[k];
    __syncthreads();
    for (uint32_t s = kBlk / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s)
            for (int k = 0; k < 6; ++k)
                red[k][threadIdx.x] = k < 3 ? fminf(red[k][threadIdx.x], red[k][threadIdx.x + s])
                                            : fmaxf(red[k][threadIdx.x], red[k][threadIdx.x + s]);
        __syncthreads();
    }
    if (threadIdx.x < 6) out[threadIdx.x] = red[threadIdx.x][0];
}

__device__ __forceinline__ uint32_t spread10(uint32_t x) {
    x &= 0x3ffu;
    x = (x | (x << 16)) & 0x030000ffu;
    x = (x | (x << 8)) & 0x0300f00fu;
    x = (x | (x << 4)) & 0x030c30c3u;
    x = (x | (x << 2)) & 0x09249249u;
    return x;
}

// 2. Morton codes
template <class Src>
__global__ void __launch_bounds__(kBlk) morton_k(Src src, uint32_t n, const float* __restrict__ cb,
                                                 uint32_t* __restrict__ keys, uint32_t* __restrict__ idx) {
    const uint32_t i = blockIdx.x * kBlk + threadIdx.x;
    if (i >= n) return;
    uint32_t q[3];
    for (int k = 0; k < 3; ++k) {
        const float ext = cb[3 + k] - cb[k];
        const float u = ext > 0.0f ? (src.center(i, k) - cb[k]) / ext : 0.5f;
        q[k] = uint32_t(fminf(fmaxf(u * 1024.0f, 0.0f), 1023.0f));
    }
    keys[i] = (spread10(q[0]) << 2) | (spread10(q[1]) << 1) | spread10(q[2]);
    idx[i] = i;
}

// common prefix length of sorted codes i and j (-1 outside the range)
__device__ __forceinline__ int delta(const uint32_t* keys, uint32_t n, int i, int j) {
    if (j < 0 || j >= int(n)) return -1;
    const uint32_t a = keys[i], b = keys[j];
    if (a == b) return 32 + __clz(uint32_t(i) ^ uint32_t(j));
    return __clz(a ^ b);
}

// 3. binary radix tree: node ids 0..n-2 internal, n-1+k leaf k
__global__ void __launch_bounds__(kBlk) karras_k(const uint32_t* __restrict__ keys, uint32_t n,
                                                 uint32_t* __restrict__ child, uint32_t* __restrict__ parent,
                                                 uint32_t* __restrict__ range) {
    const int i = int(blockIdx.x * kBlk + threadIdx.x);
    if (i >= int(n) - 1) return;
    const int d = delta(keys, n, i, i + 1) - delta(keys, n, i, i - 1) > 0 ? 1 : -1;
    const int dmin = delta(keys, n, i, i - d);
    int lmax = 2;
    while (delta(keys, n, i, i + lmax * d) > dmin) lmax *= 2;
    int l = 0;
    for (int t = lmax / 2; t >= 1; t /= 2)
        if (delta(keys, n, i, i + (l + t) * d) > dmin) l += t;
    const int j = i + l * d;
    const int dnode = delta(keys, n, i, j);
    int s = 0;
    for (int t = (l + 1) / 2;; t = (t + 1) / 2) {
        if (delta(keys, n, i, i + (s + t) * d) > dnode) s += t;
        if (t <= 1) break;
    }
    const int gamma = i + s * d + min(d, 0);
    const int first = min(i, j), last = max(i, j);
    const uint32_t lc = first == gamma ? uint32_t(n - 1 + gamma) : uint32_t(gamma);
    const uint32_t rc = last == gamma + 1 ? uint32_t(n - 1 + gamma + 1) : uint32_t(gamma + 1);
    child[2 * i] = lc;
    child[2 * i + 1] = rc;
    parent[lc] = uint32_t(i);
    parent[rc] = uint32_t(i);
    range[2 * i] = uint32_t(first);
    range[2 * i + 1] = uint32_t(last);
}

__device__ __forceinline__ void store_box(float* nb, uint32_t node, const Box& b) {
    for (int k = 0; k < 3; ++k) {
        __hip_atomic_store(nb + node * 6 + k, b.lo[k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __hip_atomic_store(nb + node * 6 + 3 + k, b.hi[k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}
__device__ __forceinline__ Box load_box(float* nb, uint32_t node) {
    Box b;
    for (int k = 0; k < 3; ++k) {
        b.lo[k] = __hip_atomic_load(nb + node * 6 + k, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        b.hi[k] = __hip_atomic_load(nb + node * 6 + 3 + k, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    return b;
}

// 4. boxes bottom-up; flags zeroed before the launch
template <class Src>
__global__ void __launch_bounds__(kBlk) boxes_k(Src src, uint32_t n, const uint32_t* __restrict__ idx,
                                                const uint32_t* __restrict__ child, const uint32_t* __restrict__ parent,
                                                uint32_t* __restrict__ flags, float* __restrict__ nb) {
    const uint32_t k = blockIdx.x * kBlk + threadIdx.x;
    if (k >= n) return;
    uint32_t node = n - 1 + k;
    store_box(nb, node, src.box(idx[k]));
    if (n == 1) return;
    for (;;) {
        const uint32_t p = parent[node];
        // the second child to arrive merges; release our box, acquire the sibling's
        const uint32_t seen = __hip_atomic_fetch_add(flags + p, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
        if (seen == 0u) return;
        const Box a = load_box(nb, child[2 * p]), b = load_box(nb, child[2 * p + 1]);
        Box m;
        for (int c = 0; c < 3; ++c) {
            m.lo[c] = fminf(a.lo[c], b.lo[c]);
            m.hi[c] = fmaxf(a.hi[c], b.hi[c]);
        }
        store_box(nb, p, m);
        if (p == 0u) return;
        node = p;
    }
}

struct Item {
    uint32_t node;  // binary node id
    uint32_t slot;  // its wide node
    uint32_t acc;   // sum of (children - 1) over the wide ancestors
};

__device__ __forceinline__ uint32_t subtree_size(const uint32_t* range, uint32_t n, uint32_t node) {
    return node >= n - 1 ? 1u : range[2 * node + 1] - range[2 * node] + 1u;
}
__device__ __forceinline__ uint32_t subtree_first(const uint32_t* range, uint32_t n, uint32_t node) {
    return node >= n - 1 ? node - (n - 1) : range[2 * node];
}
__device__ __forceinline__ float area(const Box& b) {
    const float x = b.hi[0] - b.lo[0], y = b.hi[1] - b.lo[1], z = b.hi[2] - b.lo[2];
    return x * y + y * z + z * x;
}

// 5. one level of the 4-wide collapse.  ctr[0]: next free wide slot, ctr[1]:
// items of the next level, ctr[2]: the largest (children - 1) sum of a path.
// A subtree of at most leaf_max primitives becomes a leaf: a range of the
// sorted order at leaf_base (the primitives are permuted into that order),
// or, with leaf_idx (leaf_max 1), the one primitive's own index.
__global__ void __launch_bounds__(kBlk) collapse_k(const uint32_t* __restrict__ child, const uint32_t* __restrict__ range,
                                                   float* __restrict__ nb, uint32_t n, uint32_t leaf_max,
                                                   uint32_t leaf_base, const uint32_t* __restrict__ leaf_idx,
                                                   const Item* __restrict__ in, uint32_t n_in, Item* __restrict__ out,
                                                   uint32_t* __restrict__ ctr, WNode* __restrict__ wn) {
    const uint32_t i = blockIdx.x * kBlk + threadIdx.x;
    if (i >= n_in) return;
    const Item it = in[i];
    uint32_t c[4] = {child[2 * it.node], child[2 * it.node + 1], 0u, 0u};
    uint32_t cnt = 2;
    while (cnt < 4u) {
        int best = -1;
        float ba = -1.0f;
        for (uint32_t k = 0; k < cnt; ++k) {
            if (c[k] >= n - 1 || subtree_size(range, n, c[k]) <= leaf_max) continue;
            const float a = area(load_box(nb, c[k]));
            if (a > ba) { ba = a; best = int(k); }
        }
        if (best < 0) break;
        const uint32_t open = c[best];
        c[best] = child[2 * open];
        c[cnt++] = child[2 * open + 1];
    }
    WNode w;
    float lo[3][4], hi[3][4];
    uint32_t ref[4] = {0u, 0u, 0u, 0u};
    const uint32_t acc = it.acc + cnt - 1u;
    for (uint32_t k = 0; k < 4; ++k) {
        Box b;
        if (k < cnt) {
            b = load_box(nb, c[k]);
            const uint32_t sz = subtree_size(range, n, c[k]);
            if (c[k] >= n - 1 || sz <= leaf_max) {
                const uint32_t first = subtree_first(range, n, c[k]);
                ref[k] = kLeafBit | ((sz - 1u) << 25) | (leaf_idx ? leaf_idx[first] : leaf_base + first);
                atomicMax(ctr + 2, acc);
            } else {
                const uint32_t slot = atomicAdd(ctr + 0, 1u);
                const uint32_t q = atomicAdd(ctr + 1, 1u);
                out[q] = Item{c[k], slot, acc};
                ref[k] = slot;
            }
        } else {
            for (int a = 0; a < 3; ++a) { b.lo[a] = INFINITY; b.hi[a] = -INFINITY; }
        }
        for (int a = 0; a < 3; ++a) { lo[a][k] = b.lo[a]; hi[a][k] = b.hi[a]; }
    }
    w.lox = make_float4(lo[0][0], lo[0][1], lo[0][2], lo[0][3]);
    w.loy = make_float4(lo[1][0], lo[1][1], lo[1][2], lo[1][3]);
    w.loz = make_float4(lo[2][0], lo[2][1], lo[2][2], lo[2][3]);
    w.hix = make_float4(hi[0][0], hi[0][1], hi[0][2], hi[0][3]);
    w.hiy = make_float4(hi[1][0], hi[1][1], hi[1][2], hi[1][3]);
    w.hiz = make_float4(hi[2][0], hi[2][1], hi[2][2], hi[2][3]);
    w.ref = make_uint4(ref[0], ref[1], ref[2], ref[3]);
    w.meta = make_uint4(cnt, 0u, 0u, 0u);
    wn[it.slot] = w;
}

// 6. triangle records in the sorted order, original index in the first w
__global__ void __launch_bounds__(kBlk) permute_k(const float4* __restrict__ geo, uint32_t t0, uint32_t n,
                                                  const uint32_t* __restrict__ idx, float4* __restrict__ out) {
    const uint32_t j = blockIdx.x * kBlk + threadIdx.x;
    if (j >= n) return;
    const uint32_t src = t0 + idx[j], dst = t0 + j;
    float4 a = geo[3 * src];
    a.w = __uint_as_float(src);
    out[3 * dst] = a;
    out[3 * dst + 1] = geo[3 * src + 1];
    out[3 * dst + 2] = geo[3 * src + 2];
}

// ---- the top-level tree (PT_TLAS_BUILD_DEVICE) ----
// World box of object i, conservative: the local box of its primitive
// (sphere and cube [-1, 1]^3; cone-cylinder y in [0, 1], radius max(1,
// |top|); torus major radius 1 plus |minor| by |minor|; mesh its BLAS root
// box) through the inverse of the affine inv_matrix (column-major: local =
// A world + b), inverted in double, then widened for the rounding of the
// fp32 ray transforms.  A singular or non-finite transform gets a box
// around everything.
__global__ void __launch_bounds__(kBlk) obj_box_k(const pt_object* __restrict__ objs, uint32_t n,
                                                  const uint32_t* __restrict__ obj_root, const Root* __restrict__ roots,
                                                  Box* __restrict__ out) {
    const uint32_t i = blockIdx.x * kBlk + threadIdx.x;
    if (i >= n) return;
    const pt_object& o = objs[i];
    double lo[3] = {-1.0, -1.0, -1.0}, hi[3] = {1.0, 1.0, 1.0};
    if (o.object_type == PT_OBJ_CYLINDER) {
        const double r = fmax(1.0, fabs(double(o.param0)));
        lo[0] = lo[2] = -r;
        hi[0] = hi[2] = r;
        lo[1] = 0.0;
        hi[1] = 1.0;
    } else if (o.object_type == PT_OBJ_TORUS) {
        const double r = fabs(double(o.param0)), R = 1.0 + r;
        lo[0] = lo[2] = -R;
        hi[0] = hi[2] = R;
        lo[1] = -r;
        hi[1] = r;
    } else if (o.object_type == PT_OBJ_MESH) {
        const Root& rt = roots[obj_root[i]];
        for (int k = 0; k < 3; ++k) {
            lo[k] = rt.mn[k];
            hi[k] = rt.mx[k];
        }
    }
    const float* m = o.inv_matrix;
    double A[3][3], b[3];
    for (int r = 0; r < 3; ++r) {
        for (int c = 0; c < 3; ++c) A[r][c] = m[c * 4 + r];
        b[r] = m[12 + r];
    }
    // M = A^-1 by cofactors; world = M (local - b)
    double M[3][3];
    M[0][0] = A[1][1] * A[2][2] - A[1][2] * A[2][1];
    M[0][1] = A[0][2] * A[2][1] - A[0][1] * A[2][2];
    M[0][2] = A[0][1] * A[1][2] - A[0][2] * A[1][1];
    M[1][0] = A[1][2] * A[2][0] - A[1][0] * A[2][2];
    M[1][1] = A[0][0] * A[2][2] - A[0][2] * A[2][0];
    M[1][2] = A[0][2] * A[1][0] - A[0][0] * A[1][2];
    M[2][0] = A[1][0] * A[2][1] - A[1][1] * A[2][0];
    M[2][1] = A[0][1] * A[2][0] - A[0][0] * A[2][1];
    M[2][2] = A[0][0] * A[1][1] - A[0][1] * A[1][0];
    const double det = A[0][0] * M[0][0] + A[0][1] * M[1][0] + A[0][2] * M[2][0];
    Box w;
    bool ok = isfinite(det) && fabs(det) > 1e-300;
    double cw[3], hw[3];
    if (ok) {
        for (int r = 0; r < 3; ++r) {
            double c = 0.0, h = 0.0;
            for (int k = 0; k < 3; ++k) {
                const double mk = M[r][k] / det;
                c += mk * (0.5 * (lo[k] + hi[k]) - b[k]);
                h += fabs(mk) * 0.5 * (hi[k] - lo[k]);
            }
            cw[r] = c;
            hw[r] = h;
            ok = ok && isfinite(c) && isfinite(h) && fabs(c) + h < 1e30;
        }
    }
    if (!ok) {
        for (int k = 0; k < 3; ++k) {
            w.lo[k] = -1e30f;
            w.hi[k] = 1e30f;
        }
    } else {
        double ext = 0.0;
        for (int k = 0; k < 3; ++k) ext = fmax(ext, fabs(cw[k]) + hw[k]);
        const double pad = 1e-5 * ext + 1e-6;
        for (int k = 0; k < 3; ++k) {
            w.lo[k] = float(cw[k] - hw[k] - pad);
            w.hi[k] = float(cw[k] + hw[k] + pad);
        }
    }
    out[i] = w;
}

// Child-pair nodes from the binary radix tree: node i's CNode is index i,
// a leaf child references its one object (count 1) by its own index.
__global__ void __launch_bounds__(kBlk) cnodes_k(const uint32_t* __restrict__ child, const uint32_t* __restrict__ sidx,
                                                 float* __restrict__ nb, uint32_t n, CNode* __restrict__ cn) {
    const uint32_t i = blockIdx.x * kBlk + threadIdx.x;
    if (i + 1u >= n) return;
    const uint32_t lc = child[2 * i], rc = child[2 * i + 1];
    const Box L = load_box(nb, lc), R = load_box(nb, rc);
    CNode c;
    c.a = make_float4(L.lo[0], L.lo[1], L.lo[2], L.hi[0]);
    c.b = make_float4(L.hi[1], L.hi[2], R.lo[0], R.lo[1]);
    c.c = make_float4(R.lo[2], R.hi[0], R.hi[1], R.hi[2]);
    const uint32_t lref = lc >= n - 1 ? (kLeafBit | sidx[lc - (n - 1)]) : lc;
    const uint32_t rref = rc >= n - 1 ? (kLeafBit | sidx[rc - (n - 1)]) : rc;
    c.d = make_uint4(lref, rref, 0u, 0u);
    cn[i] = c;
}

// Depth of the binary tree (edges from the root to the deepest leaf) into *depth.
__global__ void __launch_bounds__(kBlk) depth_k(const uint32_t* __restrict__ parent, uint32_t n,
                                                uint32_t* __restrict__ depth) {
    const uint32_t k = blockIdx.x * kBlk + threadIdx.x;
    if (k >= n || n < 2u) return;
    uint32_t node = n - 1 + k, d = 0;
    while (node != 0u && d < 4096u) {
        node = parent[node];
        ++d;
    }
    atomicMax(depth, d);
}

__global__ void zero_u32_k(uint32_t* p, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0u;
}

uint32_t blocks_of(uint32_t n) { return (n + kBlk - 1u) / kBlk; }

// The 4-wide collapse of a built binary tree of m > leaf_max primitives, one
// launch per level from the root: the root takes the next free slot of ctr[0].
// Returns the root slot and the largest (children - 1) sum of a path.
hipError_t collapse_all(const uint32_t* child, const uint32_t* range, float* nb, uint32_t m, uint32_t leaf_max,
                        uint32_t leaf_base, const uint32_t* leaf_idx, Item* qa, Item* qb, uint32_t* ctr, WNode* wn,
                        uint32_t* root_slot, uint32_t* need, hipStream_t s) {
    uint32_t h[3];
    hipError_t e = hipMemcpyAsync(h, ctr, 12, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) return e;
    *root_slot = h[0];
    const Item root{0u, h[0], 0u};
    const uint32_t init[3] = {h[0] + 1u, 0u, 0u};
    e = hipMemcpyAsync(qa, &root, sizeof(Item), hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hipMemcpyAsync(ctr, init, 12, hipMemcpyHostToDevice, s);
    uint32_t n_in = 1;
    Item* cur = qa;
    Item* nxt = qb;
    while (e == hipSuccess && n_in > 0) {
        hipLaunchKernelGGL(zero_u32_k, dim3(1), dim3(64), 0, s, ctr + 1, 1u);
        hipLaunchKernelGGL(collapse_k, dim3(blocks_of(n_in)), dim3(kBlk), 0, s, child, range, nb, m, leaf_max,
                           leaf_base, leaf_idx, cur, n_in, nxt, ctr, wn);
        e = hipMemcpyAsync(h, ctr, 12, hipMemcpyDeviceToHost, s);
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        n_in = h[1];
        std::swap(cur, nxt);
    }
    *need = h[2];
    return e == hipSuccess ? hipGetLastError() : e;
}

}  // namespace

size_t blas_build_scratch_bytes(uint32_t max_n) {
    size_t sort_bytes = 0;
    (void)rocprim::radix_sort_pairs(nullptr, sort_bytes, static_cast<uint32_t*>(nullptr), static_cast<uint32_t*>(nullptr),
                                    static_cast<uint32_t*>(nullptr), static_cast<uint32_t*>(nullptr), max_n, 0, 30);
    const size_t n = max_n;
    const size_t words = 4 * n            // keys, idx, sorted keys, sorted idx
                         + 2 * n + 2 * n  // child, range
                         + 2 * n          // parent
                         + n              // flags
                         + 6 * 2 * n      // node boxes (floats)
                         + 6 * (blocks_of(max_n) + 1) + 8;
    return words * 4 + 2 * n * sizeof(Item) + sort_bytes + 1024;
}

hipError_t build_blas_device(const float4* geo, const pt_triangle* tris, const BlasBuildJob* jobs, uint32_t n_jobs,
                             void* scratch, size_t scratch_bytes, WNode* wn, float4* geo_out, uint32_t* ctr,
                             BlasBuildResult* res, hipStream_t s) {
    uint32_t max_n = 1;
    for (uint32_t j = 0; j < n_jobs; ++j) max_n = std::max(max_n, jobs[j].n);
    if (scratch_bytes < blas_build_scratch_bytes(max_n)) return hipErrorInvalidValue;
    const size_t n = max_n;
    uint32_t* keys = static_cast<uint32_t*>(scratch);
    uint32_t* idx = keys + n;
    uint32_t* skeys = idx + n;
    uint32_t* sidx = skeys + n;
    uint32_t* child = sidx + n;
    uint32_t* range = child + 2 * n;
    uint32_t* parent = range + 2 * n;
    uint32_t* flags = parent + 2 * n;
    float* nb = reinterpret_cast<float*>(flags + n);
    float* part = nb + 12 * n;
    float* cb = part + 6 * blocks_of(max_n);
    Item* qa = reinterpret_cast<Item*>(reinterpret_cast<uintptr_t>(cb + 8 + 63) & ~uintptr_t(63));
    Item* qb = qa + n;
    void* sort_tmp = reinterpret_cast<void*>((reinterpret_cast<uintptr_t>(qb + n) + 255) & ~uintptr_t(255));
    size_t sort_bytes = scratch_bytes - (static_cast<char*>(sort_tmp) - static_cast<char*>(scratch));
    hipError_t e = hipSuccess;
    // ctr: [0] next wide slot (kept across jobs), [1] next level's items, [2] path need
    hipLaunchKernelGGL(zero_u32_k, dim3(1), dim3(64), 0, s, ctr, 4u);
    for (uint32_t j = 0; j < n_jobs && e == hipSuccess; ++j) {
        const BlasBuildJob& J = jobs[j];
        const uint32_t m = J.n;
        const TriSrc src{tris, J.t0};
        hipLaunchKernelGGL(cbounds_k<TriSrc>, dim3(blocks_of(m)), dim3(kBlk), 0, s, src, m, part);
        hipLaunchKernelGGL(cbounds_fin_k, dim3(1), dim3(kBlk), 0, s, part, blocks_of(m), cb);
        hipLaunchKernelGGL(morton_k<TriSrc>, dim3(blocks_of(m)), dim3(kBlk), 0, s, src, m, cb, keys, idx);
        e = rocprim::radix_sort_pairs(sort_tmp, sort_bytes, keys, skeys, idx, sidx, m, 0, 30, s);
        if (e != hipSuccess) break;
        if (m > 1)
            hipLaunchKernelGGL(karras_k, dim3(blocks_of(m - 1)), dim3(kBlk), 0, s, skeys, m, child, parent, range);
        hipLaunchKernelGGL(zero_u32_k, dim3(blocks_of(m)), dim3(kBlk), 0, s, flags, m);
        hipLaunchKernelGGL(boxes_k<TriSrc>, dim3(blocks_of(m)), dim3(kBlk), 0, s, src, m, sidx, child, parent, flags,
                           nb);
        hipLaunchKernelGGL(permute_k, dim3(blocks_of(m)), dim3(kBlk), 0, s, geo, J.t0, m, sidx, geo_out);
        res[j].leaf_root = m <= kLeafMax;
        if (m <= kLeafMax) {
            res[j].root_ref = kLeafBit | ((m - 1u) << 25) | J.leaf_base;
            res[j].need = 0;
            continue;
        }
        e = collapse_all(child, range, nb, m, kLeafMax, J.leaf_base, nullptr, qa, qb, ctr, wn, &res[j].root_ref,
                         &res[j].need, s);
    }
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    return e;
}

size_t tlas_build_scratch_bytes(uint32_t n) {
    return blas_build_scratch_bytes(std::max(n, 1u)) + size_t(std::max(n, 1u)) * sizeof(Box) + 256;
}

hipError_t build_tlas_device(const pt_object* objs, uint32_t n, const uint32_t* obj_root, const Root* roots,
                             void* scratch, size_t scratch_bytes, CNode* cn, WNode* wn, uint32_t* ctr,
                             TlasBuildResult* res, hipStream_t s) {
    if (n == 0 || scratch_bytes < tlas_build_scratch_bytes(n)) return hipErrorInvalidValue;
    const size_t bs = blas_build_scratch_bytes(n);
    uint32_t* keys = static_cast<uint32_t*>(scratch);
    uint32_t* idx = keys + n;
    uint32_t* skeys = idx + n;
    uint32_t* sidx = skeys + n;
    uint32_t* child = sidx + n;
    uint32_t* range = child + 2 * size_t(n);
    uint32_t* parent = range + 2 * size_t(n);
    uint32_t* flags = parent + 2 * size_t(n);
    float* nb = reinterpret_cast<float*>(flags + n);
    float* part = nb + 12 * size_t(n);
    float* cb = part + 6 * blocks_of(n);
    Item* qa = reinterpret_cast<Item*>(reinterpret_cast<uintptr_t>(cb + 8 + 63) & ~uintptr_t(63));
    Item* qb = qa + n;
    void* sort_tmp = reinterpret_cast<void*>((reinterpret_cast<uintptr_t>(qb + n) + 255) & ~uintptr_t(255));
    Box* boxes = reinterpret_cast<Box*>((reinterpret_cast<uintptr_t>(scratch) + bs + 63) & ~uintptr_t(63));
    size_t sort_bytes = reinterpret_cast<char*>(boxes) - static_cast<char*>(sort_tmp);
    hipLaunchKernelGGL(zero_u32_k, dim3(1), dim3(64), 0, s, ctr, 4u);
    hipLaunchKernelGGL(obj_box_k, dim3(blocks_of(n)), dim3(kBlk), 0, s, objs, n, obj_root, roots, boxes);
    const BoxSrc src{boxes};
    hipLaunchKernelGGL(cbounds_k<BoxSrc>, dim3(blocks_of(n)), dim3(kBlk), 0, s, src, n, part);
    hipLaunchKernelGGL(cbounds_fin_k, dim3(1), dim3(kBlk), 0, s, part, blocks_of(n), cb);
    hipLaunchKernelGGL(morton_k<BoxSrc>, dim3(blocks_of(n)), dim3(kBlk), 0, s, src, n, cb, keys, idx);
    hipError_t e = rocprim::radix_sort_pairs(sort_tmp, sort_bytes, keys, skeys, idx, sidx, n, 0, 30, s);
    if (e != hipSuccess) return e;
    if (n > 1)
        hipLaunchKernelGGL(karras_k, dim3(blocks_of(n - 1)), dim3(kBlk), 0, s, skeys, n, child, parent, range);
    hipLaunchKernelGGL(zero_u32_k, dim3(blocks_of(n)), dim3(kBlk), 0, s, flags, n);
    hipLaunchKernelGGL(boxes_k<BoxSrc>, dim3(blocks_of(n)), dim3(kBlk), 0, s, src, n, sidx, child, parent, flags, nb);
    Box root_box;
    if (n == 1) {
        uint32_t first;
        e = hipMemcpyAsync(&root_box, boxes, sizeof(Box), hipMemcpyDeviceToHost, s);
        if (e == hipSuccess) e = hipMemcpyAsync(&first, sidx, 4, hipMemcpyDeviceToHost, s);
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        if (e != hipSuccess) return e;
        res->root.ref = res->root.wref = kLeafBit | first;
        res->depth = res->need = 0;
    } else {
        hipLaunchKernelGGL(cnodes_k, dim3(blocks_of(n - 1)), dim3(kBlk), 0, s, child, sidx, nb, n, cn);
        hipLaunchKernelGGL(depth_k, dim3(blocks_of(n)), dim3(kBlk), 0, s, parent, n, ctr + 3);
        e = collapse_all(child, range, nb, n, 1u, 0u, sidx, qa, qb, ctr, wn, &res->root.wref, &res->need, s);
        uint32_t depth = 0;
        if (e == hipSuccess) e = hipMemcpyAsync(&root_box, nb, sizeof(Box), hipMemcpyDeviceToHost, s);
        if (e == hipSuccess) e = hipMemcpyAsync(&depth, ctr + 3, 4, hipMemcpyDeviceToHost, s);
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        if (e != hipSuccess) return e;
        res->root.ref = 0u;
        res->depth = depth;
    }
    for (int k = 0; k < 3; ++k) {
        res->root.mn[k] = root_box.lo[k];
        res->root.mx[k] = root_box.hi[k];
    }
    return hipGetLastError();
}

// ------------------------------------------------------- quantised nodes ----
__global__ void __launch_bounds__(kBlk) quantize_k(const WNode* __restrict__ in, QNode* __restrict__ out, uint32_t n) {
    const uint32_t i = blockIdx.x * kBlk + threadIdx.x;
    if (i < n) out[i] = quantize_node(in[i]);
}

hipError_t quantize_wide_nodes(const WNode* in, QNode* out, uint32_t n, hipStream_t s) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(quantize_k, dim3((n + kBlk - 1) / kBlk), dim3(kBlk), 0, s, in, out, n);
    return hipGetLastError();
}

}  // namespace pt
