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

// 1. per-block centroid bounds -> part[block][6]
__global__ void __launch_bounds__(kBlk) cbounds_k(const pt_triangle* __restrict__ tris, uint32_t t0, uint32_t n,
                                                  float* __restrict__ part) {
    __shared__ float red[6][kBlk];
    const uint32_t i = blockIdx.x * kBlk + threadIdx.x;
    float v[6] = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY};
    if (i < n) {
        const pt_triangle& t = tris[t0 + i];
        for (int k = 0; k < 3; ++k) v[k] = v[3 + k] = centroid(t, k);
    }
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
__global__ void __launch_bounds__(kBlk) morton_k(const pt_triangle* __restrict__ tris, uint32_t t0, uint32_t n,
                                                 const float* __restrict__ cb, uint32_t* __restrict__ keys,
                                                 uint32_t* __restrict__ idx) {
    const uint32_t i = blockIdx.x * kBlk + threadIdx.x;
    if (i >= n) return;
    const pt_triangle& t = tris[t0 + i];
    uint32_t q[3];
    for (int k = 0; k < 3; ++k) {
        const float ext = cb[3 + k] - cb[k];
        const float u = ext > 0.0f ? (centroid(t, k) - cb[k]) / ext : 0.5f;
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
__global__ void __launch_bounds__(kBlk) boxes_k(const pt_triangle* __restrict__ tris, uint32_t t0, uint32_t n,
                                                const uint32_t* __restrict__ idx, const uint32_t* __restrict__ child,
                                                const uint32_t* __restrict__ parent, uint32_t* __restrict__ flags,
                                                float* __restrict__ nb) {
    const uint32_t k = blockIdx.x * kBlk + threadIdx.x;
    if (k >= n) return;
    uint32_t node = n - 1 + k;
    store_box(nb, node, tri_box(tris[t0 + idx[k]]));
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
__global__ void __launch_bounds__(kBlk) collapse_k(const uint32_t* __restrict__ child, const uint32_t* __restrict__ range,
                                                   float* __restrict__ nb, uint32_t n, uint32_t leaf_base,
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
            if (c[k] >= n - 1 || subtree_size(range, n, c[k]) <= kLeafMax) continue;
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
            if (c[k] >= n - 1 || sz <= kLeafMax) {
                ref[k] = kLeafBit | ((sz - 1u) << 25) | (leaf_base + subtree_first(range, n, c[k]));
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

__global__ void zero_u32_k(uint32_t* p, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0u;
}

uint32_t blocks_of(uint32_t n) { return (n + kBlk - 1u) / kBlk; }

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
        hipLaunchKernelGGL(cbounds_k, dim3(blocks_of(m)), dim3(kBlk), 0, s, tris, J.t0, m, part);
        hipLaunchKernelGGL(cbounds_fin_k, dim3(1), dim3(kBlk), 0, s, part, blocks_of(m), cb);
        hipLaunchKernelGGL(morton_k, dim3(blocks_of(m)), dim3(kBlk), 0, s, tris, J.t0, m, cb, keys, idx);
        e = rocprim::radix_sort_pairs(sort_tmp, sort_bytes, keys, skeys, idx, sidx, m, 0, 30, s);
        if (e != hipSuccess) break;
        if (m > 1)
            hipLaunchKernelGGL(karras_k, dim3(blocks_of(m - 1)), dim3(kBlk), 0, s, skeys, m, child, parent, range);
        hipLaunchKernelGGL(zero_u32_k, dim3(blocks_of(m)), dim3(kBlk), 0, s, flags, m);
        hipLaunchKernelGGL(boxes_k, dim3(blocks_of(m)), dim3(kBlk), 0, s, tris, J.t0, m, sidx, child, parent, flags, nb);
        hipLaunchKernelGGL(permute_k, dim3(blocks_of(m)), dim3(kBlk), 0, s, geo, J.t0, m, sidx, geo_out);
        res[j].leaf_root = m <= kLeafMax;
        if (m <= kLeafMax) {
            res[j].root_ref = kLeafBit | ((m - 1u) << 25) | J.leaf_base;
            res[j].need = 0;
            continue;
        }
        // the root item takes the next wide slot
        uint32_t h[3];
        e = hipMemcpyAsync(h, ctr, 12, hipMemcpyDeviceToHost, s);
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        if (e != hipSuccess) break;
        const uint32_t root_slot = h[0];
        const Item root{0u, root_slot, 0u};
        const uint32_t init[3] = {root_slot + 1u, 0u, 0u};
        e = hipMemcpyAsync(qa, &root, sizeof(Item), hipMemcpyHostToDevice, s);
        if (e == hipSuccess) e = hipMemcpyAsync(ctr, init, 12, hipMemcpyHostToDevice, s);
        uint32_t n_in = 1;
        Item* cur = qa;
        Item* nxt = qb;
        while (e == hipSuccess && n_in > 0) {
            hipLaunchKernelGGL(zero_u32_k, dim3(1), dim3(64), 0, s, ctr + 1, 1u);
            hipLaunchKernelGGL(collapse_k, dim3(blocks_of(n_in)), dim3(kBlk), 0, s, child, range, nb, m, J.leaf_base,
                               cur, n_in, nxt, ctr, wn);
            e = hipMemcpyAsync(h, ctr, 12, hipMemcpyDeviceToHost, s);
            if (e == hipSuccess) e = hipStreamSynchronize(s);
            n_in = h[1];
            std::swap(cur, nxt);
        }
        res[j].root_ref = root_slot;
        res[j].need = h[2];
        if (e == hipSuccess) e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    return e;
}

}  // namespace pt
