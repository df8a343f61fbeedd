// scene_gen.cpp — host scene builder: tracer.js's scene preprocessing restated
// in C++ (see pt_scene.h for the list of reference functions), plus the seeded
// synthetic scenes used for measurement.
//
// JavaScript arithmetic is IEEE double and glMatrix stores every vector and
// matrix in a Float32Array, so each glMatrix-style operation below computes in
// double from float inputs and rounds once to float on store.
#include "pt_scene.h"

#include <algorithm>
#include <array>
#include <cmath>
#include <cstring>
#include <memory>
#include <new>
#include <string>
#include <unordered_map>
#include <vector>

namespace {

constexpr double kPi = 3.14159265358979323846;

// ---------------------------------------------------------------- f16 ------
// Round-to-nearest-even double -> binary16, the conversion a Float16Array
// store performs.  Values beyond the largest finite half become infinity.
uint16_t double_to_half(double v) {
    uint16_t sign = 0;
    if (std::signbit(v)) { sign = 0x8000; v = -v; }
    if (std::isnan(v)) return sign | 0x7e00;
    if (v >= 65520.0) return sign | 0x7c00;            // rounds past 65504
    if (v < std::ldexp(1.0, -25)) return sign;          // below half the min subnormal
    int e;
    std::frexp(v, &e);                                   // v = m * 2^e, m in [0.5, 1)
    int exp_unbiased = e - 1;                            // v = 1.f * 2^exp_unbiased
    if (exp_unbiased < -14) {
        // subnormal: value = k * 2^-24
        double k = v * std::ldexp(1.0, 24);
        double r = std::nearbyint(k);                    // default rounding mode is RNE
        return sign | static_cast<uint16_t>(r);          // r == 1024 becomes min normal
    }
    double scaled = std::ldexp(v, 10 - exp_unbiased);   // in [1024, 2048)
    double r = std::nearbyint(scaled);
    int mant = static_cast<int>(r);
    int ex = exp_unbiased;
    if (mant == 2048) { mant = 1024; ex += 1; }
    if (ex > 15) return sign | 0x7c00;
    return sign | static_cast<uint16_t>(((ex + 15) << 10) | (mant - 1024));
}

double half_to_double(uint16_t h) {
    int sign = (h >> 15) & 1;
    int ex = (h >> 10) & 0x1f;
    int mant = h & 0x3ff;
    double v;
    if (ex == 0) v = std::ldexp(static_cast<double>(mant), -24);
    else if (ex == 31) v = mant ? NAN : INFINITY;
    else v = std::ldexp(static_cast<double>(mant + 1024), ex - 25);
    return sign ? -v : v;
}

// -------------------------------------------------------------- rng --------
// SplitMix64 keyed by (stream seed, counter): every synthetic draw is a pure
// function of its key, so scenes do not depend on generation order.
struct Rng {
    uint64_t state;
    explicit Rng(uint64_t seed) : state(seed) {}
    uint64_t next_u64() {
        uint64_t z = (state += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    double uniform() { return static_cast<double>(next_u64() >> 11) * (1.0 / 9007199254740992.0); }
    double uniform(double a, double b) { return a + (b - a) * uniform(); }
};

uint64_t mix_seed(uint64_t a, uint64_t b) {
    Rng r(a ^ (b * 0xD1B54A32D192ED03ull));
    return r.next_u64();
}

// ----------------------------------------------------------- glMatrix ------
using V3 = std::array<float, 3>;
struct M4 { float m[16]; };

M4 from_rotation_translation_scale(const float q[4], const float t[3], const float s[3]) {
    // Column-major rotation from a unit quaternion, columns scaled by s.
    double x = q[0], y = q[1], z = q[2], w = q[3];
    double x2 = x + x, y2 = y + y, z2 = z + z;
    double xx = x * x2, xy = x * y2, xz = x * z2;
    double yy = y * y2, yz = y * z2, zz = z * z2;
    double wx = w * x2, wy = w * y2, wz = w * z2;
    M4 o;
    o.m[0] = float((1 - (yy + zz)) * s[0]);
    o.m[1] = float((xy + wz) * s[0]);
    o.m[2] = float((xz - wy) * s[0]);
    o.m[3] = 0.f;
    o.m[4] = float((xy - wz) * s[1]);
    o.m[5] = float((1 - (xx + zz)) * s[1]);
    o.m[6] = float((yz + wx) * s[1]);
    o.m[7] = 0.f;
    o.m[8] = float((xz + wy) * s[2]);
    o.m[9] = float((yz - wx) * s[2]);
    o.m[10] = float((1 - (xx + yy)) * s[2]);
    o.m[11] = 0.f;
    o.m[12] = t[0];
    o.m[13] = t[1];
    o.m[14] = t[2];
    o.m[15] = 1.f;
    return o;
}

// General 4x4 inverse by Gauss-Jordan elimination with partial pivoting in
// double, rounded to float.  Returns false for a singular matrix.
bool invert(const M4& a, M4* out) {
    double A[4][8];
    for (int r = 0; r < 4; ++r) {
        for (int c = 0; c < 4; ++c) A[r][c] = a.m[c * 4 + r];
        for (int c = 0; c < 4; ++c) A[r][4 + c] = (r == c) ? 1.0 : 0.0;
    }
    for (int col = 0; col < 4; ++col) {
        int piv = col;
        for (int r = col + 1; r < 4; ++r)
            if (std::fabs(A[r][col]) > std::fabs(A[piv][col])) piv = r;
        if (A[piv][col] == 0.0) return false;
        if (piv != col) for (int c = 0; c < 8; ++c) std::swap(A[piv][c], A[col][c]);
        double inv = 1.0 / A[col][col];
        for (int c = 0; c < 8; ++c) A[col][c] *= inv;
        for (int r = 0; r < 4; ++r) {
            if (r == col) continue;
            double f = A[r][col];
            if (f == 0.0) continue;
            for (int c = 0; c < 8; ++c) A[r][c] -= f * A[col][c];
        }
    }
    for (int r = 0; r < 4; ++r)
        for (int c = 0; c < 4; ++c) out->m[c * 4 + r] = float(A[r][4 + c]);
    return true;
}

V3 transform_point(const M4& m, const V3& p) {
    V3 o;
    for (int i = 0; i < 3; ++i)
        o[i] = float(double(m.m[i]) * p[0] + double(m.m[4 + i]) * p[1] + double(m.m[8 + i]) * p[2] +
                     double(m.m[12 + i]));
    return o;
}

struct Aabb {
    V3 mn{INFINITY, INFINITY, INFINITY};
    V3 mx{-INFINITY, -INFINITY, -INFINITY};
    void grow(const V3& p) {
        for (int i = 0; i < 3; ++i) { mn[i] = std::min(mn[i], p[i]); mx[i] = std::max(mx[i], p[i]); }
    }
    void grow(const Aabb& b) { grow(b.mn); grow(b.mx); }
};

// transformAABB (tracer.js:300-317): bounds of the 8 transformed corners.
Aabb transform_aabb(const V3& lo, const V3& hi, const M4& m) {
    Aabb b;
    for (int ix = 0; ix < 2; ++ix)
        for (int iy = 0; iy < 2; ++iy)
            for (int iz = 0; iz < 2; ++iz) {
                V3 p{ix ? hi[0] : lo[0], iy ? hi[1] : lo[1], iz ? hi[2] : lo[2]};
                b.grow(transform_point(m, p));
            }
    return b;
}

// ------------------------------------------------------------ model --------
struct Model {
    std::vector<pt_bvh_node> nodes;     // flattened DFS, relative indices
    std::vector<pt_triangle> tris;      // leaf order
    uint32_t max_depth = 0;
    uint32_t n_vertices = 0;
    std::vector<float> positions;       // kept for the precise-bounds path
    uint32_t node_offset = 0, tri_offset = 0;
};

struct RawTri {
    double c[3];                        // centroid in double, as the JS computes it
    float v[3][3];
};

// ModelData.generateBVH (tracer.js:1025-1118).  The split search evaluates
// the same candidate planes (start + i/numBins * extent, numBins =
// min(n, 12)) with the same strict `centroid < split` test and the same
// double-precision cost n_L*A_L + n_R*A_R against n*A, but gathers the
// per-candidate counts and bounds from one binning pass: a centroid is left
// of candidate i exactly when fewer than i candidates are <= it, because the
// candidates increase with i.
class BlasBuilder {
  public:
    BlasBuilder(const std::vector<RawTri>& tris, Model* out) : tris_(tris), out_(out) {}

    void build() {
        std::vector<uint32_t> idx(tris_.size());
        for (uint32_t i = 0; i < idx.size(); ++i) idx[i] = i;
        out_->nodes.clear();
        out_->tris.clear();
        out_->max_depth = 0;
        if (idx.empty()) return;
        subdivide_and_flatten(idx, 0);
    }

  private:
    struct DBox {
        double mn[3] = {INFINITY, INFINITY, INFINITY};
        double mx[3] = {-INFINITY, -INFINITY, -INFINITY};
        void grow_tri(const RawTri& t) {
            for (int v = 0; v < 3; ++v)
                for (int k = 0; k < 3; ++k) {
                    mn[k] = std::min(mn[k], double(t.v[v][k]));
                    mx[k] = std::max(mx[k], double(t.v[v][k]));
                }
        }
        void grow(const DBox& b) {
            for (int k = 0; k < 3; ++k) { mn[k] = std::min(mn[k], b.mn[k]); mx[k] = std::max(mx[k], b.mx[k]); }
        }
        double area() const {
            double d0 = mx[0] - mn[0], d1 = mx[1] - mn[1], d2 = mx[2] - mn[2];
            return 2 * (d0 * d1 + d1 * d2 + d2 * d0);
        }
    };

    // Recursion builds the node tree in the JS's DFS order directly: the
    // node is emitted, then the left subtree, then the right subtree.
    uint32_t subdivide_and_flatten(std::vector<uint32_t>& indices, uint32_t depth) {
        out_->max_depth = std::max(out_->max_depth, depth);
        DBox box;
        for (uint32_t i : indices) box.grow_tri(tris_[i]);
        const double root_area = box.area();
        double best_cost = double(indices.size()) * root_area;
        int best_axis = -1;
        double best_split = 0;
        const size_t n = indices.size();
        const int num_bins = int(std::min<size_t>(n, 12));

        if (n > 2) {
            for (int axis = 0; axis < 3; ++axis) {
                const double start = box.mn[axis], end = box.mx[axis];
                if (end - start < 1e-6) continue;
                double splits[12];
                for (int i = 1; i < num_bins; ++i)
                    splits[i] = start + (double(i) / double(num_bins)) * (end - start);
                // bin b holds centroids with exactly b candidates <= centroid
                DBox bins[12];
                size_t counts[12] = {0};
                for (uint32_t ti : indices) {
                    const double c = tris_[ti].c[axis];
                    int b = 0;
                    while (b + 1 < num_bins && splits[b + 1] <= c) ++b;
                    bins[b].grow_tri(tris_[ti]);
                    counts[b]++;
                }
                // suffix bounds for the right side
                DBox suffix[13];
                size_t suffix_n[13] = {0};
                for (int b = num_bins - 1; b >= 0; --b) {
                    suffix[b] = suffix[b + 1];
                    suffix[b].grow(bins[b]);
                    suffix_n[b] = suffix_n[b + 1] + counts[b];
                }
                DBox left;
                size_t left_n = 0;
                for (int i = 1; i < num_bins; ++i) {
                    left.grow(bins[i - 1]);
                    left_n += counts[i - 1];
                    const size_t right_n = suffix_n[i];
                    if (left_n == 0 || right_n == 0) continue;
                    const double cost = double(left_n) * left.area() + double(right_n) * suffix[i].area();
                    if (cost < best_cost) {
                        best_cost = cost;
                        best_axis = axis;
                        best_split = splits[i];
                    }
                }
            }
        }

        const uint32_t node_index = uint32_t(out_->nodes.size());
        out_->nodes.push_back(pt_bvh_node{});
        {
            pt_bvh_node& nd = out_->nodes[node_index];
            for (int k = 0; k < 3; ++k) { nd.aabb_min[k] = float(box.mn[k]); nd.aabb_max[k] = float(box.mx[k]); }
        }
        if (best_axis == -1 || n <= 2) {
            pt_bvh_node& nd = out_->nodes[node_index];
            nd.count = uint32_t(n);
            nd.next = uint32_t(out_->tris.size());
            for (uint32_t ti : indices) out_->tris.push_back(make_triangle(ti));
            return node_index;
        }
        std::vector<uint32_t> left_idx, right_idx;
        left_idx.reserve(n);
        right_idx.reserve(n);
        for (uint32_t ti : indices) {
            if (tris_[ti].c[best_axis] < best_split) left_idx.push_back(ti);
            else right_idx.push_back(ti);
        }
        indices.clear();
        indices.shrink_to_fit();
        subdivide_and_flatten(left_idx, depth + 1);
        left_idx.clear();
        left_idx.shrink_to_fit();
        const uint32_t right = subdivide_and_flatten(right_idx, depth + 1);
        out_->nodes[node_index].count = 0;
        out_->nodes[node_index].next = right;
        return node_index;
    }

    pt_triangle make_triangle(uint32_t ti) const { return shading_[ti]; }

  public:
    std::vector<pt_triangle> shading_;  // full records in input order

  private:
    const std::vector<RawTri>& tris_;
    Model* out_;
};

// ------------------------------------------------------------ scene --------
enum class Kind { Sphere, Cube, Cylinder, Torus, Plane, Instance };

struct Prim {
    Kind kind;
    int material = 0;
    float pos[3] = {0, 0, 0};
    float rot[4] = {0, 0, 0, 1};
    float scale[3] = {1, 1, 1};
    float param = 0;         // cylinder top ratio / torus inner ratio
    float plane_n[3] = {0, 1, 0};
    float plane_d = 0;
    int model = -1;
    bool in_bvh = true;
    bool enable_nee = false;
    int light_idx = -1;
    M4 matrix, inv;

    void update_matrix() {
        matrix = from_rotation_translation_scale(rot, pos, scale);
        if (!invert(matrix, &inv)) std::memset(&inv, 0, sizeof(inv));
    }
};

}  // namespace

struct ptgen_scene {
    std::string err;
    std::vector<pt_material> materials;
    std::vector<ptgen_material> material_src;
    std::vector<Prim> prims;
    std::vector<std::unique_ptr<Model>> models;
    // sky
    uint32_t sky_w = 1, sky_h = 1;
    std::vector<uint16_t> sky;       // rgba16f
    float sky_exposure = 1.f;
    bool sky_plain = true;
    // camera
    float eye[3] = {0.f, 1.5f, 4.5f};
    float target[3] = {0.f, 0.5f, 0.f};
    float fov = 45.f;
    float exposure = 1.f;
    int bounces = 8;
    // packed
    bool finalized = false;
    std::vector<pt_object> objects;
    std::vector<pt_bvh_node> tlas;
    std::vector<pt_mesh_instance> meshes;
    std::vector<pt_bvh_node> bvh;
    std::vector<pt_triangle> triangles;
    std::vector<pt_plane> planes;
    std::vector<pt_light> lights;
    std::vector<float> sky_cdf;
    std::vector<uint8_t> textures;   // 512 x 512 rgba8 layers
    uint32_t texture_layers = 0;
    double total_lum = 0;
    float total_light_power = 0;
    uint32_t flags = 0;

    ptgen_scene() {
        sky = {double_to_half(0.02), double_to_half(0.03), double_to_half(0.05), double_to_half(1.0)};
    }
    int fail(const char* m) { err = m; return PT_ERR_ARG; }
};

namespace {

float f32(double v) { return static_cast<float>(v); }

pt_material pack_material(const ptgen_material& m) {
    // Material constructor defaults (tracer.js:208-255) then getSchema
    // (tracer.js:258-297).
    auto dflt = [](float v, float d) { return v != 0.f && !std::isnan(v) ? v : d; };
    pt_material o;
    std::memset(&o, 0, sizeof(o));
    for (int i = 0; i < 3; ++i) o.color[i] = f32(std::pow(double(m.color[i]), 2.2));
    o.metallic = m.metallic;
    o.roughness = m.roughness;
    o.ior = dflt(m.ior, 1.5f);
    o.specular_tint = m.specular_tint;
    o.anisotropic = m.anisotropic;
    o.aniso_rotation = m.anisotropic_rotation;
    o.sheen = m.sheen;
    o.sheen_tint = dflt(m.sheen_tint, 0.5f);
    o.clearcoat = m.clearcoat;
    o.clearcoat_gloss = dflt(m.clearcoat_gloss, 1.0f);
    o.clearcoat_ior = dflt(m.clearcoat_ior, 1.5f);
    o.transmission = m.transmission;
    o.concentration = dflt(m.concentration, 1.0f);
    const float* tint = (m.subsurface_tint[0] < 0.f) ? m.color : m.subsurface_tint;
    for (int i = 0; i < 3; ++i) o.subsurface_tint[i] = f32(std::pow(double(tint[i]), 2.2));
    o.subsurface = m.subsurface;
    for (int i = 0; i < 3; ++i) o.emittance[i] = f32(double(m.emittance[i]) * double(m.emission_intensity));
    o.emissive_idx = o.albedo_idx = o.normal_idx = o.height_idx = o.roughness_idx = o.metallic_idx = -1;
    o.uv_scale[0] = o.uv_scale[1] = 1.f;
    o.height_params[0] = 1.f;    // normalMultiplier
    o.height_params[1] = 0.05f;  // heightMultiplier
    o.height_params[2] = 32.f;   // heightSamp
    o.height_params[3] = 0.f;    // heightOffset
    return o;
}

// getBounds per primitive type.
Aabb prim_bounds(const Prim& p, const std::vector<std::unique_ptr<Model>>& models) {
    const float* m = p.matrix.m;
    switch (p.kind) {
        case Kind::Sphere: {  // tracer.js:437-452
            double ex = std::sqrt(double(m[0]) * m[0] + double(m[4]) * m[4] + double(m[8]) * m[8]);
            double ey = std::sqrt(double(m[1]) * m[1] + double(m[5]) * m[5] + double(m[9]) * m[9]);
            double ez = std::sqrt(double(m[2]) * m[2] + double(m[6]) * m[6] + double(m[10]) * m[10]);
            Aabb b;
            b.mn = {f32(m[12] - ex), f32(m[13] - ey), f32(m[14] - ez)};
            b.mx = {f32(m[12] + ex), f32(m[13] + ey), f32(m[14] + ez)};
            return b;
        }
        case Kind::Cylinder: {  // tracer.js:645-669
            auto circle = [&](double y_local, double radius) {
                double ex = std::sqrt(std::pow(radius * m[0], 2) + std::pow(radius * m[8], 2));
                double ey = std::sqrt(std::pow(radius * m[1], 2) + std::pow(radius * m[9], 2));
                double ez = std::sqrt(std::pow(radius * m[2], 2) + std::pow(radius * m[10], 2));
                V3 c = transform_point(p.matrix, V3{0.f, f32(y_local), 0.f});
                Aabb b;
                b.mn = {f32(c[0] - ex), f32(c[1] - ey), f32(c[2] - ez)};
                b.mx = {f32(c[0] + ex), f32(c[1] + ey), f32(c[2] + ez)};
                return b;
            };
            Aabb b = circle(0, 1), t = circle(1, p.param);
            b.grow(t);
            return b;
        }
        case Kind::Torus: {  // tracer.js:731-761
            double ex = std::sqrt(std::pow(double(m[0]), 2) + std::pow(double(m[8]), 2));
            double ey = std::sqrt(std::pow(double(m[1]), 2) + std::pow(double(m[9]), 2));
            double ez = std::sqrt(std::pow(double(m[2]), 2) + std::pow(double(m[10]), 2));
            V3 c = transform_point(p.matrix, V3{0.f, 0.f, 0.f});
            double sx = std::sqrt(double(m[0]) * m[0] + double(m[1]) * m[1] + double(m[2]) * m[2]);
            double sy = std::sqrt(double(m[4]) * m[4] + double(m[5]) * m[5] + double(m[6]) * m[6]);
            double sz = std::sqrt(double(m[8]) * m[8] + double(m[9]) * m[9] + double(m[10]) * m[10]);
            double th = p.param * std::max(sx, std::max(sy, sz));
            Aabb b;
            b.mn = {f32(c[0] - ex - th), f32(c[1] - ey - th), f32(c[2] - ez - th)};
            b.mx = {f32(c[0] + ex + th), f32(c[1] + ey + th), f32(c[2] + ez + th)};
            return b;
        }
        case Kind::Instance: {  // Model.getBounds tracer.js:1360-1403
            M4 world;
            invert(p.inv, &world);
            const Model& md = *models[p.model];
            if (md.n_vertices > 0 && md.n_vertices < 5000) {
                Aabb b;
                for (uint32_t i = 0; i < md.n_vertices; ++i)
                    b.grow(transform_point(world, V3{md.positions[3 * i], md.positions[3 * i + 1],
                                                     md.positions[3 * i + 2]}));
                return b;
            }
            if (md.nodes.empty()) {
                Aabb b;
                b.mn = {-1, -1, -1};
                b.mx = {1, 1, 1};
                return b;
            }
            V3 lo{md.nodes[0].aabb_min[0], md.nodes[0].aabb_min[1], md.nodes[0].aabb_min[2]};
            V3 hi{md.nodes[0].aabb_max[0], md.nodes[0].aabb_max[1], md.nodes[0].aabb_max[2]};
            return transform_aabb(lo, hi, world);
        }
        case Kind::Cube:
        default: {  // Primitive.getBounds tracer.js:420-425
            M4 world;
            invert(p.inv, &world);
            return transform_aabb(V3{-1, -1, -1}, V3{1, 1, 1}, world);
        }
    }
}

// SceneBVHBuilder.recursiveBuild (tracer.js:1534-1590): median split on the
// widest axis after a stable sort by box centre (V8's Array.prototype.sort is
// stable), leaf at <= 1 item or depth > 20.
struct TlasItem { uint32_t index; Aabb box; };

void tlas_build(std::vector<TlasItem> items, int depth, std::vector<pt_bvh_node>& nodes) {
    const uint32_t node_idx = uint32_t(nodes.size());
    nodes.push_back(pt_bvh_node{});
    Aabb b;
    for (const auto& it : items) { b.grow(it.box.mn); b.grow(it.box.mx); }
    for (int k = 0; k < 3; ++k) { nodes[node_idx].aabb_min[k] = b.mn[k]; nodes[node_idx].aabb_max[k] = b.mx[k]; }
    if (items.size() <= 1 || depth > 20) {
        nodes[node_idx].count = uint32_t(items.size());
        nodes[node_idx].next = items[0].index;
        return;
    }
    float size[3];
    for (int k = 0; k < 3; ++k) size[k] = f32(double(b.mx[k]) - double(b.mn[k]));
    int axis = 0;
    if (size[1] > size[0]) axis = 1;
    if (size[2] > size[axis]) axis = 2;
    std::stable_sort(items.begin(), items.end(), [axis](const TlasItem& a, const TlasItem& c) {
        double ca = (double(a.box.mn[axis]) + double(a.box.mx[axis])) * 0.5;
        double cc = (double(c.box.mn[axis]) + double(c.box.mx[axis])) * 0.5;
        return ca - cc < 0;
    });
    const size_t mid = items.size() / 2;
    tlas_build(std::vector<TlasItem>(items.begin(), items.begin() + mid), depth + 1, nodes);
    nodes[node_idx].next = uint32_t(nodes.size());
    nodes[node_idx].count = 0;
    tlas_build(std::vector<TlasItem>(items.begin() + mid, items.end()), depth + 1, nodes);
}

void put_u32_as_float(float* dst, uint32_t v) { std::memcpy(dst, &v, 4); }

}  // namespace

// ------------------------------------------------------------- C ABI -------
extern "C" {

int ptgen_new(ptgen_scene** out) {
    if (!out) return PT_ERR_ARG;
    *out = new (std::nothrow) ptgen_scene();
    return *out ? PT_OK : PT_ERR_OOM;
}

void ptgen_destroy(ptgen_scene* s) { delete s; }

const char* ptgen_last_error(const ptgen_scene* s) { return s ? s->err.c_str() : "null scene"; }

int ptgen_add_material(ptgen_scene* s, const ptgen_material* m) {
    if (!s || !m) return PT_ERR_ARG;
    s->finalized = false;
    s->material_src.push_back(*m);
    s->materials.push_back(pack_material(*m));
    return int(s->materials.size() - 1);
}

int ptgen_add_texture(ptgen_scene* s, const uint8_t* rgba) {
    if (!s || !rgba) return PT_ERR_ARG;
    if (s->texture_layers >= 4096) return s->fail("at most 4096 texture layers");
    s->textures.insert(s->textures.end(), rgba, rgba + size_t(512) * 512 * 4);
    return int(s->texture_layers++);
}

int ptgen_set_material_maps(ptgen_scene* s, int material, const int32_t maps[6], const float uv_scale[2]) {
    if (!s || !maps) return PT_ERR_ARG;
    if (material < 0 || size_t(material) >= s->materials.size()) return s->fail("bad material index");
    for (int k = 0; k < 6; ++k)
        if (maps[k] >= int32_t(s->texture_layers)) return s->fail("texture index beyond the added layers");
    pt_material& m = s->materials[size_t(material)];
    m.emissive_idx = maps[0] < 0 ? -1 : maps[0];
    m.albedo_idx = maps[1] < 0 ? -1 : maps[1];
    m.normal_idx = maps[2] < 0 ? -1 : maps[2];
    m.height_idx = maps[3] < 0 ? -1 : maps[3];
    m.roughness_idx = maps[4] < 0 ? -1 : maps[4];
    m.metallic_idx = maps[5] < 0 ? -1 : maps[5];
    if (uv_scale) {
        m.uv_scale[0] = uv_scale[0];
        m.uv_scale[1] = uv_scale[1];
    }
    s->finalized = false;
    return PT_OK;
}

int ptgen_set_height_params(ptgen_scene* s, int material, const float params[4]) {
    if (!s || !params) return PT_ERR_ARG;
    if (material < 0 || size_t(material) >= s->materials.size()) return s->fail("bad material index");
    if (!(params[2] > 0.0f && std::isfinite(params[2]))) return s->fail("height map layers must be positive");
    pt_material& m = s->materials[size_t(material)];
    for (int k = 0; k < 4; ++k) m.height_params[k] = params[k];
    s->finalized = false;
    return PT_OK;
}

static int add_prim(ptgen_scene* s, Prim p) {
    s->finalized = false;
    p.update_matrix();
    s->prims.push_back(p);
    return int(s->prims.size() - 1);
}

static bool valid_material(const ptgen_scene* s, int m) { return m >= 0 && size_t(m) < s->materials.size(); }

int ptgen_add_sphere(ptgen_scene* s, int material, const float pos[3], float radius, int enable_nee) {
    if (!s || !pos || !valid_material(s, material) || !(radius > 0)) return s ? s->fail("bad sphere") : PT_ERR_ARG;
    Prim p;
    p.kind = Kind::Sphere;
    p.material = material;
    std::memcpy(p.pos, pos, 12);
    p.scale[0] = p.scale[1] = p.scale[2] = radius;
    p.enable_nee = enable_nee != 0;
    return add_prim(s, p);
}

int ptgen_add_cube(ptgen_scene* s, int material, const float pos[3], const float half[3],
                   const float rotation[4], int enable_nee) {
    if (!s || !pos || !half || !valid_material(s, material)) return s ? s->fail("bad cube") : PT_ERR_ARG;
    Prim p;
    p.kind = Kind::Cube;
    p.material = material;
    std::memcpy(p.pos, pos, 12);
    std::memcpy(p.scale, half, 12);
    if (rotation) std::memcpy(p.rot, rotation, 16);
    p.enable_nee = enable_nee != 0;
    return add_prim(s, p);
}

int ptgen_add_cylinder(ptgen_scene* s, int material, const float pos[3], float radius, float height,
                       const float rotation[4], float top_ratio) {
    if (!s || !pos || !valid_material(s, material)) return s ? s->fail("bad cylinder") : PT_ERR_ARG;
    Prim p;
    p.kind = Kind::Cylinder;
    p.material = material;
    std::memcpy(p.pos, pos, 12);
    p.scale[0] = radius;
    p.scale[1] = height;
    p.scale[2] = radius;
    if (rotation) std::memcpy(p.rot, rotation, 16);
    p.param = top_ratio;
    return add_prim(s, p);
}

int ptgen_add_torus(ptgen_scene* s, int material, const float pos[3], const float rotation[4],
                    float outer_radius, float inner_radius) {
    if (!s || !pos || !valid_material(s, material) || !(outer_radius > 0)) return s ? s->fail("bad torus") : PT_ERR_ARG;
    Prim p;
    p.kind = Kind::Torus;
    p.material = material;
    std::memcpy(p.pos, pos, 12);
    if (rotation) std::memcpy(p.rot, rotation, 16);
    p.scale[0] = p.scale[1] = p.scale[2] = outer_radius;  // Torus.setRadii tracer.js:693-700
    p.param = f32(double(inner_radius) / double(outer_radius));
    return add_prim(s, p);
}

int ptgen_add_plane(ptgen_scene* s, int material, const float normal[3], float d) {
    if (!s || !normal || !valid_material(s, material)) return s ? s->fail("bad plane") : PT_ERR_ARG;
    double len = std::sqrt(double(normal[0]) * normal[0] + double(normal[1]) * normal[1] +
                           double(normal[2]) * normal[2]);
    if (!(len > 0)) return s->fail("plane normal is zero");
    Prim p;
    p.kind = Kind::Plane;
    p.material = material;
    for (int i = 0; i < 3; ++i) p.plane_n[i] = f32(normal[i] / len);
    double dd = 0;
    for (int i = 0; i < 3; ++i) dd += double(f32(double(p.plane_n[i]) * d)) * p.plane_n[i];
    p.plane_d = f32(dd);  // getOrientation: dot(position, normal), tracer.js:805-810
    return add_prim(s, p);
}

int ptgen_add_model(ptgen_scene* s, const float* positions, const float* normals, const float* uvs,
                    uint32_t n_vertices, const uint32_t* indices, uint32_t n_triangles) {
    if (!s || !positions || !normals || !indices || n_triangles == 0) return s ? s->fail("bad model") : PT_ERR_ARG;
    for (uint64_t i = 0; i < uint64_t(n_triangles) * 3; ++i)
        if (indices[i] >= n_vertices) return s->fail("model index out of range");
    s->finalized = false;
    auto model = std::make_unique<Model>();
    model->n_vertices = n_vertices;
    model->positions.assign(positions, positions + size_t(n_vertices) * 3);
    std::vector<RawTri> raw(n_triangles);
    BlasBuilder builder(raw, model.get());
    builder.shading_.resize(n_triangles);
    for (uint32_t t = 0; t < n_triangles; ++t) {
        pt_triangle& tr = builder.shading_[t];
        std::memset(&tr, 0, sizeof(tr));
        float* vs[3] = {tr.v0, tr.v1, tr.v2};
        float* ns[3] = {tr.n0, tr.n1, tr.n2};
        float* us[3] = {tr.uv0, tr.uv1, tr.uv2};
        for (int k = 0; k < 3; ++k) {
            uint32_t vi = indices[3 * t + k];
            for (int c = 0; c < 3; ++c) {
                vs[k][c] = positions[3 * vi + c];
                ns[k][c] = normals[3 * vi + c];
                raw[t].v[k][c] = positions[3 * vi + c];
            }
            us[k][0] = uvs ? uvs[2 * vi] : 0.f;
            us[k][1] = uvs ? uvs[2 * vi + 1] : 0.f;
        }
        for (int c = 0; c < 3; ++c)
            raw[t].c[c] = (double(raw[t].v[0][c]) + double(raw[t].v[1][c]) + double(raw[t].v[2][c])) / 3;
    }
    builder.build();
    s->models.push_back(std::move(model));
    return int(s->models.size() - 1);
}

int ptgen_add_instance(ptgen_scene* s, int model, int material, const float pos[3], const float rotation[4],
                       const float scale[3], int in_scene_bvh) {
    if (!s || !pos || !valid_material(s, material) || model < 0 || size_t(model) >= s->models.size())
        return s ? s->fail("bad instance") : PT_ERR_ARG;
    Prim p;
    p.kind = Kind::Instance;
    p.material = material;
    p.model = model;
    std::memcpy(p.pos, pos, 12);
    if (rotation) std::memcpy(p.rot, rotation, 16);
    if (scale) std::memcpy(p.scale, scale, 12);
    p.in_bvh = in_scene_bvh != 0;
    return add_prim(s, p);
}

int ptgen_set_sky(ptgen_scene* s, uint32_t width, uint32_t height, const float* rgb, float exposure) {
    if (!s) return PT_ERR_ARG;
    s->finalized = false;
    s->sky_exposure = exposure;
    if (!rgb) {
        s->sky_w = s->sky_h = 1;
        s->sky = {double_to_half(0.02), double_to_half(0.03), double_to_half(0.05), double_to_half(1.0)};
        s->sky_plain = true;
        return PT_OK;
    }
    if (width == 0 || height == 0) return s->fail("sky size is zero");
    s->sky_w = width;
    s->sky_h = height;
    s->sky_plain = false;
    s->sky.resize(size_t(width) * height * 4);
    const double threshold = 65536 - 16 - 1;  // HDRTexture.threshold, tracer.js:30
    for (size_t i = 0; i < size_t(width) * height; ++i) {
        for (int c = 0; c < 3; ++c)
            s->sky[i * 4 + c] = double_to_half(std::min(double(rgb[i * 3 + c]) * exposure, threshold));
        s->sky[i * 4 + 3] = double_to_half(1.0);
    }
    return PT_OK;
}

int ptgen_set_sky_procedural(ptgen_scene* s, uint32_t width, uint32_t height, uint64_t seed, float sun_peak) {
    if (!s || width == 0 || height == 0) return PT_ERR_ARG;
    Rng r(mix_seed(seed, 0x5C1));
    const double sun_elev = r.uniform(20.0, 60.0) * kPi / 180.0;
    const double sun_az = r.uniform(-kPi, kPi);
    const double sd[3] = {std::cos(sun_elev) * std::cos(sun_az), std::sin(sun_elev),
                          std::cos(sun_elev) * std::sin(sun_az)};
    const double sigma = 2.0 * kPi / 180.0;
    std::vector<float> rgb(size_t(width) * height * 3);
    // base sky, then scale the sun so that peak / mean(base) == sun_peak
    double mean = 0;
    std::vector<double> lobe(size_t(width) * height);
    for (uint32_t y = 0; y < height; ++y) {
        const double theta = kPi * (y + 0.5) / height;
        for (uint32_t x = 0; x < width; ++x) {
            const double phi = 2 * kPi * (x + 0.5) / width - kPi;
            const double d[3] = {std::sin(theta) * std::cos(phi), std::cos(theta), std::sin(theta) * std::sin(phi)};
            double c[3];
            if (d[1] >= 0) {
                const double t = std::sqrt(d[1]);
                const double hz[3] = {1.0, 0.9, 0.8}, zn[3] = {0.25, 0.45, 1.0};
                for (int k = 0; k < 3; ++k) c[k] = hz[k] * (1 - t) + zn[k] * t;
            } else {
                const double g[3] = {0.25, 0.22, 0.18};
                for (int k = 0; k < 3; ++k) c[k] = g[k] * (1.0 + 0.5 * d[1]);
            }
            const size_t i = size_t(y) * width + x;
            for (int k = 0; k < 3; ++k) rgb[i * 3 + k] = float(c[k]);
            mean += 0.2126 * c[0] + 0.7152 * c[1] + 0.0722 * c[2];
            const double cosang = std::max(-1.0, std::min(1.0, d[0] * sd[0] + d[1] * sd[1] + d[2] * sd[2]));
            const double ang = std::acos(cosang);
            lobe[i] = std::exp(-(ang * ang) / (2 * sigma * sigma));
        }
    }
    mean /= double(width) * height;
    const double peak = double(sun_peak) * mean;
    const double sun_col[3] = {1.0, 0.95, 0.85};
    for (size_t i = 0; i < lobe.size(); ++i)
        for (int k = 0; k < 3; ++k) rgb[i * 3 + k] = float(rgb[i * 3 + k] + peak * sun_col[k] * lobe[i]);
    return ptgen_set_sky(s, width, height, rgb.data(), 1.0f);
}

int ptgen_set_camera(ptgen_scene* s, const float eye[3], const float target[3], float fov_deg, float exposure) {
    if (!s || !eye || !target) return PT_ERR_ARG;
    std::memcpy(s->eye, eye, 12);
    std::memcpy(s->target, target, 12);
    s->fov = fov_deg;
    s->exposure = exposure;
    return PT_OK;
}

int ptgen_set_bounces(ptgen_scene* s, int bounces) {
    if (!s || bounces < 1) return PT_ERR_ARG;
    s->bounces = bounces;
    return PT_OK;
}

int ptgen_build_sky_cdf(const uint16_t* data, uint32_t width, uint32_t height, float exposure, float* cdf,
                        double* total_weight) {
    // HDRTexture.buildHDRCDF (tracer.js:135-204).  Accumulation is double;
    // every stored value is a Float32Array element, and the normalisation
    // divides the stored float by the double sum.
    if (!data || !cdf || width == 0 || height == 0) return PT_ERR_ARG;
    const double threshold = 65536 - 16 - 1;
    const size_t marg = height + 1;
    float* mcdf = cdf;
    float* ccdf = cdf + marg;
    double total = 0;
    mcdf[0] = 0.f;
    for (uint32_t y = 0; y < height; ++y) {
        double row = 0;
        const double sin_theta = std::sin(kPi * (y + 0.5) / height);
        float* crow = ccdf + size_t(y) * (width + 1);
        crow[0] = 0.f;
        for (uint32_t x = 0; x < width; ++x) {
            const size_t idx = (size_t(y) * width + x) * 4;
            const double r = std::min(half_to_double(data[idx]) * exposure, threshold);
            const double g = std::min(half_to_double(data[idx + 1]) * exposure, threshold);
            const double b = std::min(half_to_double(data[idx + 2]) * exposure, threshold);
            const double lum = (0.2126 * r + 0.7152 * g + 0.0722 * b) * sin_theta;
            row += lum;
            crow[x + 1] = f32(row);
        }
        if (row > 0) {
            for (uint32_t x = 1; x <= width; ++x) crow[x] = f32(double(crow[x]) / row);
        } else {
            for (uint32_t x = 1; x <= width; ++x) crow[x] = f32(double(x) / width);
        }
        total += row;
        mcdf[y + 1] = f32(total);
    }
    if (total > 0) {
        for (uint32_t y = 1; y <= height; ++y) mcdf[y] = f32(double(mcdf[y]) / total);
    } else {
        for (uint32_t y = 1; y <= height; ++y) mcdf[y] = f32(double(y) / height);
    }
    if (total_weight) *total_weight = total;
    return PT_OK;
}

int ptgen_finalize(ptgen_scene* s) {
    if (!s) return PT_ERR_ARG;
    if (s->materials.empty()) return s->fail("scene has no materials");
    // Split objects as packSceneData does (tracer.js:1832-1836).
    std::vector<int> primitives, in_bvh_models, list_models, planes, model_objects;
    for (size_t i = 0; i < s->prims.size(); ++i) {
        const Prim& p = s->prims[i];
        switch (p.kind) {
            case Kind::Sphere: case Kind::Cube: case Kind::Cylinder: case Kind::Torus:
                primitives.push_back(int(i)); break;
            case Kind::Plane: planes.push_back(int(i)); break;
            case Kind::Instance:
                model_objects.push_back(int(i));
                (p.in_bvh ? in_bvh_models : list_models).push_back(int(i));
                break;
        }
    }
    std::vector<int> scene_objects = primitives;
    scene_objects.insert(scene_objects.end(), in_bvh_models.begin(), in_bvh_models.end());

    // Scene.getLights / lightPower / Light (tracer.js:1654-1711).
    s->lights.clear();
    s->total_light_power = 0;
    double total_power = 0;
    for (auto& p : s->prims) p.light_idx = -1;
    for (size_t i = 0; i < scene_objects.size(); ++i) {
        Prim& p = s->prims[scene_objects[i]];
        if (!p.enable_nee) continue;
        if (p.kind != Kind::Sphere && p.kind != Kind::Cube) continue;
        const ptgen_material& m = s->material_src[p.material];
        const double lum = (0.2126 * m.emittance[0] + 0.7152 * m.emittance[1] + 0.0722 * m.emittance[2]) *
                           m.emission_intensity;
        if (lum <= 0) continue;
        double area;
        if (p.kind == Kind::Sphere) {
            const double r = (double(p.scale[0]) + p.scale[1] + p.scale[2]) / 3;
            area = 4 * kPi * r * r;
        } else {
            area = 8 * (double(p.scale[0]) * p.scale[1] + double(p.scale[1]) * p.scale[2] +
                        double(p.scale[0]) * p.scale[2]);
        }
        const double power = lum * area;
        if (power < 0.01) continue;
        pt_light L;
        std::memset(&L, 0, sizeof(L));
        L.obj_idx = int32_t(i);
        L.area = f32(area);
        L.power = f32(power);
        if (p.kind == Kind::Sphere) {
            const float avg = f32((double(p.scale[0]) + p.scale[1] + p.scale[2]) / 3);
            p.scale[0] = p.scale[1] = p.scale[2] = avg;
            p.update_matrix();
            L.scale = avg;
        } else {
            L.scale = f32(std::hypot(double(p.scale[0]), double(p.scale[1]), double(p.scale[2])));
        }
        std::memcpy(L.matrix, p.matrix.m, 64);
        p.light_idx = int(s->lights.size());
        s->lights.push_back(L);
        total_power += power;
    }
    s->total_light_power = f32(total_power);

    // Unique models in order of first use (tracer.js:1846-1855).
    std::vector<int> unique_models;
    for (int oi : model_objects) {
        int m = s->prims[oi].model;
        if (std::find(unique_models.begin(), unique_models.end(), m) == unique_models.end()) unique_models.push_back(m);
    }
    s->bvh.clear();
    s->triangles.clear();
    uint64_t total_nodes = 0, total_tris = 0;
    for (int m : unique_models) {
        Model& md = *s->models[m];
        md.node_offset = uint32_t(total_nodes);
        md.tri_offset = uint32_t(total_tris);
        total_nodes += md.nodes.size();
        total_tris += md.tris.size();
    }
    if (total_nodes > 0xffffffffull || total_tris > 0xffffffffull) return s->fail("scene too large");
    s->bvh.reserve(std::max<uint64_t>(1, total_nodes));
    s->triangles.reserve(std::max<uint64_t>(1, total_tris));
    for (int m : unique_models) {
        const Model& md = *s->models[m];
        s->bvh.insert(s->bvh.end(), md.nodes.begin(), md.nodes.end());
        s->triangles.insert(s->triangles.end(), md.tris.begin(), md.tris.end());
    }
    if (s->bvh.empty()) s->bvh.push_back(pt_bvh_node{});
    if (s->triangles.empty()) { pt_triangle t; std::memset(&t, 0, sizeof(t)); s->triangles.push_back(t); }

    // Objects in scene order (primitives, then in-BVH models).
    s->objects.clear();
    for (int oi : scene_objects) {
        const Prim& p = s->prims[oi];
        pt_object o;
        std::memset(&o, 0, sizeof(o));
        std::memcpy(o.inv_matrix, p.inv.m, 64);
        o.material_idx = p.material;
        o.light_idx = p.light_idx;
        std::memcpy(o.world_position, p.pos, 12);
        switch (p.kind) {
            case Kind::Sphere: o.object_type = PT_OBJ_SPHERE; break;
            case Kind::Cube: o.object_type = PT_OBJ_CUBE; break;
            case Kind::Cylinder: o.object_type = PT_OBJ_CYLINDER; o.param0 = p.param; break;
            case Kind::Torus: o.object_type = PT_OBJ_TORUS; o.param0 = p.param; break;
            case Kind::Instance:
                o.object_type = PT_OBJ_MESH;
                put_u32_as_float(&o.param0, s->models[p.model]->node_offset);
                put_u32_as_float(&o.param1, s->models[p.model]->tri_offset);
                break;
            default: break;
        }
        s->objects.push_back(o);
    }
    // TLAS over the same objects.
    s->tlas.clear();
    if (!scene_objects.empty()) {
        std::vector<TlasItem> items;
        for (size_t i = 0; i < scene_objects.size(); ++i)
            items.push_back(TlasItem{uint32_t(i), prim_bounds(s->prims[scene_objects[i]], s->models)});
        tlas_build(items, 0, s->tlas);
    }
    if (s->objects.empty()) { pt_object o; std::memset(&o, 0, sizeof(o)); o.material_idx = -1; o.light_idx = -1; o.object_type = 1; s->objects.push_back(o); }
    if (s->tlas.empty()) s->tlas.push_back(pt_bvh_node{});

    s->planes.clear();
    for (int pi : planes) {
        const Prim& p = s->prims[pi];
        pt_plane pl;
        std::memset(&pl, 0, sizeof(pl));
        std::memcpy(pl.normal, p.plane_n, 12);
        pl.d = p.plane_d;
        pl.material_idx = p.material;
        s->planes.push_back(pl);
    }
    if (s->planes.empty()) { pt_plane pl; std::memset(&pl, 0, sizeof(pl)); pl.material_idx = -1; s->planes.push_back(pl); }
    s->meshes.clear();
    for (int mi : list_models) {
        const Prim& p = s->prims[mi];
        pt_mesh_instance mesh;
        std::memset(&mesh, 0, sizeof(mesh));
        std::memcpy(mesh.inv_matrix, p.inv.m, 64);
        mesh.material_idx = p.material;
        mesh.node_offset = s->models[p.model]->node_offset;
        mesh.tri_offset = s->models[p.model]->tri_offset;
        s->meshes.push_back(mesh);
    }
    if (s->meshes.empty()) { pt_mesh_instance m; std::memset(&m, 0, sizeof(m)); s->meshes.push_back(m); }
    if (s->lights.empty()) { pt_light L; std::memset(&L, 0, sizeof(L)); L.obj_idx = -1; s->lights.push_back(L); }

    bool has[5] = {false, false, false, false, false};
    for (int oi : primitives) has[int(s->prims[oi].kind)] = true;
    uint32_t f = 0;
    if (has[int(Kind::Sphere)]) f |= PT_HAS_SPHERES;
    if (has[int(Kind::Cube)]) f |= PT_HAS_CUBES;
    if (has[int(Kind::Cylinder)]) f |= PT_HAS_CYLINDERS;
    if (has[int(Kind::Torus)]) f |= PT_HAS_TORI;
    if (!in_bvh_models.empty()) f |= PT_HAS_MESHES;
    if (!list_models.empty()) f |= PT_HAS_LIST_MESHES;
    if (!planes.empty()) f |= PT_HAS_PLANES;
    if (total_power > 0 && !s->lights.empty() && s->lights[0].obj_idx >= 0) f |= PT_HAS_LIGHTS;
    f |= PT_HAS_SKYBOX;
    if (!s->sky_plain) f |= PT_MIS_SKYBOX;
    for (const pt_material& m : s->materials)  // packSceneData, tracer.js:1825
        if (m.height_idx >= 0) f |= PT_HAS_HEIGHTMAPS;
    s->flags = f;

    s->sky_cdf.assign(size_t(s->sky_h + 1) + size_t(s->sky_h) * (s->sky_w + 1), 0.f);
    ptgen_build_sky_cdf(s->sky.data(), s->sky_w, s->sky_h, s->sky_exposure, s->sky_cdf.data(), &s->total_lum);
    s->finalized = true;
    return PT_OK;
}

int ptgen_get_desc(const ptgen_scene* s, pt_scene_desc* d) {
    if (!s || !d) return PT_ERR_ARG;
    if (!s->finalized) return PT_ERR_STATE;
    std::memset(d, 0, sizeof(*d));
    auto arr = [](const void* p, size_t n, uint32_t stride) { pt_array a; a.ptr = p; a.count = n; a.stride = stride; a._pad = 0; return a; };
    d->materials = arr(s->materials.data(), s->materials.size(), sizeof(pt_material));
    d->objects = arr(s->objects.data(), s->objects.size(), sizeof(pt_object));
    d->tlas = arr(s->tlas.data(), s->tlas.size(), sizeof(pt_bvh_node));
    d->meshes = arr(s->meshes.data(), s->meshes.size(), sizeof(pt_mesh_instance));
    d->bvh = arr(s->bvh.data(), s->bvh.size(), sizeof(pt_bvh_node));
    d->triangles = arr(s->triangles.data(), s->triangles.size(), sizeof(pt_triangle));
    d->planes = arr(s->planes.data(), s->planes.size(), sizeof(pt_plane));
    d->lights = arr(s->lights.data(), s->lights.size(), sizeof(pt_light));
    d->sky_width = s->sky_w;
    d->sky_height = s->sky_h;
    d->sky_rgba16f = s->sky.data();
    d->sky_cdf = s->sky_cdf.data();
    d->sky_cdf_count = s->sky_cdf.size();
    d->sky_exposure = s->sky_exposure;
    d->sky_total_lum = f32(s->total_lum);
    d->textures = s->texture_layers ? s->textures.data() : nullptr;
    d->texture_layers = s->texture_layers;
    d->flags = s->flags;
    d->bounce_limit = s->bounces;
    return PT_OK;
}

int ptgen_scene_params(const ptgen_scene* s, uint32_t width, uint32_t height, uint32_t sample_number, uint32_t seed,
                       pt_scene_params* out) {
    // Camera.updateRays2 with aperture 0 (tracer.js:1449-1498) and
    // Renderer.updateUniforms (tracer.js:2207-2240).
    if (!s || !out || width == 0 || height == 0) return PT_ERR_ARG;
    std::memset(out, 0, sizeof(*out));
    auto sub = [](const float* a, const float* b, float* o) { for (int i = 0; i < 3; ++i) o[i] = f32(double(a[i]) - b[i]); };
    auto add = [](const float* a, const float* b, float* o) { for (int i = 0; i < 3; ++i) o[i] = f32(double(a[i]) + b[i]); };
    auto length = [](const float* a) { return std::sqrt(double(a[0]) * a[0] + double(a[1]) * a[1] + double(a[2]) * a[2]); };
    auto normalize = [&](const float* a, float* o) {
        double l2 = double(a[0]) * a[0] + double(a[1]) * a[1] + double(a[2]) * a[2];
        double inv = l2 > 0 ? 1 / std::sqrt(l2) : 0;
        for (int i = 0; i < 3; ++i) o[i] = f32(a[i] * inv);
    };
    auto cross = [](const float* a, const float* b, float* o) {
        double x = double(a[1]) * b[2] - double(a[2]) * b[1];
        double y = double(a[2]) * b[0] - double(a[0]) * b[2];
        double z = double(a[0]) * b[1] - double(a[1]) * b[0];
        o[0] = f32(x); o[1] = f32(y); o[2] = f32(z);
    };
    float look[3], f[3], r[3], u[3], tmp[3], hr[3], hu[3], fc[3];
    const float up[3] = {0, 1, 0};
    sub(s->target, s->eye, look);
    const double focus_dist = length(look) * 1.0;
    normalize(look, f);
    cross(f, up, tmp);
    normalize(tmp, r);
    cross(r, f, u);
    const double h = std::tan((double(s->fov) * kPi / 180) / 2) * focus_dist;
    const double w = h * (double(width) / double(height));
    for (int i = 0; i < 3; ++i) { hr[i] = f32(r[i] * w); hu[i] = f32(u[i] * h); fc[i] = f32(s->eye[i] + f[i] * focus_dist); }
    float p00[3], p10[3], p01[3], p11[3];
    sub(fc, hr, p00); sub(p00, hu, p00);
    add(fc, hr, p10); sub(p10, hu, p10);
    sub(fc, hr, p01); add(p01, hu, p01);
    add(fc, hr, p11); add(p11, hu, p11);
    const float* jp = s->eye;  // aperture 0: the lens offset is zero
    std::memcpy(out->eye, jp, 12);
    sub(p00, jp, out->ray00);
    sub(p10, jp, out->ray10);
    sub(p01, jp, out->ray01);
    sub(p11, jp, out->ray11);
    out->sample_number = sample_number;
    out->width = width;
    out->height = height;
    out->exposure = s->exposure;
    out->seed = seed;
    out->sky_width = s->sky_w;
    out->sky_height = s->sky_h;
    out->sky_total_lum = f32(s->total_lum);
    out->total_light_power = s->total_light_power;
    out->section[0] = out->section[1] = 0;
    out->sky_cdf_index = s->sky_h + 1;
    return PT_OK;
}

int ptgen_model_info(const ptgen_scene* s, int model, uint32_t* n_nodes, uint32_t* n_tris, uint32_t* max_depth) {
    if (!s || model < 0 || size_t(model) >= s->models.size()) return PT_ERR_ARG;
    const Model& m = *s->models[model];
    if (n_nodes) *n_nodes = uint32_t(m.nodes.size());
    if (n_tris) *n_tris = uint32_t(m.tris.size());
    if (max_depth) *max_depth = m.max_depth;
    return PT_OK;
}

}  // extern "C"

// ---------------------------------------------------- synthetic scenes -----
namespace {

void random_quat(Rng& r, float q[4]) {
    // Shoemake's uniform rotation from three uniforms.
    const double u1 = r.uniform(), u2 = r.uniform(), u3 = r.uniform();
    const double a = std::sqrt(1 - u1), b = std::sqrt(u1);
    q[0] = f32(a * std::sin(2 * kPi * u2));
    q[1] = f32(a * std::cos(2 * kPi * u2));
    q[2] = f32(b * std::sin(2 * kPi * u3));
    q[3] = f32(b * std::cos(2 * kPi * u3));
}

// 3-D value noise: trilinear interpolation (smoothstep weights) of hashed
// lattice values in [-1, 1].
double lattice(uint64_t seed, int64_t x, int64_t y, int64_t z) {
    uint64_t h = seed ^ (uint64_t(x) * 0x8DA6B343ull) ^ (uint64_t(y) * 0xD8163841ull) ^ (uint64_t(z) * 0xCB1AB31Full);
    Rng r(h);
    return r.uniform() * 2 - 1;
}

double value_noise(uint64_t seed, double x, double y, double z) {
    const double fx = std::floor(x), fy = std::floor(y), fz = std::floor(z);
    const int64_t ix = int64_t(fx), iy = int64_t(fy), iz = int64_t(fz);
    auto sm = [](double t) { return t * t * (3 - 2 * t); };
    const double tx = sm(x - fx), ty = sm(y - fy), tz = sm(z - fz);
    double c[2][2][2];
    for (int a = 0; a < 2; ++a)
        for (int b = 0; b < 2; ++b)
            for (int d = 0; d < 2; ++d) c[a][b][d] = lattice(seed, ix + a, iy + b, iz + d);
    auto lerp = [](double a, double b, double t) { return a + (b - a) * t; };
    const double x00 = lerp(c[0][0][0], c[1][0][0], tx), x10 = lerp(c[0][1][0], c[1][1][0], tx);
    const double x01 = lerp(c[0][0][1], c[1][0][1], tx), x11 = lerp(c[0][1][1], c[1][1][1], tx);
    return lerp(lerp(x00, x10, ty), lerp(x01, x11, ty), tz);
}

// Geodesic sphere of frequency n (class I: each icosahedron face split into
// n^2 triangles), vertices projected to the unit sphere and displaced
// radially by 3-octave value noise of amplitude `amp`; smooth vertex normals.
void geodesic_sphere(uint32_t n, uint64_t seed, double amp, std::vector<float>& pos, std::vector<float>& nrm,
                     std::vector<float>& uv, std::vector<uint32_t>& idx) {
    const double t = (1 + std::sqrt(5.0)) / 2;
    double V[12][3] = {{-1, t, 0}, {1, t, 0}, {-1, -t, 0}, {1, -t, 0}, {0, -1, t}, {0, 1, t},
                       {0, -1, -t}, {0, 1, -t}, {t, 0, -1}, {t, 0, 1}, {-t, 0, -1}, {-t, 0, 1}};
    for (auto& v : V) { double l = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]); for (double& c : v) c /= l; }
    const int F[20][3] = {{0, 11, 5}, {0, 5, 1}, {0, 1, 7}, {0, 7, 10}, {0, 10, 11}, {1, 5, 9}, {5, 11, 4},
                          {11, 10, 2}, {10, 7, 6}, {7, 1, 8}, {3, 9, 4}, {3, 4, 2}, {3, 2, 6}, {3, 6, 8},
                          {3, 8, 9}, {4, 9, 5}, {2, 4, 11}, {6, 2, 10}, {8, 6, 7}, {9, 8, 1}};
    // Shared vertices are deduplicated by their (a, b, c, i, j) lattice key
    // mapped to a canonical global key (sorted corner ids with weights).
    std::unordered_map<uint64_t, uint32_t> lut;
    pos.clear(); nrm.clear(); uv.clear(); idx.clear();
    auto vertex = [&](int fa, int fb, int fc, uint32_t wa, uint32_t wb, uint32_t wc) -> uint32_t {
        // canonical key: the (corner, weight) pairs with weight > 0, sorted
        // by corner, 4 + 8 bits each (frequency <= 255)
        std::pair<int, uint32_t> cw[3];
        int m = 0;
        if (wa) cw[m++] = {fa, wa};
        if (wb) cw[m++] = {fb, wb};
        if (wc) cw[m++] = {fc, wc};
        std::sort(cw, cw + m);
        uint64_t key = 0;
        for (int q = 0; q < m; ++q) key = (key << 12) | (uint64_t(cw[q].first) << 8) | cw[q].second;
        auto it = lut.find(key);
        if (it != lut.end()) return it->second;
        double p[3] = {0, 0, 0};
        for (int q = 0; q < m; ++q)
            for (int k = 0; k < 3; ++k) p[k] += V[cw[q].first][k] * double(cw[q].second);
        double l = std::sqrt(p[0] * p[0] + p[1] * p[1] + p[2] * p[2]);
        for (double& c : p) c /= l;
        double nsum = 0, wsum = 0, f = 1.5;
        for (int o = 0; o < 3; ++o) {
            const double w = std::pow(0.5, o);
            nsum += w * value_noise(seed + o, p[0] * f + 17.0, p[1] * f + 31.0, p[2] * f + 47.0);
            wsum += w;
            f *= 2;
        }
        const double r = 1.0 + amp * nsum / wsum;
        const uint32_t id = uint32_t(pos.size() / 3);
        for (int k = 0; k < 3; ++k) pos.push_back(f32(p[k] * r));
        nrm.insert(nrm.end(), {0.f, 0.f, 0.f});
        uv.push_back(f32(0.5 + std::atan2(p[2], p[0]) / (2 * kPi)));
        uv.push_back(f32(0.5 - std::asin(std::max(-1.0, std::min(1.0, p[1]))) / kPi));
        lut.emplace(key, id);
        return id;
    };
    for (const auto& face : F) {
        for (uint32_t i = 0; i < n; ++i) {
            for (uint32_t j = 0; j < n - i; ++j) {
                // lattice point (i, j): weights (n-i-j, i, j) on corners (a, b, c)
                auto P = [&](uint32_t ii, uint32_t jj) { return vertex(face[0], face[1], face[2], n - ii - jj, ii, jj); };
                const uint32_t a = P(i, j), b = P(i + 1, j), c = P(i, j + 1);
                idx.insert(idx.end(), {a, b, c});
                if (j + 1 < n - i) {
                    const uint32_t d = P(i + 1, j + 1);
                    idx.insert(idx.end(), {b, d, c});
                }
            }
        }
    }
    // Smooth normals: sum of (area-weighted) face normals.
    std::vector<double> acc(pos.size(), 0.0);
    for (size_t t3 = 0; t3 < idx.size(); t3 += 3) {
        const float* a = &pos[3 * idx[t3]];
        const float* b = &pos[3 * idx[t3 + 1]];
        const float* c = &pos[3 * idx[t3 + 2]];
        double e1[3], e2[3], fn[3];
        for (int k = 0; k < 3; ++k) { e1[k] = double(b[k]) - a[k]; e2[k] = double(c[k]) - a[k]; }
        fn[0] = e1[1] * e2[2] - e1[2] * e2[1];
        fn[1] = e1[2] * e2[0] - e1[0] * e2[2];
        fn[2] = e1[0] * e2[1] - e1[1] * e2[0];
        for (int v = 0; v < 3; ++v)
            for (int k = 0; k < 3; ++k) acc[3 * idx[t3 + v] + k] += fn[k];
    }
    for (size_t v = 0; v < pos.size() / 3; ++v) {
        double l = std::sqrt(acc[3 * v] * acc[3 * v] + acc[3 * v + 1] * acc[3 * v + 1] + acc[3 * v + 2] * acc[3 * v + 2]);
        for (int k = 0; k < 3; ++k) nrm[3 * v + k] = f32(l > 0 ? acc[3 * v + k] / l : 0.0);
    }
}

ptgen_material plain_material() {
    ptgen_material m;
    std::memset(&m, 0, sizeof(m));
    m.color[0] = m.color[1] = m.color[2] = 1.f;
    m.roughness = 0.5f;
    m.emittance[0] = m.emittance[1] = m.emittance[2] = 1.f;
    m.subsurface_tint[0] = -1.f;
    return m;
}

// The configuration-1 content (primitives, emissive quad, plane, sky,
// bounce limit) without finalizing the scene.
void add_config1_content(ptgen_scene* s, uint64_t seed, const ptgen_config_opts& o) {
    Rng r(mix_seed(seed, 1));
    auto disney = [&]() {
        ptgen_material m = plain_material();
        for (int k = 0; k < 3; ++k) m.color[k] = f32(r.uniform());
        m.roughness = f32(r.uniform());
        m.metallic = r.uniform() < 0.3 ? 1.f : 0.f;
        const bool trans = r.uniform() < 0.15;
        const float ior = f32(r.uniform(1.3, 1.7));
        m.transmission = trans ? 1.f : 0.f;
        m.ior = trans ? ior : 1.5f;
        return m;
    };
    auto pos = [&](float p[3]) { for (int k = 0; k < 3; ++k) p[k] = f32(r.uniform(-2, 2)); };
    for (int i = 0; i < 6; ++i) {
        ptgen_material m = disney();
        int mi = ptgen_add_material(s, &m);
        float p[3]; pos(p);
        ptgen_add_sphere(s, mi, p, f32(r.uniform(0.2, 0.8)), 0);
    }
    for (int i = 0; i < 4; ++i) {
        ptgen_material m = disney();
        int mi = ptgen_add_material(s, &m);
        float p[3], h[3], q[4]; pos(p);
        for (int k = 0; k < 3; ++k) h[k] = f32(r.uniform(0.2, 0.8));
        random_quat(r, q);
        ptgen_add_cube(s, mi, p, h, q, 0);
    }
    {
        ptgen_material m = disney();
        int mi = ptgen_add_material(s, &m);
        float p[3], q[4]; pos(p);
        random_quat(r, q);
        const float rad = f32(r.uniform(0.2, 0.8));
        ptgen_add_cylinder(s, mi, p, rad, f32(r.uniform(0.4, 1.6)), q, 1.0f);
    }
    {   // emissive quad: a cube with one half-extent 1e-3, radiance 15, NEE on
        ptgen_material m = plain_material();
        m.emission_intensity = 15.f;
        m.roughness = 1.f;
        int mi = ptgen_add_material(s, &m);
        const float p[3] = {0.f, 3.0f, 0.5f};
        const float h[3] = {1.0f, 1e-3f, 1.0f};
        ptgen_add_cube(s, mi, p, h, nullptr, 1);
    }
    {   // ground plane y = -2
        ptgen_material m = plain_material();
        m.color[0] = m.color[1] = m.color[2] = 0.8f;
        m.roughness = 0.9f;
        int mi = ptgen_add_material(s, &m);
        const float n[3] = {0.f, 1.f, 0.f};
        ptgen_add_plane(s, mi, n, -2.f);
    }
    ptgen_set_sky_procedural(s, o.sky_width ? o.sky_width : 512, o.sky_height ? o.sky_height : 256, seed, 1e3f);
    ptgen_set_bounces(s, o.bounces > 0 ? o.bounces : 4);
}

int make_config1(ptgen_scene* s, uint64_t seed, const ptgen_config_opts& o) {
    add_config1_content(s, seed, o);
    return ptgen_finalize(s);
}

// Configuration 1 plus instanced meshes: four displaced geodesic spheres of
// frequency mesh_freq[i] (default 8), n_instances instances (default 16)
// placed like the primitives (centres U[-2,2]^3, uniform scale U[0.2,0.8],
// random rotation), each with its own Disney material drawn like config 1's.
int make_config3(ptgen_scene* s, uint64_t seed, const ptgen_config_opts& o) {
    add_config1_content(s, seed, o);
    int models[4];
    for (int i = 0; i < 4; ++i) {
        const uint32_t n = o.mesh_freq[i] ? o.mesh_freq[i] : 8;
        if (n > 255) return s->fail("geodesic frequency above 255");
        std::vector<float> pos, nrm, uv;
        std::vector<uint32_t> idx;
        geodesic_sphere(n, mix_seed(seed, 300 + i), 0.1, pos, nrm, uv, idx);
        models[i] = ptgen_add_model(s, pos.data(), nrm.data(), uv.data(), uint32_t(pos.size() / 3), idx.data(),
                                    uint32_t(idx.size() / 3));
        if (models[i] < 0) return models[i];
    }
    Rng r(mix_seed(seed, 3));
    const uint32_t n_inst = o.n_instances ? o.n_instances : 16;
    for (uint32_t i = 0; i < n_inst; ++i) {
        ptgen_material m = plain_material();
        for (int k = 0; k < 3; ++k) m.color[k] = f32(r.uniform());
        m.roughness = f32(r.uniform());
        m.metallic = r.uniform() < 0.3 ? 1.f : 0.f;
        const bool trans = r.uniform() < 0.15;
        const float ior = f32(r.uniform(1.3, 1.7));
        m.transmission = trans ? 1.f : 0.f;
        m.ior = trans ? ior : 1.5f;
        const int mi = ptgen_add_material(s, &m);
        float p[3], q[4];
        for (int k = 0; k < 3; ++k) p[k] = f32(r.uniform(-2, 2));
        random_quat(r, q);
        const float sc1 = f32(r.uniform(0.2, 0.8));
        const float sc[3] = {sc1, sc1, sc1};
        ptgen_add_instance(s, models[i % 4], mi, p, q, sc, 1);
    }
    return ptgen_finalize(s);
}

int make_config2(ptgen_scene* s, uint64_t seed, const ptgen_config_opts& o) {
    const uint32_t def_freq[4] = {64, 90, 110, 127};
    int models[4];
    for (int i = 0; i < 4; ++i) {
        const uint32_t n = o.mesh_freq[i] ? o.mesh_freq[i] : def_freq[i];
        if (n > 255) return s->fail("geodesic frequency above 255");
        std::vector<float> pos, nrm, uv;
        std::vector<uint32_t> idx;
        geodesic_sphere(n, mix_seed(seed, 100 + i), 0.1, pos, nrm, uv, idx);
        models[i] = ptgen_add_model(s, pos.data(), nrm.data(), uv.data(), uint32_t(pos.size() / 3), idx.data(),
                                    uint32_t(idx.size() / 3));
        if (models[i] < 0) return models[i];
    }
    Rng r(mix_seed(seed, 2));
    const uint32_t n_inst = o.n_instances ? o.n_instances : 64;
    for (uint32_t i = 0; i < n_inst; ++i) {
        ptgen_material m = plain_material();
        for (int k = 0; k < 3; ++k) m.color[k] = f32(r.uniform());
        m.metallic = f32(r.uniform());
        m.roughness = f32(r.uniform());
        m.ior = f32(r.uniform());
        m.specular_tint = f32(r.uniform());
        m.anisotropic = f32(r.uniform());
        m.anisotropic_rotation = f32(r.uniform());
        m.sheen = f32(r.uniform());
        m.sheen_tint = f32(r.uniform());
        m.clearcoat = f32(r.uniform());
        m.clearcoat_gloss = f32(r.uniform());
        m.clearcoat_ior = f32(r.uniform());
        m.transmission = f32(r.uniform());
        m.concentration = f32(r.uniform());
        m.subsurface = f32(r.uniform());
        const int mi = ptgen_add_material(s, &m);
        float p[3], q[4];
        // keep instance centres clear of the camera so it never starts inside a mesh
        do {
            for (int k = 0; k < 3; ++k) p[k] = f32(r.uniform(-8, 8));
        } while ((p[0] - s->eye[0]) * (p[0] - s->eye[0]) + (p[1] - s->eye[1]) * (p[1] - s->eye[1]) +
                     (p[2] - s->eye[2]) * (p[2] - s->eye[2]) < 2.0f * 2.0f);
        random_quat(r, q);
        const float sc[3] = {1.f, 1.f, 1.f};
        ptgen_add_instance(s, models[i % 4], mi, p, q, sc, 1);
    }
    ptgen_set_sky_procedural(s, o.sky_width ? o.sky_width : 2048, o.sky_height ? o.sky_height : 1024, seed, 1e5f);
    ptgen_set_bounces(s, o.bounces > 0 ? o.bounces : 8);
    return ptgen_finalize(s);
}

}  // namespace

extern "C" int ptgen_make_config(ptgen_scene** out, int config, uint64_t seed, const ptgen_config_opts* opts) {
    if (!out || config < 1 || config > 3) return PT_ERR_ARG;
    ptgen_config_opts o;
    std::memset(&o, 0, sizeof(o));
    if (opts) o = *opts;
    int rc = ptgen_new(out);
    if (rc) return rc;
    rc = config == 1 ? make_config1(*out, seed, o) : config == 2 ? make_config2(*out, seed, o) : make_config3(*out, seed, o);
    if (rc < 0) { ptgen_destroy(*out); *out = nullptr; }
    return rc < 0 ? rc : PT_OK;
}
