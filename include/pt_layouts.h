/*
 * pt_layouts.h — byte layouts of every buffer the tracer consumes.
 *
 * These are the std430-style records that tracer.js packs with
 * packDataFromSchema (tracer.js:1772-1817): little-endian, strides rounded
 * up to 16 bytes, matrices column-major (glMatrix Float32Array).  The WGSL
 * declarations they must match are cited per struct.  The C ABI in pt_abi.h
 * accepts arrays of exactly these records, so a caller holding tracer.js's
 * typed arrays can hand them over without repacking.
 *
 * Plain C99 so that the HIP library, the host scene builder, the oracle and
 * the Node-API addon can all include it.
 */
#ifndef PT_LAYOUTS_H
#define PT_LAYOUTS_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Material, 144 B — shader.wgsl:35-54, packed by Material.getSchema
 * (tracer.js:258-297).  color and subsurface_tint already carry pow(x,2.2);
 * emittance already carries the emissionIntensity factor. */
typedef struct pt_material {
    float color[3];
    float metallic;
    float roughness;
    float ior;
    float specular_tint;
    float anisotropic;
    float aniso_rotation;
    float sheen;
    float sheen_tint;
    float clearcoat;
    float clearcoat_gloss;
    float clearcoat_ior;
    float transmission;
    float concentration;
    float subsurface_tint[3];
    float subsurface;
    float emittance[3];
    int32_t emissive_idx;
    int32_t albedo_idx;
    int32_t normal_idx;
    int32_t height_idx;
    int32_t roughness_idx;
    int32_t metallic_idx;
    float _pad1;
    float uv_scale[2];
    float height_params[4]; /* normal mult, height mult, samples, offset */
} pt_material;

/* TransformedObject, 96 B — shader.wgsl:56-63; schemas tracer.js:478-489
 * (sphere), 525-536 (cube), 672-683 (cylinder), 774-785 (torus),
 * 1406-1417 (in-BVH model).  param0 is the cylinder top radius, the torus
 * inner radius, or (bit pattern) the mesh's BLAS node offset; param1 is
 * (bit pattern) the mesh's triangle offset. */
typedef struct pt_object {
    float inv_matrix[16];
    int32_t material_idx;
    int32_t light_idx;
    int32_t object_type; /* PT_OBJ_* */
    float param0;
    float world_position[3];
    float param1;
} pt_object;

enum {
    PT_OBJ_MESH = 0,
    PT_OBJ_SPHERE = 1,
    PT_OBJ_CUBE = 2,
    PT_OBJ_CYLINDER = 3,
    PT_OBJ_TORUS = 4
};

/* BVHNode, 32 B — shader.wgsl:102-107.  TLAS (tracer.js:1592-1618): leaf iff
 * count>0, next = first object index; inner: left = i+1, right = next.
 * BLAS (tracer.js:1099-1115): leaf next = first triangle relative to the
 * mesh's tri_offset; inner: left = i+1 (absolute), right = node_offset+next. */
typedef struct pt_bvh_node {
    float aabb_min[3];
    uint32_t count;
    float aabb_max[3];
    uint32_t next;
} pt_bvh_node;

/* Triangle, 128 B — shader.wgsl:84-92; packed at tracer.js:1867-1872. */
typedef struct pt_triangle {
    float v0[3]; float pad0;
    float v1[3]; float pad1;
    float v2[3]; float pad2;
    float n0[3]; float pad3;
    float n1[3]; float pad4;
    float n2[3]; float pad5;
    float uv0[2];
    float uv1[2];
    float uv2[2];
    float pad6[2];
} pt_triangle;

/* MeshInstance, 80 B — the WGSL declaration shader.wgsl:94-100 (list
 * meshes traced outside the TLAS).  tracer.js writes a 96 B Model record
 * into this binding (tracer.js:1890), which the shader cannot read
 * correctly; only the 80 B WGSL layout is accepted here. */
typedef struct pt_mesh_instance {
    float inv_matrix[16];
    int32_t material_idx;
    int32_t pad;
    uint32_t node_offset;
    uint32_t tri_offset;
} pt_mesh_instance;

/* Plane, 32 B — shader.wgsl:65-70; Plane.getSchema tracer.js:902-910. */
typedef struct pt_plane {
    float normal[3];
    float d;
    int32_t material_idx;
    float pad[3];
} pt_plane;

/* Light, 80 B — shader.wgsl:109-115; Light.getSchema tracer.js:1702-1711.
 * matrix is the object's forward matrix. */
typedef struct pt_light {
    int32_t obj_idx;
    float area;
    float power;
    float scale;
    float matrix[16];
} pt_light;

/* SceneParams uniform, 112 B — shader.wgsl:23-33; written by
 * Renderer.updateUniforms tracer.js:2207-2240. */
typedef struct pt_scene_params {
    float eye[3];
    uint32_t sample_number;
    float ray00[3];
    uint32_t width;
    float ray10[3];
    uint32_t height;
    float ray01[3];
    float exposure;
    float ray11[3];
    uint32_t seed;
    uint32_t sky_width;
    uint32_t sky_height;
    float sky_total_lum;
    float total_light_power; /* written by the host, never read by the shader */
    uint32_t section[2];
    uint32_t sky_cdf_index;
    float pad0;
} pt_scene_params;

/* GBufferData, 48 B — shader.wgsl:2497-2505, denoiser.wgsl:16-24. */
typedef struct pt_gbuffer_texel {
    float normal[3];
    float roughness;
    float albedo[3];
    float metallic;
    float motion[2];
    float transmission;
    float dist;
} pt_gbuffer_texel;

/* Denoiser uniforms, 48 B — denoiser.wgsl:1-14; Denoiser.updateUniforms
 * tracer.js:2560-2581. flags bit0 = median prepass, bit1 = demodulate. */
typedef struct pt_denoise_params {
    uint32_t width;
    uint32_t height;
    uint32_t step_width;
    uint32_t passes;
    float sigma_color;
    float sigma_normal;
    float sigma_depth;
    float sigma_albedo;
    float sigma_material;
    float glossy_protect;
    float exposure;
    uint32_t flags;
} pt_denoise_params;

/* Specialisation flags — the override constants shader.wgsl:2-12, derived
 * by packSceneData tracer.js:1901-1913. */
enum {
    PT_HAS_SPHERES = 1u << 0,
    PT_HAS_CUBES = 1u << 1,
    PT_HAS_PLANES = 1u << 2,
    PT_HAS_CYLINDERS = 1u << 3,
    PT_HAS_TORI = 1u << 4,
    PT_HAS_MESHES = 1u << 5,
    PT_HAS_LIST_MESHES = 1u << 6,
    PT_HAS_HEIGHTMAPS = 1u << 7,
    PT_HAS_LIGHTS = 1u << 8,
    PT_HAS_SKYBOX = 1u << 9,
    PT_MIS_SKYBOX = 1u << 10
};

#ifdef __cplusplus
} /* extern "C" */
#define PT_LAYOUT_ASSERT(c, m) static_assert(c, m)
#else
#define PT_LAYOUT_ASSERT(c, m) _Static_assert(c, m)
#endif

PT_LAYOUT_ASSERT(sizeof(pt_material) == 144, "Material is 144 B");
PT_LAYOUT_ASSERT(offsetof(pt_material, subsurface_tint) == 64, "Material.subsurface_tint @64");
PT_LAYOUT_ASSERT(offsetof(pt_material, emittance) == 80, "Material.emittance @80");
PT_LAYOUT_ASSERT(offsetof(pt_material, albedo_idx) == 96, "Material.albedo_idx @96");
PT_LAYOUT_ASSERT(offsetof(pt_material, uv_scale) == 120, "Material.uv_scale @120");
PT_LAYOUT_ASSERT(offsetof(pt_material, height_params) == 128, "Material.height_params @128");
PT_LAYOUT_ASSERT(sizeof(pt_object) == 96, "TransformedObject is 96 B");
PT_LAYOUT_ASSERT(offsetof(pt_object, material_idx) == 64, "object.material_idx @64");
PT_LAYOUT_ASSERT(offsetof(pt_object, param0) == 76, "object.param0 @76");
PT_LAYOUT_ASSERT(offsetof(pt_object, world_position) == 80, "object.world_position @80");
PT_LAYOUT_ASSERT(offsetof(pt_object, param1) == 92, "object.param1 @92");
PT_LAYOUT_ASSERT(sizeof(pt_bvh_node) == 32, "BVHNode is 32 B");
PT_LAYOUT_ASSERT(sizeof(pt_triangle) == 128, "Triangle is 128 B");
PT_LAYOUT_ASSERT(offsetof(pt_triangle, uv0) == 96, "Triangle.uv0 @96");
PT_LAYOUT_ASSERT(sizeof(pt_mesh_instance) == 80, "MeshInstance is 80 B");
PT_LAYOUT_ASSERT(sizeof(pt_plane) == 32, "Plane is 32 B");
PT_LAYOUT_ASSERT(sizeof(pt_light) == 80, "Light is 80 B");
PT_LAYOUT_ASSERT(sizeof(pt_scene_params) == 112, "SceneParams is 112 B");
PT_LAYOUT_ASSERT(offsetof(pt_scene_params, seed) == 76, "SceneParams.seed @76");
PT_LAYOUT_ASSERT(offsetof(pt_scene_params, sky_total_lum) == 88, "SceneParams.sky_total_lum @88");
PT_LAYOUT_ASSERT(offsetof(pt_scene_params, section) == 96, "SceneParams.section @96");
PT_LAYOUT_ASSERT(offsetof(pt_scene_params, sky_cdf_index) == 104, "SceneParams.sky_cdf_index @104");
PT_LAYOUT_ASSERT(sizeof(pt_gbuffer_texel) == 48, "GBufferData is 48 B");
PT_LAYOUT_ASSERT(offsetof(pt_gbuffer_texel, transmission) == 40, "GBufferData.transmission @40");
PT_LAYOUT_ASSERT(sizeof(pt_denoise_params) == 48, "denoiser uniforms are 48 B");

#endif /* PT_LAYOUTS_H */
