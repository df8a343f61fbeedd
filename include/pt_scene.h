/*
 * pt_scene.h — host-side scene builder (libptscene.so).
 *
 * A C++ restatement of the parts of tracer.js that turn a scene into the
 * buffers the shaders read: Material.getSchema (tracer.js:258-297), the
 * primitive transforms and bounds (tracer.js:299-910), ModelData.generateBVH
 * (binned SAH BLAS, tracer.js:1025-1118), SceneBVHBuilder (median-split
 * TLAS, tracer.js:1509-1619), Scene.getLights (tracer.js:1654-1711),
 * packSceneData (tracer.js:1819-1917), HDRTexture.updateData/buildHDRCDF
 * (tracer.js:76-87,135-204) and Camera.updateRays2 (tracer.js:1449-1498).
 *
 * It also generates the seeded synthetic scenes the measurements use.  The
 * builder owns every array it returns; pointers stay valid until the scene
 * is destroyed or modified.
 */
#ifndef PT_SCENE_H
#define PT_SCENE_H

#include "pt_abi.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ptgen_scene ptgen_scene;

/* The tracer.js Material constructor fields, before getSchema's transforms
 * (pow 2.2 on color/subsurface_tint, emittance * intensity).  Zero-valued
 * fields that tracer.js defaults with `||` are defaulted the same way:
 * ior, sheen_tint, clearcoat_gloss, clearcoat_ior, concentration. */
typedef struct ptgen_material {
    float color[3];
    float roughness;
    float metallic;
    float ior;
    float emittance[3];
    float emission_intensity;
    float subsurface;
    float subsurface_tint[3]; /* negative first component: copy color */
    float specular_tint;
    float anisotropic;
    float anisotropic_rotation;
    float sheen;
    float sheen_tint;
    float clearcoat;
    float clearcoat_gloss;
    float clearcoat_ior;
    float transmission;
    float concentration;
} ptgen_material;

int ptgen_new(ptgen_scene** out);
void ptgen_destroy(ptgen_scene* s);
const char* ptgen_last_error(const ptgen_scene* s);

/* Returns the material index (>= 0) or a negative error. */
int ptgen_add_material(ptgen_scene* s, const ptgen_material* m);

/* Primitives.  rotation is a unit quaternion (x, y, z, w) as glMatrix's
 * quat; NULL means identity.  enable_nee marks a light candidate (only
 * spheres and cubes become lights, tracer.js:1655-1669).  Each returns the
 * primitive's creation index. */
int ptgen_add_sphere(ptgen_scene* s, int material, const float pos[3], float radius,
                     int enable_nee);
int ptgen_add_cube(ptgen_scene* s, int material, const float pos[3], const float half[3],
                   const float rotation[4], int enable_nee);
/* Cone-cylinder with base radius `radius` at `pos`, height `height` along
 * the rotated local +Y, top radius = top_ratio * radius. */
int ptgen_add_cylinder(ptgen_scene* s, int material, const float pos[3], float radius,
                       float height, const float rotation[4], float top_ratio);
int ptgen_add_torus(ptgen_scene* s, int material, const float pos[3], const float rotation[4],
                    float outer_radius, float inner_radius);
int ptgen_add_plane(ptgen_scene* s, int material, const float normal[3], float d);

/* A triangle mesh (ModelData): positions/normals are 3 floats per vertex,
 * uvs 2 floats per vertex (NULL: zeros), indices 3 per triangle.  Builds the
 * SAH BLAS immediately.  Returns the model index. */
int ptgen_add_model(ptgen_scene* s, const float* positions, const float* normals,
                    const float* uvs, uint32_t n_vertices, const uint32_t* indices,
                    uint32_t n_triangles);
/* An instance of a model (a Model primitive).  in_scene_bvh = 0 makes it a
 * list mesh traced outside the TLAS. */
int ptgen_add_instance(ptgen_scene* s, int model, int material, const float pos[3],
                       const float rotation[4], const float scale[3], int in_scene_bvh);

/* Sky from linear RGB floats (W*H*3): converted with HDRTexture.updateData
 * (x exposure, clamp at 65519, f16).  rgb == NULL selects the 1x1 plain
 * colour (0.02, 0.03, 0.05), which disables env NEE (plainColor). */
int ptgen_set_sky(ptgen_scene* s, uint32_t width, uint32_t height, const float* rgb,
                  float exposure);
/* Procedural sky: vertical gradient plus a Gaussian sun lobe whose peak is
 * `sun_peak` times the mean sky radiance. */
int ptgen_set_sky_procedural(ptgen_scene* s, uint32_t width, uint32_t height,
                             uint64_t seed, float sun_peak);

/* Material textures: ptgen_add_texture appends one 512x512 rgba8 layer
 * (rows top first) and returns its index; ptgen_set_material_maps sets a
 * material's texture indices in Material.getSchema order (emissive, albedo,
 * normal, height, roughness, metallic; -1 = none) and its uv_scale (NULL
 * keeps (1, 1)).  A material with a height map sets PT_HAS_HEIGHTMAPS, as
 * packSceneData does (tracer.js:1825). */
int ptgen_add_texture(ptgen_scene* s, const uint8_t* rgba);
int ptgen_set_material_maps(ptgen_scene* s, int material, const int32_t maps[6], const float uv_scale[2]);
/* Material.height_params (tracer.js:247-250, 295): normalMultiplier,
 * heightMultiplier (> 0 depth map, < 0 height map), heightSamp (parallax
 * layers, > 0) and heightOffset; defaults 1, 0.05, 32, 0. */
int ptgen_set_height_params(ptgen_scene* s, int material, const float params[4]);

/* Asset ingest.  ptgen_parse_hdr reads a Radiance RGBE (.hdr) file as
 * lib/parse-hdr.js does (parse-hdr.js:20-211, as HDRTexture._load uses it,
 * tracer.js:53-73): the header's first 20 lines (FORMAT=32-bit_rle_rgbe
 * required, "-Y H +X W" size), new-style run-length scanlines or flat RGBE,
 * each channel decoded as byte / 255 * 2^(e - 128) in double and stored as
 * float, alpha 1, rows top first.  rgba == NULL asks for the size only;
 * otherwise rgba_count must be W * H * 4.  `exposure` receives the header's
 * EXPOSURE (the reference does not apply it; ptgen_set_sky's exposure is
 * HDRTexture's).
 * ptgen_obj_triangles / ptgen_obj_read read a Wavefront OBJ text as
 * ModelData.parseOBJ + generateTriangles do (tracer.js:949-1022): v, vt, vn
 * and f lines (v, v/vt, v/vt/vn, v//vn; 1-based or negative = relative
 * indices), faces fan-triangulated around their first corner, one output
 * vertex per corner (positions and normals 9 floats per triangle, uvs 6).  A
 * corner without a normal gets its face's geometric normal and one without
 * texture coordinates (0, 0), where the reference would read undefined (NaN).
 * ptgen_add_obj adds such a mesh as a model (ptgen_add_model) and returns its
 * index.  ptgen_asset_error describes the last failure on this thread. */
int ptgen_parse_hdr(const uint8_t* bytes, uint64_t n, uint32_t* width, uint32_t* height, float* exposure,
                    float* rgba, uint64_t rgba_count);
int ptgen_obj_triangles(const char* text, uint64_t n, uint32_t* n_triangles);
int ptgen_obj_read(const char* text, uint64_t n, float* positions, float* normals, float* uvs,
                   uint32_t n_triangles);
int ptgen_add_obj(ptgen_scene* s, const char* text, uint64_t n);
const char* ptgen_asset_error(void);

int ptgen_set_camera(ptgen_scene* s, const float eye[3], const float target[3], float fov_deg,
                     float exposure);
int ptgen_set_bounces(ptgen_scene* s, int bounces);

/* packSceneData + buildHDRCDF: builds lights, TLAS, packed arrays, flags. */
int ptgen_finalize(ptgen_scene* s);
/* Valid after ptgen_finalize. */
int ptgen_get_desc(const ptgen_scene* s, pt_scene_desc* out);
/* updateRays2 (aperture 0) + updateUniforms for a width x height frame. */
int ptgen_scene_params(const ptgen_scene* s, uint32_t width, uint32_t height,
                       uint32_t sample_number, uint32_t seed, pt_scene_params* out);

/* Seeded synthetic scenes.
 *  config 1: 6 spheres, 4 cubes, 1 cylinder, ground plane, emissive quad
 *            (radiance 15, NEE), 512x256 procedural sky (sun peak/mean 1e3).
 *  config 2: 4 displaced geodesic spheres of frequency freq[i] (20 n^2
 *            triangles), 64 instances, 2048x1024 sky (sun dynamic range 1e5).
 *  config 3: config 1 plus n_instances (16) instances of 4 displaced
 *            geodesic spheres (frequency freq[i], default 8) placed and
 *            scaled like config 1's primitives (the batched-jobs scenes).
 * opts: NULL for the defaults; otherwise overrides (0 keeps a default). */
typedef struct ptgen_config_opts {
    uint32_t sky_width;
    uint32_t sky_height;
    uint32_t mesh_freq[4];   /* config 2/3 geodesic frequencies (64, 90, 110, 127 / 8) */
    uint32_t n_instances;    /* config 2/3 instance count (64 / 16) */
    int32_t bounces;         /* BOUNCE_LIMIT (4 for config 1, 8 for config 2) */
} ptgen_config_opts;
int ptgen_make_config(ptgen_scene** out, int config, uint64_t seed, const ptgen_config_opts* opts);

/* Building blocks, exposed for tests. */
/* buildHDRCDF on rgba16f texels; cdf must hold (H+1)+H*(W+1) floats. */
int ptgen_build_sky_cdf(const uint16_t* rgba16f, uint32_t width, uint32_t height, float exposure,
                        float* cdf, double* total_weight);
/* Number of BLAS nodes / max depth of a model. */
int ptgen_model_info(const ptgen_scene* s, int model, uint32_t* n_nodes, uint32_t* n_tris,
                     uint32_t* max_depth);

#ifdef __cplusplus
} /* extern "C" */
#endif

#endif /* PT_SCENE_H */
