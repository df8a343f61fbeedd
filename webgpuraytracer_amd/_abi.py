"""ctypes mirror of include/pt_layouts.h, include/pt_abi.h and include/pt_scene.h,
and the loaders for the in-tree native libraries.

The structures here are byte-for-byte the C layouts (checked against the
headers' sizes by the tests), so numpy record arrays of these dtypes can be
handed to the C ABI without copies.
"""
import ctypes as C
import os

import numpy as np

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_DIR = os.path.join(PKG, "lib")

# ----------------------------------------------------------------- layouts --
f32, i32, u32, u64 = C.c_float, C.c_int32, C.c_uint32, C.c_uint64


class Material(C.Structure):
    _fields_ = [("color", f32 * 3), ("metallic", f32), ("roughness", f32), ("ior", f32),
                ("specular_tint", f32), ("anisotropic", f32), ("aniso_rotation", f32), ("sheen", f32),
                ("sheen_tint", f32), ("clearcoat", f32), ("clearcoat_gloss", f32), ("clearcoat_ior", f32),
                ("transmission", f32), ("concentration", f32), ("subsurface_tint", f32 * 3),
                ("subsurface", f32), ("emittance", f32 * 3), ("emissive_idx", i32), ("albedo_idx", i32),
                ("normal_idx", i32), ("height_idx", i32), ("roughness_idx", i32), ("metallic_idx", i32),
                ("_pad1", f32), ("uv_scale", f32 * 2), ("height_params", f32 * 4)]


class Object(C.Structure):
    _fields_ = [("inv_matrix", f32 * 16), ("material_idx", i32), ("light_idx", i32), ("object_type", i32),
                ("param0", f32), ("world_position", f32 * 3), ("param1", f32)]


class BVHNode(C.Structure):
    _fields_ = [("aabb_min", f32 * 3), ("count", u32), ("aabb_max", f32 * 3), ("next", u32)]


class Triangle(C.Structure):
    _fields_ = [("v0", f32 * 3), ("pad0", f32), ("v1", f32 * 3), ("pad1", f32), ("v2", f32 * 3), ("pad2", f32),
                ("n0", f32 * 3), ("pad3", f32), ("n1", f32 * 3), ("pad4", f32), ("n2", f32 * 3), ("pad5", f32),
                ("uv0", f32 * 2), ("uv1", f32 * 2), ("uv2", f32 * 2), ("pad6", f32 * 2)]


class MeshInstance(C.Structure):
    _fields_ = [("inv_matrix", f32 * 16), ("material_idx", i32), ("pad", i32), ("node_offset", u32),
                ("tri_offset", u32)]


class Plane(C.Structure):
    _fields_ = [("normal", f32 * 3), ("d", f32), ("material_idx", i32), ("pad", f32 * 3)]


class Light(C.Structure):
    _fields_ = [("obj_idx", i32), ("area", f32), ("power", f32), ("scale", f32), ("matrix", f32 * 16)]


class SceneParams(C.Structure):
    _fields_ = [("eye", f32 * 3), ("sample_number", u32), ("ray00", f32 * 3), ("width", u32),
                ("ray10", f32 * 3), ("height", u32), ("ray01", f32 * 3), ("exposure", f32),
                ("ray11", f32 * 3), ("seed", u32), ("sky_width", u32), ("sky_height", u32),
                ("sky_total_lum", f32), ("total_light_power", f32), ("section", u32 * 2),
                ("sky_cdf_index", u32), ("pad0", f32)]


class GBufferTexel(C.Structure):
    _fields_ = [("normal", f32 * 3), ("roughness", f32), ("albedo", f32 * 3), ("metallic", f32),
                ("motion", f32 * 2), ("transmission", f32), ("dist", f32)]


class DenoiseParams(C.Structure):
    _fields_ = [("width", u32), ("height", u32), ("step_width", u32), ("passes", u32), ("sigma_color", f32),
                ("sigma_normal", f32), ("sigma_depth", f32), ("sigma_albedo", f32), ("sigma_material", f32),
                ("glossy_protect", f32), ("exposure", f32), ("flags", u32)]


class Array(C.Structure):
    _fields_ = [("ptr", C.c_void_p), ("count", u64), ("stride", u32), ("_pad", u32)]


class SceneDesc(C.Structure):
    _fields_ = [("materials", Array), ("objects", Array), ("tlas", Array), ("meshes", Array), ("bvh", Array),
                ("triangles", Array), ("planes", Array), ("lights", Array), ("sky_width", u32),
                ("sky_height", u32), ("sky_rgba16f", C.c_void_p), ("sky_cdf", C.c_void_p),
                ("sky_cdf_count", u64), ("sky_exposure", f32), ("sky_total_lum", f32),
                ("textures", C.c_void_p), ("texture_layers", u32), ("flags", u32), ("bounce_limit", i32),
                ("_pad", u32)]


class Stats(C.Structure):
    _fields_ = [("extension_rays", u64), ("shadow_rays", u64), ("paths", u64), ("node_pops", u64),
                ("inner_pops", u64), ("object_visits", u64), ("tri_tests", u64), ("shade_events", u64),
                ("kernel_ms", C.c_double), ("launches", u32), ("_pad", u32)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_ if not k.startswith("_")}


class Hit(C.Structure):
    _fields_ = [("t", f32), ("o_idx", i32), ("m_idx", i32), ("prim_id", i32), ("plane_idx", i32),
                ("count", u32), ("n", f32 * 3), ("uv", f32 * 2), ("_pad", f32 * 3)]


class GenMaterial(C.Structure):
    _fields_ = [("color", f32 * 3), ("roughness", f32), ("metallic", f32), ("ior", f32), ("emittance", f32 * 3),
                ("emission_intensity", f32), ("subsurface", f32), ("subsurface_tint", f32 * 3),
                ("specular_tint", f32), ("anisotropic", f32), ("anisotropic_rotation", f32), ("sheen", f32),
                ("sheen_tint", f32), ("clearcoat", f32), ("clearcoat_gloss", f32), ("clearcoat_ior", f32),
                ("transmission", f32), ("concentration", f32)]


class ConfigOpts(C.Structure):
    _fields_ = [("sky_width", u32), ("sky_height", u32), ("mesh_freq", u32 * 4), ("n_instances", u32),
                ("bounces", i32)]


SIZES = {"pt_material": (Material, 144), "pt_object": (Object, 96), "pt_bvh_node": (BVHNode, 32),
         "pt_triangle": (Triangle, 128), "pt_mesh_instance": (MeshInstance, 80), "pt_plane": (Plane, 32),
         "pt_light": (Light, 80), "pt_scene_params": (SceneParams, 112), "pt_gbuffer_texel": (GBufferTexel, 48),
         "pt_denoise_params": (DenoiseParams, 48), "pt_hit": (Hit, 56)}

# flags (pt_layouts.h)
HAS_SPHERES, HAS_CUBES, HAS_PLANES, HAS_CYLINDERS, HAS_TORI, HAS_MESHES, HAS_LIST_MESHES, HAS_HEIGHTMAPS, \
    HAS_LIGHTS, HAS_SKYBOX, MIS_SKYBOX = [1 << i for i in range(11)]

PT_OK, PT_ERR_ARG, PT_ERR_LAYOUT, PT_ERR_CAPACITY, PT_ERR_FLAGS, PT_ERR_HIP, PT_ERR_OOM, PT_ERR_STATE, \
    PT_ERR_UNSUPPORTED, PT_ERR_NO_DEVICE = 0, -1, -2, -3, -4, -5, -6, -7, -8, -9
ERR_NAMES = {0: "PT_OK", -1: "PT_ERR_ARG", -2: "PT_ERR_LAYOUT", -3: "PT_ERR_CAPACITY", -4: "PT_ERR_FLAGS",
             -5: "PT_ERR_HIP", -6: "PT_ERR_OOM", -7: "PT_ERR_STATE", -8: "PT_ERR_UNSUPPORTED",
             -9: "PT_ERR_NO_DEVICE"}

PT_OPT_BACKEND, PT_OPT_COUNT_STATS, PT_OPT_WAVEFRONT_POOL, PT_OPT_SCHEDULE, PT_OPT_TRAVERSAL, PT_OPT_LDS_STACK, \
    PT_OPT_WAVEFRONT_REFILL, PT_OPT_BLAS_BUILD, PT_OPT_TLAS_BUILD = 1, 2, 3, 4, 5, 6, 7, 8, 9
PT_OPT_KERNEL_FLAGS, PT_OPT_SPECIALISED, PT_OPT_DENOISE_MATH, PT_OPT_BOUNCE_TIMING = 10, 11, 12, 13
PT_DENOISE_MATH_FUSED, PT_DENOISE_MATH_EXACT = 0, 1
KERNEL_TEXTURES_BIT, KERNEL_TRANSLUCENT_BIT = 1 << 20, 1 << 21
PT_BLAS_BUILD_REFERENCE, PT_BLAS_BUILD_DEVICE = 0, 1
PT_TLAS_BUILD_REFERENCE, PT_TLAS_BUILD_DEVICE = 0, 1
PT_BACKEND_MEGAKERNEL, PT_BACKEND_WAVEFRONT = 0, 1
PT_UNET_INPUT, PT_UNET_TARGET = 0, 1
PT_OBJ_MESH, PT_OBJ_SPHERE, PT_OBJ_CUBE, PT_OBJ_CYLINDER, PT_OBJ_TORUS = 0, 1, 2, 3, 4
PT_UNET_BF16, PT_UNET_FP8 = 0, 1
PT_SCHEDULE_SYNC, PT_SCHEDULE_REGEN = 0, 1
PT_TRAVERSAL_REFERENCE, PT_TRAVERSAL_WIDE, PT_TRAVERSAL_WIDE_QUANTIZED = 0, 1, 2


def dtype_of(struct):
    """numpy dtype with the C structure's exact size (for zero-copy views)."""
    return np.dtype((np.void, C.sizeof(struct)))


# ----------------------------------------------------------------- loading --
P = C.POINTER
vp = C.c_void_p

PT_FUNCS = {
    "pt_create": (C.c_int, [P(vp), C.c_int]),
    "pt_destroy": (None, [vp]),
    "pt_last_error": (C.c_char_p, [vp]),
    "pt_set_option": (C.c_int, [vp, C.c_int, C.c_int64]),
    "pt_upload_scene": (C.c_int, [vp, P(SceneDesc)]),
    "pt_update_objects": (C.c_int, [vp, P(SceneDesc)]),
    "pt_resize": (C.c_int, [vp, u32, u32, C.c_int]),
    "pt_trace": (C.c_int, [vp, P(SceneParams), u32, u32, u32, P(u32), P(Stats)]),
    "pt_trace_hits": (C.c_int, [vp, P(SceneParams), u32, P(Hit)]),
    "pt_get_option": (C.c_int, [vp, C.c_int, P(C.c_int64)]),
    "pt_trace_noalloc": (C.c_int, [vp, P(SceneParams), u32, u32, u32, P(u32), P(Stats)]),
    "pt_get_bounce_stats": (C.c_int, [vp, vp, vp, vp, u32]),
    "pt_trace_hits_rays": (C.c_int, [vp, P(SceneParams), u32, P(Hit), vp]),
    "pt_gbuffer": (C.c_int, [vp, P(SceneParams)]),
    "pt_denoise": (C.c_int, [vp, P(DenoiseParams), C.c_int]),
    "pt_read_accum": (C.c_int, [vp, vp, u64]),
    "pt_write_accum": (C.c_int, [vp, vp, u64]),
    "pt_read_output": (C.c_int, [vp, vp, u64]),
    "pt_read_gbuffer": (C.c_int, [vp, vp, u64]),
    "pt_write_gbuffer": (C.c_int, [vp, vp, u64]),
    "pt_read_denoised": (C.c_int, [vp, vp, u64]),
    "pt_read_sky_cdf": (C.c_int, [vp, vp, u64, P(f32)]),
    "pt_trace_rays": (C.c_int, [vp, vp, u32, vp]),
    "pt_trace_shadow_rays": (C.c_int, [vp, vp, vp, vp, u32, vp]),
    "pt_synchronize": (C.c_int, [vp]),
    # the learned denoiser (pt_unet.h)
    "pt_unet_init": (C.c_int, [vp, u64]),
    "pt_unet_param_count": (C.c_int, [vp, P(u64)]),
    "pt_unet_get_params": (C.c_int, [vp, vp, u64]),
    "pt_unet_set_params": (C.c_int, [vp, vp, u64]),
    "pt_unet_get_adam": (C.c_int, [vp, vp, vp, u64, P(u32)]),
    "pt_unet_set_adam": (C.c_int, [vp, vp, vp, u64, u32]),
    "pt_unet_forward": (C.c_int, [vp, vp, u32, u32, u32, vp]),
    "pt_unet_loss_grad": (C.c_int, [vp, vp, vp, u32, u32, u32, P(f32), vp]),
    "pt_unet_set_dataset": (C.c_int, [vp, vp, vp, u32, u32, u32]),
    "pt_unet_dataset_reserve": (C.c_int, [vp, u32, u32, u32]),
    "pt_unet_dataset_capture": (C.c_int, [vp, u32, C.c_int, u32, u32]),
    "pt_unet_train": (C.c_int, [vp, u32, u32, f32, u64, vp]),
    "pt_unet_denoise": (C.c_int, [vp, f32]),
    "pt_unet_set_precision": (C.c_int, [vp, C.c_int]),
    "pt_unet_fp8_state": (C.c_int, [vp, vp, u64, vp, u64, vp, u64]),
    # many jobs per graph launch (pt_batch.h)
    "pt_batch_create": (C.c_int, [P(vp), C.c_int]),
    "pt_batch_destroy": (None, [vp]),
    "pt_batch_last_error": (C.c_char_p, [vp]),
    "pt_batch_add": (C.c_int, [vp, vp, P(u32)]),
    "pt_batch_size": (C.c_int, [vp, P(u32)]),
    "pt_batch_frame": (C.c_int, [vp, vp, P(u32), vp, vp]),
    "pt_batch_graph_info": (C.c_int, [vp, P(u32), P(u32), P(u32)]),
}

PTGEN_FUNCS = {
    "ptgen_new": (C.c_int, [P(vp)]),
    "ptgen_destroy": (None, [vp]),
    "ptgen_last_error": (C.c_char_p, [vp]),
    "ptgen_add_material": (C.c_int, [vp, P(GenMaterial)]),
    "ptgen_add_sphere": (C.c_int, [vp, C.c_int, P(f32), f32, C.c_int]),
    "ptgen_add_cube": (C.c_int, [vp, C.c_int, P(f32), P(f32), P(f32), C.c_int]),
    "ptgen_add_cylinder": (C.c_int, [vp, C.c_int, P(f32), f32, f32, P(f32), f32]),
    "ptgen_add_torus": (C.c_int, [vp, C.c_int, P(f32), P(f32), f32, f32]),
    "ptgen_add_plane": (C.c_int, [vp, C.c_int, P(f32), f32]),
    "ptgen_add_model": (C.c_int, [vp, vp, vp, vp, u32, vp, u32]),
    "ptgen_add_texture": (C.c_int, [vp, vp]),
    "ptgen_set_material_maps": (C.c_int, [vp, C.c_int, vp, vp]),
    "ptgen_set_height_params": (C.c_int, [vp, C.c_int, vp]),
    "ptgen_parse_hdr": (C.c_int, [vp, u64, P(u32), P(u32), P(f32), vp, u64]),
    "ptgen_obj_triangles": (C.c_int, [C.c_char_p, u64, P(u32)]),
    "ptgen_obj_read": (C.c_int, [C.c_char_p, u64, vp, vp, vp, u32]),
    "ptgen_add_obj": (C.c_int, [vp, C.c_char_p, u64]),
    "ptgen_asset_error": (C.c_char_p, []),
    "ptgen_add_instance": (C.c_int, [vp, C.c_int, C.c_int, P(f32), P(f32), P(f32), C.c_int]),
    "ptgen_set_sky": (C.c_int, [vp, u32, u32, vp, f32]),
    "ptgen_set_sky_procedural": (C.c_int, [vp, u32, u32, u64, f32]),
    "ptgen_set_camera": (C.c_int, [vp, P(f32), P(f32), f32, f32]),
    "ptgen_set_bounces": (C.c_int, [vp, C.c_int]),
    "ptgen_finalize": (C.c_int, [vp]),
    "ptgen_get_desc": (C.c_int, [vp, P(SceneDesc)]),
    "ptgen_scene_params": (C.c_int, [vp, u32, u32, u32, u32, P(SceneParams)]),
    "ptgen_make_config": (C.c_int, [P(vp), C.c_int, u64, P(ConfigOpts)]),
    "ptgen_build_sky_cdf": (C.c_int, [vp, u32, u32, f32, vp, P(C.c_double)]),
    "ptgen_model_info": (C.c_int, [vp, C.c_int, P(u32), P(u32), P(u32)]),
}


def _bind(lib, table):
    for name, (res, args) in table.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


_LIBS = {}


def load_pt():
    """libpt.so — raises if it has not been built (no fallback path exists)."""
    if "pt" not in _LIBS:
        path = os.path.join(LIB_DIR, "libpt.so")
        if not os.path.exists(path):
            raise RuntimeError("libpt.so is not built: run `python -m webgpuraytracer_amd.build`")
        _LIBS["pt"] = _bind(C.CDLL(path), PT_FUNCS)
    return _LIBS["pt"]


def load_scene_lib():
    if "scene" not in _LIBS:
        path = os.path.join(LIB_DIR, "libptscene.so")
        if not os.path.exists(path):
            raise RuntimeError("libptscene.so is not built: run `python -m webgpuraytracer_amd.build`")
        _LIBS["scene"] = _bind(C.CDLL(path), PTGEN_FUNCS)
    return _LIBS["scene"]


class PTError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("%s (%d): %s" % (ERR_NAMES.get(code, "?"), code, msg))
        self.code = code


def fvec(*vals):
    return (f32 * len(vals))(*vals)
