"""Host-side scene builder: Python face of libptscene (include/pt_scene.h).

Mirrors the scene-construction API of tracer.js (Scene.newSphere / newCube /
newCylinder / newTorus / newPlane / newModel, Material, HDRTexture, Camera)
and produces the packed buffers Renderer.setScene uploads.
"""
import ctypes as C
import os

import numpy as np

from . import _abi as A


def _mat_struct(color=(1.0, 1.0, 1.0), roughness=0.5, metallic=0.0, ior=1.5, emittance=(1.0, 1.0, 1.0),
                emission_intensity=0.0, subsurface=0.0, subsurface_tint=None, specular_tint=0.0,
                anisotropic=0.0, anisotropic_rotation=0.0, sheen=0.0, sheen_tint=0.5, clearcoat=0.0,
                clearcoat_gloss=1.0, clearcoat_ior=1.5, transmission=0.0, concentration=1.0):
    m = A.GenMaterial()
    m.color = A.fvec(*color)
    m.roughness, m.metallic, m.ior = roughness, metallic, ior
    m.emittance = A.fvec(*emittance)
    m.emission_intensity = emission_intensity
    m.subsurface = subsurface
    m.subsurface_tint = A.fvec(*(subsurface_tint if subsurface_tint is not None else (-1.0, 0.0, 0.0)))
    m.specular_tint, m.anisotropic, m.anisotropic_rotation = specular_tint, anisotropic, anisotropic_rotation
    m.sheen, m.sheen_tint, m.clearcoat, m.clearcoat_gloss = sheen, sheen_tint, clearcoat, clearcoat_gloss
    m.clearcoat_ior, m.transmission, m.concentration = clearcoat_ior, transmission, concentration
    return m


class SceneBuilder:
    """A scene under construction; finalize() packs it (packSceneData)."""

    def __init__(self, _handle=None):
        self.lib = A.load_scene_lib()
        self.h = C.c_void_p(_handle) if _handle else C.c_void_p()
        if not _handle:
            self._check(self.lib.ptgen_new(C.byref(self.h)))
        self._keep = []

    @classmethod
    def config(cls, config, seed=1, sky=None, mesh_freq=None, n_instances=0, bounces=0):
        """The seeded synthetic scenes (1: small analytic, 2: large instanced
        meshes, 3: small analytic plus instanced meshes)."""
        lib = A.load_scene_lib()
        o = A.ConfigOpts()
        if sky:
            o.sky_width, o.sky_height = sky
        if mesh_freq:
            o.mesh_freq = (A.u32 * 4)(*mesh_freq)
        o.n_instances = n_instances
        o.bounces = bounces
        h = C.c_void_p()
        rc = lib.ptgen_make_config(C.byref(h), config, seed, C.byref(o))
        if rc:
            raise A.PTError(rc, "ptgen_make_config(%d) failed" % config)
        return cls(_handle=h.value)

    def _check(self, rc):
        if rc < 0:
            raise A.PTError(rc, (self.lib.ptgen_last_error(self.h) or b"").decode())
        return rc

    def close(self):
        if self.h:
            self.lib.ptgen_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- construction (tracer.js Scene.new*) --
    def material(self, **kw):
        return self._check(self.lib.ptgen_add_material(self.h, C.byref(_mat_struct(**kw))))

    def sphere(self, material, pos, radius, enable_nee=False):
        return self._check(self.lib.ptgen_add_sphere(self.h, material, A.fvec(*pos), radius, int(enable_nee)))

    def cube(self, material, pos, half, rotation=None, enable_nee=False):
        rot = A.fvec(*rotation) if rotation is not None else None
        return self._check(self.lib.ptgen_add_cube(self.h, material, A.fvec(*pos), A.fvec(*half), rot,
                                                   int(enable_nee)))

    def cylinder(self, material, pos, radius, height, rotation=None, top_ratio=1.0):
        rot = A.fvec(*rotation) if rotation is not None else None
        return self._check(self.lib.ptgen_add_cylinder(self.h, material, A.fvec(*pos), radius, height, rot,
                                                       top_ratio))

    def torus(self, material, pos, outer, inner, rotation=None):
        rot = A.fvec(*rotation) if rotation is not None else None
        return self._check(self.lib.ptgen_add_torus(self.h, material, A.fvec(*pos), rot, outer, inner))

    def plane(self, material, normal, d):
        return self._check(self.lib.ptgen_add_plane(self.h, material, A.fvec(*normal), d))

    def model(self, positions, normals, indices, uvs=None):
        p = np.ascontiguousarray(positions, np.float32).reshape(-1, 3)
        n = np.ascontiguousarray(normals, np.float32).reshape(-1, 3)
        i = np.ascontiguousarray(indices, np.uint32).reshape(-1, 3)
        u = None if uvs is None else np.ascontiguousarray(uvs, np.float32).reshape(-1, 2)
        return self._check(self.lib.ptgen_add_model(self.h, p.ctypes.data, n.ctypes.data,
                                                    None if u is None else u.ctypes.data, p.shape[0],
                                                    i.ctypes.data, i.shape[0]))

    def instance(self, model, material, pos, rotation=None, scale=(1.0, 1.0, 1.0), in_scene_bvh=True):
        rot = A.fvec(*rotation) if rotation is not None else None
        return self._check(self.lib.ptgen_add_instance(self.h, model, material, A.fvec(*pos), rot, A.fvec(*scale),
                                                       int(in_scene_bvh)))

    def sky(self, rgb=None, exposure=1.0):
        if rgb is None:
            return self._check(self.lib.ptgen_set_sky(self.h, 1, 1, None, exposure))
        a = np.ascontiguousarray(rgb, np.float32)
        h, w = a.shape[:2]
        self._keep.append(a)
        return self._check(self.lib.ptgen_set_sky(self.h, w, h, a.ctypes.data, exposure))

    def texture(self, rgba):
        """Adds a 512x512 rgba8 layer to the material texture array; returns its index."""
        a = np.ascontiguousarray(rgba, np.uint8)
        if a.shape != (512, 512, 4):
            raise ValueError("a texture layer is 512 x 512 x 4 bytes")
        return self._check(self.lib.ptgen_add_texture(self.h, a.ctypes.data))

    def material_maps(self, material, albedo=-1, roughness=-1, metallic=-1, emissive=-1, normal=-1, height=-1,
                      uv_scale=None):
        """Texture indices of a material (Material.albedoTex etc., tracer.js:234-246)."""
        maps = (C.c_int32 * 6)(emissive, albedo, normal, height, roughness, metallic)
        uv = A.fvec(*uv_scale) if uv_scale is not None else None
        return self._check(self.lib.ptgen_set_material_maps(self.h, material, maps, uv))

    def height_params(self, material, normal_multiplier=1.0, height_multiplier=0.05, layers=32.0, offset=0.0):
        """Material.height_params (tracer.js:247-250): the normal map's x/y
        scale, the parallax depth (negative: a height map, white high), the
        number of parallax layers and the start offset."""
        p = (C.c_float * 4)(normal_multiplier, height_multiplier, layers, offset)
        return self._check(self.lib.ptgen_set_height_params(self.h, material, p))

    def sky_from_hdr(self, source, exposure=1.0):
        """The sky from a Radiance .hdr file (a path or its bytes), as
        HDRTexture loads one (tracer.js:53-87)."""
        rgba, _ = read_hdr(source)
        return self.sky(rgba[..., :3], exposure)

    def model_from_obj(self, source):
        """A model from a Wavefront OBJ text (a path, str or bytes), as
        ModelData.parseOBJ reads one (tracer.js:949-1022); returns its index."""
        text = _obj_bytes(source)
        return self._check(self.lib.ptgen_add_obj(self.h, text, len(text)))

    def procedural_sky(self, width, height, seed=1, sun_peak=1e3):
        return self._check(self.lib.ptgen_set_sky_procedural(self.h, width, height, seed, sun_peak))

    def camera(self, eye=(0.0, 1.5, 4.5), target=(0.0, 0.5, 0.0), fov=45.0, exposure=1.0):
        return self._check(self.lib.ptgen_set_camera(self.h, A.fvec(*eye), A.fvec(*target), fov, exposure))

    def bounces(self, n):
        return self._check(self.lib.ptgen_set_bounces(self.h, n))

    def finalize(self):
        self._check(self.lib.ptgen_finalize(self.h))
        return self

    # -- packed output --
    def desc(self):
        d = A.SceneDesc()
        self._check(self.lib.ptgen_get_desc(self.h, C.byref(d)))
        return d

    def params(self, width, height, sample_number=0, seed=0):
        p = A.SceneParams()
        self._check(self.lib.ptgen_scene_params(self.h, width, height, sample_number, seed, C.byref(p)))
        return p

    def model_info(self, model):
        n, t, d = A.u32(), A.u32(), A.u32()
        self._check(self.lib.ptgen_model_info(self.h, model, C.byref(n), C.byref(t), C.byref(d)))
        return {"nodes": n.value, "triangles": t.value, "max_depth": d.value}

    def arrays(self):
        """numpy views (copies) of every packed array, keyed like pt_scene_desc."""
        d = self.desc()
        out = {}
        for name, struct in (("materials", A.Material), ("objects", A.Object), ("tlas", A.BVHNode),
                             ("meshes", A.MeshInstance), ("bvh", A.BVHNode), ("triangles", A.Triangle),
                             ("planes", A.Plane), ("lights", A.Light)):
            arr = getattr(d, name)
            buf = (struct * arr.count).from_address(arr.ptr)
            out[name] = np.ctypeslib.as_array(buf).copy() if arr.count else np.zeros(0)
            out[name + "_raw"] = bytes(C.string_at(arr.ptr, arr.count * arr.stride))
        out["sky"] = np.ctypeslib.as_array((C.c_uint16 * (d.sky_width * d.sky_height * 4)).from_address(
            d.sky_rgba16f)).copy().reshape(d.sky_height, d.sky_width, 4)
        out["sky_cdf"] = np.ctypeslib.as_array((C.c_float * d.sky_cdf_count).from_address(d.sky_cdf)).copy()
        out["flags"] = d.flags
        out["bounce_limit"] = d.bounce_limit
        out["sky_total_lum"] = d.sky_total_lum
        return out


def build_sky_cdf(rgba16f, exposure=1.0):
    """HDRTexture.buildHDRCDF on f16 texels (H, W, 4) -> (cdf float32, total weight)."""
    lib = A.load_scene_lib()
    t = np.ascontiguousarray(rgba16f, np.uint16)
    h, w = t.shape[:2]
    cdf = np.zeros(h + 1 + h * (w + 1), np.float32)
    tot = C.c_double()
    rc = lib.ptgen_build_sky_cdf(t.ctypes.data, w, h, exposure, cdf.ctypes.data, C.byref(tot))
    if rc:
        raise A.PTError(rc, "ptgen_build_sky_cdf failed")
    return cdf, tot.value


def _asset_error(lib, rc, what):
    raise A.PTError(rc, "%s: %s" % (what, (lib.ptgen_asset_error() or b"").decode()))


def read_hdr(source):
    """A Radiance RGBE image (a path or its bytes) as parse-hdr.js decodes it:
    (float32 [height][width][4] with alpha 1, rows top first; the header's
    EXPOSURE, which the reference does not apply)."""
    data = source if isinstance(source, (bytes, bytearray)) else open(source, "rb").read()
    lib = A.load_scene_lib()
    buf = (C.c_uint8 * len(data)).from_buffer_copy(data)
    w, h, e = A.u32(), A.u32(), A.f32()
    rc = lib.ptgen_parse_hdr(buf, len(data), C.byref(w), C.byref(h), C.byref(e), None, 0)
    if rc:
        _asset_error(lib, rc, "hdr header")
    out = np.empty((h.value, w.value, 4), np.float32)
    rc = lib.ptgen_parse_hdr(buf, len(data), C.byref(w), C.byref(h), C.byref(e), out.ctypes.data, out.size)
    if rc:
        _asset_error(lib, rc, "hdr pixels")
    return out, float(e.value)


def _obj_bytes(source):
    if isinstance(source, (bytes, bytearray)):
        return bytes(source)
    if isinstance(source, str) and ("\n" in source or not os.path.exists(source)):
        return source.encode()
    return open(source, "rb").read()


def read_obj(source):
    """A Wavefront OBJ mesh (a path, str or bytes) as parseOBJ +
    generateTriangles read it, one vertex per triangle corner:
    (positions [3T][3], normals [3T][3], uvs [3T][2]) float32."""
    text = _obj_bytes(source)
    lib = A.load_scene_lib()
    n = A.u32()
    rc = lib.ptgen_obj_triangles(text, len(text), C.byref(n))
    if rc:
        _asset_error(lib, rc, "obj")
    t = n.value
    pos, nrm, uv = (np.empty((3 * t, 3), np.float32), np.empty((3 * t, 3), np.float32),
                    np.empty((3 * t, 2), np.float32))
    rc = lib.ptgen_obj_read(text, len(text), pos.ctypes.data, nrm.ctypes.data, uv.ctypes.data, t)
    if rc:
        _asset_error(lib, rc, "obj")
    return pos, nrm, uv
