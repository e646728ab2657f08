"""ctypes mirror of include/tinypt_gpu.h (the C ABI of the MI355X backend).

The structs here must match the header field for field; tests/test_abi.py checks sizes and
that the built library exports every declared entry point.
"""
import ctypes as C

TPT_OK = 0
TPT_ERR_INVALID_ARGUMENT = -1
TPT_ERR_DEVICE = -2
TPT_ERR_OUT_OF_MEMORY = -3
TPT_ERR_UNSUPPORTED = -4
TPT_ERR_INTERNAL = -5

# material.rs:30-43
MAT_LAMBERT = 0
MAT_METAL = 1
MAT_DIELECTRIC = 2
MAT_GGX = 3
MAT_SUBSURFACE = 4
MAT_DIFFUSE_LIGHT = 5

ENV_SKY = 0
ENV_IMAGE = 1

FP32_WAVEFRONT = 0
FP64_PLAIN = 1

GRAD_PER_MATERIAL = 4

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int32)
_lp = C.POINTER(C.c_int64)


class CameraDesc(C.Structure):
    _fields_ = [
        ("eye", C.c_double * 3),
        ("target", C.c_double * 3),
        ("up", C.c_double * 3),
        ("vfov_deg", C.c_double),
        ("aspect", C.c_double),
        ("focus_dist", C.c_double),
        ("aperture", C.c_double),
    ]


class SceneDesc(C.Structure):
    _fields_ = [
        ("camera", CameraDesc),
        ("n_materials", C.c_int32),
        ("mat_kind", _ip),
        ("mat_param", _dp),
        ("n_spheres", C.c_int32),
        ("sphere_center", _dp),
        ("sphere_radius", _dp),
        ("sphere_mat", _ip),
        ("n_meshes", C.c_int32),
        ("mesh_tri_begin", _lp),
        ("tri_vertices", _dp),
        ("tri_mat", _ip),
        ("n_instances", C.c_int32),
        ("inst_mesh", _ip),
        ("inst_xform", _dp),
        ("inst_mat_override", _ip),
        ("env_kind", C.c_int32),
        ("env_width", C.c_int32),
        ("env_height", C.c_int32),
        ("env_rgb", _dp),
    ]


class RenderConfigC(C.Structure):
    _fields_ = [
        ("width", C.c_int32),
        ("height", C.c_int32),
        ("spp", C.c_int32),
        ("max_bounces", C.c_int32),
        ("rr_start", C.c_int32),
        ("adaptive_enabled", C.c_int32),
        ("adaptive_min_spp", C.c_int32),
        ("precision", C.c_int32),
        ("adaptive_threshold", C.c_double),
        ("seed", C.c_uint64),
        ("samples_per_pass", C.c_int32),
        ("traversal_refill", C.c_int32),
    ]


class Stats(C.Structure):
    _fields_ = [
        ("paths", C.c_uint64),
        ("rays_extension", C.c_uint64),
        ("rays_shadow", C.c_uint64),
        ("elapsed_ms", C.c_double),
        ("trace_kernel_ms", C.c_double),
        ("trace_launches", C.c_uint64),
        ("trace_rays", C.c_uint64),
        ("light_samples", C.c_uint64),
        ("light_self_occluded", C.c_uint64),
    ]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


class PassPlan(C.Structure):
    _fields_ = [
        ("paths_alloc", C.c_uint64),
        ("samples_per_pass", C.c_uint32),
        ("pixels_per_pass", C.c_uint32),
        ("path_sets", C.c_uint32),
        ("reserved", C.c_uint32),
    ]


class SceneInfo(C.Structure):
    _fields_ = [
        ("n_tris", C.c_int64),
        ("n_blas_nodes", C.c_int64),
        ("n_tlas_nodes", C.c_int64),
        ("max_blas_depth", C.c_int32),
        ("max_tlas_depth", C.c_int32),
        ("n_lights", C.c_int32),
        ("traversal_stack", C.c_int32),
        ("device_bytes", C.c_int64),
    ]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


# Every entry point include/tinypt_gpu.h declares: name -> (restype, argtypes)
_vp = C.c_void_p
_fp = C.POINTER(C.c_float)
ENTRY_POINTS = {
    "tpt_abi_version": (C.c_int, []),
    "tpt_last_error": (C.c_char_p, []),
    "tpt_ctx_create": (C.c_int, [C.c_int, C.POINTER(_vp)]),
    "tpt_ctx_destroy": (C.c_int, [_vp]),
    "tpt_scene_upload": (C.c_int, [_vp, C.POINTER(SceneDesc), C.POINTER(_vp)]),
    "tpt_scene_destroy": (C.c_int, [_vp]),
    "tpt_scene_get_info": (C.c_int, [_vp, C.POINTER(SceneInfo)]),
    "tpt_scene_get_bvh": (C.c_int, [_vp, C.c_int32, _lp, _dp, _ip, _ip, C.POINTER(C.c_uint32),
                                    C.POINTER(C.c_uint32), _lp]),
    "tpt_render": (C.c_int, [_vp, _vp, C.POINTER(RenderConfigC), _dp, _dp, _dp, C.POINTER(Stats)]),
    "tpt_render_device": (C.c_int, [_vp, _vp, C.POINTER(RenderConfigC), _vp, _vp, _vp, C.POINTER(Stats)]),
    "tpt_trace_closest": (C.c_int, [_vp, _vp, C.c_int64, _fp, _fp, _ip, _ip, _ip]),
    "tpt_trace_closest_noalloc": (C.c_int, [_vp, _vp, C.c_int64, _vp, _vp, _vp, _vp, _vp]),
    "tpt_render_plan": (C.c_int, [C.POINTER(RenderConfigC), C.c_int64, C.c_uint32, C.c_uint32, C.c_uint32,
                                  C.POINTER(PassPlan)]),
    "tpt_render_adjoint": (C.c_int, [_vp, _vp, C.POINTER(RenderConfigC), _fp, _dp, C.POINTER(Stats)]),
    "tpt_scene_update_materials": (C.c_int, [_vp, _vp, _dp]),
    "tpt_scene_build_host": (C.c_int, [C.POINTER(SceneDesc), C.POINTER(_vp)]),
    "tpt_scene_get_lights": (C.c_int, [_vp, _ip, _dp]),
    "tpt_scene_get_env_tables": (C.c_int, [_vp, _dp, _fp, _ip, _fp]),
    "tpt_scene_load_xml": (C.c_int, [C.c_char_p, C.POINTER(RenderConfigC), C.POINTER(_vp)]),
    "tpt_loaded_scene_desc": (C.c_int, [_vp, C.POINTER(C.POINTER(SceneDesc))]),
    "tpt_loaded_scene_warnings": (C.c_char_p, [_vp]),
    "tpt_loaded_scene_destroy": (C.c_int, [_vp]),
    "tpt_obj_load": (C.c_int, [C.c_char_p, _lp, _dp]),
    "tpt_render_pixels": (C.c_int, [_vp, _vp, C.POINTER(RenderConfigC), C.c_int64, _ip, _dp, _dp, _dp,
                                    C.POINTER(Stats)]),
    "tpt_resolve_pixels": (C.c_int, [C.c_int64, _dp, _dp, _dp]),
    "tpt_write_image": (C.c_int, [C.c_char_p, C.c_int32, C.c_int32, _dp, _dp, C.c_double, C.c_int32]),
    "tpt_checkpoint_save": (C.c_int, [C.c_char_p, C.c_uint64, C.c_int32, C.c_int32, C.c_uint64, _dp, _dp]),
    "tpt_checkpoint_load": (C.c_int, [C.c_char_p, C.c_uint64, C.c_int32, C.c_int32, C.POINTER(C.c_uint64), _dp, _dp,
                                      _ip]),
}
ABI_VERSION = 3  # TPT_ABI_VERSION of include/tinypt_gpu.h
TONEMAP_NONE = 0
TONEMAP_ACES = 1
