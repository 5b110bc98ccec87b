"""ctypes binding of the C ABI declared in include/trt.h (libtrt.so, built in-tree).

Loading fails loudly when the library has not been built: there is no fallback path.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libtrt.so")

c_float_p = ctypes.POINTER(ctypes.c_float)
c_u32_p = ctypes.POINTER(ctypes.c_uint32)
c_u8_p = ctypes.POINTER(ctypes.c_uint8)


class TrtMaterial(ctypes.Structure):
    _fields_ = [
        ("base_color", ctypes.c_float * 3), ("roughness", ctypes.c_float),
        ("emission", ctypes.c_float * 3), ("metallic", ctypes.c_float),
        ("specular", ctypes.c_float), ("specular_tint", ctypes.c_float),
        ("sheen", ctypes.c_float), ("sheen_tint", ctypes.c_float),
        ("clearcoat", ctypes.c_float), ("clearcoat_gloss", ctypes.c_float),
        ("subsurface", ctypes.c_float), ("anisotropic", ctypes.c_float),
        ("model", ctypes.c_uint32),
    ]


class TrtInstance(ctypes.Structure):
    _fields_ = [("transform", ctypes.c_float * 12), ("mesh", ctypes.c_uint32), ("material_offset", ctypes.c_uint32)]


class TrtCamera(ctypes.Structure):
    _fields_ = [
        ("position", ctypes.c_float * 3), ("target", ctypes.c_float * 3), ("up", ctypes.c_float * 3),
        ("fov", ctypes.c_float), ("aperture", ctypes.c_float), ("focal_dist", ctypes.c_float),
        ("distortion", ctypes.c_float), ("bokeh_sides", ctypes.c_int32), ("bokeh_rotation", ctypes.c_float),
    ]


class TrtSceneFileInfo(ctypes.Structure):
    _fields_ = [
        ("n_models", ctypes.c_uint32), ("n_instances", ctypes.c_uint32), ("n_materials", ctypes.c_uint32),
        ("n_triangles", ctypes.c_uint32), ("has_camera", ctypes.c_uint32), ("camera", TrtCamera),
        ("spp", ctypes.c_uint32), ("max_depth", ctypes.c_uint32), ("ray_epsilon", ctypes.c_float),
        ("ao_dist", ctypes.c_float), ("zdepth_max", ctypes.c_float),
        ("sky_enabled", ctypes.c_uint32), ("sky_draw_sun", ctypes.c_uint32), ("sky_dir", ctypes.c_float * 3),
        ("sky_angular_diameter", ctypes.c_float), ("sky_intensity", ctypes.c_float),
        ("sky_turbidity", ctypes.c_float), ("sky_bias", ctypes.c_float), ("n_warnings", ctypes.c_uint32),
    ]

    def as_dict(self):
        out = {}
        for k, _ in self._fields_:
            v = getattr(self, k)
            if k == "camera":
                v = dict(position=tuple(v.position), target=tuple(v.target), up=tuple(v.up), fov=v.fov,
                         aperture=v.aperture, focal_dist=v.focal_dist, distortion=v.distortion,
                         bokeh_sides=v.bokeh_sides, bokeh_rotation=v.bokeh_rotation)
            elif k == "sky_dir":
                v = tuple(v)
            out[k] = v
        return out


class TrtSky(ctypes.Structure):
    _fields_ = [("enabled", ctypes.c_uint32), ("draw_sun", ctypes.c_uint32), ("sun_dir", ctypes.c_float * 3),
                ("angular_diameter", ctypes.c_float), ("intensity", ctypes.c_float), ("turbidity", ctypes.c_float),
                ("sun_nee", ctypes.c_uint32)]


class TrtTonemapOpts(ctypes.Structure):
    _fields_ = [("method", ctypes.c_uint32), ("exposure", ctypes.c_float), ("gamma", ctypes.c_float),
                ("source", ctypes.c_uint32)]


class TrtBuildOpts(ctypes.Structure):
    _fields_ = [("bins", ctypes.c_int32), ("max_leaf", ctypes.c_int32), ("threads", ctypes.c_int32),
                ("presplit", ctypes.c_float)]


class TrtRenderOpts(ctypes.Structure):
    _fields_ = [
        ("width", ctypes.c_uint32), ("height", ctypes.c_uint32), ("spp", ctypes.c_uint32),
        ("max_depth", ctypes.c_uint32), ("rr_depth", ctypes.c_uint32), ("seed", ctypes.c_uint32),
        ("ray_epsilon", ctypes.c_float), ("clamp", ctypes.c_float),
        ("mode", ctypes.c_uint32), ("mode_flags", ctypes.c_uint32), ("ao_dist", ctypes.c_float),
        ("zdepth_max", ctypes.c_float),
    ]


class TrtStats(ctypes.Structure):
    _fields_ = [
        ("paths", ctypes.c_uint64), ("extend_rays", ctypes.c_uint64), ("shadow_rays", ctypes.c_uint64),
        ("primary_rays", ctypes.c_uint64),
        ("raygen_ms", ctypes.c_float), ("extend_ms", ctypes.c_float), ("shade_ms", ctypes.c_float),
        ("shadow_ms", ctypes.c_float), ("finalize_ms", ctypes.c_float), ("frame_ms", ctypes.c_float),
        ("extend_launches", ctypes.c_uint32), ("shadow_launches", ctypes.c_uint32),
        ("sample_count", ctypes.c_uint32), ("overflow", ctypes.c_uint32), ("trace_ms", ctypes.c_float),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class TrtBvhInfo(ctypes.Structure):
    _fields_ = [
        ("nodes", ctypes.c_uint64), ("leaves", ctypes.c_uint64), ("depth", ctypes.c_uint32),
        ("stack_entries", ctypes.c_uint32), ("sah_cost", ctypes.c_double), ("build_ms", ctypes.c_double),
        ("device_bytes", ctypes.c_uint64), ("two_level", ctypes.c_uint32), ("tlas_depth", ctypes.c_uint32),
        ("references", ctypes.c_uint64),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class TrtPickResult(ctypes.Structure):
    _fields_ = [("origin", ctypes.c_float * 3), ("direction", ctypes.c_float * 3), ("t", ctypes.c_float),
                ("instance", ctypes.c_uint32), ("primitive", ctypes.c_uint32)]


class TrtDenoiseStats(ctypes.Structure):
    _fields_ = [("features_ms", ctypes.c_float), ("network_ms", ctypes.c_float), ("total_ms", ctypes.c_float)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class TrtTraceCounts(ctypes.Structure):
    _fields_ = [
        ("closest_nodes_per_ray", ctypes.c_double), ("closest_tris_per_ray", ctypes.c_double),
        ("any_nodes_per_ray", ctypes.c_double), ("any_tris_per_ray", ctypes.c_double),
        ("closest_rays", ctypes.c_uint64), ("any_rays", ctypes.c_uint64),
        ("node_bytes", ctypes.c_uint32), ("tri_bytes", ctypes.c_uint32),
        ("closest_spills", ctypes.c_uint64), ("any_spills", ctypes.c_uint64),
        ("closest_distinct_nodes", ctypes.c_uint64), ("closest_distinct_tris", ctypes.c_uint64),
        ("any_distinct_nodes", ctypes.c_uint64), ("any_distinct_tris", ctypes.c_uint64),
        ("launches", ctypes.c_uint32),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


# (name, restype, argtypes) for every entry point of include/trt.h
P = ctypes.c_void_p
S = ctypes.c_int
PROTOTYPES = [
    ("trt_create", S, [ctypes.c_int, ctypes.POINTER(P)]),
    ("trt_destroy", None, [P]),
    ("trt_last_error", ctypes.c_char_p, [P]),
    ("trt_version", ctypes.c_char_p, []),
    ("trt_clear_scene", S, [P]),
    ("trt_set_materials", S, [P, ctypes.POINTER(TrtMaterial), ctypes.c_uint32]),
    ("trt_add_mesh", S, [P, c_float_p, c_float_p, c_u32_p, ctypes.c_uint32, c_u32_p]),
    ("trt_set_instances", S, [P, ctypes.POINTER(TrtInstance), ctypes.c_uint32]),
    ("trt_build", S, [P, ctypes.POINTER(TrtBuildOpts)]),
    ("trt_bvh_info_get", S, [P, ctypes.POINTER(TrtBvhInfo)]),
    ("trt_scene_file_read", S, [ctypes.c_char_p, ctypes.POINTER(TrtSceneFileInfo), c_float_p, c_float_p, c_u32_p,
                                c_u32_p, c_u32_p, c_u32_p, ctypes.POINTER(TrtMaterial), ctypes.POINTER(TrtInstance),
                                ctypes.c_char_p, ctypes.c_size_t]),
    ("trt_load_scene_file", S, [P, ctypes.c_char_p, ctypes.POINTER(TrtBuildOpts), ctypes.POINTER(TrtSceneFileInfo)]),
    ("trt_set_camera", S, [P, ctypes.POINTER(TrtCamera)]),
    ("trt_set_sky", S, [P, ctypes.POINTER(TrtSky)]),
    ("trt_render", S, [P, ctypes.POINTER(TrtRenderOpts), ctypes.POINTER(TrtStats)]),
    ("trt_set_phase_timing", S, [P, ctypes.c_int]),
    ("trt_reset", S, [P]),
    ("trt_sample_count", ctypes.c_uint32, [P]),
    ("trt_frame_id", ctypes.c_uint64, [P]),
    ("trt_trace_closest_noalloc", S, [P, c_float_p, ctypes.c_uint32, c_float_p, c_u32_p]),
    ("trt_write_accum", S, [P, ctypes.c_uint32, ctypes.c_uint32, c_float_p]),
    ("trt_read_accum", S, [P, c_float_p]),
    ("trt_read_color", S, [P, c_float_p]),
    ("trt_read_albedo", S, [P, c_float_p]),
    ("trt_read_normal", S, [P, c_float_p]),
    ("trt_read_path_radiance", S, [P, c_float_p, ctypes.c_uint64]),
    ("trt_trace_closest", S, [P, c_float_p, ctypes.c_uint32, c_float_p, c_u32_p]),
    ("trt_trace_any", S, [P, c_float_p, ctypes.c_uint32, c_u8_p]),
    ("trt_pick", S, [P, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32,
                     ctypes.POINTER(TrtPickResult)]),
    ("trt_measure_traversal", S, [P, ctypes.POINTER(TrtRenderOpts), ctypes.POINTER(TrtTraceCounts)]),
    ("trt_measure_gather", S, [ctypes.c_int, ctypes.c_uint64, ctypes.c_uint32, ctypes.POINTER(ctypes.c_double)]),
    ("trt_measure_gather_lanes", S, [ctypes.c_int, ctypes.c_uint64, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32,
                                     ctypes.POINTER(ctypes.c_double)]),
    ("trt_set_material_params", S, [P, c_float_p, ctypes.c_uint32, ctypes.c_uint32]),
    ("trt_backward", S, [P, ctypes.POINTER(TrtRenderOpts), c_float_p, c_float_p]),
    ("trt_denoiser_init", S, [P, ctypes.c_uint32]),
    ("trt_denoiser_num_params", ctypes.c_uint64, []),
    ("trt_denoiser_get_params", S, [P, c_float_p, ctypes.c_uint64]),
    ("trt_denoiser_set_params", S, [P, c_float_p, ctypes.c_uint64]),
    ("trt_denoiser_save", S, [P, ctypes.c_char_p]),
    ("trt_denoiser_load", S, [P, ctypes.c_char_p]),
    ("trt_denoiser_dataset", S, [P, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32]),
    ("trt_denoiser_capture", S, [P, ctypes.c_uint32, ctypes.c_uint32]),
    ("trt_denoiser_train", S, [P, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_float, ctypes.c_uint32,
                               c_float_p]),
    ("trt_denoise", S, [P, ctypes.POINTER(TrtDenoiseStats)]),
    ("trt_denoiser_set_policy", S, [P, ctypes.c_uint32, ctypes.c_uint32]),
    ("trt_denoise_frame", S, [P, ctypes.POINTER(TrtDenoiseStats), ctypes.POINTER(ctypes.c_uint32)]),
    ("trt_read_denoised", S, [P, c_float_p]),
    ("trt_tonemap", S, [P, ctypes.POINTER(TrtTonemapOpts), c_u8_p]),
    ("trt_write_pfm", S, [ctypes.c_char_p, ctypes.c_uint32, ctypes.c_uint32, c_float_p]),
    ("trt_write_png", S, [ctypes.c_char_p, ctypes.c_uint32, ctypes.c_uint32, c_u8_p]),
    ("trt_denoiser_forward", S, [P, c_float_p, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, c_float_p]),
    ("trt_denoiser_activation", S, [P, ctypes.c_uint32, c_float_p, ctypes.c_uint64]),
    ("trt_denoiser_apply", S, [P, c_float_p, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, c_float_p]),
    ("trt_denoiser_backward", S, [P, c_float_p, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, c_float_p]),
]

# render modes (TRT_MODE_* of include/trt.h)
MODES = {"path_tracing": 0, "ambient_occlusion": 1, "ambient_occlusion_shading": 2, "bitangent": 3,
         "geometric_normal": 4, "material_id": 5, "material_property": 6, "object_id": 7, "shading_normal": 8,
         "tangent": 9, "texture_coordinate": 10, "wireframe": 11, "zdepth": 12}
MATERIAL_PROPERTIES = ["anisotropic", "clearcoat", "clearcoatgloss", "diffuse", "emissive", "metallic", "normal",
                       "roughness", "sheen", "sheentint", "specular", "speculartint", "subsurface"]

STATUS = {0: "TRT_OK", 1: "TRT_INVALID_ARG", 2: "TRT_OUT_OF_MEMORY", 3: "TRT_DEVICE", 4: "TRT_NOT_BUILT",
          5: "TRT_UNSUPPORTED"}

_lib = None


TRT_NOALLOC_MAX_RAYS = 65536  # include/trt.h


class TrtError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(f"{STATUS.get(status, status)}: {msg}")
        self.status = status


def load(path=LIB_PATH):
    """Loads libtrt.so and declares every prototype; raises if it is missing."""
    global _lib
    if _lib is not None and path == LIB_PATH:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(f"{path} is not built: run `python -c 'import __graft_entry__ as g; g.build()'` "
                           "or `python trayracer_amd/build.py`")
    lib = ctypes.CDLL(path)
    for name, res, args in PROTOTYPES:
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if path == LIB_PATH:
        _lib = lib
    return lib
