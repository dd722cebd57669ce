"""ctypes bindings of libspray_gpu.so (the C ABI declared in include/spray_gpu.h).

The library is built in-tree (``spray_amd/libspray_gpu.so``); importing this
module fails loudly when it is missing -- there is no CPU fallback.
"""
import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
# SPRAY_LIB names another build of the same library (e.g. the host-sanitized
# one the CPU suite can run against)
LIB_PATH = os.environ.get("SPRAY_LIB") or os.path.join(_HERE, "libspray_gpu.so")

SPRAY_OK = 0
SPRAY_ERR_INVALID = 1
SPRAY_ERR_IO = 2
SPRAY_ERR_FORMAT = 3
SPRAY_ERR_OOM = 4
SPRAY_ERR_CAPACITY = 5
SPRAY_ERR_DEVICE = 6
SPRAY_ERR_STATE = 7

SCHEDULE_PLAIN = 0
SCHEDULE_BATCHED = 1
SCHEDULE_LEARNED = 2

QUERY_CLOSEST = 0
QUERY_ANY = 1


class spray_config(C.Structure):
    _fields_ = [
        ("image_w", C.c_int), ("image_h", C.c_int),
        ("has_camera", C.c_int),
        ("camera_pos", C.c_float * 3), ("camera_lookat", C.c_float * 3),
        ("camera_up", C.c_float * 3),
        ("fov", C.c_float),
        ("pixel_samples", C.c_int), ("ao_samples", C.c_int), ("ao_mode", C.c_int),
        ("bounces", C.c_int),
        ("ks", C.c_float * 3), ("shininess", C.c_float),
        ("max_samples_per_rank", C.c_longlong),
        ("nthreads", C.c_int),
        ("ao_max_dist", C.c_float),
        ("rr_start", C.c_int),
        ("schedule", C.c_int),
        ("device", C.c_int),
    ]


class spray_stats(C.Structure):
    _fields_ = [
        ("useful_rays", C.c_ulonglong), ("eye_rays", C.c_ulonglong),
        ("radiance_visits", C.c_ulonglong), ("shadow_visits", C.c_ulonglong),
        ("shades", C.c_ulonglong), ("final_hits", C.c_ulonglong),
        ("passes", C.c_ulonglong), ("windows", C.c_ulonglong), ("tiles", C.c_ulonglong),
        ("queue_entries", C.c_ulonglong), ("deferred_entries", C.c_ulonglong),
        ("pruned_entries", C.c_ulonglong),
        ("node_visits", C.c_ulonglong), ("tri_tests", C.c_ulonglong),
        ("trace_ms", C.c_double), ("build_ms", C.c_double),
        ("cache_hits", C.c_ulonglong), ("cache_misses", C.c_ulonglong), ("cache_bytes", C.c_ulonglong),
        ("regrows", C.c_ulonglong),
    ]


class spray_rays_soa(C.Structure):
    _fields_ = [(n, C.POINTER(C.c_float)) for n in ("ox", "oy", "oz", "dx", "dy", "dz", "tnear", "tfar")]


class spray_hits_soa(C.Structure):
    _fields_ = [("t", C.POINTER(C.c_float)), ("u", C.POINTER(C.c_float)),
                ("v", C.POINTER(C.c_float)), ("prim", C.POINTER(C.c_int))]


class spray_scenegen_params(C.Structure):
    _fields_ = [
        ("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int),
        ("res", C.c_int),
        ("a", C.c_double), ("b", C.c_double), ("phase", C.c_double),
        ("per_domain_random", C.c_int),
        ("seed", C.c_ulonglong),
        ("tri_total", C.c_double),
        ("rgb", C.c_uint),
    ]


_P = C.c_void_p
_FP = C.POINTER(C.c_float)
_IP = C.POINTER(C.c_int)

# name -> (restype, argtypes)
PROTOTYPES = {
    "spray_config_defaults": (None, [C.POINTER(spray_config)]),
    "spray_create": (C.c_int, [C.POINTER(spray_config), C.POINTER(_P)]),
    "spray_destroy": (None, [_P]),
    "spray_last_error": (C.c_char_p, [_P]),
    "spray_set_config": (C.c_int, [_P, C.POINTER(spray_config)]),
    "spray_load_scene": (C.c_int, [_P, C.c_char_p, C.c_char_p]),
    "spray_num_domains": (C.c_int, [_P]),
    "spray_domain_info": (C.c_int, [_P, C.c_int, C.POINTER(C.c_longlong), C.POINTER(C.c_longlong), _FP, _FP]),
    "spray_domain_normals": (C.c_int, [_P, C.c_int, _FP]),
    "spray_set_camera": (C.c_int, [_P, _FP, _FP, _FP, C.c_float]),
    "spray_trace_frame": (C.c_int, [_P, _FP, C.POINTER(spray_stats)]),
    "spray_trace_rays": (C.c_int, [_P, C.POINTER(spray_rays_soa), C.c_size_t, C.POINTER(spray_hits_soa), C.c_int, C.c_int]),
    "spray_domain_list": (C.c_int, [_P, _FP, _FP, C.c_size_t, C.c_int, _IP, _IP, _FP]),
    "spray_set_schedule": (C.c_int, [_P, C.c_int]),
    "spray_autotune": (C.c_int, [_P, _IP, _IP]),
    "spray_set_batch_params": (C.c_int, [_P, C.c_int, C.c_int]),
    "spray_set_tiles_in_flight": (C.c_int, [_P, C.c_int]),
    "spray_set_bvh_builder": (C.c_int, [_P, C.c_int]),
    "spray_set_cache_size": (C.c_int, [_P, C.c_int]),
    "spray_set_pool_headroom": (C.c_int, [_P, C.c_double]),
    "spray_cache_state": (C.c_int, [_P, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_ulonglong)]),
    "spray_profile_kernels": (C.c_int, [_P, C.c_int]),
    "spray_kernel_time": (C.c_int, [_P, C.POINTER(C.c_longlong), C.POINTER(C.c_double), C.c_int]),
    "spray_schedule_log": (C.c_int, [_P, C.c_int]),
    "spray_schedule_read": (C.c_int, [_P, _IP, _IP, C.POINTER(C.c_ulonglong), _IP]),
    "spray_predictor_log": (C.c_int, [_P, C.c_int, C.c_longlong]),
    "spray_predictor_num_samples": (C.c_longlong, [_P]),
    "spray_predictor_train": (C.c_int, [_P, C.c_int, C.c_int, C.c_float, _FP]),
    "spray_predictor_eval": (C.c_int, [_P, _FP, _FP]),
    "spray_predictor_save": (C.c_int, [_P, C.c_char_p]),
    "spray_predictor_load": (C.c_int, [_P, C.c_char_p]),
    "spray_mlp_num_params": (C.c_int, [_P]),
    "spray_mlp_grad": (C.c_int, [_P, _FP, C.c_int, _FP, _FP, _FP]),
    "spray_mlp_logits": (C.c_int, [_P, _FP, C.c_int, _FP, _FP]),
    "spray_write_ppm": (C.c_int, [_FP, C.c_int, C.c_int, C.c_char_p]),
    "spray_scenegen_write": (C.c_int, [C.POINTER(spray_scenegen_params), C.c_char_p, C.c_char_p, C.c_char_p, C.POINTER(C.c_longlong)]),
    "spray_trace_rays_device": (C.c_int, [_P, _FP, _FP, _FP, _FP, C.c_size_t, _FP, _FP, _FP, _IP, C.c_int, C.c_int]),
    "spray_scenegen_domain": (C.c_int, [C.POINTER(spray_scenegen_params), C.c_int, _FP, C.POINTER(C.c_uint),
                                        C.POINTER(C.c_uint), C.POINTER(C.c_longlong), C.POINTER(C.c_longlong)]),
    "spray_scenegen_load": (C.c_int, [_P, C.POINTER(spray_scenegen_params), C.c_char_p, C.c_char_p]),
    "spray_remesh": (C.c_int, [_P, C.c_double]),
    "spray_domain_mesh": (C.c_int, [_P, C.c_int, _FP, C.POINTER(C.c_uint), C.POINTER(C.c_uint),
                                    C.POINTER(C.c_longlong), C.POINTER(C.c_longlong)]),
    "spray_write_scene": (C.c_int, [_P, C.c_char_p, C.c_char_p]),
    "spray_ply_info": (C.c_int, [_P, C.c_char_p, C.POINTER(C.c_longlong), C.POINTER(C.c_longlong), _FP, _FP]),
}

_lib = None


def load():
    """Loads libspray_gpu.so once; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"{LIB_PATH} is missing: build it with `python -c 'import __graft_entry__ as g; g.build()'`")
    lib = C.CDLL(LIB_PATH)
    for name, (res, args) in PROTOTYPES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib
