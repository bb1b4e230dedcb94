"""ctypes mirrors of the intersector plugin interface (include/obvh/
korangar_ads.h, the reference's src/common/korangar.h and ads.h layouts) and
helpers that drive a plugin the way the reference's host does (ads.c:6-37:
dlopen, dlsym("korangar_ads_action"), CREATE with a NULL ads, then actions
on the kr_ads that holds the context)."""
import ctypes as C
import os

import numpy as np

from obvh_amd import build

ACTION = {"NONE": 0, "INTERSECTION_QUERY": 1, "OCCLUSION_QUERY": 2, "PREFERRED_QUERY_TYPE": 3, "GET_BLAS": 4,
          "GET_TLAS": 5, "CREATE": 6, "COMMIT": 7, "INIT": 8, "UPDATE": 9, "SETTINGS_UPDATE": 10, "DESTROY": 11,
          "CSV_EXPORT": 12, "NAME": 17}
QUERY_TYPE_CUDA = 2
OBJECT_MESH, OBJECT_SPHERE, OBJECT_AABB = 1, 2, 3


class Mat4(C.Structure):
    _fields_ = [("v", C.c_float * 16)]


class Transform(C.Structure):
    _fields_ = [("to", Mat4), ("from_", Mat4)]


class Camera(C.Structure):
    _fields_ = [("type", C.c_int), ("view", Transform), ("viewport", C.c_uint32 * 4),
                ("projection", Transform), ("fov", C.c_float)]


class Mesh(C.Structure):
    _fields_ = [("vertices", C.c_void_p), ("normals", C.c_void_p), ("uvs", C.c_void_p), ("faces", C.c_void_p),
                ("face_count", C.c_uint64), ("attr_count", C.c_uint64)]


class Sphere(C.Structure):
    _fields_ = [("center", C.c_float * 3), ("radius", C.c_float)]


class ObjectUnion(C.Union):
    _fields_ = [("as_sphere", Sphere), ("as_mesh", Mesh)]


class Object(C.Structure):
    _fields_ = [("aabb", C.c_float * 6), ("type", C.c_int), ("u", ObjectUnion)]


class Instance(C.Structure):
    _fields_ = [("model", Transform), ("object_id", C.c_uint64)]


class Scene(C.Structure):
    _fields_ = [("camera", Camera), ("aabb", C.c_float * 6),
                ("objects", C.POINTER(Object)), ("object_count", C.c_uint64),
                ("instances", C.POINTER(Instance)), ("instance_count", C.c_uint64),
                ("cameras", C.c_void_p), ("camera_count", C.c_uint64),
                ("materials", C.c_void_p), ("material_count", C.c_uint64),
                ("textures", C.c_void_p), ("texture_count", C.c_uint64),
                ("samplers", C.c_void_p), ("sampler_count", C.c_uint64),
                ("frame", C.c_uint32 * 2), ("tile", C.c_uint32 * 2), ("benchmark_file", C.c_char_p)]


class Ads(C.Structure):
    _fields_ = [("callback", C.c_void_p), ("library", C.c_void_p), ("context", C.c_void_p),
                ("runtime", C.c_void_p)]


class Descriptor(C.Structure):
    _fields_ = [("key", C.c_char_p), ("value", C.c_char_p)]


class DescriptorContainer(C.Structure):
    _fields_ = [("descriptors", C.POINTER(Descriptor)), ("descriptor_count", C.c_uint64)]


class Query(C.Structure):
    _fields_ = [("type", C.c_int), ("rays", C.c_void_p), ("isects", C.c_void_p), ("ray_count", C.c_uint64)]


SIZES = {Transform: 128, Camera: 280, Object: 80, Instance: 136, Scene: 424, Ads: 32, Query: 32,
         DescriptorContainer: 16}


def load(name: str):
    """The plugin library lib<name>.so and its entry point (restype: raw
    pointer, since results are strings, contexts or small integers)."""
    path = build.plugin_path(name)
    if not os.path.exists(path):
        raise RuntimeError(f"{path} is not built")
    lib = C.CDLL(path)
    fn = lib.korangar_ads_action
    fn.restype = C.c_void_p
    fn.argtypes = [C.c_void_p, C.c_void_p, C.c_int]
    return lib, fn


def message(ptr) -> str | None:
    return None if not ptr else C.string_at(ptr).decode()


def settings(**kv):
    """kr_descriptor_container of key = value strings (kept alive by the caller)."""
    items = [(k.encode(), (v if isinstance(v, str) else str(v)).encode()) for k, v in kv.items()]
    arr = (Descriptor * max(len(items), 1))(*[Descriptor(k, v) for k, v in items])
    return DescriptorContainer(C.cast(arr, C.POINTER(Descriptor)), len(items)), arr


def mesh_scene(V: np.ndarray, F: np.ndarray, benchmark_file: str | None = None):
    """A kr_scene with one mesh object and no instances; host arrays are
    returned with it so they stay alive."""
    V = np.ascontiguousarray(V, dtype=np.float32)
    F = np.ascontiguousarray(F, dtype=np.uint32)
    objs = (Object * 1)()
    objs[0].type = OBJECT_MESH
    m = objs[0].u.as_mesh
    m.vertices = V.ctypes.data
    m.faces = F.ctypes.data
    m.face_count = F.shape[0]
    m.attr_count = V.shape[0]
    scene = Scene()
    scene.objects = C.cast(objs, C.POINTER(Object))
    scene.object_count = 1
    scene.benchmark_file = benchmark_file.encode() if benchmark_file else None
    return scene, (V, F, objs)
