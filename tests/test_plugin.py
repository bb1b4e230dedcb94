"""The intersector plugins (liblbvh_gpu.so, libatrbvh_gpu.so) on the CPU:
layouts, the exported entry point, and the actions that need no GPU (NAME,
PREFERRED_QUERY_TYPE, CREATE, INIT, DESTROY, COMMIT's scene checks).  The
reference's behaviour is src/ads/lbvh_gpu/module.c:6-41 and lbvh.cu:1764-1800
(empty scene message)."""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

import korangar_abi as K

PLUGINS = {"lbvh_gpu": "LBVH GPU", "atrbvh_gpu": "ATRBVH"}


@pytest.mark.parametrize("cls", sorted(K.SIZES, key=lambda c: c.__name__))
def test_layouts_match_the_reference(cls):
    assert C.sizeof(cls) == K.SIZES[cls]


def test_scene_field_offsets():
    assert K.Scene.objects.offset == 304
    assert K.Scene.instances.offset == 320
    assert K.Scene.benchmark_file.offset == 416
    assert K.Object.u.offset == 32
    assert K.Instance.object_id.offset == 128


@pytest.mark.parametrize("name", sorted(PLUGINS))
def test_exports_only_the_entry_point(name):
    K.load(name)
    out = subprocess.run(["nm", "-D", "--defined-only", K.build.plugin_path(name)], capture_output=True,
                         text=True, check=True).stdout
    syms = {line.split()[-1] for line in out.splitlines() if line.split()[1] in "TDB"}
    assert {s for s in syms if not s.startswith("_")} == {"korangar_ads_action"}


@pytest.mark.parametrize("name", sorted(PLUGINS))
def test_name_and_query_type(name):
    _, fn = K.load(name)
    assert K.message(fn(None, None, K.ACTION["NAME"])) == PLUGINS[name]
    assert fn(None, None, K.ACTION["PREFERRED_QUERY_TYPE"]) == K.QUERY_TYPE_CUDA
    assert fn(None, None, K.ACTION["CSV_EXPORT"]) is None  # unhandled actions succeed


@pytest.mark.parametrize("name", sorted(PLUGINS))
def test_create_init_commit_checks_destroy(name):
    _, fn = K.load(name)
    ads = K.Ads()
    ads.context = fn(None, None, K.ACTION["CREATE"])
    assert ads.context
    cont, keep = K.settings(collapse="y", use_obbs="y", use_persistent_kernel="y", ct=1.0, ci=1.2)
    assert fn(C.byref(ads), C.byref(cont), K.ACTION["INIT"]) is None
    # an empty scene: the reference's message, no GPU work
    empty = K.Scene()
    assert K.message(fn(C.byref(ads), C.byref(empty), K.ACTION["COMMIT"])) == \
        "The scene has no intersectable primitives"
    # a mesh object without faces counts no primitives either
    V = np.zeros((3, 3), np.float32)
    scene, keep2 = K.mesh_scene(V, np.zeros((0, 4), np.uint32))
    assert K.message(fn(C.byref(ads), C.byref(scene), K.ACTION["COMMIT"])) == \
        "The scene has no intersectable primitives"
    # an AABB object reached through an instance counts one primitive but is not a mesh
    objs = (K.Object * 1)()
    objs[0].type = K.OBJECT_AABB
    inst = (K.Instance * 1)()
    inst[0].object_id = 0
    sc = K.Scene()
    sc.objects = C.cast(objs, C.POINTER(K.Object))
    sc.object_count = 1
    sc.instances = C.cast(inst, C.POINTER(K.Instance))
    sc.instance_count = 1
    assert "only mesh objects" in K.message(fn(C.byref(ads), C.byref(sc), K.ACTION["COMMIT"]))
    inst[0].object_id = 5
    assert "missing object" in K.message(fn(C.byref(ads), C.byref(sc), K.ACTION["COMMIT"]))
    # a query before any tree
    q = K.Query(K.QUERY_TYPE_CUDA, None, None, 0)
    assert "before a successful COMMIT" in K.message(fn(C.byref(ads), C.byref(q), K.ACTION["INTERSECTION_QUERY"]))
    assert fn(C.byref(ads), None, K.ACTION["DESTROY"]) is None
    assert ads.context is None


@pytest.mark.parametrize("name", sorted(PLUGINS))
def test_persistent_aabb_mode_is_refused(name):
    _, fn = K.load(name)
    ads = K.Ads()
    ads.context = fn(None, None, K.ACTION["CREATE"])
    cont, keep = K.settings(use_obbs="n", use_persistent_kernel="y")
    assert fn(C.byref(ads), C.byref(cont), K.ACTION["INIT"]) is None
    V = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], np.float32)
    scene, keep2 = K.mesh_scene(V, np.array([[0, 1, 2, 0]], np.uint32))
    assert "AABB traversal" in K.message(fn(C.byref(ads), C.byref(scene), K.ACTION["COMMIT"]))
    assert fn(C.byref(ads), None, K.ACTION["DESTROY"]) is None


def test_actions_without_a_context_fail():
    _, fn = K.load("lbvh_gpu")
    ads = K.Ads()
    for a in ("INIT", "COMMIT", "INTERSECTION_QUERY"):
        assert "no context" in K.message(fn(C.byref(ads), None, K.ACTION[a]))


HOST_C = r"""
#include <dlfcn.h>
#include <stdio.h>
#include <string.h>
#include "obvh/korangar_ads.h"

/* the reference's loader, src/common/ads.c:6-37 */
int main(int argc, char** argv) {
    void* lib = dlopen(argv[1], RTLD_NOW);
    if (!lib) { printf("dlopen: %s\n", dlerror()); return 1; }
    kr_ads ads;
    memset(&ads, 0, sizeof(ads));
    ads.callback = (kr_ads_callback)dlsym(lib, "korangar_ads_action");
    if (!ads.callback) return 2;
    ads.context = ads.callback(NULL, NULL, KR_ADS_ACTION_CREATE);
    ads.library = lib;
    kr_descriptor d[2] = {{"use_obbs", "y"}, {"collapse", "y"}};
    kr_descriptor_container c = {d, 2};
    if (ads.callback(&ads, &c, KR_ADS_ACTION_INIT) != NULL) return 3;
    kr_scene scene;
    memset(&scene, 0, sizeof(scene));
    const char* e = (const char*)ads.callback(&ads, &scene, KR_ADS_ACTION_COMMIT);
    printf("%s|%s|%d\n", (const char*)ads.callback(&ads, NULL, KR_ADS_ACTION_NAME), e ? e : "(null)",
           (int)(size_t)ads.callback(&ads, NULL, KR_ADS_ACTION_PREFERRED_QUERY_TYPE));
    ads.callback(&ads, NULL, KR_ADS_ACTION_DESTROY);
    return ads.context == NULL ? 0 : 4;
}
"""


@pytest.mark.parametrize("name", sorted(PLUGINS))
def test_c_host_loads_the_plugin_like_the_reference(tmp_path, name):
    """A C program using only include/obvh/korangar_ads.h and dlopen drives
    the plugin the way the reference's application does."""
    src = tmp_path / "host.c"
    src.write_text(HOST_C)
    exe = tmp_path / "host"
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    subprocess.run(["gcc", "-std=c11", "-Wall", "-Werror", "-I", os.path.join(root, "include"), str(src), "-o",
                    str(exe), "-ldl"], check=True)
    r = subprocess.run([str(exe), K.build.plugin_path(name)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.strip() == f"{PLUGINS[name]}|The scene has no intersectable primitives|2"
