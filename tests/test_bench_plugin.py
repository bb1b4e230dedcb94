"""libbench.so, the reference's bench integrator (src/integrators/bench/
module.c:6-36, bench.cu:20-135), loaded the way the reference's application
loads integrators (src/common/integrator.c:8-38: dlopen, dlsym
"korangar_action", CREATE with a NULL integrator).  CPU tests cover the
entry point, the context actions and the error paths; the GPU test runs
RENDER, whose host and GPU boxes must be equal."""
import ctypes as C
import os
import re
import subprocess

import numpy as np
import pytest

import korangar_abi as K
from obvh_amd import build, synth

ACT = {"NONE": 0, "RENDER": 1, "CREATE": 2, "INIT": 3, "COMMIT": 4, "CAMERA_UPDATE": 5, "UPDATE": 6,
       "SETTINGS_UPDATE": 7, "DESTROY": 8, "ADS_ACTION": 9}
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


class Integrator(C.Structure):
    _fields_ = [("callback", C.c_void_p), ("library", C.c_void_p), ("context", C.c_void_p)]


class Payload(C.Structure):
    _fields_ = [("action", C.c_int), ("data", C.c_void_p)]


def load():
    lib = C.CDLL(build.plugin_path("bench"))
    fn = lib.korangar_action
    fn.restype = C.c_void_p
    fn.argtypes = [C.c_void_p, C.c_void_p, C.c_int]
    return lib, fn


def test_exports_only_the_entry_point():
    out = subprocess.run(["nm", "-D", "--defined-only", build.plugin_path("bench")], capture_output=True, text=True,
                         check=True).stdout
    syms = {line.split()[-1] for line in out.splitlines() if line.split()[1] in "TDB"}
    assert {s for s in syms if not s.startswith("_")} == {"korangar_action"}


def test_context_actions_and_errors():
    _, fn = load()
    it = Integrator()
    it.context = fn(None, None, ACT["CREATE"])
    assert it.context
    cont, keep = K.settings(benchmark_dito="y", benhcmark_sample_count=1000, intersector="lbvh")
    assert fn(C.byref(it), C.byref(cont), ACT["INIT"]) is None
    assert K.message(fn(C.byref(it), None, ACT["RENDER"])) == "bench: RENDER before COMMIT"
    empty = K.Scene()
    assert fn(C.byref(it), C.byref(empty), ACT["COMMIT"]) is None
    assert K.message(fn(C.byref(it), None, ACT["RENDER"])) == "bench: the scene has no instances"
    # an instance naming a sphere object
    objs = (K.Object * 1)()
    objs[0].type = K.OBJECT_SPHERE
    inst = (K.Instance * 1)()
    sc = K.Scene()
    sc.objects, sc.object_count = C.cast(objs, C.POINTER(K.Object)), 1
    sc.instances, sc.instance_count = C.cast(inst, C.POINTER(K.Instance)), 1
    assert fn(C.byref(it), C.byref(sc), ACT["COMMIT"]) is None
    assert K.message(fn(C.byref(it), None, ACT["RENDER"])) == "bench: instance 0 is not a mesh"
    inst[0].object_id = 3
    assert K.message(fn(C.byref(it), None, ACT["RENDER"])) == "bench: instance 0 names a missing object"
    p = Payload(1, None)
    assert K.message(fn(C.byref(it), C.byref(p), ACT["ADS_ACTION"])) == "bench: no intersector attached"
    for a in ("UPDATE", "CAMERA_UPDATE", "SETTINGS_UPDATE", "NONE"):
        assert fn(C.byref(it), None, ACT[a]) is None
    assert fn(C.byref(it), None, ACT["DESTROY"]) is None and not it.context
    assert "no context" in K.message(fn(C.byref(it), None, ACT["RENDER"]))


HOST_C = r"""
#include <dlfcn.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "obvh/korangar_integrator.h"

/* the reference's integrator loader and call sequence
   (src/common/integrator.c:8-38, src/app/main.cpp:383-384,141) */
int main(int argc, char** argv) {
    void* lib = dlopen(argv[1], RTLD_NOW);
    if (!lib) { printf("dlopen: %s\n", dlerror()); return 1; }
    kr_integrator it;
    memset(&it, 0, sizeof(it));
    it.callback = (kr_callback)dlsym(lib, "korangar_action");
    if (!it.callback) return 2;
    it.context = it.callback(NULL, NULL, KR_ACTION_CREATE);
    it.library = lib;
    kr_descriptor d[1] = {{"benchmark_dito", "y"}};
    kr_descriptor_container c = {d, 1};
    if (it.callback(&it, &c, KR_ACTION_INIT) != NULL) return 3;
    int n = atoi(argv[2]);
    kr_vec3* v = (kr_vec3*)malloc(sizeof(kr_vec3) * n);
    unsigned s = 12345u;
    for (int i = 0; i < n; ++i) {  /* an anisotropic cloud, rotated */
        float a = 0, b = 0, g = 0;
        for (int k = 0; k < 12; ++k) { s = s * 1664525u + 1013904223u; a += (float)(s >> 8) / 16777216.0f; }
        for (int k = 0; k < 12; ++k) { s = s * 1664525u + 1013904223u; b += (float)(s >> 8) / 16777216.0f; }
        for (int k = 0; k < 12; ++k) { s = s * 1664525u + 1013904223u; g += (float)(s >> 8) / 16777216.0f; }
        a = 4.0f * (a - 6.0f); b = b - 6.0f; g = 0.25f * (g - 6.0f);
        v[i].x = 0.6f * a - 0.8f * b + 3.0f; v[i].y = 0.8f * a + 0.6f * b - 1.0f; v[i].z = g + 0.5f;
    }
    kr_object obj;
    memset(&obj, 0, sizeof(obj));
    obj.type = KR_OBJECT_MESH;
    obj.as_mesh.vertices = v;
    obj.as_mesh.attr_count = (kr_size)n;
    kr_object_instance inst;
    memset(&inst, 0, sizeof(inst));
    kr_scene scene;
    memset(&scene, 0, sizeof(scene));
    scene.objects = &obj; scene.object_count = 1;
    scene.instances = &inst; scene.instance_count = 1;
    if (it.callback(&it, &scene, KR_ACTION_COMMIT) != NULL) return 4;
    kr_target target;
    memset(&target, 0, sizeof(target));
    const char* r = (const char*)it.callback(&it, &target, KR_ACTION_RENDER);
    printf("RENDER: %s\n", r ? r : "(null)");
    it.callback(&it, NULL, KR_ACTION_DESTROY);
    free(v);
    return it.context == NULL ? 0 : 5;
}
"""


def run_host(tmp_path, n):
    src = tmp_path / "host.c"
    src.write_text(HOST_C)
    exe = tmp_path / "host"
    subprocess.run(["gcc", "-std=c11", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"), str(src), "-o",
                    str(exe), "-ldl"], check=True)
    return subprocess.run([str(exe), build.plugin_path("bench"), str(n)], capture_output=True, text=True, timeout=120)


def boxes(out):
    """The CPU and GPU boxes the plugin printed, as float arrays."""
    res = {}
    for title in ("CPU", "GPU"):
        m = re.search(rf"Computed OBB {title}:\n(.*?)Area: ([^\n]+)\n", out, re.S)
        assert m, out
        nums = re.findall(r"[-+]?(?:\d+\.\d+|nan|inf)", m.group(1))
        res[title] = np.array([float(x) for x in nums], dtype=np.float64)
    return res


def test_c_host_drives_the_plugin_like_the_reference_cpu_box(tmp_path, oracle_lib):
    """Without a GPU the host fit still runs (and equals the oracle's box as
    printed); RENDER then reports the GPU fit's error."""
    import torch

    if torch.cuda.is_available():
        pytest.skip("GPU present: covered by the GPU test")
    r = run_host(tmp_path, 5000)
    assert r.returncode == 0, r.stdout + r.stderr
    b = boxes(r.stdout)
    assert len(b["CPU"]) == 15 and np.all(np.isfinite(b["CPU"]))
    assert "RENDER: " in r.stdout and "Simulation is done" not in r.stdout


@pytest.mark.gpu
def test_render_prints_equal_host_and_gpu_boxes(tmp_path):
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    print("[bench plugin] RENDER of a 131072-point cloud", flush=True)
    r = run_host(tmp_path, 131072)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "RENDER: Simulation is done" in r.stdout
    b = boxes(r.stdout)
    assert np.array_equal(b["CPU"], b["GPU"]) and np.all(np.isfinite(b["CPU"]))
