"""The C ABI (include/eos_c.h) as shipped in eos_amd/libeos_amd.so, checked without a GPU."""
import ctypes as C
import os
import re
import subprocess

import pytest

from eos_amd import _lib as E

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "eos_c.h")


def _declared():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^(?:eos_status|const char\s*\*|int32_t)\s+(eos_\w+)\s*\(", text, re.M)))


def test_every_declared_symbol_is_exported():
    names = _declared()
    assert len(names) >= 15
    assert sorted(E.EXPORTS) == names
    lib = E.lib()
    for n in names:
        assert hasattr(lib, n), n


def test_struct_layouts_match_header(tmp_path):
    """sizeof/offsetof from the C compiler equal the ctypes mirrors."""
    structs = {"eos_scene_desc": E.SceneDesc, "eos_build_stats": E.BuildStats, "eos_build_opts": E.BuildOpts,
               "eos_render_opts": E.RenderOpts,
               "eos_render_stats": E.RenderStats, "eos_adjoint_opts": E.AdjointOpts,
               "eos_adjoint_stats": E.AdjointStats, "eos_batch_opts": E.BatchOpts,
               "eos_batch_stats": E.BatchStats, "eos_launch_config": E.LaunchConfig, "eos_trace_opts": E.TraceOpts}
    src = ['#include <stdio.h>', '#include <stddef.h>', f'#include "{HEADER}"', "int main(void){"]
    for cname, py in structs.items():
        src.append(f'printf("{cname} %zu\\n", sizeof({cname}));')
        for f, _ in py._fields_:
            src.append(f'printf("{cname}.{f} %zu\\n", offsetof({cname}, {f}));')
    src += ['printf("eos_camera %zu\\n", sizeof(eos_camera));', 'printf("eos_material %zu\\n", sizeof(eos_material));',
            'printf("eos_light %zu\\n", sizeof(eos_light));', 'printf("eos_ray_hit %zu\\n", sizeof(eos_ray_hit));']
    for f in E.RAY_HIT.names:
        src.append(f'printf("eos_ray_hit.{f} %zu\\n", offsetof(eos_ray_hit, {f}));')
    src.append("return 0;}")
    c = tmp_path / "probe.c"
    c.write_text("\n".join(src))
    exe = tmp_path / "probe"
    subprocess.check_call(["gcc", "-std=c99", "-o", str(exe), str(c)])
    got = dict(line.split() for line in subprocess.check_output([str(exe)]).decode().splitlines())
    for cname, py in structs.items():
        assert int(got[cname]) == C.sizeof(py), cname
        for f, _ in py._fields_:
            assert int(got[f"{cname}.{f}"]) == getattr(py, f).offset, f"{cname}.{f}"
    assert int(got["eos_camera"]) == 56 and int(got["eos_material"]) == 64 and int(got["eos_light"]) == 32
    assert int(got["eos_ray_hit"]) == E.RAY_HIT.itemsize == 48
    for f in E.RAY_HIT.names:
        assert int(got[f"eos_ray_hit.{f}"]) == E.RAY_HIT.fields[f][1], f


def test_null_arguments_are_rejected():
    L = E.lib()
    h = C.c_void_p()
    assert L.eos_create_scene(None, C.byref(h)) == E.EOS_ERR_INVALID
    assert b"null" in L.eos_last_error()
    assert L.eos_render(None, None, None, None, None) == E.EOS_ERR_INVALID
    assert L.eos_build_lbvh(None, 4, None) == E.EOS_ERR_INVALID
    assert L.eos_render_adjoint(None, None, None, None, None, None, None) == E.EOS_ERR_INVALID
    assert L.eos_render_batch(None, None, 0, None, None, None) == E.EOS_ERR_INVALID
    assert L.eos_debug_read(None, 1, None, 0) == E.EOS_ERR_INVALID
    assert L.eos_destroy_scene(None) == E.EOS_OK
    assert L.eos_abi_version() == E.ABI_VERSION == 7
    assert L.eos_set_transforms(None, 0, None, None, 0) == E.EOS_ERR_INVALID
    assert L.eos_render_rows(None, None, 0, 1, None, None, None) == E.EOS_ERR_INVALID
    assert L.eos_render_frames(None, None, 1, None, None, None) == E.EOS_ERR_INVALID
    assert L.eos_set_ambient_refractive_index(None, 1.0) == E.EOS_ERR_INVALID
    assert L.eos_trace_rays(None, 1, None, None, None, None, None) == E.EOS_ERR_INVALID


def test_no_device_is_reported():
    try:
        E.device_name()
        pytest.skip("a HIP device is present")
    except E.EosError as e:
        assert e.status == E.EOS_ERR_NODEVICE


def test_scene_limits_checked_before_device_use():
    from eos_amd import scenes
    import numpy as np
    a = scenes.config_a(8, 8)
    lights = scenes.make_lights(np.zeros((33, 3)))
    d = E.SceneDesc(0, None, None, None, 0, 1, a.materials.ctypes.data, 33, lights.ctypes.data)
    h = C.c_void_p()
    assert E.lib().eos_create_scene(C.byref(d), C.byref(h)) == E.EOS_ERR_LIMIT
    d = E.SceneDesc(0, None, None, None, 0, 0, None, 0, None)
    assert E.lib().eos_create_scene(C.byref(d), C.byref(h)) == E.EOS_ERR_INVALID


def test_reference_api_symbols_exported():
    """The extern "C" launchers and utilities of src/kernelWrapper.h / src/cudaWrapper.h."""
    lib = E.lib()
    for n in ["rayTrace_MegaKernel", "rayTrace_WarpShuffle_MegaKernel", "rayTrace_MultiMegaKernel",
              "resetDevice", "setGLDevice", "gpuSynch", "cudaPreferL1Cache"]:
        assert hasattr(lib, n), n
