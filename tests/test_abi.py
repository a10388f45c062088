"""The C ABI: libfenix.so loads without a GPU, exports every function that
include/fenix.h declares, and the ctypes mirror matches the C struct layouts."""
import ctypes as C
import os
import re
import subprocess
import tempfile

import pytest

from fenix_amd import abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "fenix.h")


def declared_functions():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\s*\*?\s*(fx_\w+)\s*\(", text, re.M)))


def test_header_declares_entry_points():
    names = declared_functions()
    assert "fx_create" in names and "fx_step" in names and "fx_raymarch" in names
    assert len(names) >= 28


def test_library_exports_every_declared_symbol():
    lib = abi.load()
    for name in declared_functions():
        assert hasattr(lib, name), f"libfenix.so does not export {name}"
    assert set(declared_functions()) == set(abi.FUNCTIONS), "ctypes mirror and header disagree"


STRUCTS = {
    "fx_quality": (abi.Quality, []),
    "fx_runtime_params": (abi.RuntimeParams, ["world_size", "vorticity_strength"]),
    "fx_sim_params": (abi.SimParams, ["dx", "vorticity_mask"]),
    "fx_vcycle": (abi.VCycle, ["omega"]),
    "fx_create_desc": (abi.CreateDesc, ["runtime", "quality", "domain_mode", "vcycle", "ignition_lo", "device"]),
    "fx_camera": (abi.Camera, ["aspect", "tan_half_fov_y"]),
    "fx_light": (abi.Light, ["intensity"]),
    "fx_render_params": (abi.RenderParams, ["lights", "scattering_back", "elapsed_s"]),
    "fx_diff_desc": (abi.DiffDesc, ["slab", "dt", "confinement", "vcycle", "device"]),
    "fx_bloom_params": (abi.BloomParams, ["threshold", "radius"]),
}


def test_struct_layouts_match_c():
    src = ['#include <stdio.h>', '#include <stddef.h>', f'#include "{HEADER}"', 'int main(void) {']
    for cname, (_, fields) in STRUCTS.items():
        src.append(f'printf("{cname} %zu\\n", sizeof({cname}));')
        for f in fields:
            src.append(f'printf("{cname}.{f} %zu\\n", offsetof({cname}, {f}));')
    src.append("return 0; }")
    with tempfile.TemporaryDirectory() as d:
        cfile = os.path.join(d, "layout.c")
        exe = os.path.join(d, "layout")
        open(cfile, "w").write("\n".join(src))
        subprocess.run(["gcc", "-o", exe, cfile], check=True)
        out = subprocess.run([exe], check=True, capture_output=True, text=True).stdout
    got = dict(line.split() for line in out.strip().splitlines())
    for cname, (cls, fields) in STRUCTS.items():
        assert int(got[cname]) == C.sizeof(cls), cname
        for f in fields:
            assert int(got[f"{cname}.{f}"]) == getattr(cls, f).offset, f"{cname}.{f}"


def test_sim_params_is_the_80_byte_uniform():
    # simulation-params.wgsl.ts:2-14, ORCH:57-61 (80-byte uniform)
    assert C.sizeof(abi.SimParams) == 80


def test_default_desc_matches_python_defaults():
    from fenix_amd.simulation import make_desc
    d_c = abi.default_desc()
    d_py = make_desc()
    assert bytes(d_c) == bytes(d_py)


def test_create_without_gpu_fails_loudly_or_rejects_bad_args():
    lib = abi.load()
    h = C.c_void_p()
    d = abi.default_desc(lib)
    d.width = 2  # below the minimum resolution
    rc = lib.fx_create(C.byref(d), C.byref(h))
    assert rc == abi.FX_ERR_INVALID_ARG
    assert b"resolution" in lib.fx_last_error(None)
    assert lib.fx_create(None, C.byref(h)) == abi.FX_ERR_INVALID_ARG


def test_null_handle_calls_return_errors():
    lib = abi.load()
    assert lib.fx_step(None, 0.0, 1 / 60) == abi.FX_ERR_INVALID_ARG
    assert lib.fx_reset(None) == abi.FX_ERR_INVALID_ARG
    assert lib.fx_synchronize(None) == abi.FX_ERR_INVALID_ARG
    lib.fx_destroy(None)  # accepted


@pytest.mark.parametrize("path", ["oracle", "product"])
def test_product_does_not_reference_the_oracle(path):
    """The shipped library and host never link or import the oracle."""
    if path == "product":
        for name in os.listdir(os.path.join(ROOT, "fenix_amd")):
            if name.endswith(".py"):
                text = open(os.path.join(ROOT, "fenix_amd", name)).read()
                assert "oracle" not in text.lower() or name == "build.py", name
        out = subprocess.run(["nm", "-D", abi.LIB_PATH], capture_output=True, text=True).stdout
        assert "orc_" not in out
    else:
        assert os.path.exists(os.path.join(ROOT, "oracle", "oracle.cpp"))


def test_diff_create_rejects_bad_args_without_a_gpu():
    lib = abi.load()
    from fenix_amd.differentiable import diff_desc
    h = C.c_void_p()
    d = diff_desc(lib=lib)
    d.vcycle.post_sweeps = d.vcycle.pre_sweeps + 1  # the adjoint projection needs a symmetric V-cycle
    assert lib.fx_diff_create(C.byref(d), C.byref(h)) == abi.FX_ERR_INVALID_ARG
    assert b"symmetric" in lib.fx_diff_last_error(None)
    d = diff_desc(lib=lib, slab=0)
    assert lib.fx_diff_create(C.byref(d), C.byref(h)) == abi.FX_ERR_INVALID_ARG
    assert lib.fx_diff_create(None, C.byref(h)) == abi.FX_ERR_INVALID_ARG
    lib.fx_diff_destroy(None)  # accepted
