"""The C ABI: the library loads without a GPU, exports every function declared in
include/marsmd.h, and the ctypes structure layouts match the C compiler's."""
import ctypes as C
import re
import subprocess
from pathlib import Path

import pytest

from marsmd import _capi as api
from marsmd import configs

ROOT = Path(__file__).resolve().parent.parent
HEADER = ROOT / "include" / "marsmd.h"


def declared_functions():
    text = HEADER.read_text()
    return sorted(set(re.findall(r"^\s*(?:marsmd_status|const char\*|uint64_t)\s+(marsmd_\w+)\s*\(", text, re.M)))


def test_every_declared_symbol_is_exported():
    lib = api.lib()
    names = declared_functions()
    assert len(names) >= 40
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert set(names) == set(api.PROTOTYPES), set(names) ^ set(api.PROTOTYPES)


def test_version_and_error_strings_without_gpu():
    lib = api.lib()
    assert lib.marsmd_version().decode().startswith("marsmd-hip")
    assert lib.marsmd_last_error() is not None


def test_null_arguments_are_rejected_without_touching_the_device():
    with pytest.raises(api.MarsError):
        api.call("marsmd_zorder_create", 10, None)
    with pytest.raises(api.MarsError):
        api.call("marsmd_pairlist_get_num_pairs", None, None)
    with pytest.raises(api.MarsError):
        api.call("marsmd_patch_run", None, None, 0, 1, None)


def test_radix_sort_rejects_overlapping_buffers_without_touching_the_device():
    # the checks run on the host before any launch; the addresses are never dereferenced
    n = 1000
    nbytes = api.lib().marsmd_radix_sort_scratch_bytes(n)
    a, b, c, d, scratch = (0x10000000 * k for k in range(1, 6))
    for keys_in, vals_in, keys_out, vals_out in [(a, b, a, d), (a, b, c, b), (a, b, c, c), (a, b, a + 4 * (n - 1), d)]:
        with pytest.raises(api.MarsError, match="overlap"):
            api.call("marsmd_radix_sort_pairs", keys_in, vals_in, keys_out, vals_out, n, 30, scratch, nbytes, None)


STRUCTS = {
    "marsmd_box": api.Box,
    "marsmd_table": api.Table,
    "marsmd_exclusion_view": api.ExclusionView,
    "marsmd_grid_term": api.GridTerm,
    "marsmd_grid_view": api.GridView,
    "marsmd_particle_types": api.ParticleTypes,
    "marsmd_particle_view": api.ParticleView,
    "marsmd_pairlist_stats": api.PairlistStats,
    "marsmd_patch_desc": api.PatchDesc,
    "marsmd_run_params": api.RunParams,
    "marsmd_fm_desc": api.FMDesc,
    "marsmd_batch_desc": api.BatchDesc,
    "marsmd_batch_launch": api.BatchLaunch,
    "marsmd_batch_cells": api.BatchCells,
    "marsmd_patch_launch": api.PatchLaunch,
    "marsmd_pair_params": api.PairParams,
}


def test_struct_layouts_match_c(tmp_path):
    lines = ['#include <stdio.h>', '#include <stddef.h>', f'#include "{HEADER}"', "int main(void) {"]
    for cname, py in STRUCTS.items():
        lines.append(f'printf("{cname} %zu\\n", sizeof({cname}));')
        for fname, _ in py._fields_:
            lines.append(f'printf("{cname}.{fname} %zu\\n", offsetof({cname}, {fname}));')
    lines.append("return 0; }")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c11", "-o", str(exe), str(src)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n")
    got = dict(l.rsplit(" ", 1) for l in out if l)
    for cname, py in STRUCTS.items():
        assert int(got[cname]) == C.sizeof(py), cname
        for fname, _ in py._fields_:
            assert int(got[f"{cname}.{fname}"]) == getattr(py, fname).offset, f"{cname}.{fname}"


def test_cpp_interface_compiles(tmp_path):
    """mars-md_amd/host/MARS.hpp (the reference's MARS:: classes over the C ABI)
    compiles with the host compiler alone, and its enums keep the reference's values
    (IntegratorType SimParam.h:160, InterpolationOrder GridTerm.h:36)."""
    src = tmp_path / "t.cpp"
    src.write_text(
        '#include "MARS.hpp"\n'
        "static_assert(int(MARS::IntegratorType::Langevin) == 0, \"\");\n"
        "static_assert(int(MARS::IntegratorType::Brownian) == 1, \"\");\n"
        "static_assert(int(MARS::InterpolationOrder::Linear) == MARSMD_INTERP_LINEAR, \"\");\n"
        "static_assert(int(MARS::InterpolationOrder::Cubic) == MARSMD_INTERP_CUBIC, \"\");\n"
        "static_assert(int(MARS::ExceptionType::RuntimeError) == 9, \"\");\n"
        "int main() { return 0; }\n")
    r = subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-Wall", "-Wextra", "-Werror",
                        "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include", f"-I{ROOT / 'mars-md_amd' / 'host'}",
                        str(src)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


def test_batch_rejects_replicas_with_different_seeds():
    # checked on the host before the library is called
    from marsmd.batch import Batch

    a = configs.config0(n=100, L=60.0, seed=1)
    b = configs.config0(n=120, L=60.0, seed=2)
    with pytest.raises(ValueError, match="base seed"):
        Batch([a, b])


def test_static_workspace_sort_limits_without_touching_the_device():
    lib = api.lib()
    limit = 1 << 21
    # the workspace holds the allocating entry point's scratch for the largest sort
    assert lib.marsmd_radix_sort_static_scratch_bytes() >= lib.marsmd_radix_sort_scratch_bytes(limit)
    # more keys than the workspace serves: rejected on the host before any device call
    a, b, c, d = (0x10000000 * k for k in range(1, 5))
    with pytest.raises(api.MarsError, match="maximum"):
        api.call("marsmd_radix_sort_pairs_static", a, b, c, d, limit + 1, 30, None)
