"""The C-ABI library loads and exports every symbol include/*.h declares; the
Python mirror fails loudly without a GPU (no CPU fallback). No compute calls."""
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    inc = os.path.join(ROOT, "include")
    for fn in os.listdir(inc):
        if not fn.endswith(".h"):
            continue
        text = open(os.path.join(inc, fn)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^[A-Za-z_][\w \*]*?\b(\w+)\s*\(", text, flags=re.M):
            name = m.group(1)
            if m.group(0).startswith("typedef"):
                continue  # a function-pointer type, not a symbol
            if name not in ("if", "while", "for", "sizeof", "typedef"):
                names.add(name)
    return names


def test_header_declares_reference_entry_points():
    names = declared_functions()
    reference = {
        "initialise_ocl_", "initialise_chunk_kernel_ocl_", "generate_chunk_kernel_ocl_",
        "set_field_kernel_ocl_", "field_summary_kernel_ocl_", "update_halo_kernel_ocl_",
        "tea_leaf_common_init_kernel_ocl_", "tea_leaf_calc_residual_ocl_",
        "tea_leaf_calc_2norm_kernel_ocl_", "tea_leaf_common_finalise_kernel_ocl_",
        "tea_leaf_cg_init_kernel_ocl_", "tea_leaf_cg_calc_w_kernel_ocl_",
        "tea_leaf_cg_calc_ur_kernel_ocl_", "tea_leaf_cg_calc_p_kernel_ocl_",
        "tea_leaf_cheby_init_kernel_ocl_", "tea_leaf_cheby_iterate_kernel_ocl_",
        "tea_leaf_ppcg_init_ocl_", "tea_leaf_ppcg_init_sd_kernel_ocl_",
        "tea_leaf_ppcg_inner_kernel_ocl_", "tea_leaf_jacobi_solve_kernel_ocl_",
        "ocl_pack_buffers_", "ocl_unpack_buffers_", "print_opencl_profiling_info_",
    }
    assert reference <= names


def test_library_exports_every_declared_symbol(hiplib):
    from tealeaf_opencl_amd.build import lib_path
    out = subprocess.run(["nm", "-D", "--defined-only", lib_path()], stdout=subprocess.PIPE,
                         text=True, check=True).stdout
    exported = {l.split()[-1] for l in out.splitlines() if l.strip()}
    missing = declared_functions() - exported
    assert not missing, sorted(missing)
    lib = hiplib.load_library()
    for name in declared_functions():
        getattr(lib, name)


def test_no_gpu_means_loud_failure(hiplib):
    if hiplib.device_count() > 0:
        pytest.skip("a GPU is present")
    with pytest.raises(hiplib.TealeafError):
        hiplib.Chunk(16, 16)


def test_python_constants_match_header(hiplib):
    text = open(os.path.join(ROOT, "include", "tealeaf_mi355x.h")).read()
    m = re.search(r"TL_NOALLOC_MAX_STATES\s*=\s*(\d+)", text)
    assert m and int(m.group(1)) == hiplib.NOALLOC_MAX_STATES
