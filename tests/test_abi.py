"""C-ABI boundary checks that need no GPU: the library loads, exports every
symbol include/sunny_geneig.h declares, keeps the reference's enum values,
and handles NULL / invalid arguments the way the reference's
tests/common/test_c_api.cpp:89-120 expects. hipRTC compilation into the
on-disk cache is exercised without a device."""
import ctypes
import os
import re
import subprocess
import sys

import pytest

import sunny_gen_eigensolver_amd as sg

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(REPO, "include", "sunny_geneig.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(sunny_geneig_[a-z_0-9]+)\s*\(", text)))


def test_header_declares_reference_symbols():
    ref = {"sunny_geneig_create", "sunny_geneig_create_ex", "sunny_geneig_destroy",
           "sunny_geneig_launch", "sunny_geneig_matrix_size", "sunny_geneig_batches_per_block",
           "sunny_geneig_block_dim_x", "sunny_geneig_shared_mem_bytes", "sunny_geneig_last_error",
           "sunny_geneig_path_name", "sunny_geneig_kernel_mode"}
    assert ref <= set(declared_functions())


def test_library_exports_every_declared_symbol():
    L = sg.lib()
    missing = [f for f in declared_functions() if not hasattr(L, f)]
    assert not missing, missing
    assert set(declared_functions()) == set(sg.EXPORTED_SYMBOLS)


def test_enum_values_match_reference():
    text = open(HEADER).read()
    vals = dict((k, int(v)) for k, v in re.findall(r"(SUNNY_GENEIG_[A-Z_]+)\s*=\s*(-?\d+)", text))
    assert vals == {
        "SUNNY_GENEIG_OK": 0, "SUNNY_GENEIG_INVALID_ARG": -1, "SUNNY_GENEIG_DEVICE_ERROR": -2,
        "SUNNY_GENEIG_COMPILE_ERROR": -3, "SUNNY_GENEIG_LAUNCH_ERROR": -4,
        "SUNNY_GENEIG_CACHE_AUTO": 0, "SUNNY_GENEIG_CACHE_NODISK": 1, "SUNNY_GENEIG_CACHE_NOCACHE": 2,
        "SUNNY_GENEIG_TUNING_SUGGESTED": 0, "SUNNY_GENEIG_TUNING_AUTOTUNE": 1,
        "SUNNY_GENEIG_KERNEL_UNIFIED": 0, "SUNNY_GENEIG_KERNEL_HYBRID": 1,
        # additive: no counterpart in the reference
        "SUNNY_GENEIG_KERNEL_MIXED": 2}


def test_null_handle_accessors():
    L = sg.lib()
    assert L.sunny_geneig_matrix_size(None) == -1
    assert L.sunny_geneig_batches_per_block(None) == -1
    assert L.sunny_geneig_block_dim_x(None) == -1
    assert L.sunny_geneig_shared_mem_bytes(None) == 0
    assert L.sunny_geneig_path_name(None) == b""
    assert L.sunny_geneig_kernel_mode(None) == -1
    assert L.sunny_geneig_cache_layer(None) == b""
    assert L.sunny_geneig_set_stage_timing(None, 1) == sg.STATUS_INVALID_ARG
    assert L.sunny_geneig_stage_times(None, None, None, 0) == -1
    assert L.sunny_geneig_set_graphs(None, 1) == sg.STATUS_INVALID_ARG
    assert L.sunny_geneig_graph_stats(None, None, None) == sg.STATUS_INVALID_ARG
    L.sunny_geneig_destroy(None)


def test_chunked_launch_null_and_bad_arguments():
    L = sg.lib()
    one = (ctypes.c_uint64 * 1)(0)
    assert L.sunny_geneig_launch_chunked(None, 1, 1, 1, 1, 1, 8, 4, one, 1) == sg.STATUS_INVALID_ARG
    assert "NULL" in sg.last_error()


def test_null_launch_is_invalid_arg_with_message():
    L = sg.lib()
    st = L.sunny_geneig_launch(None, 0, 0, 0, 0, 0, 8, 0)
    assert st == sg.STATUS_INVALID_ARG
    assert len(sg.last_error()) > 0


def test_create_rejects_bad_arguments():
    L = sg.lib()
    for n, kernel in ((0, 0), (-3, 0), (54, 7), (54, 3)):
        h = L.sunny_geneig_create_ex(n, 0, 0, 0, 0, kernel)
        assert not h
        assert len(sg.last_error()) > 0


def test_precompile_writes_code_object(tmp_path):
    """hipRTC compiles offline; the cache file is an ELF code object."""
    env = dict(os.environ, SUNNY_GENEIG_CACHE_DIR=str(tmp_path))
    code = ("import sunny_gen_eigensolver_amd as sg; sg.precompile(8, 64, 'gfx950'); "
            "print(sg.cache_dir())")
    out = subprocess.run([sys.executable, "-c", code], env=env, cwd=REPO, check=True,
                         capture_output=True, text=True)
    assert out.stdout.strip() == str(tmp_path)
    files = sorted(f for f in os.listdir(tmp_path) if f.endswith(".hsaco"))
    assert len(files) == 2, files
    assert files[0].startswith("geneig_n8_wg64_gfx950_") and files[1].startswith("geneig_n8_wg64_gfx950_fp32_")
    with open(os.path.join(tmp_path, files[0]), "rb") as f:
        assert f.read(4) == b"\x7fELF"


def test_precompile_rejects_out_of_range_order():
    with pytest.raises(sg.GeneigError):
        sg.precompile(0)


def test_product_path_never_touches_the_oracle():
    pkg = os.path.join(REPO, "sunny_gen_eigensolver_amd")
    for root, _, files in os.walk(pkg):
        for f in files:
            if f.endswith((".py", ".cpp", ".hpp", ".hip", ".h")):
                text = open(os.path.join(root, f)).read()
                assert "oracle" not in text.replace("# oracle-free", ""), f


def test_varlen_null_and_bad_arguments():
    L = sg.lib()
    assert L.sunny_geneig_varlen_num_buckets(None) == -1
    assert L.sunny_geneig_varlen_bucket_order(None, 0) == -1
    L.sunny_geneig_varlen_destroy(None)
    orders = (ctypes.c_int * 1)(8)
    st = L.sunny_geneig_varlen_launch(None, 0, 0, 0, 0, 0, orders, 1, 0)
    assert st == sg.STATUS_INVALID_ARG and "NULL" in sg.last_error()
    for max_n, buckets in ((0, None), (65, None), (16, [8, 4, 16]), (16, [4, 8]), (16, [0, 16])):
        arr = (ctypes.c_int * len(buckets))(*buckets) if buckets else None
        h = L.sunny_geneig_varlen_create(max_n, 0, arr, len(buckets) if buckets else 0, 0)
        assert not h
        assert "sunny_geneig_varlen_create" in sg.last_error()


def test_session_null_and_bad_arguments():
    L = sg.lib()
    L.sunny_geneig_session_destroy(None)
    assert L.sunny_geneig_session_submit(None, 8, 1, 0, 0, 0, 0, 0, 1, 0) == sg.STATUS_INVALID_ARG
    assert L.sunny_geneig_session_set_policy(None, 0, 0) == sg.STATUS_INVALID_ARG
    assert L.sunny_geneig_session_poll(None, 0) == sg.STATUS_INVALID_ARG
    assert L.sunny_geneig_session_wait(None, 0, None, None) == sg.STATUS_INVALID_ARG
    assert L.sunny_geneig_session_stats(None, None, None, None, None) == sg.STATUS_INVALID_ARG
    h = L.sunny_geneig_session_create(0, 0, None, 0, 16, 2, 1 << 20, 0)
    assert not h and "orders" in sg.last_error()
    orders = (ctypes.c_int * 1)(200)
    h = L.sunny_geneig_session_create(0, 0, orders, 1, 16, 2, 1 << 20, 0)
    assert not h and "outside [1, 192]" in sg.last_error()


def test_session_request_bytes():
    """H, S and U (16 n^2 each) plus W (8 n) and info (4), each 256-aligned."""
    def al(x):
        return (x + 255) // 256 * 256
    for n, c in ((54, 1), (24, 100), (54, 4096)):
        want = 3 * al(16 * n * n * c) + al(8 * n * c) + al(4 * c)
        assert sg.ServingSession.request_bytes(n, c) == want
    assert sg.ServingSession.request_bytes(0, 5) == 0
