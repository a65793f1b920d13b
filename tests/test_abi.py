"""The C-ABI library loads and exports every symbol include/*.h declares,
with the reference's by-value struct layouts.  No compute calls (CPU only)."""
import ctypes
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "zeknox_amd", "libzeknox.so")
HEADERS = [os.path.join(ROOT, "include", h) for h in ("zeknox.h", "zeknox_ext.h")]

# native/lib.h:21-51, lib.cu:137-155, all_gpus.cu:68-71 and merkle/merkle.h:30-62
REFERENCE_SYMBOLS = [
    "get_number_of_gpus", "list_devices_info", "init_cuda", "init_cuda_degree", "init_cuda_degree_and_generator",
    "init_twiddle_factors", "init_coset", "compute_batched_ntt", "compute_batched_lde",
    "compute_batched_lde_multi_gpu", "compute_transpose_rev", "fill_digests_buf_linear_gpu_with_gpu_ptr",
    "fill_digests_buf_linear_multigpu_with_gpu_ptr", "fill_digests_buf_linear_cpu", "cuda_available",
]


def declared_functions(path):
    text = open(path).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"^(?:ZK_EXTERN\s+)?[A-Za-z_][\w\s\*]*?\b([a-z_][a-z0-9_]*)\s*\(", text, flags=re.M)
    return {n for n in names if n not in ("if", "sizeof")}


@pytest.fixture(scope="module")
def exported():
    if not os.path.exists(LIB):
        pytest.fail("libzeknox.so is not built (run __graft_entry__.build())")
    out = subprocess.run(["nm", "-D", "--defined-only", LIB], capture_output=True, text=True, check=True).stdout
    return {ln.split()[-1] for ln in out.splitlines() if " T " in ln}


def test_library_loads():
    ctypes.CDLL(LIB)


def test_reference_symbols_exported(exported):
    missing = [s for s in REFERENCE_SYMBOLS if s not in exported]
    assert not missing, missing


@pytest.mark.parametrize("header", HEADERS)
def test_every_declared_function_is_exported(exported, header):
    declared = declared_functions(header)
    assert declared, header
    missing = sorted(declared - exported)
    assert not missing, missing


def test_struct_layouts():
    import zeknox_amd as zk
    # ntt.h: u32, enum, enum, u32, 4 x bool, u32 -> 24 bytes
    assert ctypes.sizeof(zk.NTTConfig) == 24
    assert zk.NTTConfig.salt_size.offset == 20
    assert zk.NTTConfig.are_inputs_on_device.offset == 16
    assert ctypes.sizeof(zk.TransposeConfig) == 8
    assert ctypes.sizeof(zk.RustError) == 16
    assert ctypes.sizeof(zk.CommitJob) == 56


def test_invalid_arguments_without_device():
    """Argument errors are reported before any device work (code 1, as the
    reference's cudaErrorInvalidValue)."""
    import zeknox_amd as zk
    out = (ctypes.c_uint64 * 8)()
    inp = (ctypes.c_uint64 * 8)()
    cfg = zk.NTTConfig(batches=1, extension_rate_bits=0)
    err = zk.lib.compute_batched_lde(0, ctypes.addressof(out), ctypes.addressof(inp), 3, 0, cfg)
    assert err.code == 1
    err = zk.lib.compute_batched_ntt(0, ctypes.addressof(out), 0, 0, zk.NTTConfig())
    assert err.code == 1


# ---- NTT pass planner (host-only logic) --------------------------------------

zk = pytest.importorskip("zeknox_amd")

def _check_plan(lg, sizes, smax):
    assert sum(sizes) == lg
    assert sizes[-1] <= 12
    below = lg
    for m in sizes[:-1]:  # strided passes
        below -= m
        assert 1 <= m <= min(smax, 9)
        assert below >= 12 - m  # stride at least the tile's column count


@pytest.mark.parametrize("even", [0, 1])
@pytest.mark.parametrize("smax", [1, 3, 4, 6, 8, 9])
def test_ntt_pass_plans_are_valid(even, smax):
    for lg in range(0, 33):
        sizes = zk.ntt_pass_sizes(lg, zk.Tuning(smax, even, 8))
        _check_plan(lg, sizes, smax)


def test_ntt_pass_plan_defaults():
    # greedy: strided passes of up to 9 stages, the last up to 12
    assert zk.ntt_pass_sizes(10) == [10]
    assert zk.ntt_pass_sizes(15) == [3, 12]
    assert zk.ntt_pass_sizes(21) == [9, 12]
    assert zk.ntt_pass_sizes(24) == [9, 3, 12]
    # even: the same number of passes, stages spread
    assert zk.ntt_pass_sizes(15, zk.Tuning(9, 1, 8)) == [8, 7]
    assert zk.ntt_pass_sizes(24, zk.Tuning(9, 1, 8)) == [8, 8, 8]


def test_tuning_out_of_range_rejected():
    with pytest.raises(zk.ZeknoxError):
        zk.ntt_pass_sizes(20, zk.Tuning(10, 0, 8))
    with pytest.raises(zk.ZeknoxError):
        zk.ntt_pass_sizes(20, zk.Tuning(9, 2, 8))


def test_init_cuda_degree_reports_bad_degree():
    """init_cuda_degree returns nothing (lib.h:28): an invalid degree is
    reported through zkx_last_error instead of being ignored."""
    import zeknox_amd as zk
    with pytest.raises(zk.ZeknoxError):
        zk.init_cuda_degree_rs(40)
    assert "max_degree" in zk.last_error()
