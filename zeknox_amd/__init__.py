"""Python binding of libzeknox (the MI355X LDE + Merkle-commit library).

Mirrors the reference's Rust wrapper (wrappers/rust/src/lib.rs:99-260): the
same function names and argument meaning, with failures raised as
ZeknoxError where the Rust wrapper panics.  Buffers are numpy uint64 arrays
(host) or DeviceBuffer objects (device memory owned by the library's HIP
runtime, like the wrapper's cudart-backed device slices).

The library must have been built (zeknox_amd.build.build()); importing this
module loads it and fails loudly if it is missing -- there is no CPU
fallback.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libzeknox.so")

P = 0xFFFFFFFF00000001
HASH_POSEIDON = 0
HASH_KECCAK = 1
HASH_POSEIDON_BN128 = 2
HASH_POSEIDON2 = 3
HASH_MONOLITH = 4
FORWARD, INVERSE = 0, 1


class ZeknoxError(RuntimeError):
    def __init__(self, code, message):
        super().__init__(f"zeknox error {code}: {message}")
        self.code = code
        self.message = message


class RustError(ctypes.Structure):
    """native/utils/rusterror.h:15-41 -- {int code; char *message}."""
    _fields_ = [("code", ctypes.c_int), ("message", ctypes.c_void_p)]


class NTTConfig(ctypes.Structure):
    """native/ntt/ntt.h:43-55 (24 bytes, passed by value)."""
    _fields_ = [
        ("batches", ctypes.c_uint32),
        ("order", ctypes.c_int),
        ("ntt_type", ctypes.c_int),
        ("extension_rate_bits", ctypes.c_uint32),
        ("are_inputs_on_device", ctypes.c_bool),
        ("are_outputs_on_device", ctypes.c_bool),
        ("with_coset", ctypes.c_bool),
        ("is_multi_gpu", ctypes.c_bool),
        ("salt_size", ctypes.c_uint32),
    ]

    def __init__(self, batches=1, extension_rate_bits=0, are_inputs_on_device=False,
                 are_outputs_on_device=False, with_coset=False, salt_size=0):
        super().__init__(batches, 0, 0, extension_rate_bits, are_inputs_on_device,
                         are_outputs_on_device, with_coset, False, salt_size)


class TransposeConfig(ctypes.Structure):
    """native/ntt/ntt.h:57-62."""
    _fields_ = [
        ("batches", ctypes.c_uint32),
        ("are_inputs_on_device", ctypes.c_bool),
        ("are_outputs_on_device", ctypes.c_bool),
    ]

    def __init__(self, batches=1, are_inputs_on_device=False, are_outputs_on_device=False):
        super().__init__(batches, are_inputs_on_device, are_outputs_on_device)


class CommitJob(ctypes.Structure):
    """zeknox_ext.h ZkxCommitJob (device pointers)."""
    _fields_ = [
        ("coeffs", ctypes.c_void_p),
        ("leaves", ctypes.c_void_p),
        ("digests", ctypes.c_void_p),
        ("caps", ctypes.c_void_p),
        ("cols", ctypes.c_uint32),
        ("lg_n", ctypes.c_uint32),
        ("rate", ctypes.c_uint32),
        ("cap_height", ctypes.c_uint32),
        ("hash_type", ctypes.c_uint32),
        ("reserved", ctypes.c_uint32),
    ]


class Tuning(ctypes.Structure):
    """zeknox_ext.h ZkxTuning: schedule parameters of a commit shape."""
    _fields_ = [("ntt_strided_max", ctypes.c_uint32), ("ntt_even", ctypes.c_uint32),
                ("merkle_tile_log", ctypes.c_uint32), ("merkle_valu", ctypes.c_uint32)]

    def __init__(self, ntt_strided_max=9, ntt_even=0, merkle_tile_log=8, merkle_valu=0):
        super().__init__(ntt_strided_max, ntt_even, merkle_tile_log, merkle_valu)

    def astuple(self):
        return (self.ntt_strided_max, self.ntt_even, self.merkle_tile_log, self.merkle_valu)


class IssueRates(ctypes.Structure):
    """zeknox_ext.h ZkxIssueRates."""
    _fields_ = [("valu_wave_insts_per_s", ctypes.c_double), ("mfma_wave_insts_per_s", ctypes.c_double),
                ("mfma_issue_slots", ctypes.c_double), ("ms_valu", ctypes.c_double), ("ms_mfma", ctypes.c_double),
                ("ms_mixed", ctypes.c_double), ("simds", ctypes.c_uint32), ("waves_per_simd", ctypes.c_uint32),
                ("mads_per_mfma", ctypes.c_uint32), ("reserved", ctypes.c_uint32)]


class ServeResult(ctypes.Structure):
    """zeknox_ext.h ZkxServeResult."""
    _fields_ = [("leaves", ctypes.c_void_p), ("digests", ctypes.c_void_p), ("caps", ctypes.c_void_p),
                ("t_submit", ctypes.c_double), ("t_start", ctypes.c_double), ("t_done", ctypes.c_double),
                ("state", ctypes.c_int), ("reserved", ctypes.c_int), ("t_enqueued", ctypes.c_double)]


class ServeStats(ctypes.Structure):
    """zeknox_ext.h ZkxServeStats."""
    _fields_ = [("submitted", ctypes.c_uint64), ("completed", ctypes.c_uint64), ("failed", ctypes.c_uint64),
                ("batches", ctypes.c_uint64), ("max_batch", ctypes.c_uint64), ("bytes_in_use", ctypes.c_size_t),
                ("peak_bytes", ctypes.c_size_t)]


def _load():
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is not built; run zeknox_amd.build.build() (hipcc, gfx950)")
    lib = ctypes.CDLL(LIB_PATH)
    u64, sz, u32, vp = ctypes.c_uint64, ctypes.c_size_t, ctypes.c_uint32, ctypes.c_void_p
    sig = {
        "get_number_of_gpus": (RustError, [ctypes.POINTER(sz)]),
        "list_devices_info": (RustError, []),
        "init_cuda": (None, []),
        "init_cuda_degree": (None, [u32]),
        "init_cuda_degree_and_generator": (None, [u32, u64]),
        "init_twiddle_factors": (RustError, [sz, sz]),
        "init_coset": (RustError, [sz, sz, u64]),
        "compute_batched_ntt": (RustError, [sz, vp, u32, ctypes.c_int, NTTConfig]),
        "compute_batched_lde": (RustError, [sz, vp, vp, u32, ctypes.c_int, NTTConfig]),
        "compute_batched_lde_multi_gpu": (RustError, [vp, vp, u32, ctypes.c_int, NTTConfig, u32, sz, sz]),
        "compute_transpose_rev": (RustError, [sz, vp, vp, u32, TransposeConfig]),
        "fill_digests_buf_linear_gpu_with_gpu_ptr": (None, [vp, vp, vp, u64, u64, u64, u64, u64, u64, u64]),
        "fill_digests_buf_linear_multigpu_with_gpu_ptr": (None, [vp, vp, vp, u64, u64, u64, u64, u64, u64]),
        "fill_digests_buf_linear_cpu": (None, [vp, vp, vp, u64, u64, u64, u64, u64, u64]),
        "zkx_fill_digests_buf_linear_gpu_noalloc": (RustError, [vp, vp, vp, u64, u64, u64, u64, u64, u64, u64]),
        "cuda_available": (ctypes.c_bool, []),
        "zkx_last_error": (ctypes.c_char_p, []),
        "zkx_device_malloc": (RustError, [sz, ctypes.POINTER(vp), sz]),
        "zkx_device_free": (RustError, [sz, vp]),
        "zkx_memcpy_htod": (RustError, [sz, vp, vp, sz]),
        "zkx_memcpy_dtoh": (RustError, [sz, vp, vp, sz]),
        "zkx_memset": (RustError, [sz, vp, ctypes.c_int, sz]),
        "zkx_fill_random": (RustError, [sz, vp, sz, u64]),
        "zkx_gl_ops": (RustError, [sz, ctypes.c_int, vp, vp, vp, sz]),
        "zkx_commit_trim": (RustError, [sz]),
        "zkx_autotune": (RustError, [sz, vp, sz, ctypes.c_int]),
        "zkx_get_tuning": (RustError, [sz, vp, ctypes.POINTER(Tuning)]),
        "zkx_set_tuning": (RustError, [sz, vp, ctypes.POINTER(Tuning)]),
        "zkx_set_autotune": (RustError, [sz, ctypes.c_int]),
        "zkx_set_commit_graphs": (RustError, [sz, ctypes.c_int]),
        "zkx_set_level_launches": (RustError, [sz, u64]),
        "zkx_set_lone_level_launches": (RustError, [sz, u64]),
        "zkx_commit_graphs_captured": (sz, [sz]),
        "zkx_tuning_clear": (RustError, [sz]),
        "zkx_tuning_save": (RustError, [sz, ctypes.c_char_p]),
        "zkx_tuning_load": (RustError, [sz, ctypes.c_char_p]),
        "zkx_ntt_pass_sizes": (ctypes.c_int, [u32, ctypes.POINTER(Tuning), ctypes.POINTER(u32), ctypes.c_int]),
        "zkx_merkle_plan": (ctypes.c_int, [u64, u64, u32, ctypes.POINTER(u64), ctypes.POINTER(u32),
                                           ctypes.POINTER(u32), ctypes.POINTER(u32), ctypes.c_int,
                                           ctypes.POINTER(u64)]),
        "zkx_compute_batched_lde_seeded": (RustError, [sz, vp, vp, u32, ctypes.c_int, NTTConfig, u64]),
        "zkx_device_synchronize": (RustError, [sz]),
        "zkx_issue_probe": (RustError, [sz, u32, ctypes.POINTER(IssueRates)]),
        "zkx_stream": (vp, [sz]),
        "zkx_profile_enable": (RustError, [ctypes.c_int]),
        "zkx_kernel_stats": (ctypes.c_int, [vp, vp, vp, ctypes.c_int]),
        "zkx_kernel_stats_reset": (None, []),
        "zkx_commit_batch": (RustError, [sz, ctypes.POINTER(CommitJob), sz]),
        "zkx_commit_batch_pieces": (RustError, [sz, ctypes.POINTER(CommitJob), sz, ctypes.c_uint64, ctypes.c_uint64]),
        "zkx_commit_scratch_bytes": (sz, [ctypes.POINTER(CommitJob), sz]),
        "zkx_server_create": (RustError, [sz, sz, ctypes.c_int, sz, ctypes.POINTER(vp)]),
        "zkx_server_destroy": (RustError, [vp]),
        "zkx_server_footprint": (sz, [ctypes.POINTER(CommitJob), ctypes.c_int]),
        "zkx_server_submit": (RustError, [vp, ctypes.POINTER(CommitJob), ctypes.c_int, ctypes.POINTER(u64)]),
        "zkx_server_wait": (RustError, [vp, u64, ctypes.POINTER(ServeResult)]),
        "zkx_server_wait_any": (RustError, [vp, ctypes.POINTER(u64), ctypes.POINTER(ServeResult)]),
        "zkx_server_release": (RustError, [vp, u64]),
        "zkx_server_stats": (RustError, [vp, ctypes.POINTER(ServeStats)]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


lib = _load()
_libc = ctypes.CDLL(None)
_libc.free.argtypes = [ctypes.c_void_p]


def _check(err):
    if err.code != 0:
        msg = ctypes.string_at(err.message).decode() if err.message else f"sppark::Error #{err.code}"
        if err.message:
            _libc.free(err.message)
        raise ZeknoxError(err.code, msg)


def last_error():
    return lib.zkx_last_error().decode()


class DeviceBuffer:
    """Device allocation of `count` u64 on `device_id`."""

    def __init__(self, count, device_id=0):
        self.count = int(count)
        self.device_id = device_id
        p = ctypes.c_void_p()
        _check(lib.zkx_device_malloc(device_id, ctypes.byref(p), self.count * 8))
        self.ptr = p.value

    @classmethod
    def from_host(cls, arr, device_id=0):
        a = np.ascontiguousarray(arr, dtype=np.uint64).ravel()
        buf = cls(max(a.size, 1), device_id)
        buf.copy_from_host(a)
        return buf

    def copy_from_host(self, arr):
        a = np.ascontiguousarray(arr, dtype=np.uint64).ravel()
        assert a.size <= self.count
        _check(lib.zkx_memcpy_htod(self.device_id, self.ptr, a.ctypes.data, a.size * 8))

    def to_host(self, count=None):
        n = self.count if count is None else int(count)
        out = np.empty(n, dtype=np.uint64)
        _check(lib.zkx_memcpy_dtoh(self.device_id, out.ctypes.data, self.ptr, n * 8))
        return out

    def zero(self):
        _check(lib.zkx_memset(self.device_id, self.ptr, 0, self.count * 8))

    def fill_random(self, seed, count=None):
        """Uniform canonical field elements generated on the device."""
        n = self.count if count is None else int(count)
        _check(lib.zkx_fill_random(self.device_id, self.ptr, n, seed))

    def free(self):
        if self.ptr:
            _check(lib.zkx_device_free(self.device_id, self.ptr))
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def _addr(x):
    if isinstance(x, DeviceBuffer):
        return x.ptr
    if isinstance(x, np.ndarray):
        assert x.dtype == np.uint64 and x.flags["C_CONTIGUOUS"]
        return x.ctypes.data
    return x  # raw address


# ---- wrappers/rust/src/lib.rs:99-260 mirror -------------------------------

def get_number_of_gpus_rs():
    n = ctypes.c_size_t()
    _check(lib.get_number_of_gpus(ctypes.byref(n)))
    return n.value


def list_devices_info_rs():
    _check(lib.list_devices_info())


def _check_last(what):
    e = last_error()
    if e:
        raise ZeknoxError(1, f"{what}: {e}")


def init_cuda_rs():
    lib.init_cuda()
    _check_last("init_cuda")


def init_cuda_degree_rs(max_degree):
    lib.init_cuda_degree(max_degree)
    _check_last("init_cuda_degree")


def init_cuda_degree_and_generator_rs(max_degree, generator):
    lib.init_cuda_degree_and_generator(max_degree, generator)
    _check_last("init_cuda_degree_and_generator")


def init_twiddle_factors_rs(device_id, lg_n):
    _check(lib.init_twiddle_factors(device_id, lg_n))


def init_coset_rs(device_id, lg_n, coset_gen):
    _check(lib.init_coset(device_id, lg_n, coset_gen))


def lde_batch(device_id, output, input, log_n_size, cfg, direction=FORWARD):
    _check(lib.compute_batched_lde(device_id, _addr(output), _addr(input), log_n_size, direction, cfg))


def lde_batch_seeded(device_id, output, input, log_n_size, cfg, salt_seed, direction=FORWARD):
    """compute_batched_lde whose salt columns come from `salt_seed`
    (zkx_compute_batched_lde_seeded)."""
    _check(lib.zkx_compute_batched_lde_seeded(device_id, _addr(output), _addr(input), log_n_size, direction, cfg,
                                              salt_seed))


def lde_batch_multi_gpu(output, input, num_gpu, cfg, log_n_size, total_in, total_out):
    _check(lib.compute_batched_lde_multi_gpu(_addr(output), _addr(input), num_gpu, FORWARD, cfg, log_n_size,
                                             total_in, total_out))


def ntt_batch(device_id, inout, log_n_size, cfg):
    _check(lib.compute_batched_ntt(device_id, _addr(inout), log_n_size, FORWARD, cfg))


def intt_batch(device_id, inout, log_n_size, cfg):
    _check(lib.compute_batched_ntt(device_id, _addr(inout), log_n_size, INVERSE, cfg))


def transpose_rev_batch(device_id, output, input, log_n_size, cfg):
    _check(lib.compute_transpose_rev(device_id, _addr(output), _addr(input), log_n_size, cfg))


def fill_digests_buf_linear_gpu_with_gpu_ptr(digests, caps, leaves, digests_size, caps_size, leaves_size,
                                             leaf_size, cap_height, hash_type, gpu_id=0):
    """merkle.h:30-40; raises ZeknoxError on the argument errors the
    reference asserts on."""
    lib.fill_digests_buf_linear_gpu_with_gpu_ptr(_addr(digests), _addr(caps), _addr(leaves), digests_size,
                                                 caps_size, leaves_size, leaf_size, cap_height, hash_type, gpu_id)
    e = last_error()
    if e:
        raise ZeknoxError(1, e)


def fill_digests_buf_linear_gpu_noalloc(digests, caps, leaves, digests_size, caps_size, leaves_size,
                                        leaf_size, cap_height, hash_type, gpu_id=0):
    """fill_digests_buf_linear_gpu_with_gpu_ptr without device allocation
    (zkx_fill_digests_buf_linear_gpu_noalloc); raises ZeknoxError."""
    _check(lib.zkx_fill_digests_buf_linear_gpu_noalloc(_addr(digests), _addr(caps), _addr(leaves), digests_size,
                                                       caps_size, leaves_size, leaf_size, cap_height, hash_type,
                                                       gpu_id))


def fill_digests_buf_linear_cpu(digests, caps, leaves, digests_size, caps_size, leaves_size, leaf_size,
                                cap_height, hash_type):
    lib.fill_digests_buf_linear_cpu(_addr(digests), _addr(caps), _addr(leaves), digests_size, caps_size,
                                    leaves_size, leaf_size, cap_height, hash_type)
    e = last_error()
    if e:
        raise ZeknoxError(1, e)


def merkle_tree(leaves, leaf_size, cap_height, hash_type=HASH_POSEIDON, device_id=0):
    """Convenience: host leaves (n_leaves x leaf_size) -> (digests, caps) via
    the device entry point."""
    leaves = np.ascontiguousarray(leaves, dtype=np.uint64).ravel()
    nl = leaves.size // leaf_size
    nc = 1 << cap_height
    nd = 2 * (nl - nc) if nc != nl else 0
    dl = DeviceBuffer.from_host(leaves, device_id)
    dc = DeviceBuffer(nc * 4, device_id)
    dd = DeviceBuffer(max(nd, 1) * 4, device_id)
    fill_digests_buf_linear_gpu_with_gpu_ptr(dd, dc, dl, nd, nc, nl, leaf_size, cap_height, hash_type, device_id)
    return (dd.to_host(nd * 4) if nd else np.zeros(0, np.uint64)), dc.to_host()


# ---- batched commits --------------------------------------------------------

class Commit:
    """Device buffers of one commit job: coefficients in, leaf rows, digests
    and caps out (zeknox_ext.h ZkxCommitJob)."""

    def __init__(self, coeffs, cols, lg_n, rate, cap_height, hash_type=HASH_POSEIDON, leaves=True, device_id=0):
        self.cols, self.lg_n, self.rate = cols, lg_n, rate
        self.cap_height, self.hash_type = cap_height, hash_type
        self.N = 1 << (lg_n + rate)
        self.num_caps = 1 << cap_height
        self.num_digests = 0 if self.num_caps == self.N else 2 * (self.N - self.num_caps)
        self.d_coeffs = coeffs if isinstance(coeffs, DeviceBuffer) else DeviceBuffer.from_host(coeffs, device_id)
        self.d_leaves = DeviceBuffer(cols * self.N, device_id) if leaves else None
        self.d_digests = DeviceBuffer(max(self.num_digests, 1) * 4, device_id)
        self.d_caps = DeviceBuffer(self.num_caps * 4, device_id)

    def job(self):
        return CommitJob(self.d_coeffs.ptr, self.d_leaves.ptr if self.d_leaves else None, self.d_digests.ptr,
                         self.d_caps.ptr, self.cols, self.lg_n, self.rate, self.cap_height, self.hash_type, 0)

    def results(self):
        leaves = self.d_leaves.to_host() if self.d_leaves else None
        dig = self.d_digests.to_host(self.num_digests * 4) if self.num_digests else np.zeros(0, np.uint64)
        return leaves, dig, self.d_caps.to_host()


GL_MUL_LAZY, GL_MUL, GL_REDUCE128, GL_REDUCE96, GL_POW7, GL_ADD, GL_SUB, GL_CANON = range(8)


def gl_ops(op, a, b, device_id=0):
    """Elementwise Goldilocks op on the device (zeknox_ext.h zkx_gl_ops):
    numpy uint64 arrays in, canonical numpy uint64 array out."""
    a = np.ascontiguousarray(a, dtype=np.uint64)
    b = np.ascontiguousarray(b, dtype=np.uint64)
    assert a.shape == b.shape
    da, db = DeviceBuffer.from_host(a, device_id), DeviceBuffer.from_host(b, device_id)
    out = DeviceBuffer(a.size, device_id)
    _check(lib.zkx_gl_ops(device_id, op, da.ptr, db.ptr, out.ptr, a.size))
    return out.to_host()


def commit_batch(commits, device_id=0, ntt_tiles=0, merkle_groups=0):
    """Run a list of Commit objects as one batch (synchronous); ntt_tiles /
    merkle_groups > 0 split the long launches into ranges of that size
    (zkx_commit_batch_pieces)."""
    arr = (CommitJob * len(commits))(*[c.job() for c in commits])
    if ntt_tiles or merkle_groups:
        _check(lib.zkx_commit_batch_pieces(device_id, arr, len(commits), ntt_tiles, merkle_groups))
    else:
        _check(lib.zkx_commit_batch(device_id, arr, len(commits)))


def commit_scratch_bytes(commits):
    arr = (CommitJob * len(commits))(*[c.job() for c in commits])
    return lib.zkx_commit_scratch_bytes(arr, len(commits))


# ---- schedule autotuning ---------------------------------------------------------

def autotune(commits, iters=3, device_id=0):
    """Tune every distinct shape among the Commit objects (their outputs are
    written with their results)."""
    arr = (CommitJob * len(commits))(*[c.job() for c in commits])
    _check(lib.zkx_autotune(device_id, arr, len(commits), iters))


def get_tuning(commit=None, device_id=0):
    t = Tuning()
    job = ctypes.byref(commit.job()) if commit is not None else None
    _check(lib.zkx_get_tuning(device_id, job, ctypes.byref(t)))
    return t


def set_tuning(tuning, commit=None, device_id=0):
    """Schedule of the commit's shape, or the default (commit None)."""
    job = ctypes.byref(commit.job()) if commit is not None else None
    _check(lib.zkx_set_tuning(device_id, job, ctypes.byref(tuning)))


def set_autotune(on, device_id=0):
    _check(lib.zkx_set_autotune(device_id, int(bool(on))))


def set_commit_graphs(on, device_id=0):
    """Replay repeated batch launch lists as captured HIP graphs (default on)."""
    _check(lib.zkx_set_commit_graphs(device_id, 1 if on else 0))


LEVEL_LAUNCHES_DEFAULT = 1 << 18


def set_level_launches(min_leaves, device_id=0):
    """Batches whose trees hold at least min_leaves leaves build the tree
    levels with one launch per level (0: always, None: never)."""
    v = (1 << 64) - 1 if min_leaves is None else int(min_leaves)
    _check(lib.zkx_set_level_launches(device_id, v))


LONE_LEVEL_LAUNCHES_DEFAULT = 1 << 20


def set_lone_level_launches(min_leaves, device_id=0):
    """One tree through the C-ABI Merkle entry points builds its levels with
    one launch per level from min_leaves leaves for Poseidon and Poseidon2,
    min_leaves / 4 for Keccak-256 and Monolith, min_leaves / 32 for
    Poseidon-BN128 (0: always, None: never)."""
    v = (1 << 64) - 1 if min_leaves is None else int(min_leaves)
    _check(lib.zkx_set_lone_level_launches(device_id, v))


def commit_graphs_captured(device_id=0):
    return lib.zkx_commit_graphs_captured(device_id)


def tuning_clear(device_id=0):
    _check(lib.zkx_tuning_clear(device_id))


def tuning_save(path, device_id=0):
    _check(lib.zkx_tuning_save(device_id, os.fsencode(path)))


def tuning_load(path, device_id=0):
    _check(lib.zkx_tuning_load(device_id, os.fsencode(path)))


def ntt_pass_sizes(lg, tuning=None):
    """Stage counts of the DIF passes over 2^lg points (host-only)."""
    out = (ctypes.c_uint32 * 40)()
    n = lib.zkx_ntt_pass_sizes(lg, ctypes.byref(tuning) if tuning is not None else None, out, 40)
    if n < 0:
        raise ZeknoxError(1, "schedule parameters out of range")
    return list(out[:n])


def merkle_plan(num_leaves, num_caps, tile_log):
    """The one-launch Merkle forest's pass plan (host-only): a list of
    (nodes, lg_w, tile_log, first_counter) per pass and the counter total."""
    n = 8
    count = (ctypes.c_uint64 * n)()
    lg_w, tl, fc = (ctypes.c_uint32 * n)(), (ctypes.c_uint32 * n)(), (ctypes.c_uint32 * n)()
    total = ctypes.c_uint64()
    k = lib.zkx_merkle_plan(num_leaves, num_caps, tile_log, count, lg_w, tl, fc, n, ctypes.byref(total))
    if k < 0:
        raise ZeknoxError(1, "not a linear Merkle forest")
    return [(count[i], lg_w[i], tl[i], fc[i]) for i in range(k)], total.value


# ---- commit server -----------------------------------------------------------

class Server:
    """Continuous-batching commit server (zeknox_ext.h zkx_server_*)."""

    def __init__(self, budget_bytes, streams=6, max_batch=64, device_id=0):
        h = ctypes.c_void_p()
        _check(lib.zkx_server_create(device_id, int(budget_bytes), streams, max_batch, ctypes.byref(h)))
        self.h = h.value
        self.device_id = device_id

    def submit(self, coeffs, cols, lg_n, rate, cap_height, hash_type=HASH_POSEIDON, want_leaves=True):
        """coeffs: DeviceBuffer (kept alive by the caller until done)."""
        job = CommitJob(_addr(coeffs), None, None, None, cols, lg_n, rate, cap_height, hash_type, 0)
        rid = ctypes.c_uint64()
        _check(lib.zkx_server_submit(self.h, ctypes.byref(job), 1 if want_leaves else 0, ctypes.byref(rid)))
        return rid.value

    def wait(self, rid):
        r = ServeResult()
        _check(lib.zkx_server_wait(self.h, rid, ctypes.byref(r)))
        return r

    def wait_any(self):
        """Next finished request not yet returned by wait/wait_any: (id, result)."""
        rid, r = ctypes.c_uint64(), ServeResult()
        _check(lib.zkx_server_wait_any(self.h, ctypes.byref(rid), ctypes.byref(r)))
        return rid.value, r

    def results(self, r, cols, lg_n, rate, cap_height, leaves=True):
        """Copy a finished request's outputs to the host: (leaves, digests,
        caps); leaves None if not kept or not asked for."""
        N = 1 << (lg_n + rate)
        nc = 1 << cap_height
        nd = 0 if nc == N else 2 * (N - nc)

        def fetch(ptr, n):
            out = np.empty(n, np.uint64)
            if n:
                _check(lib.zkx_memcpy_dtoh(self.device_id, out.ctypes.data, ptr, n * 8))
            return out
        rows = fetch(r.leaves, cols * N) if (r.leaves and leaves) else None
        return rows, fetch(r.digests, nd * 4), fetch(r.caps, nc * 4)

    def release(self, rid):
        _check(lib.zkx_server_release(self.h, rid))

    def stats(self):
        st = ServeStats()
        _check(lib.zkx_server_stats(self.h, ctypes.byref(st)))
        return {f: getattr(st, f) for f, _ in ServeStats._fields_}

    @staticmethod
    def footprint(cols, lg_n, rate, cap_height, want_leaves=True):
        job = CommitJob(None, None, None, None, cols, lg_n, rate, cap_height, 0, 0)
        return lib.zkx_server_footprint(ctypes.byref(job), 1 if want_leaves else 0)

    def close(self):
        if self.h:
            _check(lib.zkx_server_destroy(self.h))
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---- timing ---------------------------------------------------------------

def issue_probe(iters=4096, device_id=0):
    """zkx_issue_probe: the measured v_mad_u64_u32 rate and the issue slots
    one i8 MFMA takes beside it, as a dict."""
    r = IssueRates()
    _check(lib.zkx_issue_probe(device_id, iters, ctypes.byref(r)))
    return {f: getattr(r, f) for f, _ in IssueRates._fields_ if f != "reserved"}


def profile_enable(on=True):
    _check(lib.zkx_profile_enable(1 if on else 0))


def kernel_stats():
    """{family: (launches, total_ms)} of library kernels timed with HIP
    events on their launch stream since the last reset."""
    cap = 64
    names = (ctypes.c_char * 32 * cap)()
    launches = (ctypes.c_uint64 * cap)()
    totals = (ctypes.c_double * cap)()
    n = lib.zkx_kernel_stats(names, launches, totals, cap)
    return {names[i].value.decode(): (launches[i], totals[i]) for i in range(min(n, cap))}


def kernel_stats_reset():
    lib.zkx_kernel_stats_reset()
