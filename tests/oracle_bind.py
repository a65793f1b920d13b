"""ctypes binding of the CPU oracle (oracle/liboracle.so) and, where it has
been built in this tree, of the reference's own CPU hashers (oracle/_ref).

Test infrastructure only: the checker the HIP library is compared against.
"""
import ctypes
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(ROOT, "oracle", "liboracle.so")
REF_SO = os.path.join(ROOT, "oracle", "_ref", "libzkref.so")
P = 0xFFFFFFFF00000001

_u64p = ctypes.POINTER(ctypes.c_uint64)


def _ptr(a):
    return a.ctypes.data_as(_u64p)


def _load_oracle():
    if not os.path.exists(ORACLE_SO):
        import subprocess
        subprocess.run(["make", "-C", os.path.join(ROOT, "oracle"), "liboracle.so"], check=True,
                       capture_output=True)
    lib = ctypes.CDLL(ORACLE_SO)
    u64, u32, i = ctypes.c_uint64, ctypes.c_uint, ctypes.c_int
    for name, res, args in [
        ("orc_gl_add", u64, [u64, u64]), ("orc_gl_sub", u64, [u64, u64]), ("orc_gl_mul", u64, [u64, u64]),
        ("orc_gl_pow", u64, [u64, u64]), ("orc_gl_inv", u64, [u64]), ("orc_gl_omega", u64, [u32]),
        ("orc_poseidon_permute", None, [_u64p]), ("orc_poseidon2_permute", None, [_u64p]),
        ("orc_monolith_permute", None, [_u64p]), ("orc_poseidon_bn128_permute", None, [_u64p]),
        ("orc_bn_mont_mul", None, [_u64p, _u64p, _u64p]), ("orc_bn_add", None, [_u64p, _u64p, _u64p]),
        ("orc_keccak", None, [ctypes.c_char_p, u64, ctypes.c_char_p, u32, u32, u32]),
        ("orc_hash_one", None, [_u64p, u64, _u64p, i]), ("orc_hash_two", None, [_u64p, _u64p, _u64p, i]),
        ("orc_fill_digests", i, [_u64p, _u64p, _u64p, u64, u64, u64, u64, u64, i]),
        ("orc_ntt", None, [_u64p, u32, u32, i, i, u64]),
        ("orc_lde", None, [_u64p, _u64p, u32, u32, u32, i, i, u64]),
        ("orc_transpose_rev", None, [_u64p, _u64p, u32, u32]),
        ("orc_fill_canonical", None, [_u64p, u64, u64]),
        ("orc_salt", None, [_u64p, u64, u64]), ("orc_splitmix64", None, [_u64p, u64, u64]),
        ("orc_num_threads", i, []),
    ]:
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


orc = _load_oracle()


def ref_lib():
    """The reference's CPU Merkle/Poseidon built from its own sources, or None."""
    if not os.path.exists(REF_SO):
        return None
    lib = ctypes.CDLL(REF_SO)
    u64, i = ctypes.c_uint64, ctypes.c_int
    lib.ref_hash_one.argtypes = [_u64p, u64, _u64p, i]
    lib.ref_hash_two.argtypes = [_u64p, _u64p, _u64p, i]
    lib.fill_digests_buf_linear_cpu.argtypes = [_u64p, _u64p, _u64p, u64, u64, u64, u64, u64, u64]
    return lib


# ---- numpy-level helpers ----------------------------------------------------

def canonical(count, seed):
    """`count` canonical field elements from the shared splitmix64 stream."""
    out = np.empty(int(count), dtype=np.uint64)
    orc.orc_fill_canonical(_ptr(out), out.size, seed)
    return out


def hash_one(x, hash_type=0):
    x = np.ascontiguousarray(x, dtype=np.uint64)
    out = np.zeros(4, np.uint64)
    orc.orc_hash_one(_ptr(x) if x.size else _ptr(np.zeros(1, np.uint64)), x.size, _ptr(out), hash_type)
    return out


def hash_two(l, r, hash_type=0):
    l = np.ascontiguousarray(l, dtype=np.uint64)
    r = np.ascontiguousarray(r, dtype=np.uint64)
    out = np.zeros(4, np.uint64)
    orc.orc_hash_two(_ptr(l), _ptr(r), _ptr(out), hash_type)
    return out


def permute(state, hash_type=0):
    s = np.ascontiguousarray(state, dtype=np.uint64).copy()
    fn = {2: orc.orc_poseidon_bn128_permute, 3: orc.orc_poseidon2_permute,
          4: orc.orc_monolith_permute}.get(hash_type, orc.orc_poseidon_permute)
    fn(_ptr(s))
    return s


def keccak(data, outlen=32, rate=136, ds=0x01):
    """Keccak sponge over bytes (ds 0x01: original Keccak padding, 0x06: SHA-3)."""
    out = ctypes.create_string_buffer(outlen)
    orc.orc_keccak(bytes(data), len(data), out, outlen, rate, ds)
    return out.raw


def merkle(leaves, leaf_size, cap_height, hash_type=0):
    leaves = np.ascontiguousarray(leaves, dtype=np.uint64).ravel()
    nl = leaves.size // leaf_size
    nc = 1 << cap_height
    nd = 2 * (nl - nc) if nc != nl else 0
    dig = np.zeros(max(nd, 1) * 4, np.uint64)
    cap = np.zeros(nc * 4, np.uint64)
    rc = orc.orc_fill_digests(_ptr(dig), _ptr(cap), _ptr(leaves), nd, nc, nl, leaf_size, cap_height, hash_type)
    assert rc == 0
    return dig[: nd * 4], cap


def ntt(data, lg, batches, inverse=False, coset=False, gen=7):
    d = np.ascontiguousarray(data, dtype=np.uint64).copy()
    orc.orc_ntt(_ptr(d), lg, batches, int(inverse), int(coset), gen)
    return d


def lde(coeffs, lg_n, rate, batches, inverse=False, coset=True, gen=7):
    c = np.ascontiguousarray(coeffs, dtype=np.uint64)
    out = np.empty(batches << (lg_n + rate), np.uint64)
    orc.orc_lde(_ptr(out), _ptr(c), lg_n, rate, batches, int(inverse), int(coset), gen)
    return out


def transpose_rev(data, lg_n, batches):
    d = np.ascontiguousarray(data, dtype=np.uint64)
    out = np.empty_like(d)
    orc.orc_transpose_rev(_ptr(out), _ptr(d), lg_n, batches)
    return out


def bitrev(x, bits):
    return int(format(x, f"0{bits}b")[::-1], 2) if bits else 0


def salt(count, seed):
    """The seeded salt / zkx_fill_random stream: element g from its own seed."""
    out = np.empty(int(count), dtype=np.uint64)
    orc.orc_salt(_ptr(out), out.size, seed)
    return out


def splitmix64(count, seed):
    out = np.empty(int(count), dtype=np.uint64)
    orc.orc_splitmix64(_ptr(out), out.size, seed)
    return out
