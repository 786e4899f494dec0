"""ctypes binding of the CPU oracle (oracle/spx_oracle.c).

TEST INFRASTRUCTURE ONLY.  The oracle is what the HIP engine in
``cuspx_amd`` is tested against: only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import this module.  The product
path never loads it.

The library is built by ``oracle/Makefile`` into ``oracle/_build``.
"""
import ctypes
import os
import subprocess

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "libspx_oracle.so")

PARAM_NAMES = [
    "sha2-128s", "sha2-128f", "sha2-192s", "sha2-192f", "sha2-256s", "sha2-256f",
    "shake256-128s", "shake256-128f", "shake256-192s", "shake256-192f",
    "shake256-256s", "shake256-256f",
]


class Params(ctypes.Structure):
    _fields_ = [(f, ctypes.c_int) for f in (
        "hash", "n", "full_h", "d", "a", "k", "w", "logw", "len1", "len2", "len",
        "tree_h", "wots_bytes", "fors_msg_bytes", "fors_bytes", "sig_bytes",
        "pk_bytes", "sk_bytes", "tree_bits", "tree_bytes", "leaf_bits",
        "leaf_bytes", "dgst_bytes")] + [("name", ctypes.c_char * 32)]


def build():
    """Compile the oracle with gcc (make -C oracle)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = ctypes.CDLL(_LIB_PATH)
        u8p = ctypes.c_char_p
        _lib.spxo_count_sha256_blocks.restype = ctypes.c_uint64
        _lib.spxo_count_keccak_perms.restype = ctypes.c_uint64
        _lib.spxo_randombytes.argtypes = [ctypes.c_void_p, ctypes.c_uint64]
        _lib.spxo_sha256.argtypes = [ctypes.c_void_p, u8p, ctypes.c_size_t]
        _lib.spxo_shake256.argtypes = [ctypes.c_void_p, ctypes.c_size_t, u8p, ctypes.c_size_t]
        _lib.spxo_mgf1.argtypes = [ctypes.c_void_p, ctypes.c_size_t, u8p, ctypes.c_size_t]
        _lib.spxo_hmac_sha256.argtypes = [ctypes.c_void_p, u8p, ctypes.c_size_t, u8p,
                                          ctypes.c_size_t]
        _lib.spxo_gen_message_random.argtypes = [ctypes.c_int, ctypes.c_void_p, u8p, u8p, u8p,
                                                 ctypes.c_uint64]
        _lib.spxo_hash_message.argtypes = [ctypes.c_int, ctypes.c_void_p,
                                           ctypes.POINTER(ctypes.c_uint64),
                                           ctypes.POINTER(ctypes.c_uint32), u8p, u8p, u8p,
                                           ctypes.c_uint64]
        _lib.spxo_addr_set.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_uint64,
                                       ctypes.c_int]
        _lib.spxo_signature_with_optrand.argtypes = [ctypes.c_int, ctypes.c_void_p, u8p,
                                                     ctypes.c_size_t, u8p, u8p]
        _lib.spxo_verify.argtypes = [ctypes.c_int, u8p, ctypes.c_size_t, u8p, ctypes.c_size_t,
                                     u8p]
        _lib.spxo_sign.argtypes = [ctypes.c_int, ctypes.c_void_p,
                                   ctypes.POINTER(ctypes.c_uint64), u8p, ctypes.c_uint64, u8p]
        _lib.spxo_open.argtypes = [ctypes.c_int, ctypes.c_void_p,
                                   ctypes.POINTER(ctypes.c_uint64), u8p, ctypes.c_uint64, u8p]
        _lib.spxo_sign_trace.argtypes = [ctypes.c_int, u8p, ctypes.c_size_t, u8p, u8p,
                                         ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                         ctypes.c_void_p]
        _lib.spxo_compute_root.argtypes = [ctypes.c_int, ctypes.c_void_p, u8p, ctypes.c_uint32,
                                           ctypes.c_uint32, u8p, ctypes.c_uint32, u8p,
                                           ctypes.c_void_p]
    return _lib


def params(pid):
    p = Params()
    if lib().spxo_get_params(pid, ctypes.byref(p)) != 0:
        raise ValueError("bad parameter id %r" % pid)
    return p


def param_id(name):
    pid = lib().spxo_param_id(name.encode())
    if pid < 0:
        raise KeyError(name)
    return pid


def set_debug_mode(on):
    lib().spxo_set_debug_mode(1 if on else 0)


def reset_counts():
    lib().spxo_reset_counts()


def counts():
    return lib().spxo_count_sha256_blocks(), lib().spxo_count_keccak_perms()


def _buf(n):
    return ctypes.create_string_buffer(max(n, 1))


def sha256(data):
    out = _buf(32)
    lib().spxo_sha256(out, data, len(data))
    return out.raw[:32]


def sha256_compress_words(state, words):
    st = (ctypes.c_uint32 * 8)(*state)
    w = (ctypes.c_uint32 * 16)(*words)
    lib().spxo_sha256_compress_words(st, w)
    return list(st)


def shake256(data, outlen):
    out = _buf(outlen)
    lib().spxo_shake256(out, outlen, data, len(data))
    return out.raw[:outlen]


def keccakf1600(lanes):
    st = (ctypes.c_uint64 * 25)(*lanes)
    lib().spxo_keccakf1600(st)
    return list(st)


def mgf1(seed, outlen):
    out = _buf(outlen)
    lib().spxo_mgf1(out, outlen, seed, len(seed))
    return out.raw[:outlen]


def hmac_sha256(key, msg):
    out = _buf(32)
    lib().spxo_hmac_sha256(out, key, len(key), msg, len(msg))
    return out.raw[:32]


def aes256_ecb(key, block):
    out = _buf(16)
    lib().spxo_aes256_ecb(key, block, out)
    return out.raw[:16]


def randombytes_init(entropy, pers=None):
    lib().spxo_randombytes_init(entropy, pers, 256)


def randombytes(n):
    out = _buf(n)
    lib().spxo_randombytes(out, n)
    return out.raw[:n]


ADDR_FIELDS = {"layer": 0, "tree": 1, "type": 2, "keypair": 3, "chain": 4, "hash": 5,
               "height": 6, "index": 7}


def make_addr(pid, **fields):
    a = (ctypes.c_uint32 * 8)()
    for k, v in fields.items():
        lib().spxo_addr_set(a, ADDR_FIELDS[k], v, pid)
    return a


def addr_bytes(a):
    return bytes(a)


def thash(pid, data, pub_seed, addr):
    p = params(pid)
    out = _buf(p.n)
    lib().spxo_thash(pid, out, data, len(data) // p.n, pub_seed, addr)
    return out.raw[:p.n]


def prf_addr(pid, key, addr):
    p = params(pid)
    out = _buf(p.n)
    lib().spxo_prf_addr(pid, out, key, addr)
    return out.raw[:p.n]


def gen_message_random(pid, sk_prf, optrand, m):
    p = params(pid)
    out = _buf(p.n)
    lib().spxo_gen_message_random(pid, out, sk_prf, optrand, m, len(m))
    return out.raw[:p.n]


def hash_message(pid, R, pk, m):
    p = params(pid)
    dg = _buf(p.fors_msg_bytes)
    tree = ctypes.c_uint64()
    leaf = ctypes.c_uint32()
    lib().spxo_hash_message(pid, dg, ctypes.byref(tree), ctypes.byref(leaf), R, pk, m, len(m))
    return dg.raw[:p.fors_msg_bytes], tree.value, leaf.value


def chain_lengths(pid, msg):
    p = params(pid)
    out = (ctypes.c_uint * p.len)()
    lib().spxo_chain_lengths(pid, out, msg)
    return list(out)


def gen_chain(pid, x, start, steps, pub_seed, addr):
    p = params(pid)
    out = _buf(p.n)
    lib().spxo_gen_chain(pid, out, x, start, steps, pub_seed, addr)
    return out.raw[:p.n]


def wots_gen_pk(pid, sk_seed, pub_seed, addr):
    p = params(pid)
    out = _buf(p.wots_bytes)
    lib().spxo_wots_gen_pk(pid, out, sk_seed, pub_seed, addr)
    return out.raw[:p.wots_bytes]


def wots_sign(pid, msg, sk_seed, pub_seed, addr):
    p = params(pid)
    out = _buf(p.wots_bytes)
    lib().spxo_wots_sign(pid, out, msg, sk_seed, pub_seed, addr)
    return out.raw[:p.wots_bytes]


def wots_pk_from_sig(pid, sig, msg, pub_seed, addr):
    p = params(pid)
    out = _buf(p.wots_bytes)
    lib().spxo_wots_pk_from_sig(pid, out, sig, msg, pub_seed, addr)
    return out.raw[:p.wots_bytes]


def wots_gen_leaf(pid, sk_seed, pub_seed, addr_idx, tree_addr):
    p = params(pid)
    out = _buf(p.n)
    lib().spxo_wots_gen_leaf(pid, out, sk_seed, pub_seed, addr_idx, tree_addr)
    return out.raw[:p.n]


def xmss_treehash(pid, sk_seed, pub_seed, leaf_idx, tree_addr):
    p = params(pid)
    root = _buf(p.n)
    auth = _buf(p.n * p.tree_h)
    lib().spxo_xmss_treehash(pid, root, auth, sk_seed, pub_seed, leaf_idx, tree_addr)
    return root.raw[:p.n], auth.raw[:p.n * p.tree_h]


def compute_root(pid, leaf, leaf_idx, idx_offset, auth, height, pub_seed, addr):
    p = params(pid)
    out = _buf(p.n)
    lib().spxo_compute_root(pid, out, leaf, leaf_idx, idx_offset, auth, height, pub_seed, addr)
    return out.raw[:p.n]


def message_to_indices(pid, m):
    p = params(pid)
    out = (ctypes.c_uint32 * p.k)()
    lib().spxo_message_to_indices(pid, out, m)
    return list(out)


def fors_sign(pid, m, sk_seed, pub_seed, fors_addr):
    p = params(pid)
    sig = _buf(p.fors_bytes)
    pk = _buf(p.n)
    lib().spxo_fors_sign(pid, sig, pk, m, sk_seed, pub_seed, fors_addr)
    return sig.raw[:p.fors_bytes], pk.raw[:p.n]


def fors_pk_from_sig(pid, sig, m, pub_seed, fors_addr):
    p = params(pid)
    pk = _buf(p.n)
    lib().spxo_fors_pk_from_sig(pid, pk, sig, m, pub_seed, fors_addr)
    return pk.raw[:p.n]


def seed_keypair(pid, seed):
    p = params(pid)
    pk = _buf(p.pk_bytes)
    sk = _buf(p.sk_bytes)
    lib().spxo_seed_keypair(pid, pk, sk, seed)
    return pk.raw[:p.pk_bytes], sk.raw[:p.sk_bytes]


def keypair(pid):
    p = params(pid)
    pk = _buf(p.pk_bytes)
    sk = _buf(p.sk_bytes)
    lib().spxo_keypair(pid, pk, sk)
    return pk.raw[:p.pk_bytes], sk.raw[:p.sk_bytes]


def signature_with_optrand(pid, m, sk, optrand):
    p = params(pid)
    sig = _buf(p.sig_bytes)
    lib().spxo_signature_with_optrand(pid, sig, m, len(m), sk, optrand)
    return sig.raw[:p.sig_bytes]


def verify(pid, sig, m, pk):
    return lib().spxo_verify(pid, sig, len(sig), m, len(m), pk)


def sign(pid, m, sk):
    p = params(pid)
    sm = _buf(p.sig_bytes + len(m))
    smlen = ctypes.c_uint64()
    lib().spxo_sign(pid, sm, ctypes.byref(smlen), m, len(m), sk)
    return sm.raw[:smlen.value]


def open_(pid, sm, pk):
    m = _buf(len(sm))
    mlen = ctypes.c_uint64()
    r = lib().spxo_open(pid, m, ctypes.byref(mlen), sm, len(sm), pk)
    return r, m.raw[:mlen.value], m.raw[:len(sm)]


def sign_trace(pid, m, sk, optrand):
    p = params(pid)
    fpk = _buf(p.n)
    roots = _buf(p.n * p.d)
    trees = (ctypes.c_uint64 * p.d)()
    leaves = (ctypes.c_uint32 * p.d)()
    lib().spxo_sign_trace(pid, m, len(m), sk, optrand, fpk, roots, trees, leaves)
    return fpk.raw[:p.n], roots.raw[:p.n * p.d], list(trees), list(leaves)
