"""cuspx_amd -- MI355X SPHINCS+ engine (host-side binding of libcuspx.so).

The compute path is the HIP library ``cuspx_amd/libcuspx.so`` (C ABI in
``include/cuspx.h``).  This module is a thin ctypes mirror of the
reference's interface:

* the api.h surface -- ``crypto_sign_keypair``, ``crypto_sign``,
  ``crypto_sign_open``, ... (api.h:19-183 of the reference) acting on the
  default parameter set;
* the batched data-parallel entry points ``face_mdp_crypto_sign_keypair`` /
  ``face_mdp_crypto_sign`` / ``face_mdp_crypto_sign_open``
  (sign.cu:1208, 5098, 6856);
* runtime-parameterised batches with per-item status (``batch_keygen``,
  ``batch_sign``, ``batch_verify``);
* device-resident entry points taking raw device pointers (``sign_device``
  ...), used by bench.py with buffers already in HBM.

Importing raises if the library has not been built: there is no CPU
fallback for any part of the engine.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("CUSPX_LIBRARY") or os.path.join(_HERE, "libcuspx.so")

OK, ERR_VERIFY, ERR_PARAM, ERR_HIP, ERR_ARG, ERR_BUDGET = 0, -1, -2, -3, -4, -5
OP_SIGN, OP_VERIFY, OP_KEYGEN = 0, 1, 2

PARAM_NAMES = [
    "sha2-128s", "sha2-128f", "sha2-192s", "sha2-192f", "sha2-256s", "sha2-256f",
    "shake256-128s", "shake256-128f", "shake256-192s", "shake256-192f",
    "shake256-256s", "shake256-256f",
]


class CuspxError(RuntimeError):
    """A libcuspx call failed; ``code`` is its negative status."""

    def __init__(self, msg, code=None):
        super().__init__(msg)
        self.code = code


class ParamInfo(ctypes.Structure):
    _fields_ = [(f, ctypes.c_int) for f in (
        "pid", "hash", "n", "full_h", "d", "a", "k", "len", "sig_bytes", "pk_bytes",
        "sk_bytes", "seed_bytes")] + [("name", ctypes.c_char * 32)]


_lib = None


def lib():
    """The loaded libcuspx.so (raises if it was not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError("libcuspx.so is not built: run `python -m cuspx_amd.build` "
                              "(or __graft_entry__.build())")
        L = ctypes.CDLL(LIB_PATH)
        vp, u8p, sz = ctypes.c_void_p, ctypes.c_char_p, ctypes.c_size_t
        L.cuspx_last_error.restype = ctypes.c_char_p
        L.cuspx_workspace_bytes.restype = ctypes.c_size_t
        L.cuspx_workspace_bytes.argtypes = [ctypes.c_int, ctypes.c_uint32, ctypes.c_int]
        L.crypto_sign_secretkeybytes.restype = ctypes.c_ulonglong
        L.crypto_sign_publickeybytes.restype = ctypes.c_ulonglong
        L.crypto_sign_bytes.restype = ctypes.c_ulonglong
        L.crypto_sign_seedbytes.restype = ctypes.c_ulonglong
        L.randombytes.argtypes = [vp, ctypes.c_ulonglong]
        L.randombytes_init.argtypes = [vp, vp, ctypes.c_int]
        L.crypto_sign_seed_keypair.argtypes = [vp, vp, vp]
        L.crypto_sign_keypair.argtypes = [vp, vp]
        L.crypto_sign_signature.argtypes = [vp, ctypes.POINTER(sz), vp, sz, vp]
        L.crypto_sign_verify.argtypes = [vp, sz, vp, sz, vp]
        L.crypto_sign.argtypes = [vp, ctypes.POINTER(ctypes.c_ulonglong), vp,
                                  ctypes.c_ulonglong, vp]
        L.crypto_sign_open.argtypes = [vp, ctypes.POINTER(ctypes.c_ulonglong), vp,
                                       ctypes.c_ulonglong, vp]
        L.face_mdp_crypto_sign_keypair.argtypes = [vp, vp, ctypes.c_uint]
        L.face_mdp_crypto_sign.argtypes = [vp, ctypes.POINTER(ctypes.c_ulonglong), vp,
                                           ctypes.c_ulonglong, vp, ctypes.c_uint]
        L.face_mdp_crypto_sign_open.argtypes = [vp, ctypes.POINTER(ctypes.c_ulonglong), vp,
                                                ctypes.c_ulonglong, vp, ctypes.c_uint]
        L.face_ms_mdp_crypto_sign_keypair.argtypes = L.face_mdp_crypto_sign_keypair.argtypes
        L.face_ms_mdp_crypto_sign.argtypes = L.face_mdp_crypto_sign.argtypes
        L.face_ms_mdp_crypto_sign_open.argtypes = L.face_mdp_crypto_sign_open.argtypes
        L.spx_batch_keygen.argtypes = [ctypes.c_int, ctypes.c_uint32, vp, vp, vp]
        L.spx_batch_sign.argtypes = [ctypes.c_int, ctypes.c_uint32, vp, vp, vp, vp, vp, vp]
        L.spx_batch_verify.argtypes = [ctypes.c_int, ctypes.c_uint32, vp, vp, vp, vp, vp, vp]
        L.cuspx_sign_device.argtypes = [ctypes.c_int, ctypes.c_uint32, vp, ctypes.c_uint32, vp,
                                        vp, vp, vp, vp, ctypes.c_uint64, vp, vp]
        L.cuspx_verify_device.argtypes = [ctypes.c_int, ctypes.c_uint32, vp, ctypes.c_uint32, vp,
                                          ctypes.c_uint64, vp, vp, vp, vp, vp, vp]
        L.cuspx_keygen_device.argtypes = [ctypes.c_int, ctypes.c_uint32, vp, ctypes.c_uint32, vp,
                                          vp, vp, vp]
        for f in ("cuspx_dev_free", "cuspx_stream_destroy", "cuspx_stream_sync",
                  "cuspx_event_destroy"):
            getattr(L, f).argtypes = [vp]
        L.cuspx_dev_alloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        L.cuspx_memcpy_htod.argtypes = [vp, vp, ctypes.c_size_t]
        L.cuspx_memcpy_dtoh.argtypes = [vp, vp, ctypes.c_size_t]
        L.cuspx_memcpy_dtod.argtypes = [vp, vp, ctypes.c_size_t]
        L.cuspx_memset_device.argtypes = [vp, ctypes.c_int, ctypes.c_size_t]
        L.cuspx_stream_create.argtypes = [ctypes.POINTER(ctypes.c_void_p)]
        L.cuspx_event_create.argtypes = [ctypes.POINTER(ctypes.c_void_p)]
        L.cuspx_event_record.argtypes = [vp, vp]
        L.cuspx_event_elapsed_ms.argtypes = [vp, vp, ctypes.POINTER(ctypes.c_float)]
        L.cuspx_device_info.argtypes = [ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int),
                                        ctypes.POINTER(ctypes.c_size_t), ctypes.c_char_p,
                                        ctypes.c_int]
        L.cuspx_profile_read.argtypes = [ctypes.c_char_p, ctypes.c_int,
                                         ctypes.POINTER(ctypes.c_float),
                                         ctypes.POINTER(ctypes.c_int), ctypes.c_int]
        _lib = L
    return _lib


def _check(rc, what):
    if rc not in (OK, ERR_VERIFY):
        raise CuspxError("%s failed (%d): %s" % (what, rc, lib().cuspx_last_error().decode()), rc)
    return rc


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p) if a is not None else None


# ------------------------------------------------------------ parameters

def param_id(name):
    pid = lib().cuspx_param_id(name.encode())
    if pid < 0:
        raise KeyError(name)
    return pid


def params(pid):
    p = ParamInfo()
    if lib().cuspx_get_param_info(pid, ctypes.byref(p)) != 0:
        raise KeyError(pid)
    return p


def set_default_param(pid):
    if lib().cuspx_set_default_param(pid) != 0:
        raise KeyError(pid)


def set_split(pid, op, lanes_per_leaf):
    """Force a work split (0 = automatic / tuned): lanes per WOTS leaf of the
    hypertree kernel for OP_SIGN / OP_KEYGEN (1, 2, 4, 8, 16); mean lanes per
    request of the verify kernel for OP_VERIFY (16, 32, 64; 128 when len = 67),
    which runs 256 / lanes requests per workgroup, or 1 for speculative
    verification (batches up to 256 requests)."""
    rc = lib().cuspx_set_split(pid, op, lanes_per_leaf)
    if rc != OK:
        raise CuspxError("cuspx_set_split failed (%d)" % rc, rc)


def get_split(pid, op, num):
    return lib().cuspx_get_split(pid, op, num)


def set_debug_mode(on):
    lib().cuspx_set_debug_mode(1 if on else 0)


def get_debug_mode():
    return bool(lib().cuspx_get_debug_mode())


# ------------------------------------------------------- api.h surface

def randombytes_init(entropy, personalization=None):
    e = ctypes.create_string_buffer(bytes(entropy), 48)
    p = ctypes.create_string_buffer(bytes(personalization), 48) if personalization else None
    lib().randombytes_init(e, p, 256)


def randombytes(n):
    b = ctypes.create_string_buffer(max(n, 1))
    lib().randombytes(b, n)
    return b.raw[:n]


def crypto_sign_seed_keypair(seed):
    L = lib()
    pk = ctypes.create_string_buffer(L.crypto_sign_publickeybytes())
    sk = ctypes.create_string_buffer(L.crypto_sign_secretkeybytes())
    _check(L.crypto_sign_seed_keypair(pk, sk, ctypes.create_string_buffer(bytes(seed))),
           "crypto_sign_seed_keypair")
    return pk.raw, sk.raw


def crypto_sign_keypair():
    L = lib()
    pk = ctypes.create_string_buffer(L.crypto_sign_publickeybytes())
    sk = ctypes.create_string_buffer(L.crypto_sign_secretkeybytes())
    _check(L.crypto_sign_keypair(pk, sk), "crypto_sign_keypair")
    return pk.raw, sk.raw


def crypto_sign(m, sk):
    L = lib()
    sm = ctypes.create_string_buffer(L.crypto_sign_bytes() + len(m))
    smlen = ctypes.c_ulonglong()
    _check(L.crypto_sign(sm, ctypes.byref(smlen), ctypes.create_string_buffer(bytes(m), max(len(m), 1)),
                         len(m), ctypes.create_string_buffer(bytes(sk))), "crypto_sign")
    return sm.raw[:smlen.value]


def crypto_sign_open(sm, pk):
    """Returns (rc, message); rc is 0 or -1 as in the reference."""
    L = lib()
    m = ctypes.create_string_buffer(max(len(sm), 1))
    mlen = ctypes.c_ulonglong()
    rc = _check(L.crypto_sign_open(m, ctypes.byref(mlen), ctypes.create_string_buffer(bytes(sm)),
                                   len(sm), ctypes.create_string_buffer(bytes(pk))),
                "crypto_sign_open")
    return rc, m.raw[:mlen.value]


def crypto_sign_signature(m, sk):
    L = lib()
    sig = ctypes.create_string_buffer(L.crypto_sign_bytes())
    siglen = ctypes.c_size_t()
    _check(L.crypto_sign_signature(sig, ctypes.byref(siglen),
                                   ctypes.create_string_buffer(bytes(m), max(len(m), 1)), len(m),
                                   ctypes.create_string_buffer(bytes(sk))), "crypto_sign_signature")
    return sig.raw[:siglen.value]


def crypto_sign_verify(sig, m, pk):
    return _check(lib().crypto_sign_verify(ctypes.create_string_buffer(bytes(sig)), len(sig),
                                           ctypes.create_string_buffer(bytes(m), max(len(m), 1)),
                                           len(m), ctypes.create_string_buffer(bytes(pk))),
                  "crypto_sign_verify")


# ------------------------------------------------- face_mdp (sign.cu)

def face_mdp_crypto_sign_keypair(num, ms=False):
    """ms: the multi-stream form face_ms_mdp_crypto_sign_keypair (same bytes)."""
    L = lib()
    pk = np.zeros((num, L.crypto_sign_publickeybytes()), np.uint8)
    sk = np.zeros((num, L.crypto_sign_secretkeybytes()), np.uint8)
    name = "face_ms_mdp_crypto_sign_keypair" if ms else "face_mdp_crypto_sign_keypair"
    _check(getattr(L, name)(_ptr(pk), _ptr(sk), num), name)
    return pk, sk


def face_mdp_crypto_sign(m, mlen, sk, num, ms=False):
    """m: uint8 array of num*32 bytes (SPX_MLEN stride); returns (sm, smlen)."""
    L = lib()
    stride = L.crypto_sign_bytes() + 32
    sm = np.zeros((num, stride), np.uint8)
    smlen = ctypes.c_ulonglong()
    m = np.ascontiguousarray(m, np.uint8)
    sk = np.ascontiguousarray(sk, np.uint8)
    name = "face_ms_mdp_crypto_sign" if ms else "face_mdp_crypto_sign"
    _check(getattr(L, name)(_ptr(sm), ctypes.byref(smlen), _ptr(m), mlen, _ptr(sk), num), name)
    return sm, smlen.value


def face_mdp_crypto_sign_open(sm, smlen, pk, num, ms=False):
    """Returns (rc, m[num][32], mlen)."""
    L = lib()
    m = np.zeros((num, 32), np.uint8)
    mlen = ctypes.c_ulonglong()
    sm = np.ascontiguousarray(sm, np.uint8)
    pk = np.ascontiguousarray(pk, np.uint8)
    name = "face_ms_mdp_crypto_sign_open" if ms else "face_mdp_crypto_sign_open"
    rc = _check(getattr(L, name)(_ptr(m), ctypes.byref(mlen), _ptr(sm), smlen, _ptr(pk), num),
                name)
    return rc, m, mlen.value


# ------------------------------------- runtime-parameterised batches

def _pack(msgs):
    mlen = np.array([len(m) for m in msgs], np.uint32)
    moff = np.zeros(len(msgs), np.uint64)
    if len(msgs):
        moff[1:] = np.cumsum(mlen[:-1], dtype=np.uint64)
    buf = np.frombuffer(b"".join(bytes(m) for m in msgs) + b"\0", np.uint8).copy()
    return buf, moff, mlen


def batch_keygen(pid, seeds):
    p = params(pid)
    seeds = np.ascontiguousarray(seeds, np.uint8).reshape(-1, 3 * p.n)
    num = seeds.shape[0]
    pk = np.zeros((num, p.pk_bytes), np.uint8)
    sk = np.zeros((num, p.sk_bytes), np.uint8)
    _check(lib().spx_batch_keygen(pid, num, _ptr(seeds), _ptr(pk), _ptr(sk)), "spx_batch_keygen")
    return pk, sk


def batch_sign(pid, sks, msgs, optrand=None):
    p = params(pid)
    sks = np.ascontiguousarray(sks, np.uint8).reshape(-1, p.sk_bytes)
    num = sks.shape[0]
    buf, moff, mlen = _pack(msgs)
    opt = None if optrand is None else np.ascontiguousarray(optrand, np.uint8).reshape(num, p.n)
    sig = np.zeros((num, p.sig_bytes), np.uint8)
    _check(lib().spx_batch_sign(pid, num, _ptr(sks), _ptr(buf), _ptr(moff), _ptr(mlen), _ptr(opt),
                                _ptr(sig)), "spx_batch_sign")
    return sig


def batch_verify(pid, pks, sigs, msgs):
    p = params(pid)
    pks = np.ascontiguousarray(pks, np.uint8).reshape(-1, p.pk_bytes)
    sigs = np.ascontiguousarray(sigs, np.uint8).reshape(-1, p.sig_bytes)
    num = pks.shape[0]
    buf, moff, mlen = _pack(msgs)
    st = np.full(num, -7, np.int32)
    _check(lib().spx_batch_verify(pid, num, _ptr(pks), _ptr(sigs), _ptr(buf), _ptr(moff),
                                  _ptr(mlen), _ptr(st)), "spx_batch_verify")
    return st


# --------------------------------------------- device-resident batches

def workspace_bytes(pid, num, op):
    return lib().cuspx_workspace_bytes(pid, num, op)


def sign_device(pid, num, sk, sk_stride, msg, moff, mlen, optrand, sig, sig_stride, ws,
                stream=None):
    """All buffers are raw device pointers (ints); asynchronous on stream."""
    return _check(lib().cuspx_sign_device(pid, num, sk, sk_stride, msg, moff, mlen, optrand, sig,
                                          sig_stride, ws, stream), "cuspx_sign_device")


def verify_device(pid, num, pk, pk_stride, sig, sig_stride, msg, moff, mlen, status, ws,
                  stream=None):
    return _check(lib().cuspx_verify_device(pid, num, pk, pk_stride, sig, sig_stride, msg, moff,
                                            mlen, status, ws, stream), "cuspx_verify_device")


def keygen_device(pid, num, seed, seed_stride, pk, sk, ws, stream=None):
    return _check(lib().cuspx_keygen_device(pid, num, seed, seed_stride, pk, sk, ws, stream),
                  "cuspx_keygen_device")


def profile_enable(on=True, by_set=False):
    lib().cuspx_profile_enable((2 if by_set else 1) if on else 0)


def profile_reset():
    lib().cuspx_profile_reset()


def profile_read():
    """{kernel: (total_ms, launches)} since the last reset."""
    cap = 64
    names = ctypes.create_string_buffer(4096)
    ms = (ctypes.c_float * cap)()
    cnt = (ctypes.c_int * cap)()
    k = lib().cuspx_profile_read(names, 4096, ms, cnt, cap)
    keys = names.value.decode().split("\n")[:k]
    return {keys[i]: (ms[i], cnt[i]) for i in range(k)}


# ------------------------------------------------------ device plumbing
# The library carries its own HIP runtime; these helpers keep data resident
# in HBM without a second runtime in the process.

class DeviceBuffer:
    """A device allocation owned by libcuspx's HIP runtime."""

    def __init__(self, nbytes):
        self.nbytes = int(nbytes)
        p = ctypes.c_void_p()
        _check(lib().cuspx_dev_alloc(ctypes.byref(p), self.nbytes), "cuspx_dev_alloc")
        self.ptr = p.value

    @classmethod
    def from_array(cls, arr):
        arr = np.ascontiguousarray(arr)
        b = cls(max(arr.nbytes, 1))
        b.upload(arr)
        return b

    def upload(self, arr):
        arr = np.ascontiguousarray(arr)
        _check(lib().cuspx_memcpy_htod(self.ptr, arr.ctypes.data_as(ctypes.c_void_p), arr.nbytes),
               "cuspx_memcpy_htod")

    def download(self, dtype=np.uint8, count=None):
        dtype = np.dtype(dtype)
        count = self.nbytes // dtype.itemsize if count is None else count
        out = np.empty(count, dtype)
        _check(lib().cuspx_memcpy_dtoh(out.ctypes.data_as(ctypes.c_void_p), self.ptr, out.nbytes),
               "cuspx_memcpy_dtoh")
        return out

    def fill(self, value=0):
        _check(lib().cuspx_memset_device(self.ptr, value, self.nbytes), "cuspx_memset_device")

    def free(self):
        if self.ptr:
            lib().cuspx_dev_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Stream:
    def __init__(self):
        p = ctypes.c_void_p()
        _check(lib().cuspx_stream_create(ctypes.byref(p)), "cuspx_stream_create")
        self.handle = p.value

    def sync(self):
        _check(lib().cuspx_stream_sync(self.handle), "cuspx_stream_sync")


class Event:
    def __init__(self):
        p = ctypes.c_void_p()
        _check(lib().cuspx_event_create(ctypes.byref(p)), "cuspx_event_create")
        self.handle = p.value

    def record(self, stream=None):
        _check(lib().cuspx_event_record(self.handle, stream.handle if stream else None),
               "cuspx_event_record")

    def elapsed_ms(self, stop):
        ms = ctypes.c_float()
        _check(lib().cuspx_event_elapsed_ms(self.handle, stop.handle, ctypes.byref(ms)),
               "cuspx_event_elapsed_ms")
        return ms.value


def device_info():
    cu, clk, mem = ctypes.c_int(), ctypes.c_int(), ctypes.c_size_t()
    name = ctypes.create_string_buffer(128)
    _check(lib().cuspx_device_info(ctypes.byref(cu), ctypes.byref(clk), ctypes.byref(mem), name,
                                   128), "cuspx_device_info")
    return {"compute_units": cu.value, "clock_mhz": clk.value / 1000.0,
            "hbm_gib": mem.value / 2**30, "name": name.value.decode()}


# ---------------------------------------------------------------- serving

class ServerConfig(ctypes.Structure):
    _fields_ = [("hbm_budget_bytes", ctypes.c_uint64), ("max_batch", ctypes.c_uint32),
                ("slots", ctypes.c_uint32), ("max_msg_bytes", ctypes.c_uint32),
                ("batch_window_us", ctypes.c_uint32), ("max_keys_per_set", ctypes.c_uint32),
                ("long_slots", ctypes.c_uint32)]


class ServerStats(ctypes.Structure):
    _fields_ = [("signs", ctypes.c_uint64), ("verifies", ctypes.c_uint64),
                ("batches", ctypes.c_uint64), ("pending", ctypes.c_uint64),
                ("mean_batch", ctypes.c_double), ("sign_p50_us", ctypes.c_float),
                ("sign_p99_us", ctypes.c_float), ("verify_p50_us", ctypes.c_float),
                ("verify_p99_us", ctypes.c_float), ("hbm_bytes", ctypes.c_uint64),
                ("last_error", ctypes.c_int32)]


class Server:
    """Continuous-batching sign/verify service with keys resident on the GPU
    (libcuspx spx_server_*)."""

    def __init__(self, hbm_budget_bytes=0, max_batch=0, slots=0, max_msg_bytes=0,
                 batch_window_us=200, max_keys_per_set=0, long_slots=0):
        L = lib()
        vp = ctypes.c_void_p
        L.spx_server_create.argtypes = [ctypes.POINTER(ServerConfig), ctypes.POINTER(vp)]
        L.spx_server_add_key.argtypes = [vp, ctypes.c_int, vp, vp, ctypes.POINTER(ctypes.c_int32)]
        L.spx_server_submit.argtypes = [vp, ctypes.c_int, ctypes.c_int32, vp, ctypes.c_uint32, vp,
                                        ctypes.POINTER(ctypes.c_uint64)]
        L.spx_server_wait.argtypes = [vp, ctypes.c_uint64, ctypes.c_int, vp,
                                      ctypes.POINTER(ctypes.c_int32)]
        L.spx_server_cancel.argtypes = [vp, ctypes.c_uint64]
        L.spx_server_stats.argtypes = [vp, ctypes.POINTER(ServerStats)]
        L.spx_server_reset_stats.argtypes = [vp]
        L.spx_server_destroy.argtypes = [vp]
        cfg = ServerConfig(hbm_budget_bytes, max_batch, slots, max_msg_bytes, batch_window_us,
                           max_keys_per_set, long_slots)
        h = ctypes.c_void_p()
        _check(L.spx_server_create(ctypes.byref(cfg), ctypes.byref(h)), "spx_server_create")
        self.h = h.value
        self._sig_bytes = {}

    def add_key(self, pid, sk, pk):
        kid = ctypes.c_int32()
        skb = None if sk is None else ctypes.create_string_buffer(bytes(sk))
        _check(lib().spx_server_add_key(self.h, pid, skb, ctypes.create_string_buffer(bytes(pk)),
                                        ctypes.byref(kid)), "spx_server_add_key")
        self._sig_bytes[kid.value] = params(pid).sig_bytes
        return kid.value

    def remove_key(self, key):
        """Refuse new requests on the key; queued / running ones finish and
        its row is reused afterwards (under a new key id)."""
        L = lib()
        L.spx_server_remove_key.argtypes = [ctypes.c_void_p, ctypes.c_int32]
        return _check(L.spx_server_remove_key(self.h, key), "spx_server_remove_key")

    def _submit(self, op, key, msg, aux):
        t = ctypes.c_uint64()
        m = np.frombuffer(msg, np.uint8) if len(msg) else np.zeros(1, np.uint8)
        a = None if aux is None else np.frombuffer(bytes(aux), np.uint8)
        rc = lib().spx_server_submit(self.h, op, key, _ptr(m), len(msg), _ptr(a), ctypes.byref(t))
        if rc != OK:
            raise CuspxError("spx_server_submit failed (%d)" % rc, rc)
        return t.value

    def submit_sign(self, key, msg, optrand=None):
        return self._submit(0, key, msg, optrand)

    def submit_verify(self, key, msg, sig):
        return self._submit(1, key, msg, sig)

    def wait(self, ticket, timeout_ms=-1, key=None):
        """(status, signature-or-None); status 0 ok / -1 invalid signature."""
        st = ctypes.c_int32()
        out = ctypes.create_string_buffer(64 * 1024)
        rc = lib().spx_server_wait(self.h, ticket, timeout_ms, out, ctypes.byref(st))
        if rc == 1:
            raise TimeoutError("request %d still running" % ticket)
        _check(rc, "spx_server_wait")
        n = self._sig_bytes.get(key, 0) if key is not None else 0
        return st.value, (out.raw[:n] if n else out.raw)

    def cancel(self, ticket):
        return lib().spx_server_cancel(self.h, ticket)

    def stats(self):
        s = ServerStats()
        _check(lib().spx_server_stats(self.h, ctypes.byref(s)), "spx_server_stats")
        return {f: getattr(s, f) for f, _ in ServerStats._fields_}

    def reset_stats(self):
        lib().spx_server_reset_stats(self.h)

    def latency(self, pid, op):
        """(p50_us, p99_us, count) of one (parameter set, op)."""
        L = lib()
        L.spx_server_latency.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int,
                                         ctypes.POINTER(ctypes.c_float),
                                         ctypes.POINTER(ctypes.c_float),
                                         ctypes.POINTER(ctypes.c_uint64)]
        a, b, c = ctypes.c_float(), ctypes.c_float(), ctypes.c_uint64()
        _check(L.spx_server_latency(self.h, pid, op, ctypes.byref(a), ctypes.byref(b),
                                    ctypes.byref(c)), "spx_server_latency")
        return a.value, b.value, c.value

    def close(self):
        if self.h:
            lib().spx_server_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


# ------------------------------------------------------- hipGraph replay

class SignGroup(ctypes.Structure):
    _fields_ = [("pid", ctypes.c_int), ("num", ctypes.c_uint32), ("sk", ctypes.c_void_p),
                ("sk_stride", ctypes.c_uint32), ("msg", ctypes.c_void_p),
                ("moff", ctypes.c_void_p), ("mlen", ctypes.c_void_p),
                ("optrand", ctypes.c_void_p), ("sig", ctypes.c_void_p),
                ("sig_stride", ctypes.c_uint64), ("ws", ctypes.c_void_p)]


class Graph:
    """An instantiated hipGraph of one batch's launch sequence."""

    def __init__(self, handle):
        self.h = handle

    def launch(self, stream=None):
        L = lib()
        L.cuspx_graph_launch.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
        _check(L.cuspx_graph_launch(self.h, stream.handle if stream else None),
               "cuspx_graph_launch")

    def close(self):
        if self.h:
            L = lib()
            L.cuspx_graph_destroy.argtypes = [ctypes.c_void_p]
            L.cuspx_graph_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def graph_sign(pid, num, sk, sk_stride, msg, moff, mlen, optrand, sig, sig_stride, ws):
    L = lib()
    vp = ctypes.c_void_p
    L.cuspx_graph_sign.argtypes = [ctypes.c_int, ctypes.c_uint32, vp, ctypes.c_uint32, vp, vp,
                                   vp, vp, vp, ctypes.c_uint64, vp, ctypes.POINTER(vp)]
    h = vp()
    _check(L.cuspx_graph_sign(pid, num, sk, sk_stride, msg, moff, mlen, optrand, sig,
                              sig_stride, ws, ctypes.byref(h)), "cuspx_graph_sign")
    return Graph(h.value)


def graph_verify(pid, num, pk, pk_stride, sig, sig_stride, msg, moff, mlen, status, ws):
    L = lib()
    vp = ctypes.c_void_p
    L.cuspx_graph_verify.argtypes = [ctypes.c_int, ctypes.c_uint32, vp, ctypes.c_uint32, vp,
                                     ctypes.c_uint64, vp, vp, vp, vp, vp, ctypes.POINTER(vp)]
    h = vp()
    _check(L.cuspx_graph_verify(pid, num, pk, pk_stride, sig, sig_stride, msg, moff, mlen,
                                status, ws, ctypes.byref(h)), "cuspx_graph_verify")
    return Graph(h.value)


def graph_sign_mixed(groups):
    """groups: list of SignGroup; one graph launch signs every group."""
    L = lib()
    vp = ctypes.c_void_p
    L.cuspx_graph_sign_mixed.argtypes = [ctypes.POINTER(SignGroup), ctypes.c_int,
                                         ctypes.POINTER(vp)]
    arr = (SignGroup * len(groups))(*groups)
    h = vp()
    _check(L.cuspx_graph_sign_mixed(arr, len(groups), ctypes.byref(h)), "cuspx_graph_sign_mixed")
    return Graph(h.value)


def sign_mixed_device(groups, stream=None):
    """Launch a mixed batch (list of SignGroup) directly on `stream`: one
    persistent kernel when the groups are distinct fast sets of one hash
    family, else the groups one after another."""
    L = lib()
    L.cuspx_sign_mixed_device.argtypes = [ctypes.POINTER(SignGroup), ctypes.c_int,
                                          ctypes.c_void_p]
    arr = (SignGroup * len(groups))(*groups)
    return _check(L.cuspx_sign_mixed_device(arr, len(groups), stream.handle if stream else None),
                  "cuspx_sign_mixed_device")


class MixedSignBatch:
    """A resident mixed-parameter-set signing batch: requests are grouped by
    set, each group gets device buffers and a workspace, and the whole batch
    is one hipGraph (one launch per replay).  Distinct fast sets of one hash
    family, each large enough for one or two lanes per WOTS leaf, run as one
    persistent kernel; other mixes run each set's pipeline on its own forked
    stream.  launch(eager=True) launches the same batch
    without the graph.

    items: list of (pid, sk bytes, message bytes, optrand bytes)."""

    def __init__(self, items):
        self.groups = {}
        for i, (pid, sk, m, opt) in enumerate(items):
            self.groups.setdefault(pid, []).append((i, sk, m, opt))
        self.n = len(items)
        self.bufs = []
        descs = []
        self.layout = []
        for pid, rows in sorted(self.groups.items()):
            p = params(pid)
            num = len(rows)
            lens = [len(m) for _, _, m, _ in rows]
            offs = np.zeros(num, np.uint64)
            pos = 0
            for j, ln in enumerate(lens):
                offs[j] = pos
                pos += (ln + 15) & ~15
            msgbuf = np.zeros(pos + 16, np.uint8)
            for j, (_, _, m, _) in enumerate(rows):
                msgbuf[int(offs[j]):int(offs[j]) + len(m)] = np.frombuffer(m, np.uint8)
            sk = np.frombuffer(b"".join(r[1] for r in rows), np.uint8)
            opt = np.frombuffer(b"".join(r[3] for r in rows), np.uint8)
            d = dict(sk=DeviceBuffer.from_array(sk), msg=DeviceBuffer.from_array(msgbuf),
                     moff=DeviceBuffer.from_array(offs),
                     mlen=DeviceBuffer.from_array(np.array(lens, np.uint32)),
                     opt=DeviceBuffer.from_array(opt), sig=DeviceBuffer(num * p.sig_bytes),
                     ws=DeviceBuffer(workspace_bytes(pid, num, OP_SIGN)))
            self.bufs.append(d)
            self.layout.append((pid, [r[0] for r in rows], p.sig_bytes))
            descs.append(SignGroup(pid, num, d["sk"].ptr, p.sk_bytes, d["msg"].ptr, d["moff"].ptr,
                                   d["mlen"].ptr, d["opt"].ptr, d["sig"].ptr, p.sig_bytes,
                                   d["ws"].ptr))
        self.descs = descs
        self.graph = graph_sign_mixed(descs)

    def launch(self, stream=None, eager=False):
        if eager:
            sign_mixed_device(self.descs, stream)
        else:
            self.graph.launch(stream)

    def signatures(self):
        out = [None] * self.n
        for d, (pid, idx, sb) in zip(self.bufs, self.layout):
            s = d["sig"].download().reshape(len(idx), sb)
            for j, i in enumerate(idx):
                out[i] = bytes(s[j])
        return out
