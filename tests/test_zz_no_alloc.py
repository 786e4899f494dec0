"""cuspx_sign_device_noalloc (workspace in the library's fixed device array)
against cuspx_sign_device (workspace from the caller) on the same inputs:
byte-identical signatures on two calls in a row, the second one smaller, so
the array's contents from the first call must not leak into the second; and
a batch whose workspace exceeds the array is refused with -5 before any
launch."""
import ctypes
import random

import numpy as np
import pytest

import cuspx_amd as C

pytestmark = pytest.mark.gpu

SCRATCH_BYTES = 128 << 20


def _noalloc():
    f = C.lib().cuspx_sign_device_noalloc
    vp = ctypes.c_void_p
    f.argtypes = [ctypes.c_int, ctypes.c_uint32, vp, ctypes.c_uint32, vp, vp, vp, vp, vp,
                  ctypes.c_uint64, vp]
    return f


def _device_batch(pid, num, seed):
    p = C.params(pid)
    rng = random.Random(seed)
    pk, sk = C.batch_keygen(pid, np.frombuffer(rng.randbytes(3 * p.n * num), np.uint8))
    msgs = [rng.randbytes(rng.choice((0, 1, 32, 77, 1000, 5000))) for _ in range(num)]
    offs = np.zeros(num, np.uint64)
    pos = 0
    for i, m in enumerate(msgs):
        offs[i] = pos
        pos += (len(m) + 15) & ~15
    buf = np.zeros(pos + 16, np.uint8)
    for i, m in enumerate(msgs):
        buf[int(offs[i]):int(offs[i]) + len(m)] = np.frombuffer(m, np.uint8)
    D = C.DeviceBuffer
    return dict(sk=D.from_array(sk), msg=D.from_array(buf), moff=D.from_array(offs),
                mlen=D.from_array(np.array([len(m) for m in msgs], np.uint32)),
                opt=D.from_array(np.frombuffer(rng.randbytes(p.n * num), np.uint8)))


@pytest.mark.parametrize("name,sizes", [("sha2-128f", (96, 17)), ("shake256-256f", (40, 3)),
                                        ("sha2-128s", (9, 2))])
def test_noalloc_matches_sign_device_twice(name, sizes):
    pid = C.param_id(name)
    p = C.params(pid)
    f = _noalloc()
    st = C.Stream()
    for k, num in enumerate(sizes):
        assert C.workspace_bytes(pid, num, C.OP_SIGN) <= SCRATCH_BYTES
        print("%s: call %d with %d requests" % (name, k, num), flush=True)
        b = _device_batch(pid, num, 100 * k + pid)
        ref = C.DeviceBuffer(num * p.sig_bytes)
        got = C.DeviceBuffer(num * p.sig_bytes)
        ref.fill(0)
        got.fill(0xEE)
        ws = C.DeviceBuffer(C.workspace_bytes(pid, num, C.OP_SIGN))
        C.sign_device(pid, num, b["sk"].ptr, p.sk_bytes, b["msg"].ptr, b["moff"].ptr,
                      b["mlen"].ptr, b["opt"].ptr, ref.ptr, p.sig_bytes, ws.ptr, st.handle)
        rc = f(pid, num, b["sk"].ptr, p.sk_bytes, b["msg"].ptr, b["moff"].ptr, b["mlen"].ptr,
               b["opt"].ptr, got.ptr, p.sig_bytes, st.handle)
        assert rc == C.OK, C.lib().cuspx_last_error()
        st.sync()
        assert bytes(got.download()) == bytes(ref.download()), (name, num)


def test_noalloc_refuses_batches_beyond_the_array():
    pid = C.param_id("sha2-128s")
    num = 1
    while C.workspace_bytes(pid, num, C.OP_SIGN) <= SCRATCH_BYTES:
        num *= 2
    f = _noalloc()
    dummy = C.DeviceBuffer(256)
    rc = f(pid, num, dummy.ptr, 64, dummy.ptr, dummy.ptr, dummy.ptr, dummy.ptr, dummy.ptr,
           7856, None)
    assert rc == C.ERR_BUDGET
    assert b"scratch" in C.lib().cuspx_last_error()
