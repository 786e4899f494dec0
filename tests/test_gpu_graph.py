"""hipGraph capture and replay: every replay of a captured batch must equal
the eager launch of the same inputs, with the program's own eager work and
synchronisations between replays and new inputs written between them."""
import random

import numpy as np
import pytest

import cuspx_amd as C
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _dev_batch(pid, num, seed):
    p = C.params(pid)
    rng = random.Random(seed)
    seeds = np.frombuffer(rng.randbytes(3 * p.n * num), np.uint8)
    pk, sk = C.batch_keygen(pid, seeds)
    msgs = np.frombuffer(rng.randbytes(32 * num), np.uint8)
    opts = np.frombuffer(rng.randbytes(p.n * num), np.uint8)
    return pk, sk, msgs, opts


@pytest.mark.parametrize("name,num", [("sha2-128s", 3), ("sha2-128f", 64), ("shake256-128f", 5)])
def test_graph_replays_equal_eager(name, num):
    pid = C.param_id(name)
    p = C.params(pid)
    pk, sk, msgs, opts = _dev_batch(pid, num, 7 + pid)
    D = C.DeviceBuffer
    d_sk, d_msg, d_opt = D.from_array(sk), D.from_array(msgs), D.from_array(opts)
    d_moff = D.from_array(np.arange(num, dtype=np.uint64) * 32)
    d_mlen = D.from_array(np.full(num, 32, np.uint32))
    d_sig = D(num * p.sig_bytes)
    ws = D(C.workspace_bytes(pid, num, C.OP_SIGN))
    g = C.graph_sign(pid, num, d_sk.ptr, p.sk_bytes, d_msg.ptr, d_moff.ptr, d_mlen.ptr,
                     d_opt.ptr, d_sig.ptr, p.sig_bytes, ws.ptr)
    st = C.Stream()
    rng = random.Random(99)
    first = None
    for rep in range(3):
        m = np.frombuffer(rng.randbytes(32 * num), np.uint8)
        d_msg.upload(m)
        # eager work between replays: a batched sign of other data
        C.batch_sign(pid, sk[:1], [b"eager %d" % rep], opts[:p.n])
        g.launch(st)
        st.sync()
        got = d_sig.download().reshape(num, p.sig_bytes)
        want = C.batch_sign(pid, sk, [bytes(m[32 * i:32 * i + 32]) for i in range(num)],
                            opts)
        assert (got == want).all(), rep
        if first is None:
            first = (bytes(m[:32]), bytes(got[0]))
    g.close()
    msg0, sig0 = first
    assert sig0 == O.signature_with_optrand(pid, msg0, bytes(sk[0]), bytes(opts[:p.n]))


def test_mixed_set_graph_equals_per_set_batches():
    rng = random.Random(5)
    items = []
    for name, cnt in (("shake256-128f", 5), ("shake256-192f", 4), ("sha2-128f", 3)):
        pid = C.param_id(name)
        p = C.params(pid)
        pk, sk = C.batch_keygen(pid, np.frombuffer(rng.randbytes(3 * p.n * cnt), np.uint8))
        for i in range(cnt):
            mlen = int(round(np.exp(rng.uniform(0, np.log(4096)))))
            items.append((pid, bytes(sk[i]), rng.randbytes(mlen), rng.randbytes(p.n)))
    rng.shuffle(items)
    mb = C.MixedSignBatch(items)
    for rep in range(2):
        mb.launch()
        C.lib().cuspx_stream_sync(None)
        sigs = mb.signatures()
        for (pid, sk, m, opt), s in zip(items, sigs):
            want = C.batch_sign(pid, np.frombuffer(sk, np.uint8), [m],
                                np.frombuffer(opt, np.uint8))
            assert s == bytes(want[0])


def test_graph_replays_with_message_lengths_changing():
    """A captured batch does not know its message lengths (they sit in device
    memory), so its prep launch keeps the two-wave workgroups of long SHA-256
    messages: replays whose messages move across the 512-byte threshold and
    up to 6000 bytes equal the eager batch every time."""
    pid = C.param_id("sha2-128f")
    p = C.params(pid)
    num, cap = 8, 6000
    pk, sk, _, opts = _dev_batch(pid, num, 31)
    D = C.DeviceBuffer
    d_sk, d_opt = D.from_array(sk), D.from_array(opts)
    d_msg = D(num * cap + 16)
    d_moff = D.from_array(np.arange(num, dtype=np.uint64) * cap)
    d_mlen = D.from_array(np.full(num, 32, np.uint32))
    d_sig = D(num * p.sig_bytes)
    ws = D(C.workspace_bytes(pid, num, C.OP_SIGN))
    g = C.graph_sign(pid, num, d_sk.ptr, p.sk_bytes, d_msg.ptr, d_moff.ptr, d_mlen.ptr,
                     d_opt.ptr, d_sig.ptr, p.sig_bytes, ws.ptr)
    st = C.Stream()
    rng = random.Random(5)
    for rep, lens in enumerate([[32] * num, [511, 512, 600, 5000, 1, 2047, 2048, 6000],
                                [6000] * num, [0, 3, 40, 100, 513, 4208, 5999, 64]]):
        msgs = [rng.randbytes(n) for n in lens]
        buf = bytearray(num * cap + 16)
        for i, m in enumerate(msgs):
            buf[i * cap:i * cap + len(m)] = m
        d_msg.upload(np.frombuffer(bytes(buf), np.uint8))
        d_mlen.upload(np.array(lens, np.uint32))
        C.batch_sign(pid, sk[:1], [b"eager %d" % rep], opts[:p.n])
        print("replay %d, lengths %s" % (rep, lens), flush=True)
        g.launch(st)
        st.sync()
        got = d_sig.download().reshape(num, p.sig_bytes)
        want = C.batch_sign(pid, sk, msgs, opts)
        assert (got == want).all(), rep
    g.close()
