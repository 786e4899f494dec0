"""Host threads calling the library at the same time.

The reference's calls are not reentrant (one global midstate, the global
DRBG context and global device scratch, e.g. gpu_sha256.cu:1189,
rng.cu:14).  Here the host-buffer calls share a lock per device, the DRBG
has its own, and the device-resident calls only touch the caller's buffers
and stream.  Python threads calling through ctypes run the library
concurrently (ctypes releases the GIL); every result must equal the same
calls made one after another.  optrand is passed explicitly, so the order in
which threads reach the DRBG cannot change any byte.
"""
import random
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

import cuspx_amd as C
from oracle import oracle as O

pytestmark = pytest.mark.gpu

SETS = ["sha2-128f", "shake256-128f", "sha2-192f", "sha2-128s"]


def _inputs(name, num, seed):
    p = C.params(C.param_id(name))
    rng = random.Random(seed)
    seeds = np.frombuffer(rng.randbytes(3 * p.n * num), np.uint8)
    msgs = [rng.randbytes(rng.choice((0, 1, 33, 200, 700, 3000))) for _ in range(num)]
    opts = np.frombuffer(rng.randbytes(p.n * num), np.uint8)
    return seeds, msgs, opts


def _host_job(name, seed, num=24):
    pid = C.param_id(name)
    seeds, msgs, opts = _inputs(name, num, seed)
    pk, sk = C.batch_keygen(pid, seeds)
    sig = C.batch_sign(pid, sk, msgs, opts)
    bad = [m + b"!" for m in msgs[:4]] + msgs[4:]
    st = C.batch_verify(pid, pk, sig, bad)
    return pk, sk, sig, st


def test_host_buffer_calls_from_four_threads():
    want = {name: _host_job(name, 100 + k) for k, name in enumerate(SETS)}
    for name, (pk, sk, sig, st) in want.items():
        pid = C.param_id(name)
        seeds, msgs, opts = _inputs(name, 24, 100 + SETS.index(name))
        p = O.params(pid)
        assert bytes(sig[5]) == O.signature_with_optrand(pid, msgs[5], bytes(sk[5]),
                                                          bytes(opts[5 * p.n:6 * p.n])), name
        assert list(st[:4]) != [0] * 4 and list(st[4:]) == [0] * 20, name
    print("four threads x two rounds of keygen / sign / verify", flush=True)
    jobs = [(name, 100 + k) for _ in range(2) for k, name in enumerate(SETS)]
    with ThreadPoolExecutor(4) as ex:
        got = list(ex.map(lambda a: (a[0], _host_job(*a)), jobs))
    for name, res in got:
        for a, b in zip(res, want[name]):
            assert np.array_equal(a, b), name


def test_server_submitters_and_waiters_on_four_threads():
    """Four threads each submit 40 sign and 10 verify requests to one server
    (three parameter sets, messages 0 B .. 5 KiB, explicit optrand) and wait
    for them; every signature equals the host-buffer batch call's for the
    same inputs, and every verify status is exact."""
    names = ["sha2-128f", "sha2-128s", "shake256-256f"]
    keys = []
    for k, name in enumerate(names):
        pid = C.param_id(name)
        seeds, _, _ = _inputs(name, 2, 300 + k)
        pk, sk = C.batch_keygen(pid, seeds)
        keys += [(pid, bytes(sk[i]), bytes(pk[i])) for i in range(2)]
    work = []
    for t in range(4):
        rng = random.Random(400 + t)
        reqs = []
        for _ in range(40):
            k = rng.randrange(len(keys))
            n = C.params(keys[k][0]).n
            reqs.append((k, rng.randbytes(rng.choice((0, 17, 300, 2100, 5000))),
                         rng.randbytes(n)))
        work.append(reqs)
    want = {}
    for t, reqs in enumerate(work):
        for j, (k, m, opt) in enumerate(reqs):
            pid, sk, _ = keys[k]
            want[(t, j)] = bytes(C.batch_sign(pid, np.frombuffer(sk, np.uint8), [m],
                                              np.frombuffer(opt, np.uint8))[0])
    with C.Server(hbm_budget_bytes=4 << 30, slots=4, batch_window_us=300) as srv:
        kids = [srv.add_key(pid, sk, pk) for pid, sk, pk in keys]

        def client(t):
            reqs = work[t]
            ts = [srv.submit_sign(kids[k], m, opt) for k, m, opt in reqs]
            sigs = [srv.wait(x, key=kids[k]) for x, (k, _, _) in zip(ts, reqs)]
            # verify ten of them, every other one with its last byte flipped
            vs = []
            for j in range(10):
                k, m, _ = reqs[j]
                s = sigs[j][1] if j % 2 == 0 else sigs[j][1][:-1] + bytes([sigs[j][1][-1] ^ 1])
                vs.append(srv.submit_verify(kids[k], m, s))
            return sigs, [srv.wait(v)[0] for v in vs]

        print("four server clients", flush=True)
        with ThreadPoolExecutor(4) as ex:
            res = list(ex.map(client, range(4)))
        assert srv.stats()["last_error"] == 0
    for t, (sigs, vst) in enumerate(res):
        for j, (st, sig) in enumerate(sigs):
            assert st == 0 and sig == want[(t, j)], (t, j)
        assert vst == [0 if j % 2 == 0 else -1 for j in range(10)], t


def _device_job(seed, num=64):
    pid = C.param_id("sha2-128f")
    p = C.params(pid)
    seeds, msgs, opts = _inputs("sha2-128f", num, seed)
    pk, sk = C.batch_keygen(pid, seeds)
    D = C.DeviceBuffer
    moff = np.zeros(num, np.uint64)
    buf = bytearray()
    for i, m in enumerate(msgs):
        moff[i] = len(buf)
        buf += m + bytes(-len(m) % 4)
    d_msg = D.from_array(np.frombuffer(bytes(buf) + bytes(4), np.uint8))
    d_moff = D.from_array(moff)
    d_mlen = D.from_array(np.array([len(m) for m in msgs], np.uint32))
    d_sk = D.from_array(np.ascontiguousarray(sk).reshape(-1))
    d_opt = D.from_array(opts)
    d_sig = D(num * p.sig_bytes)
    ws = D(C.workspace_bytes(pid, num, C.OP_SIGN))
    st = C.Stream()
    for _ in range(3):
        C.sign_device(pid, num, d_sk.ptr, p.sk_bytes, d_msg.ptr, d_moff.ptr, d_mlen.ptr,
                      d_opt.ptr, d_sig.ptr, p.sig_bytes, ws.ptr, st.handle)
    st.sync()
    return d_sig.download().reshape(num, p.sig_bytes)


def test_device_calls_from_two_threads_on_their_own_streams():
    want = [_device_job(7), _device_job(8)]
    print("two threads signing on their own streams", flush=True)
    with ThreadPoolExecutor(2) as ex:
        got = list(ex.map(_device_job, [7, 8, 7, 8]))
    for k, g in enumerate(got):
        assert np.array_equal(g, want[k % 2]), k
