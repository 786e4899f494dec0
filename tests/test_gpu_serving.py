"""GPU tests of the continuous-batching server: keys resident on the GPU,
mixed parameter sets and message sizes; every signature must be identical
whether its request ran alone or inside a mixed stream of concurrent
requests, and equal to the oracle's."""
import random

import numpy as np
import pytest

import cuspx_amd as C
from oracle import oracle as O

pytestmark = pytest.mark.gpu

SETS = ["sha2-128f", "sha2-128s", "shake256-256f"]


def _keys():
    out = []
    rng = random.Random(31)
    for name in SETS:
        pid = C.param_id(name)
        p = O.params(pid)
        seeds = [rng.randbytes(3 * p.n) for _ in range(2)]
        pk, sk = C.batch_keygen(pid, np.frombuffer(b"".join(seeds), np.uint8))
        for i in range(2):
            out.append((pid, bytes(sk[i]), bytes(pk[i])))
    return out


def test_server_alone_equals_mixed_and_oracle():
    keys = _keys()
    rng = random.Random(32)
    with C.Server(hbm_budget_bytes=2 << 30, slots=3, batch_window_us=500) as srv:
        kids = [srv.add_key(pid, sk, pk) for pid, sk, pk in keys]
        reqs = []
        for i in range(24):
            k = i % len(keys)
            pid = keys[k][0]
            p = O.params(pid)
            mlen = int(round(np.exp(rng.uniform(np.log(16), np.log(4096)))))
            reqs.append((k, rng.randbytes(mlen), rng.randbytes(p.n)))
        alone = []
        for k, m, opt in reqs[:6]:
            st, sig = srv.wait(srv.submit_sign(kids[k], m, opt), key=kids[k])
            assert st == 0
            alone.append(sig)
        # oracle for one request of each parameter set (requests 0, 2, 4)
        for (k, m, opt), sig in list(zip(reqs[:6], alone))[:6:2]:
            pid, sk, pk = keys[k]
            assert sig == O.signature_with_optrand(pid, m, sk, opt)
        # the same requests inside a mixed stream, plus verifies in flight
        tickets = [(srv.submit_sign(kids[k], m, opt), k) for k, m, opt in reqs]
        vt = [(srv.submit_verify(kids[k], reqs[j][1], alone[j]), j)
              for j, (k, _, _) in enumerate(reqs[:6])]
        bad = bytearray(alone[0])
        bad[40] ^= 1
        vbad = srv.submit_verify(kids[reqs[0][0]], reqs[0][1], bytes(bad))
        mixed = [srv.wait(t, key=kids[k]) for t, k in tickets]
        assert all(st == 0 for st, _ in mixed)
        for j in range(6):
            assert mixed[j][1] == alone[j], j
        for t, j in vt:
            assert srv.wait(t)[0] == 0
        assert srv.wait(vbad)[0] == -1
        # every mixed signature verifies with the batch API
        for (k, m, opt), (st, sig) in zip(reqs, mixed):
            pid, sk, pk = keys[k]
            assert list(C.batch_verify(pid, np.frombuffer(pk, np.uint8),
                                       np.frombuffer(sig, np.uint8), [m])) == [0]
        s = srv.stats()
        assert s["signs"] >= 30 and s["verifies"] == 7
        assert s["sign_p99_us"] > 0


def test_server_long_message_slot_cap():
    """With long_slots = 1 of 2 slots, requests over 16 KiB queue behind one
    another while short ones keep the other slot; every result still equals
    the batch API's signature for the same key, message and optrand."""
    keys = _keys()
    rng = random.Random(77)
    reqs = []
    for i in range(12):
        k = i % len(keys)
        p = O.params(keys[k][0])
        mlen = rng.randrange(17 << 10, 40 << 10) if i % 2 == 0 else rng.randrange(1, 3000)
        reqs.append((k, rng.randbytes(mlen), rng.randbytes(p.n)))
    with C.Server(hbm_budget_bytes=2 << 30, slots=2, batch_window_us=300, long_slots=1) as srv:
        kids = [srv.add_key(pid, sk, pk) for pid, sk, pk in keys]
        tickets = [(srv.submit_sign(kids[k], m, opt), k) for k, m, opt in reqs]
        got = [srv.wait(t, key=kids[k]) for t, k in tickets]
    for (k, m, opt), (st, sig) in zip(reqs, got):
        pid, sk, pk = keys[k]
        assert st == 0
        ref = C.batch_sign(pid, np.frombuffer(sk, np.uint8), [m], np.frombuffer(opt, np.uint8))
        assert sig == bytes(ref[0])


def test_server_admission_and_cancel():
    pid = C.param_id("sha2-128f")
    p = O.params(pid)
    pk, sk = C.batch_keygen(pid, np.zeros(3 * p.n, np.uint8))
    with C.Server(hbm_budget_bytes=1 << 30, slots=2, max_msg_bytes=1024) as srv:
        kid = srv.add_key(pid, bytes(sk[0]), bytes(pk[0]))
        with pytest.raises(C.CuspxError):
            srv.submit_sign(kid, b"x" * 2048)  # above max_msg_bytes: refused
        with pytest.raises(C.CuspxError):
            srv.submit_sign(kid + 7, b"x")  # unknown key row
        ts = [srv.submit_sign(kid, b"m%d" % i) for i in range(20)]
        srv.cancel(ts[5])
        for i, t in enumerate(ts):
            if i == 5:
                continue
            assert srv.wait(t, key=kid)[0] == 0


def test_server_refusals():
    """A sign request on a key added without a secret key is refused (-4),
    verifies on it run; a request that can never fit one slot's share of the
    budget is refused (-5)."""
    pid = C.param_id("sha2-128f")
    p = O.params(pid)
    pk, sk = C.batch_keygen(pid, np.frombuffer(bytes(range(3 * p.n)), np.uint8))
    sig = C.batch_sign(pid, sk, [b"verify-only"], np.zeros(p.n, np.uint8))
    with C.Server(hbm_budget_bytes=1 << 30, slots=2) as srv:
        vk = srv.add_key(pid, None, bytes(pk[0]))
        with pytest.raises(C.CuspxError) as e:
            srv.submit_sign(vk, b"m")
        assert e.value.code == C.ERR_ARG
        assert srv.wait(srv.submit_verify(vk, b"verify-only", bytes(sig[0])))[0] == 0
        assert srv.wait(srv.submit_verify(vk, b"verify-Only", bytes(sig[0])))[0] == -1
    # 4 MiB over 4 slots: a SHA2-128s signing workspace (about 2 MiB) never fits
    pid_s = C.param_id("sha2-128s")
    ps = O.params(pid_s)
    assert C.workspace_bytes(pid_s, 1, C.OP_SIGN) > (1 << 20)
    pk_s, sk_s = C.batch_keygen(pid_s, np.frombuffer(bytes(3 * ps.n), np.uint8))
    with C.Server(hbm_budget_bytes=4 << 20, slots=4) as srv:
        k = srv.add_key(pid_s, bytes(sk_s[0]), bytes(pk_s[0]))
        with pytest.raises(C.CuspxError) as e:
            srv.submit_sign(k, b"m")
        assert e.value.code == C.ERR_BUDGET


def test_server_wait_timeout_and_cancel_before_launch():
    """One slot runs a long SHAKE256-256f request (a 64 KiB message); a
    second request queues behind it.  A wait with timeout 0 on the running
    request returns 1 (TimeoutError) and leaves it live; cancelling the
    queued request removes it from the queue at once (pending drops to 0)
    and wakes the thread already waiting on it with -4; the running request
    then completes normally."""
    import threading
    pid_l = C.param_id("shake256-256f")
    pl = O.params(pid_l)
    pk_l, sk_l = C.batch_keygen(pid_l, np.frombuffer(bytes(range(3 * pl.n)), np.uint8))
    pid = C.param_id("sha2-128f")
    p = O.params(pid)
    pk, sk = C.batch_keygen(pid, np.frombuffer(bytes(range(3 * p.n)), np.uint8))
    long_msg = bytes(range(256)) * 256
    with C.Server(hbm_budget_bytes=1 << 30, slots=1, batch_window_us=0) as srv:
        kl = srv.add_key(pid_l, bytes(sk_l[0]), bytes(pk_l[0]))
        ks = srv.add_key(pid, bytes(sk[0]), bytes(pk[0]))
        import time
        t_long = srv.submit_sign(kl, long_msg, bytes(pl.n))
        deadline = time.time() + 5
        while srv.stats()["pending"] and time.time() < deadline:
            time.sleep(1e-4)  # the long request has left the queue for the slot
        t_q = srv.submit_sign(ks, b"queued", bytes(p.n))
        assert srv.stats()["pending"] == 1
        with pytest.raises(TimeoutError):
            srv.wait(t_long, timeout_ms=0)
        got = {}

        def waiter():
            try:
                got["r"] = srv.wait(t_q, timeout_ms=20000)
            except C.CuspxError as e:
                got["code"] = e.code

        th = threading.Thread(target=waiter)
        th.start()
        assert srv.cancel(t_q) == 0  # still queued behind the long request
        th.join(timeout=10)
        assert not th.is_alive() and got.get("code") == C.ERR_ARG
        assert srv.stats()["pending"] == 0
        st, sig = srv.wait(t_long, key=kl)
        assert st == 0
        assert sig == O.signature_with_optrand(pid_l, long_msg, bytes(sk_l[0]), bytes(pl.n))
        # a finished request cannot be cancelled any more
        t2 = srv.submit_sign(ks, b"after", bytes(p.n))
        assert srv.wait(t2, key=ks)[0] == 0
        assert srv.cancel(t2) == C.ERR_ARG


def test_server_remove_key_reuses_its_row():
    """A removed key refuses new requests (-4) while its queued request still
    completes; with one row per set, the next key can only be added once that
    request has drained, and it reuses the row under a new id, which the old
    id cannot reach."""
    pid = C.param_id("sha2-128f")
    p = O.params(pid)
    pk, sk = C.batch_keygen(pid, np.frombuffer(bytes(range(2 * 3 * p.n)), np.uint8))
    with C.Server(hbm_budget_bytes=1 << 30, slots=1, max_keys_per_set=1) as srv:
        k0 = srv.add_key(pid, bytes(sk[0]), bytes(pk[0]))
        t = srv.submit_sign(k0, b"before removal", bytes(p.n))
        srv.remove_key(k0)
        with pytest.raises(C.CuspxError) as e:
            srv.submit_sign(k0, b"after removal", bytes(p.n))
        assert e.value.code == C.ERR_ARG
        st, sig = srv.wait(t, key=k0)
        assert st == 0
        assert sig == O.signature_with_optrand(pid, b"before removal", bytes(sk[0]), bytes(p.n))
        k1 = srv.add_key(pid, bytes(sk[1]), bytes(pk[1]))  # the drained row
        assert k1 != k0
        with pytest.raises(C.CuspxError):
            srv.submit_sign(k0, b"stale id", bytes(p.n))
        with pytest.raises(C.CuspxError):
            srv.remove_key(k0)
        st, sig = srv.wait(srv.submit_sign(k1, b"new key", bytes(p.n)), key=k1)
        assert st == 0
        assert sig == O.signature_with_optrand(pid, b"new key", bytes(sk[1]), bytes(p.n))
        with pytest.raises(C.CuspxError) as e:  # no free row for a third key
            srv.add_key(pid, bytes(sk[0]), bytes(pk[0]))
        assert e.value.code == C.ERR_BUDGET
