"""GPU parity: the HIP engine against the CPU oracle, byte for byte.

Every case runs the HIP kernels through the C ABI (cuspx_amd -> libcuspx.so)
on the same seeded inputs as the oracle: distinct keys per item, messages
of uneven lengths, fixed optrand.  Invalid signatures must be rejected item
by item (test/spx.cu:159-166 of the reference flips a message bit).
"""
import random

import numpy as np
import pytest

import cuspx_amd as C
from oracle import oracle as O

pytestmark = pytest.mark.gpu

F_SETS = ["sha2-128f", "sha2-192f", "sha2-256f", "shake256-128f", "shake256-192f",
          "shake256-256f"]
S_SETS = ["sha2-128s", "sha2-192s", "sha2-256s", "shake256-128s", "shake256-192s",
          "shake256-256s"]


def _inputs(pid, num, seed, lens):
    p = O.params(pid)
    rng = random.Random(seed)
    seeds = [rng.randbytes(3 * p.n) for _ in range(num)]
    msgs = [rng.randbytes(lens[i % len(lens)]) for i in range(num)]
    opts = [rng.randbytes(p.n) for _ in range(num)]
    return seeds, msgs, opts


def _check_set(name, num, lens):
    pid = C.param_id(name)
    p = O.params(pid)
    seeds, msgs, opts = _inputs(pid, num, 1000 + pid, lens)
    pk, sk = C.batch_keygen(pid, np.frombuffer(b"".join(seeds), np.uint8))
    for i in range(num):
        opk, osk = O.seed_keypair(pid, seeds[i])
        assert bytes(pk[i]) == opk, (name, "pk", i)
        assert bytes(sk[i]) == osk, (name, "sk", i)
    sig = C.batch_sign(pid, sk, msgs, np.frombuffer(b"".join(opts), np.uint8))
    for i in range(num):
        want = O.signature_with_optrand(pid, msgs[i], bytes(sk[i]), opts[i])
        got = bytes(sig[i])
        if got != want:
            diff = [j for j in range(p.sig_bytes) if got[j] != want[j]]
            pytest.fail("%s item %d: %d bytes differ, first at %d" % (name, i, len(diff), diff[0]))
    st = C.batch_verify(pid, pk, sig, msgs)
    assert list(st) == [0] * num
    bad_msgs = list(msgs)
    bad_msgs[0] = bytes([bad_msgs[0][0] ^ 1]) + bad_msgs[0][1:] if bad_msgs[0] else b"\x01"
    bad_sig = sig.copy()
    if num > 1:
        bad_sig[1, p.n + 3] ^= 0x40  # a FORS secret-key byte
    st = C.batch_verify(pid, pk, bad_sig, bad_msgs)
    assert st[0] == -1
    if num > 1:
        assert st[1] == -1
    assert list(st[2:]) == [0] * (num - 2)


@pytest.mark.parametrize("name", F_SETS)
def test_fast_sets_keygen_sign_verify(name):
    _check_set(name, 6, [0, 1, 32, 33, 200, 4096])


@pytest.mark.parametrize("name", S_SETS)
def test_small_sets_keygen_sign_verify(name):
    _check_set(name, 2, [32, 5])


def test_sha2_128f_trace_roots_and_lengths_edge():
    # message lengths around the SHA-256 block boundaries of PRF_msg / H_msg
    pid = C.param_id("sha2-128f")
    p = O.params(pid)
    lens = [0, 3, 4, 31, 47, 48, 55, 56, 63, 64, 65, 127, 128, 1000, 65536]
    seeds, msgs, opts = _inputs(pid, len(lens), 77, lens)
    pk, sk = C.batch_keygen(pid, np.frombuffer(b"".join(seeds), np.uint8))
    sig = C.batch_sign(pid, sk, msgs, np.frombuffer(b"".join(opts), np.uint8))
    for i in range(len(lens)):
        assert bytes(sig[i]) == O.signature_with_optrand(pid, msgs[i], bytes(sk[i]), opts[i]), lens[i]
    assert list(C.batch_verify(pid, pk, sig, msgs)) == [0] * len(lens)


@pytest.mark.parametrize("name,lens", [
    # n = 16: PRF_msg absorbs 32 prefix bytes, H_msg 48, so their first
    # 136-byte blocks end after 104 / 88 message bytes; whole blocks follow
    ("shake256-128f", [0, 3, 4, 87, 88, 89, 103, 104, 105, 135, 136, 137, 223, 224, 225,
                       239, 240, 241, 511, 65536]),
    # n = 32: prefixes of 64 / 96 bytes, first blocks end after 72 / 40
    ("shake256-256f", [1, 39, 40, 41, 71, 72, 73, 175, 176, 177, 208, 4097]),
])
def test_shake_message_lengths_edge(name, lens):
    """Message lengths around the SHAKE256 rate-block boundaries of PRF_msg and
    H_msg (the prep kernels absorb on a lane pair, whole blocks directly and
    the head and tail word by word): every signature must match the oracle
    and verify."""
    pid = C.param_id(name)
    seeds, msgs, opts = _inputs(pid, len(lens), 78, lens)
    pk, sk = C.batch_keygen(pid, np.frombuffer(b"".join(seeds), np.uint8))
    sig = C.batch_sign(pid, sk, msgs, np.frombuffer(b"".join(opts), np.uint8))
    for i in range(len(lens)):
        assert bytes(sig[i]) == O.signature_with_optrand(pid, msgs[i], bytes(sk[i]), opts[i]), lens[i]
    assert list(C.batch_verify(pid, pk, sig, msgs)) == [0] * len(lens)


def test_api_surface_matches_oracle_debug_mode():
    # as shipped (DEBUG_MODE): keygen seed 0..3n-1, optrand 0xad (sign.cu:345-349, 1950-1954)
    pid = C.param_id("sha2-128f")
    C.set_default_param(pid)
    C.set_debug_mode(True)
    O.set_debug_mode(True)
    C.randombytes_init(bytes(range(48)))
    O.randombytes_init(bytes(range(48)))
    pk, sk = C.crypto_sign_keypair()
    opk, osk = O.keypair(pid)
    assert (pk, sk) == (opk, osk)
    m = C.randombytes(33)
    assert m == O.randombytes(33)
    sm = C.crypto_sign(m, sk)
    assert sm == O.sign(pid, m, osk)
    rc, mo = C.crypto_sign_open(sm, pk)
    assert rc == 0 and mo == m
    bad = bytearray(sm)
    bad[-1] ^= 1
    rc, mo = C.crypto_sign_open(bytes(bad), pk)
    assert rc == -1 and mo == b""
    rc, mo = C.crypto_sign_open(sm[:100], pk)  # smlen < SPX_BYTES
    assert rc == -1


def test_api_surface_drbg_mode_kat_count0():
    # KAT count 0 protocol (PQCgenKAT_sign.cu:55-131) in DRBG mode
    pid = C.param_id("sha2-128f")
    C.set_default_param(pid)
    C.set_debug_mode(False)
    O.set_debug_mode(False)
    try:
        C.randombytes_init(bytes(range(48)))
        seed = C.randombytes(48)
        msg = C.randombytes(33)
        C.randombytes_init(seed)
        O.randombytes_init(seed)
        pk, sk = C.crypto_sign_keypair()
        opk, osk = O.keypair(pid)
        assert (pk, sk) == (opk, osk)
        sm = C.crypto_sign(msg, sk)
        assert sm == O.sign(pid, msg, osk)
        assert C.crypto_sign_open(sm, pk) == (0, msg)
    finally:
        C.set_debug_mode(True)
        O.set_debug_mode(True)


def test_face_mdp_batch_surface():
    pid = C.param_id("sha2-128f")
    p = O.params(pid)
    C.set_default_param(pid)
    C.set_debug_mode(True)
    num = 5
    pk, sk = C.face_mdp_crypto_sign_keypair(num)
    opk, osk = O.seed_keypair(pid, bytes(range(3 * p.n)))
    assert all(bytes(pk[i]) == opk and bytes(sk[i]) == osk for i in range(num))
    rng = random.Random(5)
    m = np.frombuffer(rng.randbytes(32 * num), np.uint8)
    sm, smlen = C.face_mdp_crypto_sign(m, 32, sk, num)
    assert smlen == p.sig_bytes + 32
    for i in range(num):
        mi = bytes(m[32 * i:32 * i + 32])
        assert bytes(sm[i]) == O.sign(pid, mi, osk)
    rc, mo, mlen = C.face_mdp_crypto_sign_open(sm, smlen, pk, num)
    assert rc == 0 and mlen == 32
    assert bytes(mo.reshape(-1)) == bytes(m)
    sm2 = sm.copy()
    sm2[2, -1] ^= 1
    rc, mo, _ = C.face_mdp_crypto_sign_open(sm2, smlen, pk, num)
    assert rc == -1
    assert bytes(mo[2]) == bytes(32)
    assert bytes(mo[1]) == bytes(m[32:64])


def test_empty_messages_with_a_null_message_pointer():
    """A batch whose messages are all empty may pass msg = NULL (nothing is
    read from it): the signatures equal the oracle's for the empty message,
    and they verify with msg = NULL as well."""
    import ctypes
    pid = C.param_id("sha2-128f")
    p = O.params(pid)
    num = 3
    seeds, _, opts = _inputs(pid, num, 777, [0])
    pk, sk = C.batch_keygen(pid, np.frombuffer(b"".join(seeds), np.uint8))
    opt = np.frombuffer(b"".join(opts), np.uint8).reshape(num, p.n)
    moff = np.zeros(num, np.uint64)
    mlen = np.zeros(num, np.uint32)
    sig = np.zeros((num, p.sig_bytes), np.uint8)
    ptr = lambda a: a.ctypes.data_as(ctypes.c_void_p)
    assert C.lib().spx_batch_sign(pid, num, ptr(sk), None, ptr(moff), ptr(mlen), ptr(opt),
                                  ptr(sig)) == 0
    for i in range(num):
        assert bytes(sig[i]) == O.signature_with_optrand(pid, b"", bytes(sk[i]), opts[i]), i
    st = np.full(num, -7, np.int32)
    assert C.lib().spx_batch_verify(pid, num, ptr(pk), ptr(sig), None, ptr(moff), ptr(mlen),
                                    ptr(st)) == 0
    assert list(st) == [0] * num


def test_batch_determinism_and_order_independence():
    pid = C.param_id("sha2-128f")
    seeds, msgs, opts = _inputs(pid, 8, 9, [7, 32, 100])
    pk, sk = C.batch_keygen(pid, np.frombuffer(b"".join(seeds), np.uint8))
    opt = np.frombuffer(b"".join(opts), np.uint8).reshape(8, -1)
    s1 = C.batch_sign(pid, sk, msgs, opt)
    s2 = C.batch_sign(pid, sk, msgs, opt)
    assert (s1 == s2).all()
    perm = [3, 1, 7, 0, 5, 2, 6, 4]
    s3 = C.batch_sign(pid, sk[perm], [msgs[i] for i in perm], opt[perm])
    assert (s3 == s1[perm]).all()


@pytest.mark.parametrize("name", ["sha2-128f", "sha2-128s", "shake256-128f", "sha2-192f",
                                  "sha2-256f", "shake256-192f", "shake256-256f", "sha2-256s",
                                  "shake256-128s"])
def test_every_leaf_split_matches_oracle(name):
    pid = C.param_id(name)
    p = O.params(pid)
    num = 2
    seeds, msgs, opts = _inputs(pid, num, 555 + pid, [45, 3])
    want_pk = [O.seed_keypair(pid, s) for s in seeds]
    want_sig = [O.signature_with_optrand(pid, msgs[i], want_pk[i][1], opts[i]) for i in range(num)]
    try:
        for S in (1, 2, 4, 8, 16):
            C.set_split(pid, C.OP_SIGN, S)
            C.set_split(pid, C.OP_KEYGEN, S)
            # subtrees wider than one workgroup (s sets at S >= 4) run split in parts
            assert C.get_split(pid, C.OP_SIGN, num) == S
            pk, sk = C.batch_keygen(pid, np.frombuffer(b"".join(seeds), np.uint8))
            assert [bytes(x) for x in pk] == [w[0] for w in want_pk], S
            sig = C.batch_sign(pid, sk, msgs, np.frombuffer(b"".join(opts), np.uint8))
            assert [bytes(x) for x in sig] == want_sig, S
    finally:
        C.set_split(pid, C.OP_SIGN, 0)
        C.set_split(pid, C.OP_KEYGEN, 0)


@pytest.mark.gpu
@pytest.mark.parametrize("name", F_SETS + S_SETS)
def test_every_verify_width(name):
    """k_vht at every width (256 / G requests per workgroup, their chains dealt
    over its lanes) and the speculative pair (G = 1) accept the valid
    signatures and reject one tampered byte in each part of the hypertree
    section; six requests leave most of each workgroup's request slots empty."""
    pid = C.param_id(name)
    p = O.params(pid)
    num = 6
    seeds, msgs, opts = _inputs(pid, num, 777 + pid, [32, 0, 100])
    pk, sk = C.batch_keygen(pid, np.frombuffer(b"".join(seeds), np.uint8))
    sig = C.batch_sign(pid, sk, msgs, np.frombuffer(b"".join(opts), np.uint8))
    assert bytes(sig[0]) == O.signature_with_optrand(pid, msgs[0], bytes(sk[0]), opts[0])
    th = p.full_h // p.d
    ht = p.n + p.k * (p.a + 1) * p.n                 # first hypertree byte
    layer_bytes = (p.len + th) * p.n
    bad = sig.copy()
    bad[1, ht + 5] ^= 1                              # layer 0, first WOTS chain
    bad[2, ht + (p.len - 1) * p.n] ^= 2              # layer 0, last (checksum) chain
    bad[3, ht + p.len * p.n + 1] ^= 4                # layer 0, auth path
    bad[4, ht + (p.d - 1) * layer_bytes + 7] ^= 8    # top layer, WOTS
    widths = [1] + [G for G in (16, 32, 64, 128) if G <= (128 if p.len > 64 else 64)]
    try:
        for G in widths:
            C.set_split(pid, C.OP_VERIFY, G)
            assert C.get_split(pid, C.OP_VERIFY, num) == G
            assert list(C.batch_verify(pid, pk, sig, msgs)) == [0] * num, G
            assert list(C.batch_verify(pid, pk, bad, msgs)) == [0, -1, -1, -1, -1, 0], G
    finally:
        C.set_split(pid, C.OP_VERIFY, 0)


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["sha2-128f", "sha2-128s", "sha2-256f", "shake256-128f",
                                  "shake256-192s"])
def test_every_signature_hash_is_checked(name):
    """One bit flipped in each n-byte hash of a signature (R, every FORS
    secret and auth node, every WOTS chain value, every hypertree auth node)
    must fail verification: the test the reference keeps behind the never
    defined SPX_TEST_INVALIDSIG (test/spx.cu:168-184), here as one batch of
    sig_bytes / n tampered copies plus the original, which must verify.  A
    sample of the copies is also rejected by the oracle."""
    pid = C.param_id(name)
    p = O.params(pid)
    seeds, msgs, opts = _inputs(pid, 1, 4321 + pid, [57])
    pk, sk = C.batch_keygen(pid, np.frombuffer(seeds[0], np.uint8))
    sig = C.batch_sign(pid, sk, msgs, np.frombuffer(opts[0], np.uint8))[0]
    segs = p.sig_bytes // p.n
    bad = np.tile(sig, (segs + 1, 1))
    for k in range(segs):  # segment k: byte k % n of it, bit k % 8
        bad[k, k * p.n + k % p.n] ^= 1 << (k % 8)
    st = C.batch_verify(pid, np.tile(pk[0], (segs + 1, 1)), bad, msgs * (segs + 1))
    assert st[segs] == 0
    passed = [k for k in range(segs) if st[k] == 0]
    assert passed == [], "%s: %d of %d tampered hashes verified, first %s" % (
        name, len(passed), segs, passed[:8])
    for k in range(0, segs, max(1, segs // 12)):
        assert O.verify(pid, bytes(bad[k]), msgs[0], bytes(pk[0])) != 0, k


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["sha2-128f", "shake256-256f"])
def test_verify_ragged_workgroups(name):
    """37 requests: at G = 16 two full workgroups and one with 5 of 16 request
    slots, at G = 32 four full and one with 5 of 8, plus the speculative pair.
    Tampered requests sit in every workgroup; each status must be exact."""
    pid = C.param_id(name)
    p = O.params(pid)
    num = 37
    seeds, msgs, opts = _inputs(pid, num, 4711 + pid, [32, 1, 77])
    pk, sk = C.batch_keygen(pid, np.frombuffer(b"".join(seeds), np.uint8))
    sig = C.batch_sign(pid, sk, msgs, np.frombuffer(b"".join(opts), np.uint8))
    ht = p.n + p.k * (p.a + 1) * p.n
    layer_bytes = (p.len + p.full_h // p.d) * p.n
    bad = sig.copy()
    want = [0] * num
    for i, layer in ((3, 0), (15, 5), (16, p.d - 1), (20, 2), (36, 1)):
        bad[i, ht + layer * layer_bytes + 3 * p.n + 9] ^= 0x40
        want[i] = -1
    try:
        for G in (1, 16, 32):
            C.set_split(pid, C.OP_VERIFY, G)
            assert list(C.batch_verify(pid, pk, sig, msgs)) == [0] * num, G
            assert list(C.batch_verify(pid, pk, bad, msgs)) == want, G
    finally:
        C.set_split(pid, C.OP_VERIFY, 0)


@pytest.mark.gpu
@pytest.mark.parametrize("name,num,S", [("sha2-128f", 1600, 1), ("shake256-128f", 1600, 1),
                                        ("sha2-192f", 300, 1), ("sha2-256f", 300, 1),
                                        ("sha2-128s", 300, 2)])
def test_large_batch_launch_paths(name, num, S):
    """Batches big enough for the full-chip launch shapes: the fused persistent
    signing kernel (SHA-256; at S = 1 its FORS tiles hold 32-64 trees, so a
    tile spans request boundaries), the separate FORS and hypertree kernels
    (SHAKE at a full chip, 128s with 512/1024-lane subtrees).  A sample is
    compared with the oracle byte for byte; every signature must verify."""
    pid = C.param_id(name)
    p = O.params(pid)
    rng = random.Random(4242 + pid)
    seeds = [rng.randbytes(3 * p.n) for _ in range(num)]
    msgs = [rng.randbytes(rng.randrange(0, 200)) for _ in range(num)]
    opts = [rng.randbytes(p.n) for _ in range(num)]
    try:
        C.set_split(pid, C.OP_SIGN, S)
        pk, sk = C.batch_keygen(pid, np.frombuffer(b"".join(seeds), np.uint8))
        sig = C.batch_sign(pid, sk, msgs, np.frombuffer(b"".join(opts), np.uint8))
    finally:
        C.set_split(pid, C.OP_SIGN, 0)
    for i in (0, num // 2, num - 1):
        opk, osk = O.seed_keypair(pid, seeds[i])
        assert bytes(pk[i]) == opk and bytes(sk[i]) == osk, i
        assert bytes(sig[i]) == O.signature_with_optrand(pid, msgs[i], osk, opts[i]), i
    assert list(C.batch_verify(pid, pk, sig, msgs)) == [0] * num


@pytest.mark.gpu
@pytest.mark.parametrize("name,S", [("sha2-128f", 2), ("shake256-192f", 1), ("sha2-128s", 2)])
def test_fors_pk_inside_and_outside_the_fors_tiles(name, S):
    """Sign batches of up to 256 requests compute each FORS public key on the
    lane that stores the request's last FORS root; larger batches launch a
    separate T_k kernel.  The same 257 requests signed as one batch (separate
    T_k) and as batches of 256 and 1 (T_k in the tiles; 128s at S = 2 runs
    the separate FORS and hypertree kernels) must give identical bytes, and
    a sample must match the oracle."""
    pid = C.param_id(name)
    num = 257
    seeds, msgs, opts = _inputs(pid, num, 9090 + pid, [32, 5, 300])
    opt = np.frombuffer(b"".join(opts), np.uint8)
    try:
        C.set_split(pid, C.OP_SIGN, S)
        pk, sk = C.batch_keygen(pid, np.frombuffer(b"".join(seeds), np.uint8))
        whole = C.batch_sign(pid, sk, msgs, opt)
        p = O.params(pid)
        head = C.batch_sign(pid, sk[:256], msgs[:256], opt[:256 * p.n])
        tail = C.batch_sign(pid, sk[256:], msgs[256:], opt[256 * p.n:])
    finally:
        C.set_split(pid, C.OP_SIGN, 0)
    assert np.array_equal(whole[:256], head)
    assert np.array_equal(whole[256:], tail)
    for i in (0, 255, 256):
        assert bytes(whole[i]) == O.signature_with_optrand(pid, msgs[i], bytes(sk[i]), opts[i]), i


def _device_sign_verify(pid, sk, pk, msgs, opts):
    """Sign and verify through the device entry points, whose message lengths
    the library does not see on the host (every prep path stays possible)."""
    p = O.params(pid)
    num = len(msgs)
    moff = np.zeros(num, np.uint64)
    buf = bytearray()
    for i, m in enumerate(msgs):  # 4-byte aligned records
        moff[i] = len(buf)
        buf += m + bytes(-len(m) % 4)
    D = C.DeviceBuffer
    d_msg = D.from_array(np.frombuffer(bytes(buf) + bytes(4), np.uint8))
    d_moff, d_mlen = D.from_array(moff), D.from_array(np.array([len(m) for m in msgs], np.uint32))
    d_sk = D.from_array(np.ascontiguousarray(sk, np.uint8).reshape(-1))
    d_pk = D.from_array(np.ascontiguousarray(pk, np.uint8).reshape(-1))
    d_opt = D.from_array(np.frombuffer(b"".join(opts), np.uint8))
    d_sig, d_st = D(num * p.sig_bytes), D(4 * num)
    ws = D(max(C.workspace_bytes(pid, num, op) for op in (0, 1)))
    st = C.Stream()
    C.sign_device(pid, num, d_sk.ptr, p.sk_bytes, d_msg.ptr, d_moff.ptr, d_mlen.ptr, d_opt.ptr,
                  d_sig.ptr, p.sig_bytes, ws.ptr, st.handle)
    d_st.fill(0x7F)
    C.verify_device(pid, num, d_pk.ptr, p.pk_bytes, d_sig.ptr, p.sig_bytes, d_msg.ptr,
                    d_moff.ptr, d_mlen.ptr, d_st.ptr, ws.ptr, st.handle)
    st.sync()
    return d_sig.download().reshape(num, p.sig_bytes), d_st.download(np.int32)


@pytest.mark.parametrize("name", ["sha2-128f", "sha2-192f", "sha2-256f"])
def test_sha2_long_messages_two_wave_prep(name):
    """SHA-256 messages of 512 bytes or more are hashed by the prep kernels'
    two-wave workgroups (wave 1 computes the schedules of the message-only
    blocks, 32 per stage, while wave 0 runs the rounds).  Lengths around that
    threshold, around 32 / 33 and 64 / 65 message-only blocks in either
    message hash (n = 16, 24, 32 put 4 / 6 / 8 prefix words before the
    message in PRF_msg and 12 / 18 / 24 in H_msg), with partial last words,
    mixed with short messages in one batch: every signature equals the
    oracle's on the host-buffer path (message lengths known to the library)
    and the device path (not known), every one verifies, and a flipped
    message byte fails exactly its item."""
    pid = C.param_id(name)
    p = O.params(pid)
    lens = [511, 512, 513, 515, 0, 2064, 2096, 2127, 33, 2159, 2160, 4144, 4207, 4208,
            1000, 8195, 65535]
    seeds, msgs, opts = _inputs(pid, len(lens), 91 + pid, lens)
    pk, sk = C.batch_keygen(pid, np.frombuffer(b"".join(seeds), np.uint8))
    print("%s: host-buffer sign of %d (lengths %s)" % (name, len(lens), lens), flush=True)
    sig = C.batch_sign(pid, sk, msgs, np.frombuffer(b"".join(opts), np.uint8))
    for i, n in enumerate(lens):
        assert bytes(sig[i]) == O.signature_with_optrand(pid, msgs[i], bytes(sk[i]), opts[i]), n
    assert list(C.batch_verify(pid, pk, sig, msgs)) == [0] * len(lens)
    bad = [bytearray(m) for m in msgs]
    flip = [i for i, n in enumerate(lens) if n >= 512][::3]
    for i in flip:
        bad[i][lens[i] // 2] ^= 0x10
    st = C.batch_verify(pid, pk, sig, [bytes(m) for m in bad])
    assert [i for i in range(len(lens)) if st[i]] == flip
    print("%s: device-path sign + verify" % name, flush=True)
    dsig, dst = _device_sign_verify(pid, sk, pk, msgs, opts)
    assert all(bytes(dsig[i]) == bytes(sig[i]) for i in range(len(lens)))
    assert list(dst) == [0] * len(lens)


def test_sha2_long_messages_beyond_the_long_workgroups():
    """At most 512 two-wave workgroups run per launch, each taking the long
    requests i, i + 512, ...: a 600-request batch whose long messages sit at
    indices 3, 515 (the same workgroup as 3) and 599, among 1-byte ones."""
    pid = C.param_id("sha2-128f")
    p = O.params(pid)
    num = 600
    lens = [1] * num
    for i, n in ((3, 5000), (515, 512), (599, 3001)):
        lens[i] = n
    seeds, msgs, opts = _inputs(pid, num, 92, lens)
    pk, sk = C.batch_keygen(pid, np.frombuffer(b"".join(seeds), np.uint8))
    print("sha2-128f: device-path sign + verify of %d" % num, flush=True)
    dsig, dst = _device_sign_verify(pid, sk, pk, msgs, opts)
    for i in (0, 3, 4, 514, 515, 598, 599):
        assert bytes(dsig[i]) == O.signature_with_optrand(pid, msgs[i], bytes(sk[i]), opts[i]), i
    assert list(dst) == [0] * num
    sig = C.batch_sign(pid, sk, msgs, np.frombuffer(b"".join(opts), np.uint8))
    assert np.array_equal(sig, dsig)


@pytest.mark.parametrize("name", ["sha2-128f", "shake256-128f"])
def test_one_mebibyte_messages(name):
    """Messages of 1 MiB and 1 MiB + 1 bytes (the reference hashes a 1 MiB
    patterned input, test/sha256.cu:142-182): PRF_msg and H_msg each absorb
    16384 SHA-256 blocks (the two-wave path) or 7711 SHAKE256 rate blocks,
    through the host-buffer and the device entry points; signatures equal
    the oracle's and verify, and a flipped last message byte fails."""
    pid = C.param_id(name)
    p = O.params(pid)
    pattern = bytes((i * 7 + 3) & 0xFF for i in range(256))
    msgs = [pattern * 4096, pattern * 4096 + b"\x5a"]
    rng = random.Random(1 << 20)
    seeds = [rng.randbytes(3 * p.n) for _ in msgs]
    opts = [rng.randbytes(p.n) for _ in msgs]
    pk, sk = C.batch_keygen(pid, np.frombuffer(b"".join(seeds), np.uint8))
    print("%s: sign of two 1 MiB messages" % name, flush=True)
    sig = C.batch_sign(pid, sk, msgs, np.frombuffer(b"".join(opts), np.uint8))
    for i, m in enumerate(msgs):
        assert bytes(sig[i]) == O.signature_with_optrand(pid, m, bytes(sk[i]), opts[i]), i
    dsig, dst = _device_sign_verify(pid, sk, pk, msgs, opts)
    assert np.array_equal(dsig, sig) and list(dst) == [0, 0]
    bad = [m[:-1] + bytes([m[-1] ^ 1]) for m in msgs]
    assert all(s != 0 for s in C.batch_verify(pid, pk, sig, bad))


def test_sha2_small_set_long_messages():
    """The two-wave prep path on a small set (SHA2-128s: its own digest split
    into FORS indices and a 54-bit tree index): messages at the 512-byte
    threshold, at a stage boundary and at 64 KiB, on the host-buffer and the
    device path, equal to the oracle's and verified."""
    pid = C.param_id("sha2-128s")
    lens = [511, 512, 4208, 65535]
    seeds, msgs, opts = _inputs(pid, len(lens), 93, lens)
    pk, sk = C.batch_keygen(pid, np.frombuffer(b"".join(seeds), np.uint8))
    print("sha2-128s: host-buffer sign of %d (lengths %s)" % (len(lens), lens), flush=True)
    sig = C.batch_sign(pid, sk, msgs, np.frombuffer(b"".join(opts), np.uint8))
    for i, n in enumerate(lens):
        assert bytes(sig[i]) == O.signature_with_optrand(pid, msgs[i], bytes(sk[i]), opts[i]), n
    assert list(C.batch_verify(pid, pk, sig, msgs)) == [0] * len(lens)
    dsig, dst = _device_sign_verify(pid, sk, pk, msgs, opts)
    assert np.array_equal(dsig, sig)
    assert list(dst) == [0] * len(lens)
