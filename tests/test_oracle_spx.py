"""Pins the oracle's SPHINCS+ layer.

Closed forms with hashlib for the tweakable hashes (the SHA-256 address is
22 bytes, sha256_offsets.h:9-17; SHAKE absorbs all 32 address bytes,
thash_shake256.cu:20-24), brute-force Merkle trees for treehash and
compute_root, and the reference's own round-trip invariants
(test/wots.cu:39-48, test/fors.cu:35-42, test/spx.cu:127-166).
"""
import hashlib
import hmac
import random
import struct

import pytest

from oracle import oracle as O

SHA2_F = O.param_id("sha2-128f")
ALL = list(range(12))
FAST = [O.param_id(x) for x in ("sha2-128f", "sha2-192f", "sha2-256f", "shake256-128f",
                                 "shake256-192f", "shake256-256f")]


def rb(rng, n):
    return rng.randbytes(n)


def test_param_table_sizes():
    # signature sizes of the six round-3 sets (params.h:132-134)
    expect = {"128s": 7856, "128f": 17088, "192s": 16224, "192f": 35664,
              "256s": 29792, "256f": 49856}
    for pid in ALL:
        p = O.params(pid)
        size = p.name.decode().split("-")[1]
        assert p.sig_bytes == expect[size]
        assert p.pk_bytes == 2 * p.n and p.sk_bytes == 4 * p.n


def test_address_layout_table():
    # byte offsets from sha256_offsets.h:9-17, big-endian multi-byte fields
    pid = O.param_id("sha2-128s")  # tree height 9 > 8: keypair uses 2 bytes
    a = O.make_addr(pid, layer=5, tree=0x0102030405060708, type=3, keypair=0x1A2,
                    chain=0x11, hash=0x22)
    b = O.addr_bytes(a)
    assert b[0] == 5
    assert b[1:9] == bytes.fromhex("0102030405060708")
    assert b[9] == 3
    assert b[12] == 0x01 and b[13] == 0xA2
    assert b[17] == 0x11 and b[21] == 0x22
    a2 = O.make_addr(pid, height=7, index=0xDEADBEEF)
    b2 = O.addr_bytes(a2)
    assert b2[17] == 7 and b2[18:22] == bytes.fromhex("DEADBEEF")
    # tree height <= 8: the high keypair byte is not written (address.cu:81-86)
    a3 = O.make_addr(SHA2_F, keypair=0x1A2)
    assert O.addr_bytes(a3)[12] == 0 and O.addr_bytes(a3)[13] == 0xA2


@pytest.mark.parametrize("pid", ALL)
def test_thash_prf_closed_forms(pid):
    p = O.params(pid)
    rng = random.Random(pid)
    for inblocks in (1, 2, p.len, p.k):
        pub_seed, data = rb(rng, p.n), rb(rng, p.n * inblocks)
        a = O.make_addr(pid, layer=rng.randrange(256), tree=rng.getrandbits(64),
                        type=rng.randrange(5), keypair=rng.randrange(512),
                        chain=rng.randrange(256), hash=rng.randrange(16))
        ab = O.addr_bytes(a)
        got = O.thash(pid, data, pub_seed, a)
        if p.hash == 0:
            want = hashlib.sha256(pub_seed + bytes(64 - p.n) + ab[:22] + data).digest()[:p.n]
        else:
            want = hashlib.shake_256(pub_seed + ab + data).digest(p.n)
        assert got == want
        key = rb(rng, p.n)
        got = O.prf_addr(pid, key, a)
        if p.hash == 0:
            want = hashlib.sha256(key + ab[:22]).digest()[:p.n]
        else:
            want = hashlib.shake_256(key + ab).digest(p.n)
        assert got == want


@pytest.mark.parametrize("pid", ALL)
def test_prf_msg_and_h_msg_closed_forms(pid):
    p = O.params(pid)
    rng = random.Random(100 + pid)
    for mlen in (0, 1, 31, 32, 33, 64 - 2 * p.n, 64 - p.n - 1, 64 - p.n, 64 - p.n + 1,
                 64 - 3 * p.n, 128 - 3 * p.n, 128 - 3 * p.n + 1, 200, 4096, 65536):
        if mlen < 0:
            continue
        sk_prf, optrand, m = rb(rng, p.n), rb(rng, p.n), rb(rng, mlen)
        R = O.gen_message_random(pid, sk_prf, optrand, m)
        if p.hash == 0:
            want = hmac.new(sk_prf, optrand + m, hashlib.sha256).digest()[:p.n]
        else:
            want = hashlib.shake_256(sk_prf + optrand + m).digest(p.n)
        assert R == want, mlen
        pk = rb(rng, 2 * p.n)
        dg, tree, leaf = O.hash_message(pid, R, pk, m)
        if p.hash == 0:
            seed = hashlib.sha256(R + pk + m).digest()
            buf = b""
            c = 0
            while len(buf) < p.dgst_bytes:
                buf += hashlib.sha256(seed + struct.pack(">I", c)).digest()
                c += 1
            buf = buf[:p.dgst_bytes]
        else:
            buf = hashlib.shake_256(R + pk + m).digest(p.dgst_bytes)
        assert dg == buf[:p.fors_msg_bytes]
        tb = buf[p.fors_msg_bytes:p.fors_msg_bytes + p.tree_bytes]
        lb = buf[p.fors_msg_bytes + p.tree_bytes:p.dgst_bytes]
        assert tree == int.from_bytes(tb, "big") & ((1 << p.tree_bits) - 1)
        assert leaf == int.from_bytes(lb, "big") & ((1 << p.leaf_bits) - 1)


@pytest.mark.parametrize("pid", [O.param_id("sha2-128f"), O.param_id("sha2-192f"),
                                 O.param_id("sha2-256f")])
def test_chain_lengths_special_cases_and_checksum(pid):
    p = O.params(pid)
    # all-zero message: every digit 0; checksum len1*15, whose 12 bits (after
    # the reference's << 4 into two bytes, wots.cu:135-136) are the last 3 digits
    z = O.chain_lengths(pid, bytes(p.n))
    csum = p.len1 * 15
    assert z[:p.len1] == [0] * p.len1
    assert z[p.len1:] == [(csum >> 8) & 15, (csum >> 4) & 15, csum & 15]
    # all-ones message: digits 15, checksum 0
    f = O.chain_lengths(pid, b"\xff" * p.n)
    assert f == [15] * p.len1 + [0, 0, 0]
    rng = random.Random(7)
    for _ in range(200):
        m = rb(rng, p.n)
        L = O.chain_lengths(pid, m)
        # digits are the nibbles of m, most significant first
        assert L[:p.len1] == [x for b in m for x in (b >> 4, b & 15)]
        cs = sum(15 - x for x in L[:p.len1])
        assert (L[p.len1] << 8 | L[p.len1 + 1] << 4 | L[p.len1 + 2]) == cs


@pytest.mark.parametrize("pid", FAST)
def test_gen_chain_composition(pid):
    p = O.params(pid)
    rng = random.Random(8)
    pub_seed, x = rb(rng, p.n), rb(rng, p.n)
    a = O.make_addr(pid, layer=3, tree=77, keypair=5, chain=9)
    for s1 in (0, 1, 7, 15):
        for s2 in (0, 1, 8):
            y = O.gen_chain(pid, x, 0, s1, pub_seed, a)
            z = O.gen_chain(pid, y, s1, s2, pub_seed, a)
            assert z == O.gen_chain(pid, x, 0, s1 + s2, pub_seed, a)
    # the loop stops at hash address w (wots.cu:56: i < start+steps && i < w)
    assert O.gen_chain(pid, x, 0, 40, pub_seed, a) == O.gen_chain(pid, x, 0, 16, pub_seed, a)
    assert O.gen_chain(pid, x, 0, 16, pub_seed, a) != O.gen_chain(pid, x, 0, 15, pub_seed, a)


@pytest.mark.parametrize("pid", FAST)
def test_wots_round_trip(pid):
    p = O.params(pid)
    rng = random.Random(9)
    sk_seed, pub_seed = rb(rng, p.n), rb(rng, p.n)
    for _ in range(2):
        m = rb(rng, p.n)
        a = O.make_addr(pid, layer=2, tree=rng.getrandbits(60), keypair=3)
        pk = O.wots_gen_pk(pid, sk_seed, pub_seed, a)
        sig = O.wots_sign(pid, m, sk_seed, pub_seed, a)
        assert O.wots_pk_from_sig(pid, sig, m, pub_seed, a) == pk
        m2 = bytes([m[0] ^ 1]) + m[1:]
        assert O.wots_pk_from_sig(pid, sig, m2, pub_seed, a) != pk


def _h_closed(p, pub_seed, ab, left, right):
    if p.hash == 0:
        return hashlib.sha256(pub_seed + bytes(64 - p.n) + ab[:22] + left + right).digest()[:p.n]
    return hashlib.shake_256(pub_seed + ab + left + right).digest(p.n)


def _node_addr(pid, base_fields, height, index):
    return O.addr_bytes(O.make_addr(pid, height=height, index=index, **base_fields))


@pytest.mark.parametrize("pid", [O.param_id("sha2-128f"), O.param_id("shake256-128f"),
                                 O.param_id("sha2-256f")])
def test_xmss_treehash_vs_brute_force(pid):
    p = O.params(pid)
    rng = random.Random(10)
    sk_seed, pub_seed = rb(rng, p.n), rb(rng, p.n)
    layer, tree = 4, rng.getrandbits(40)
    base = dict(layer=layer, tree=tree, type=2)
    ta = O.make_addr(pid, **base)
    leaves = [O.wots_gen_leaf(pid, sk_seed, pub_seed, i, ta) for i in range(1 << p.tree_h)]
    # brute force: level by level with hashlib H at (height, index)
    levels = [leaves]
    for h in range(1, p.tree_h + 1):
        prev = levels[-1]
        levels.append([_h_closed(p, pub_seed, _node_addr(pid, base, h, i),
                                 prev[2 * i], prev[2 * i + 1]) for i in range(len(prev) // 2)])
    root_bf = levels[-1][0]
    for leaf_idx in range(1 << p.tree_h):
        ta = O.make_addr(pid, **base)
        root, auth = O.xmss_treehash(pid, sk_seed, pub_seed, leaf_idx, ta)
        assert root == root_bf
        want_auth = b"".join(levels[h][(leaf_idx >> h) ^ 1] for h in range(p.tree_h))
        assert auth == want_auth
        ca = O.make_addr(pid, **base)
        assert O.compute_root(pid, leaves[leaf_idx], leaf_idx, 0, auth, p.tree_h, pub_seed,
                              ca) == root_bf


def test_message_to_indices_lsb_first():
    pid = SHA2_F
    p = O.params(pid)
    # exhaustive over single set bits: bit b of the stream lands in index b // a,
    # bit position b % a, with the stream read LSB-first in each byte
    for b in range(p.a * p.k):
        m = bytearray(p.fors_msg_bytes)
        m[b >> 3] |= 1 << (b & 7)
        idx = O.message_to_indices(pid, bytes(m))
        want = [0] * p.k
        want[b // p.a] = 1 << (b % p.a)
        assert idx == want


@pytest.mark.parametrize("pid", [O.param_id("sha2-128f"), O.param_id("shake256-128f")])
def test_fors_vs_brute_force_and_round_trip(pid):
    p = O.params(pid)
    rng = random.Random(11)
    sk_seed, pub_seed = rb(rng, p.n), rb(rng, p.n)
    tree, kp = rng.getrandbits(63), rng.randrange(1 << p.tree_h)
    fa = O.make_addr(pid, tree=tree, keypair=kp, type=0)
    m = rb(rng, p.fors_msg_bytes)
    sig, pk = O.fors_sign(pid, m, sk_seed, pub_seed, fa)
    assert O.fors_pk_from_sig(pid, sig, m, pub_seed, fa) == pk
    # brute force tree 0 and the FORS pk from hashlib closed forms
    idx = O.message_to_indices(pid, m)
    base = dict(tree=tree, keypair=kp, type=3)
    roots = []
    for t in range(p.k):
        off = t << p.a
        lv = []
        for i in range(1 << p.a):
            ab = _node_addr(pid, base, 0, off + i)
            if p.hash == 0:
                sk = hashlib.sha256(sk_seed + ab[:22]).digest()[:p.n]
                lf = hashlib.sha256(pub_seed + bytes(64 - p.n) + ab[:22] + sk).digest()[:p.n]
            else:
                sk = hashlib.shake_256(sk_seed + ab).digest(p.n)
                lf = hashlib.shake_256(pub_seed + ab + sk).digest(p.n)
            if i == idx[t]:
                sk_sel = sk
            lv.append(lf)
        levels = [lv]
        for h in range(1, p.a + 1):
            prev = levels[-1]
            levels.append([_h_closed(p, pub_seed, _node_addr(pid, base, h, (off >> h) + i),
                                     prev[2 * i], prev[2 * i + 1])
                           for i in range(len(prev) // 2)])
        roots.append(levels[-1][0])
        rec = sig[t * (p.a + 1) * p.n:(t + 1) * (p.a + 1) * p.n]
        assert rec[:p.n] == sk_sel
        assert rec[p.n:] == b"".join(levels[h][(idx[t] >> h) ^ 1] for h in range(p.a))
    pk_ab = O.addr_bytes(O.make_addr(pid, tree=tree, keypair=kp, type=4))
    if p.hash == 0:
        want = hashlib.sha256(pub_seed + bytes(64 - p.n) + pk_ab[:22] + b"".join(roots)).digest()
    else:
        want = hashlib.shake_256(pub_seed + pk_ab + b"".join(roots)).digest(p.n)
    assert pk == want[:p.n]


@pytest.mark.parametrize("pid", FAST)
def test_sign_verify_round_trip_and_negative(pid):
    p = O.params(pid)
    O.set_debug_mode(False)
    O.randombytes_init(bytes(range(48)))
    pk, sk = O.keypair(pid)
    assert pk == sk[2 * p.n:]
    m = O.randombytes(33)
    sm = O.sign(pid, m, sk)
    assert len(sm) == p.sig_bytes + 33 and sm[p.sig_bytes:] == m
    r, mo, _ = O.open_(pid, sm, pk)
    assert r == 0 and mo == m
    bad = bytearray(sm)
    bad[-1] ^= 1  # test/spx.cu:159-166
    r, mo, full = O.open_(pid, bytes(bad), pk)
    assert r == -1 and mo == b"" and full == bytes(len(sm))
    bad = bytearray(sm)
    bad[p.n + 5] ^= 0x10  # a FORS secret-key byte
    assert O.open_(pid, bytes(bad), pk)[0] == -1
    r, _, full = O.open_(pid, sm[:p.sig_bytes - 1], pk)  # smlen < SPX_BYTES
    assert r == -1 and full == bytes(p.sig_bytes - 1)
    O.set_debug_mode(True)


def test_debug_mode_semantics():
    pid = SHA2_F
    p = O.params(pid)
    O.set_debug_mode(True)
    O.randombytes_init(bytes(range(48)))
    pk, sk = O.keypair(pid)
    assert sk[:3 * p.n] == bytes(range(3 * p.n))  # sign.cu:345-349
    # optrand = 0xad * n (sign.cu:1950-1954), after n DRBG bytes were drawn
    m = b"hello"
    sig = O.sign(pid, m, sk)[:p.sig_bytes]
    assert sig == O.signature_with_optrand(pid, m, sk, b"\xad" * p.n)
    O.set_debug_mode(False)
    O.randombytes_init(bytes(range(48)))
    seed = O.randombytes(3 * p.n)
    opt = O.randombytes(p.n)
    O.randombytes_init(bytes(range(48)))
    pk2, sk2 = O.keypair(pid)
    assert sk2[:3 * p.n] == seed
    sig2 = O.sign(pid, m, sk2)[:p.sig_bytes]
    assert sig2 == O.signature_with_optrand(pid, m, sk2, opt)
    O.set_debug_mode(True)


def test_sign_work_counts_sha2_128f():
    # exact per-signature compression count has the expected magnitude
    # (22 layers x 8 leaves x 35 chains x 16 hashes dominate)
    pid = SHA2_F
    p = O.params(pid)
    rng = random.Random(12)
    sk = rb(rng, 3 * p.n)
    pk, sk = O.seed_keypair(pid, sk)
    O.reset_counts()
    O.signature_with_optrand(pid, rb(rng, 32), sk, rb(rng, p.n))
    c = O.counts()[0]
    wots_pk = p.d * (1 << p.tree_h) * p.len * 16
    assert wots_pk < c < wots_pk * 1.25
