"""Pins the oracle's primitives against library routines and published vectors.

Checks are independent of the oracle's own code: hashlib/hmac (OpenSSL),
FIPS 180-4 / FIPS 202 / FIPS-197 examples and the NIST PQC KAT DRBG output
(count 0 of PQCgenKAT_sign.cu:55-63, entropy 00..2f).
"""
import hashlib
import hmac
import random
import struct

import pytest

from oracle import oracle as O


def test_sha256_fips_examples():
    assert O.sha256(b"abc").hex() == (
        "ba7816bf8f01cfea414140de5dae2223b00361a396177a9cb410ff61f20015ad")
    assert O.sha256(b"abcdbcdecdefdefgefghfghighijhijkijkljklmklmnlmnomnopnopq").hex() == (
        "248d6a61d20638b8e5c026930c3e6039a33ce45964ff2167f6ecedd419db06c1")
    assert O.sha256(b"").hex() == (
        "e3b0c44298fc1c149afbf4c8996fb92427ae41e4649b934ca495991b7852b855")


def test_sha256_random_lengths_vs_hashlib():
    rng = random.Random(1)
    lens = list(range(0, 200)) + [rng.randrange(0, 200 * 1024) for _ in range(20)]
    for n in lens:
        data = rng.randbytes(n)
        assert O.sha256(data) == hashlib.sha256(data).digest(), n


def test_sha256_block_count():
    O.reset_counts()
    O.sha256(b"x" * 55)
    assert O.counts()[0] == 1
    O.reset_counts()
    O.sha256(b"x" * 56)
    assert O.counts()[0] == 2


def test_sha256_compress_words_is_one_block_of_hashlib():
    # one block "abc" padded by hand; state from the FIPS 180-4 IV
    iv = [0x6a09e667, 0xbb67ae85, 0x3c6ef372, 0xa54ff53a,
          0x510e527f, 0x9b05688c, 0x1f83d9ab, 0x5be0cd19]
    block = b"abc" + b"\x80" + b"\x00" * 52 + struct.pack(">Q", 24)
    words = list(struct.unpack(">16I", block))
    out = O.sha256_compress_words(iv, words)
    assert struct.pack(">8I", *out) == hashlib.sha256(b"abc").digest()


def test_hmac_vs_library():
    rng = random.Random(2)
    for n in (16, 24, 32):
        for mlen in (0, 1, 31, 32, 47, 48, 63, 64, 65, 200, 5000):
            key, msg = rng.randbytes(n), rng.randbytes(mlen)
            assert O.hmac_sha256(key, msg) == hmac.new(key, msg, hashlib.sha256).digest()


def _mgf1_ref(seed, outlen):
    out = b""
    c = 0
    while len(out) < outlen:
        out += hashlib.sha256(seed + struct.pack(">I", c)).digest()
        c += 1
    return out[:outlen]


def test_mgf1_vs_closed_form():
    rng = random.Random(3)
    for outlen in (1, 30, 32, 34, 42, 49, 64, 65, 100):
        seed = rng.randbytes(32)
        assert O.mgf1(seed, outlen) == _mgf1_ref(seed, outlen)


def test_shake256_vs_hashlib():
    rng = random.Random(4)
    for n in list(range(0, 300)) + [1000, 4096, 65536]:
        data = rng.randbytes(n)
        for outlen in (16, 24, 32, 49, 136, 300):
            assert O.shake256(data, outlen) == hashlib.shake_256(data).digest(outlen)


def test_keccak_zero_state_first_lane():
    # Keccak-f[1600] applied to the all-zero state (published test value)
    out = O.keccakf1600([0] * 25)
    assert out[0] == 0xF1258F7940E1DDE7
    assert out[1] == 0x84D5CCF933C0478A


def test_aes256_fips197_c3():
    key = bytes(range(32))
    pt = bytes.fromhex("00112233445566778899aabbccddeeff")
    assert O.aes256_ecb(key, pt).hex() == "8ea2b7ca516745bfeafc49904b496089"


def test_drbg_nist_kat_count0():
    # PQCgenKAT_sign.cu:55-63: entropy 00..2f, then seed = randombytes(48),
    # msg = randombytes(33) for count 0
    O.randombytes_init(bytes(range(48)))
    seed = O.randombytes(48)
    msg = O.randombytes(33)
    assert seed.hex().upper() == (
        "061550234D158C5EC95595FE04EF7A25767F2E24CC2BC479D09D86DC9ABCFDE7"
        "056A8C266F9EF97ED08541DBD2E1FFA1")
    assert msg.hex().upper() == (
        "D81C4D8D734FCBFBEADE3D3F8A039FAA2A2C9957E835AD55B22E75BF57BB556AC8")


def test_drbg_matches_aes_ctr_closed_form():
    # After init with entropy E: (K, V) = Update(E, 0, 0); each randombytes
    # call returns AES_K(V+1), AES_K(V+2), ... then runs Update(NULL).
    ent = bytes(range(48))

    def inc(v):
        return ((int.from_bytes(v, "big") + 1) % (1 << 128)).to_bytes(16, "big")

    def update(data, K, V):
        t = b""
        for _ in range(3):
            V = inc(V)
            t += O.aes256_ecb(K, V)
        if data is not None:
            t = bytes(a ^ b for a, b in zip(t, data))
        return t[:32], t[32:]

    K, V = update(ent, bytes(32), bytes(16))
    V1 = inc(V)
    expect = O.aes256_ecb(K, V1)[:10]
    O.randombytes_init(ent)
    assert O.randombytes(10) == expect
