"""Synthetic inputs of the benchmark configurations.

One definition of each workload, used by ``bench.py`` (timing), the GPU
tests (bit-exact checks at the full sizes) and the fixture script
``tests/golden/make_configs.py`` (expected outputs).  Every generator is
deterministic: keys and optrand come from the NIST AES-256-CTR DRBG
(``drbg`` is any object with ``randombytes_init(entropy)`` and
``randombytes(n)``: the library's host DRBG or the oracle's, which draw the
same stream), sizes and message bytes from seeded ``random.Random`` /
SHAKE256 streams.

* ``kat_batch`` -- SHA2-128f batched signing: B keypairs by the NIST KAT
  protocol in DRBG mode (PQCgenKAT_sign.cu:55-63: master entropy 00..2f,
  count i draws a 48-byte seed and a 32-byte message; DRBG(seed_i) then
  yields the 3n-byte key seed, sign.cu:345, and the n-byte optrand,
  sign.cu:1950).
* ``mixed_shake_batch`` -- SHAKE256-128f / -192f in one batch, messages
  log-uniform in 1 B .. 4 KiB.
* ``sha2_128s_batches`` -- SHA2-128s batches of 1 / 64 / 1024 requests with
  32-byte messages.
* ``serving_keys`` / ``serving_requests`` -- 256 resident keys over
  SHA2-128f / SHA2-128s / SHAKE256-256f; 80% sign / 20% verify, messages
  log-uniform in 16 B .. 64 KiB.
"""
import hashlib
import math
import random

SERVE_SETS = ("sha2-128f", "sha2-128s", "shake256-256f")
SERVE_KEYS = (86, 86, 84)
SET_N = {"sha2-128f": 16, "sha2-128s": 16, "sha2-192f": 24, "sha2-256f": 32,
         "shake256-128f": 16, "shake256-192f": 24, "shake256-256f": 32}


def log_uniform(rng, lo, hi):
    """Integer length log-uniform in [lo, hi]."""
    return int(round(math.exp(rng.uniform(math.log(lo), math.log(hi)))))


def kat_batch(drbg, n, count):
    """(key seeds [3n each], messages [32 B each], optrand [n each]) of KAT
    counts 0 .. count-1."""
    drbg.randombytes_init(bytes(range(48)))
    seeds48, msgs = [], []
    for _ in range(count):
        seeds48.append(drbg.randombytes(48))
        msgs.append(drbg.randombytes(32))
    kseeds, opts = [], []
    for s in seeds48:
        drbg.randombytes_init(s)
        kseeds.append(drbg.randombytes(3 * n))
        opts.append(drbg.randombytes(n))
    return kseeds, msgs, opts


def mixed_shake_batch(drbg, count, seed=3):
    """[(set name, key seed, message, optrand)] for count requests, half
    SHAKE256-128f and half SHAKE256-192f, each with its own key, in a shuffled
    order."""
    rng = random.Random(seed)
    drbg.randombytes_init(bytes([seed]) * 48)
    items = []
    for name in ("shake256-128f", "shake256-192f"):
        n = SET_N[name]
        cnt = count // 2
        seeds = drbg.randombytes(3 * n * cnt)
        for i in range(cnt):
            m = rng.randbytes(log_uniform(rng, 1, 4096))
            items.append((name, seeds[3 * n * i:3 * n * (i + 1)], m, drbg.randombytes(n)))
    rng.shuffle(items)
    return items


def sha2_128s_batches(sizes=(1, 64, 1024), seed=4):
    """{B: (key seeds, messages, optrand)} of SHA2-128s batches."""
    rng = random.Random(seed)
    out = {}
    for B in sizes:
        seeds = [rng.randbytes(48) for _ in range(B)]
        msgs = [rng.randbytes(32) for _ in range(B)]
        opts = [rng.randbytes(16) for _ in range(B)]
        out[B] = (seeds, msgs, opts)
    return out


def serving_keys(drbg, seed=5):
    """[(set name, key seed)] of the 256 resident serving keys."""
    drbg.randombytes_init(bytes([seed]) * 48)
    keys = []
    for name, cnt in zip(SERVE_SETS, SERVE_KEYS):
        n = SET_N[name]
        blob = drbg.randombytes(3 * n * cnt)
        keys += [(name, blob[3 * n * i:3 * n * (i + 1)]) for i in range(cnt)]
    return keys


class MessagePool:
    """Messages log-uniform in 16 B .. 64 KiB, cut from a 1 MiB SHAKE256
    stream at random offsets."""

    def __init__(self, seed):
        self.pool = hashlib.shake_256(b"serving message pool %d" % seed).digest(1 << 20)

    def draw(self, rng, lo=16, hi=65536):
        n = log_uniform(rng, lo, hi)
        o = rng.randrange(0, len(self.pool) - n)
        return self.pool[o:o + n]


def serving_requests(count, seed=6, nkeys=sum(SERVE_KEYS), verify_frac=0.2, tamper_frac=0.25):
    """The mixed request stream of the serving workload.

    Returns (sources, requests).  ``sources`` are (key index, message,
    optrand) triples signed before the stream starts; verify requests check
    them.  ``requests`` are ("sign", key, message, optrand) or ("verify",
    source index, tamper) with tamper None (valid), ("msg", byte, bit) or
    ("sig", byte, bit).  Sizes of the sign / verify parts follow from
    verify_frac."""
    rng = random.Random(seed)
    pool = MessagePool(seed)
    n_ver = int(round(count * verify_frac))
    sources = []
    for _ in range(n_ver):
        k = rng.randrange(nkeys)
        sources.append((k, pool.draw(rng), rng.randbytes(32)))
    ops = ["sign"] * (count - n_ver) + ["verify"] * n_ver
    rng.shuffle(ops)
    reqs, vi = [], 0
    for op in ops:
        if op == "sign":
            reqs.append(("sign", rng.randrange(nkeys), pool.draw(rng), rng.randbytes(32)))
        else:
            t = None
            if rng.random() < tamper_frac:
                part = rng.choice(("msg", "sig"))
                t = (part, rng.randrange(1 << 30), rng.randrange(8))
            reqs.append(("verify", vi, t))
            vi += 1
    return sources, reqs


def apply_tamper(tamper, msg, sig):
    """Flip one bit of the message or the signature (a copy); byte index
    reduced modulo the length."""
    if tamper is None:
        return msg, sig
    part, byte, bit = tamper
    if part == "msg":
        b = bytearray(msg)
        b[byte % len(b)] ^= 1 << bit
        return bytes(b), sig
    b = bytearray(sig)
    b[byte % len(b)] ^= 1 << bit
    return msg, bytes(b)


def digest(b, nbytes=8):
    """Short SHA-256 digest of one output, as stored in the fixtures."""
    return hashlib.sha256(bytes(b)).digest()[:nbytes]
