"""Independent closed-form pins of the oracle's WOTS+ layer and of the whole
keypair / signature composition.

A second restatement of SPHINCS+ round-3 "simple" written here from the
specification with Python's hashlib (no code shared with
oracle/spx_oracle.c), following the reference's choices:

* 32-byte address in the SHA-256 layout for both hash families
  (sha256_offsets.h:9-17): layer @0, tree @1..8, type @9, keypair @10..13,
  chain / tree height @14..17, hash / tree index @18..21; the SHA-256
  functions absorb its first 22 bytes, SHAKE256 all 32
  (thash_shake256.cu:20-24, hash_shake256.cu:44);
* thash = SHA-256(pub_seed || 0^(64-n) || addr22 || in)[:n] or
  SHAKE256(pub_seed || addr32 || in, n); PRF without pub_seed;
  PRF_msg = HMAC-SHA-256 / SHAKE256, H_msg = MGF1(SHA-256(R || pk || m))
  / SHAKE256 (hash_sha256.cu:65-262, hash_shake256.cu:38-130);
* wots_gen_leaf (sign.cu:87-104, wots.cu:27-186): sk_j = PRF(WOTS_HASH
  address of chain j with hash address 0), 15 chain steps at hash
  addresses 0..14, T_len under the WOTS_PK address of the keypair;
* WOTS signature (wots.cu:614-626): chain j of the message digit d_j
  (base-16 digits, MSB first, then the 3-digit checksum << 4);
* FORS indices LSB first (fors.cu:77-91); node addresses (height, index +
  offset >> height) (utils.cu:189-241);
* keypair (sign.cu:154-182) and signature (sign.cu:1924-1994): layer i
  signs the previous root with leaf idx_leaf of tree idx_tree, then
  idx_leaf = idx_tree mod 2^h', idx_tree >>= h'.

A slip in the oracle that sign and verify would share (a wrong address type
for the WOTS PRF or T_len, the hash address set after the PRF, the layer
wiring) changes these bytes, which the round-trip tests cannot see.
"""
import hashlib
import hmac
import random
import struct

import pytest

from oracle import oracle as O

WOTS, WOTS_PK, TREE, FORS_TREE, FORS_PK = 0, 1, 2, 3, 4


class Spx:
    def __init__(self, name):
        self.name = name
        self.shake = name.startswith("shake")
        size = name.split("-")[1]
        self.n, self.h, self.d, self.a, self.k = {
            "128s": (16, 63, 7, 12, 14), "128f": (16, 66, 22, 6, 33),
            "192s": (24, 63, 7, 14, 17), "192f": (24, 66, 22, 8, 33),
            "256s": (32, 64, 8, 14, 22), "256f": (32, 68, 17, 9, 35)}[size]
        self.hp = self.h // self.d
        self.len1 = 2 * self.n
        self.len = self.len1 + 3
        self.md_bytes = (self.k * self.a + 7) // 8
        self.tree_bits = self.h - self.hp
        self.tree_bytes = (self.tree_bits + 7) // 8
        self.leaf_bytes = (self.hp + 7) // 8
        self.dgst = self.md_bytes + self.tree_bytes + self.leaf_bytes

    # ------------------------------------------------------------ address
    @staticmethod
    def adrs(layer=0, tree=0, typ=0, kp=0, w1=0, w2=0):
        """w1: chain address or tree height; w2: hash address or tree index"""
        return (bytes([layer]) + tree.to_bytes(8, "big") + bytes([typ]) + kp.to_bytes(4, "big") +
                w1.to_bytes(4, "big") + w2.to_bytes(4, "big") + bytes(10))

    # ------------------------------------------------------------ hashes
    def thash(self, pub, adrs, data):
        if self.shake:
            return hashlib.shake_256(pub + adrs + data).digest(self.n)
        return hashlib.sha256(pub + bytes(64 - self.n) + adrs[:22] + data).digest()[:self.n]

    def prf(self, key, adrs):
        if self.shake:
            return hashlib.shake_256(key + adrs).digest(self.n)
        return hashlib.sha256(key + adrs[:22]).digest()[:self.n]

    def prf_msg(self, sk_prf, opt, m):
        if self.shake:
            return hashlib.shake_256(sk_prf + opt + m).digest(self.n)
        return hmac.new(sk_prf, opt + m, hashlib.sha256).digest()[:self.n]

    def h_msg(self, R, pk, m):
        if self.shake:
            buf = hashlib.shake_256(R + pk + m).digest(self.dgst)
        else:
            seed = hashlib.sha256(R + pk + m).digest()
            buf = b"".join(hashlib.sha256(seed + struct.pack(">I", c)).digest()
                           for c in range((self.dgst + 31) // 32))[:self.dgst]
        md = buf[:self.md_bytes]
        t = int.from_bytes(buf[self.md_bytes:self.md_bytes + self.tree_bytes], "big")
        lf = int.from_bytes(buf[self.md_bytes + self.tree_bytes:self.dgst], "big")
        return md, t & ((1 << self.tree_bits) - 1), lf & ((1 << self.hp) - 1)

    # ------------------------------------------------------------ WOTS+
    def digits(self, msg):
        d = []
        for b in msg:
            d += [b >> 4, b & 15]
        csum = sum(15 - x for x in d) << 4  # 12 significant bits in 2 bytes
        return d + [(csum >> 12) & 15, (csum >> 8) & 15, (csum >> 4) & 15]

    def chain(self, x, start, steps, pub, layer, tree, kp, j):
        for s in range(start, start + steps):
            x = self.thash(pub, self.adrs(layer, tree, WOTS, kp, j, s), x)
        return x

    def wots_sk(self, seed, layer, tree, kp, j):
        return self.prf(seed, self.adrs(layer, tree, WOTS, kp, j, 0))

    def wots_leaf(self, seed, pub, layer, tree, kp):
        ends = [self.chain(self.wots_sk(seed, layer, tree, kp, j), 0, 15, pub, layer, tree, kp, j)
                for j in range(self.len)]
        return self.thash(pub, self.adrs(layer, tree, WOTS_PK, kp), b"".join(ends))

    def wots_sign(self, msg, seed, pub, layer, tree, kp):
        return b"".join(self.chain(self.wots_sk(seed, layer, tree, kp, j), 0, dj, pub, layer,
                                   tree, kp, j)
                        for j, dj in enumerate(self.digits(msg)))

    # ------------------------------------------------------------ Merkle
    def merkle(self, leaves, pub, layer, tree, typ, kp, offset, target):
        """root and the auth path of leaf `target` over `leaves`; node x at
        height h gets the address (h, x + (offset >> h))"""
        auth, level, h = [], leaves, 0
        while len(level) > 1:
            auth.append(level[(target >> h) ^ 1])
            h += 1
            level = [self.thash(pub, self.adrs(layer, tree, typ, kp, h, i + (offset >> h)),
                                level[2 * i] + level[2 * i + 1]) for i in range(len(level) // 2)]
        return level[0], b"".join(auth)

    def xmss(self, seed, pub, layer, tree, leaf):
        leaves = [self.wots_leaf(seed, pub, layer, tree, i) for i in range(1 << self.hp)]
        return self.merkle(leaves, pub, layer, tree, TREE, 0, 0, leaf)

    # ------------------------------------------------------------ FORS
    def fors_sign(self, md, seed, pub, tree, kp):
        bits = int.from_bytes(md, "little")
        sig, roots = b"", []
        for t in range(self.k):
            idx = (bits >> (t * self.a)) & ((1 << self.a) - 1)
            off = t << self.a
            sks = [self.prf(seed, self.adrs(0, tree, FORS_TREE, kp, 0, off + i))
                   for i in range(1 << self.a)]
            leaves = [self.thash(pub, self.adrs(0, tree, FORS_TREE, kp, 0, off + i), sks[i])
                      for i in range(1 << self.a)]
            root, auth = self.merkle(leaves, pub, 0, tree, FORS_TREE, kp, off, idx)
            sig += sks[idx] + auth
            roots.append(root)
        return sig, self.thash(pub, self.adrs(0, tree, FORS_PK, kp), b"".join(roots))

    # ------------------------------------------------------------ SPHINCS+
    def keypair(self, seed3n):
        sk_seed, pub = seed3n[:self.n], seed3n[2 * self.n:]
        root, _ = self.xmss(sk_seed, pub, self.d - 1, 0, 0)
        return pub + root, seed3n + root

    def sign(self, m, sk, opt):
        n = self.n
        sk_seed, sk_prf, pub, root = sk[:n], sk[n:2 * n], sk[2 * n:3 * n], sk[3 * n:]
        R = self.prf_msg(sk_prf, opt, m)
        md, tree, leaf = self.h_msg(R, pub + root, m)
        fsig, node = self.fors_sign(md, sk_seed, pub, tree, leaf)
        sig = R + fsig
        for layer in range(self.d):
            sig += self.wots_sign(node, sk_seed, pub, layer, tree, leaf)
            node, auth = self.xmss(sk_seed, pub, layer, tree, leaf)
            sig += auth
            leaf, tree = tree & ((1 << self.hp) - 1), tree >> self.hp
        return sig


def _rng_bytes(rng, n):
    return bytes(rng.randrange(256) for _ in range(n))


@pytest.mark.parametrize("name", ["sha2-128f", "shake256-128f", "sha2-192f", "sha2-128s",
                                  "shake256-256f"])
def test_wots_gen_leaf_closed_form(name):
    S, pid = Spx(name), O.param_id(name)
    rng = random.Random(41)
    seed, pub = _rng_bytes(rng, S.n), _rng_bytes(rng, S.n)
    for layer, tree, kp in ((0, 0, 0), (3, rng.getrandbits(40), 5), (S.d - 1, 1, (1 << S.hp) - 1)):
        ta = O.make_addr(pid, layer=layer, tree=tree, type=TREE)
        assert O.wots_gen_leaf(pid, seed, pub, kp, ta) == S.wots_leaf(seed, pub, layer, tree, kp)


@pytest.mark.parametrize("name", ["sha2-128f", "shake256-128f", "sha2-256f"])
def test_wots_sign_closed_form(name):
    """The WOTS signature of chosen messages (all-zero digits: every chain at
    its secret; all-ones: 15 steps; random) and of the checksum digits."""
    S, pid = Spx(name), O.param_id(name)
    rng = random.Random(42)
    seed, pub = _rng_bytes(rng, S.n), _rng_bytes(rng, S.n)
    layer, tree, kp = 2, rng.getrandbits(50), 3
    for msg in (bytes(S.n), b"\xff" * S.n, _rng_bytes(rng, S.n), _rng_bytes(rng, S.n)):
        a = O.make_addr(pid, layer=layer, tree=tree, type=WOTS, keypair=kp)
        got = O.wots_sign(pid, msg, seed, pub, a)
        assert got == S.wots_sign(msg, seed, pub, layer, tree, kp)
        assert O.chain_lengths(pid, msg) == S.digits(msg)
    # the all-zero message signs with the secrets themselves
    a = O.make_addr(pid, layer=layer, tree=tree, type=WOTS, keypair=kp)
    sig0 = O.wots_sign(pid, bytes(S.n), seed, pub, a)
    assert sig0[:S.n] == S.wots_sk(seed, layer, tree, kp, 0)


@pytest.mark.parametrize("name", ["sha2-128f", "shake256-128f"])
def test_keypair_and_signature_closed_form(name):
    S, pid = Spx(name), O.param_id(name)
    rng = random.Random(43)
    seed = _rng_bytes(rng, 3 * S.n)
    pk, sk = O.seed_keypair(pid, seed)
    assert (pk, sk) == S.keypair(seed)
    for mlen in (0, 33, 200):
        m, opt = _rng_bytes(rng, mlen), _rng_bytes(rng, S.n)
        sig = O.signature_with_optrand(pid, m, sk, opt)
        want = S.sign(m, sk, opt)
        assert len(want) == O.params(pid).sig_bytes
        assert sig == want, mlen


@pytest.mark.slow
def test_signature_closed_form_other_sets():
    """The composition with 2-byte keypair fields (h' = 9 > 8) and n = 24 / 32."""
    for name in ("sha2-192f", "shake256-256f"):
        S, pid = Spx(name), O.param_id(name)
        rng = random.Random(44)
        seed = _rng_bytes(rng, 3 * S.n)
        pk, sk = O.seed_keypair(pid, seed)
        assert (pk, sk) == S.keypair(seed)
        m, opt = _rng_bytes(rng, 77), _rng_bytes(rng, S.n)
        assert O.signature_with_optrand(pid, m, sk, opt) == S.sign(m, sk, opt), name


def test_keypair_closed_form_two_byte_keypair_field():
    """SHA2-128s: 2^9 leaves per subtree, so keypair addresses need two bytes."""
    S, pid = Spx("sha2-128s"), O.param_id("sha2-128s")
    rng = random.Random(45)
    seed, pub = _rng_bytes(rng, S.n), _rng_bytes(rng, S.n)
    ta = O.make_addr(pid, layer=6, tree=0, type=TREE)
    assert O.wots_gen_leaf(pid, seed, pub, 300, ta) == S.wots_leaf(seed, pub, 6, 0, 300)
