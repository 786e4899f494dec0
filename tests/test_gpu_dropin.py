"""The drop-in boundary on the GPU.

* The C++ callers of tests/dropin (shaped like the reference's test/spx.cu
  and test/dp_spx.cu, built with g++ against include/cuspx_api.h only and
  linked to libcuspx.so) run and pass their own checks; their keypairs and
  signed messages equal the CPU oracle's for the same DRBG stream.  So do
  the same single-item callers built against the plain declarations and
  linked to the per-set library libspx_<set>.so alone.
* The batched api.h calls in DRBG mode (the reference draws curand there,
  sign.cu:1208 / rng.cu:196; this library draws the host DRBG item by item)
  equal the oracle's keypair / signature sequence, and the multi-stream
  face_ms_mdp_* calls equal face_mdp_* byte for byte.
"""
import os
import subprocess

import numpy as np
import pytest

import cuspx_amd as C
from oracle import oracle as O

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "tests", "dropin", "bin")


def _run(name, *args):
    exe = os.path.join(BIN, name)
    assert os.path.exists(exe), "build the drop-in callers first (make -C tests/dropin)"
    print("running %s %s" % (name, " ".join(args)), flush=True)
    p = subprocess.run([exe] + list(args), capture_output=True, text=True, timeout=300)
    ok = p.stdout.startswith("OK ") or "\nOK " in p.stdout
    assert p.returncode == 0 and ok, p.stdout[-2000:] + p.stderr[-2000:]
    return p.stdout


@pytest.mark.parametrize("name,prog", [("sha2-128f", "spx_sha2_128f"),
                                       ("shake256-128f", "spx_shake256_128f"),
                                       ("sha2-128f", "spx_sha2_128f_perset"),
                                       ("shake256-128f", "spx_shake256_128f_perset")])
def test_dropin_single_item_caller(name, prog):
    """The *_perset callers declare the api.h names without cuspx_api.h and
    link libspx_<set>.so alone: the per-set library picks the set."""
    out = {}
    for line in _run(prog).splitlines():
        tag, _, val = line.partition(" ")
        if tag in ("pk", "sk", "m", "sm"):
            out[tag] = val
    pid = O.param_id(name)
    O.set_debug_mode(True)
    O.randombytes_init(bytes(range(48)))
    m = O.randombytes(32)
    pk, sk = O.keypair(pid)
    sm = O.sign(pid, m, sk)
    assert bytes.fromhex(out["m"]) == m
    assert bytes.fromhex(out["pk"]) == pk and bytes.fromhex(out["sk"]) == sk
    assert bytes.fromhex(out["sm"]) == sm


@pytest.mark.parametrize("name,prog,num", [("sha2-128f", "dp_sha2_128f", 4096),
                                           ("shake256-128f", "dp_shake256_128f", 1100),
                                           ("shake256-128f", "dp_shake256_128f_perset", 300)])
def test_dropin_batched_caller(name, prog, num, tmp_path):
    path = str(tmp_path / "dp.bin")
    _run(prog, str(num), path)
    pid = O.param_id(name)
    p = O.params(pid)
    raw = open(path, "rb").read()
    sm_bytes = p.sig_bytes + 32
    assert len(raw) == num * (32 + p.pk_bytes + p.sk_bytes + sm_bytes)
    m = raw[:32 * num]
    pk = raw[32 * num:32 * num + p.pk_bytes * num]
    sk = raw[32 * num + p.pk_bytes * num:32 * num + (p.pk_bytes + p.sk_bytes) * num]
    sm = raw[32 * num + (p.pk_bytes + p.sk_bytes) * num:]
    O.set_debug_mode(True)
    O.randombytes_init(bytes(range(48)))
    assert m == O.randombytes(32 * num)
    # DEBUG_MODE: every key from seed 0..3n-1, optrand 0xad (sign.cu:345-349, 1950-1954)
    opk, osk = O.seed_keypair(pid, bytes(range(3 * p.n)))
    assert pk == opk * num and sk == osk * num
    for i in sorted({0, 1, num // 2, num - 1}):
        rec = sm[i * sm_bytes:(i + 1) * sm_bytes]
        mi = m[32 * i:32 * (i + 1)]
        assert rec[:p.sig_bytes] == O.signature_with_optrand(pid, mi, osk, b"\xad" * p.n), i
        assert rec[p.sig_bytes:] == mi


def test_face_mdp_drbg_mode_matches_oracle_sequence():
    """Non-debug batched keygen and signing draw 3n / n DRBG bytes per item,
    in item order, like consecutive crypto_sign_keypair / crypto_sign calls
    (sign.cu:342-352, 1950-1954)."""
    pid = C.param_id("sha2-128f")
    p = O.params(pid)
    num = 24
    ent = bytes((5 * i + 1) & 0xFF for i in range(48))
    C.set_default_param(pid)
    C.set_debug_mode(False)
    O.set_debug_mode(False)
    try:
        C.randombytes_init(ent)
        pk, sk = C.face_mdp_crypto_sign_keypair(num)
        m = np.frombuffer(C.randombytes(32 * num), np.uint8)
        sm, smlen = C.face_mdp_crypto_sign(m, 32, sk, num)
        O.randombytes_init(ent)
        okeys = [O.keypair(pid) for _ in range(num)]
        om = O.randombytes(32 * num)
        assert bytes(m) == om
        opts = [O.randombytes(p.n) for _ in range(num)]
        assert len({bytes(sk[i]) for i in range(num)}) == num  # distinct keys
        for i in range(num):
            assert (bytes(pk[i]), bytes(sk[i])) == okeys[i], i
            want = O.signature_with_optrand(pid, om[32 * i:32 * i + 32], okeys[i][1], opts[i])
            assert bytes(sm[i, :p.sig_bytes]) == want, i
        rc, mo, mlen = C.face_mdp_crypto_sign_open(sm, smlen, pk, num)
        assert rc == 0 and mlen == 32 and bytes(mo.reshape(-1)) == om
    finally:
        C.set_debug_mode(True)
        O.set_debug_mode(True)


def test_face_ms_mdp_equals_face_mdp_drbg_mode():
    """The multi-stream forms (4 slices at 4096 items for verify, 2 for keygen
    and signing) give the same bytes as the single-stream ones for the same
    DRBG stream, and report a tampered item the same way."""
    pid = C.param_id("sha2-128f")
    C.set_default_param(pid)
    C.set_debug_mode(False)
    num = 4096
    p = C.params(pid)
    ent = bytes(range(1, 49))
    try:
        outs = []
        for ms in (False, True):
            print("face_%smdp_*: %d items" % ("ms_" if ms else "", num), flush=True)
            C.randombytes_init(ent)
            pk, sk = C.face_mdp_crypto_sign_keypair(num, ms=ms)
            m = np.frombuffer(C.randombytes(32 * num), np.uint8).copy()
            sm, smlen = C.face_mdp_crypto_sign(m, 32, sk, num, ms=ms)
            assert smlen == p.sig_bytes + 32
            rc, mo, mlen = C.face_mdp_crypto_sign_open(sm, smlen, pk, num, ms=ms)
            assert rc == 0 and mlen == 32 and bytes(mo.reshape(-1)) == bytes(m)
            bad = sm.copy()
            bad[[7, 3000], p.sig_bytes - 1] ^= 1  # last auth-path byte of two records
            rc, mo, _ = C.face_mdp_crypto_sign_open(bad, smlen, pk, num, ms=ms)
            assert rc == -1
            assert [i for i in range(num) if not mo[i].any()] == [7, 3000]
            outs.append((pk.tobytes(), sk.tobytes(), sm.tobytes()))
        assert outs[0] == outs[1]
    finally:
        C.set_debug_mode(True)
