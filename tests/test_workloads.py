"""CPU checks of the benchmark workloads (cuspx_amd/workloads.py) and of the
full-size fixtures in tests/golden/configs.json.

* The generators draw the same inputs from the library's host DRBG and from
  the oracle's, so the fixtures (written with the oracle's) and the GPU
  tests / bench.py (the library's) run the same workloads.
* The CPU oracle still reproduces sampled items of every fixture (first and
  last of a batch), so the committed digests are the oracle's outputs.
* Fixture invariants: every batch has one digest per item; the serving
  stream's verify statuses are -1 exactly for the tampered requests.
"""
import json
import os

import pytest

import cuspx_amd as C
from cuspx_amd import workloads as W
from oracle import oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))
G = json.load(open(os.path.join(HERE, "golden", "configs.json")))


class LibDrbg:
    randombytes_init = staticmethod(C.randombytes_init)
    randombytes = staticmethod(C.randombytes)


class OracleDrbg:
    randombytes_init = staticmethod(O.randombytes_init)
    randombytes = staticmethod(O.randombytes)


def _dg(hexstr, i):
    return bytes.fromhex(hexstr[16 * i:16 * (i + 1)])


def test_generators_agree_between_the_two_drbgs():
    assert W.kat_batch(LibDrbg, 16, 64) == W.kat_batch(OracleDrbg, 16, 64)
    assert W.mixed_shake_batch(LibDrbg, 64) == W.mixed_shake_batch(OracleDrbg, 64)
    assert W.serving_keys(LibDrbg) == W.serving_keys(OracleDrbg)


def test_generators_are_deterministic_and_shaped():
    k1 = W.kat_batch(OracleDrbg, 16, 8)
    assert k1 == W.kat_batch(OracleDrbg, 16, 8)
    kseeds, msgs, opts = k1
    assert [len(x) for x in kseeds] == [48] * 8 and [len(x) for x in msgs] == [32] * 8
    # KAT protocol: count 0's message is the second DRBG draw of entropy 00..2f
    O.randombytes_init(bytes(range(48)))
    O.randombytes(48)
    assert msgs[0] == O.randombytes(32)
    items = W.mixed_shake_batch(OracleDrbg, 2000)
    assert sum(1 for it in items if it[0] == "shake256-128f") == 1000
    lens = [len(it[2]) for it in items]
    assert min(lens) >= 1 and max(lens) <= 4096 and 30 < sorted(lens)[1000] < 130
    b = W.sha2_128s_batches()
    assert sorted(b) == [1, 64, 1024] and len(b[1024][0]) == 1024
    keys = W.serving_keys(OracleDrbg)
    assert len(keys) == 256 and [k[0] for k in keys].count("shake256-256f") == 84
    src, reqs = W.serving_requests(640)
    assert len(reqs) == 640 and len(src) == 128
    assert sum(r[0] == "verify" for r in reqs) == 128
    ml = [len(r[2]) for r in reqs if r[0] == "sign"]
    assert min(ml) >= 16 and max(ml) <= 65536


def test_fixture_shapes_and_verify_statuses():
    assert len(G["config2"]["sig"]) == 16 * 4096
    assert len(G["config3"]["sig"]) == 16 * 8192
    for B in (1, 64, 1024):
        assert len(G["config4"]["batches"][str(B)]["sig"]) == 16 * B
    g5 = G["config5"]
    src, reqs = W.serving_requests(g5["requests"])
    tampered = [r[2] is not None for r in reqs if r[0] == "verify"]
    assert g5["verify_status"] == [-1 if t else 0 for t in tampered]
    assert len(g5["source_sig"]) == 16 * len(src)
    assert len(g5["sign_sig"]) == 16 * sum(r[0] == "sign" for r in reqs)


def test_oracle_reproduces_config2_items():
    kseeds, msgs, opts = W.kat_batch(OracleDrbg, 16, 4096)
    pid = O.param_id("sha2-128f")
    for i in (0, 4095):
        pk, sk = O.seed_keypair(pid, kseeds[i])
        assert W.digest(O.signature_with_optrand(pid, msgs[i], sk, opts[i])) == \
            _dg(G["config2"]["sig"], i)


@pytest.mark.slow
def test_oracle_reproduces_config3_to_5_items():
    items = W.mixed_shake_batch(OracleDrbg, 8192)
    for i in (0, 8191):
        name, seed, m, opt = items[i]
        pid = O.param_id(name)
        pk, sk = O.seed_keypair(pid, seed)
        assert W.digest(O.signature_with_optrand(pid, m, sk, opt)) == _dg(G["config3"]["sig"], i)
    seeds, msgs, opts = W.sha2_128s_batches()[1]
    pid = O.param_id("sha2-128s")
    pk, sk = O.seed_keypair(pid, seeds[0])
    assert W.digest(O.signature_with_optrand(pid, msgs[0], sk, opts[0])) == \
        _dg(G["config4"]["batches"]["1"]["sig"], 0)
    kdef = W.serving_keys(OracleDrbg)
    src, reqs = W.serving_requests(G["config5"]["requests"])
    signs = [r for r in reqs if r[0] == "sign"]
    j = next(i for i, r in enumerate(signs) if kdef[r[1]][0] == "sha2-128f")
    _, k, m, o = signs[j]
    pid = O.param_id("sha2-128f")
    pk, sk = O.seed_keypair(pid, kdef[k][1])
    assert W.digest(O.signature_with_optrand(pid, m, sk, o[:16])) == \
        _dg(G["config5"]["sign_sig"], j)
