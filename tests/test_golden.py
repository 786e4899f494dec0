"""Committed fixtures (tests/golden/golden.json, written by
tests/golden/make_golden.py) and the NIST KAT file format (cuspx_amd/kat.py,
PQCgenKAT_sign.cu:28-158).

CPU: the oracle still reproduces the fixtures; the library's DRBG draws the
KAT request entries; .req/.rsp files survive a write/parse round trip.
GPU: the engine's keypairs, signatures and KAT count-0 responses equal the
fixtures byte for byte."""
import hashlib
import json
import os

import numpy as np
import pytest

import cuspx_amd as C
from cuspx_amd import kat
from oracle import oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))
G = json.load(open(os.path.join(HERE, "golden", "golden.json")))
CPU_SIG_SETS = ("sha2-128f", "sha2-192f", "shake256-128f", "sha2-128s")


@pytest.mark.parametrize("name", sorted(G["sets"]))
def test_oracle_keypair_matches_fixture(name):
    g = G["sets"][name]
    pk, sk = O.seed_keypair(O.param_id(name), bytes.fromhex(g["seed"]))
    assert (pk.hex(), sk.hex()) == (g["pk"], g["sk"])


@pytest.mark.parametrize("name", CPU_SIG_SETS)
def test_oracle_signature_matches_fixture(name):
    g = G["sets"][name]
    sig = O.signature_with_optrand(O.param_id(name), bytes.fromhex(g["msg"]),
                                   bytes.fromhex(g["sk"]), bytes.fromhex(g["optrand"]))
    assert hashlib.sha256(sig).hexdigest() == g["sig_sha256"]
    assert sig[:64].hex() == g["sig_head"] and sig[-64:].hex() == g["sig_tail"]


def test_kat_requests_use_the_nist_drbg():
    e = kat.make_requests(3)
    assert [x["mlen"] for x in e] == [33, 66, 99]
    for mode in ("debug", "drbg"):
        assert e[0]["seed"].hex() == G["kat0"][mode]["seed"]
        assert e[0]["msg"].hex() == G["kat0"][mode]["msg"]


def test_kat_file_round_trip(tmp_path):
    req = kat.make_requests(4)
    kat.write_req(tmp_path / "a.req", req)
    text = (tmp_path / "a.req").read_text()
    assert text.startswith("count = 0\nseed = " + req[0]["seed"].hex().upper())
    assert "pk =\nsk =\nsmlen =\nsm =\n" in text
    back = kat.parse(tmp_path / "a.req")
    assert [(b["count"], b["seed"], b["mlen"], b["msg"]) for b in back] == \
        [(r["count"], r["seed"], r["mlen"], r["msg"]) for r in req]
    assert back[0]["pk"] is None and back[0]["smlen"] is None
    rsp = [dict(r, pk=b"\x01" * 32, sk=b"\x02" * 64, smlen=r["mlen"] + 3,
                sm=b"\xab\xcd\xef" + r["msg"]) for r in req]
    rsp.append({"count": 4, "seed": b"\x00" * 48, "mlen": 0, "msg": b"", "pk": b"\x03" * 32,
                "sk": b"\x04" * 64, "smlen": 1, "sm": b"\x05"})
    kat.write_rsp(tmp_path / "a.rsp", rsp, "SPHINCS+-sha2-128f-simple")
    text = (tmp_path / "a.rsp").read_text()
    assert text.startswith("# SPHINCS+-sha2-128f-simple\n\ncount = 0\n")
    assert "mlen = 0\nmsg = 00\n" in text  # fprintBstr prints 00 for an empty string
    assert kat.parse(tmp_path / "a.rsp") == rsp


@pytest.mark.gpu
@pytest.mark.parametrize("name", sorted(G["sets"]))
def test_gpu_matches_fixture(name):
    g = G["sets"][name]
    pid = C.param_id(name)
    pks, sks = C.batch_keygen(pid, np.frombuffer(bytes.fromhex(g["seed"]), np.uint8))
    assert (bytes(pks[0]).hex(), bytes(sks[0]).hex()) == (g["pk"], g["sk"])
    msg = bytes.fromhex(g["msg"])
    sk = np.frombuffer(bytes.fromhex(g["sk"]), np.uint8)
    opt = np.frombuffer(bytes.fromhex(g["optrand"]), np.uint8)
    sig = bytes(C.batch_sign(pid, sk, [msg], opt)[0])
    assert hashlib.sha256(sig).hexdigest() == g["sig_sha256"]
    pk = np.frombuffer(bytes.fromhex(g["pk"]), np.uint8)
    assert C.batch_verify(pid, pk, np.frombuffer(sig, np.uint8), [msg]).tolist() == [0]


@pytest.mark.gpu
@pytest.mark.parametrize("mode", ["debug", "drbg"])
def test_gpu_kat_count0_matches_fixture(mode):
    g = G["kat0"][mode]
    out = kat.respond(kat.make_requests(1), C.param_id("sha2-128f"), mode == "debug")
    C.set_debug_mode(True)
    e = out[0]
    assert (e["pk"].hex(), e["sk"].hex(), e["smlen"]) == (g["pk"], g["sk"], g["smlen"])
    assert hashlib.sha256(e["sm"]).hexdigest() == g["sm_sha256"]
