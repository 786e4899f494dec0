"""CPU-side checks of the C ABI (no GPU needed): the library loads, exports
every symbol include/cuspx.h declares, its parameter table agrees with the
oracle's, and its host DRBG reproduces the NIST PQC KAT count-0 seed/msg
(PQCgenKAT_sign.cu:55-63)."""
import os
import re

import pytest

import cuspx_amd as C
from oracle import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared_symbols():
    src = open(os.path.join(ROOT, "include", "cuspx.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    names = re.findall(r"\b([A-Za-z_][A-Za-z0-9_]*)\s*\(", src)
    skip = {"extern", "if", "sizeof", "typedef"}
    return sorted(set(n for n in names if n not in skip))


def test_library_exports_every_declared_symbol():
    L = C.lib()
    syms = _declared_symbols()
    assert len(syms) >= 30
    missing = [s for s in syms if not hasattr(L, s)]
    assert missing == []


def test_param_table_matches_oracle():
    for pid in range(12):
        a, b = C.params(pid), O.params(pid)
        assert a.name == b.name
        for f in ("n", "full_h", "d", "a", "k", "len", "sig_bytes", "pk_bytes", "sk_bytes"):
            assert getattr(a, f) == getattr(b, f), (a.name, f)
        assert C.param_id(a.name.decode()) == pid


def test_unknown_param_is_an_error():
    assert C.lib().cuspx_param_id(b"sha2-129f") == C.ERR_PARAM
    with pytest.raises(KeyError):
        C.params(12)
    assert C.lib().spx_batch_sign(99, 1, None, None, None, None, None, None) == C.ERR_PARAM


def test_host_drbg_nist_kat():
    C.randombytes_init(bytes(range(48)))
    assert C.randombytes(48).hex().upper() == (
        "061550234D158C5EC95595FE04EF7A25767F2E24CC2BC479D09D86DC9ABCFDE7"
        "056A8C266F9EF97ED08541DBD2E1FFA1")
    assert C.randombytes(33).hex().upper() == (
        "D81C4D8D734FCBFBEADE3D3F8A039FAA2A2C9957E835AD55B22E75BF57BB556AC8")


def test_host_drbg_matches_oracle_stream():
    ent = bytes((7 * i + 3) & 0xFF for i in range(48))
    C.randombytes_init(ent)
    O.randombytes_init(ent)
    for n in (1, 15, 16, 17, 48, 100, 1000):
        assert C.randombytes(n) == O.randombytes(n)


def test_workspace_sizes():
    for pid in range(12):
        for op in (C.OP_SIGN, C.OP_VERIFY, C.OP_KEYGEN):
            assert C.workspace_bytes(pid, 64, op) > 0
    assert C.workspace_bytes(99, 64, 0) == 0
