"""CPU-side checks of the C ABI (no GPU needed): the library loads, exports
every symbol include/cuspx.h declares, its parameter table agrees with the
oracle's, and its host DRBG reproduces the NIST PQC KAT count-0 seed/msg
(PQCgenKAT_sign.cu:55-63)."""
import os
import re
import subprocess
import sys

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


def test_workspace_grows_with_the_batch():
    for pid in range(12):
        for op in (C.OP_SIGN, C.OP_VERIFY, C.OP_KEYGEN):
            sizes = [C.workspace_bytes(pid, n, op) for n in (1, 2, 7, 64, 1000, 4096)]
            assert sizes == sorted(sizes), (pid, op)


def test_verify_workspace_holds_the_speculative_candidates():
    """Verify workspaces carry the candidate chain ends of min(n, 256)
    requests (d x len x 16 values of n bytes each), so the size steps by that
    amount per request up to 256 requests and by far less beyond; and a
    forced speculative split falls back above 256 requests."""
    L = C.lib()
    for pid in range(12):
        p = C.params(pid)
        cand = p.d * p.len * 16 * p.n
        w = lambda n: C.workspace_bytes(pid, n, C.OP_VERIFY)
        assert w(2) - w(1) >= cand
        assert w(256) - w(255) >= cand
        assert w(258) - w(257) < cand
        try:
            assert L.cuspx_set_split(pid, C.OP_VERIFY, 1) == C.OK
            assert L.cuspx_get_split(pid, C.OP_VERIFY, 256) == 1
            assert L.cuspx_get_split(pid, C.OP_VERIFY, 257) != 1
        finally:
            L.cuspx_set_split(pid, C.OP_VERIFY, 0)


def test_split_validation_and_override():
    """cuspx_set_split accepts exactly the documented values (host logic only)."""
    L = C.lib()
    for pid in range(12):
        p = C.params(pid)
        hi = 128 if p.len > 64 else 64
        good = {C.OP_SIGN: [1, 2, 4, 8, 16], C.OP_KEYGEN: [1, 2, 4, 8, 16],
                C.OP_VERIFY: [1] + [g for g in (16, 32, 64, 128) if g <= hi]}
        bad = {C.OP_SIGN: [3, 32, -1, 1016], C.OP_KEYGEN: [5, 64],
               C.OP_VERIFY: [8, 2, 24, 256, 1016, 1032, 2 * hi]}
        try:
            for op in good:
                for v in good[op]:
                    assert L.cuspx_set_split(pid, op, v) == C.OK, (pid, op, v)
                    assert L.cuspx_get_split(pid, op, 5) == v, (pid, op, v)
                for v in bad[op]:
                    assert L.cuspx_set_split(pid, op, v) == C.ERR_ARG, (pid, op, v)
        finally:
            for op in good:
                L.cuspx_set_split(pid, op, 0)
    assert L.cuspx_set_split(99, 0, 1) == C.ERR_PARAM
    assert L.cuspx_set_split(1, 7, 1) == C.ERR_ARG


def test_tuning_table_is_read(tmp_path):
    """A table given by $CUSPX_TUNING drives the automatic choice: the entry
    with the largest threshold not above the launch size applies."""
    table = tmp_path / "t.txt"
    table.write_text("# test\nsha2-128f sign 0 4\nsha2-128f sign 2200 16\n"
                     "sha2-128f verify 0 32\nsha2-128f verify 50 16\nbogus line\n")
    code = ("import cuspx_amd as C\n"
            "L = C.lib()\n"
            "print(L.cuspx_get_split(1, 0, 1), L.cuspx_get_split(1, 0, 99),"
            " L.cuspx_get_split(1, 0, 100), L.cuspx_get_split(1, 1, 49), L.cuspx_get_split(1, 1, 50))\n")
    env = dict(os.environ, CUSPX_TUNING=str(table))
    out = subprocess.run([sys.executable, "-c", code], env=env, cwd=ROOT, capture_output=True,
                         text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    # sign thresholds count subtrees: batch * d (d = 22 for SHA2-128f)
    assert out.stdout.split() == ["4", "4", "16", "32", "16"]


def test_autotune_table_compaction():
    from cuspx_amd.autotune import compact
    rows = [("s", "sign", 22, 8), ("s", "sign", 88, 8), ("s", "sign", 352, 4),
            ("s", "sign", 1408, 4), ("s", "sign", 5632, 1), ("s", "verify", 1, 64),
            ("s", "verify", 4096, 16)]
    assert compact(rows) == [("s", "sign", 0, 8), ("s", "sign", 352, 4), ("s", "sign", 5632, 1),
                             ("s", "verify", 0, 64), ("s", "verify", 4096, 16)]


def test_server_refuses_more_rows_than_a_key_id_holds():
    """Key ids keep the row in 16 bits: a server asked for more resident keys
    per set is refused before it touches the GPU."""
    with pytest.raises(C.CuspxError) as e:
        C.Server(hbm_budget_bytes=1 << 30, max_keys_per_set=(1 << 16) + 1)
    assert e.value.code == C.ERR_ARG
