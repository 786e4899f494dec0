"""CPU checks of the drop-in boundary (no GPU work is started).

* include/cuspx_params.h gives the reference's SPX_* / CRYPTO_* sizes
  (params.h:33-151, api.h:9-14) for all twelve -DSHA256 / -DSHAKE256 x
  -DSPX_<size> selections: a C program built with gcc prints them and they
  must equal the library's and the oracle's parameter tables.
* The drop-in callers of tests/dropin (shaped like the reference's
  test/spx.cu and test/dp_spx.cu) compile with g++ against
  include/cuspx_api.h alone, without warnings, and link to libcuspx.so;
  cuspx_api.h makes the compiled-in set the library's default at load.
* The batched calls refuse a message longer than the 32-byte stride
  (face_mdp_crypto_sign / face_ms_mdp_crypto_sign) or a record longer than
  SPX_BYTES + 32 (the _open forms), and the single-item calls a message of
  2^32 bytes or more, with -4 before anything is copied or launched.
"""
import ctypes
import os
import shutil
import subprocess

import pytest

import cuspx_amd as C
from oracle import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
INC = os.path.join(ROOT, "include")
SIZES = ["128S", "128F", "192S", "192F", "256S", "256F"]

PRINTER = r"""
#include <stdio.h>
#include "cuspx_params.h"
int main(void) {
    printf("%s %d %d %d %d %d %d %d %d %d %d %d %d %d %d %d %d %d %s %d\n", CUSPX_PARAM_NAME,
           SPX_N, SPX_FULL_HEIGHT, SPX_D, SPX_FORS_HEIGHT, SPX_FORS_TREES, SPX_WOTS_LEN,
           SPX_WOTS_BYTES, SPX_TREE_HEIGHT, SPX_FORS_MSG_BYTES, SPX_FORS_BYTES, SPX_BYTES,
           SPX_PK_BYTES, SPX_SK_BYTES, SPX_SM_BYTES, CRYPTO_BYTES, CRYPTO_SECRETKEYBYTES,
           CRYPTO_SEEDBYTES, CRYPTO_ALGNAME, SPX_MLEN);
    return 0;
}
"""


@pytest.fixture(scope="module")
def printer_src(tmp_path_factory):
    d = tmp_path_factory.mktemp("params")
    src = d / "p.c"
    src.write_text(PRINTER)
    return d, src


@pytest.mark.parametrize("hash_macro", ["SHA256", "SHAKE256"])
@pytest.mark.parametrize("size", SIZES)
def test_size_macros_match_parameter_tables(printer_src, hash_macro, size):
    d, src = printer_src
    exe = d / ("p_%s_%s" % (hash_macro, size))
    subprocess.run(["gcc", "-std=c99", "-Wall", "-Werror", "-D" + hash_macro, "-DSPX_" + size,
                    "-I" + INC, str(src), "-o", str(exe)], check=True)
    f = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()
    name = f[0]
    v = [int(x) for x in f[1:18]] + [int(f[19])]
    pid = C.param_id(name)
    assert name == ("sha2-" if hash_macro == "SHA256" else "shake256-") + size.lower()
    a, o = C.params(pid), O.params(pid)
    (n, h, dd, fa, fk, wlen, wbytes, th, fmsg, fbytes, sig, pkb, skb, smb, cb, csk, cseed,
     mlen) = v
    assert (n, h, dd, fa, fk) == (a.n, a.full_h, a.d, a.a, a.k)
    assert (wlen, wbytes, th) == (o.len, o.wots_bytes, o.tree_h)
    assert (fmsg, fbytes) == (o.fors_msg_bytes, o.fors_bytes)
    assert (sig, pkb, skb) == (a.sig_bytes, a.pk_bytes, a.sk_bytes) == (o.sig_bytes, o.pk_bytes,
                                                                         o.sk_bytes)
    assert (smb, cb, csk, cseed, mlen) == (sig + 32, sig, skb, 3 * n, 32)
    assert f[18] == "SPHINCS+"


def test_params_header_requires_a_selection(tmp_path):
    src = tmp_path / "x.c"
    src.write_text('#include "cuspx_params.h"\nint main(void){return SPX_N;}\n')
    for flags in ([], ["-DSHA256"], ["-DSPX_128F"], ["-DSHA256", "-DSHAKE256", "-DSPX_128F"]):
        p = subprocess.run(["gcc", "-I" + INC, "-c", str(src), "-o", str(tmp_path / "x.o")] +
                           flags, capture_output=True, text=True)
        assert p.returncode != 0, flags


@pytest.mark.parametrize("prog", ["spx_main.cpp", "dp_main.cpp"])
@pytest.mark.parametrize("sel", [("SHA256", "SPX_128F"), ("SHAKE256", "SPX_128F"),
                                 ("SHA256", "SPX_256S")])
def test_dropin_callers_build_with_gxx(tmp_path, prog, sel):
    src = os.path.join(ROOT, "tests", "dropin", prog)
    exe = tmp_path / "a.out"
    p = subprocess.run(["g++", "-O1", "-std=c++11", "-Wall", "-Wextra", "-Werror",
                        "-D" + sel[0], "-D" + sel[1], "-I" + INC, src, "-o", str(exe),
                        "-L" + os.path.join(ROOT, "cuspx_amd"), "-lcuspx"],
                       capture_output=True, text=True)
    assert p.returncode == 0, p.stderr[-3000:]


API_NAMES = ["randombytes_init", "randombytes", "crypto_sign_secretkeybytes",
             "crypto_sign_publickeybytes", "crypto_sign_bytes", "crypto_sign_seedbytes",
             "crypto_sign_seed_keypair", "crypto_sign_keypair", "crypto_sign_signature",
             "crypto_sign_verify", "crypto_sign", "crypto_sign_open",
             "face_mdp_crypto_sign_keypair", "face_mdp_crypto_sign", "face_mdp_crypto_sign_open",
             "face_ms_mdp_crypto_sign_keypair", "face_ms_mdp_crypto_sign",
             "face_ms_mdp_crypto_sign_open"]


def test_perset_libraries_export_the_api_names():
    """One libspx_<set>.so per parameter set (the reference builds one
    library per set), each defining every api.h / rng.h name and depending
    on libcuspx.so."""
    from cuspx_amd import build as B
    names = B.perset_names()
    assert len(names) == 12 and sorted(names) == sorted(C.params(i).name.decode()
                                                         for i in range(12))
    for name in names:
        path = B.perset_path(name)
        assert os.path.exists(path), path
        syms = subprocess.run(["nm", "-D", "--defined-only", path], capture_output=True,
                              text=True, check=True).stdout.split()
        assert [s for s in API_NAMES if s not in syms] == [], name
        dyn = subprocess.run(["readelf", "-d", path], capture_output=True, text=True,
                             check=True).stdout
        assert "[libcuspx.so]" in dyn, name


def test_perset_library_alone_picks_its_set(tmp_path):
    """A C caller with the plain declarations (no cuspx_api.h, so no
    constructor of its own) linked to libspx_<set>.so alone sees that set's
    sizes through the api.h getters; so does one that also links libcuspx.so
    after it (host code only: no device call).  Listed after libcuspx.so,
    the per-set library would be dropped by --as-needed, every name being
    found in libcuspx.so first."""
    src = tmp_path / "plain.c"
    src.write_text('#include <stdio.h>\n#include "cuspx.h"\n'
                   'int main(void){printf("%llu %llu %llu %llu\\n", crypto_sign_bytes(),'
                   ' crypto_sign_publickeybytes(), crypto_sign_secretkeybytes(),'
                   ' crypto_sign_seedbytes());return 0;}\n')
    lib = os.path.join(ROOT, "cuspx_amd")
    for name in ("shake256-192f", "sha2-128s", "sha2-256f"):
        p = C.params(C.param_id(name))
        want = [p.sig_bytes, p.pk_bytes, p.sk_bytes, 3 * p.n]
        stem = "spx_" + name.replace("-", "_")
        for order in (["-l" + stem], ["-l" + stem, "-lcuspx"]):
            exe = tmp_path / ("plain_%s_%d" % (stem, len(order)))
            subprocess.run(["gcc", "-Wall", "-Werror", "-I" + INC, str(src), "-o", str(exe),
                            "-L" + lib] + order + ["-Wl,-rpath," + lib], check=True)
            out = subprocess.run([str(exe)], capture_output=True, text=True, check=True)
            assert [int(x) for x in out.stdout.split()] == want, (name, order)


def test_dropin_caller_builds_against_plain_declarations(tmp_path):
    src = os.path.join(ROOT, "tests", "dropin", "spx_main.cpp")
    exe = tmp_path / "plain.out"
    p = subprocess.run(["g++", "-O1", "-std=c++11", "-Wall", "-Wextra", "-Werror",
                        "-DSHAKE256", "-DSPX_128F", "-DPLAIN_API_H", "-I" + INC, src, "-o",
                        str(exe), "-L" + os.path.join(ROOT, "cuspx_amd"), "-lspx_shake256_128f"],
                       capture_output=True, text=True)
    assert p.returncode == 0, p.stderr[-3000:]


def test_api_header_selects_the_default_set(tmp_path):
    """cuspx_api.h's load-time constructor sets the library's default to the
    compiled-in set (host code only: no device call)."""
    src = tmp_path / "sel.c"
    src.write_text('#include <stdio.h>\n#include "cuspx_api.h"\n'
                   'int main(void){printf("%d %llu %llu\\n", cuspx_get_default_param(),'
                   ' crypto_sign_bytes(), (unsigned long long)SPX_BYTES);return 0;}\n')
    lib = os.path.join(ROOT, "cuspx_amd")
    for hm, sm, name in (("SHAKE256", "SPX_192F", "shake256-192f"), ("SHA256", "SPX_128S",
                                                                     "sha2-128s")):
        exe = tmp_path / ("sel_" + sm)
        subprocess.run(["gcc", "-D" + hm, "-D" + sm, "-I" + INC, str(src), "-o", str(exe),
                        "-L" + lib, "-lcuspx", "-Wl,-rpath," + lib], check=True)
        out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()
        assert int(out[0]) == C.param_id(name)
        assert out[1] == out[2]


def test_batched_calls_refuse_long_messages():
    L = C.lib()
    pid = C.param_id("sha2-128f")
    C.set_default_param(pid)
    p = C.params(pid)
    buf = ctypes.create_string_buffer(64)
    n = ctypes.c_ulonglong()
    for name in ("face_mdp_crypto_sign", "face_ms_mdp_crypto_sign"):
        assert getattr(L, name)(buf, ctypes.byref(n), buf, 33, buf, 4) == C.ERR_ARG, name
        assert b"SPX_MLEN" in L.cuspx_last_error()
    for name in ("face_mdp_crypto_sign_open", "face_ms_mdp_crypto_sign_open"):
        assert getattr(L, name)(buf, ctypes.byref(n), buf, p.sig_bytes + 33, buf, 4) == C.ERR_ARG
    siglen = ctypes.c_size_t()
    assert L.crypto_sign_signature(buf, ctypes.byref(siglen), buf, 1 << 32, buf) == C.ERR_ARG
    assert L.crypto_sign_verify(buf, p.sig_bytes, buf, 1 << 32, buf) == C.ERR_ARG


def test_gxx_is_available():
    assert shutil.which("g++") and shutil.which("gcc")
