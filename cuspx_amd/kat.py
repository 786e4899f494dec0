"""NIST PQC known-answer files (.req / .rsp) for the engine.

Follows the reference's generator PQCgenKAT_sign.cu:28-158 and its hex helpers
(fprintBstr, ReadHex, FindMarker at PQCgenKAT_sign.cu:163-261). The request file
comes from the master DRBG seeded with entropy 00..2f: for each count, a
48-byte seed and a message of 33*(count+1) bytes. The response file reseeds
the DRBG with each seed, then runs crypto_sign_keypair, crypto_sign and
crypto_sign_open through libcuspx (the GPU path).

    python -m cuspx_amd.kat req PQCsignKAT_64.req [--count 100]
    python -m cuspx_amd.kat rsp PQCsignKAT_64.req PQCsignKAT_64.rsp --param sha2-128f [--drbg]
    python -m cuspx_amd.kat check PQCsignKAT_64.rsp --param sha2-128f [--drbg]
"""
import argparse
import sys

import cuspx_amd as C

FIELDS = ("count", "seed", "mlen", "msg", "pk", "sk", "smlen", "sm")
INT_FIELDS = ("count", "mlen", "smlen")


def _hex(b):
    # fprintBstr: upper-case hex, "00" for an empty string
    return b.hex().upper() if len(b) else "00"


def make_requests(count=100):
    """Request entries exactly as PQCgenKAT_sign.cu:55-70 draws them."""
    C.randombytes_init(bytes(range(48)))
    out = []
    for i in range(count):
        seed = C.randombytes(48)
        mlen = 33 * (i + 1)
        out.append({"count": i, "seed": seed, "mlen": mlen, "msg": C.randombytes(mlen)})
    return out


def write_req(path, entries):
    with open(path, "w") as f:
        for e in entries:
            f.write("count = %d\n" % e["count"])
            f.write("seed = %s\n" % _hex(e["seed"]))
            f.write("mlen = %d\n" % e["mlen"])
            f.write("msg = %s\n" % _hex(e["msg"]))
            f.write("pk =\nsk =\nsmlen =\nsm =\n\n")


def parse(path):
    """Entries of a .req or .rsp file (missing / empty fields are None)."""
    entries, cur = [], None
    with open(path) as f:
        for line in f:
            line = line.strip()
            if not line or line.startswith("#") or "=" not in line:
                continue
            key, _, val = (x.strip() for x in line.partition("="))
            if key == "count":
                cur = {k: None for k in FIELDS}
                entries.append(cur)
            if cur is None or key not in FIELDS:
                continue
            if key in INT_FIELDS:
                cur[key] = int(val) if val else None
            else:
                cur[key] = bytes.fromhex(val) if val else None
    for e in entries:  # ReadHex: a zero length reads as empty
        if e["mlen"] == 0:
            e["msg"] = b""
    return entries


def respond(entries, pid, debug):
    """Fill pk / sk / sm of each entry with the GPU engine (PQCgenKAT_sign.cu:103-156)."""
    C.set_default_param(pid)
    C.set_debug_mode(debug)
    out = []
    for e in entries:
        C.randombytes_init(e["seed"])
        pk, sk = C.crypto_sign_keypair()
        sm = C.crypto_sign(e["msg"], sk)
        rc, m1 = C.crypto_sign_open(sm, pk)
        if rc != 0 or m1 != e["msg"]:
            raise RuntimeError("crypto_sign_open failed for count %d" % e["count"])
        out.append(dict(e, pk=pk, sk=sk, smlen=len(sm), sm=sm))
    return out


def write_rsp(path, entries, algname):
    with open(path, "w") as f:
        f.write("# %s\n\n" % algname)
        for e in entries:
            f.write("count = %d\n" % e["count"])
            f.write("seed = %s\n" % _hex(e["seed"]))
            f.write("mlen = %d\n" % e["mlen"])
            f.write("msg = %s\n" % _hex(e["msg"]))
            f.write("pk = %s\n" % _hex(e["pk"]))
            f.write("sk = %s\n" % _hex(e["sk"]))
            f.write("smlen = %d\n" % e["smlen"])
            f.write("sm = %s\n\n" % _hex(e["sm"]))


def check(entries, pid, debug):
    """Regenerate every entry and compare; returns the mismatching counts."""
    fresh = respond([{k: e[k] for k in ("count", "seed", "mlen", "msg")} for e in entries],
                    pid, debug)
    return [e["count"] for e, g in zip(entries, fresh)
            if (e["pk"], e["sk"], e["sm"]) != (g["pk"], g["sk"], g["sm"])]


def main(argv=None):
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="cmd", required=True)
    r = sub.add_parser("req")
    r.add_argument("out")
    r.add_argument("--count", type=int, default=100)
    s = sub.add_parser("rsp")
    s.add_argument("req")
    s.add_argument("out")
    s.add_argument("--param", default="sha2-128f")
    s.add_argument("--drbg", action="store_true", help="DRBG seeds/optrand instead of DEBUG_MODE")
    c = sub.add_parser("check")
    c.add_argument("rsp")
    c.add_argument("--param", default="sha2-128f")
    c.add_argument("--drbg", action="store_true")
    a = ap.parse_args(argv)
    if a.cmd == "req":
        write_req(a.out, make_requests(a.count))
    elif a.cmd == "rsp":
        pid = C.param_id(a.param)
        write_rsp(a.out, respond(parse(a.req), pid, not a.drbg),
                  "SPHINCS+-%s-simple" % a.param)
    else:
        bad = check(parse(a.rsp), C.param_id(a.param), not a.drbg)
        print("mismatching counts:", bad)
        return 1 if bad else 0
    return 0


if __name__ == "__main__":
    sys.exit(main())
