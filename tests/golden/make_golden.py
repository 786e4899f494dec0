"""Writes tests/golden/golden.json from the CPU oracle.

Regression fixtures (the reference ships no known-answer vectors): for every
parameter set, a keypair from a fixed 3n-byte seed and one signature over a
fixed message with fixed optrand; plus the KAT count-0 protocol
(PQCgenKAT_sign.cu:55-131) for SHA2-128f in both DEBUG_MODE and DRBG mode.
Full signatures are stored as SHA-256 digests to keep the file small.

    python tests/golden/make_golden.py
"""
import hashlib
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import oracle as O  # noqa: E402


def main():
    out = {"sets": {}, "kat0": {}}
    for pid in range(12):
        p = O.params(pid)
        seed = bytes((7 * i + pid) & 0xFF for i in range(3 * p.n))
        msg = bytes((3 * i + 1) & 0xFF for i in range(77))
        opt = bytes((11 * i + 5) & 0xFF for i in range(p.n))
        pk, sk = O.seed_keypair(pid, seed)
        sig = O.signature_with_optrand(pid, msg, sk, opt)
        out["sets"][p.name.decode()] = {
            "seed": seed.hex(), "msg": msg.hex(), "optrand": opt.hex(), "pk": pk.hex(),
            "sk": sk.hex(), "sig_sha256": hashlib.sha256(sig).hexdigest(),
            "sig_head": sig[:64].hex(), "sig_tail": sig[-64:].hex()}
    pid = O.param_id("sha2-128f")
    for mode, debug in (("debug", True), ("drbg", False)):
        O.set_debug_mode(debug)
        O.randombytes_init(bytes(range(48)))
        seed48 = O.randombytes(48)
        msg = O.randombytes(33)
        O.randombytes_init(seed48)
        pk, sk = O.keypair(pid)
        sm = O.sign(pid, msg, sk)
        out["kat0"][mode] = {"seed": seed48.hex(), "msg": msg.hex(), "pk": pk.hex(),
                             "sk": sk.hex(), "smlen": len(sm),
                             "sm_sha256": hashlib.sha256(sm).hexdigest()}
    O.set_debug_mode(True)
    with open(os.path.join(HERE, "golden.json"), "w") as f:
        json.dump(out, f, indent=1, sort_keys=True)


if __name__ == "__main__":
    main()
