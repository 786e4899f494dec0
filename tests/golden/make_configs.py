"""Writes tests/golden/configs.json: expected outputs of the benchmark
configurations at their full sizes, computed by the CPU oracle.

    python tests/golden/make_configs.py [--procs 8]

Inputs come from cuspx_amd/workloads.py with the oracle's DRBG (the same
stream as the library's).  Every signature is stored as the first 8 bytes
of its SHA-256 (``workloads.digest``), in item order, as one hex string per
batch; public keys as one SHA-256 over the batch's keys.  Signing runs in
forked processes, one oracle library each.

* config2: SHA2-128f, 4096 KAT-protocol keypairs (DRBG mode), 32-byte
  messages.
* config3: 8192 mixed SHAKE256-128f / -192f requests, messages 1 B .. 4 KiB.
* config4: SHA2-128s batches of 1, 64 and 1024.
* config5: 256 serving keys (SHA2-128f / SHA2-128s / SHAKE256-256f), the
  signatures of the verify sources, a 640-request stream (80% sign, 20%
  verify with a quarter tampered) and the exact verify statuses.
"""
import argparse
import hashlib
import json
import multiprocessing as mp
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from cuspx_amd import workloads as W  # noqa: E402
from oracle import oracle as O  # noqa: E402

SERVE_REQUESTS = 640


class OracleDrbg:
    randombytes_init = staticmethod(O.randombytes_init)
    randombytes = staticmethod(O.randombytes)


def _keygen(task):
    name, seed = task
    return O.seed_keypair(O.param_id(name), seed)


def _sign(task):
    name, sk, msg, opt = task
    pid = O.param_id(name)
    return W.digest(O.signature_with_optrand(pid, msg, sk, opt[:O.params(pid).n]))


def _sign_full(task):
    name, sk, msg, opt = task
    pid = O.param_id(name)
    return O.signature_with_optrand(pid, msg, sk, opt[:O.params(pid).n])


def _verify(task):
    name, pk, msg, sig = task
    return O.verify(O.param_id(name), sig, msg, pk)


def _run(pool, fn, tasks, what):
    t0 = time.time()
    out = list(pool.imap(fn, tasks, chunksize=max(1, len(tasks) // 256)))
    print("  %s: %d items in %.0f s" % (what, len(tasks), time.time() - t0), flush=True)
    return out


def config2(pool):
    kseeds, msgs, opts = W.kat_batch(OracleDrbg, 16, 4096)
    keys = _run(pool, _keygen, [("sha2-128f", s) for s in kseeds], "config2 keygen")
    sigs = _run(pool, _sign, [("sha2-128f", sk, m, o) for (pk, sk), m, o in
                              zip(keys, msgs, opts)], "config2 sign")
    return {"set": "sha2-128f", "count": 4096,
            "pk_sha256": hashlib.sha256(b"".join(pk for pk, _ in keys)).hexdigest(),
            "sig": b"".join(sigs).hex()}


def config3(pool):
    items = W.mixed_shake_batch(OracleDrbg, 8192)
    keys = _run(pool, _keygen, [(name, s) for name, s, _, _ in items], "config3 keygen")
    sigs = _run(pool, _sign, [(it[0], sk, it[2], it[3]) for it, (pk, sk) in zip(items, keys)],
                "config3 sign")
    return {"sets": ["shake256-128f", "shake256-192f"], "count": len(items),
            "pk_sha256": hashlib.sha256(b"".join(pk for pk, _ in keys)).hexdigest(),
            "sig": b"".join(sigs).hex()}


def config4(pool):
    out = {}
    for B, (seeds, msgs, opts) in W.sha2_128s_batches().items():
        keys = _run(pool, _keygen, [("sha2-128s", s) for s in seeds], "config4 keygen %d" % B)
        sigs = _run(pool, _sign, [("sha2-128s", sk, m, o) for (pk, sk), m, o in
                                  zip(keys, msgs, opts)], "config4 sign %d" % B)
        out[str(B)] = {"pk_sha256": hashlib.sha256(b"".join(pk for pk, _ in keys)).hexdigest(),
                       "sig": b"".join(sigs).hex()}
    return {"set": "sha2-128s", "batches": out}


def config5(pool):
    kdef = W.serving_keys(OracleDrbg)
    keys = _run(pool, _keygen, kdef, "config5 keygen")
    sources, reqs = W.serving_requests(SERVE_REQUESTS)
    src_sigs = _run(pool, _sign_full, [(kdef[k][0], keys[k][1], m, o) for k, m, o in sources],
                    "config5 sources")
    signs = [r for r in reqs if r[0] == "sign"]
    sign_dg = _run(pool, _sign, [(kdef[k][0], keys[k][1], m, o) for _, k, m, o in signs],
                   "config5 sign")
    vt = []
    for r in reqs:
        if r[0] != "verify":
            continue
        k, m, _ = sources[r[1]]
        m2, s2 = W.apply_tamper(r[2], m, src_sigs[r[1]])
        vt.append((kdef[k][0], keys[k][0], m2, s2))
    status = _run(pool, _verify, vt, "config5 verify")
    return {"sets": list(W.SERVE_SETS), "keys": len(kdef), "requests": len(reqs),
            "pk_sha256": hashlib.sha256(b"".join(pk for pk, _ in keys)).hexdigest(),
            "source_sig": b"".join(W.digest(s) for s in src_sigs).hex(),
            "sign_sig": b"".join(sign_dg).hex(),
            "verify_status": status}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--procs", type=int, default=os.cpu_count() or 1)
    ap.add_argument("--only", default="")
    args = ap.parse_args()
    O.set_debug_mode(False)
    path = os.path.join(HERE, "configs.json")
    out = json.load(open(path)) if os.path.exists(path) else {}
    with mp.get_context("fork").Pool(args.procs) as pool:
        for name, fn in (("config2", config2), ("config3", config3), ("config4", config4),
                         ("config5", config5)):
            if args.only and name not in args.only.split(","):
                continue
            print(name, flush=True)
            out[name] = fn(pool)
            with open(path, "w") as f:
                json.dump(out, f, indent=1, sort_keys=True)


if __name__ == "__main__":
    main()
