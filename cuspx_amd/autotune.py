"""Autotuner for the work splits of the hypertree and verify kernels.

k_ht gives each WOTS leaf S lanes (S in 1, 2, 4, 8, 16; each lane walks
ceil(len/S) chains). A small batch fills the chip only with a wider split; a
large one runs fastest at S=1, which needs no cross-lane reduction. k_vht
runs 256 / G requests per workgroup (G = 16 to 128 lanes per request on
average): wider groups shorten each layer's chain phase down to the longest
chain, narrower ones keep more requests busy per lane; G = 1 is speculative
verification (every chain from every start first, batches up to 256). The
crossovers depend
on the set and on the batch, so they are measured: for every set, operation
and batch size this times device-resident launches at each split the set
allows and writes the winners as

    <set> <sign|keygen|verify> <min_units> <lanes>

lines (a sign launch covers batch*d subtrees, keygen and verify launches
cover batch), the table libcuspx reads from tuning.txt beside itself or from
$CUSPX_TUNING.

    python -m cuspx_amd.autotune [--sets sha2-128f,...] [--out cuspx_amd/tuning.txt]
"""
import argparse
import os
import sys
import time

import numpy as np

import cuspx_amd as C

OPS = {"sign": 0, "keygen": 2, "verify": 1}
# powers of two: a winner holds from its batch up to the next grid point, so a
# 4x grid left e.g. 2048 requests on the split timed best at 1024
BATCHES_FAST = tuple(1 << i for i in range(13))   # 1 .. 4096
BATCHES_SMALL = tuple(1 << i for i in range(11))  # 1 .. 1024


def _time(fn, stream, reps):
    fn()
    stream.sync()
    e0, e1 = C.Event(), C.Event()
    best = float("inf")
    for _ in range(reps):
        e0.record(stream)
        fn()
        e1.record(stream)
        stream.sync()
        best = min(best, e0.elapsed_ms(e1))
    return best


def tune_set(name, batches, reps, log):
    pid = C.param_id(name)
    p = C.params(pid)
    B = max(batches)
    rng = np.random.default_rng(pid)
    D = C.DeviceBuffer
    stream = C.Stream()
    sh = stream.handle
    d_seed = D.from_array(rng.integers(0, 256, B * 3 * p.n, dtype=np.uint8))
    d_opt = D.from_array(rng.integers(0, 256, B * p.n, dtype=np.uint8))
    d_msg = D.from_array(rng.integers(0, 256, B * 32, dtype=np.uint8))
    d_moff = D.from_array(np.arange(B, dtype=np.uint64) * 32)
    d_mlen = D.from_array(np.full(B, 32, np.uint32))
    d_pk, d_sk, d_sig = D(B * p.pk_bytes), D(B * p.sk_bytes), D(B * p.sig_bytes)
    d_st = D(4 * B)
    ws = D(max(C.workspace_bytes(pid, B, op) for op in (0, 1, 2)))
    C.keygen_device(pid, B, d_seed.ptr, 3 * p.n, d_pk.ptr, d_sk.ptr, ws.ptr, sh)
    C.sign_device(pid, B, d_sk.ptr, p.sk_bytes, d_msg.ptr, d_moff.ptr, d_mlen.ptr, d_opt.ptr,
                  d_sig.ptr, p.sig_bytes, ws.ptr, sh)
    stream.sync()
    hi = 128 if p.len > 64 else 64
    widths = {0: (1, 2, 4, 8, 16), 2: (1, 2, 4, 8, 16),
              1: (1,) + tuple(G for G in (16, 32, 64, 128) if G <= hi)}
    rows = []
    for op_name, op in OPS.items():
        for b in batches:
            def run():
                if op == 0:
                    C.sign_device(pid, b, d_sk.ptr, p.sk_bytes, d_msg.ptr, d_moff.ptr,
                                  d_mlen.ptr, d_opt.ptr, d_sig.ptr, p.sig_bytes, ws.ptr, sh)
                elif op == 1:
                    C.verify_device(pid, b, d_pk.ptr, p.pk_bytes, d_sig.ptr, p.sig_bytes,
                                    d_msg.ptr, d_moff.ptr, d_mlen.ptr, d_st.ptr, ws.ptr, sh)
                else:
                    C.keygen_device(pid, b, d_seed.ptr, 3 * p.n, d_pk.ptr, d_sk.ptr, ws.ptr, sh)
            times, seen = {}, set()
            for S in widths[op]:
                C.set_split(pid, op, S)
                eff = C.get_split(pid, op, b)
                if eff in seen:
                    continue
                seen.add(eff)
                times[eff] = _time(run, stream, reps)
            C.set_split(pid, op, 0)
            default = C.get_split(pid, op, b)
            win = min(times, key=times.get)
            if op == 1:
                st = d_st.download(np.int32, b)
                if np.count_nonzero(st):
                    raise RuntimeError("%s: verify rejected a valid signature" % name)
            trees = b * p.d if op == 0 else b
            rows.append((name, op_name, trees, win))
            log("%-14s %-6s batch %5d  %s  -> S=%d (heuristic S=%d, %.3f vs %.3f ms)" % (
                name, op_name, b, " ".join("S%d %.3f" % kv for kv in sorted(times.items())), win,
                default, times[win], times.get(default, float("nan"))))
    for buf in (d_seed, d_opt, d_msg, d_moff, d_mlen, d_pk, d_sk, d_sig, d_st, ws):
        buf.free()
    return rows


def compact(rows):
    """Keep only the lines where a set's choice changes as the batch grows."""
    out, last = [], {}
    for name, op, trees, S in sorted(rows, key=lambda r: (r[0], r[1], r[2])):
        key = (name, op)
        if last.get(key) != S:
            out.append((name, op, 0 if key not in last else trees, S))
            last[key] = S
    return out


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--sets", default="all")
    ap.add_argument("--reps", type=int, default=3)
    ap.add_argument("--out", default=os.path.join(os.path.dirname(os.path.abspath(__file__)),
                                                  "tuning.txt"))
    a = ap.parse_args(argv)
    names = [C.params(i).name.decode() for i in range(12)] if a.sets == "all" else \
        a.sets.split(",")
    log = lambda s: print(s, flush=True)
    rows, t0 = [], time.perf_counter()
    for name in names:
        small = name.endswith("s") and ("192" in name or "256" in name)
        rows += tune_set(name, BATCHES_SMALL if small else BATCHES_FAST, a.reps, log)
    with open(a.out, "w") as f:
        f.write("# k_ht lanes per leaf (sign, keygen; min_units = subtrees) and k_vht lanes per\n"
                "# request (verify; min_units = requests): <set> <op> <min_units> <lanes>\n")
        f.write("# measured by python -m cuspx_amd.autotune on %s (%.0f s)\n"
                % (C.device_info()["name"], time.perf_counter() - t0))
        for r in compact(rows):
            f.write("%s %s %d %d\n" % r)
    log("wrote %s" % a.out)
    return 0


if __name__ == "__main__":
    sys.exit(main())
