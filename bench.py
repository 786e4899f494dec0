"""Benchmark: SPHINCS+-SHA2-128f-simple batched signing and verification on
one MI355X (inputs resident in HBM), with the CPU oracle timed on the host
cores beside it.

Workload (one "step"): one batched sign launch over B distinct keypairs
(default 4096) with one 32-byte message each, then one batched verify launch
over the B signatures.  Keys follow the NIST KAT protocol in DRBG mode:
master DRBG entropy 00..2f, count i draws seed_i (48 B) and msg_i (32 B);
DRBG(seed_i) then yields the 3n-byte key seed and the n-byte optrand.

Prints one JSON line.  `value` is signs/s (B / mean batched-sign time);
verifies/s is reported beside it.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
# The serving phase keeps several batches in flight on their own streams; give
# the process more hardware queues than streams, or streams share queues and
# serialise (read by the HIP runtime when it initialises).
os.environ.setdefault("GPU_MAX_HW_QUEUES", "16")

# MI355X VALU issue rates (measured by an issue-rate probe, profiles/valu_issue_rates.md):
# the simple ops (v_xor/v_add/v_lshrrev ..._e32, v_bitop3_b32) issue a wave64
# instruction every 2 cycles per SIMD, i.e. 32 lanes/clk -- the rate behind the
# published 157.3 TFLOP/s fp32 vector peak; the other 3-operand ops (v_alignbit,
# v_add3, v_perm, 64-bit shifts) every 4 cycles.
CLK_HZ = 2.4e9
SIMDS = 256 * 4
PEAK_VALU_T = SIMDS * 32 * CLK_HZ / 1e12    # 78.64 T lane-instr/s (fast class)
PEAK_VOP3_T = SIMDS * 16 * CLK_HZ / 1e12    # 39.32 T lane-instr/s (4-cycle class)
# VALU lane-instructions per unit when profiles/pmc_counters.json has no
# SQ_INSTS_VALU count for the kernel and batch (that count takes precedence)
SHA256_OPS = 1237
KECCAK_OPS = 4163
# a generic SHA-256 compression priced at 1450 VALU ops (F needs fewer); the
# compression rate against 39.32 T / 1450 is reported as frac_at_1450_ops
SHA256_OPS_UNFUSED = 1450
# Issue rate of independent instruction blocks in the proportions of the unit's
# loop body (wave-instructions per SIMD per ns, "mix" rows of profiles/valu_issue_rates.md):
# SHA-256 F = 717 four-cycle (508 alignbit, 205 add3, 4 perm) : 506 two-cycle
# (312 bitop3, 194 VOP2) ~ 7:5; Keccak-f round of a WOTS chain step (bit-interleaved)
# = 52 alignbit : 122 bitop3/xor ~ 3:7.
# Mixing the two classes issues slower than the sum of their separate rates,
# so this measured rate, not a 4c/2c model, is the ceiling of a unit.
MIX_RATE_SHA256 = 0.6235
MIX_RATE_KECCAK = 0.7501


def mix_ceiling_units(rate, ops_per_unit):
    """Units/s if every SIMD issued the unit's instruction mix back to back."""
    return rate * 1e9 * SIMDS * 64 / ops_per_unit


def kernel_work(p, batch):
    """Compressions (SHA-2) or permutations (SHAKE) of the hypertree tiles of
    one sign launch: every layer's WOTS leaves (PRF + 15 F per chain, T_len)
    and subtree nodes."""
    th = p.full_h // p.d
    leaves = batch * p.d * (1 << th)
    nodes = batch * p.d * ((1 << th) - 1)
    if p.hash == 0:
        tlen = -(-(22 + p.len * p.n + 9) // 64)
        hblk = -(-(22 + 2 * p.n + 9) // 64)
        return leaves * (p.len * 16 + tlen) + nodes * hblk
    tlen = (p.n + 32 + p.len * p.n) // 136 + 1
    return leaves * (p.len * 16 + tlen) + nodes


def fors_work(p, batch):
    """Compressions / permutations of the FORS tiles of one sign launch:
    k trees of 2^a leaves (PRF + F each) and 2^a - 1 nodes."""
    hblk = -(-(22 + 2 * p.n + 9) // 64) if p.hash == 0 else 1
    return batch * p.k * ((1 << p.a) * 2 + ((1 << p.a) - 1) * hblk)


def host_cpus():
    """(CPU model, logical CPUs this process may use, one of them per
    physical core) from /proc/cpuinfo and the sysfs topology."""
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    allowed = sorted(os.sched_getaffinity(0))
    seen, phys = set(), []
    for c in allowed:
        base = "/sys/devices/system/cpu/cpu%d/topology/" % c
        try:
            key = (open(base + "physical_package_id").read().strip(),
                   open(base + "core_id").read().strip())
        except OSError:
            key = ("?", str(c))
        if key not in seen:
            seen.add(key)
            phys.append(c)
    return model, allowed, phys


def cpu_baseline(pid, seconds, max_procs):
    """The oracle (single-threaded C restatement of the reference host path)
    timed on the host cores:

    * one core: one process pinned to one physical core, 2 warm-up then 10
      timed signatures (32-byte messages) and 10 verifications; medians on
      CLOCK_MONOTONIC and CLOCK_PROCESS_CPUTIME_ID;
    * all cores: one process per physical core (at most max_procs), each
      pinned to its core, signing for 0.8 x `seconds` then verifying for
      0.2 x `seconds`; aggregate rate = sum of the per-process rates."""
    import multiprocessing as mp
    import statistics

    def setup(k):
        from oracle import oracle as O
        import random
        p = O.params(pid)
        rng = random.Random(k)
        pk, sk = O.seed_keypair(pid, rng.randbytes(3 * p.n))
        return O, p, rng, pk, sk

    def single(q, cpu):
        os.sched_setaffinity(0, {cpu})
        O, p, rng, pk, sk = setup(0)
        mono, proc, vmono, vproc, sigs = [], [], [], [], []
        for it in range(12):
            m, opt = rng.randbytes(32), rng.randbytes(p.n)
            a, b = time.clock_gettime_ns(time.CLOCK_MONOTONIC), time.process_time_ns()
            s = O.signature_with_optrand(pid, m, sk, opt)
            a2, b2 = time.clock_gettime_ns(time.CLOCK_MONOTONIC), time.process_time_ns()
            assert O.verify(pid, s, m, pk) == 0
            a3, b3 = time.clock_gettime_ns(time.CLOCK_MONOTONIC), time.process_time_ns()
            if it >= 2:
                mono.append(a2 - a)
                proc.append(b2 - b)
                vmono.append(a3 - a2)
                vproc.append(b3 - b2)
        md = lambda x: statistics.median(x) / 1e6
        q.put({"sign_ms_monotonic": md(mono), "sign_ms_process_cpu": md(proc),
               "verify_ms_monotonic": md(vmono), "verify_ms_process_cpu": md(vproc)})

    def work(q, k, cpu):
        os.sched_setaffinity(0, {cpu})
        O, p, rng, pk, sk = setup(k)
        sigs, n_sign, t0 = [], 0, time.perf_counter()
        while time.perf_counter() - t0 < 0.8 * seconds or n_sign == 0:
            m = rng.randbytes(32)
            sigs.append((m, O.signature_with_optrand(pid, m, sk, rng.randbytes(p.n))))
            n_sign += 1
        t_sign = time.perf_counter() - t0
        n_ver, t0 = 0, time.perf_counter()
        while time.perf_counter() - t0 < 0.2 * seconds or n_ver == 0:
            m, s = sigs[n_ver % len(sigs)]
            assert O.verify(pid, s, m, pk) == 0
            n_ver += 1
        q.put((n_sign / t_sign, n_ver / (time.perf_counter() - t0)))

    model, allowed, phys = host_cpus()
    cores = phys[:max_procs]
    ctx = mp.get_context("fork")
    q = ctx.Queue()
    p1 = ctx.Process(target=single, args=(q, cores[0]))
    p1.start()
    one = q.get()
    p1.join()
    ps = [ctx.Process(target=work, args=(q, k, c)) for k, c in enumerate(cores)]
    for p in ps:
        p.start()
    res = [q.get() for _ in ps]
    for p in ps:
        p.join()
    return {"signs_per_s": sum(r[0] for r in res), "verifies_per_s": sum(r[1] for r in res),
            "processes": len(cores), "single_core": one, "cpu_model": model,
            "physical_cores_visible": len(phys), "logical_cpus_visible": len(allowed)}


class LibDrbg:
    """The library's host DRBG as the workload generators' draw source."""

    def __init__(self, C):
        self.randombytes_init = C.randombytes_init
        self.randombytes = C.randombytes


def make_inputs(C, pid, batch):
    """KAT-protocol inputs (workloads.kat_batch): per-key 3n-byte seeds,
    optrand, 32-byte messages."""
    import numpy as np
    from cuspx_amd import workloads as W
    kseeds, msgs, opts = W.kat_batch(LibDrbg(C), C.params(pid).n, batch)
    to = lambda xs: np.frombuffer(b"".join(xs), np.uint8).copy()
    return to(kseeds), to(opts), to(msgs)


def host_phase(C, pid, sk, pk, m, num, reps):
    """The reference's batched entry points on host buffers: one
    face_mdp_crypto_sign and one face_mdp_crypto_sign_open call per rep
    (sign.cu:5098, 6856 signatures; fixed 32-byte message stride), so the
    rates include the PCIe copies, the host staging and each call's
    synchronisation.  Outputs are preallocated and touched first."""
    import ctypes
    import numpy as np
    p = C.params(pid)
    L = C.lib()
    C.set_default_param(pid)
    sm = np.ones((num, p.sig_bytes + 32), np.uint8)
    mo = np.ones((num, 32), np.uint8)
    smlen, mlen = ctypes.c_ulonglong(0), ctypes.c_ulonglong(0)
    ptr = lambda a: a.ctypes.data_as(ctypes.c_void_p)

    out = {}
    for pre in ("face_mdp", "face_ms_mdp"):
        fs, fo = getattr(L, pre + "_crypto_sign"), getattr(L, pre + "_crypto_sign_open")

        def sign():
            assert fs(ptr(sm), ctypes.byref(smlen), ptr(m), 32, ptr(sk), num) == 0

        def open_():
            assert fo(ptr(mo), ctypes.byref(mlen), ptr(sm), smlen.value, ptr(pk), num) == 0

        sign()
        open_()  # grows the library's device arenas and pinned staging once
        ts, tv = [], []
        for _ in range(reps):
            t0 = time.perf_counter()
            sign()
            ts.append(time.perf_counter() - t0)
            mo.fill(1)
            t0 = time.perf_counter()
            open_()
            tv.append(time.perf_counter() - t0)
            assert bytes(mo.reshape(-1)) == bytes(m)
        key = "" if pre == "face_mdp" else "ms_"
        out.update({key + "signs_per_s": round(num / min(ts), 1),
                    key + "verifies_per_s": round(num / min(tv), 1),
                    key + "sign_call_ms": round(min(ts) * 1e3, 3),
                    key + "open_call_ms": round(min(tv) * 1e3, 3)})
    out["calls"] = ("face_mdp_crypto_sign / face_mdp_crypto_sign_open and the multi-stream "
                    "face_ms_mdp_* (ms_ keys) on pageable host buffers (%d items, best of %d); "
                    "PCIe-inclusive, never `value`" % (num, reps))
    return out


SERVE_SETS = ["sha2-128f", "sha2-128s", "shake256-256f"]


def serve_phase(C, seconds, qps, sat_requests, single_requests, slots=4, window_us=1000,
                sweep=(), sweep_seconds=5.0, long_slots=0,
                seed=5):
    """Serving workload: 256 resident keys over SHA2-128f / SHA2-128s /
    SHAKE256-256f, Poisson arrivals at `qps`, 80% sign / 20% verify, message
    sizes log-uniform in 16 B .. 64 KiB.  Returns latency percentiles at the
    fixed QPS, the saturated throughput of the batched server and the
    throughput of plain per-request launches on the same mix."""
    import random
    import numpy as np
    from cuspx_amd import workloads as W
    rng = random.Random(seed)
    keys = []
    kdef = W.serving_keys(LibDrbg(C), seed)
    for name in SERVE_SETS:
        pid = C.param_id(name)
        seeds = [s for nm, s in kdef if nm == name]
        pk, sk = C.batch_keygen(pid, np.frombuffer(b"".join(seeds), np.uint8))
        keys += [(pid, bytes(sk[i]), bytes(pk[i])) for i in range(len(seeds))]
    pool = W.MessagePool(seed)

    def msg():
        return pool.draw(rng)

    # signatures for the verify requests: 2 messages per key
    signed = []
    for pid in sorted(set(k[0] for k in keys)):
        idx = [i for i, k in enumerate(keys) if k[0] == pid]
        ms = [msg() for _ in idx for _ in range(2)]
        sks = np.frombuffer(b"".join(keys[i][1] for i in idx for _ in range(2)), np.uint8)
        sigs = C.batch_sign(pid, sks, ms)
        signed += [(idx[j // 2], ms[j], bytes(sigs[j])) for j in range(len(ms))]

    def mix(n):
        out = []
        for _ in range(n):
            if rng.random() < 0.8:
                out.append((0, rng.randrange(len(keys)), msg(), None))
            else:
                k, m, s = signed[rng.randrange(len(signed))]
                out.append((1, k, m, s))
        return out

    def run(srv, kids, reqs, gaps):
        tickets = []
        t0 = time.perf_counter()
        t_next = t0
        for (op, k, m, s), g in zip(reqs, gaps):
            t_next += g
            while True:
                d = t_next - time.perf_counter()
                if d <= 0:
                    break
                if d > 2e-4:
                    time.sleep(d - 1.5e-4)
            t = srv.submit_sign(kids[k], m) if op == 0 else srv.submit_verify(kids[k], m, s)
            tickets.append((t, op))
        bad = 0
        for t, op in tickets:
            st, _ = srv.wait(t)
            bad += st != 0
        el = time.perf_counter() - t0
        return el, bad

    out = {"sets": SERVE_SETS, "keys": len(keys), "mix": "80% sign / 20% verify",
           "message_bytes": "log-uniform 16..65536", "slots": slots, "batch_window_us": window_us,
           "long_slots": min(long_slots, slots) or slots}
    with C.Server(hbm_budget_bytes=32 << 30, slots=slots, batch_window_us=window_us,
                  long_slots=long_slots) as srv:
        kids = [srv.add_key(pid, sk, pk) for pid, sk, pk in keys]
        n = int(qps * seconds)
        reqs = mix(n)
        gaps = [rng.expovariate(qps) for _ in range(n)]
        el, bad = run(srv, kids, reqs, gaps)
        st = srv.stats()
        out.update({"qps_offered": qps, "requests": n, "seconds": round(el, 3),
                    "throughput_rps": round(n / el, 1), "invalid": bad,
                    "sign_p50_ms": round(st["sign_p50_us"] / 1e3, 3),
                    "sign_p99_ms": round(st["sign_p99_us"] / 1e3, 3),
                    "verify_p50_ms": round(st["verify_p50_us"] / 1e3, 3),
                    "verify_p99_ms": round(st["verify_p99_us"] / 1e3, 3),
                    "mean_batch": round(st["mean_batch"], 2)})
        for name in SERVE_SETS:
            a, z, c = srv.latency(C.param_id(name), 0)
            out["sign_ms_" + name] = [round(a / 1e3, 3), round(z / 1e3, 3), c]
        for k, op in (("sign_queue_wait_ms", 0), ("sign_execution_ms", 1)):
            a, z, c = srv.latency(-1, op)
            out[k] = [round(a / 1e3, 3), round(z / 1e3, 3)]
        # higher offered rates: the highest rate up to which sign p99 stays <= 50 ms
        out["sweep"] = {}
        for q in sorted(sweep):
            srv.reset_stats()
            m = int(q * sweep_seconds)
            el, b = run(srv, kids, mix(m), [rng.expovariate(q) for _ in range(m)])
            st = srv.stats()
            bad += b
            out["sweep"][str(q)] = {"throughput_rps": round(m / el, 1),
                                    "sign_p50_ms": round(st["sign_p50_us"] / 1e3, 3),
                                    "sign_p99_ms": round(st["sign_p99_us"] / 1e3, 3),
                                    "verify_p99_ms": round(st["verify_p99_us"] / 1e3, 3)}
            if st["sign_p99_us"] > 50e3:
                break  # higher rates only queue longer
        # walk up from the fixed rate; the first rate over 50 ms ends the walk.
        # The speedup below uses the throughput achieved at that rate.
        best = best_rps = None
        for q, p99, rps in sorted([(qps, out["sign_p99_ms"], out["throughput_rps"])] +
                                  [(float(k), v["sign_p99_ms"], v["throughput_rps"])
                                   for k, v in out["sweep"].items()]):
            if p99 > 50.0:
                break
            best, best_rps = q, rps
        out["max_qps_p99_le_50ms"] = best
        out["achieved_rps_p99_le_50ms"] = best_rps
        srv.reset_stats()
        reqs = mix(sat_requests)
        el, bad2 = run(srv, kids, reqs, [0.0] * len(reqs))
        st = srv.stats()
        out.update({"saturation_rps": round(sat_requests / el, 1),
                    "saturation_sign_p99_ms": round(st["sign_p99_us"] / 1e3, 3),
                    "saturation_mean_batch": round(st["mean_batch"], 2)})
        bad += bad2
    with C.Server(hbm_budget_bytes=4 << 30, slots=1, max_batch=1) as one:
        kids = [one.add_key(pid, sk, pk) for pid, sk, pk in keys]
        reqs = mix(single_requests)
        t0 = time.perf_counter()
        for op, k, m, s in reqs:
            t = one.submit_sign(kids[k], m) if op == 0 else one.submit_verify(kids[k], m, s)
            bad += one.wait(t)[0] != 0
        el = time.perf_counter() - t0
        out["per_request_launch_rps"] = round(single_requests / el, 1)
    out["invalid"] = bad
    out["speedup_vs_per_request"] = round(out["saturation_rps"] / out["per_request_launch_rps"], 2)
    if out["achieved_rps_p99_le_50ms"]:
        out["speedup_at_p99_le_50ms"] = round(out["achieved_rps_p99_le_50ms"] /
                                              out["per_request_launch_rps"], 2)
    return out


def mixed_phase(C, n, reps=3, seed=3):
    """One mixed batch of n SHAKE256-128f / -192f signing requests (half
    each, distinct DRBG keys, messages log-uniform 1 B .. 4 KiB) replayed as
    one hipGraph (workloads.mixed_shake_batch); returns signs/s of the whole
    mixed batch."""
    import numpy as np
    from cuspx_amd import workloads as W
    work = W.mixed_shake_batch(LibDrbg(C), n, seed)
    sks = [None] * len(work)
    for name in ("shake256-128f", "shake256-192f"):
        idx = [i for i, it in enumerate(work) if it[0] == name]
        pk, sk = C.batch_keygen(C.param_id(name),
                                np.frombuffer(b"".join(work[i][1] for i in idx), np.uint8))
        for j, i in enumerate(idx):
            sks[i] = bytes(sk[j])
    items = [(C.param_id(name), sk, m, opt) for (name, _, m, opt), sk in zip(work, sks)]
    mb = C.MixedSignBatch(items)
    st = C.Stream()
    # which path the batch takes: the graph is captured from the same calls
    C.profile_reset()
    C.profile_enable(True)
    mb.launch(st, eager=True)
    st.sync()
    fused = "k_sign_family" in C.profile_read()
    C.profile_enable(False)
    mb.launch(st)
    st.sync()
    ts = []
    for _ in range(reps):
        a, z = C.Event(), C.Event()
        a.record(st)
        mb.launch(st)
        z.record(st)
        st.sync()
        ts.append(a.elapsed_ms(z))
    sigs = mb.signatures()
    bad = 0
    for k in range(0, len(items), max(1, len(items) // 16)):
        pid, sk, m, _ = items[k]
        p = C.params(pid)
        pk = np.frombuffer(sk[2 * p.n:], np.uint8)
        bad += int(C.batch_verify(pid, pk, np.frombuffer(sigs[k], np.uint8), [m])[0] != 0)
    ms = sum(ts) / len(ts)
    return {"sets": ["shake256-128f", "shake256-192f"], "requests": len(items),
            "message_bytes": "log-uniform 1..4096", "ms_per_batch": round(ms, 3),
            "signs_per_s": round(len(items) / ms * 1e3, 1), "spot_check_invalid": bad,
            "launch": ("one hipGraph per replay: both sets' prep kernels, one persistent "
                       "family kernel over every set's tiles, then T_k and the WOTS gather "
                       "per set") if fused else
                      ("one hipGraph per replay: each set's pipeline on its own forked "
                       "stream")}


def graph_128s_phase(C, sizes=(1, 64, 1024), reps=3, seed=4):
    """SHA2-128s signing batches of 1 / 64 / 1024 captured once as hipGraphs
    and replayed (workloads.sha2_128s_batches); returns ms per batch and
    signs/s per size."""
    import numpy as np
    from cuspx_amd import workloads as W
    pid = C.param_id("sha2-128s")
    p = C.params(pid)
    out = {}
    for B, (seeds, msgs, opts) in W.sha2_128s_batches(sizes, seed).items():
        pk, sk = C.batch_keygen(pid, np.frombuffer(b"".join(seeds), np.uint8))
        D = C.DeviceBuffer
        d_sk = D.from_array(sk)
        d_msg = D.from_array(np.frombuffer(b"".join(msgs), np.uint8))
        d_opt = D.from_array(np.frombuffer(b"".join(opts), np.uint8))
        d_moff = D.from_array(np.arange(B, dtype=np.uint64) * 32)
        d_mlen = D.from_array(np.full(B, 32, np.uint32))
        d_sig = D(B * p.sig_bytes)
        ws = D(C.workspace_bytes(pid, B, C.OP_SIGN))
        g = C.graph_sign(pid, B, d_sk.ptr, p.sk_bytes, d_msg.ptr, d_moff.ptr, d_mlen.ptr,
                         d_opt.ptr, d_sig.ptr, p.sig_bytes, ws.ptr)
        st = C.Stream()
        g.launch(st)
        st.sync()
        ts = []
        for _ in range(reps):
            a, z = C.Event(), C.Event()
            a.record(st)
            g.launch(st)
            z.record(st)
            st.sync()
            ts.append(a.elapsed_ms(z))
        ms = sum(ts) / len(ts)
        out[str(B)] = {"ms_per_batch": round(ms, 3), "signs_per_s": round(B / ms * 1e3, 1),
                       "lanes_per_leaf": C.get_split(pid, C.OP_SIGN, B)}
        g.close()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--batch", type=int, default=4096)
    ap.add_argument("--param", default="sha2-128f")
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--cpu-procs", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--serve-seconds", type=float, default=5.0)
    ap.add_argument("--serve-qps", type=float, default=2500.0)
    ap.add_argument("--serve-saturation", type=int, default=8000)
    ap.add_argument("--serve-single", type=int, default=150)
    ap.add_argument("--mixed", type=int, default=8192)
    ap.add_argument("--graph-128s", type=int, default=1)
    ap.add_argument("--host-reps", type=int, default=3,
                    help="face_mdp_* calls on host buffers (PCIe-inclusive rate; 0 skips)")
    ap.add_argument("--serve-slots", type=int, default=4)
    ap.add_argument("--serve-long-slots", type=int, default=0,
                    help="slots open to messages over 16 KiB at once (0: all)")
    ap.add_argument("--serve-window-us", type=int, default=1000)
    ap.add_argument("--serve-sweep", default="4000,6000,8000,10000,12000,14000,16000,20000",
                    help="offered rates (req/s) tried beyond --serve-qps in ascending order, "
                         "5 s each, until sign p99 passes 50 ms")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: signatures shard by request with no exchange "
              "between GPUs, so one MI355X's throughput is the figure. Use --gpus 1.",
              file=sys.stderr)
        sys.exit(2)
    if args.steps < 1 or args.warmup < 0 or args.batch < 1:
        ap.error("--steps and --batch must be at least 1 and --warmup at least 0")

    import numpy as np
    import cuspx_amd as C

    pid = C.param_id(args.param)
    p = C.params(pid)
    B = args.batch

    cpu = None
    if not args.no_cpu:  # before the GPU is touched: the baseline forks
        # one process per physical core, capped at the 16 CPUs a one-GPU
        # share of the host allows
        max_procs = args.cpu_procs or 16
        print("cpu baseline: up to %d pinned processes, %.0f s" % (max_procs, args.cpu_seconds),
              file=sys.stderr, flush=True)
        cb = cpu_baseline(pid, args.cpu_seconds, max_procs)
        one = cb["single_core"]
        cpu = {"value": round(cb["signs_per_s"], 3), "unit": "signs/s",
               "cores": cb["processes"], "kind": "port",
               "verifies_per_s": round(cb["verifies_per_s"], 2),
               "cpu_model": cb["cpu_model"],
               "physical_cores_visible": cb["physical_cores_visible"],
               "logical_cpus_visible": cb["logical_cpus_visible"],
               "single_core_median_ms": {k: round(v, 3) for k, v in one.items()},
               "signs_per_s_1core": round(1e3 / one["sign_ms_monotonic"], 3),
               "sample": "oracle/spx_oracle.c %s, 32-byte messages: one pinned process, 2 warm-up "
                         "+ 10 timed sign / verify (medians, CLOCK_MONOTONIC and "
                         "CLOCK_PROCESS_CPUTIME_ID); then %d processes pinned one per physical "
                         "core, signing %.0f s then verifying %.0f s each, rates summed"
                         % (args.param, cb["processes"], 0.8 * args.cpu_seconds,
                            0.2 * args.cpu_seconds)}

    info = C.device_info()
    print("device: %s" % info, file=sys.stderr, flush=True)
    stream = C.Stream()
    sh = stream.handle
    print("inputs: %d keys (KAT protocol, DRBG mode)" % B, file=sys.stderr, flush=True)
    kseeds, opts, msgs = make_inputs(C, pid, B)
    D = C.DeviceBuffer
    d_seed, d_opt, d_msg = D.from_array(kseeds), D.from_array(opts), D.from_array(msgs)
    d_moff = D.from_array(np.arange(B, dtype=np.uint64) * 32)
    d_mlen = D.from_array(np.full(B, 32, np.uint32))
    d_pk, d_sk, d_sig = D(B * p.pk_bytes), D(B * p.sk_bytes), D(B * p.sig_bytes)
    d_st = D(4 * B)
    ws = D(max(C.workspace_bytes(pid, B, op) for op in (0, 1, 2)))
    C.keygen_device(pid, B, d_seed.ptr, 3 * p.n, d_pk.ptr, d_sk.ptr, ws.ptr, sh)

    def sign():
        C.sign_device(pid, B, d_sk.ptr, p.sk_bytes, d_msg.ptr, d_moff.ptr, d_mlen.ptr, d_opt.ptr,
                      d_sig.ptr, p.sig_bytes, ws.ptr, sh)

    def verify():
        C.verify_device(pid, B, d_pk.ptr, p.pk_bytes, d_sig.ptr, p.sig_bytes, d_msg.ptr,
                        d_moff.ptr, d_mlen.ptr, d_st.ptr, ws.ptr, sh)

    def all_valid():
        return int(np.count_nonzero(d_st.download(np.int32))) == 0

    print("warmup: %d steps" % args.warmup, file=sys.stderr, flush=True)
    d_st.fill(0xFF)
    for _ in range(args.warmup):
        sign()
        verify()
    stream.sync()
    if args.warmup > 0:  # with no warm-up, the statuses still hold the 0xFF fill
        assert all_valid(), "batched verify rejected a batched signature"

    C.profile_reset()
    C.profile_enable(True)
    ev = [[C.Event() for _ in range(3)] for _ in range(args.steps)]
    stream.sync()
    t0 = time.perf_counter()
    for k in range(args.steps):
        ev[k][0].record(stream)
        sign()
        ev[k][1].record(stream)
        verify()
        ev[k][2].record(stream)
    stream.sync()
    wall = time.perf_counter() - t0
    C.profile_enable(False)
    prof = C.profile_read()
    sign_ms = [e[0].elapsed_ms(e[1]) for e in ev]
    ver_ms = [e[1].elapsed_ms(e[2]) for e in ev]
    assert all_valid(), "batched verify rejected a batched signature"
    ms_step = wall * 1e3 / args.steps
    sign_s = sum(sign_ms) / len(sign_ms) / 1e3
    ver_s = sum(ver_ms) / len(ver_ms) / 1e3

    # the dominant kernel: the fused FORS + hypertree launch, or k_ht alone
    kname = "k_sign_fused" if "k_sign_fused" in prof else "k_ht"
    pmc = None
    try:
        tj = json.load(open(os.path.join(ROOT, "profiles", "pmc_counters.json")))
        pmc = tj.get("%s/%s/%d" % (kname, p.name.decode(), B))
    except (OSError, ValueError):
        pass
    traffic = None
    if pmc:
        traffic = int(1024 * (pmc["fetch_kib"] * pmc["fetch_scale"] +
                              pmc["write_kib"] * pmc["write_scale"]))
    host = None
    if args.host_reps > 0:
        print("host-buffer entry points: %d reps" % args.host_reps, file=sys.stderr, flush=True)
        host = host_phase(C, pid, d_sk.download(np.uint8), d_pk.download(np.uint8), msgs, B,
                          args.host_reps)
    mixed = graph128s = None
    if args.mixed > 0:
        print("mixed SHAKE batch: %d requests" % args.mixed, file=sys.stderr, flush=True)
        mixed = mixed_phase(C, args.mixed)
    if args.graph_128s:
        print("sha2-128s graph replay: batches 1/64/1024", file=sys.stderr, flush=True)
        graph128s = graph_128s_phase(C)
    serve = None
    if args.serve_seconds > 0:
        print("serving: %.0f s at %.0f req/s" % (args.serve_seconds, args.serve_qps),
              file=sys.stderr, flush=True)
        serve = serve_phase(C, args.serve_seconds, args.serve_qps, args.serve_saturation,
                            args.serve_single, args.serve_slots, args.serve_window_us,
                            [float(x) for x in args.serve_sweep.split(",") if x],
                            long_slots=args.serve_long_slots)
    kt, kn = prof.get(kname, (0.0, 0))
    k_avg = kt / max(kn, 1) / 1e3
    work = kernel_work(p, B) + (fors_work(p, B) if kname == "k_sign_fused" else 0)
    # lane-instructions per unit: the launch's SQ_INSTS_VALU x 64 over its
    # units, as counted for this kernel and batch (profiles/pmc_counters.json)
    if pmc and pmc.get("sq_insts_valu"):
        ops_unit = round(pmc["sq_insts_valu"] * 64 / work)
        ops_source = "SQ_INSTS_VALU x 64 / units, profiles/pmc_counters.json"
    else:
        ops_unit = SHA256_OPS if p.hash == 0 else KECCAK_OPS
        ops_source = "constant"
    units_s = work / k_avg if k_avg > 0 else 0.0
    achieved = units_s * ops_unit / 1e12
    ceiling = mix_ceiling_units(MIX_RATE_SHA256 if p.hash == 0 else MIX_RATE_KECCAK, ops_unit)
    out = {
        "metric": "SHA2-128f signs/sec + verifies/sec on 1 GPU; p50/p99 sign latency at fixed QPS",
        "value": round(B / sign_s, 1),
        "unit": "signs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32",
        "data": "synthetic",
        "config": {"workload": "SPHINCS+-%s-simple batched sign then batched verify, %d distinct "
                               "DRBG keypairs (KAT counts 0..%d), one 32-byte message each"
                               % (p.name.decode(), B, B - 1),
                   "param": p.name.decode(), "batch": B},
        "verifies_per_s": round(B / ver_s, 1),
        "sign_p50_ms": serve["sign_p50_ms"] if serve else None,
        "sign_p99_ms": serve["sign_p99_ms"] if serve else None,
        "serve_qps": serve["qps_offered"] if serve else None,
        "serve_max_qps_p99_le_50ms": serve.get("max_qps_p99_le_50ms") if serve else None,
        "sign_batch_ms": round(sign_s * 1e3, 3),
        "verify_batch_ms": round(ver_s * 1e3, 3),
        "kernels_ms_per_launch": {k: round(v[0] / max(v[1], 1), 4) for k, v in prof.items()},
        "roofline": {"bound": "valu", "kernel": kname, "achieved": round(achieved, 3),
                     "peak": round(PEAK_VALU_T, 2), "unit": "T VALU lane-instr/s",
                     "frac": round(achieved / PEAK_VALU_T, 4), "traffic": traffic,
                     "peak_vop3": round(PEAK_VOP3_T, 2),
                     "frac_of_vop3": round(achieved / PEAK_VOP3_T, 4),
                     "mix_ceiling_units_per_s": round(ceiling, 1),
                     "frac_of_mix_ceiling": round(units_s / ceiling, 4),
                     "work_per_launch": work,
                     "work_unit": "sha256 compressions" if p.hash == 0 else "keccak-f permutations",
                     "ops_per_unit": ops_unit,
                     "ops_per_unit_source": ops_source,
                     "kernel_ms": round(k_avg * 1e3, 3),
                     "units_per_s": round(units_s, 1),
                     "frac_at_1450_ops": round(units_s / (PEAK_VOP3_T * 1e12 / SHA256_OPS_UNFUSED), 4)
                     if p.hash == 0 else None},
        "cpu_baseline": cpu,
        "host_buffers": host,
        "serve": serve,
        "mixed_shake_batch": mixed,
        "sha2_128s_graph": graph128s,
        "device": info,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
