"""bench.py on the CPU: its argument checks, and the work counts its roofline
divides by, pinned to the hash blocks the oracle counts while it runs the
algorithm (spxo_count_sha256_blocks / spxo_count_keccak_perms).

* Hypertree: one layer's work is exactly a keygen (the top subtree's WOTS
  leaves and nodes, sign.cu:154-182), so kernel_work(p, 1) = d x the
  oracle's keygen count, less the one SHA-256 block that seeds the pub_seed
  midstate per key (sha256.cu:385; SHAKE256 has no midstate).
* FORS: fors_sign (fors.cu:107) also derives each tree's signed secret leaf
  (one PRF per tree) and hashes the k roots into the FORS public key (T_k);
  the signing kernel takes the secret leaf from its own leaf pass and runs
  T_k in k_fors_pk, so fors_work(p, 1) = the oracle's count - k - T_k.
"""
import os
import subprocess
import sys

import pytest

from oracle import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
import bench  # noqa: E402

SETS = ["sha2-128f", "sha2-128s", "sha2-192f", "sha2-256f",
        "shake256-128f", "shake256-192f", "shake256-256f", "shake256-256s"]


def _blocks(p):
    sha, keccak = O.counts()
    return sha if p.hash == 0 else keccak


def _tk_blocks(p):
    """hash blocks of T_k over the k FORS roots (address 22 / 32 bytes)"""
    if p.hash == 0:
        return -(-(22 + p.k * p.n + 9) // 64)
    return (p.n + 32 + p.k * p.n) // 136 + 1


@pytest.mark.parametrize("name", SETS)
def test_work_counts_match_the_oracles_block_counts(name):
    pid = O.param_id(name)
    p = O.params(pid)
    O.reset_counts()
    O.seed_keypair(pid, bytes(range(3 * p.n)))
    keygen = _blocks(p)
    midstate = 1 if p.hash == 0 else 0
    assert bench.kernel_work(p, 1) == p.d * (keygen - midstate)
    assert bench.kernel_work(p, 4096) == 4096 * bench.kernel_work(p, 1)

    md = bytes((7 * i + 3) & 0xFF for i in range((p.k * p.a + 7) // 8))
    O.reset_counts()
    O.fors_sign(pid, md, bytes(range(p.n)), bytes(range(1, p.n + 1)), O.make_addr(pid))
    assert bench.fors_work(p, 1) == _blocks(p) - p.k - _tk_blocks(p)


def test_headline_work_per_launch():
    """SHA2-128f, 4096 requests: the work_per_launch figure of the bench
    line (hypertree + FORS tiles of one k_sign_fused launch)."""
    p = O.params(O.param_id("sha2-128f"))
    assert bench.kernel_work(p, 4096) + bench.fors_work(p, 4096) == 437358592


def _run(*args):
    return subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), *args],
                          capture_output=True, text=True, timeout=120)


def test_more_than_one_gpu_is_refused_with_a_message():
    r = _run("--gpus", "2")
    assert r.returncode == 2 and "one GPU" in r.stderr and r.stdout == ""


@pytest.mark.parametrize("args", [("--steps", "0"), ("--warmup", "-1"), ("--batch", "0")])
def test_empty_runs_are_refused(args):
    r = _run(*args)
    assert r.returncode == 2 and "at least" in r.stderr and r.stdout == ""
