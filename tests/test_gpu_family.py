"""Mixed batches of the fast sets of one hash family in ONE persistent
kernel (k_sign_family): every signature equals the CPU oracle's, launched
directly and as a graph replayed twice; mixes outside one family keep the
per-set pipelines and are checked the same way.

The fused path applies when every set's batch runs at one or two lanes
per WOTS leaf; small test batches force such a split (cuspx_set_split) so
the fused kernel itself is what runs, with batches whose last workgroups are partly
empty and with empty sets."""
import random

import numpy as np
import pytest

import cuspx_amd as C
from oracle import oracle as O

pytestmark = pytest.mark.gpu

FAMILIES = {"sha2": ["sha2-128f", "sha2-192f", "sha2-256f"],
            "shake256": ["shake256-128f", "shake256-192f", "shake256-256f"]}


def _items(names, counts, seed):
    rng = random.Random(seed)
    items = []
    for name, cnt in zip(names, counts):
        pid = C.param_id(name)
        p = O.params(pid)
        seeds = [rng.randbytes(3 * p.n) for _ in range(cnt)]
        if not cnt:
            continue
        pk, sk = C.batch_keygen(pid, np.frombuffer(b"".join(seeds), np.uint8))
        for i in range(cnt):
            m = rng.randbytes(rng.choice((0, 1, 33, 136, 300, 2000)))
            items.append((pid, bytes(sk[i]), m, rng.randbytes(p.n)))
    rng.shuffle(items)
    return items


def _check(items, sigs):
    for i, (pid, sk, m, opt) in enumerate(items):
        assert sigs[i] == O.signature_with_optrand(pid, m, sk, opt), (C.params(pid).name, i)


@pytest.mark.parametrize("family", sorted(FAMILIES))
@pytest.mark.parametrize("counts", [(3, 2, 2), (0, 3, 1), (5, 0, 0)])
def test_family_kernel_matches_oracle(family, counts):
    names = FAMILIES[family]
    pids = [C.param_id(n) for n in names]
    try:
        for pid in pids:
            C.set_split(pid, C.OP_SIGN, 1)  # one lane per WOTS leaf: the fused path
        items = _items(names, counts, 7 + sum(counts))
        print("%s family batch %s" % (family, counts), flush=True)
        mb = C.MixedSignBatch(items)
        st = C.Stream()
        mb.launch(st, eager=True)
        st.sync()
        eager = mb.signatures()
        _check(items, eager)
        for _ in range(2):  # graph replays give the same bytes
            for d in mb.bufs:
                d["sig"].fill(0)
            mb.launch(st)
            st.sync()
            assert mb.signatures() == eager
    finally:
        for pid in pids:
            C.set_split(pid, C.OP_SIGN, 0)


def test_family_kernel_taken_up_to_two_lanes_per_leaf():
    """The split rule picks the path: at two lanes per WOTS leaf the batch
    still runs as k_sign_family, at four it runs the per-set pipelines; the
    bytes are the oracle's either way."""
    names = FAMILIES["shake256"]
    pids = [C.param_id(n) for n in names]
    items = _items(names, (3, 2, 0), 41)
    st = C.Stream()
    try:
        for split, fused in ((2, True), (4, False)):
            for pid in pids:
                C.set_split(pid, C.OP_SIGN, split)
            print("shake256 family batch at %d lanes per leaf" % split, flush=True)
            mb = C.MixedSignBatch(items)
            C.profile_reset()
            C.profile_enable(True)
            mb.launch(st, eager=True)
            st.sync()
            prof = C.profile_read()
            C.profile_enable(False)
            assert ("k_sign_family" in prof) == fused, (split, sorted(prof))
            _check(items, mb.signatures())
    finally:
        C.profile_enable(False)
        for pid in pids:
            C.set_split(pid, C.OP_SIGN, 0)


def test_mix_across_families_uses_per_set_pipelines():
    items = _items(["sha2-128f", "shake256-128f", "sha2-128s"], (2, 2, 1), 99)
    mb = C.MixedSignBatch(items)
    st = C.Stream()
    mb.launch(st)
    st.sync()
    graph = mb.signatures()
    _check(items, graph)
    for d in mb.bufs:
        d["sig"].fill(0)
    mb.launch(st, eager=True)
    st.sync()
    assert mb.signatures() == graph
