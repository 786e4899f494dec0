// vht_loop.hpp -- hypertree verification as one hash-unit loop per lane
// (one SHA-256 compression or one Keccak-f[1600] permutation per step).
//
// Semantics: crypto_sign_verify, sign.cu:3077-3113 (per layer: WOTS
// pk_from_sig, wots.cu:777-790; T_len; compute_root, utils.cu:94-135), then
// the root compare against pk (sign.cu:3115-3121).
//
// G lanes per request (one group never spans a wave).  Lanes 0..G-2 walk the
// WOTS chains of the current layer, dealt longest-first in snake order;
// lane G-1 (the absorber) hashes T_len block by block as soon as the chains a
// block covers are done -- overlapping T_len with the chain phase -- and then
// climbs the auth path.  Every lane runs the same loop body: pick a job,
// assemble one block and the chaining state for it, run ONE
// SHA-256 compression (SHAKE256: one Keccak-f permutation; the absorber's
// T_len sponge state lives in LDS), fold the result back.  F steps, T_len
// blocks and Merkle nodes therefore share a single hash unit in the code and
// issue together instead of one after the other.  Each group moves to its
// next layer the iteration after its absorber publishes the layer root in
// LDS, independently of the other groups of the wave.  Lanes of one group talk only through LDS
// within one wave, so no workgroup barrier is needed inside the loop.
#pragma once
#include "spx_kernels.hpp"

namespace spx {

template <class P, int G>
__global__ void __launch_bounds__(256) k_vht_loop(BatchArgs A) {
    static_assert(G >= 16 && G <= 64 && (64 % G) == 0, "a group stays inside one wave");
    constexpr bool SHA2 = P::hash == kSha2;
    using Hh = HashT<P>;
    constexpr int NW = P::nw;
    constexpr int LEN = P::len;
    constexpr int TH = P::th;
    constexpr int R = 256 / G;                      // requests per workgroup
    constexpr int GC = G - 1;                       // chain lanes per request
    constexpr int MAXC = (LEN + GC - 1) / GC;       // chains per chain lane
    constexpr int RC = (LEN + G - 1) / G;           // chains per lane when ranking
    constexpr int TLW = LEN * NW;                   // T_len payload words
    // SHA-256: 64-byte blocks of addr22 || payload; SHAKE256: 136-byte blocks
    // of pub_seed || addr32 || payload (the pad always fits the last block)
    constexpr int TLT = NW + 8 + TLW;               // SHAKE: absorbed words
    constexpr int TLB = SHA2 ? (6 + TLW + 2 + 15) / 16 : TLT / 34 + 1;  // T_len blocks
    constexpr int TAIL = TLT - 34 * (TLB - 1);      // SHAKE: words in the last block
    constexpr uint32_t TLBITS = (64u + 22u + 4u * TLW) * 8u;
    constexpr int HB = SHA2 ? (6 + 2 * NW + 2 + 15) / 16 : 1;  // units per Merkle node
    constexpr uint32_t MAX_IT = P::d * (LEN * 15 + TLB + TH * HB + 8) + 16;

    __shared__ uint32_t pkb[R][TLW];     // chain ends of the current layer
    __shared__ uint32_t rootb[R][NW];    // root of the last finished layer
    __shared__ int epoch[R];             // layers finished
    __shared__ uint8_t steps[R][LEN];
    __shared__ uint8_t order[R][LEN];
    __shared__ uint8_t done[R][LEN];
    __shared__ uint32_t hist[R][16];     // chains per step count (ranking)
    __shared__ uint32_t sponge[SHA2 ? 1 : R][50];  // SHAKE: absorber's T_len state

    const int tid = threadIdx.x;
    const int grp = tid / G;
    const int l = tid % G;
    const uint32_t req = blockIdx.x * R + grp;
    const bool active = req < A.n_req;
    const uint32_t rq = active ? req : 0;
    const Plan& pl = A.plan[rq];
    uint32_t mid[8];  // SHA-256: the pub_seed midstate; SHAKE256: pub_seed words
#pragma unroll
    for (int i = 0; i < 8; i++) mid[i] = SHA2 ? pl.mid[i] : pl.pubseed[i];
    const bool absorber = l == G - 1;
    const uint8_t* sig = A.sig + (size_t)rq * A.sig_stride;
    if (l == 0) epoch[grp] = 0;
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");

    bool finished = !active;
    int my_epoch = -1;
    int layer = 0;
    uint64_t tree = pl.tree;
    uint32_t leaf = pl.leaf;
    const uint8_t* lsig = sig;
    // chain lane
    int mine[MAXC];
    uint32_t pre[MAXC][NW];
    int nmine = 0, ci = 0, cj = 0;
    uint32_t s = 0;
    uint32_t v[NW];
    // absorber
    int phase = 2;  // 0 T_len, 1 auth path, 2 idle
    int blk = 0, prefix = 0, h = 0, hb = 0;
    uint32_t tl[8], node[NW], ap[NW];
    uint32_t li = 0;
#pragma unroll
    for (int c = 0; c < MAXC; c++) mine[c] = 0;
#pragma unroll
    for (int w = 0; w < NW; w++) v[w] = node[w] = ap[w] = 0;
#pragma unroll
    for (int i = 0; i < 8; i++) tl[i] = 0;

    // finish chains that need no (more) steps: park the end value
    auto settle = [&]() {
#pragma unroll 1
        while (ci < nmine && s >= (uint32_t)P::w - 1) {
#pragma unroll
            for (int w = 0; w < NW; w++) pkb[grp][cj * NW + w] = v[w];
            done[grp][cj] = 1;
            ci++;
            if (ci < nmine) {
#pragma unroll
                for (int c = 1; c < MAXC; c++)
                    if (c == ci) {
                        cj = mine[c];
#pragma unroll
                        for (int w = 0; w < NW; w++) v[w] = pre[c][w];
                    }
                s = P::w - 1 - steps[grp][cj];
            }
        }
    };

    uint32_t it = 0;
#pragma unroll 1
    while (true) {
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
        // ---- layer change: every lane of the group sees the new epoch at once
        if (!finished && epoch[grp] != my_epoch) {
            const int ep = epoch[grp];
            if (ep > 0) {  // the layer above is signed by the root just found
                leaf = (uint32_t)(tree & ((1u << TH) - 1));
                tree >>= TH;
            }
            my_epoch = ep;
            layer = ep;
            if (layer == P::d) {
                if (l == 0) {
                    bool ok = true;
#pragma unroll
                    for (int w = 0; w < NW; w++) ok = ok && (rootb[grp][w] == pl.root[w]);
                    A.status[rq] = ok ? 0 : -1;
                }
                finished = true;
            } else {
                lsig = sig + P::n + P::fors_bytes + (size_t)layer * P::layer_bytes;
                uint32_t msg[NW];
                if (layer == 0) {
                    load_val<P::hash, NW>(A.fors_pk + (size_t)rq * P::n, msg);
                } else {
#pragma unroll
                    for (int w = 0; w < NW; w++) msg[w] = rootb[grp][w];
                }
                // remaining steps of every chain, ranked longest first by a
                // counting sort over the 16 step counts (ties in any order)
                if (l < 16) hist[grp][l] = 0;
                __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
                uint32_t st_c[RC], slot[RC];
#pragma unroll
                for (int c = 0; c < RC; c++) {
                    const int j = l + c * G;
                    st_c[c] = j < LEN ? P::w - 1 - chain_len<P>(msg, j) : 0u;
                    slot[c] = 0;
                    if (j < LEN) {
                        steps[grp][j] = (uint8_t)st_c[c];
                        done[grp][j] = 0;
                        slot[c] = atomicAdd(&hist[grp][15 - st_c[c]], 1u);
                    }
                }
                __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
#pragma unroll
                for (int c = 0; c < RC; c++) {
                    const int j = l + c * G;
                    if (j < LEN) {
                        uint32_t start = 0;
#pragma unroll 1
                        for (uint32_t b = 0; b < 15 - st_c[c]; b++) start += hist[grp][b];
                        order[grp][start + slot[c]] = (uint8_t)j;
                    }
                }
                __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
                // chain lanes take ranks l, 2GC-1-l, 2GC+l, ... (snake order)
                nmine = 0;
                if (!absorber) {
#pragma unroll
                    for (int c = 0; c < MAXC; c++) {
                        const int r = (c & 1) ? (c + 1) * GC - 1 - l : c * GC + l;
                        mine[c] = r < LEN ? order[grp][r] : 0;
                        nmine += r < LEN;
                        load_val<P::hash, NW>(lsig + (size_t)mine[c] * P::n, pre[c]);
                    }
                }
                ci = 0;
                cj = mine[0];
                if (nmine > 0) {
                    s = P::w - 1 - steps[grp][cj];
#pragma unroll
                    for (int w = 0; w < NW; w++) v[w] = pre[0][w];
                }
                settle();
                if (absorber) {
                    phase = 0;
                    blk = 0;
                    prefix = 0;
                    h = 0;
                    hb = 0;
#pragma unroll
                    for (int i = 0; i < 8; i++) tl[i] = mid[i];
                    if constexpr (!SHA2) {
#pragma unroll
                        for (int i = 0; i < 50; i++) sponge[grp][i] = 0;
                    }
                    load_val<P::hash, NW>(lsig + P::wots_bytes, ap);  // auth[0]
                }
            }
        }
        // ---- pick a job: 1 = F step, 2 = T_len block, 3 = Merkle node block
        int job = 0;
        if (!finished) {
            if (absorber) {
                if (phase == 0) {
#pragma unroll 1
                    while (prefix < LEN && done[grp][prefix]) prefix++;
                    // chain holding the last payload word of the block
                    const int need = SHA2 ? (16 * blk + 10) / NW : (34 * blk + 33 - NW - 8) / NW;
                    if (prefix > (need < LEN - 1 ? need : LEN - 1)) job = 2;
                } else if (phase == 1) {
                    job = 3;
                }
            } else if (ci < nmine) {
                job = 1;
            }
        }
        if constexpr (SHA2) {
            // ---- assemble the chaining state and the block
            uint32_t sin[8], w[16];
#pragma unroll
            for (int i = 0; i < 8; i++) sin[i] = (job == 2 || (job == 3 && hb > 0)) ? tl[i] : mid[i];
            if (job == 1) {
                Adr wa = adr_make(layer, tree, kWots);
                adr_keypair<TH>(wa, leaf);
                adr_chain(wa, cj);
                adr_hash(wa, s);
#pragma unroll
                for (int i = 0; i < 16; i++) w[i] = Hh::template block_word<NW>(wa, v, 0, i);
            } else if (job == 2) {
                const uint32_t* pk = pkb[grp];
                if (blk > 0 && 16 * blk + 15 < 5 + TLW) {  // a block of payload words only
                    const int base = 16 * blk - 6;
#pragma unroll
                    for (int i = 0; i < 16; i++)
                        w[i] = __builtin_amdgcn_alignbit(pk[base + i], pk[base + i + 1], 16);
                } else {
                    Adr pa = adr_make(layer, tree, kWotsPk);
                    adr_keypair<TH>(pa, leaf);
#pragma unroll
                    for (int i = 0; i < 16; i++) {
                        const int m = 16 * blk + i;
                        uint32_t x;
                        if (m < 5) x = pa.w[m < 5 ? m : 0];
                        else if (m == 5) x = pa.w[5] | (pk[0] >> 16);
                        else if (m < 5 + TLW) x = __builtin_amdgcn_alignbit(pk[m - 6], pk[m - 5], 16);
                        else if (m == 5 + TLW) x = (pk[TLW - 1] << 16) | 0x8000u;
                        else if (m == 16 * TLB - 1) x = TLBITS;
                        else x = 0;
                        w[i] = x;
                    }
                }
            } else if (job == 3) {
                Adr ha = adr_make(layer, tree, kHashTree);
                adr_node(ha, h + 1, li >> 1);
                uint32_t d[2 * NW];
                const bool odd = li & 1u;
#pragma unroll
                for (int k = 0; k < NW; k++) {
                    d[k] = odd ? ap[k] : node[k];
                    d[NW + k] = odd ? node[k] : ap[k];
                }
                if (HB == 1 || hb == 0) {
#pragma unroll
                    for (int i = 0; i < 16; i++) w[i] = Hh::template block_word<2 * NW>(ha, d, 0, i);
                } else {
#pragma unroll
                    for (int i = 0; i < 16; i++) w[i] = Hh::template block_word<2 * NW>(ha, d, 1, i);
                }
            }
            // ---- the one compression of this iteration
            uint32_t out[8];
#pragma unroll
            for (int i = 0; i < 8; i++) out[i] = sin[i];
            if (job != 0) sha256_compress(out, w);  // out = sin + rounds(sin, w)
            // ---- fold the result back
            if (job == 1) {
#pragma unroll
                for (int k = 0; k < NW; k++) v[k] = out[k];
                s++;
                settle();
            } else if (job == 2) {
#pragma unroll
                for (int i = 0; i < 8; i++) tl[i] = out[i];
                blk++;
                if (blk == TLB) {  // WOTS public key compressed: the leaf
#pragma unroll
                    for (int k = 0; k < NW; k++) node[k] = tl[k];
                    phase = 1;
                    li = leaf;
                }
            } else if (job == 3) {
                if (hb + 1 < HB) {
#pragma unroll
                    for (int i = 0; i < 8; i++) tl[i] = out[i];
                    hb++;
                } else {
#pragma unroll
                    for (int k = 0; k < NW; k++) node[k] = out[k];
                    hb = 0;
                    li >>= 1;
                    h++;
                    if (h == TH) {  // layer root: publish, the group moves on next iteration
#pragma unroll
                        for (int k = 0; k < NW; k++) rootb[grp][k] = node[k];
                        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
                        epoch[grp] = layer + 1;
                        phase = 2;
                    } else {
                        load_val<P::hash, NW>(lsig + P::wots_bytes + (size_t)h * P::n, ap);
                    }
                }
            }
        } else {
            // ---- assemble the sponge state of this step (Keccak-f words, LE)
            uint32_t st[50];
            if (job == 1) {
                Adr wa = adr_make(layer, tree, kWots);
                adr_keypair<TH>(wa, leaf);
                adr_chain(wa, cj);
                adr_hash(wa, s);
#pragma unroll
                for (int i = 0; i < 50; i++) {
                    uint32_t x = 0;
                    if (i < NW) x = mid[i];
                    else if (i < NW + 8) x = Hh::aw(wa, i - NW);
                    else if (i < 2 * NW + 8) x = v[i - NW - 8];
                    if (i == 2 * NW + 8) x ^= 0x1Fu;
                    if (i == 33) x ^= 0x80000000u;
                    st[i] = x;
                }
            } else if (job == 2) {
                Adr pa = adr_make(layer, tree, kWotsPk);
                adr_keypair<TH>(pa, leaf);
                const uint32_t* pk = pkb[grp];
#pragma unroll
                for (int i = 0; i < 50; i++) {
                    uint32_t x = sponge[grp][i];
                    if (i < 34) {
                        const int k = 34 * blk + i;  // absorbed word index
                        uint32_t m = 0;
                        if (k < NW) m = mid[k < NW ? k : 0];
                        else if (k < NW + 8) m = Hh::aw(pa, k - NW < 8 ? k - NW : 0);
                        else if (k < TLT) m = pk[k - NW - 8];
                        if (blk == TLB - 1 && i == TAIL) m ^= 0x1Fu;
                        if (blk == TLB - 1 && i == 33) m ^= 0x80000000u;
                        x ^= m;
                    }
                    st[i] = x;
                }
            } else if (job == 3) {
                Adr ha = adr_make(layer, tree, kHashTree);
                adr_node(ha, h + 1, li >> 1);
                const bool odd = li & 1u;
#pragma unroll
                for (int i = 0; i < 50; i++) {
                    uint32_t x = 0;
                    if (i < NW) x = mid[i];
                    else if (i < NW + 8) x = Hh::aw(ha, i - NW);
                    else if (i < 2 * NW + 8) x = odd ? ap[i - NW - 8] : node[i - NW - 8];
                    else if (i < 3 * NW + 8) x = odd ? node[i - 2 * NW - 8] : ap[i - 2 * NW - 8];
                    if (i == 3 * NW + 8) x ^= 0x1Fu;
                    if (i == 33) x ^= 0x80000000u;
                    st[i] = x;
                }
            }
            // ---- the one permutation of this iteration
            if (job != 0) keccak_f(st);
            // ---- fold the result back
            if (job == 1) {
#pragma unroll
                for (int k = 0; k < NW; k++) v[k] = st[k];
                s++;
                settle();
            } else if (job == 2) {
                blk++;
                if (blk == TLB) {  // WOTS public key squeezed: the leaf
#pragma unroll
                    for (int k = 0; k < NW; k++) node[k] = st[k];
                    phase = 1;
                    li = leaf;
                } else {
#pragma unroll
                    for (int i = 0; i < 50; i++) sponge[grp][i] = st[i];
                }
            } else if (job == 3) {
#pragma unroll
                for (int k = 0; k < NW; k++) node[k] = st[k];
                li >>= 1;
                h++;
                if (h == TH) {  // layer root: publish, the group moves on next iteration
#pragma unroll
                    for (int k = 0; k < NW; k++) rootb[grp][k] = node[k];
                    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
                    epoch[grp] = layer + 1;
                    phase = 2;
                } else {
                    load_val<P::hash, NW>(lsig + P::wots_bytes + (size_t)h * P::n, ap);
                }
            }
        }
        if (__ballot(!finished) == 0) break;
        if (++it > MAX_IT) {  // unreachable when every job progresses; bounds the loop
            if (active && l == 0 && !finished) A.status[rq] = -1;
            break;
        }
    }
}

}  // namespace spx
