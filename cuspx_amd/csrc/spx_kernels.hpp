// spx_kernels.hpp -- batched SPHINCS+ sign / verify / keygen kernels.
//
// One signature is split across many lanes (MI355X needs >> 256 workgroups
// and prefers uniform control flow inside a wave):
//
//   sign:   k_prep       lane pair per request: midstate, PRF_msg, H_msg, plan
//           k_fors       L lanes per FORS tree (2^J leaves per lane), LDS reduce
//           k_fors_pk    lane pair per request: T_k over the k roots (small
//                        batches: run by the lane storing the last root)
//           k_ht         lane per WOTS leaf of every hypertree layer (all d
//                        layers at once: every layer's tree/leaf comes from
//                        H_msg alone), LDS reduce, auth path, layer root;
//                        the signing leaf's chain values are captured
//           k_wots_gather lane per (request, layer, chain): picks the
//                        captured chain value for the layer's message
//   verify: k_vprep, k_vfors, k_fors_pk, k_vht (the chains of a workgroup's
//           requests dealt over its lanes, layers in sequence); small
//           batches: k_vspec (every chain from every start) + k_vht_spec
//           (T_len and the climb per layer)
//   keygen: k_kprep, k_ht in keygen mode (top layer only)
//
// Semantics follow sign.cu:1924-1994 (sign), sign.cu:3048-3121 (verify),
// sign.cu:154-182 (keygen), utils.cu:94-241 (compute_root / treehash),
// wots.cu:27-790 and fors.cu:24-633 of the reference.
#pragma once
#include <type_traits>

#include "sha2_dev.hpp"
#include "shake_dev.hpp"

namespace spx {

template <class P>
using HashT = typename std::conditional<P::hash == kSha2, Sha2<P::n>, Shake<P::n>>::type;

// ------------------------------------------------------------- arguments

struct BatchArgs {
    // inputs
    const uint8_t* sk;        // sign: secret keys, stride sk_stride
    uint32_t sk_stride;
    const uint8_t* pk;        // verify: public keys, stride pk_stride
    uint32_t pk_stride;
    const uint8_t* seed;      // keygen: 3n-byte seeds, stride seed_stride
    uint32_t seed_stride;
    const uint8_t* msg;       // messages: msg + moff[i], mlen[i] bytes, 4-byte aligned
    const uint64_t* moff;
    const uint32_t* mlen;
    const uint8_t* optrand;   // sign: [N][n]
    const uint32_t* key_idx;  // optional: request i uses key row key_idx[i] of sk / pk
    const uint32_t* key_mid;  // optional: precomputed SHA-256 midstate per key row (8 words)
    // outputs
    uint8_t* sig;             // signatures (sign: written, verify: read)
    uint64_t sig_stride;
    int32_t* status;          // verify: 0 ok / -1 fail per request
    uint8_t* pk_out;          // keygen outputs
    uint8_t* sk_out;
    // workspace
    Plan* plan;
    uint8_t* fors_roots;      // [N][k][n]
    uint8_t* fors_pk;         // [N][n]
    uint8_t* roots;           // [N][d][n]
    uint8_t* cap;             // [N][d][len][16][n]
    uint32_t* tl_buf;         // SHAKE T_len parking: [len*nw][lanes]
    uint32_t tl_lanes;
    uint32_t* ht_parts;       // split subtrees: part roots [tree][part][nw]
    uint32_t* ht_count;       // split subtrees: parts finished per tree (zeroed by the prep kernel)
    uint32_t* tile_ctr;       // fused signing: next tile (zeroed by the prep kernel)
    uint32_t* fors_count;     // FORS roots finished per request (zeroed by the prep kernels)
    uint32_t n_req;
    // prep kernels: workgroups of the lane-pair part; messages of long_min
    // bytes or more (0: none) are hashed by the two-wave part after them
    uint32_t prep_wgs;
    uint32_t long_min;
};

// ----------------------------------------------------- message / indices

template <class P>
__device__ __forceinline__ void plan_from_digest(Plan& pl, const uint8_t* dg) {
    // message_to_indices (fors.cu:77-91): LSB-first bits within each byte
#pragma unroll
    for (int t = 0; t < P::k; t++) {
        uint32_t v = 0;
#pragma unroll
        for (int j = 0; j < P::a; j++) {
            const int off = t * P::a + j;
            v |= (uint32_t)((dg[off >> 3] >> (off & 7)) & 1) << j;
        }
        pl.fors_idx[t] = v;
    }
    // tree / leaf (hash_sha256.cu:249-261): big-endian, masked
    uint64_t tree = 0;
#pragma unroll
    for (int i = 0; i < P::tree_bytes; i++)
        tree = (tree << 8) | dg[P::fors_msg_bytes + i];
    if (P::tree_bits < 64) tree &= (~0ull) >> (64 - P::tree_bits);
    uint32_t leaf = 0;
#pragma unroll
    for (int i = 0; i < P::leaf_bytes; i++)
        leaf = (leaf << 8) | dg[P::fors_msg_bytes + P::tree_bytes + i];
    leaf &= (~0u) >> (32 - P::leaf_bits);
    pl.tree = tree;
    pl.leaf = leaf;
}

// The prep kernels (128-lane workgroups).  Workgroups [0, prep_wgs): every
// request hashes its message on a lane pair (sha256_compress_dup /
// Shake::stream_msg_pair; both lanes compute everything, the even lane
// stores), 64 requests per workgroup.  SHA-256 messages of at least
// long_min bytes are left to the workgroups after those: each takes long
// requests in turn and runs one with both waves (Sha2::stream_msg_long), the
// schedules of its message-only blocks computed by wave 1 ahead of wave 0's
// rounds.
constexpr uint32_t kPrepLanes = 2;
constexpr uint32_t kPrepWg = 128;
// the two-wave path from this length: a single SHA2-128f / 256f request of
// 512 B signs in 0.329 / 0.624 ms against 0.336 / 0.634 on the lane pair,
// 2 KiB in 0.396 / 0.686 against 0.432 / 0.723, 64 KiB in 3.28 / 3.56
// against 4.60 / 5.07 (below 512 B the two are equal)
constexpr uint32_t kLongMsgBytes = 512;
constexpr uint32_t kLongWgs = 512;         // at most this many two-wave workgroups

template <class P>
__host__ __device__ constexpr bool prep_has_long() { return P::hash == kSha2; }

__device__ __forceinline__ bool prep_is_long(const BatchArgs& A, uint32_t i) {
    return A.long_min && A.mlen[i] >= A.long_min;
}

template <class P>
__device__ __forceinline__ void prep_plan_init(const BatchArgs& A, uint32_t row,
                                               const uint32_t pub[P::nw], Plan& pl) {
    if (A.key_mid) {
#pragma unroll
        for (int j = 0; j < 8; j++) pl.mid[j] = A.key_mid[(size_t)row * 8 + j];
    } else {
        HashT<P>::template midstate<true>(pub, pl.mid);
    }
}

// sign, one request: midstate, R = PRF_msg, digest = H_msg, plan; LONG: the
// two-wave form over kw_s (prod: wave 1), else a lane pair
template <class P, bool LONG>
__device__ __forceinline__ void prep_sign_one(const BatchArgs& A, uint32_t i, bool store,
                                              lds_u32* kw_s, bool prod) {
    using Hh = HashT<P>;
    constexpr int NW = P::nw;
    const uint32_t row = A.key_idx ? A.key_idx[i] : i;
    const uint8_t* sk = A.sk + (size_t)row * A.sk_stride;
    uint32_t skseed[NW], skprf[NW], pub[NW], pk[2 * NW], opt[NW], R[NW];
    load_val<P::hash, NW>(sk, skseed);
    load_val<P::hash, NW>(sk + P::n, skprf);
    load_val<P::hash, NW>(sk + 2 * P::n, pub);
    load_val<P::hash, 2 * NW>(sk + 2 * P::n, pk);
    load_val<P::hash, NW>(A.optrand + (size_t)i * P::n, opt);
    const uint8_t* m = A.msg + A.moff[i];
    const uint32_t ml = A.mlen[i];
    Plan pl;
    prep_plan_init<P>(A, row, pub, pl);
#pragma unroll
    for (int j = 0; j < 8; j++) {
        pl.skseed[j] = j < NW ? skseed[j] : 0u;
        pl.pubseed[j] = j < NW ? pub[j] : 0u;
        pl.root[j] = 0;
    }
    uint8_t dg[P::dgst_bytes];
    if constexpr (LONG) {
        Hh::template prf_msg<true>(skprf, opt, m, ml, R, kw_s, prod);
        Hh::template h_msg<P::dgst_bytes, true>(R, pk, m, ml, dg, kw_s, prod);
    } else {
        (void)kw_s;
        (void)prod;
        Hh::prf_msg(skprf, opt, m, ml, R);
        Hh::template h_msg<P::dgst_bytes>(R, pk, m, ml, dg);
    }
    if (!store) return;
    uint8_t* sig = A.sig + (size_t)i * A.sig_stride;
    store_val<P::hash, NW>(sig, R);
    plan_from_digest<P>(pl, dg);
    pl.status = 0;
    A.plan[i] = pl;
}

// sign: midstate, R = PRF_msg, digest = H_msg, plan
template <class P>
__global__ void __launch_bounds__(kPrepWg) k_prep(BatchArgs A) {
    // one serial hash chain per request over the whole message: ask the SIMD
    // arbiter to favour these waves over concurrent bulk hypertree waves
    __builtin_amdgcn_s_setprio(3);
    if (blockIdx.x >= A.prep_wgs) {
        if constexpr (prep_has_long<P>()) {
            __shared__ __align__(16) uint32_t kw_s[HashT<P>::kLongLdsWords];
            const bool prod = threadIdx.x >= 64;
#pragma unroll 1
            for (uint32_t i = blockIdx.x - A.prep_wgs; i < A.n_req; i += gridDim.x - A.prep_wgs)
                if (prep_is_long(A, i))
                    prep_sign_one<P, true>(A, i, threadIdx.x == 0, (lds_u32*)kw_s, prod);
        }
        return;
    }
    const uint32_t i = (blockIdx.x * kPrepWg + threadIdx.x) / kPrepLanes;
    if (i >= A.n_req) return;
    if (A.ht_count) {
#pragma unroll 1
        for (int layer = 0; layer < P::d; layer++)
            if (threadIdx.x % kPrepLanes == 0) A.ht_count[(size_t)i * P::d + layer] = 0;
    }
    if (A.tile_ctr && i == 0 && threadIdx.x == 0) *A.tile_ctr = 0;
    if (A.fors_count && threadIdx.x % kPrepLanes == 0) A.fors_count[i] = 0;
    if (prep_is_long(A, i)) return;
    prep_sign_one<P, false>(A, i, (threadIdx.x & 1u) == 0, nullptr, false);
}

// verify, one request: midstate, digest = H_msg(R from the signature), plan
// (LONG as for prep_sign_one)
template <class P, bool LONG>
__device__ __forceinline__ void prep_verify_one(const BatchArgs& A, uint32_t i, bool store,
                                                lds_u32* kw_s, bool prod) {
    using Hh = HashT<P>;
    constexpr int NW = P::nw;
    const uint32_t row = A.key_idx ? A.key_idx[i] : i;
    const uint8_t* pkp = A.pk + (size_t)row * A.pk_stride;
    uint32_t pub[NW], pk[2 * NW], R[NW];
    load_val<P::hash, NW>(pkp, pub);
    load_val<P::hash, 2 * NW>(pkp, pk);
    const uint8_t* sig = A.sig + (size_t)i * A.sig_stride;
    load_val<P::hash, NW>(sig, R);
    Plan pl;
    prep_plan_init<P>(A, row, pub, pl);
#pragma unroll
    for (int j = 0; j < 8; j++) {
        pl.skseed[j] = 0;
        pl.pubseed[j] = j < NW ? pub[j] : 0u;
        pl.root[j] = j < NW ? pk[NW + j] : 0u;
    }
    uint8_t dg[P::dgst_bytes];
    if constexpr (LONG) {
        Hh::template h_msg<P::dgst_bytes, true>(R, pk, A.msg + A.moff[i], A.mlen[i], dg, kw_s,
                                                prod);
    } else {
        (void)kw_s;
        (void)prod;
        Hh::template h_msg<P::dgst_bytes>(R, pk, A.msg + A.moff[i], A.mlen[i], dg);
    }
    if (!store) return;
    plan_from_digest<P>(pl, dg);
    pl.status = 0;
    A.plan[i] = pl;
}

// verify: midstate, digest = H_msg(R from the signature), plan
template <class P>
__global__ void __launch_bounds__(kPrepWg) k_vprep(BatchArgs A) {
    __builtin_amdgcn_s_setprio(3);  // serial message hash per request, as in k_prep
    if (blockIdx.x >= A.prep_wgs) {
        if constexpr (prep_has_long<P>()) {
            __shared__ __align__(16) uint32_t kw_s[HashT<P>::kLongLdsWords];
            const bool prod = threadIdx.x >= 64;
#pragma unroll 1
            for (uint32_t i = blockIdx.x - A.prep_wgs; i < A.n_req; i += gridDim.x - A.prep_wgs)
                if (prep_is_long(A, i))
                    prep_verify_one<P, true>(A, i, threadIdx.x == 0, (lds_u32*)kw_s, prod);
        }
        return;
    }
    const uint32_t i = (blockIdx.x * kPrepWg + threadIdx.x) / kPrepLanes;
    if (i >= A.n_req || prep_is_long(A, i)) return;
    prep_verify_one<P, false>(A, i, (threadIdx.x & 1u) == 0, nullptr, false);
}

// grid of the prep kernels for n requests whose messages are at most
// mlen_bound bytes (0: not known); sets A.prep_wgs / A.long_min
template <class P>
inline uint32_t prep_grid(BatchArgs& A, uint32_t n, uint32_t mlen_bound) {
    A.prep_wgs = (uint32_t)(((uint64_t)n * kPrepLanes + kPrepWg - 1) / kPrepWg);
    const bool lng = prep_has_long<P>() && (mlen_bound == 0 || mlen_bound >= kLongMsgBytes);
    A.long_min = lng ? kLongMsgBytes : 0;
    return A.prep_wgs + (lng ? (n < kLongWgs ? n : kLongWgs) : 0);
}

// resident-key midstates: mid[i] = state after pub_seed || 0^(64-n) for the
// public key row i (seed_state, sha256.cu:391-405); zero for SHAKE sets
template <class P>
__global__ void __launch_bounds__(64) k_midstate(const uint8_t* pk, uint32_t pk_stride,
                                                 uint32_t* mid, uint32_t n) {
    using Hh = HashT<P>;
    constexpr int NW = P::nw;
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint32_t pub[NW], m[8];
    load_val<P::hash, NW>(pk + (size_t)i * pk_stride, pub);
    Hh::midstate(pub, m);
#pragma unroll
    for (int j = 0; j < 8; j++) mid[(size_t)i * 8 + j] = m[j];
}

// keygen: sk = seed (3n) || root, plan with the key material
template <class P>
__global__ void __launch_bounds__(64) k_kprep(BatchArgs A) {
    using Hh = HashT<P>;
    constexpr int NW = P::nw;
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= A.n_req) return;
    const uint8_t* s = A.seed + (size_t)i * A.seed_stride;
    uint32_t skseed[NW], pub[NW];
    load_val<P::hash, NW>(s, skseed);
    load_val<P::hash, NW>(s + 2 * P::n, pub);
    Plan pl;
    Hh::midstate(pub, pl.mid);
#pragma unroll
    for (int j = 0; j < 8; j++) {
        pl.skseed[j] = j < NW ? skseed[j] : 0u;
        pl.pubseed[j] = j < NW ? pub[j] : 0u;
        pl.root[j] = 0;
    }
    pl.tree = 0;
    pl.leaf = 0;
    pl.status = 0;
    A.plan[i] = pl;
    if (A.ht_count) A.ht_count[i] = 0;
    // sk = SK_SEED || SK_PRF || PUB_SEED (|| root, written by k_ht); pk = PUB_SEED || root
    uint8_t* sko = A.sk_out + (size_t)i * P::sk_bytes;
    uint8_t* pko = A.pk_out + (size_t)i * P::pk_bytes;
    for (int j = 0; j < 3 * P::n; j++) sko[j] = s[j];
    for (int j = 0; j < P::n; j++) pko[j] = s[2 * P::n + j];
}

// ------------------------------------------------------------------ FORS

template <class P, bool PAIR>
__device__ __forceinline__ void fors_pk_of(const BatchArgs& A, uint32_t i);

// a FORS root of request req is stored.  Small sign batches (fors_count
// set): count it, and the lane that stores the request's last root computes
// the FORS public key, so T_k overlaps the hypertree tiles instead of taking
// a launch of its own (the k roots of one request come from several
// workgroups; every counter was zeroed by the prep kernel).  Large batches
// launch k_fors_pk instead: there the per-tile release fences (an L2
// write-back each) cost more than the launch.
template <class P>
__device__ __forceinline__ void fors_root_done(const BatchArgs& A, uint32_t req) {
    if (!A.fors_count) return;
    __threadfence();  // release this root before counting it
    if (atomicAdd(&A.fors_count[req], 1u) == (uint32_t)P::k - 1) {
        __threadfence();  // acquire: every root of the request is visible
        fors_pk_of<P, false>(A, req);
    }
}

// LDS words a FORS tile of set P uses (two levels of 256 nodes)
template <class P>
constexpr int fors_lds_words() {
    return 2 * 256 * P::nw;
}

// sign: FORS trees, L = 2^(a-J) lanes per tree, 2^J leaves per lane; `blk`
// is the workgroup's tile (TPW trees), so the same body serves k_fors and the
// fused persistent signing kernels.  smem: fors_lds_words<P>() words of LDS
// owned by the kernel (shared with the hypertree tiles of a persistent
// kernel, which never run at the same time in one workgroup)
template <class P, int J>
__device__ __forceinline__ void fors_tile(const BatchArgs& A, uint32_t blk, uint32_t* smem) {
    using Hh = HashT<P>;
    using Key = typename Hh::Key;
    constexpr int NW = P::nw;
    constexpr int LOGL = P::a - J;
    constexpr int L = 1 << LOGL;
    constexpr int WG = 256;
    static_assert(L <= WG, "tree wider than a workgroup");
    constexpr int TPW = WG / L;
    auto& lds = *reinterpret_cast<uint32_t (*)[2][WG][NW]>(smem);

    const int tid = threadIdx.x;
    const uint32_t g = blk * TPW + tid / L;  // global tree id
    const uint32_t u = tid % L;
    const bool active = g < A.n_req * P::k;
    const uint32_t req = active ? g / P::k : 0;
    const uint32_t t = active ? g % P::k : 0;
    const Plan& pl = A.plan[req];
    Key key;
    Hh::load_key(pl, key);
    const uint32_t target = pl.fors_idx[t];
    Adr adr = adr_make(0, pl.tree, kForsTree);
    adr_keypair<P::th>(adr, pl.leaf);
    uint8_t* rec = A.sig + (size_t)req * A.sig_stride + P::n + (size_t)t * (P::a + 1) * P::n;

    uint32_t node[NW];
    uint32_t stk[J > 0 ? J : 1][NW];
    const uint32_t base = (t << P::a) + (u << J);  // global leaf index of this lane's first leaf
#pragma unroll 1
    for (uint32_t i = 0; i < (1u << J); i++) {
        const uint32_t li = (u << J) + i;  // leaf index within the tree
        uint32_t sk[NW];
        adr_node(adr, 0, base + i);
        Hh::PRF(key, adr, sk);
        Hh::F(key, adr, sk, node);
        if (active && li == target) store_val<P::hash, NW>(rec, sk);
        if (active && li == (target ^ 1u)) store_val<P::hash, NW>(rec + P::n, node);
        // merge with pending left siblings (i is uniform across the wave)
        uint32_t h = 0;
#pragma unroll 1
        while (h < (uint32_t)J && ((i >> h) & 1u)) {
            uint32_t left[NW];
#pragma unroll
            for (int w = 0; w < NW; w++) {
                uint32_t x = stk[0][w];
#pragma unroll
                for (int s = 1; s < J; s++) x = (h == (uint32_t)s) ? stk[s][w] : x;
                left[w] = x;
            }
            h++;
            const uint32_t idx = li >> h;  // node index at height h within the tree
            adr_node(adr, h, ((t << P::a) >> h) + idx);
            Hh::H(key, adr, left, node, node);
            if (active && idx == ((target >> h) ^ 1u))
                store_val<P::hash, NW>(rec + P::n + (size_t)h * P::n, node);
        }
        if (h < (uint32_t)J) {
#pragma unroll
            for (int s = 0; s < J; s++)
#pragma unroll
                for (int w = 0; w < NW; w++) stk[s][w] = (h == (uint32_t)s) ? node[w] : stk[s][w];
        }
    }
    // cross-lane reduction: level l holds WG >> l nodes, node x = H(2x, 2x+1)
#pragma unroll
    for (int w = 0; w < NW; w++) lds[0][tid][w] = node[w];
#pragma unroll 1
    for (int l = 1; l <= LOGL; l++) {
        __syncthreads();
        const int src = (l - 1) & 1, dst = l & 1;
        if (tid < (WG >> l)) {
            const uint32_t tau = tid >> (LOGL - l);
            const uint32_t x = tid & ((L >> l) - 1);
            const uint32_t g2 = blk * TPW + tau;
            const bool act2 = g2 < A.n_req * P::k;
            const uint32_t req2 = act2 ? g2 / P::k : 0, t2 = act2 ? g2 % P::k : 0;
            const Plan& p2 = A.plan[req2];
            Key k2;
            Hh::load_key(p2, k2);
            uint32_t lft[NW], rgt[NW], out[NW];
#pragma unroll
            for (int w = 0; w < NW; w++) {
                lft[w] = lds[src][2 * tid][w];
                rgt[w] = lds[src][2 * tid + 1][w];
            }
            const uint32_t hgt = J + l;
            Adr a2 = adr_make(0, p2.tree, kForsTree);
            adr_keypair<P::th>(a2, p2.leaf);
            adr_node(a2, hgt, ((t2 << P::a) >> hgt) + x);
            Hh::H(k2, a2, lft, rgt, out);
#pragma unroll
            for (int w = 0; w < NW; w++) lds[dst][tid][w] = out[w];
            if (act2) {
                uint8_t* rec2 =
                    A.sig + (size_t)req2 * A.sig_stride + P::n + (size_t)t2 * (P::a + 1) * P::n;
                const uint32_t tg = p2.fors_idx[t2];
                if (hgt < (uint32_t)P::a && x == ((tg >> hgt) ^ 1u))
                    store_val<P::hash, NW>(rec2 + P::n + (size_t)hgt * P::n, out);
                if (hgt == (uint32_t)P::a) {
                    store_val<P::hash, NW>(A.fors_roots + ((size_t)req2 * P::k + t2) * P::n, out);
                    fors_root_done<P>(A, req2);
                }
            }
        }
    }
}

template <class P, int J>
__global__ void __launch_bounds__(256) k_fors(BatchArgs A) {
    __shared__ uint32_t smem[fors_lds_words<P>()];
    fors_tile<P, J>(A, blockIdx.x, smem);
}

// thash (SHA-256) over M payload words after the 22-byte address,
// thash_sha256.cu:16-34: one compression per loop step; rd(i) returns payload
// word i (native order) from LDS or global memory.  The words of a block are
// read together (one memory latency per block), and blocks of payload words
// only cost one v_alignbit per word.
template <int NW, int M, class Rd>
__device__ __forceinline__ void thash_sha2_rd(const uint32_t mid[8], const Adr& a, const Rd& rd,
                                              uint32_t out[NW]) {
    constexpr int B = (6 + M + 2 + 15) / 16;             // compression blocks
    constexpr int FT = (M + 5) / 16;                     // first block that is not payload only
    constexpr uint32_t bits = (64u + 22u + 4u * M) * 8u;
    static_assert(FT >= 1 && FT < B, "a head, payload and tail blocks");
    uint32_t st[8];
#pragma unroll
    for (int i = 0; i < 8; i++) st[i] = mid[i];
#pragma unroll 1
    for (int blk = 0; blk < B; blk++) {
        uint32_t w[16], x[17];
        if (blk == 0) {
#pragma unroll
            for (int j = 0; j < 11; j++) x[j] = rd(j);
#pragma unroll
            for (int i = 0; i < 16; i++)
                w[i] = i < 5 ? a.w[i < 5 ? i : 0]
                     : i == 5 ? (a.w[5] | (x[0] >> 16))
                              : __builtin_amdgcn_alignbit(x[i - 6 > 0 ? i - 6 : 0], x[i - 5], 16);
        } else if (blk < FT) {
            const int base = 16 * blk - 6;
#pragma unroll
            for (int j = 0; j < 17; j++) x[j] = rd(base + j);
#pragma unroll
            for (int i = 0; i < 16; i++) w[i] = __builtin_amdgcn_alignbit(x[i], x[i + 1], 16);
        } else {
#pragma unroll
            for (int tb = FT; tb < B; tb++) {
                if (blk == tb) {
#pragma unroll
                    for (int j = 0; j < 17; j++) {
                        const int k = 16 * tb - 6 + j;
                        x[j] = rd(k < M ? k : M - 1);
                    }
#pragma unroll
                    for (int i = 0; i < 16; i++) {
                        const int g = 16 * tb + i;
                        uint32_t y = 0;
                        if (g < 5 + M) y = __builtin_amdgcn_alignbit(x[i], x[i + 1], 16);
                        else if (g == 5 + M) y = (x[i] << 16) | 0x8000u;
                        else if (g == 16 * B - 1) y = bits;
                        w[i] = y;
                    }
                }
            }
        }
        sha256_compress(st, w);
    }
#pragma unroll
    for (int i = 0; i < NW; i++) out[i] = st[i];
}

// thash_sha2_rd on a lane pair (both lanes active with the same mid, a and
// payload; sha256_compress_pair): rd2(j) returns payload word j + (lane & 1)
// clamped to 0..M-1, so each lane reads the words of its own half of every
// block.  The steps are separate so that a caller can feed the payload-only
// blocks from precomputed K + W (sha256_rounds_pair_kw).
template <int NW, int M>
struct Sha2Pair {
    static constexpr int B = (6 + M + 2 + 15) / 16;  // compression blocks
    static constexpr int FT = (M + 5) / 16;          // first block that is not payload only
    static constexpr uint32_t bits = (64u + 22u + 4u * M) * 8u;

    __device__ static __forceinline__ uint32_t mask() { return (threadIdx.x & 1u) ? ~0u : 0u; }

    // chaining value halves: e..h (even lane), a..d (odd lane)
    __device__ static __forceinline__ void init(const uint32_t mid[8], uint32_t st[4]) {
        const uint32_t m = mask();
#pragma unroll
        for (int i = 0; i < 4; i++) st[i] = lane_sel(m, mid[i], mid[4 + i]);
    }
    // word pairs of block tb (a constant once unrolled): word g = 16 tb + i is
    // address, payload re-aligned by 2 bytes (lo / hi = this lane's payload
    // words g - 6 and g - 5), the 0x80 pad, zero or the bit length
    template <class Rd2>
    __device__ static __forceinline__ void block(int tb, const Adr& a, const Rd2& rd2, uint32_t v[8]) {
        const uint32_t m = mask();
#pragma unroll
        for (int k = 0; k < 8; k++) {
            const uint32_t lo = rd2(16 * tb - 6 + 2 * k), hi = rd2(16 * tb - 5 + 2 * k);
            uint32_t y[2];
#pragma unroll
            for (int p = 0; p < 2; p++) {
                const int g = 16 * tb + 2 * k + p;
                y[p] = g < 5        ? a.w[g < 5 ? g : 0]
                     : g == 5       ? (a.w[5] | (lo >> 16))
                     : g < 5 + M    ? __builtin_amdgcn_alignbit(lo, hi, 16)
                     : g == 5 + M   ? ((lo << 16) | 0x8000u)
                     : g == 16 * B - 1 ? bits : 0u;
            }
            v[k] = lane_sel(m, y[1], y[0]);
        }
    }
    // payload-only block blk (runtime): this lane's word pairs
    template <class Rd2>
    __device__ static __forceinline__ void payload(int blk, const Rd2& rd2, uint32_t v[8]) {
        const int base = 16 * blk - 6;
#pragma unroll
        for (int k = 0; k < 8; k++)
            v[k] = __builtin_amdgcn_alignbit(rd2(base + 2 * k), rd2(base + 2 * k + 1), 16);
    }
    // blocks from..to-1
    template <class Rd2>
    __device__ static __forceinline__ void absorb(uint32_t st[4], int from, int to, const Adr& a,
                                                  const Rd2& rd2) {
#pragma unroll 1
        for (int blk = from; blk < to; blk++) {
            uint32_t v[8];
            if (blk == 0) {
                block(0, a, rd2, v);
            } else if (blk < FT) {
                payload(blk, rd2, v);
            } else {
#pragma unroll
                for (int tb = 1; tb < B; tb++)
                    if (tb >= FT && blk == tb) block(tb, a, rd2, v);
            }
            sha256_compress_pair(st, v);
        }
    }
    // a..d are on the odd lane, e..h on the even one; the output in both
    __device__ static __forceinline__ void output(const uint32_t st[4], uint32_t out[NW]) {
        const uint32_t m = mask();
#pragma unroll
        for (int w = 0; w < NW; w++) {
            const uint32_t own = st[w & 3], other = pair_swap(st[w & 3]);
            out[w] = w < 4 ? lane_sel(m, own, other) : lane_sel(m, other, own);
        }
    }
};

// the whole thash on a lane pair; the output is in both lanes
template <int NW, int M, class Rd2>
__device__ __forceinline__ void thash_sha2_pair(const uint32_t mid[8], const Adr& a, const Rd2& rd2,
                                                uint32_t out[NW]) {
    using SP = Sha2Pair<NW, M>;
    uint32_t st[4];
    SP::init(mid, st);
    SP::absorb(st, 0, SP::B, a, rd2);
    SP::output(st, out);
}

// K[t] + W[t] (t = 0..63) of payload-only block b of a thash_sha2_pair
// payload held in LDS as native words (pk), on one lane: the message
// schedule of a block does not depend on the chaining value, so idle lanes
// can run it ahead of the pair that compresses (sha256_rounds_pair_kw)
__device__ __forceinline__ void sha2_kw_block(const uint32_t* pk, int b, uint32_t kw[64]) {
    uint32_t w[16];
    const int base = 16 * b - 6;
#pragma unroll
    for (int i = 0; i < 16; i++) w[i] = __builtin_amdgcn_alignbit(pk[base + i], pk[base + i + 1], 16);
#pragma unroll
    for (int t = 0; t < 64; t++) {
        if (t >= 16) {
            const uint32_t x = w[(t - 15) & 15], y = w[(t - 2) & 15];
            w[t & 15] = xor3(rotr32(y, 17), rotr32(y, 19), y >> 10) + w[(t - 7) & 15] +
                        xor3(rotr32(x, 7), rotr32(x, 18), x >> 3) + w[t & 15];
        }
        kw[t] = kSha256K[t] + w[t & 15];
    }
}

// SHAKE256 thash over M payload words (thash_shake256.cu:15-25: pub_seed ||
// addr32 || payload), 136 bytes per permutation; rd(i) returns payload word
// i, read a block at a time.  IL: the payload words are bit-interleaved and
// the sponge runs in the interleaved domain (keccak_f_il); else plain.  The
// output is a plain value either way.
template <class P, int M, bool IL, class Rd>
__device__ __forceinline__ void thash_shake_rd(const typename Shake<P::n>::Key& key, const Adr& a,
                                               const Rd& rd, uint32_t out[P::nw]) {
    using Hh = Shake<P::n>;
    constexpr int NW = P::nw;
    constexpr int T = NW + 8 + M;             // absorbed words
    constexpr int NB = T / 34 + 1;            // permutations (the pad always fits)
    constexpr int TAIL = T - 34 * (NB - 1);   // words in the last block: the 0x1F byte follows
    static_assert(NB >= 2, "at least two blocks");
    uint32_t pre[NW + 8];                     // pub_seed || address
    if constexpr (IL) {
        typename Hh::ChainIl c;
        Hh::chain_il_begin(key, a, c);
#pragma unroll
        for (int i = 0; i < NW + 8; i++) pre[i] = c.w[i];
    } else {
#pragma unroll
        for (int i = 0; i < NW + 8; i++) pre[i] = i < NW ? key.pub[i < NW ? i : 0] : Hh::aw(a, i - NW);
    }
    uint32_t s[50];
#pragma unroll
    for (int i = 0; i < 50; i++) s[i] = 0;
#pragma unroll 1
    for (int blk = 0; blk < NB; blk++) {
        if (blk == 0) {
            uint32_t x[34 - NW - 8];
#pragma unroll
            for (int j = 0; j < 34 - NW - 8; j++) x[j] = rd(j);
#pragma unroll
            for (int i = 0; i < 34; i++) s[i] ^= i < NW + 8 ? pre[i < NW + 8 ? i : 0] : x[i - NW - 8];
        } else if (blk < NB - 1) {
            const int base = 34 * blk - NW - 8;
            uint32_t x[34];
#pragma unroll
            for (int j = 0; j < 34; j++) x[j] = rd(base + j);
#pragma unroll
            for (int i = 0; i < 34; i++) s[i] ^= x[i];
        } else {
            constexpr int base = 34 * (NB - 1) - NW - 8;
            uint32_t x[TAIL > 0 ? TAIL : 1];
#pragma unroll
            for (int j = 0; j < TAIL; j++) x[j] = rd(base + j);
#pragma unroll
            for (int i = 0; i < TAIL; i++) s[i] ^= x[i];
            if constexpr (!IL) {
                s[TAIL] ^= 0x1Fu;
            } else if constexpr ((TAIL & 1) == 0) {  // 0x1F at bit 0 of lane TAIL / 2
                s[TAIL] ^= 7u;
                s[TAIL + 1] ^= 3u;
            } else {                                  // 0x1F at bit 32 of lane TAIL / 2
                s[TAIL - 1] ^= 7u << 16;
                s[TAIL] ^= 3u << 16;
            }
            s[33] ^= 0x80000000u;  // 0x80 at byte 135 (lane 16, bit 63: odd, so the same word)
        }
        if constexpr (IL) keccak_f_il(s);
        else keccak_f(s);
    }
    if constexpr (IL) Hh::from_il(s, out);
    else {
#pragma unroll
        for (int i = 0; i < NW; i++) out[i] = s[i];
    }
}

// SHAKE256 thash on a lane pair in the interleaved domain (both lanes
// active, same key, address and payload; keccak_f_il_pair): rd2(j), j even,
// returns interleaved payload word j + (lane & 1), clamped to 0..M-1, i.e.
// this lane's half of payload lane j / 2.  The output (plain) is in both
// lanes.
template <class P, int M, class Rd2>
__device__ __forceinline__ void thash_shake_pair(const typename Shake<P::n>::Key& key,
                                                 const Adr& a, const Rd2& rd2,
                                                 uint32_t out[P::nw]) {
    using Hh = Shake<P::n>;
    constexpr int NW = P::nw;
    constexpr int PRE = NW + 8;                // pub_seed || address, interleaved words
    constexpr int T = PRE + M;                 // absorbed words
    constexpr int NB = T / 34 + 1;             // permutations (the pad always fits)
    constexpr int TAIL = T - 34 * (NB - 1);    // words in the last block
    static_assert((PRE & 1) == 0, "the prefix fills whole lanes");
    const bool odd = threadIdx.x & 1u;
    const uint32_t m = odd ? ~0u : 0u;
    typename Hh::ChainIl c;
    Hh::chain_il_begin(key, a, c);
    uint32_t s[25];
#pragma unroll
    for (int i = 0; i < 25; i++) s[i] = 0;
    // interleaved word w of the last block: the 0x1F pad (byte 4 TAIL) and
    // 0x80 (byte 135) in their interleaved bits
    auto pad = [](int w) -> uint32_t {
        uint32_t v = 0;
        if ((TAIL & 1) == 0) v ^= w == TAIL ? 7u : w == TAIL + 1 ? 3u : 0u;
        else v ^= w == TAIL - 1 ? 7u << 16 : w == TAIL ? 3u << 16 : 0u;
        return v ^ (w == 33 ? 0x80000000u : 0u);
    };
    // block tb (a constant once unrolled): stream word g = 34 tb + 2 L + half
    // is prefix, payload (this lane's half from rd2), zero, and the pad in
    // the last block
    auto block = [&](int tb) {
#pragma unroll
        for (int L = 0; L < 17; L++) {
            const int g0 = 34 * tb + 2 * L;
            const uint32_t x = (g0 >= PRE && g0 < T) ? rd2(g0 - PRE) : 0u;
            uint32_t y[2];
#pragma unroll
            for (int p = 0; p < 2; p++) {
                const int g = g0 + p;
                uint32_t v = g < PRE ? c.w[g < PRE ? g : 0] : g < T ? x : 0u;
                if (tb == NB - 1) v ^= pad(g - 34 * (NB - 1));
                y[p] = v;
            }
            s[L] ^= lane_sel(m, y[1], y[0]);
        }
    };
#pragma unroll 1
    for (int blk = 0; blk < NB; blk++) {
        if (blk == 0) {
            block(0);
        } else if (blk < NB - 1) {  // payload only
            const int base = 34 * blk - PRE;
#pragma unroll
            for (int L = 0; L < 17; L++) s[L] ^= rd2(base + 2 * L);
        } else {
            block(NB - 1);
        }
        keccak_f_il_pair(s);
    }
    uint32_t il[NW];
#pragma unroll
    for (int L = 0; L < NW / 2; L++) {
        const uint32_t other = pair_swap(s[L]);
        il[2 * L] = lane_sel(m, other, s[L]);
        il[2 * L + 1] = lane_sel(m, s[L], other);
    }
    Hh::from_il(il, out);
}

// FORS public key of request i: T_k over its k roots (fors.cu:142); the
// roots are read past the caches, as another workgroup may have stored them.
// PAIR: on a lane pair (both lanes active, the even lane stores).
template <class P, bool PAIR>
__device__ __forceinline__ void fors_pk_of(const BatchArgs& A, uint32_t i) {
    using Hh = HashT<P>;
    using Key = typename Hh::Key;
    constexpr int NW = P::nw;
    constexpr int KW = P::k * NW;
    const Plan& pl = A.plan[i];
    Key key;
    Hh::load_key(pl, key);
    Adr adr = adr_make(0, pl.tree, kForsPk);
    adr_keypair<P::th>(adr, pl.leaf);
    const uint32_t* rw = reinterpret_cast<const uint32_t*>(A.fors_roots + (size_t)i * P::k * P::n);
    auto root = [&](int u) { return __builtin_nontemporal_load(rw + u); };
    const int h = threadIdx.x & 1;
    uint32_t out[NW];
    if constexpr (P::hash == kSha2) {
        if constexpr (PAIR) {
            thash_sha2_pair<NW, KW>(key.mid, adr, [&](int j) {
                const int u = j + h;
                return bswap32(root(u < 0 ? 0 : u < KW ? u : KW - 1));
            }, out);
        } else {
            thash_sha2_rd<NW, KW>(key.mid, adr, [&](int u) { return bswap32(root(u)); }, out);
        }
    } else {
        if constexpr (PAIR) {
            thash_shake_pair<P, KW>(key, adr, [&](int j) {  // this lane's half of root lane j / 2
                int l = j >> 1;
                l = l < 0 ? 0 : l < KW / 2 ? l : KW / 2 - 1;
                uint32_t e, o;
                il_to(root(2 * l), root(2 * l + 1), e, o);
                return h ? o : e;
            }, out);
        } else {
            thash_shake_rd<P, KW, false>(key, adr, root, out);
        }
    }
    if (!PAIR || h == 0) store_val<P::hash, NW>(A.fors_pk + (size_t)i * P::n, out);
}

// FORS public keys of a whole batch, a lane pair per request (large sign
// batches and verify)
template <class P>
__global__ void __launch_bounds__(64) k_fors_pk(BatchArgs A) {
    const uint32_t i = (blockIdx.x * blockDim.x + threadIdx.x) >> 1;
    if (i < A.n_req) fors_pk_of<P, true>(A, i);
}

// verify: one FORS tree per lane: leaf from the revealed sk, climb with the
// auth path (fors_pk_from_sig, fors.cu:596-633; compute_root utils.cu:94-135)
template <class P>
__global__ void __launch_bounds__(64) k_vfors(BatchArgs A) {
    using Hh = HashT<P>;
    using Key = typename Hh::Key;
    constexpr int NW = P::nw;
    const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= A.n_req * P::k) return;
    const uint32_t req = g / P::k, t = g % P::k;
    const Plan& pl = A.plan[req];
    Key key;
    Hh::load_key(pl, key);
    const uint8_t* rec = A.sig + (size_t)req * A.sig_stride + P::n + (size_t)t * (P::a + 1) * P::n;
    uint32_t li = pl.fors_idx[t];
    uint32_t off = t << P::a;
    Adr adr = adr_make(0, pl.tree, kForsTree);
    adr_keypair<P::th>(adr, pl.leaf);
    adr_node(adr, 0, li + off);
    uint32_t sk[NW], node[NW];
    load_val<P::hash, NW>(rec, sk);
    Hh::F(key, adr, sk, node);
#pragma unroll 1
    for (int h = 0; h < P::a; h++) {
        uint32_t ap[NW];
        load_val<P::hash, NW>(rec + P::n + (size_t)h * P::n, ap);
        const bool odd = li & 1u;
        li >>= 1;
        off >>= 1;
        adr_node(adr, h + 1, li + off);
        uint32_t l[NW], r[NW];
#pragma unroll
        for (int w = 0; w < NW; w++) {
            l[w] = odd ? ap[w] : node[w];
            r[w] = odd ? node[w] : ap[w];
        }
        Hh::H(key, adr, l, r, node);
    }
    store_val<P::hash, NW>(A.fors_roots + (size_t)g * P::n, node);
}

// ------------------------------------------------------------- hypertree

// WOTS chain lengths of an n-byte message (wots.cu:84-162): nibbles MSB
// first, then the 3-nibble checksum (sum of 15 - d_i, shifted left by 4)
template <class P>
__device__ __forceinline__ uint32_t chain_len(const uint32_t msg[P::nw], int j) {
    if (j < P::len1) {
        const uint32_t byte = val_byte<P::hash>(msg, j >> 1);
        return (j & 1) ? (byte & 15u) : (byte >> 4);
    }
    uint32_t csum = 0;
#pragma unroll
    for (int i = 0; i < P::n; i++) {
        const uint32_t byte = val_byte<P::hash>(msg, i);
        csum += 30u - (byte >> 4) - (byte & 15u);
    }
    const uint32_t v = csum << 4;  // 16-bit field, digits are its top 12 bits
    const int c = j - P::len1;
    return (v >> (12 - 4 * c)) & 15u;
}

// tree index of hypertree layer `layer`: the reference shifts by h' once per
// layer (sign.cu:1987-1988), so a total shift of 64 bits (256f, layer 16)
// yields 0 -- a single 64-bit shift by 64 would be undefined
template <class P>
__device__ __forceinline__ uint64_t layer_tree(uint64_t tree, uint32_t layer) {
    const uint32_t sh = layer * P::th;
    return sh >= 64 ? 0ull : (tree >> sh);
}

// One Merkle node of a hypertree subtree (treehash, utils.cu:189-241): node x
// of level l = H(left, right) under the subtree's HASHTREE address; writes the
// node into the signature's auth path when it is the sibling of the signing
// leaf's ancestor, and the subtree root (sign: roots[], keygen: pk / sk).
template <class P, int MODE>
__device__ __forceinline__ void ht_node(const BatchArgs& A, uint32_t g, uint32_t ntrees, int l,
                                        uint32_t x, const uint32_t lft[P::nw],
                                        const uint32_t rgt[P::nw], uint32_t out[P::nw]) {
    using Hh = HashT<P>;
    constexpr int NW = P::nw;
    constexpr int L = 1 << P::th;
    const bool act = g < ntrees;
    const uint32_t req = act ? (MODE == 0 ? g / P::d : g) : 0;
    const uint32_t lay = MODE == 0 ? (act ? g % P::d : 0) : P::d - 1;
    const Plan& pl = A.plan[req];
    typename Hh::Key key;
    Hh::load_key(pl, key);
    uint64_t tree = 0;
    uint32_t target = 0;
    if (MODE == 0) {
        tree = layer_tree<P>(pl.tree, lay);
        target = lay == 0 ? pl.leaf : (uint32_t)((pl.tree >> ((lay - 1) * P::th)) & (L - 1));
    }
    Adr ha = adr_make(lay, tree, kHashTree);
    adr_node(ha, l, x);
    Hh::H(key, ha, lft, rgt, out);
    if (!act) return;
    if (MODE == 0) {
        uint8_t* a2 = A.sig + (size_t)req * A.sig_stride + P::n + P::fors_bytes +
                      (size_t)lay * P::layer_bytes + P::wots_bytes;
        if (l < P::th && x == ((target >> l) ^ 1u)) store_val<P::hash, NW>(a2 + (size_t)l * P::n, out);
        if (l == P::th) store_val<P::hash, NW>(A.roots + ((size_t)req * P::d + lay) * P::n, out);
    } else if (l == P::th) {
        store_val<P::hash, NW>(A.sk_out + (size_t)req * P::sk_bytes + 3 * P::n, out);
        store_val<P::hash, NW>(A.pk_out + (size_t)req * P::pk_bytes + P::n, out);
    }
}

// Hypertree leaves.  MODE 0 = sign: every (request, layer) subtree, auth
// path into the signature, layer root into roots[], chain values of the
// signing leaf captured.  MODE 1 = keygen: the top subtree only, root into
// the key outputs.
//
// S lanes share one WOTS leaf (lane q takes chains q, q+S, ...): S = 1 gives
// the most work per lane (throughput), larger S shortens the per-lane chain
// of dependent compressions (latency of small batches); the launcher picks S.
// With S > 1 (and always for SHAKE) the chain outputs are parked in tl_buf
// ([word][leaf], coalesced) and absorbed into T_len by the lane pair q = 0, 1
// (sha256_compress_pair / keccak_f_il_pair; SHAKE256 at S = 1: lane q = 0).
//
// A subtree whose 2^h' x S lanes exceed one 1024-lane workgroup (the s sets
// at S >= 4) is split over PARTS workgroups of 256 lanes.  Each part reduces
// its leaves to a part root in LDS; the part that finishes last (device-scope
// counter, release/acquire fences, no waiting) climbs the remaining log2(PARTS)
// levels from the part roots.
template <class P, int S>
struct HtShape {
    static constexpr int L = 1 << P::th;              // leaves per subtree
    static constexpr int LT = L * S;                  // lanes per subtree
    static constexpr int PARTS = LT > 1024 ? LT / 256 : 1;
    static constexpr int LP = L / PARTS;              // leaves per part
    static constexpr int LOGP = P::th - ilog2(PARTS); // levels reduced inside a part
    static constexpr int WG = PARTS > 1 ? 256 : (LT > 256 ? LT : 256);
    static constexpr int TPW = PARTS > 1 ? 1 : WG / LT;  // subtrees per workgroup
    static constexpr int NL = WG / S;                 // leaves per workgroup
    static constexpr bool park = S > 1 || P::hash == kShake;
    static constexpr int NLDS = NL > PARTS ? NL : PARTS;  // the last part climbs from PARTS roots
    static constexpr int lds_words = 2 * NLDS * P::nw;    // two levels of nodes
};

// smem: HtShape<P, S>::lds_words words of LDS owned by the kernel
template <class P, int MODE, int S>
__device__ __forceinline__ void ht_tile(const BatchArgs& A, uint32_t blk, uint32_t* smem) {
    using Hh = HashT<P>;
    using Key = typename Hh::Key;
    using Sh = HtShape<P, S>;
    constexpr int NW = P::nw;
    constexpr int LOGL = P::th;
    constexpr int L = Sh::L, TPW = Sh::TPW, NL = Sh::NL;
    constexpr int PARTS = Sh::PARTS, LP = Sh::LP, LOGP = Sh::LOGP;
    constexpr int LPT = PARTS > 1 ? LP * S : Sh::LT;  // lanes per subtree in this workgroup
    constexpr int NLDS = Sh::NLDS;
    auto& lds = *reinterpret_cast<uint32_t (*)[2][NLDS][NW]>(smem);

    const int tid = threadIdx.x;
    const int tau = tid / LPT;
    const uint32_t part = PARTS > 1 ? blk % PARTS : 0;
    const uint32_t u = part * LP + (tid % LPT) / S;  // leaf index = keypair address
    const int q = tid % S;                            // chain phase of this lane
    const int slot = tid / S;                         // leaf slot in the workgroup
    const uint32_t g = (blk / PARTS) * TPW + tau;
    const uint32_t ntrees = MODE == 0 ? A.n_req * P::d : A.n_req;
    const bool active = g < ntrees;
    const uint32_t req = active ? (MODE == 0 ? g / P::d : g) : 0;
    const uint32_t layer = MODE == 0 ? (active ? g % P::d : 0) : P::d - 1;
    const Plan& pl = A.plan[req];
    Key key;
    Hh::load_key(pl, key);
    uint64_t tree = 0;
    uint32_t target = 0;
    if (MODE == 0) {
        tree = layer_tree<P>(pl.tree, layer);
        target = layer == 0 ? pl.leaf
                            : (uint32_t)((pl.tree >> ((layer - 1) * P::th)) & (L - 1));
    }
    const bool capture = MODE == 0 && active && u == target;
    uint8_t* cap = A.cap + ((size_t)req * P::d + layer) * P::len * 16 * P::n;
    const uint32_t leaf_g = blk * NL + slot;  // leaf column in tl_buf

    // ---- WOTS public key of leaf u (wots_gen_leaf, sign.cu:87-104)
    Adr wa = adr_make(layer, tree, kWots);
    adr_keypair<P::th>(wa, u);
    Adr pa = adr_make(layer, tree, kWotsPk);
    adr_keypair<P::th>(pa, u);
    // SHA-256 at S = 1 streams T_len in registers; otherwise the chain ends
    // are parked in tl_buf and absorbed afterwards (by the leaf's lane pair
    // when S > 1)
    typename Sha2<P::n>::Stream ts;
    if constexpr (P::hash == kSha2 && !Sh::park) ts.begin(key, pa);
    auto parked = [&](int w) { return A.tl_buf[(size_t)w * A.tl_lanes + leaf_g]; };
#pragma unroll 1
    for (int j = q; j < P::len; j += S) {
        uint32_t v[NW];
        adr_chain(wa, j);
        adr_hash(wa, 0);
        Hh::PRF(key, wa, v);  // wots_gen_sk, wots.cu:27-33
        if constexpr (P::hash == kSha2) {
            if (capture) store_val<P::hash, NW>(cap + (size_t)(j * 16) * P::n, v);
#pragma unroll 1
            for (int s = 0; s < P::w - 1; s++) {  // gen_chain(0, w-1), wots.cu:51-62
                adr_hash(wa, s);
                Hh::F(key, wa, v, v);
                if (capture) store_val<P::hash, NW>(cap + (size_t)(j * 16 + s + 1) * P::n, v);
            }
        } else {
            // SHAKE: the chain runs bit-interleaved; the capture keeps that
            // form (k_wots_gather converts the values it picks)
            typename Hh::ChainIl ch;
            Hh::chain_il_begin(key, wa, ch);
            uint32_t vi[NW];
            Hh::to_il(v, vi);
            if (capture) store_val<P::hash, NW>(cap + (size_t)(j * 16) * P::n, vi);
#pragma unroll 1
            for (int s = 0; s < P::w - 1; s++) {  // gen_chain(0, w-1), wots.cu:51-62
                Hh::F_il(ch, s, vi, vi);
                if (capture) store_val<P::hash, NW>(cap + (size_t)(j * 16 + s + 1) * P::n, vi);
            }
#pragma unroll
            for (int w = 0; w < NW; w++) v[w] = vi[w];  // parked interleaved
        }
        if constexpr (!Sh::park) {
#pragma unroll 1
            for (int i = 0; i < NW; i++) ts.word(pick<NW>(v, i));
        } else {
#pragma unroll
            for (int i = 0; i < NW; i++) A.tl_buf[(size_t)(j * NW + i) * A.tl_lanes + leaf_g] = v[i];
        }
    }
    if constexpr (S > 1) __syncthreads();  // all chains of the leaf parked
    uint32_t node[NW];
    // T_len: with S > 1 on the lane pair q = 0, 1 of the leaf (S is even, so
    // the pair is lanes 2i, 2i + 1), else on lane q = 0
    constexpr bool pair = S > 1;
    if (q < (pair ? 2 : 1)) {
        constexpr int TLW = P::len * NW;
        auto parked2 = [&](int j) {
            const int i = j + q;
            return parked(i < 0 ? 0 : i < TLW ? i : TLW - 1);
        };
        if constexpr (P::hash == kSha2) {
            if constexpr (pair) thash_sha2_pair<NW, TLW>(key.mid, pa, parked2, node);
            else ts.template finish<TLW>(node);
        } else {
            if constexpr (pair) thash_shake_pair<P, TLW>(key, pa, parked2, node);
            else thash_shake_rd<P, TLW, true>(key, pa, parked, node);
        }
        if (q == 0) {
            uint8_t* auth = A.sig + (size_t)req * A.sig_stride + P::n + P::fors_bytes +
                            (size_t)layer * P::layer_bytes + P::wots_bytes;
            if (MODE == 0 && active && u == (target ^ 1u)) store_val<P::hash, NW>(auth, node);
#pragma unroll
            for (int w = 0; w < NW; w++) lds[0][slot][w] = node[w];
        }
    }

    // ---- subtree: level l has NL >> l nodes here, node x = H(2x, 2x+1)
#pragma unroll 1
    for (int l = 1; l <= LOGP; l++) {
        __syncthreads();
        const int src = (l - 1) & 1, dst = l & 1;
        if (tid < (NL >> l)) {
            const uint32_t t2 = tid >> (LOGP - l);
            const uint32_t x = part * (LP >> l) + (tid & ((LP >> l) - 1));
            uint32_t lft[NW], rgt[NW];
#pragma unroll
            for (int w = 0; w < NW; w++) {
                lft[w] = lds[src][2 * tid][w];
                rgt[w] = lds[src][2 * tid + 1][w];
            }
            uint32_t out[NW];
            ht_node<P, MODE>(A, (blk / PARTS) * TPW + t2, ntrees, l, x, lft, rgt, out);
#pragma unroll
            for (int w = 0; w < NW; w++) lds[dst][tid][w] = out[w];
        }
    }
    if constexpr (PARTS > 1) {
        // part root -> ht_parts; the last part of the subtree climbs the rest
        __shared__ uint32_t last;
        if (tid == 0) {
            uint32_t* dst = A.ht_parts + ((size_t)g * PARTS + part) * NW;
#pragma unroll
            for (int w = 0; w < NW; w++) dst[w] = lds[LOGP & 1][0][w];
            __threadfence();  // release the part root before counting this part
            last = active ? (atomicAdd(&A.ht_count[g], 1u) == (uint32_t)PARTS - 1) : 0u;
        }
        __syncthreads();
        if (last) {
            __threadfence();  // acquire: every part root is visible
            if (tid < PARTS) {
                const uint32_t* srcp = A.ht_parts + ((size_t)g * PARTS + tid) * NW;
#pragma unroll
                for (int w = 0; w < NW; w++)
                    lds[LOGP & 1][tid][w] = __builtin_nontemporal_load(srcp + w);
            }
#pragma unroll 1
            for (int l = LOGP + 1; l <= LOGL; l++) {
                __syncthreads();
                const int src = (l - 1) & 1, dst = l & 1;
                if (tid < (L >> l)) {
                    uint32_t lft[NW], rgt[NW], out[NW];
#pragma unroll
                    for (int w = 0; w < NW; w++) {
                        lft[w] = lds[src][2 * tid][w];
                        rgt[w] = lds[src][2 * tid + 1][w];
                    }
                    ht_node<P, MODE>(A, g, ntrees, l, tid, lft, rgt, out);
#pragma unroll
                    for (int w = 0; w < NW; w++) lds[dst][tid][w] = out[w];
                }
            }
        }
    }
}

template <class P, int MODE, int S>
__global__ void __launch_bounds__((HtShape<P, S>::WG)) k_ht(BatchArgs A) {
    __shared__ uint32_t smem[HtShape<P, S>::lds_words];
    ht_tile<P, MODE, S>(A, blockIdx.x, smem);
}

constexpr int max_int(int a, int b) { return a > b ? a : b; }

// Signing as one persistent kernel: a grid of resident 256-lane workgroups
// pulls tiles from a device counter (zeroed by k_prep) until the queue is
// empty -- first the hypertree tiles (long), then the FORS tiles (short),
// which fill the tail.  The tiles are independent given the plans, so no
// workgroup ever waits on another; every workgroup exits once the counter
// passes the last tile.
// SHA-256: 4 waves per SIMD, as k_ht (5, at 96 VGPRs, keeps the F loop free
// of spills but spills in T_len, and took 13.87-13.90 ms at SHA2-128f / 4096
// against 13.73-13.81, both with 16 lanes per FORS tree); SHAKE256 (50-word
// sponge): 3, as k_ht -- at 4 (128 VGPRs) every fast set signed slower, also
// with the chain-step loop kept free of spills (SHAKE256-128f at 4096 40.4
// against 39.95 ms)
template <class P, int J, int S>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(P::hash == kSha2 ? 4 : 3)))
k_sign_fused(BatchArgs A, uint32_t n_ht, uint32_t n_fors) {
    static_assert(HtShape<P, S>::WG == 256, "hypertree tiles of 256 lanes");
    __shared__ uint32_t tile;
    // one LDS buffer for both tile kinds (a workgroup runs one tile at a time)
    __shared__ uint32_t smem[max_int(HtShape<P, S>::lds_words, fors_lds_words<P>())];
#pragma unroll 1
    while (true) {
        if (threadIdx.x == 0) tile = atomicAdd(A.tile_ctr, 1u);
        __syncthreads();
        const uint32_t t = tile;
        __syncthreads();  // every lane has the tile (and left the previous one)
        if (t >= n_ht + n_fors) break;
        if (t < n_ht) ht_tile<P, 0, S>(A, t, smem);
        else fors_tile<P, J>(A, t - n_ht, smem);
    }
}

// WOTS signatures: lane per (request, layer, chain); the layer signs the FORS
// public key (layer 0) or the root of the layer below (sign.cu:1970-1989)
template <class P>
__global__ void __launch_bounds__(256) k_wots_gather(BatchArgs A) {
    constexpr int NW = P::nw;
    const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= A.n_req * P::d * P::len) return;
    const uint32_t j = g % P::len;
    const uint32_t rl = g / P::len;
    const uint32_t layer = rl % P::d, req = rl / P::d;
    const uint8_t* mp = layer == 0 ? A.fors_pk + (size_t)req * P::n
                                   : A.roots + ((size_t)req * P::d + layer - 1) * P::n;
    uint32_t msg[NW];
    load_val<P::hash, NW>(mp, msg);
    const uint32_t len = chain_len<P>(msg, j);
    // the capture holds n-byte values at multiples of n from a 256-byte
    // aligned base: 16-byte loads (8-byte for n = 24), all before the stores
    // (word-by-word load / store pairs could alias and stayed 4-byte
    // accesses)
    const uint8_t* src = A.cap + ((((size_t)req * P::d + layer) * P::len + j) * 16 + len) * P::n;
    uint8_t* dst = A.sig + (size_t)req * A.sig_stride + P::n + P::fors_bytes +
                   (size_t)layer * P::layer_bytes + (size_t)j * P::n;
    uint32_t v[NW];
    if constexpr (NW % 4 == 0) {
#pragma unroll
        for (int k = 0; k < NW / 4; k++) {
            const uint4 x = reinterpret_cast<const uint4*>(src)[k];
            v[4 * k] = x.x;
            v[4 * k + 1] = x.y;
            v[4 * k + 2] = x.z;
            v[4 * k + 3] = x.w;
        }
    } else {
#pragma unroll
        for (int k = 0; k < NW / 2; k++) {
            const uint2 x = reinterpret_cast<const uint2*>(src)[k];
            v[2 * k] = x.x;
            v[2 * k + 1] = x.y;
        }
    }
    if constexpr (P::hash == kShake) {  // SHAKE chains are captured bit-interleaved
        uint32_t vi[NW];
#pragma unroll
        for (int w = 0; w < NW; w++) vi[w] = v[w];
        Shake<P::n>::from_il(vi, v);
    }
#pragma unroll
    for (int w = 0; w < NW; w++) reinterpret_cast<uint32_t*>(dst)[w] = v[w];
}

// The climb of one layer's subtree on a lane pair (compute_root,
// utils.cu:94-135): node (in both lanes) is the WOTS public key's T_len on
// entry and the layer root on exit; the auth path is in LDS (raw words) and
// nd is NW words of LDS for the node.
template <class P>
__device__ __forceinline__ void climb_pair(const typename HashT<P>::Key& key, uint32_t layer,
                                           uint64_t tree, uint32_t leaf, const uint32_t* auth,
                                           uint32_t* nd, uint32_t node[P::nw]) {
    constexpr int NW = P::nw;
    constexpr bool SHA2 = P::hash == kSha2;
    const int half = threadIdx.x & 1;
    uint32_t li = leaf;
    Adr ha = adr_make(layer, tree, kHashTree);
#pragma unroll 1
    for (int h = 0; h < P::th; h++) {
        const bool right = li & 1u;  // node is the right child
        li >>= 1;
        adr_node(ha, h + 1, li);
        // payload node || auth or auth || node from LDS (nd: the node as raw
        // words, like the auth path)
#pragma unroll
        for (int w = 0; w < NW; w++) nd[w] = SHA2 ? bswap32(node[w]) : node[w];
        const uint32_t* ap = auth + h * NW;
        auto word = [&](int i) {  // payload word i (raw)
            const uint32_t* src = (i < NW) == right ? ap : nd;
            return src[i < NW ? i : i - NW];
        };
        if constexpr (SHA2) {
            thash_sha2_pair<NW, 2 * NW>(key.mid, ha, [&](int j) {
                int i = j + half;
                i = i < 0 ? 0 : i < 2 * NW ? i : 2 * NW - 1;
                return bswap32(word(i));
            }, node);
        } else {
            // this lane's interleaved half of payload lane j / 2
            thash_shake_pair<P, 2 * NW>(key, ha, [&](int j) {
                int l = j >> 1;
                l = l < 0 ? 0 : l < NW ? l : NW - 1;
                uint32_t e, o;
                il_to(word(2 * l), word(2 * l + 1), e, o);
                return half ? o : e;
            }, node);
        }
    }
}

// reader of a WOTS public key held in LDS for the lane-pair thashes
template <class P>
struct PkPair {
    static constexpr int TLW = P::len * P::nw;
    const uint32_t* pk;
    int half;
    __device__ uint32_t operator()(int j) const {
        const int i = j + half;
        return pk[i < 0 ? 0 : i < TLW ? i : TLW - 1];
    }
};

// Verification of one layer's subtree on a lane pair (both lanes active):
// T_len over the WOTS public key in LDS (pk: native words for SHA-256,
// interleaved for SHAKE256), then climb_pair.  node is the layer root, in
// both lanes.
template <class P>
__device__ __forceinline__ void tlen_climb_pair(const typename HashT<P>::Key& key,
                                                uint32_t layer, uint64_t tree, uint32_t leaf,
                                                const uint32_t* pk, const uint32_t* auth,
                                                uint32_t* nd, uint32_t node[P::nw]) {
    constexpr int NW = P::nw;
    constexpr int TLW = P::len * NW;
    Adr pa = adr_make(layer, tree, kWotsPk);
    adr_keypair<P::th>(pa, leaf);
    const PkPair<P> rpk{pk, (int)(threadIdx.x & 1)};
    if constexpr (P::hash == kSha2) thash_sha2_pair<NW, TLW>(key.mid, pa, rpk, node);
    else thash_shake_pair<P, TLW>(key, pa, rpk, node);
    climb_pair<P>(key, layer, tree, leaf, auth, nd, node);
}

// verify: the d layers of R = 256 / G requests per workgroup in order
// (sign.cu:3077-3113), so G is the mean number of lanes per request.  Per
// layer the R * len WOTS chains of the workgroup (15 - len_j steps each,
// data dependent) are ranked by remaining steps with a counting sort in LDS
// and dealt to the 256 lanes in snake order: a longest-first schedule that
// evens the per-lane F counts across all requests of the workgroup, not per
// request.  Each lane runs its chains in one flattened loop (one F site;
// SHAKE256 chains run bit-interleaved).  The lane pair 2r, 2r + 1 then
// computes T_len of request r block by block from LDS and climbs its subtree
// (compute_root, utils.cu:94-135; tlen_climb_pair, or for SHA-256 with n <=
// 24 the rounds of T_len's payload-only blocks on K + W that waves 1-3
// computed meanwhile).  The next layer's WOTS signature words and this
// layer's auth path are fetched into registers while the chains run, so the
// layer loop waits on no global load.
template <class P, int G>
__global__ void __launch_bounds__(256) k_vht(BatchArgs A) {
    using Hh = HashT<P>;
    using Key = typename Hh::Key;
    constexpr bool SHA2 = P::hash == kSha2;
    constexpr int NW = P::nw;
    constexpr int LEN = P::len;
    constexpr int TH = P::th;
    constexpr int R = 256 / G;                   // requests per workgroup
    constexpr int TLW = LEN * NW;                // WOTS signature / public key words
    constexpr int NCH = R * LEN;                 // chains per workgroup and layer
    constexpr int MAXC = (NCH + 255) / 256;      // chains per lane
    constexpr int MAXW = (R * TLW + 255) / 256;  // prefetched WOTS words per lane
    constexpr int AW = TH * NW;                  // auth path words per request
    constexpr int MAXA = (R * AW + 255) / 256;   // prefetched auth words per lane
    static_assert(NCH < 4096, "chain ids take 12 bits of an order entry");
    // SHA-256 with n <= 24: T_len's payload-only blocks take their message
    // schedules from waves 1-3 (the LDS for n = 32 would not fit)
    constexpr int MB = (TLW + 5) / 16 - 1;  // payload-only blocks of T_len
    constexpr bool PRE = SHA2 && NW <= 6;
    __builtin_amdgcn_s_setprio(2);  // d dependent layers per request: latency-bound
    __shared__ uint32_t key_s[R][8];    // SHA-256 midstate / SHAKE256 pub_seed
    __shared__ uint32_t root_s[R][NW];  // root of the last finished layer
    __shared__ uint64_t tree_s[R];
    __shared__ uint32_t leaf_s[R];
    __shared__ uint32_t sig_s[R][TLW];  // WOTS signature of the layer, raw words
    __shared__ uint32_t pk_s[R][TLW];   // chain ends (SHAKE256: interleaved)
    __shared__ uint32_t auth_s[R][AW];  // auth path of the layer, raw words
    __shared__ uint32_t node_s[R][NW];  // climb: the node, raw words
    // K + W of T_len's payload-only blocks: 64-word rows, row (r, b) skewed
    // by 4 words per request so the pairs' 16-byte reads of one t (one row
    // per request) fall into distinct bank groups
    __shared__ __align__(16) uint32_t kw_s[PRE ? R * MB * 64 + 64 : 1];
    auto kw_row = [&](int r, int b) { return &kw_s[(r * MB + b) * 64 + 4 * (r & 15)]; };
    __shared__ uint16_t order_s[NCH];   // rank -> chain id | remaining steps << 12
    __shared__ uint32_t hist_s[16];     // chains per remaining-step count

    const int tid = threadIdx.x;
    const uint32_t req0 = blockIdx.x * R;
    const int nact = (int)min((uint32_t)R, A.n_req - req0);  // >= 1 by the grid
    const int nch = nact * LEN;
    auto sig_of = [&](int r) { return A.sig + (size_t)(req0 + r) * A.sig_stride; };
    auto wots_of = [&](int r, int layer) {
        return reinterpret_cast<const uint32_t*>(sig_of(r) + P::n + P::fors_bytes +
                                                 (size_t)layer * P::layer_bytes);
    };

    if (tid < nact) {
        const Plan& pl = A.plan[req0 + tid];
#pragma unroll
        for (int i = 0; i < 8; i++) key_s[tid][i] = SHA2 ? pl.mid[i] : pl.pubseed[i];
        uint32_t fp[NW];
        load_val<P::hash, NW>(A.fors_pk + (size_t)(req0 + tid) * P::n, fp);
#pragma unroll
        for (int w = 0; w < NW; w++) root_s[tid][w] = fp[w];
        tree_s[tid] = pl.tree;
        leaf_s[tid] = pl.leaf;
    }
    if (tid < 16) hist_s[tid] = 0;
#pragma unroll
    for (int k = 0; k < MAXW; k++) {
        const int u = tid + 256 * k;
        if (u < nact * TLW) sig_s[u / TLW][u % TLW] = wots_of(u / TLW, 0)[u % TLW];
    }

#pragma unroll 1
    for (int layer = 0; layer < P::d; layer++) {
        __syncthreads();  // roots, tree / leaf, signature words and histogram ready
        // remaining steps of every chain and its slot within its step count
        uint32_t cst[MAXC], cslot[MAXC];
#pragma unroll
        for (int k = 0; k < MAXC; k++) {
            const int q = tid + 256 * k;
            cst[k] = 0;
            cslot[k] = 0;
            if (q < nch) {
                uint32_t msg[NW];
#pragma unroll
                for (int w = 0; w < NW; w++) msg[w] = root_s[q / LEN][w];
                cst[k] = P::w - 1 - chain_len<P>(msg, q % LEN);
                cslot[k] = atomicAdd(&hist_s[15 - cst[k]], 1u);
            }
        }
        __syncthreads();
        // rank = chains with more steps + slot (longest first)
        {
            uint32_t h[16];
#pragma unroll
            for (int b = 0; b < 16; b++) h[b] = hist_s[b];
#pragma unroll
            for (int k = 0; k < MAXC; k++) {
                const int q = tid + 256 * k;
                const uint32_t bk = 15 - cst[k];
                uint32_t start = 0;
#pragma unroll
                for (int b = 0; b < 15; b++) start += (uint32_t)b < bk ? h[b] : 0u;
                if (q < nch) order_s[start + cslot[k]] = (uint16_t)(q | (cst[k] << 12));
            }
        }
        __syncthreads();
        // my chains: ranks tid, 511 - tid, 512 + tid, ... (snake order; the
        // ranks grow with c, so the valid ones are a prefix)
        int nm = 0;
#pragma unroll
        for (int c = 0; c < MAXC; c++) nm += ((c & 1) ? (c + 1) * 256 - 1 - tid : c * 256 + tid) < nch;
        // fetch the next layer's WOTS signature and this layer's auth path
        const bool more = layer + 1 < P::d;
        uint32_t nx[MAXW], au[MAXA];
#pragma unroll
        for (int k = 0; k < MAXW; k++) {
            const int u = tid + 256 * k;
            nx[k] = (more && u < nact * TLW) ? wots_of(u / TLW, layer + 1)[u % TLW] : 0u;
        }
#pragma unroll
        for (int k = 0; k < MAXA; k++) {
            const int u = tid + 256 * k;
            au[k] = u < nact * AW ? wots_of(u / AW, layer)[TLW + u % AW] : 0u;
        }
        // walk the chains: one flattened loop, one F site
        int c = 0, r = 0, j = 0;
        uint32_t s = P::w - 1;
        Key key;
        Adr wa;
        uint32_t v[NW];
        typename std::conditional<SHA2, int, typename Shake<P::n>::ChainIl>::type ch{};
        uint32_t pre[SHA2 ? 8 : 1];  // SHA-256: the chain's rounds 0-4 (Sha2::F_prefix)
#pragma unroll
        for (int w = 0; w < NW; w++) v[w] = 0;
#pragma unroll
        for (int i = 0; i < NW; i++) key.sk[i] = 0;
        auto take = [&](int cc) {  // start chain cc of this lane (everything from LDS)
            const int rank = (cc & 1) ? (cc + 1) * 256 - 1 - tid : cc * 256 + tid;
            const uint32_t e = order_s[rank];
            const int q = e & 0xFFF;
            r = q / LEN;
            j = q % LEN;
            s = P::w - 1 - (e >> 12);
#pragma unroll
            for (int w = 0; w < NW; w++) v[w] = to_native<P::hash>(sig_s[r][j * NW + w]);
            wa = adr_make(layer, tree_s[r], kWots);
            adr_keypair<TH>(wa, leaf_s[r]);
            adr_chain(wa, j);
            if constexpr (SHA2) {
#pragma unroll
                for (int i = 0; i < 8; i++) key.mid[i] = key_s[r][i];
                Hh::F_prefix(key, wa, pre);
            } else {
#pragma unroll
                for (int i = 0; i < NW; i++) key.pub[i] = key_s[r][i];
                Hh::chain_il_begin(key, wa, ch);
                uint32_t t[NW];
                Hh::to_il(v, t);
#pragma unroll
                for (int w = 0; w < NW; w++) v[w] = t[w];
            }
        };
        if (nm > 0) take(0);
#pragma unroll 1
        while (c < nm) {
            if (s >= (uint32_t)P::w - 1) {  // chain done: park its end, take the next
#pragma unroll
                for (int w = 0; w < NW; w++) pk_s[r][j * NW + w] = v[w];
                c++;
                if (c < nm) take(c);
                continue;
            }
            if constexpr (SHA2) {
                adr_hash(wa, s);
                Hh::F_step(key, pre, wa, v, v);
            } else {
                Hh::F_il(ch, s, v, v);
            }
            s++;
        }
#pragma unroll
        for (int k = 0; k < MAXA; k++) {
            const int u = tid + 256 * k;
            if (u < nact * AW) auth_s[u / AW][u % AW] = au[k];
        }
        __syncthreads();  // every chain end parked; this layer's sig_s reads done
#pragma unroll
        for (int k = 0; k < MAXW; k++) {
            const int u = tid + 256 * k;
            if (more && u < nact * TLW) sig_s[u / TLW][u % TLW] = nx[k];
        }
        if (tid < 16) hist_s[tid] = 0;
        // T_len and the subtree climb of request rq on the lane pair 2 rq,
        // 2 rq + 1 (sha256_compress_pair / keccak_f_il_pair)
        const int rq = tid >> 1;
        if constexpr (PRE) {
            // SHA-256: waves 1-3 run the message schedules of T_len's
            // payload-only blocks (they depend on the chain ends only) while
            // the pairs of wave 0 compress block 0; the pairs then take
            // K + W of those blocks from LDS
            using SP = Sha2Pair<NW, TLW>;
            uint32_t st[4];
            Key k2;
            uint64_t tree = 0;
            uint32_t leaf = 0;
            Adr pa{};
            const PkPair<P> rpk{&pk_s[rq < R ? rq : 0][0], tid & 1};
            if (tid < 2 * nact) {
                tree = tree_s[rq];
                leaf = leaf_s[rq];
#pragma unroll
                for (int i = 0; i < NW; i++) k2.sk[i] = 0;
#pragma unroll
                for (int i = 0; i < 8; i++) k2.mid[i] = key_s[rq][i];
                pa = adr_make(layer, tree, kWotsPk);
                adr_keypair<TH>(pa, leaf);
                SP::init(k2.mid, st);
                SP::absorb(st, 0, 1, pa, rpk);
            }
            if (tid >= 64) {  // consecutive lanes take consecutive requests
                for (int it = tid - 64; it < nact * MB; it += 192)
                    sha2_kw_block(&pk_s[it % nact][0], 1 + it / nact, kw_row(it % nact, it / nact));
            }
            __syncthreads();  // K + W of every payload-only block in LDS
            if (tid < 2 * nact) {
#pragma unroll 1
                for (int b = 0; b < MB; b++) sha256_rounds_pair_kw(st, kw_row(rq, b));
                SP::absorb(st, SP::FT, SP::B, pa, rpk);
                uint32_t node[NW];
                SP::output(st, node);
                climb_pair<P>(k2, layer, tree, leaf, &auth_s[rq][0], &node_s[rq][0], node);
                if ((tid & 1) == 0) {
#pragma unroll
                    for (int w = 0; w < NW; w++) root_s[rq][w] = node[w];
                    leaf_s[rq] = (uint32_t)(tree & ((1u << TH) - 1));
                    tree_s[rq] = tree >> TH;
                }
            }
        } else if (tid < 2 * nact) {
            const uint64_t tree = tree_s[rq];
            const uint32_t leaf = leaf_s[rq];
            Key k2;
#pragma unroll
            for (int i = 0; i < NW; i++) k2.sk[i] = 0;
            if constexpr (SHA2) {
#pragma unroll
                for (int i = 0; i < 8; i++) k2.mid[i] = key_s[rq][i];
            } else {
#pragma unroll
                for (int i = 0; i < NW; i++) k2.pub[i] = key_s[rq][i];
            }
            uint32_t node[NW];
            tlen_climb_pair<P>(k2, layer, tree, leaf, &pk_s[rq][0], &auth_s[rq][0], &node_s[rq][0],
                               node);
            if ((tid & 1) == 0) {
#pragma unroll
                for (int w = 0; w < NW; w++) root_s[rq][w] = node[w];
                leaf_s[rq] = (uint32_t)(tree & ((1u << TH) - 1));
                tree_s[rq] = tree >> TH;
            }
        }
    }
    __syncthreads();
    if (tid < nact) {
        const Plan& pl = A.plan[req0 + tid];
        bool ok = true;
#pragma unroll
        for (int w = 0; w < NW; w++) ok = ok && (root_s[tid][w] == pl.root[w]);
        A.status[req0 + tid] = ok ? 0 : -1;
    }
}

// Speculative verify for small batches, part 1: every WOTS chain of every
// layer from every possible start.  The start of chain j in layer L is the
// digit d_j of layer L-1's root, known only after that layer, so the chains
// of all 16 starts are computed up front: cap[req][L][j][d] = F_14(..F_d(sig_j))
// (wots_pk_from_sig, wots.cu:777-790, with d = 15 the signature value
// itself).  Lane p of a chain walks the starts p and 15 - p, 15 F steps for
// every lane.  The per-layer pass that remains (k_vht_spec) is T_len and the
// climb only.  SHAKE256 values are stored bit-interleaved.
template <class P>
__global__ void __launch_bounds__(256) k_vspec(BatchArgs A) {
    using Hh = HashT<P>;
    using Key = typename Hh::Key;
    constexpr bool SHA2 = P::hash == kSha2;
    constexpr int NW = P::nw;
    const uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= (uint64_t)A.n_req * P::d * P::len * 8) return;
    const int p = (int)(g & 7);
    const uint64_t c = g >> 3;                       // (req * d + layer) * len + j
    const int j = (int)(c % P::len);
    const uint32_t layer = (uint32_t)((c / P::len) % P::d);
    const uint32_t req = (uint32_t)(c / ((uint64_t)P::len * P::d));
    const Plan& pl = A.plan[req];
    Key key;
    Hh::load_key(pl, key);
    const uint64_t tree = layer_tree<P>(pl.tree, layer);
    const uint32_t leaf = layer == 0 ? pl.leaf
                                     : (uint32_t)((pl.tree >> ((layer - 1) * P::th)) &
                                                  ((1u << P::th) - 1));
    Adr wa = adr_make(layer, tree, kWots);
    adr_keypair<P::th>(wa, leaf);
    adr_chain(wa, j);
    uint32_t sg[NW], v[NW];
    load_val<P::hash, NW>(A.sig + (size_t)req * A.sig_stride + P::n + P::fors_bytes +
                              (size_t)layer * P::layer_bytes + (size_t)j * P::n,
                          sg);
    typename std::conditional<SHA2, int, typename Shake<P::n>::ChainIl>::type ch{};
    if constexpr (!SHA2) {
        Hh::chain_il_begin(key, wa, ch);
        uint32_t t[NW];
        Hh::to_il(sg, t);
#pragma unroll
        for (int w = 0; w < NW; w++) sg[w] = t[w];
    }
    uint32_t* out = reinterpret_cast<uint32_t*>(A.cap) + c * 16 * NW;
    if (p == 0) {  // start 15: no step
#pragma unroll
        for (int w = 0; w < NW; w++) out[15 * NW + w] = sg[w];
    }
#pragma unroll
    for (int w = 0; w < NW; w++) v[w] = sg[w];
    uint32_t s = p;     // chain A: start p, 15 - p steps; then chain B: start 15 - p, p steps
    bool second = false;
#pragma unroll 1
    for (int it = 0; it < 15; it++) {
        if constexpr (SHA2) {
            adr_hash(wa, s);
            Hh::F(key, wa, v, v);
        } else {
            Hh::F_il(ch, s, v, v);
        }
        s++;
        if (s == (uint32_t)P::w - 1) {
            const int d = second ? 15 - p : p;
#pragma unroll
            for (int w = 0; w < NW; w++) out[d * NW + w] = v[w];
            second = true;
            s = 15 - p;
#pragma unroll
            for (int w = 0; w < NW; w++) v[w] = sg[w];
        }
    }
}

// Speculative verify, part 2: one wave per request walks the d layers
// (sign.cu:3077-3113): lanes pick each chain's end for the layer's message
// from the candidates of k_vspec and load the auth path, then lanes 0 and 1
// run T_len and compute_root (utils.cu:94-135) as a lane pair (SHA-256: with
// the message schedules of T_len's payload-only blocks from lanes 2..).  The serial
// work per layer is T_len + h' nodes instead of the longest chain + T_len +
// h' nodes.
template <class P>
__global__ void __launch_bounds__(64) k_vht_spec(BatchArgs A) {
    using Hh = HashT<P>;
    using Key = typename Hh::Key;
    constexpr int NW = P::nw;
    constexpr int LEN = P::len;
    constexpr int TH = P::th;
    constexpr int TLW = LEN * NW;
    constexpr int AW = TH * NW;
    __shared__ uint32_t pk_s[TLW];    // chain ends (SHAKE256: interleaved)
    __shared__ uint32_t auth_s[AW];   // auth path, raw words
    __shared__ uint32_t root_s[NW];
    __shared__ uint32_t node_s[NW];   // climb: the node, raw words
    // SHA-256: K + W of T_len's payload-only blocks
    constexpr int MB = P::hash == kSha2 ? (TLW + 5) / 16 - 1 : 1;
    static_assert(MB <= 62, "one lane per block");
    __shared__ uint32_t kw_s[MB][64];
    const int tid = threadIdx.x;
    const uint32_t req = blockIdx.x;  // grid = n_req
    const Plan& pl = A.plan[req];
    const uint8_t* sig = A.sig + (size_t)req * A.sig_stride;
    const uint32_t* cand = reinterpret_cast<const uint32_t*>(A.cap) + (size_t)req * P::d * LEN * 16 * NW;
    if (tid == 0) {
        uint32_t fp[NW];
        load_val<P::hash, NW>(A.fors_pk + (size_t)req * P::n, fp);
#pragma unroll
        for (int w = 0; w < NW; w++) root_s[w] = fp[w];
    }
    uint64_t tree = pl.tree;
    uint32_t leaf = pl.leaf;
#pragma unroll 1
    for (int layer = 0; layer < P::d; layer++) {
        __syncthreads();
        uint32_t msg[NW];
#pragma unroll
        for (int w = 0; w < NW; w++) msg[w] = root_s[w];
        const uint32_t* lc = cand + (size_t)layer * LEN * 16 * NW;
#pragma unroll 1
        for (int j = tid; j < LEN; j += 64) {
            const uint32_t d = chain_len<P>(msg, j);
#pragma unroll
            for (int w = 0; w < NW; w++) pk_s[j * NW + w] = lc[((size_t)j * 16 + d) * NW + w];
        }
        const uint32_t* auth = reinterpret_cast<const uint32_t*>(
            sig + P::n + P::fors_bytes + (size_t)layer * P::layer_bytes + P::wots_bytes);
        for (int u = tid; u < AW; u += 64) auth_s[u] = auth[u];
        __syncthreads();
        uint32_t node[NW];
        if constexpr (P::hash == kSha2) {
            // lanes 2.. compute K + W of T_len's payload-only blocks (one
            // pass for all of them), then the pair runs only their rounds
            if (tid >= 2 && tid < 2 + MB) sha2_kw_block(pk_s, tid - 1, kw_s[tid - 2]);
            __syncthreads();
            if (tid < 2) {  // T_len and the climb on the lane pair 0, 1
                using SP = Sha2Pair<NW, TLW>;
                Key key;
                Hh::load_key(pl, key);
                Adr pa = adr_make(layer, tree, kWotsPk);
                adr_keypair<TH>(pa, leaf);
                const PkPair<P> rpk{pk_s, tid};
                uint32_t st[4];
                SP::init(key.mid, st);
                SP::absorb(st, 0, 1, pa, rpk);
#pragma unroll 1
                for (int b = 0; b < MB; b++) sha256_rounds_pair_kw(st, kw_s[b]);
                SP::absorb(st, SP::FT, SP::B, pa, rpk);
                SP::output(st, node);
                climb_pair<P>(key, layer, tree, leaf, auth_s, node_s, node);
            }
        } else if (tid < 2) {  // T_len and the climb on the lane pair 0, 1
            Key key;
            Hh::load_key(pl, key);
            tlen_climb_pair<P>(key, layer, tree, leaf, pk_s, auth_s, node_s, node);
        }
        if (tid == 0) {
#pragma unroll
            for (int w = 0; w < NW; w++) root_s[w] = node[w];
        }
        leaf = (uint32_t)(tree & ((1u << TH) - 1));
        tree >>= TH;
    }
    __syncthreads();
    if (tid == 0) {
        bool ok = true;
#pragma unroll
        for (int w = 0; w < NW; w++) ok = ok && (root_s[w] == pl.root[w]);
        A.status[req] = ok ? 0 : -1;
    }
}

}  // namespace spx
