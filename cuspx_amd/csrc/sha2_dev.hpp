// sha2_dev.hpp -- SHA-256 tweakable-hash family for CDNA4 (gfx950).
//
// Semantics follow the reference host path:
//   thash  T_l(x) = SHA256(mid(pub_seed) ; addr[0..22) || x)[:n]   thash_sha256.cu:16-34
//   PRF(a)        = SHA256(sk_seed || addr[0..22))[:n]             hash_sha256.cu:65-76
//   PRF_msg       = HMAC-SHA256(sk_prf, optrand || m)[:n]          hash_sha256.cu:99-144
//   H_msg         = MGF1(SHA256(R || pk || m), dgst)               hash_sha256.cu:200-262
// where mid(pub_seed) is the state after the single block pub_seed||0^(64-n)
// (seed_state, sha256.cu:391-405) -- computed once per key and carried in
// the per-request plan, never shared between keys.
//
// Values are big-endian words.  The 22-byte address puts every payload at a
// 2-byte offset inside the block, so payload words are re-aligned with
// v_alignbit (one op per word).  The compression is fully unrolled with
// v_alignbit rotates and v_bitop3 for Ch / Maj / 3-way xor: 14 VALU ops per
// round and 10 per schedule word.
#pragma once
#include "spx_common.hpp"

namespace spx {

static constexpr uint32_t kSha256K[64] = {
    0x428a2f98, 0x71374491, 0xb5c0fbcf, 0xe9b5dba5, 0x3956c25b, 0x59f111f1, 0x923f82a4, 0xab1c5ed5,
    0xd807aa98, 0x12835b01, 0x243185be, 0x550c7dc3, 0x72be5d74, 0x80deb1fe, 0x9bdc06a7, 0xc19bf174,
    0xe49b69c1, 0xefbe4786, 0x0fc19dc6, 0x240ca1cc, 0x2de92c6f, 0x4a7484aa, 0x5cb0a9dc, 0x76f988da,
    0x983e5152, 0xa831c66d, 0xb00327c8, 0xbf597fc7, 0xc6e00bf3, 0xd5a79147, 0x06ca6351, 0x14292967,
    0x27b70a85, 0x2e1b2138, 0x4d2c6dfc, 0x53380d13, 0x650a7354, 0x766a0abb, 0x81c2c92e, 0x92722c85,
    0xa2bfe8a1, 0xa81a664b, 0xc24b8b70, 0xc76c51a3, 0xd192e819, 0xd6990624, 0xf40e3585, 0x106aa070,
    0x19a4c116, 0x1e376c08, 0x2748774c, 0x34b0bcb5, 0x391c0cb3, 0x4ed8aa4a, 0x5b9cca4f, 0x682e6ff3,
    0x748f82ee, 0x78a5636f, 0x84c87814, 0x8cc70208, 0x90befffa, 0xa4506ceb, 0xbef9a3f7, 0xc67178f2};

static constexpr uint32_t kSha256IV[8] = {0x6a09e667, 0xbb67ae85, 0x3c6ef372, 0xa54ff53a,
                                          0x510e527f, 0x9b05688c, 0x1f83d9ab, 0x5be0cd19};

// One FIPS 180-4 compression.  w[] is used as the rolling schedule.
__device__ __forceinline__ void sha256_compress(uint32_t st[8], uint32_t w[16]) {
    uint32_t a = st[0], b = st[1], c = st[2], d = st[3];
    uint32_t e = st[4], f = st[5], g = st[6], h = st[7];
#pragma unroll
    for (int i = 0; i < 64; i++) {
        uint32_t wi;
        if (i < 16) {
            wi = w[i];
        } else {
            uint32_t x = w[(i - 15) & 15], y = w[(i - 2) & 15];
            uint32_t s0 = xor3(rotr32(x, 7), rotr32(x, 18), x >> 3);
            uint32_t s1 = xor3(rotr32(y, 17), rotr32(y, 19), y >> 10);
            wi = s1 + w[(i - 7) & 15] + s0 + w[i & 15];
            w[i & 15] = wi;
        }
        uint32_t t1 = h + xor3(rotr32(e, 6), rotr32(e, 11), rotr32(e, 25)) + bit_ch(e, f, g) +
                      kSha256K[i] + wi;
        uint32_t t2 = xor3(rotr32(a, 2), rotr32(a, 13), rotr32(a, 22)) + bit_maj(a, b, c);
        h = g;
        g = f;
        f = e;
        e = d + t1;
        d = c;
        c = b;
        b = a;
        a = t1 + t2;
    }
    st[0] += a;
    st[1] += b;
    st[2] += c;
    st[3] += d;
    st[4] += e;
    st[5] += f;
    st[6] += g;
    st[7] += h;
}

// Rounds [R0, R1) of a compression from the working variables v (a..h after
// round R0 - 1, or the chaining value for R0 = 0); w holds the whole block,
// whose rolling schedule needs W[0..16) whatever R0 is.
template <int R0, int R1>
__device__ __forceinline__ void sha256_rounds(uint32_t v[8], uint32_t w[16]) {
    uint32_t a = v[0], b = v[1], c = v[2], d = v[3], e = v[4], f = v[5], g = v[6], h = v[7];
#pragma unroll
    for (int i = R0; i < R1; i++) {
        uint32_t wi;
        if (i < 16) {
            wi = w[i];
        } else {
            uint32_t x = w[(i - 15) & 15], y = w[(i - 2) & 15];
            uint32_t s0 = xor3(rotr32(x, 7), rotr32(x, 18), x >> 3);
            uint32_t s1 = xor3(rotr32(y, 17), rotr32(y, 19), y >> 10);
            wi = s1 + w[(i - 7) & 15] + s0 + w[i & 15];
            w[i & 15] = wi;
        }
        uint32_t t1 = h + xor3(rotr32(e, 6), rotr32(e, 11), rotr32(e, 25)) + bit_ch(e, f, g) +
                      kSha256K[i] + wi;
        uint32_t t2 = xor3(rotr32(a, 2), rotr32(a, 13), rotr32(a, 22)) + bit_maj(a, b, c);
        h = g;
        g = f;
        f = e;
        e = d + t1;
        d = c;
        c = b;
        b = a;
        a = t1 + t2;
    }
    v[0] = a;
    v[1] = b;
    v[2] = c;
    v[3] = d;
    v[4] = e;
    v[5] = f;
    v[6] = g;
    v[7] = h;
}

// ------------------------------------------------------------ lane pairs
//
// One compression on a lane pair (lanes 2i and 2i+1 of a wave, both active,
// same chaining value and message): the even lane holds e, f, g, h and the
// odd lane a, b, c, d in s[0..3]; the message is held as word pairs v[k] =
// W[2k] (even lane) / W[2k+1] (odd lane).  Each round is one instruction
// stream: Sigma1(e) and Ch(e, f, g) on the even lane, Sigma0(a) and
// Maj(a, b, c) = Ch(~(a ^ b), b, c) on the odd one (per-lane rotate counts),
// then a quad_perm DPP swap gives the odd lane T1 and the even lane d.  The
// schedule makes a word pair per step.  11 instructions per round and 14
// per schedule pair against 14 and 20 on one lane: for serial hashes whose
// lane would otherwise run alone (T_len and Merkle climbs of verification,
// the parked T_len of signing, the message hashes of the prep kernels).

__device__ __forceinline__ void sha256_compress_pair(uint32_t s[4], const uint32_t vin[8]) {
    const bool odd = threadIdx.x & 1u;
    const uint32_t m = odd ? ~0u : 0u;
    const uint32_t r1 = odd ? 2u : 6u, r2 = odd ? 13u : 11u, r3 = odd ? 22u : 25u;
    uint32_t v[8], z[8];  // pairs k - 8 .. k - 1 (index k & 7) and their swaps
#pragma unroll
    for (int k = 0; k < 8; k++) {
        v[k] = vin[k];
        z[k] = pair_swap(vin[k]);
    }
    uint32_t q0 = s[0], q1 = s[1], q2 = s[2], q3 = s[3];
#pragma unroll
    for (int t = 0; t < 64; t++) {
        if (t >= 16 && (t & 1) == 0) {
            // W[2k] = s1(W[2k-2]) + W[2k-7] + s0(W[2k-15]) + W[2k-16] (even lane)
            // W[2k+1] = s1(W[2k-1]) + W[2k-6] + s0(W[2k-14]) + W[2k-15] (odd lane)
            const int k = t >> 1;
            const uint32_t y = v[(k - 1) & 7];
            const uint32_t x = lane_sel(m, z[(k - 7) & 7], z[k & 7]);
            const uint32_t w7 = lane_sel(m, z[(k - 3) & 7], z[(k - 4) & 7]);
            const uint32_t s1 = xor3(rotr32(y, 17), rotr32(y, 19), y >> 10);
            const uint32_t s0 = xor3(rotr32(x, 7), rotr32(x, 18), x >> 3);
            const uint32_t nv = s1 + w7 + s0 + v[k & 7];
            v[k & 7] = nv;
            z[k & 7] = pair_swap(nv);
        }
        const uint32_t wt = (t & 1) ? z[(t >> 1) & 7] : v[(t >> 1) & 7];  // W[t] on the even lane
        const uint32_t hk = (q3 + kSha256K[t] + wt) & ~m;                 // h + K + W | 0
        const uint32_t sg = xor3(__builtin_amdgcn_alignbit(q0, q0, r1),
                                 __builtin_amdgcn_alignbit(q0, q0, r2),
                                 __builtin_amdgcn_alignbit(q0, q0, r3));
        const uint32_t fn = bit_ch(bit_chi(q0, q1, m), q1, q2);  // Ch(e,f,g) | Maj(a,b,c)
        const uint32_t tt = sg + fn + hk;                          // T1 | T2
        const uint32_t nw = tt + pair_swap(lane_sel(m, q3, tt));   // d + T1 | T1 + T2
        q3 = q2;
        q2 = q1;
        q1 = q0;
        q0 = nw;
    }
    s[0] += q0;
    s[1] += q1;
    s[2] += q2;
    s[3] += q3;
}

// a word pointer into LDS (address space 3), so that code reached through a
// call still accesses it with ds_* instructions rather than flat ones
typedef uint32_t __attribute__((address_space(3))) lds_u32;
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
typedef u32x4 __attribute__((address_space(3))) lds_u32x4;

// the 64 rounds of sha256_compress_pair with K[t] + W[t] precomputed (kw,
// read by both lanes; the odd lane masks it out)
// The 64 rounds of a lane pair with K[t] + W[t] precomputed (kw, read by both
// lanes), with the odd (a-side) lane two rounds behind the even (e-side) one.
// In sha256_compress_pair every round waits for its exchange: e' = d + T1
// needs the odd lane's d and a' = T1 + T2 the even lane's T1, so each round
// is alignbit, xor3, add3, select, DPP swap and add in a row.  Here each
// lane's own recurrence is alignbit, xor3, add3:
//   even, iteration i:  E[i+1] = S1(E[i]) + Ch(E[i], E[i-1], E[i-2]) + pre[i],
//                       pre[i] = E[i-3] + KW[i] + A[i-3]
//   odd,  iteration i:  A[i-1] = S0(A[i-2]) + Maj(A[i-2], A[i-3], A[i-4]) + T1[i-2],
//                       T1[i-2] = E[i-1] - A[i-5]
// and the swap of iteration i (each lane sends its new value) is first used
// in iteration i + 1, to form pre[i+1] (even) and T1[i-1] (odd).  66
// iterations; the odd lane's outputs of the first two are b and a, the even
// lane's state is taken after iteration 63.  K + W is read about 16 rounds
// ahead of its use (two groups of 16 registers).
template <class KW>
__device__ __forceinline__ void sha256_rounds_pair_kw(uint32_t s[4], KW kw) {
    const bool odd = threadIdx.x & 1u;
    const uint32_t m = odd ? ~0u : 0u;
    const uint32_t r1 = odd ? 2u : 6u, r2 = odd ? 13u : 11u, r3 = odd ? 22u : 25u;
    uint32_t c0[16], c1[16];  // K + W [16g, 16g + 16) and the group after it
#pragma unroll
    for (int j = 0; j < 16; j++) c0[j] = kw[j];
#pragma unroll
    for (int j = 0; j < 16; j++) c1[j] = kw[16 + j];
    // even: e f g h; odd: a b c d.  The odd lane starts at c, d; the even
    // lane needs d (pre[0]) and c (A[-2], for pre[1]) from it.
    const uint32_t d_ev = pair_swap(s[3]), c_ev = pair_swap(s[2]);
    const uint32_t a_od = s[0], b_od = s[1];
    uint32_t q0 = lane_sel(m, s[2], s[0]), q1 = lane_sel(m, s[3], s[1]), q2 = s[2], q3 = s[3];
    uint32_t x = q3 + c0[0] + d_ev;  // even: pre[0]; odd: unused
    uint32_t rp = c_ev;              // the previous iteration's swap
    uint32_t e0 = 0, e1 = 0, e2 = 0, e3 = 0;
#pragma unroll
    for (int i = 0; i < 66; i++) {
        if (i == 15 || i == 31 || i == 47) {  // next group of K + W
#pragma unroll
            for (int j = 0; j < 16; j++) c0[j] = c1[j];
            if (i < 47) {
#pragma unroll
                for (int j = 0; j < 16; j++) c1[j] = kw[i + 17 + j];
            }
        }
        const uint32_t sg = xor3(__builtin_amdgcn_alignbit(q0, q0, r1),
                                 __builtin_amdgcn_alignbit(q0, q0, r2),
                                 __builtin_amdgcn_alignbit(q0, q0, r3));
        const uint32_t fn = bit_ch(bit_chi(q0, q1, m), q1, q2);  // Ch | Maj
        uint32_t nw = sg + fn + x;                                // E[i+1] | A[i-1]
        if (i == 0) nw = lane_sel(m, b_od, nw);
        if (i == 1) nw = lane_sel(m, a_od, nw);
        const uint32_t r = pair_swap(nw);
        const uint32_t kn = i + 1 < 64 ? c0[(i + 1) & 15] : 0u;
        x = lane_sel(m, rp - q2, q2 + kn + rp);  // T1[i-1] | pre[i+1]
        q3 = q2;
        q2 = q1;
        q1 = q0;
        q0 = nw;
        rp = r;
        if (i == 63) {
            e0 = q0;
            e1 = q1;
            e2 = q2;
            e3 = q3;
        }
    }
    s[0] += lane_sel(m, q0, e0);
    s[1] += lane_sel(m, q1, e1);
    s[2] += lane_sel(m, q2, e2);
    s[3] += lane_sel(m, q3, e3);
}

// sha256_compress on a lane pair whose lanes both hold the whole chaining
// value st and block w (the message-hash paths): the halves are split for
// sha256_compress_pair and st is whole again in both lanes afterwards
__device__ __forceinline__ void sha256_compress_dup(uint32_t st[8], const uint32_t w[16]) {
    const uint32_t m = (threadIdx.x & 1u) ? ~0u : 0u;
    uint32_t s4[4], v[8];
#pragma unroll
    for (int i = 0; i < 4; i++) s4[i] = lane_sel(m, st[i], st[4 + i]);
#pragma unroll
    for (int k = 0; k < 8; k++) v[k] = lane_sel(m, w[2 * k + 1], w[2 * k]);
    sha256_compress_pair(s4, v);
#pragma unroll
    for (int i = 0; i < 4; i++) {
        const uint32_t other = pair_swap(s4[i]);
        st[i] = lane_sel(m, s4[i], other);
        st[4 + i] = lane_sel(m, other, s4[i]);
    }
}

// sha256_compress_dup with K[t] + W[t] precomputed (kw, LDS): the message
// hashes' blocks made only of message words, whose schedules another wave
// computed ahead
__device__ __forceinline__ void sha256_rounds_dup_kw(uint32_t st[8], const lds_u32* kw) {
    const uint32_t m = (threadIdx.x & 1u) ? ~0u : 0u;
    uint32_t s4[4];
#pragma unroll
    for (int i = 0; i < 4; i++) s4[i] = lane_sel(m, st[i], st[4 + i]);
    sha256_rounds_pair_kw(s4, kw);
#pragma unroll
    for (int i = 0; i < 4; i++) {
        const uint32_t other = pair_swap(s4[i]);
        st[i] = lane_sel(m, s4[i], other);
        st[4 + i] = lane_sel(m, other, s4[i]);
    }
}

// K[t] + W[t], t < 64, of the block whose words are bswap32(src[0..16))
// (a block made only of message words) into kw[0..64), 16-byte aligned
__device__ __forceinline__ void sha256_kw_msg_block(const uint32_t* src, lds_u32* kw) {
    const gbl_u32* g = (gbl_u32*)src;  // address-space cast
    uint32_t w[16];
#pragma unroll
    for (int i = 0; i < 16; i++) w[i] = bswap32(g[i]);
#pragma unroll
    for (int q = 0; q < 16; q++) {
        uint32_t o[4];
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const int t = 4 * q + j;
            if (t >= 16) {
                const uint32_t x = w[(t - 15) & 15], y = w[(t - 2) & 15];
                w[t & 15] = xor3(rotr32(y, 17), rotr32(y, 19), y >> 10) + w[(t - 7) & 15] +
                            xor3(rotr32(x, 7), rotr32(x, 18), x >> 3) + w[t & 15];
            }
            o[j] = kSha256K[t] + w[t & 15];
        }
        *reinterpret_cast<lds_u32x4*>(kw + 4 * q) = u32x4{o[0], o[1], o[2], o[3]};
    }
}

// one compression: on one lane, or on a lane pair holding identical inputs
template <bool PAIR>
__device__ __forceinline__ void sha256_compress_by(uint32_t st[8], uint32_t w[16]) {
    if constexpr (PAIR) sha256_compress_dup(st, w);
    else sha256_compress(st, w);
}

// write x into b[q] for a wave-uniform runtime q (v_cndmask per slot)
__device__ __forceinline__ void put16(uint32_t b[16], int q, uint32_t x) {
#pragma unroll
    for (int i = 0; i < 16; i++) b[i] = (q == i) ? x : b[i];
}

template <int N>
struct Sha2 {
    static constexpr int hash = kSha2;
    static constexpr int NW = N / 4;

    struct Key {
        uint32_t mid[8];
        uint32_t sk[NW];
    };

    __device__ static __forceinline__ void load_key(const Plan& p, Key& k) {
#pragma unroll
        for (int i = 0; i < 8; i++) k.mid[i] = p.mid[i];
#pragma unroll
        for (int i = 0; i < NW; i++) k.sk[i] = p.skseed[i];
    }

    // Word i of compression block blk of T over M payload words d after the
    // 22-byte address (the payload sits at a 2-byte offset: one v_alignbit
    // per word); blk and i may be runtime values.
    template <int M>
    __device__ static __forceinline__ uint32_t block_word(const Adr& a, const uint32_t d[M],
                                                          int blk, int i) {
        constexpr int P = 6 + M;                      // words up to and incl. the 0x80 byte
        constexpr int B = (P + 2 + 15) / 16;          // compression blocks
        constexpr uint32_t bits = (64u + 22u + 4u * M) * 8u;
        const int g = 16 * blk + i;
        if (g < 5) return a.w[g];
        if (g == 5) return a.w[5] | (d[0] >> 16);
        if (g < 5 + M) return __builtin_amdgcn_alignbit(d[g - 6], d[g - 5], 16);
        if (g == 5 + M) return (d[M - 1] << 16) | 0x8000u;
        if (g == 16 * B - 1) return bits;
        return 0;
    }

    // T over M payload words after the 22-byte address, all block indices
    // known at compile time (used for F, H and short T_l).
    template <int M>
    __device__ static __forceinline__ void thash_words(const Key& k, const Adr& a,
                                                       const uint32_t d[M], uint32_t out[NW]) {
        constexpr int B = (6 + M + 2 + 15) / 16;
        uint32_t st[8];
#pragma unroll
        for (int i = 0; i < 8; i++) st[i] = k.mid[i];
#pragma unroll
        for (int blk = 0; blk < B; blk++) {
            uint32_t w[16];
#pragma unroll
            for (int i = 0; i < 16; i++) w[i] = block_word<M>(a, d, blk, i);
            sha256_compress(st, w);
        }
#pragma unroll
        for (int i = 0; i < NW; i++) out[i] = st[i];
    }

    // F: one input value (wots.cu gen_chain step, fors leaf)
    __device__ static __forceinline__ void F(const Key& k, const Adr& a, const uint32_t in[NW],
                                             uint32_t out[NW]) {
        thash_words<NW>(k, a, in, out);
    }

    // F along one chain: its block's first five words (the address up to the
    // chain index; the hash index sits in word 5) do not change from step to
    // step, so rounds 0-4 run once per chain (F_prefix) and each step runs
    // rounds 5-63 from there (F_step).  F is one block for every n.
    __device__ static __forceinline__ void F_prefix(const Key& k, const Adr& a, uint32_t pre[8]) {
        uint32_t w[16];
#pragma unroll
        for (int i = 0; i < 16; i++) w[i] = i < 5 ? a.w[i] : 0u;
#pragma unroll
        for (int i = 0; i < 8; i++) pre[i] = k.mid[i];
        sha256_rounds<0, 5>(pre, w);
    }
    __device__ static __forceinline__ void F_step(const Key& k, const uint32_t pre[8],
                                                  const Adr& a, const uint32_t in[NW],
                                                  uint32_t out[NW]) {
        static_assert((6 + NW + 2 + 15) / 16 == 1, "F is one block");
        uint32_t w[16], v[8];
#pragma unroll
        for (int i = 0; i < 16; i++) w[i] = block_word<NW>(a, in, 0, i);
#pragma unroll
        for (int i = 0; i < 8; i++) v[i] = pre[i];
        sha256_rounds<5, 64>(v, w);
#pragma unroll
        for (int i = 0; i < NW; i++) out[i] = k.mid[i] + v[i];
    }

    // H: two input values (Merkle node)
    __device__ static __forceinline__ void H(const Key& k, const Adr& a, const uint32_t l[NW],
                                             const uint32_t r[NW], uint32_t out[NW]) {
        uint32_t d[2 * NW];
#pragma unroll
        for (int i = 0; i < NW; i++) {
            d[i] = l[i];
            d[NW + i] = r[i];
        }
        thash_words<2 * NW>(k, a, d, out);
    }

    // PRF(sk_seed, addr) from the IV: sk_seed || addr22 || pad, one block
    __device__ static __forceinline__ void PRF(const Key& k, const Adr& a, uint32_t out[NW]) {
        uint32_t w[16], st[8];
#pragma unroll
        for (int i = 0; i < 16; i++) {
            uint32_t x;
            if (i < NW) x = k.sk[i];
            else if (i < NW + 5) x = a.w[i - NW];
            else if (i == NW + 5) x = a.w[5] | 0x8000u;
            else if (i == 15) x = (uint32_t)(N + 22) * 8u;
            else x = 0;
            w[i] = x;
        }
#pragma unroll
        for (int i = 0; i < 8; i++) st[i] = kSha256IV[i];
        sha256_compress(st, w);
#pragma unroll
        for (int i = 0; i < NW; i++) out[i] = st[i];
    }

    // Streaming T_l: items of NW words appended one word at a time with a
    // wave-uniform position; used for T_len (leaf) and T_k (FORS pk).
    struct Stream {
        uint32_t st[8];
        uint32_t b[16];
        int q;
        uint32_t carry;

        __device__ __forceinline__ void begin(const Key& k, const Adr& a) {
#pragma unroll
            for (int i = 0; i < 8; i++) st[i] = k.mid[i];
#pragma unroll
            for (int i = 0; i < 16; i++) b[i] = i < 5 ? a.w[i] : 0u;
            q = 5;
            carry = a.w[5] >> 16;
        }
        // append one payload word (must be called from a single site in a loop)
        __device__ __forceinline__ void word(uint32_t dw) {
            put16(b, q, __builtin_amdgcn_alignbit(carry, dw, 16));
            carry = dw;
            q++;
            if (q == 16) {
                sha256_compress(st, b);
                q = 0;
            }
        }
        // D = total payload words appended; finishes with compile-time layout
        template <int D>
        __device__ __forceinline__ void finish(uint32_t out[NW]) {
            constexpr int qe = (5 + D) % 16;
            constexpr uint32_t bits = (64u + 22u + 4u * D) * 8u;
            b[qe] = (carry << 16) | 0x8000u;
#pragma unroll
            for (int i = qe + 1; i < 16; i++) b[i] = 0;
            if (qe <= 13) {
                b[15] = bits;
                sha256_compress(st, b);
            } else {
                sha256_compress(st, b);
#pragma unroll
                for (int i = 0; i < 15; i++) b[i] = 0;
                b[15] = bits;
                sha256_compress(st, b);
            }
#pragma unroll
            for (int i = 0; i < NW; i++) out[i] = st[i];
        }
    };

    // ------------------------------------------------ per-request hashing

    // midstate: compress(IV, pub_seed || 0^(64-n)) (seed_state, sha256.cu:391)
    // PAIR: run by a lane pair holding identical inputs (the prep kernels)
    template <bool PAIR = false>
    __device__ static __forceinline__ void midstate(const uint32_t pub[NW], uint32_t mid[8]) {
        uint32_t w[16];
#pragma unroll
        for (int i = 0; i < 16; i++) w[i] = i < NW ? pub[i] : 0u;
#pragma unroll
        for (int i = 0; i < 8; i++) mid[i] = kSha256IV[i];
        sha256_compress_by<PAIR>(mid, w);
    }

    // The message hashes run on a lane pair holding identical inputs (the
    // prep kernels give every request two lanes; sha256_compress_dup).
    //
    // SHA-256 continuing from st over (pre[0..NPRE) || m[0..mlen)), where
    // `before` bytes were already absorbed.  m is 4-byte aligned; bytes past
    // mlen are never read beyond the word holding the last byte.  Blocks made
    // only of whole message words are loaded as blocks (stream_full, or
    // stream_msg_long's producing wave); the head (prefix) and the tail (last
    // partial word, padding, length) go through stream_range.
    //
    // stream_range runs the blocks of words [kk, kk_end) of that stream (kk a
    // multiple of 16; kk_end clipped to the padded end).  Each block's 16
    // words are formed in one unrolled pass, so the loads of its message
    // words are issued together (word by word, each load stood on the serial
    // path: a short message's two passes cost one load latency per word).
    template <int NPRE>
    __device__ static void stream_range(uint32_t st[8], const uint32_t pre[NPRE],
                                        const uint8_t* m, uint32_t mlen, uint32_t before,
                                        uint32_t kk, uint32_t kk_end) {
        static_assert(NPRE <= 32, "the prefix ends within the first two blocks");
        const uint64_t bits = ((uint64_t)before + 4u * NPRE + mlen) * 8u;
        const uint32_t ndata = NPRE + (mlen + 4) / 4;  // words holding data or 0x80
        const uint32_t total = ((ndata + 2 + 15) / 16) * 16;
        if (kk_end > total) kk_end = total;
        const gbl_u32* mw = (gbl_u32*)m;  // address-space cast
        uint32_t sl[8];  // the chaining value stays in registers
#pragma unroll
        for (int i = 0; i < 8; i++) sl[i] = st[i];
#pragma unroll 1
        for (; kk < kk_end; kk += 16) {
            uint32_t b[16];
#pragma unroll
            for (int j = 0; j < 16; j++) {
                const uint32_t w = kk + (uint32_t)j;
                uint32_t x;
                if (w < (uint32_t)NPRE) {
                    // the prefix lies in block 0 (kk = 0) or block 1 (kk = 16)
                    uint32_t pv = j < NPRE ? pre[j < NPRE ? j : 0] : 0u;
                    if (16 + j < NPRE) pv = kk ? pre[16 + j < NPRE ? 16 + j : 0] : pv;
                    x = pv;
                } else if (w < ndata) {
                    const uint32_t i = w - NPRE;
                    const int valid = (int)mlen - 4 * (int)i;  // 0 .. mlen
                    const uint32_t raw = valid > 0 ? bswap32(mw[i]) : 0u;
                    if (valid >= 4) {
                        x = raw;
                    } else {
                        const uint32_t keep =
                            valid <= 0 ? 0u : (0xFFFFFFFFu << (8 * (4 - valid)));
                        x = (raw & keep) | (0x80000000u >> (8 * (valid < 0 ? 0 : valid)));
                    }
                } else if (w == total - 2) {
                    x = (uint32_t)(bits >> 32);
                } else if (w == total - 1) {
                    x = (uint32_t)bits;
                } else {
                    x = 0;
                }
                b[j] = x;
            }
            sha256_compress_dup(sl, b);
        }
#pragma unroll
        for (int i = 0; i < 8; i++) st[i] = sl[i];
    }

    // nfb blocks made only of message words, block b = src[16 b .. 16 b + 16),
    // on the lane pair; block b + 1 is loaded while block b compresses
    __device__ static void stream_full(uint32_t st[8], const uint32_t* src, uint32_t nfb) {
        const gbl_u32* g = (gbl_u32*)src;  // address-space cast
        uint32_t sl[8], cur[16], nxt[16];
#pragma unroll
        for (int i = 0; i < 8; i++) sl[i] = st[i];
#pragma unroll
        for (int i = 0; i < 16; i++) cur[i] = g[i];
#pragma unroll 1
        for (uint32_t b = 0; b < nfb; b++) {
            if (b + 1 < nfb) {
#pragma unroll
                for (int i = 0; i < 16; i++) nxt[i] = g[16 * (b + 1) + i];
            }
            uint32_t w[16];
#pragma unroll
            for (int i = 0; i < 16; i++) w[i] = bswap32(cur[i]);
            sha256_compress_dup(sl, w);
#pragma unroll
            for (int i = 0; i < 16; i++) cur[i] = nxt[i];
        }
#pragma unroll
        for (int i = 0; i < 8; i++) st[i] = sl[i];
    }

    // the whole stream on the lane pair: head words, the blocks made only of
    // message words (from the first block boundary after the prefix), tail
    template <int NPRE>
    __device__ static void stream_msg(uint32_t st[8], const uint32_t pre[NPRE],
                                      const uint8_t* m, uint32_t mlen, uint32_t before) {
        constexpr uint32_t kk0 = (NPRE + 15) / 16 * 16;
        const uint32_t nfull = NPRE + mlen / 4;
        const uint32_t nfb = nfull >= kk0 + 16 ? (nfull - kk0) / 16 : 0;
        if (!nfb) {
            stream_range<NPRE>(st, pre, m, mlen, before, 0, ~0u);
            return;
        }
        stream_range<NPRE>(st, pre, m, mlen, before, 0, kk0);
        stream_full(st, reinterpret_cast<const uint32_t*>(m) + (kk0 - NPRE), nfb);
        stream_range<NPRE>(st, pre, m, mlen, before, kk0 + 16 * nfb, ~0u);
    }

    // The same stream run by a two-wave workgroup (long messages).  The
    // blocks made only of message words (from the first block boundary
    // after the prefix) have schedules that do not depend on the chaining
    // value: wave 1 (prod) computes their K + W, kLongStage blocks per stage
    // (one per lane) into one half of kw_s, while wave 0's lane pairs run
    // the rounds of the previous stage from the other half.  Head and tail
    // blocks go through stream_range.  Both waves call this with the same
    // arguments and pass the same 1 + stages barriers.
    static constexpr uint32_t kLongStage = 32;
    static constexpr uint32_t kLongLdsWords = 2 * kLongStage * 64;

    template <int NPRE>
    __device__ static void stream_msg_long(uint32_t st[8], const uint32_t pre[NPRE],
                                           const uint8_t* m, uint32_t mlen, uint32_t before,
                                           lds_u32* kw_s, bool prod) {
        constexpr uint32_t kk0 = (NPRE + 15) / 16 * 16;  // first block boundary after pre
        const uint32_t nfull = NPRE + mlen / 4;
        const uint32_t nfb = nfull >= kk0 + 16 ? (nfull - kk0) / 16 : 0;
        const uint32_t nst = (nfb + kLongStage - 1) / kLongStage;
        const uint32_t lane = threadIdx.x & 63u;
        const uint32_t* mw = reinterpret_cast<const uint32_t*>(m) + (kk0 - NPRE);
        auto produce = [&](uint32_t c) {
            const uint32_t blk = c * kLongStage + lane;
            if (lane < kLongStage && blk < nfb)
                sha256_kw_msg_block(mw + 16 * blk, kw_s + ((c & 1) * kLongStage + lane) * 64);
        };
        if (prod) produce(0);
        else if (nfb) stream_range<NPRE>(st, pre, m, mlen, before, 0, kk0);
        __syncthreads();
        uint32_t sl[8];  // in registers across the stages
#pragma unroll
        for (int i = 0; i < 8; i++) sl[i] = st[i];
#pragma unroll 1
        for (uint32_t c = 0; c < nst; c++) {
            if (prod) {
                if (c + 1 < nst) produce(c + 1);
            } else {
                const uint32_t nb = nfb - c * kLongStage < kLongStage ? nfb - c * kLongStage
                                                                      : kLongStage;
                const lds_u32* rows = kw_s + (c & 1) * kLongStage * 64;
#pragma unroll 1
                for (uint32_t b = 0; b < nb; b++) sha256_rounds_dup_kw(sl, rows + 64 * b);
            }
            __syncthreads();
        }
#pragma unroll
        for (int i = 0; i < 8; i++) st[i] = sl[i];
        if (!prod) stream_range<NPRE>(st, pre, m, mlen, before, nfb ? kk0 + 16 * nfb : 0, ~0u);
    }

    // R = HMAC-SHA256(sk_prf, optrand || m)[:n]  (hash_sha256.cu:99-144).
    // LONG: the two-wave long-message form (stream_msg_long over kw_s; prod =
    // this thread is in the producing wave), else one lane pair.  (The form
    // is a template argument: kw_s, an LDS address, may well be 0.)
    template <bool LONG = false>
    __device__ static void prf_msg(const uint32_t skprf[NW], const uint32_t opt[NW],
                                   const uint8_t* m, uint32_t mlen, uint32_t R[NW],
                                   lds_u32* kw_s = nullptr, bool prod = false) {
        uint32_t st[8], w[16], inner[8];
#pragma unroll
        for (int i = 0; i < 8; i++) st[i] = kSha256IV[i];
#pragma unroll
        for (int i = 0; i < 16; i++) w[i] = (i < NW ? skprf[i] : 0u) ^ 0x36363636u;
        sha256_compress_dup(st, w);
        if constexpr (LONG) stream_msg_long<NW>(st, opt, m, mlen, 64, kw_s, prod);
        else stream_msg<NW>(st, opt, m, mlen, 64);
#pragma unroll
        for (int i = 0; i < 8; i++) inner[i] = st[i];
#pragma unroll
        for (int i = 0; i < 8; i++) st[i] = kSha256IV[i];
#pragma unroll
        for (int i = 0; i < 16; i++) w[i] = (i < NW ? skprf[i] : 0u) ^ 0x5c5c5c5cu;
        sha256_compress_dup(st, w);
#pragma unroll
        for (int i = 0; i < 16; i++)
            w[i] = i < 8 ? inner[i] : (i == 8 ? 0x80000000u : (i == 15 ? 768u : 0u));
        sha256_compress_dup(st, w);
#pragma unroll
        for (int i = 0; i < NW; i++) R[i] = st[i];
    }

    // digest bytes = MGF1(SHA256(R || pk || m), DG)  (hash_sha256.cu:200-247)
    template <int DG, bool LONG = false>
    __device__ static void h_msg(const uint32_t R[NW], const uint32_t pk[2 * NW],
                                 const uint8_t* m, uint32_t mlen, uint8_t dg[DG],
                                 lds_u32* kw_s = nullptr, bool prod = false) {
        uint32_t pre[3 * NW], st[8], seed[8], w[16];
#pragma unroll
        for (int i = 0; i < NW; i++) pre[i] = R[i];
#pragma unroll
        for (int i = 0; i < 2 * NW; i++) pre[NW + i] = pk[i];
#pragma unroll
        for (int i = 0; i < 8; i++) st[i] = kSha256IV[i];
        if constexpr (LONG) stream_msg_long<3 * NW>(st, pre, m, mlen, 0, kw_s, prod);
        else stream_msg<3 * NW>(st, pre, m, mlen, 0);
#pragma unroll
        for (int i = 0; i < 8; i++) seed[i] = st[i];
#pragma unroll
        for (int c = 0; c < (DG + 31) / 32; c++) {  // mgf1, sha256.cu:362-383
#pragma unroll
            for (int i = 0; i < 16; i++)
                w[i] = i < 8 ? seed[i]
                             : (i == 8 ? (uint32_t)c
                                       : (i == 9 ? 0x80000000u : (i == 15 ? 288u : 0u)));
#pragma unroll
            for (int i = 0; i < 8; i++) st[i] = kSha256IV[i];
            sha256_compress_dup(st, w);
#pragma unroll
            for (int i = 0; i < 32; i++)
                if (32 * c + i < DG) dg[32 * c + i] = (uint8_t)(st[i >> 2] >> (24 - 8 * (i & 3)));
        }
    }
};

}  // namespace spx
