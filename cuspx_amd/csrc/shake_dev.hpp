// shake_dev.hpp -- SHAKE256 tweakable-hash family for CDNA4 (gfx950).
//
// Semantics follow the reference host path:
//   thash  T_l(x) = SHAKE256(pub_seed || addr[0..32) || x, n)   thash_shake256.cu:15-25
//   PRF(a)        = SHAKE256(sk_seed || addr[0..32), n)         hash_shake256.cu:38-47
//   PRF_msg       = SHAKE256(sk_prf || optrand || m, n)         hash_shake256.cu:64-75
//   H_msg         = SHAKE256(R || pk || m, dgst)                hash_shake256.cu:96-130
// The address uses the SHA-256 byte layout (params.h:148), so bytes 22..31
// are zero.
//
// Keccak-f[1600] runs on 32-bit lane halves: the 200-byte state is 50
// little-endian words s[2i] (low) / s[2i+1] (high), so sponge byte j sits in
// word j/4.  64-bit rotations are two v_alignbit, theta's column parity two
// v_bitop3 (3-way xor) per half, theta application one 3-way xor per word and
// chi one v_bitop3 (a ^ (~b & c)) per word.
#pragma once
#include "spx_common.hpp"

namespace spx {

__constant__ uint32_t kKeccakRC[48] = {
    0x00000001u, 0x00000000u, 0x00008082u, 0x00000000u, 0x0000808au, 0x80000000u,
    0x80008000u, 0x80000000u, 0x0000808bu, 0x00000000u, 0x80000001u, 0x00000000u,
    0x80008081u, 0x80000000u, 0x00008009u, 0x80000000u, 0x0000008au, 0x00000000u,
    0x00000088u, 0x00000000u, 0x80008009u, 0x00000000u, 0x8000000au, 0x00000000u,
    0x8000808bu, 0x00000000u, 0x0000008bu, 0x80000000u, 0x00008089u, 0x80000000u,
    0x00008003u, 0x80000000u, 0x00008002u, 0x80000000u, 0x00000080u, 0x80000000u,
    0x0000800au, 0x00000000u, 0x8000000au, 0x80000000u, 0x80008081u, 0x80000000u,
    0x00008080u, 0x80000000u, 0x80000001u, 0x00000000u, 0x80008008u, 0x80000000u};

// rotation offsets r[x + 5y]
static constexpr int kRho[25] = {0,  1,  62, 28, 27, 36, 44, 6,  55, 20, 3,  10, 43,
                                 25, 39, 41, 45, 15, 21, 8,  18, 2,  61, 56, 14};

template <int R>
__device__ __forceinline__ void rol64(uint32_t lo, uint32_t hi, uint32_t& olo, uint32_t& ohi) {
    if (R == 0) {
        olo = lo;
        ohi = hi;
    } else if (R < 32) {
        ohi = __builtin_amdgcn_alignbit(hi, lo, 32 - R);
        olo = __builtin_amdgcn_alignbit(lo, hi, 32 - R);
    } else if (R == 32) {
        olo = hi;
        ohi = lo;
    } else {
        ohi = __builtin_amdgcn_alignbit(lo, hi, 64 - R);
        olo = __builtin_amdgcn_alignbit(hi, lo, 64 - R);
    }
}

// one round of Keccak-f[1600] on s[50] with round constant (rc_lo, rc_hi)
__device__ __forceinline__ void keccak_round(uint32_t s[50], uint32_t rc_lo, uint32_t rc_hi) {
    uint32_t C[10];
#pragma unroll
    for (int x = 0; x < 5; x++) {
#pragma unroll
        for (int h = 0; h < 2; h++) {
            C[2 * x + h] = xor3(xor3(s[2 * x + h], s[2 * (x + 5) + h], s[2 * (x + 10) + h]),
                                s[2 * (x + 15) + h], s[2 * (x + 20) + h]);
        }
    }
#pragma unroll
    for (int x = 0; x < 5; x++) {
        uint32_t rlo, rhi;
        const int xp = (x + 1) % 5, xm = (x + 4) % 5;
        rol64<1>(C[2 * xp], C[2 * xp + 1], rlo, rhi);
#pragma unroll
        for (int y = 0; y < 5; y++) {
            s[2 * (x + 5 * y)] = xor3(s[2 * (x + 5 * y)], C[2 * xm], rlo);
            s[2 * (x + 5 * y) + 1] = xor3(s[2 * (x + 5 * y) + 1], C[2 * xm + 1], rhi);
        }
    }
    uint32_t B[50];
#pragma unroll
    for (int x = 0; x < 5; x++) {
#pragma unroll
        for (int y = 0; y < 5; y++) {
            // pi: B[y, 2x + 3y] = rol(A[x, y], r[x, y])
            const int src = x + 5 * y;
            const int dst = y + 5 * ((2 * x + 3 * y) % 5);
            uint32_t lo, hi;
            switch (kRho[src]) {
#define SPX_ROT_CASE(R) \
    case R: rol64<R>(s[2 * src], s[2 * src + 1], lo, hi); break;
                SPX_ROT_CASE(0) SPX_ROT_CASE(1) SPX_ROT_CASE(62) SPX_ROT_CASE(28)
                SPX_ROT_CASE(27) SPX_ROT_CASE(36) SPX_ROT_CASE(44) SPX_ROT_CASE(6)
                SPX_ROT_CASE(55) SPX_ROT_CASE(20) SPX_ROT_CASE(3) SPX_ROT_CASE(10)
                SPX_ROT_CASE(43) SPX_ROT_CASE(25) SPX_ROT_CASE(39) SPX_ROT_CASE(41)
                SPX_ROT_CASE(45) SPX_ROT_CASE(15) SPX_ROT_CASE(21) SPX_ROT_CASE(8)
                SPX_ROT_CASE(18) SPX_ROT_CASE(2) SPX_ROT_CASE(61) SPX_ROT_CASE(56)
                SPX_ROT_CASE(14)
#undef SPX_ROT_CASE
                default: lo = s[2 * src]; hi = s[2 * src + 1]; break;
            }
            B[2 * dst] = lo;
            B[2 * dst + 1] = hi;
        }
    }
#pragma unroll
    for (int y = 0; y < 5; y++) {
#pragma unroll
        for (int x = 0; x < 5; x++) {
#pragma unroll
            for (int h = 0; h < 2; h++) {
                s[2 * (x + 5 * y) + h] = bit_chi(B[2 * (x + 5 * y) + h],
                                                 B[2 * ((x + 1) % 5 + 5 * y) + h],
                                                 B[2 * ((x + 2) % 5 + 5 * y) + h]);
            }
        }
    }
    s[0] ^= rc_lo;
    s[1] ^= rc_hi;
}

// Keccak-f[1600]: round 0 unrolled (so constant state words fold after an
// absorb), rounds 1..23 in a loop with the constants from constant memory.
__device__ __forceinline__ void keccak_f(uint32_t s[50]) {
    keccak_round(s, 0x00000001u, 0x00000000u);
#pragma unroll 1
    for (int r = 1; r < 24; r++) keccak_round(s, kKeccakRC[2 * r], kKeccakRC[2 * r + 1]);
}

// ---------------------------------------------------- bit-interleaved form
//
// A 64-bit lane (lo, hi) can also be held as (E, O): E = its even-numbered
// bits, O = its odd-numbered bits.  A 64-bit rotation by 2k is then a 32-bit
// rotation of both halves by k, and by 2k+1 a swap plus rotations by k+1 and
// k -- theta's rotation by 1 costs one v_alignbit instead of two, and so does
// rho's rotation by 1: 52 instead of 58 rotates per round.  The WOTS chain
// steps of the hypertree kernel run in this form (values stay interleaved
// from step to step); everything else uses the plain form.

// bits 0, 2, 4, ... of x into bits 0..15
__device__ __forceinline__ uint32_t il_even16(uint32_t x) {
    x &= 0x55555555u;
    x = (x | (x >> 1)) & 0x33333333u;
    x = (x | (x >> 2)) & 0x0F0F0F0Fu;
    x = (x | (x >> 4)) & 0x00FF00FFu;
    return (x | (x >> 8)) & 0x0000FFFFu;
}
// bits 0..15 of x into bits 0, 2, 4, ...
__device__ __forceinline__ uint32_t il_spread16(uint32_t x) {
    x &= 0x0000FFFFu;
    x = (x | (x << 8)) & 0x00FF00FFu;
    x = (x | (x << 4)) & 0x0F0F0F0Fu;
    x = (x | (x << 2)) & 0x33333333u;
    return (x | (x << 1)) & 0x55555555u;
}
// (lo, hi) -> (E, O)
__device__ __forceinline__ void il_to(uint32_t lo, uint32_t hi, uint32_t& e, uint32_t& o) {
    e = il_even16(lo) | (il_even16(hi) << 16);
    o = il_even16(lo >> 1) | (il_even16(hi >> 1) << 16);
}
// (E, O) -> (lo, hi)
__device__ __forceinline__ void il_from(uint32_t e, uint32_t o, uint32_t& lo, uint32_t& hi) {
    lo = il_spread16(e) | (il_spread16(o) << 1);
    hi = il_spread16(e >> 16) | (il_spread16(o >> 16) << 1);
}

// rotl32 by M (M in 0..32)
template <int M>
__device__ __forceinline__ uint32_t rotl32c(uint32_t x) {
    if constexpr (M == 0 || M == 32) return x;
    else return __builtin_amdgcn_alignbit(x, x, 32 - M);
}

// 64-bit rotation by R of an interleaved lane
template <int R>
__device__ __forceinline__ void rol64_il(uint32_t e, uint32_t o, uint32_t& oe, uint32_t& oo) {
    if constexpr ((R & 1) == 0) {
        oe = rotl32c<R / 2>(e);
        oo = rotl32c<R / 2>(o);
    } else {
        oe = rotl32c<(R + 1) / 2>(o);
        oo = rotl32c<(R - 1) / 2>(e);
    }
}

// round constants in interleaved form: (E, O) of each 64-bit constant
struct KeccakRcIl {
    uint32_t w[48];
    constexpr KeccakRcIl() : w{} {
        constexpr uint64_t rc[24] = {
            0x0000000000000001ull, 0x0000000000008082ull, 0x800000000000808aull,
            0x8000000080008000ull, 0x000000000000808bull, 0x0000000080000001ull,
            0x8000000080008081ull, 0x8000000000008009ull, 0x000000000000008aull,
            0x0000000000000088ull, 0x0000000080008009ull, 0x000000008000000aull,
            0x000000008000808bull, 0x800000000000008bull, 0x8000000000008089ull,
            0x8000000000008003ull, 0x8000000000008002ull, 0x8000000000000080ull,
            0x000000000000800aull, 0x800000008000000aull, 0x8000000080008081ull,
            0x8000000000008080ull, 0x0000000080000001ull, 0x8000000080008008ull};
        for (int r = 0; r < 24; r++) {
            uint32_t e = 0, o = 0;
            for (int b = 0; b < 32; b++) {
                e |= (uint32_t)((rc[r] >> (2 * b)) & 1u) << b;
                o |= (uint32_t)((rc[r] >> (2 * b + 1)) & 1u) << b;
            }
            w[2 * r] = e;
            w[2 * r + 1] = o;
        }
    }
};
__constant__ KeccakRcIl kKeccakRcIl = KeccakRcIl();

// one round on the interleaved state (s[2i] = E, s[2i+1] = O of lane i)
__device__ __forceinline__ void keccak_round_il(uint32_t s[50], uint32_t rc_e, uint32_t rc_o) {
    uint32_t C[10];
#pragma unroll
    for (int x = 0; x < 5; x++) {
#pragma unroll
        for (int h = 0; h < 2; h++) {
            C[2 * x + h] = xor3(xor3(s[2 * x + h], s[2 * (x + 5) + h], s[2 * (x + 10) + h]),
                                s[2 * (x + 15) + h], s[2 * (x + 20) + h]);
        }
    }
#pragma unroll
    for (int x = 0; x < 5; x++) {
        const int xp = (x + 1) % 5, xm = (x + 4) % 5;
        const uint32_t re = rotl32c<1>(C[2 * xp + 1]), ro = C[2 * xp];  // rol64 by 1
#pragma unroll
        for (int y = 0; y < 5; y++) {
            s[2 * (x + 5 * y)] = xor3(s[2 * (x + 5 * y)], C[2 * xm], re);
            s[2 * (x + 5 * y) + 1] = xor3(s[2 * (x + 5 * y) + 1], C[2 * xm + 1], ro);
        }
    }
    uint32_t B[50];
#pragma unroll
    for (int x = 0; x < 5; x++) {
#pragma unroll
        for (int y = 0; y < 5; y++) {
            const int src = x + 5 * y;
            const int dst = y + 5 * ((2 * x + 3 * y) % 5);
            uint32_t e, o;
            switch (kRho[src]) {
#define SPX_ROT_CASE(R) \
    case R: rol64_il<R>(s[2 * src], s[2 * src + 1], e, o); break;
                SPX_ROT_CASE(0) SPX_ROT_CASE(1) SPX_ROT_CASE(62) SPX_ROT_CASE(28)
                SPX_ROT_CASE(27) SPX_ROT_CASE(36) SPX_ROT_CASE(44) SPX_ROT_CASE(6)
                SPX_ROT_CASE(55) SPX_ROT_CASE(20) SPX_ROT_CASE(3) SPX_ROT_CASE(10)
                SPX_ROT_CASE(43) SPX_ROT_CASE(25) SPX_ROT_CASE(39) SPX_ROT_CASE(41)
                SPX_ROT_CASE(45) SPX_ROT_CASE(15) SPX_ROT_CASE(21) SPX_ROT_CASE(8)
                SPX_ROT_CASE(18) SPX_ROT_CASE(2) SPX_ROT_CASE(61) SPX_ROT_CASE(56)
                SPX_ROT_CASE(14)
#undef SPX_ROT_CASE
                default: e = s[2 * src]; o = s[2 * src + 1]; break;
            }
            B[2 * dst] = e;
            B[2 * dst + 1] = o;
        }
    }
#pragma unroll
    for (int y = 0; y < 5; y++) {
#pragma unroll
        for (int x = 0; x < 5; x++) {
#pragma unroll
            for (int h = 0; h < 2; h++) {
                s[2 * (x + 5 * y) + h] = bit_chi(B[2 * (x + 5 * y) + h],
                                                 B[2 * ((x + 1) % 5 + 5 * y) + h],
                                                 B[2 * ((x + 2) % 5 + 5 * y) + h]);
            }
        }
    }
    s[0] ^= rc_e;
    s[1] ^= rc_o;
}

// Keccak-f[1600] on the interleaved state; round 0 unrolled as in keccak_f
__device__ __forceinline__ void keccak_f_il(uint32_t s[50]) {
    keccak_round_il(s, 1u, 0u);  // RC[0] = 1: bit 0 is even
#pragma unroll 1
    for (int r = 1; r < 24; r++)
        keccak_round_il(s, kKeccakRcIl.w[2 * r], kKeccakRcIl.w[2 * r + 1]);
}

// Keccak-f[1600] on a lane pair in the interleaved form: the even lane holds
// the E halves of the 25 lanes, the odd lane the O halves (s[i] = this
// lane's half of Keccak lane i).  A rotation by 2k rotates both halves by k
// (no exchange); a rotation by 2k + 1 swaps them (one DPP move) and rotates
// by k + 1 (even lane) or k (odd lane), a per-lane count.  Theta's column
// sums, chi and iota stay within a half.  128 instructions per round on
// each lane (in the built ISA) against 174 for the whole state on one lane:
// for serial sponges whose lane would otherwise run alone.
__device__ __forceinline__ void keccak_round_il_pair(uint32_t s[25], uint32_t m, uint32_t rc) {
    uint32_t C[5];
#pragma unroll
    for (int x = 0; x < 5; x++) C[x] = xor3(xor3(s[x], s[x + 5], s[x + 10]), s[x + 15], s[x + 20]);
    const uint32_t sh1 = lane_sel(m, 0u, 31u);  // rotl by 1 (E) or 0 (O) after the swap
#pragma unroll
    for (int x = 0; x < 5; x++) {
        const uint32_t t = pair_swap(C[(x + 1) % 5]);
        const uint32_t r = __builtin_amdgcn_alignbit(t, t, sh1);  // rol64(C[x+1], 1), this half
#pragma unroll
        for (int y = 0; y < 5; y++) s[x + 5 * y] = xor3(s[x + 5 * y], C[(x + 4) % 5], r);
    }
    uint32_t B[25];
#pragma unroll
    for (int x = 0; x < 5; x++) {
#pragma unroll
        for (int y = 0; y < 5; y++) {
            const int src = x + 5 * y;
            const int dst = y + 5 * ((2 * x + 3 * y) % 5);
            const int R = kRho[src];
            uint32_t v;
            if (R == 0) {
                v = s[src];
            } else if ((R & 1) == 0) {
                v = __builtin_amdgcn_alignbit(s[src], s[src], (32 - R / 2) & 31);
            } else {
                const uint32_t t = pair_swap(s[src]);
                const uint32_t sh = lane_sel(m, (uint32_t)((32 - R / 2) & 31),
                                             (uint32_t)((31 - R / 2) & 31));
                v = __builtin_amdgcn_alignbit(t, t, sh);
            }
            B[dst] = v;
        }
    }
#pragma unroll
    for (int y = 0; y < 5; y++)
#pragma unroll
        for (int x = 0; x < 5; x++)
            s[x + 5 * y] = bit_chi(B[x + 5 * y], B[(x + 1) % 5 + 5 * y], B[(x + 2) % 5 + 5 * y]);
    s[0] ^= rc;
}

__device__ __forceinline__ void keccak_f_il_pair(uint32_t s[25]) {
    const uint32_t m = (threadIdx.x & 1u) ? ~0u : 0u;
    keccak_round_il_pair(s, m, lane_sel(m, 0u, 1u));  // RC[0] = 1: bit 0 is even
#pragma unroll 1
    for (int r = 1; r < 24; r++)
        keccak_round_il_pair(s, m, lane_sel(m, kKeccakRcIl.w[2 * r + 1], kKeccakRcIl.w[2 * r]));
}

template <int N>
struct Shake {
    static constexpr int hash = kShake;
    static constexpr int NW = N / 4;

    struct Key {
        uint32_t pub[NW];
        uint32_t sk[NW];
    };

    __device__ static __forceinline__ void load_key(const Plan& p, Key& k) {
#pragma unroll
        for (int i = 0; i < NW; i++) {
            k.pub[i] = p.pubseed[i];
            k.sk[i] = p.skseed[i];
        }
    }

    // address byte-words (big-endian) as sponge words (little-endian)
    __device__ static __forceinline__ uint32_t aw(const Adr& a, int i) {
        return i < 6 ? bswap32(a.w[i]) : 0u;
    }

    // single-permutation SHAKE256 over pre(NW) || addr32 || d(M), output n
    template <int M>
    __device__ static __forceinline__ void one_block(const uint32_t pre[NW], const Adr& a,
                                                     const uint32_t* d, uint32_t out[NW]) {
        constexpr int L = NW + 8 + M;  // input words
        static_assert(4 * L <= 135, "input must fit one block");
        uint32_t s[50];
#pragma unroll
        for (int i = 0; i < 50; i++) {
            uint32_t x = 0;
            if (i < NW) x = pre[i];
            else if (i < NW + 8) x = aw(a, i - NW);
            else if (i < L) x = d[i - NW - 8];
            if (i == L) x ^= 0x1Fu;
            if (i == 33) x ^= 0x80000000u;
            s[i] = x;
        }
        keccak_f(s);
#pragma unroll
        for (int i = 0; i < NW; i++) out[i] = s[i];
    }

    __device__ static __forceinline__ void F(const Key& k, const Adr& a, const uint32_t in[NW],
                                             uint32_t out[NW]) {
        one_block<NW>(k.pub, a, in, out);
    }

    // WOTS chain in the interleaved form: pub_seed and the chain's address
    // (hash field 0) are interleaved once; each step adds the hash field --
    // address byte 21, bits 40..43 of sponge lane (n + 21) / 8 -- by xor,
    // which commutes with interleaving.  Values in and out are interleaved.
    static constexpr int kHashLane = (N + 21) / 8;
    static constexpr int kPadLane = (2 * N + 32) / 8;  // the 0x1F byte, at a lane start
    struct ChainIl {
        uint32_t w[NW + 8];  // (E, O) of the pub_seed and address lanes
    };
    __device__ static __forceinline__ void chain_il_begin(const Key& k, const Adr& a,
                                                          ChainIl& c) {
#pragma unroll
        for (int l = 0; l < (NW + 8) / 2; l++) {
            const uint32_t lo = 2 * l < NW ? k.pub[2 * l] : aw(a, 2 * l - NW);
            const uint32_t hi = 2 * l + 1 < NW ? k.pub[2 * l + 1] : aw(a, 2 * l + 1 - NW);
            il_to(lo, hi, c.w[2 * l], c.w[2 * l + 1]);
        }
    }
    // one chain step at hash address h (h < 16), in and out interleaved
    __device__ static __forceinline__ void F_il(const ChainIl& c, uint32_t h,
                                                const uint32_t in[NW], uint32_t out[NW]) {
        uint32_t s[50];
#pragma unroll
        for (int i = 0; i < 50; i++) {
            uint32_t x = 0;
            if (i < NW + 8) x = c.w[i];
            else if (i < 2 * NW + 8) x = in[i - NW - 8];
            if (i == 2 * kPadLane) x ^= 7u;      // 0x1F: even bits 0, 2, 4
            if (i == 2 * kPadLane + 1) x ^= 3u;  //       odd bits 1, 3
            if (i == 33) x ^= 0x80000000u;       // 0x80 at byte 135: lane 16, bit 63
            s[i] = x;
        }
        s[2 * kHashLane] ^= ((h & 1u) << 20) | (((h >> 2) & 1u) << 21);
        s[2 * kHashLane + 1] ^= (((h >> 1) & 1u) << 20) | (((h >> 3) & 1u) << 21);
        keccak_f_il(s);
#pragma unroll
        for (int i = 0; i < NW; i++) out[i] = s[i];
    }
    // NW plain words (NW / 2 lanes) <-> interleaved
    __device__ static __forceinline__ void to_il(const uint32_t in[NW], uint32_t out[NW]) {
#pragma unroll
        for (int l = 0; l < NW / 2; l++) il_to(in[2 * l], in[2 * l + 1], out[2 * l], out[2 * l + 1]);
    }
    __device__ static __forceinline__ void from_il(const uint32_t in[NW], uint32_t out[NW]) {
#pragma unroll
        for (int l = 0; l < NW / 2; l++)
            il_from(in[2 * l], in[2 * l + 1], out[2 * l], out[2 * l + 1]);
    }

    __device__ static __forceinline__ void H(const Key& k, const Adr& a, const uint32_t l[NW],
                                             const uint32_t r[NW], uint32_t out[NW]) {
        uint32_t d[2 * NW];
#pragma unroll
        for (int i = 0; i < NW; i++) {
            d[i] = l[i];
            d[NW + i] = r[i];
        }
        one_block<2 * NW>(k.pub, a, d, out);
    }

    __device__ static __forceinline__ void PRF(const Key& k, const Adr& a, uint32_t out[NW]) {
        one_block<0>(k.sk, a, nullptr, out);
    }

    // ------------------------------------------------ per-request hashing

    // SHAKE256(pre || m) squeezed to OUTB bytes (OUTB <= 136), on a lane pair
    // (both lanes active with identical inputs; keccak_f_il_pair).  m is
    // 4-byte aligned; the word holding the last message byte is the last one
    // read.  The state is held interleaved, this lane's half of every Keccak
    // lane; each absorbed plain word is split into its even or odd bits on
    // the way in.  Whole 136-byte blocks of message words are absorbed
    // directly, the head after the prefix and the tail word by word.  The
    // output is in both lanes.
    template <int NPRE, int OUTB>
    __device__ static void stream_msg_pair(const uint32_t pre[NPRE], const uint8_t* m,
                                           uint32_t mlen, uint8_t out[OUTB]) {
        static_assert(NPRE < 34 && OUTB <= 136, "prefix within one block");
        const uint32_t h = threadIdx.x & 1u;
        const uint32_t msk = h ? ~0u : 0u;
        // this lane's interleaved half of plain word x at stream word q
        auto half = [&](uint32_t x, int q) { return il_even16(x >> h) << (16 * (q & 1)); };
        const uint32_t* mw = reinterpret_cast<const uint32_t*>(m);
        uint32_t s[25];
#pragma unroll
        for (int i = 0; i < 25; i++) s[i] = 0;
#pragma unroll
        for (int q = 0; q < NPRE; q++) s[q >> 1] ^= half(pre[q], q);
        int q = NPRE;
        const uint32_t full = mlen / 4, rem = mlen & 3;
        uint32_t i = 0;
        // xor this lane's half of word x at a wave-uniform runtime q
        auto put = [&](int qq, uint32_t x) {
            const uint32_t c = il_even16(x >> h) << (16 * (qq & 1));
#pragma unroll
            for (int l = 0; l < 17; l++) s[l] ^= (qq >> 1) == l ? c : 0u;
        };
        // the message words up to the first block boundary (stream positions
        // NPRE .. 33, fixed at compile time), then the blocks made only of
        // message words (block b + 1 loads while block b permutes), then the
        // words left over (positions 0 .. 32 of the last block); the head's
        // and the tail's words are each loaded together (one word per step
        // had put a load on the serial path for every word)
        const gbl_u32* mg = (gbl_u32*)mw;  // address-space cast
        constexpr int HEAD = NPRE == 0 ? 0 : 34 - NPRE;
        if constexpr (HEAD > 0) {
            const uint32_t nh = (uint32_t)HEAD < full ? (uint32_t)HEAD : full;
            uint32_t x[HEAD];
#pragma unroll
            for (int k = 0; k < HEAD; k++) x[k] = (uint32_t)k < nh ? mg[k] : 0u;
#pragma unroll
            for (int k = 0; k < HEAD; k++) s[(NPRE + k) >> 1] ^= half(x[k], NPRE + k);
            i = nh;
            q = NPRE + (int)nh;
            if (q == 34) {
                keccak_f_il_pair(s);
                q = 0;
            }
        }
        const uint32_t nfb = q == 0 ? (full - i) / 34 : 0u;
        if (nfb) {
            const gbl_u32* g = (gbl_u32*)mw + i;  // address-space cast
            uint32_t cur[34];
#pragma unroll
            for (int l = 0; l < 34; l++) cur[l] = g[l];
#pragma unroll 1
            for (uint32_t b = 0; b < nfb; b++) {
#pragma unroll
                for (int l = 0; l < 17; l++)
                    s[l] ^= il_even16(cur[2 * l] >> h) | (il_even16(cur[2 * l + 1] >> h) << 16);
                if (b + 1 < nfb) {
#pragma unroll
                    for (int l = 0; l < 34; l++) cur[l] = g[34 * (b + 1) + l];
                }
                keccak_f_il_pair(s);
            }
            i += 34 * nfb;
        }
        if (q == 0) {  // else the message ended inside the first block: i == full
            const uint32_t r = full - i;  // < 34
            uint32_t x[34];
#pragma unroll
            for (int k = 0; k < 34; k++) x[k] = (uint32_t)k < r ? mg[i + k] : 0u;
#pragma unroll
            for (int k = 0; k < 34; k++) s[k >> 1] ^= half(x[k], k);
            q = (int)r;
            i = full;
        }
        uint32_t last = rem ? (mw[full] & (0xFFFFFFFFu >> (8 * (4 - rem)))) : 0u;
        last |= 0x1Fu << (8 * rem);
        put(q, last);
        s[16] ^= half(0x80000000u, 33);
        keccak_f_il_pair(s);
#pragma unroll
        for (int l = 0; l < (OUTB + 7) / 8; l++) {
            const uint32_t other = pair_swap(s[l]);
            uint32_t lo, hi;
            il_from(lane_sel(msk, other, s[l]), lane_sel(msk, s[l], other), lo, hi);
#pragma unroll
            for (int o = 0; o < 8; o++)
                if (8 * l + o < OUTB) out[8 * l + o] = (uint8_t)((o < 4 ? lo : hi) >> (8 * (o & 3)));
        }
    }

    // R = SHAKE256(sk_prf || optrand || m, n) (hash_shake256.cu:64-75), on a
    // lane pair holding identical inputs
    __device__ static void prf_msg(const uint32_t skprf[NW], const uint32_t opt[NW],
                                   const uint8_t* m, uint32_t mlen, uint32_t R[NW]) {
        uint32_t pre[2 * NW];
        uint8_t o[N];
#pragma unroll
        for (int i = 0; i < NW; i++) {
            pre[i] = skprf[i];
            pre[NW + i] = opt[i];
        }
        stream_msg_pair<2 * NW, N>(pre, m, mlen, o);
#pragma unroll
        for (int i = 0; i < NW; i++)
            R[i] = (uint32_t)o[4 * i] | (uint32_t)o[4 * i + 1] << 8 |
                   (uint32_t)o[4 * i + 2] << 16 | (uint32_t)o[4 * i + 3] << 24;
    }

    // digest = SHAKE256(R || pk || m, DG) (hash_shake256.cu:96-130), on a
    // lane pair holding identical inputs
    template <int DG>
    __device__ static void h_msg(const uint32_t R[NW], const uint32_t pk[2 * NW],
                                 const uint8_t* m, uint32_t mlen, uint8_t dg[DG]) {
        uint32_t pre[3 * NW];
#pragma unroll
        for (int i = 0; i < NW; i++) pre[i] = R[i];
#pragma unroll
        for (int i = 0; i < 2 * NW; i++) pre[NW + i] = pk[i];
        stream_msg_pair<3 * NW, DG>(pre, m, mlen, dg);
    }

    // no midstate for SHAKE256 (zeros); PAIR as for Sha2::midstate
    template <bool PAIR = false>
    __device__ static __forceinline__ void midstate(const uint32_t*, uint32_t mid[8]) {
#pragma unroll
        for (int i = 0; i < 8; i++) mid[i] = 0;
    }
};

}  // namespace spx
