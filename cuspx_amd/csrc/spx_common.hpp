// spx_common.hpp -- compile-time parameter sets, address words and small
// device helpers shared by the SHA-256 and SHAKE256 kernel families.
//
// Parameter sets follow params.h:33-141 of the reference; the address layout
// follows sha256_offsets.h:9-17, which the reference uses for BOTH hash
// families (params.h:148).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace spx {

enum HashKind { kSha2 = 0, kShake = 1 };

constexpr int ilog2(int x) { return x <= 1 ? 0 : 1 + ilog2(x / 2); }

// Compile-time parameter set (params.h:33-141).
template <int HASH, int N_, int FULLH, int D_, int A_, int K_>
struct Params {
    static constexpr int hash = HASH;
    static constexpr int n = N_;
    static constexpr int nw = N_ / 4;          // hash value in 32-bit words
    static constexpr int full_h = FULLH;
    static constexpr int d = D_;
    static constexpr int a = A_;               // FORS tree height
    static constexpr int k = K_;               // FORS trees
    static constexpr int w = 16;
    static constexpr int len1 = 2 * N_;
    static constexpr int len2 = 3;
    static constexpr int len = len1 + len2;
    static constexpr int th = FULLH / D_;      // XMSS subtree height h'
    static constexpr int wots_bytes = len * N_;
    static constexpr int fors_msg_bytes = (A_ * K_ + 7) / 8;
    static constexpr int fors_bytes = (A_ + 1) * K_ * N_;
    static constexpr int sig_bytes = N_ + fors_bytes + D_ * wots_bytes + FULLH * N_;
    static constexpr int pk_bytes = 2 * N_;
    static constexpr int sk_bytes = 4 * N_;
    static constexpr int tree_bits = th * (D_ - 1);
    static constexpr int tree_bytes = (tree_bits + 7) / 8;
    static constexpr int leaf_bits = th;
    static constexpr int leaf_bytes = (th + 7) / 8;
    static constexpr int dgst_bytes = fors_msg_bytes + tree_bytes + leaf_bytes;
    static constexpr int layer_bytes = wots_bytes + th * N_;  // WOTS sig + auth path
    static_assert(tree_bits <= 64, "tree index must fit 64 bits");
};

// Per-request plan written by the message kernel and read by every later
// kernel of the same batch (one 512-byte record per request).
struct alignas(16) Plan {
    uint32_t mid[8];       // SHA-256 state after absorbing pub_seed||0^(64-n)
    uint32_t skseed[8];    // hash-native words (BE for SHA-2, LE for SHAKE)
    uint32_t pubseed[8];
    uint32_t root[8];      // public root (verify) / unused (sign)
    uint64_t tree;         // hypertree tree index from H_msg
    uint32_t leaf;         // leaf index in the bottom tree
    uint32_t status;       // verify: 0 ok / 1 fail
    uint32_t fors_idx[64]; // FORS leaf indices (message_to_indices)
};
static_assert(sizeof(Plan) <= 512, "plan record too large");

// ---------------------------------------------------------------- helpers

__device__ __forceinline__ uint32_t bswap32(uint32_t x) { return __builtin_bswap32(x); }

// a word pointer into global memory (address space 1): global_load rather
// than flat, so message loads are not counted with LDS accesses
typedef const uint32_t __attribute__((address_space(1))) gbl_u32;

__device__ __forceinline__ uint32_t rotr32(uint32_t x, int c) {
    return __builtin_amdgcn_alignbit(x, x, c);
}

// gfx950 v_bitop3_b32: any 3-input bitwise function in one op.  The 8-bit
// table is indexed with S0 = 0xF0, S1 = 0xCC, S2 = 0xAA.  Constant operands
// fall back to plain C so the compiler can fold them.
__device__ __forceinline__ uint32_t xor3(uint32_t a, uint32_t b, uint32_t c) {
    if (__builtin_constant_p(a) || __builtin_constant_p(b) || __builtin_constant_p(c))
        return a ^ b ^ c;
    return __builtin_amdgcn_bitop3_b32(a, b, c, 0x96);
}
__device__ __forceinline__ uint32_t bit_ch(uint32_t e, uint32_t f, uint32_t g) {
    return __builtin_amdgcn_bitop3_b32(e, f, g, 0xCA);  // (e & f) ^ (~e & g)
}
__device__ __forceinline__ uint32_t bit_maj(uint32_t a, uint32_t b, uint32_t c) {
    return __builtin_amdgcn_bitop3_b32(a, b, c, 0xE8);  // majority
}
__device__ __forceinline__ uint32_t bit_chi(uint32_t a, uint32_t b, uint32_t c) {
    return __builtin_amdgcn_bitop3_b32(a, b, c, 0xD2);  // a ^ (~b & c)
}

// Lane pairs (lanes 2i, 2i + 1 of a wave, both active) share one serial
// hash: the other lane's value of x (DPP quad_perm [1, 0, 3, 2])
__device__ __forceinline__ uint32_t pair_swap(uint32_t x) {
    return (uint32_t)__builtin_amdgcn_mov_dpp((int)x, 0xB1, 0xF, 0xF, true);
}

// m ? a : b bitwise, for a lane mask m.  Written as a bit select so that a
// choice between two elements of one array never becomes an indexed read
// (which the compiler lowers to a compare-and-select chain over the array).
__device__ __forceinline__ uint32_t lane_sel(uint32_t m, uint32_t a, uint32_t b) {
    return __builtin_amdgcn_bitop3_b32(m, a, b, 0xCA);
}

// ------------------------------------------------------------- addresses
//
// The 32-byte address as 8 big-endian words: word i holds bytes 4i..4i+3.
// Offsets (sha256_offsets.h:9-17): layer@0, tree@1..8 (BE), type@9,
// keypair hi@12 (only when h' > 8, address.cu:81-86) lo@13, chain@17,
// hash@21, tree height@17, tree index@18..21 (BE).  Bytes 22..31 stay zero.
enum AdrType { kWots = 0, kWotsPk = 1, kHashTree = 2, kForsTree = 3, kForsPk = 4 };

struct Adr {
    uint32_t w[6];  // bytes 0..23; words 6,7 are always zero
};

__device__ __forceinline__ Adr adr_make(uint32_t layer, uint64_t tree, uint32_t type) {
    Adr a;
    a.w[0] = (layer << 24) | (uint32_t)((tree >> 40) & 0xFFFFFFu);
    a.w[1] = (uint32_t)(tree >> 8);
    a.w[2] = ((uint32_t)(tree & 0xFF) << 24) | (type << 16);
    a.w[3] = 0;
    a.w[4] = 0;
    a.w[5] = 0;
    return a;
}
template <int TH>
__device__ __forceinline__ void adr_keypair(Adr& a, uint32_t kp) {
    a.w[3] = (TH > 8 ? ((kp >> 8) & 0xFF) << 24 : 0u) | ((kp & 0xFF) << 16);
}
__device__ __forceinline__ void adr_chain(Adr& a, uint32_t c) { a.w[4] = c << 16; }
__device__ __forceinline__ void adr_hash(Adr& a, uint32_t h) { a.w[5] = h << 16; }
__device__ __forceinline__ void adr_node(Adr& a, uint32_t height, uint32_t index) {
    a.w[4] = (height << 16) | (index >> 16);
    a.w[5] = index << 16;
}

// ------------------------------------------------------- word/byte access
//
// Hash values travel as n/4 32-bit words in the family's native order:
// big-endian words for SHA-256, little-endian words for SHAKE256 (the
// Keccak lane order).  Device buffers hold plain bytes.

template <int HASH>
__device__ __forceinline__ uint32_t to_native(uint32_t le_word) {
    return HASH == kSha2 ? bswap32(le_word) : le_word;
}

template <int HASH, int NW>
__device__ __forceinline__ void load_val(const uint8_t* p, uint32_t v[NW]) {
    const uint32_t* q = reinterpret_cast<const uint32_t*>(p);
#pragma unroll
    for (int i = 0; i < NW; i++) v[i] = to_native<HASH>(q[i]);
}

template <int HASH, int NW>
__device__ __forceinline__ void store_val(uint8_t* p, const uint32_t v[NW]) {
    uint32_t* q = reinterpret_cast<uint32_t*>(p);
#pragma unroll
    for (int i = 0; i < NW; i++) q[i] = to_native<HASH>(v[i]);
}

// byte i (0-based, in message order) of a native-order value
template <int HASH>
__device__ __forceinline__ uint32_t val_byte(const uint32_t* v, int i) {
    uint32_t w = v[i >> 2];
    int sh = HASH == kSha2 ? (24 - 8 * (i & 3)) : 8 * (i & 3);
    return (w >> sh) & 0xFF;
}

// select v[i] for a wave-uniform runtime i without local-memory indexing
template <int NW>
__device__ __forceinline__ uint32_t pick(const uint32_t v[NW], int i) {
    uint32_t r = v[0];
#pragma unroll
    for (int j = 1; j < NW; j++) r = (i == j) ? v[j] : r;
    return r;
}

}  // namespace spx
