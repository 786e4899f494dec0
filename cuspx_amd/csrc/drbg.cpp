// drbg.cpp -- NIST AES-256-CTR DRBG (no derivation function), the input
// generator the reference uses for keys, optrand and test messages
// (rng.cu:116-241, PQCgenKAT_sign.cu:55-63).  The reference calls OpenSSL's
// EVP AES-256-ECB; this file carries its own FIPS-197 AES-256 encryptor so
// the library has no OpenSSL dependency: the x86 AES instructions where the
// CPU has them, else 32-bit round tables.  State is process-global like the
// reference's DRBG_ctx (rng.cu:14) but guarded by a mutex.
#include <atomic>
#include <mutex>
#include <stdint.h>
#include <string.h>

#if defined(__x86_64__)
#include <immintrin.h>
#endif

#include "drbg.hpp"

namespace spx {
namespace {

struct Aes256 {
    uint8_t sbox[256];
    Aes256() {
        // S-box from its definition (FIPS-197 5.1.1): GF(2^8) inverse, affine map
        uint8_t p = 1, q = 1;
        do {
            p = (uint8_t)(p ^ (uint8_t)(p << 1) ^ ((p & 0x80) ? 0x1B : 0));  // p *= 3
            q ^= (uint8_t)(q << 1);                                             // q /= 3
            q ^= (uint8_t)(q << 2);
            q ^= (uint8_t)(q << 4);
            if (q & 0x80) q ^= 0x09;
            const uint8_t x = (uint8_t)(q ^ (uint8_t)((q << 1) | (q >> 7)) ^
                                        (uint8_t)((q << 2) | (q >> 6)) ^
                                        (uint8_t)((q << 3) | (q >> 5)) ^
                                        (uint8_t)((q << 4) | (q >> 4)));
            sbox[p] = x ^ 0x63;
        } while (p != 1);
        sbox[0] = 0x63;
        tables();
    }
    // 32-bit round tables (FIPS-197 5.1.1-5.1.3 folded per column): entry x
    // of te0 is the MixColumns column of S(x) in row 0, (2S, S, S, 3S) with
    // row 0 in the top byte; rows 1-3 are its byte rotations
    uint32_t te0[256], te1[256], te2[256], te3[256];
    void tables() {
        for (int x = 0; x < 256; x++) {
            const uint32_t s1 = sbox[x], s2 = xt(sbox[x]), s3 = s2 ^ s1;
            const uint32_t w = (s2 << 24) | (s1 << 16) | (s1 << 8) | s3;
            te0[x] = w;
            te1[x] = (w >> 8) | (w << 24);
            te2[x] = (w >> 16) | (w << 16);
            te3[x] = (w >> 24) | (w << 8);
        }
    }
    static uint8_t xt(uint8_t a) { return (uint8_t)((a << 1) ^ ((a & 0x80) ? 0x1B : 0)); }
    // key schedule (FIPS-197 5.2), 60 words with row 0 in the top byte
    void expand(const uint8_t key[32], uint32_t rk[60]) const {
        for (int i = 0; i < 8; i++)
            rk[i] = ((uint32_t)key[4 * i] << 24) | ((uint32_t)key[4 * i + 1] << 16) |
                    ((uint32_t)key[4 * i + 2] << 8) | key[4 * i + 3];
        uint8_t rcon = 1;
        for (int i = 8; i < 60; i++) {
            uint32_t t = rk[i - 1];
            if (i % 8 == 0) {
                t = (t << 8) | (t >> 24);  // RotWord
                t = sub(t) ^ ((uint32_t)rcon << 24);
                rcon = xt(rcon);
            } else if (i % 8 == 4) {
                t = sub(t);
            }
            rk[i] = rk[i - 8] ^ t;
        }
    }
    uint32_t sub(uint32_t t) const {
        return ((uint32_t)sbox[t >> 24] << 24) | ((uint32_t)sbox[(t >> 16) & 255] << 16) |
               ((uint32_t)sbox[(t >> 8) & 255] << 8) | sbox[t & 255];
    }
    void encrypt_rk(const uint32_t rk[60], const uint8_t in[16], uint8_t out[16]) const {
        uint32_t w[4];
        for (int c = 0; c < 4; c++)
            w[c] = (((uint32_t)in[4 * c] << 24) | ((uint32_t)in[4 * c + 1] << 16) |
                    ((uint32_t)in[4 * c + 2] << 8) | in[4 * c + 3]) ^ rk[c];
        for (int r = 1; r < 14; r++) {  // SubBytes, ShiftRows, MixColumns, AddRoundKey
            uint32_t t[4];
            for (int c = 0; c < 4; c++)
                t[c] = te0[w[c] >> 24] ^ te1[(w[(c + 1) & 3] >> 16) & 255] ^
                       te2[(w[(c + 2) & 3] >> 8) & 255] ^ te3[w[(c + 3) & 3] & 255] ^ rk[4 * r + c];
            for (int c = 0; c < 4; c++) w[c] = t[c];
        }
        for (int c = 0; c < 4; c++) {  // last round: no MixColumns
            const uint32_t t = ((uint32_t)sbox[w[c] >> 24] << 24) |
                               ((uint32_t)sbox[(w[(c + 1) & 3] >> 16) & 255] << 16) |
                               ((uint32_t)sbox[(w[(c + 2) & 3] >> 8) & 255] << 8) |
                               sbox[w[(c + 3) & 3] & 255];
            const uint32_t o = t ^ rk[56 + c];
            out[4 * c] = (uint8_t)(o >> 24);
            out[4 * c + 1] = (uint8_t)(o >> 16);
            out[4 * c + 2] = (uint8_t)(o >> 8);
            out[4 * c + 3] = (uint8_t)o;
        }
    }
};

const Aes256& aes() {
    static const Aes256 a;
    return a;
}

#if defined(__x86_64__)
// AES-256 on the AES instructions.  The key schedule (FIPS-197 5.2) as 15
// round-key blocks of 4 words: block k >= 2 is the running xor of block k - 2
// (w[i] = w[i-8] ^ w[i-1] within a block) xored with one word broadcast to
// all four lanes, SubWord(RotWord(w)) ^ Rcon of block k - 1's last word for
// even k and SubWord of it for odd k.  AESKEYGENASSIST returns exactly
// those two words in its lanes 3 and 2.
#define SPX_AESNI __attribute__((target("aes,sse2")))

SPX_AESNI static inline __m128i ks_next(__m128i two_back, __m128i word) {
    __m128i x = two_back;
    x = _mm_xor_si128(x, _mm_slli_si128(x, 4));
    x = _mm_xor_si128(x, _mm_slli_si128(x, 4));
    x = _mm_xor_si128(x, _mm_slli_si128(x, 4));
    return _mm_xor_si128(x, word);
}

template <int RCON>
SPX_AESNI static inline void ks_pair(__m128i* rk, int k) {
    rk[k] = ks_next(rk[k - 2], _mm_shuffle_epi32(_mm_aeskeygenassist_si128(rk[k - 1], RCON), 0xff));
    if (k + 1 < 15)
        rk[k + 1] = ks_next(rk[k - 1],
                            _mm_shuffle_epi32(_mm_aeskeygenassist_si128(rk[k], 0), 0xaa));
}

SPX_AESNI static void ni_expand(const uint8_t key[32], uint32_t w[60]) {
    __m128i rk[15];
    rk[0] = _mm_loadu_si128((const __m128i*)key);
    rk[1] = _mm_loadu_si128((const __m128i*)(key + 16));
    ks_pair<0x01>(rk, 2);
    ks_pair<0x02>(rk, 4);
    ks_pair<0x04>(rk, 6);
    ks_pair<0x08>(rk, 8);
    ks_pair<0x10>(rk, 10);
    ks_pair<0x20>(rk, 12);
    ks_pair<0x40>(rk, 14);
    memcpy(w, rk, sizeof rk);
}

SPX_AESNI static void ni_encrypt(const uint32_t w[60], const uint8_t in[16], uint8_t out[16]) {
    __m128i rk[15];
    memcpy(rk, w, sizeof rk);
    __m128i x = _mm_xor_si128(_mm_loadu_si128((const __m128i*)in), rk[0]);
    for (int r = 1; r < 14; r++) x = _mm_aesenc_si128(x, rk[r]);
    _mm_storeu_si128((__m128i*)out, _mm_aesenclast_si128(x, rk[14]));
}

static bool cpu_has_aes() { return __builtin_cpu_supports("aes"); }
#else
static void ni_expand(const uint8_t*, uint32_t*) {}
static void ni_encrypt(const uint32_t*, const uint8_t*, uint8_t*) {}
static bool cpu_has_aes() { return false; }
#endif

std::atomic<bool> g_portable{false};

// A key schedule in the form of the encryptor that made it (the AES
// instructions' 15 blocks of key bytes, or the round tables' row-0-first
// words); both are 60 words.
struct Sched {
    alignas(16) uint32_t w[60];
    bool ni;
};

void expand(const uint8_t key[32], Sched& s) {
    static const bool has = cpu_has_aes();
    s.ni = has && !g_portable.load(std::memory_order_relaxed);
    if (s.ni) ni_expand(key, s.w);
    else aes().expand(key, s.w);
}

void encrypt(const Sched& s, const uint8_t in[16], uint8_t out[16]) {
    if (s.ni) ni_encrypt(s.w, in, out);
    else aes().encrypt_rk(s.w, in, out);
}

struct Drbg {
    std::mutex mu;
    uint8_t key[32] = {0};
    uint8_t v[16] = {0};
    int reseed_counter = 0;

    static void inc(uint8_t v[16]) {
        for (int j = 15; j >= 0; j--) {
            if (v[j] == 0xff) v[j] = 0;
            else {
                v[j]++;
                break;
            }
        }
    }
    // AES_CTR_DRBG_Update (rng.cu:217-241) with the current key's schedule
    void update(const Sched& rk, const uint8_t* provided) {
        uint8_t temp[48];
        for (int i = 0; i < 3; i++) {
            inc(v);
            encrypt(rk, v, temp + 16 * i);
        }
        if (provided)
            for (int i = 0; i < 48; i++) temp[i] ^= provided[i];
        memcpy(key, temp, 32);
        memcpy(v, temp + 32, 16);
    }
};

Drbg& drbg() {
    static Drbg d;
    return d;
}

}  // namespace

void aes256_ecb(const uint8_t key[32], const uint8_t in[16], uint8_t out[16]) {
    Sched s;
    expand(key, s);
    encrypt(s, in, out);
}

void aes_force_portable(bool on) { g_portable = on; }

void drbg_init(const uint8_t entropy[48], const uint8_t* pers) {
    Drbg& d = drbg();
    std::lock_guard<std::mutex> lk(d.mu);
    uint8_t seed[48];
    memcpy(seed, entropy, 48);
    if (pers)
        for (int i = 0; i < 48; i++) seed[i] ^= pers[i];
    memset(d.key, 0, 32);
    memset(d.v, 0, 16);
    Sched rk;
    expand(d.key, rk);
    d.update(rk, seed);
    d.reseed_counter = 1;
}

void drbg_bytes(uint8_t* x, uint64_t xlen) {
    Drbg& d = drbg();
    std::lock_guard<std::mutex> lk(d.mu);
    uint8_t block[16];
    Sched rk;  // the key is fixed until the update below: expand it once
    expand(d.key, rk);
    uint64_t i = 0;
    while (xlen > 0) {
        Drbg::inc(d.v);
        encrypt(rk, d.v, block);
        const uint64_t take = xlen > 15 ? 16 : xlen;
        memcpy(x + i, block, take);
        i += take;
        xlen -= take;
    }
    d.update(rk, nullptr);
    d.reseed_counter++;
}

}  // namespace spx
