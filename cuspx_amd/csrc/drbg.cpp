// drbg.cpp -- NIST AES-256-CTR DRBG (no derivation function), the input
// generator the reference uses for keys, optrand and test messages
// (rng.cu:116-241, PQCgenKAT_sign.cu:55-63).  The reference calls OpenSSL's
// EVP AES-256-ECB; this file carries its own FIPS-197 AES-256 encryptor so
// the library has no OpenSSL dependency.  State is process-global like the
// reference's DRBG_ctx (rng.cu:14) but guarded by a mutex.
#include <mutex>
#include <stdint.h>
#include <string.h>

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
    }
    static uint8_t xt(uint8_t a) { return (uint8_t)((a << 1) ^ ((a & 0x80) ? 0x1B : 0)); }
    void encrypt(const uint8_t key[32], const uint8_t in[16], uint8_t out[16]) const {
        uint8_t rk[240];
        memcpy(rk, key, 32);
        uint8_t rcon = 1;
        for (int i = 8; i < 60; i++) {
            uint8_t t[4];
            memcpy(t, rk + 4 * (i - 1), 4);
            if (i % 8 == 0) {
                const uint8_t t0 = t[0];
                t[0] = sbox[t[1]] ^ rcon;
                t[1] = sbox[t[2]];
                t[2] = sbox[t[3]];
                t[3] = sbox[t0];
                rcon = xt(rcon);
            } else if (i % 8 == 4) {
                for (int j = 0; j < 4; j++) t[j] = sbox[t[j]];
            }
            for (int j = 0; j < 4; j++) rk[4 * i + j] = rk[4 * (i - 8) + j] ^ t[j];
        }
        uint8_t s[16];
        for (int i = 0; i < 16; i++) s[i] = in[i] ^ rk[i];
        for (int r = 1; r <= 14; r++) {
            uint8_t t[16];
            for (int c = 0; c < 4; c++)
                for (int row = 0; row < 4; row++) t[4 * c + row] = sbox[s[4 * ((c + row) & 3) + row]];
            if (r != 14) {
                for (int c = 0; c < 4; c++) {
                    uint8_t* col = t + 4 * c;
                    const uint8_t a0 = col[0], a1 = col[1], a2 = col[2], a3 = col[3];
                    const uint8_t all = a0 ^ a1 ^ a2 ^ a3;
                    col[0] = a0 ^ all ^ xt(a0 ^ a1);
                    col[1] = a1 ^ all ^ xt(a1 ^ a2);
                    col[2] = a2 ^ all ^ xt(a2 ^ a3);
                    col[3] = a3 ^ all ^ xt(a3 ^ a0);
                }
            }
            for (int i = 0; i < 16; i++) s[i] = t[i] ^ rk[16 * r + i];
        }
        memcpy(out, s, 16);
    }
};

const Aes256& aes() {
    static const Aes256 a;
    return a;
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
    void update(const uint8_t* provided) {
        uint8_t temp[48];
        for (int i = 0; i < 3; i++) {
            inc(v);
            aes().encrypt(key, v, temp + 16 * i);
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
    aes().encrypt(key, in, out);
}

void drbg_init(const uint8_t entropy[48], const uint8_t* pers) {
    Drbg& d = drbg();
    std::lock_guard<std::mutex> lk(d.mu);
    uint8_t seed[48];
    memcpy(seed, entropy, 48);
    if (pers)
        for (int i = 0; i < 48; i++) seed[i] ^= pers[i];
    memset(d.key, 0, 32);
    memset(d.v, 0, 16);
    d.update(seed);
    d.reseed_counter = 1;
}

void drbg_bytes(uint8_t* x, uint64_t xlen) {
    Drbg& d = drbg();
    std::lock_guard<std::mutex> lk(d.mu);
    uint8_t block[16];
    uint64_t i = 0;
    while (xlen > 0) {
        Drbg::inc(d.v);
        aes().encrypt(d.key, d.v, block);
        const uint64_t take = xlen > 15 ? 16 : xlen;
        memcpy(x + i, block, take);
        i += take;
        xlen -= take;
    }
    d.update(nullptr);
    d.reseed_counter++;
}

}  // namespace spx
