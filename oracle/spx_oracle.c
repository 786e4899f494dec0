/*
 * spx_oracle.c -- CPU oracle (test infrastructure, never the product path).
 *
 * A from-scratch C restatement of the reference host path of CUSPX
 * (SPHINCS+ round-3 "simple", SHA-256 and SHAKE256 families), with the
 * parameter set chosen at run time.  Behaviour that must match the reference
 * byte for byte, including its unusual choices:
 *   - both hash families use the SHA-256 22-byte address offsets
 *     (params.h:148 includes sha256_offsets.h unconditionally); SHAKE absorbs
 *     all 32 address bytes, of which bytes 22..31 stay zero;
 *   - DEBUG_MODE (all_option.h:74) replaces the keygen seed with 0..3n-1
 *     (sign.cu:345-349) and optrand with 0xad (sign.cu:1950-1954) after the
 *     DRBG bytes were drawn; switchable here with spxo_set_debug_mode;
 *   - round-3 SHA-256 constructions: PRF = SHA256(key||addr22), PRF_msg =
 *     HMAC-SHA256, H_msg = MGF1(SHA256(R||pk||m)), SHA-256 for every n;
 *   - FORS indices LSB-first within each byte (fors.cu:77-91).
 *
 * Parity is pinned per function by the checks listed in spx_oracle.h.
 * End-to-end parity with the reference binary is unpinned: the reference
 * ships no known-answer vector and cannot be compiled here.
 */
#include "spx_oracle.h"

#include <stdio.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------ */
/* parameter registry (params.h:33-141)                                */
/* ------------------------------------------------------------------ */

static const struct { const char *sz; int n, h, d, a, k; } kSizes[6] = {
    {"128s", 16, 63, 7, 12, 14},  /* params.h:33-39 */
    {"128f", 16, 66, 22, 6, 33},  /* params.h:41-47 */
    {"192s", 24, 63, 7, 14, 17},  /* params.h:49-55 */
    {"192f", 24, 66, 22, 8, 33},  /* params.h:57-63 */
    {"256s", 32, 64, 8, 14, 22},  /* params.h:65-71 */
    {"256f", 32, 68, 17, 9, 35},  /* params.h:73-79 */
};

static spxo_params g_params[12];
static int g_params_ready = 0;

static void init_params(void)
{
    if (g_params_ready) return;
    for (int f = 0; f < 2; f++) {
        for (int s = 0; s < 6; s++) {
            spxo_params *p = &g_params[f * 6 + s];
            memset(p, 0, sizeof(*p));
            p->hash = f;
            p->n = kSizes[s].n;
            p->full_h = kSizes[s].h;
            p->d = kSizes[s].d;
            p->a = kSizes[s].a;
            p->k = kSizes[s].k;
            p->w = 16;
            p->logw = 4;                                   /* params.h:91-93 */
            p->len1 = 8 * p->n / p->logw;                  /* params.h:101 */
            p->len2 = p->n <= 8 ? 2 : (p->n <= 136 ? 3 : 4); /* params.h:114-123 */
            p->len = p->len1 + p->len2;
            p->tree_h = p->full_h / p->d;                  /* params.h:130 */
            p->wots_bytes = p->len * p->n;
            p->fors_msg_bytes = (p->a * p->k + 7) / 8;     /* params.h:137 */
            p->fors_bytes = (p->a + 1) * p->k * p->n;      /* params.h:138 */
            p->sig_bytes = p->n + p->fors_bytes + p->d * p->wots_bytes + p->full_h * p->n;
            p->pk_bytes = 2 * p->n;
            p->sk_bytes = 4 * p->n;
            p->tree_bits = p->tree_h * (p->d - 1);         /* hash_sha256.cu:209 */
            p->tree_bytes = (p->tree_bits + 7) / 8;
            p->leaf_bits = p->tree_h;
            p->leaf_bytes = (p->leaf_bits + 7) / 8;
            p->dgst_bytes = p->fors_msg_bytes + p->tree_bytes + p->leaf_bytes;
            snprintf(p->name, sizeof(p->name), "%s-%s", f ? "shake256" : "sha2", kSizes[s].sz);
        }
    }
    g_params_ready = 1;
}

int spxo_num_params(void) { return 12; }

int spxo_get_params(int pid, spxo_params *out)
{
    init_params();
    if (pid < 0 || pid >= 12) return -1;
    *out = g_params[pid];
    return 0;
}

int spxo_param_id(const char *name)
{
    init_params();
    for (int i = 0; i < 12; i++)
        if (strcmp(name, g_params[i].name) == 0) return i;
    return -1;
}

static const spxo_params *P(int pid)
{
    init_params();
    if (pid < 0 || pid >= 12) {
        fprintf(stderr, "spx_oracle: bad parameter id %d\n", pid);
        abort();
    }
    return &g_params[pid];
}

static int g_debug_mode = 1; /* as shipped: all_option.h:74 defines DEBUG_MODE */
void spxo_set_debug_mode(int on) { g_debug_mode = on ? 1 : 0; }
int spxo_get_debug_mode(void) { return g_debug_mode; }

static uint64_t g_cnt_sha = 0, g_cnt_keccak = 0;
void spxo_reset_counts(void) { g_cnt_sha = 0; g_cnt_keccak = 0; }
uint64_t spxo_count_sha256_blocks(void) { return g_cnt_sha; }
uint64_t spxo_count_keccak_perms(void) { return g_cnt_keccak; }

/* ------------------------------------------------------------------ */
/* byte helpers (utils.cu:33-89)                                       */
/* ------------------------------------------------------------------ */

static void ull_to_bytes(uint8_t *out, unsigned outlen, uint64_t in)
{
    for (int i = (int)outlen - 1; i >= 0; i--) { /* utils.cu:33-41 */
        out[i] = (uint8_t)(in & 0xff);
        in >>= 8;
    }
}

static void u32_to_bytes(uint8_t *out, uint32_t in)
{
    out[0] = (uint8_t)(in >> 24); /* utils.cu:53-58 */
    out[1] = (uint8_t)(in >> 16);
    out[2] = (uint8_t)(in >> 8);
    out[3] = (uint8_t)in;
}

static uint64_t bytes_to_ull(const uint8_t *in, unsigned inlen)
{
    uint64_t r = 0; /* utils.cu:70-78 */
    for (unsigned i = 0; i < inlen; i++) r |= ((uint64_t)in[i]) << (8 * (inlen - 1 - i));
    return r;
}

/* ------------------------------------------------------------------ */
/* SHA-256 (sha256.cu:14-405), FIPS 180-4                              */
/* ------------------------------------------------------------------ */

static const uint32_t K256[64] = {
    0x428a2f98, 0x71374491, 0xb5c0fbcf, 0xe9b5dba5, 0x3956c25b, 0x59f111f1, 0x923f82a4, 0xab1c5ed5,
    0xd807aa98, 0x12835b01, 0x243185be, 0x550c7dc3, 0x72be5d74, 0x80deb1fe, 0x9bdc06a7, 0xc19bf174,
    0xe49b69c1, 0xefbe4786, 0x0fc19dc6, 0x240ca1cc, 0x2de92c6f, 0x4a7484aa, 0x5cb0a9dc, 0x76f988da,
    0x983e5152, 0xa831c66d, 0xb00327c8, 0xbf597fc7, 0xc6e00bf3, 0xd5a79147, 0x06ca6351, 0x14292967,
    0x27b70a85, 0x2e1b2138, 0x4d2c6dfc, 0x53380d13, 0x650a7354, 0x766a0abb, 0x81c2c92e, 0x92722c85,
    0xa2bfe8a1, 0xa81a664b, 0xc24b8b70, 0xc76c51a3, 0xd192e819, 0xd6990624, 0xf40e3585, 0x106aa070,
    0x19a4c116, 0x1e376c08, 0x2748774c, 0x34b0bcb5, 0x391c0cb3, 0x4ed8aa4a, 0x5b9cca4f, 0x682e6ff3,
    0x748f82ee, 0x78a5636f, 0x84c87814, 0x8cc70208, 0x90befffa, 0xa4506ceb, 0xbef9a3f7, 0xc67178f2};

static const uint8_t IV256[32] = { /* sha256.cu:261-266 */
    0x6a, 0x09, 0xe6, 0x67, 0xbb, 0x67, 0xae, 0x85, 0x3c, 0x6e, 0xf3, 0x72, 0xa5, 0x4f, 0xf5, 0x3a,
    0x51, 0x0e, 0x52, 0x7f, 0x9b, 0x05, 0x68, 0x8c, 0x1f, 0x83, 0xd9, 0xab, 0x5b, 0xe0, 0xcd, 0x19};

#define ROR32(x, c) (((x) >> (c)) | ((x) << (32 - (c))))

void spxo_sha256_compress_words(uint32_t st[8], const uint32_t win[16])
{
    uint32_t w[64];
    for (int i = 0; i < 16; i++) w[i] = win[i];
    for (int i = 16; i < 64; i++) {
        uint32_t s0 = ROR32(w[i - 15], 7) ^ ROR32(w[i - 15], 18) ^ (w[i - 15] >> 3);
        uint32_t s1 = ROR32(w[i - 2], 17) ^ ROR32(w[i - 2], 19) ^ (w[i - 2] >> 10);
        w[i] = s1 + w[i - 7] + s0 + w[i - 16];
    }
    uint32_t a = st[0], b = st[1], c = st[2], d = st[3], e = st[4], f = st[5], g = st[6], h = st[7];
    for (int i = 0; i < 64; i++) {
        uint32_t t1 = h + (ROR32(e, 6) ^ ROR32(e, 11) ^ ROR32(e, 25)) + ((e & f) ^ (~e & g)) +
                      K256[i] + w[i];
        uint32_t t2 = (ROR32(a, 2) ^ ROR32(a, 13) ^ ROR32(a, 22)) + ((a & b) ^ (a & c) ^ (b & c));
        h = g; g = f; f = e; e = d + t1; d = c; c = b; b = a; a = t1 + t2;
    }
    st[0] += a; st[1] += b; st[2] += c; st[3] += d;
    st[4] += e; st[5] += f; st[6] += g; st[7] += h;
    g_cnt_sha++;
}

/* crypto_hashblocks_sha256 (sha256.cu:103-259): state as 32 BE bytes */
static void hashblocks(uint8_t *statebytes, const uint8_t *in, size_t inlen)
{
    uint32_t st[8], w[16];
    for (int i = 0; i < 8; i++)
        st[i] = (uint32_t)statebytes[4 * i] << 24 | (uint32_t)statebytes[4 * i + 1] << 16 |
                (uint32_t)statebytes[4 * i + 2] << 8 | statebytes[4 * i + 3];
    while (inlen >= 64) {
        for (int i = 0; i < 16; i++)
            w[i] = (uint32_t)in[4 * i] << 24 | (uint32_t)in[4 * i + 1] << 16 |
                   (uint32_t)in[4 * i + 2] << 8 | in[4 * i + 3];
        spxo_sha256_compress_words(st, w);
        in += 64;
        inlen -= 64;
    }
    for (int i = 0; i < 8; i++) u32_to_bytes(statebytes + 4 * i, st[i]);
}

/* sha256_inc_init / _blocks / _finalize (sha256.cu:268-347): 40-byte state =
 * 32-byte BE chaining value || 8-byte BE byte count */
static void sha256_inc_init(uint8_t st[40])
{
    memcpy(st, IV256, 32);
    memset(st + 32, 0, 8);
}

static void sha256_inc_blocks(uint8_t st[40], const uint8_t *in, size_t inblocks)
{
    uint64_t bytes = bytes_to_ull(st + 32, 8);
    hashblocks(st, in, 64 * inblocks);
    bytes += 64 * inblocks;
    ull_to_bytes(st + 32, 8, bytes);
}

static void sha256_inc_finalize(uint8_t out[32], uint8_t st[40], const uint8_t *in, size_t inlen)
{
    uint8_t padded[128];
    uint64_t bytes = bytes_to_ull(st + 32, 8) + inlen;
    hashblocks(st, in, inlen);
    in += inlen;
    inlen &= 63;
    in -= inlen;
    memset(padded, 0, sizeof(padded));
    memcpy(padded, in, inlen);
    padded[inlen] = 0x80;
    if (inlen < 56) {
        ull_to_bytes(padded + 56, 8, bytes << 3);
        hashblocks(st, padded, 64);
    } else {
        ull_to_bytes(padded + 120, 8, bytes << 3);
        hashblocks(st, padded, 128);
    }
    memcpy(out, st, 32);
}

void spxo_sha256(uint8_t out[32], const uint8_t *in, size_t inlen)
{
    uint8_t st[40]; /* sha256.cu:349-355 */
    sha256_inc_init(st);
    sha256_inc_finalize(out, st, in, inlen);
}

void spxo_mgf1(uint8_t *out, size_t outlen, const uint8_t *in, size_t inlen)
{
    /* sha256.cu:362-383: 4-byte BE counter appended to the seed */
    uint8_t *inbuf = (uint8_t *)malloc(inlen + 4);
    uint8_t outbuf[32];
    size_t i;
    memcpy(inbuf, in, inlen);
    for (i = 0; (i + 1) * 32 <= outlen; i++) {
        u32_to_bytes(inbuf + inlen, (uint32_t)i);
        spxo_sha256(out, inbuf, inlen + 4);
        out += 32;
    }
    if (outlen > i * 32) {
        u32_to_bytes(inbuf + inlen, (uint32_t)i);
        spxo_sha256(outbuf, inbuf, inlen + 4);
        memcpy(out, outbuf, outlen - i * 32);
    }
    free(inbuf);
}

void spxo_hmac_sha256(uint8_t out[32], const uint8_t *key, size_t keylen, const uint8_t *msg,
                      size_t msglen)
{
    /* textbook HMAC (key <= 64 bytes), used to cross-check gen_message_random */
    uint8_t k0[64], st[40], inner[32], buf[96];
    memset(k0, 0, 64);
    memcpy(k0, key, keylen);
    for (int i = 0; i < 64; i++) buf[i] = k0[i] ^ 0x36;
    sha256_inc_init(st);
    sha256_inc_blocks(st, buf, 1);
    sha256_inc_finalize(inner, st, msg, msglen);
    for (int i = 0; i < 64; i++) buf[i] = k0[i] ^ 0x5c;
    memcpy(buf + 64, inner, 32);
    spxo_sha256(out, buf, 96);
}

/* seed_state (sha256.cu:391-405): absorb pub_seed || 0^(64-n) once */
void spxo_seed_state(int pid, uint8_t state[40], const uint8_t *pub_seed)
{
    const spxo_params *p = P(pid);
    uint8_t block[64];
    memset(block, 0, 64);
    memcpy(block, pub_seed, p->n);
    sha256_inc_init(state);
    sha256_inc_blocks(state, block, 1);
}

/* ------------------------------------------------------------------ */
/* Keccak-f[1600] and SHAKE256 (fips202.cu:25-676), FIPS 202            */
/* ------------------------------------------------------------------ */

static const uint64_t KeccakRC[24] = {
    0x0000000000000001ULL, 0x0000000000008082ULL, 0x800000000000808aULL, 0x8000000080008000ULL,
    0x000000000000808bULL, 0x0000000080000001ULL, 0x8000000080008081ULL, 0x8000000000008009ULL,
    0x000000000000008aULL, 0x0000000000000088ULL, 0x0000000080008009ULL, 0x000000008000000aULL,
    0x000000008000808bULL, 0x800000000000008bULL, 0x8000000000008089ULL, 0x8000000000008003ULL,
    0x8000000000008002ULL, 0x8000000000000080ULL, 0x000000000000800aULL, 0x800000008000000aULL,
    0x8000000080008081ULL, 0x8000000000008080ULL, 0x0000000080000001ULL, 0x8000000080008008ULL};

static const int KeccakRho[25] = {0,  1,  62, 28, 27, 36, 44, 6,  55, 20, 3,  10, 43,
                                  25, 39, 41, 45, 15, 21, 8,  18, 2,  61, 56, 14};

#define ROL64(x, c) ((c) ? (((x) << (c)) | ((x) >> (64 - (c)))) : (x))

void spxo_keccakf1600(uint64_t s[25])
{
    /* straightforward theta / rho+pi / chi / iota on A[x + 5y] */
    for (int r = 0; r < 24; r++) {
        uint64_t C[5], D[5], B[25];
        for (int x = 0; x < 5; x++) C[x] = s[x] ^ s[x + 5] ^ s[x + 10] ^ s[x + 15] ^ s[x + 20];
        for (int x = 0; x < 5; x++) D[x] = C[(x + 4) % 5] ^ ROL64(C[(x + 1) % 5], 1);
        for (int i = 0; i < 25; i++) s[i] ^= D[i % 5];
        for (int x = 0; x < 5; x++)
            for (int y = 0; y < 5; y++) {
                /* pi: B[y, 2x+3y] = rot(A[x,y], rho[x,y]) */
                int nx = y, ny = (2 * x + 3 * y) % 5;
                B[nx + 5 * ny] = ROL64(s[x + 5 * y], KeccakRho[x + 5 * y]);
            }
        for (int y = 0; y < 5; y++)
            for (int x = 0; x < 5; x++)
                s[x + 5 * y] = B[x + 5 * y] ^ ((~B[(x + 1) % 5 + 5 * y]) & B[(x + 2) % 5 + 5 * y]);
        s[0] ^= KeccakRC[r];
    }
    g_cnt_keccak++;
}

/* incremental SHAKE256 (fips202.cu:415-496), rate 136, domain byte 0x1F */
typedef struct { uint64_t s[25]; size_t pos; } shake_ctx;

static void shake_init(shake_ctx *c) { memset(c, 0, sizeof(*c)); }

static void shake_absorb(shake_ctx *c, const uint8_t *m, size_t mlen)
{
    for (size_t i = 0; i < mlen; i++) {
        c->s[c->pos >> 3] ^= (uint64_t)m[i] << (8 * (c->pos & 7));
        if (++c->pos == 136) {
            spxo_keccakf1600(c->s);
            c->pos = 0;
        }
    }
}

static void shake_finalize(shake_ctx *c)
{
    c->s[c->pos >> 3] ^= (uint64_t)0x1F << (8 * (c->pos & 7));
    c->s[135 >> 3] ^= (uint64_t)0x80 << (8 * (135 & 7));
    c->pos = 136; /* nothing squeezed yet */
}

static void shake_squeeze(shake_ctx *c, uint8_t *out, size_t outlen)
{
    for (size_t i = 0; i < outlen; i++) {
        if (c->pos == 136) {
            spxo_keccakf1600(c->s);
            c->pos = 0;
        }
        out[i] = (uint8_t)(c->s[c->pos >> 3] >> (8 * (c->pos & 7)));
        c->pos++;
    }
}

void spxo_shake256(uint8_t *out, size_t outlen, const uint8_t *in, size_t inlen)
{
    shake_ctx c; /* fips202.cu:658-676 */
    shake_init(&c);
    shake_absorb(&c, in, inlen);
    shake_finalize(&c);
    shake_squeeze(&c, out, outlen);
}

/* ------------------------------------------------------------------ */
/* AES-256 (FIPS-197) and the NIST AES-256-CTR DRBG (rng.cu:116-241)   */
/* ------------------------------------------------------------------ */

static uint8_t aes_sbox[256];
static int aes_ready = 0;

static uint8_t gf_mul(uint8_t a, uint8_t b)
{
    uint8_t r = 0;
    while (b) {
        if (b & 1) r ^= a;
        a = (uint8_t)((a << 1) ^ ((a & 0x80) ? 0x1b : 0));
        b >>= 1;
    }
    return r;
}

static void aes_init(void)
{
    /* S-box from its definition: multiplicative inverse in GF(2^8) then the
     * affine map b ^ rotl(b,1..4) ^ 0x63 (FIPS-197 5.1.1) */
    if (aes_ready) return;
    for (int x = 0; x < 256; x++) {
        uint8_t inv = 0;
        if (x) {
            for (int y = 1; y < 256; y++)
                if (gf_mul((uint8_t)x, (uint8_t)y) == 1) { inv = (uint8_t)y; break; }
        }
        uint8_t b = inv, s = inv;
        for (int i = 0; i < 4; i++) {
            b = (uint8_t)((b << 1) | (b >> 7));
            s ^= b;
        }
        aes_sbox[x] = s ^ 0x63;
    }
    aes_ready = 1;
}

void spxo_aes256_ecb(const uint8_t key[32], const uint8_t in[16], uint8_t out[16])
{
    uint8_t rk[240], st[16];
    aes_init();
    memcpy(rk, key, 32);
    uint8_t rcon = 1;
    for (int i = 8; i < 60; i++) { /* FIPS-197 5.2 key expansion, Nk = 8 */
        uint8_t t[4];
        memcpy(t, rk + 4 * (i - 1), 4);
        if (i % 8 == 0) {
            uint8_t t0 = t[0];
            t[0] = aes_sbox[t[1]] ^ rcon;
            t[1] = aes_sbox[t[2]];
            t[2] = aes_sbox[t[3]];
            t[3] = aes_sbox[t0];
            rcon = gf_mul(rcon, 2);
        } else if (i % 8 == 4) {
            for (int j = 0; j < 4; j++) t[j] = aes_sbox[t[j]];
        }
        for (int j = 0; j < 4; j++) rk[4 * i + j] = rk[4 * (i - 8) + j] ^ t[j];
    }
    for (int i = 0; i < 16; i++) st[i] = in[i] ^ rk[i];
    for (int r = 1; r <= 14; r++) {
        uint8_t t[16];
        for (int i = 0; i < 16; i++) st[i] = aes_sbox[st[i]];
        for (int c = 0; c < 4; c++) /* ShiftRows: row r shifts left by r */
            for (int rr = 0; rr < 4; rr++) t[4 * c + rr] = st[4 * ((c + rr) % 4) + rr];
        if (r != 14) {
            for (int c = 0; c < 4; c++) { /* MixColumns */
                uint8_t *col = t + 4 * c, a0 = col[0], a1 = col[1], a2 = col[2], a3 = col[3];
                col[0] = gf_mul(a0, 2) ^ gf_mul(a1, 3) ^ a2 ^ a3;
                col[1] = a0 ^ gf_mul(a1, 2) ^ gf_mul(a2, 3) ^ a3;
                col[2] = a0 ^ a1 ^ gf_mul(a2, 2) ^ gf_mul(a3, 3);
                col[3] = gf_mul(a0, 3) ^ a1 ^ a2 ^ gf_mul(a3, 2);
            }
        }
        for (int i = 0; i < 16; i++) st[i] = t[i] ^ rk[16 * r + i];
    }
    memcpy(out, st, 16);
}

static struct { uint8_t key[32], v[16]; int reseed_counter; } g_drbg;

static void drbg_inc_v(uint8_t v[16])
{
    for (int j = 15; j >= 0; j--) { /* rng.cu:224-231 */
        if (v[j] == 0xff) v[j] = 0x00;
        else { v[j]++; break; }
    }
}

static void drbg_update(const uint8_t *provided, uint8_t key[32], uint8_t v[16])
{
    uint8_t temp[48]; /* AES256_CTR_DRBG_Update, rng.cu:217-241 */
    for (int i = 0; i < 3; i++) {
        drbg_inc_v(v);
        spxo_aes256_ecb(key, v, temp + 16 * i);
    }
    if (provided)
        for (int i = 0; i < 48; i++) temp[i] ^= provided[i];
    memcpy(key, temp, 32);
    memcpy(v, temp + 32, 16);
}

void spxo_randombytes_init(const uint8_t entropy[48], const uint8_t *pers, int strength)
{
    uint8_t seed[48]; /* rng.cu:139-154 */
    (void)strength;
    memcpy(seed, entropy, 48);
    if (pers)
        for (int i = 0; i < 48; i++) seed[i] ^= pers[i];
    memset(g_drbg.key, 0, 32);
    memset(g_drbg.v, 0, 16);
    drbg_update(seed, g_drbg.key, g_drbg.v);
    g_drbg.reseed_counter = 1;
}

int spxo_randombytes(uint8_t *x, uint64_t xlen)
{
    uint8_t block[16]; /* rng.cu:156-184 */
    uint64_t i = 0;
    while (xlen > 0) {
        drbg_inc_v(g_drbg.v);
        spxo_aes256_ecb(g_drbg.key, g_drbg.v, block);
        if (xlen > 15) {
            memcpy(x + i, block, 16);
            i += 16;
            xlen -= 16;
        } else {
            memcpy(x + i, block, xlen);
            xlen = 0;
        }
    }
    drbg_update(NULL, g_drbg.key, g_drbg.v);
    g_drbg.reseed_counter++;
    return 0;
}

/* ------------------------------------------------------------------ */
/* addresses (address.cu:10-175, sha256_offsets.h:9-17)                */
/* ------------------------------------------------------------------ */

enum { OFF_LAYER = 0, OFF_TREE = 1, OFF_TYPE = 9, OFF_KP2 = 12, OFF_KP1 = 13, OFF_CHAIN = 17,
       OFF_HASH = 21, OFF_HGT = 17, OFF_IDX = 18 };
enum { T_WOTS = 0, T_WOTSPK = 1, T_HASHTREE = 2, T_FORSTREE = 3, T_FORSPK = 4 };

static uint8_t *AB(uint32_t a[8]) { return (uint8_t *)a; }
static const uint8_t *ABc(const uint32_t a[8]) { return (const uint8_t *)a; }

static void set_layer(uint32_t a[8], uint32_t l) { AB(a)[OFF_LAYER] = (uint8_t)l; }
static void set_tree(uint32_t a[8], uint64_t t) { ull_to_bytes(AB(a) + OFF_TREE, 8, t); }
static void set_type(uint32_t a[8], uint32_t t) { AB(a)[OFF_TYPE] = (uint8_t)t; }
static void copy_subtree(uint32_t o[8], const uint32_t i[8]) { memcpy(o, i, OFF_TREE + 8); }
static void set_keypair(const spxo_params *p, uint32_t a[8], uint32_t kp)
{
    if (p->tree_h > 8) AB(a)[OFF_KP2] = (uint8_t)(kp >> 8); /* address.cu:81-86 */
    AB(a)[OFF_KP1] = (uint8_t)kp;
}
static void copy_keypair(const spxo_params *p, uint32_t o[8], const uint32_t i[8])
{
    memcpy(o, i, OFF_TREE + 8); /* address.cu:103-111 */
    if (p->tree_h > 8) AB(o)[OFF_KP2] = ABc(i)[OFF_KP2];
    AB(o)[OFF_KP1] = ABc(i)[OFF_KP1];
}
static void set_chain(uint32_t a[8], uint32_t c) { AB(a)[OFF_CHAIN] = (uint8_t)c; }
static void set_hash(uint32_t a[8], uint32_t h) { AB(a)[OFF_HASH] = (uint8_t)h; }
static void set_height(uint32_t a[8], uint32_t h) { AB(a)[OFF_HGT] = (uint8_t)h; }
static void set_index(uint32_t a[8], uint32_t i) { u32_to_bytes(AB(a) + OFF_IDX, i); }

void spxo_addr_set(uint32_t addr[8], int field, uint64_t value, int pid)
{
    const spxo_params *p = P(pid);
    switch (field) {
    case 0: set_layer(addr, (uint32_t)value); break;
    case 1: set_tree(addr, value); break;
    case 2: set_type(addr, (uint32_t)value); break;
    case 3: set_keypair(p, addr, (uint32_t)value); break;
    case 4: set_chain(addr, (uint32_t)value); break;
    case 5: set_hash(addr, (uint32_t)value); break;
    case 6: set_height(addr, (uint32_t)value); break;
    case 7: set_index(addr, (uint32_t)value); break;
    default: break;
    }
}

/* ------------------------------------------------------------------ */
/* tweakable hash layer                                                 */
/* ------------------------------------------------------------------ */

/* The reference keeps one pub_seed midstate (state_seeded, sha256.cu:385),
 * set by initialize_hash_function at the top of each operation
 * (sign.cu:173,1942,3069) and read by every thash (thash_sha256.cu:23-26).
 * Here the midstate is cached per pub_seed; its one compression per
 * operation is counted by count_midstate() below, not here. */
static struct { int n; uint8_t seed[32]; uint8_t st[40]; int valid; } g_mid;

static void cached_midstate(int pid, uint8_t st[40], const uint8_t *pub_seed)
{
    const spxo_params *p = P(pid);
    if (!g_mid.valid || g_mid.n != p->n || memcmp(g_mid.seed, pub_seed, p->n) != 0) {
        uint64_t saved = g_cnt_sha;
        spxo_seed_state(pid, g_mid.st, pub_seed);
        g_cnt_sha = saved;
        g_mid.n = p->n;
        memcpy(g_mid.seed, pub_seed, p->n);
        g_mid.valid = 1;
    }
    memcpy(st, g_mid.st, 40);
}

/* thash: SHA2 (thash_sha256.cu:16-34) = SHA256(midstate(pub_seed); addr22||in)[:n];
 *        SHAKE (thash_shake256.cu:15-25) = SHAKE256(pub_seed||addr32||in, n) */
void spxo_thash(int pid, uint8_t *out, const uint8_t *in, unsigned inblocks,
                const uint8_t *pub_seed, const uint32_t addr[8])
{
    const spxo_params *p = P(pid);
    size_t inl = (size_t)inblocks * p->n;
    if (p->hash == SPXO_SHA2) {
        uint8_t st[40], outbuf[32];
        uint8_t *buf = (uint8_t *)malloc(22 + inl);
        cached_midstate(pid, st, pub_seed);
        memcpy(buf, addr, 22);
        memcpy(buf + 22, in, inl);
        sha256_inc_finalize(outbuf, st, buf, 22 + inl);
        memcpy(out, outbuf, p->n);
        free(buf);
    } else {
        uint8_t *buf = (uint8_t *)malloc(p->n + 32 + inl);
        memcpy(buf, pub_seed, p->n);
        memcpy(buf + p->n, addr, 32);
        memcpy(buf + p->n + 32, in, inl);
        spxo_shake256(out, p->n, buf, p->n + 32 + inl);
        free(buf);
    }
}

/* prf_addr: SHA2 (hash_sha256.cu:65-76) = SHA256(key||addr22)[:n];
 *           SHAKE (hash_shake256.cu:38-47) = SHAKE256(key||addr32, n) */
void spxo_prf_addr(int pid, uint8_t *out, const uint8_t *key, const uint32_t addr[8])
{
    const spxo_params *p = P(pid);
    uint8_t buf[32 + 32], outbuf[32];
    if (p->hash == SPXO_SHA2) {
        memcpy(buf, key, p->n);
        memcpy(buf + p->n, addr, 22);
        spxo_sha256(outbuf, buf, p->n + 22);
        memcpy(out, outbuf, p->n);
    } else {
        memcpy(buf, key, p->n);
        memcpy(buf + p->n, addr, 32);
        spxo_shake256(out, p->n, buf, p->n + 32);
    }
}

/* gen_message_random: SHA2 (hash_sha256.cu:99-144) HMAC-SHA256(sk_prf, optrand||m)[:n]
 * following the reference's block split; SHAKE (hash_shake256.cu:64-75)
 * SHAKE256(sk_prf||optrand||m, n) */
void spxo_gen_message_random(int pid, uint8_t *R, const uint8_t *sk_prf, const uint8_t *optrand,
                             const uint8_t *m, uint64_t mlen)
{
    const spxo_params *p = P(pid);
    if (p->hash == SPXO_SHA2) {
        uint8_t buf[64 + 32], st[40];
        for (int i = 0; i < p->n; i++) buf[i] = 0x36 ^ sk_prf[i];
        memset(buf + p->n, 0x36, 64 - p->n);
        sha256_inc_init(st);
        sha256_inc_blocks(st, buf, 1);
        memcpy(buf, optrand, p->n);
        if ((uint64_t)p->n + mlen < 64) {
            memcpy(buf + p->n, m, mlen);
            sha256_inc_finalize(buf + 64, st, buf, mlen + p->n);
        } else {
            memcpy(buf + p->n, m, 64 - p->n);
            sha256_inc_blocks(st, buf, 1);
            m += 64 - p->n;
            mlen -= 64 - p->n;
            sha256_inc_finalize(buf + 64, st, m, mlen);
        }
        for (int i = 0; i < p->n; i++) buf[i] = 0x5c ^ sk_prf[i];
        memset(buf + p->n, 0x5c, 64 - p->n);
        spxo_sha256(buf, buf, 96);
        memcpy(R, buf, p->n);
    } else {
        shake_ctx c;
        shake_init(&c);
        shake_absorb(&c, sk_prf, p->n);
        shake_absorb(&c, optrand, p->n);
        shake_absorb(&c, m, mlen);
        shake_finalize(&c);
        shake_squeeze(&c, R, p->n);
    }
}

/* hash_message: SHA2 (hash_sha256.cu:200-262), SHAKE (hash_shake256.cu:96-130) */
void spxo_hash_message(int pid, uint8_t *digest, uint64_t *tree, uint32_t *leaf_idx,
                       const uint8_t *R, const uint8_t *pk, const uint8_t *m, uint64_t mlen)
{
    const spxo_params *p = P(pid);
    uint8_t buf[80];
    if (p->hash == SPXO_SHA2) {
        uint8_t seed[32], st[40];
        int inblocks = (p->n + p->pk_bytes + 63) / 64; /* SPX_INBLOCKS */
        uint8_t inbuf[128];
        sha256_inc_init(st);
        memcpy(inbuf, R, p->n);
        memcpy(inbuf + p->n, pk, p->pk_bytes);
        if ((uint64_t)p->n + p->pk_bytes + mlen < (uint64_t)inblocks * 64) {
            memcpy(inbuf + p->n + p->pk_bytes, m, mlen);
            sha256_inc_finalize(seed, st, inbuf, p->n + p->pk_bytes + mlen);
        } else {
            size_t fill = (size_t)inblocks * 64 - p->n - p->pk_bytes;
            memcpy(inbuf + p->n + p->pk_bytes, m, fill);
            sha256_inc_blocks(st, inbuf, inblocks);
            sha256_inc_finalize(seed, st, m + fill, mlen - fill);
        }
        spxo_mgf1(buf, p->dgst_bytes, seed, 32);
    } else {
        shake_ctx c;
        shake_init(&c);
        shake_absorb(&c, R, p->n);
        shake_absorb(&c, pk, p->pk_bytes);
        shake_absorb(&c, m, mlen);
        shake_finalize(&c);
        shake_squeeze(&c, buf, p->dgst_bytes);
    }
    memcpy(digest, buf, p->fors_msg_bytes);
    uint8_t *bp = buf + p->fors_msg_bytes;
    uint64_t t = bytes_to_ull(bp, p->tree_bytes);
    t &= (~(uint64_t)0) >> (64 - p->tree_bits);
    *tree = t;
    bp += p->tree_bytes;
    uint32_t l = (uint32_t)bytes_to_ull(bp, p->leaf_bytes);
    l &= (~(uint32_t)0) >> (32 - p->leaf_bits);
    *leaf_idx = l;
}

/* ------------------------------------------------------------------ */
/* WOTS+ (wots.cu:27-176, 614-790)                                      */
/* ------------------------------------------------------------------ */

static void wots_gen_sk(int pid, uint8_t *sk, const uint8_t *sk_seed, uint32_t addr[8])
{
    set_hash(addr, 0); /* wots.cu:27-33 */
    spxo_prf_addr(pid, sk, sk_seed, addr);
}

void spxo_gen_chain(int pid, uint8_t *out, const uint8_t *in, unsigned start, unsigned steps,
                    const uint8_t *pub_seed, uint32_t addr[8])
{
    const spxo_params *p = P(pid); /* wots.cu:51-62 */
    memmove(out, in, p->n);
    for (unsigned i = start; i < start + steps && i < (unsigned)p->w; i++) {
        set_hash(addr, i);
        spxo_thash(pid, out, out, 1, pub_seed, addr);
    }
}

static void base_w(const spxo_params *p, unsigned *output, int out_len, const uint8_t *input)
{
    int in = 0, bits = 0; /* wots.cu:84-101 */
    uint8_t total = 0;
    for (int o = 0; o < out_len; o++) {
        if (bits == 0) {
            total = input[in++];
            bits += 8;
        }
        bits -= p->logw;
        output[o] = (total >> bits) & (p->w - 1);
    }
}

void spxo_chain_lengths(int pid, unsigned *lengths, const uint8_t *msg)
{
    const spxo_params *p = P(pid); /* wots.cu:123-162 */
    base_w(p, lengths, p->len1, msg);
    unsigned csum = 0;
    for (int i = 0; i < p->len1; i++) csum += p->w - 1 - lengths[i];
    csum <<= (8 - ((p->len2 * p->logw) % 8)) % 8;
    uint8_t cb[4];
    int cbl = (p->len2 * p->logw + 7) / 8;
    ull_to_bytes(cb, cbl, csum);
    base_w(p, lengths + p->len1, p->len2, cb);
}

void spxo_wots_gen_pk(int pid, uint8_t *pk, const uint8_t *sk_seed, const uint8_t *pub_seed,
                      uint32_t addr[8])
{
    const spxo_params *p = P(pid); /* wots.cu:176-186 */
    for (int i = 0; i < p->len; i++) {
        set_chain(addr, i);
        wots_gen_sk(pid, pk + i * p->n, sk_seed, addr);
        spxo_gen_chain(pid, pk + i * p->n, pk + i * p->n, 0, p->w - 1, pub_seed, addr);
    }
}

void spxo_wots_sign(int pid, uint8_t *sig, const uint8_t *msg, const uint8_t *sk_seed,
                    const uint8_t *pub_seed, uint32_t addr[8])
{
    const spxo_params *p = P(pid); /* wots.cu:614-626 */
    unsigned lengths[80];
    spxo_chain_lengths(pid, lengths, msg);
    for (int i = 0; i < p->len; i++) {
        set_chain(addr, i);
        wots_gen_sk(pid, sig + i * p->n, sk_seed, addr);
        spxo_gen_chain(pid, sig + i * p->n, sig + i * p->n, 0, lengths[i], pub_seed, addr);
    }
}

void spxo_wots_pk_from_sig(int pid, uint8_t *pk, const uint8_t *sig, const uint8_t *msg,
                           const uint8_t *pub_seed, uint32_t addr[8])
{
    const spxo_params *p = P(pid); /* wots.cu:777-790 */
    unsigned lengths[80];
    spxo_chain_lengths(pid, lengths, msg);
    for (int i = 0; i < p->len; i++) {
        set_chain(addr, i);
        spxo_gen_chain(pid, pk + i * p->n, sig + i * p->n, lengths[i], p->w - 1 - lengths[i],
                       pub_seed, addr);
    }
}

/* wots_gen_leaf (sign.cu:87-104) */
void spxo_wots_gen_leaf(int pid, uint8_t *leaf, const uint8_t *sk_seed, const uint8_t *pub_seed,
                        uint32_t addr_idx, const uint32_t tree_addr[8])
{
    const spxo_params *p = P(pid);
    uint8_t pk[80 * 32];
    uint32_t wots_addr[8] = {0}, wots_pk_addr[8] = {0};
    set_type(wots_addr, T_WOTS);
    set_type(wots_pk_addr, T_WOTSPK);
    copy_subtree(wots_addr, tree_addr);
    set_keypair(p, wots_addr, addr_idx);
    spxo_wots_gen_pk(pid, pk, sk_seed, pub_seed, wots_addr);
    copy_keypair(p, wots_pk_addr, wots_addr);
    spxo_thash(pid, leaf, pk, p->len, pub_seed, wots_pk_addr);
}

/* ------------------------------------------------------------------ */
/* Merkle (utils.cu:94-241)                                            */
/* ------------------------------------------------------------------ */

typedef void (*gen_leaf_fn)(int pid, uint8_t *leaf, const uint8_t *sk_seed,
                            const uint8_t *pub_seed, uint32_t addr_idx, const uint32_t addr[8]);

void spxo_compute_root(int pid, uint8_t *root, const uint8_t *leaf, uint32_t leaf_idx,
                       uint32_t idx_offset, const uint8_t *auth_path, uint32_t tree_height,
                       const uint8_t *pub_seed, uint32_t addr[8])
{
    const spxo_params *p = P(pid); /* utils.cu:94-135 */
    int n = p->n;
    uint8_t buffer[64];
    if (leaf_idx & 1) {
        memcpy(buffer + n, leaf, n);
        memcpy(buffer, auth_path, n);
    } else {
        memcpy(buffer, leaf, n);
        memcpy(buffer + n, auth_path, n);
    }
    auth_path += n;
    for (uint32_t i = 0; i + 1 < tree_height; i++) {
        leaf_idx >>= 1;
        idx_offset >>= 1;
        set_height(addr, i + 1);
        set_index(addr, leaf_idx + idx_offset);
        if (leaf_idx & 1) {
            spxo_thash(pid, buffer + n, buffer, 2, pub_seed, addr);
            memcpy(buffer, auth_path, n);
        } else {
            spxo_thash(pid, buffer, buffer, 2, pub_seed, addr);
            memcpy(buffer + n, auth_path, n);
        }
        auth_path += n;
    }
    leaf_idx >>= 1;
    idx_offset >>= 1;
    set_height(addr, tree_height);
    set_index(addr, leaf_idx + idx_offset);
    spxo_thash(pid, root, buffer, 2, pub_seed, addr);
}

static void treehash(int pid, uint8_t *root, uint8_t *auth_path, const uint8_t *sk_seed,
                     const uint8_t *pub_seed, uint32_t leaf_idx, uint32_t idx_offset,
                     uint32_t tree_height, gen_leaf_fn gen_leaf, uint32_t tree_addr[8])
{
    const spxo_params *p = P(pid); /* utils.cu:189-241 */
    int n = p->n;
    uint8_t stack[(24 + 1) * 32];
    unsigned heights[24 + 1];
    unsigned offset = 0;
    for (uint32_t idx = 0; idx < (1u << tree_height); idx++) {
        gen_leaf(pid, stack + offset * n, sk_seed, pub_seed, idx + idx_offset, tree_addr);
        offset++;
        heights[offset - 1] = 0;
        if ((leaf_idx ^ 1u) == idx) memcpy(auth_path, stack + (offset - 1) * n, n);
        while (offset >= 2 && heights[offset - 1] == heights[offset - 2]) {
            uint32_t tree_idx = idx >> (heights[offset - 1] + 1);
            set_height(tree_addr, heights[offset - 1] + 1);
            set_index(tree_addr, tree_idx + (idx_offset >> (heights[offset - 1] + 1)));
            spxo_thash(pid, stack + (offset - 2) * n, stack + (offset - 2) * n, 2, pub_seed,
                       tree_addr);
            offset--;
            heights[offset - 1]++;
            if (((leaf_idx >> heights[offset - 1]) ^ 1u) == tree_idx)
                memcpy(auth_path + heights[offset - 1] * n, stack + (offset - 1) * n, n);
        }
    }
    memcpy(root, stack, n);
}

void spxo_xmss_treehash(int pid, uint8_t *root, uint8_t *auth_path, const uint8_t *sk_seed,
                        const uint8_t *pub_seed, uint32_t leaf_idx, uint32_t tree_addr[8])
{
    treehash(pid, root, auth_path, sk_seed, pub_seed, leaf_idx, 0, P(pid)->tree_h,
             spxo_wots_gen_leaf, tree_addr);
}

/* ------------------------------------------------------------------ */
/* FORS (fors.cu:24-145, 596-633)                                       */
/* ------------------------------------------------------------------ */

static void fors_gen_leaf(int pid, uint8_t *leaf, const uint8_t *sk_seed, const uint8_t *pub_seed,
                          uint32_t addr_idx, const uint32_t fors_tree_addr[8])
{
    const spxo_params *p = P(pid); /* fors.cu:44-56 */
    uint32_t fors_leaf_addr[8] = {0};
    copy_keypair(p, fors_leaf_addr, fors_tree_addr);
    set_type(fors_leaf_addr, T_FORSTREE);
    set_index(fors_leaf_addr, addr_idx);
    spxo_prf_addr(pid, leaf, sk_seed, fors_leaf_addr);
    spxo_thash(pid, leaf, leaf, 1, pub_seed, fors_leaf_addr);
}

void spxo_message_to_indices(int pid, uint32_t *indices, const uint8_t *m)
{
    const spxo_params *p = P(pid); /* fors.cu:77-91, LSB-first within bytes */
    unsigned offset = 0;
    for (int i = 0; i < p->k; i++) {
        indices[i] = 0;
        for (int j = 0; j < p->a; j++) {
            indices[i] ^= (uint32_t)((m[offset >> 3] >> (offset & 7)) & 1) << j;
            offset++;
        }
    }
}

void spxo_fors_sign(int pid, uint8_t *sig, uint8_t *pk, const uint8_t *m, const uint8_t *sk_seed,
                    const uint8_t *pub_seed, const uint32_t fors_addr[8])
{
    const spxo_params *p = P(pid); /* fors.cu:107-145 */
    uint32_t indices[64];
    uint8_t roots[64 * 32];
    uint32_t fors_tree_addr[8] = {0}, fors_pk_addr[8] = {0};
    copy_keypair(p, fors_tree_addr, fors_addr);
    copy_keypair(p, fors_pk_addr, fors_addr);
    set_type(fors_tree_addr, T_FORSTREE);
    set_type(fors_pk_addr, T_FORSPK);
    spxo_message_to_indices(pid, indices, m);
    for (int i = 0; i < p->k; i++) {
        uint32_t idx_offset = (uint32_t)i * (1u << p->a);
        set_height(fors_tree_addr, 0);
        set_index(fors_tree_addr, indices[i] + idx_offset);
        spxo_prf_addr(pid, sig, sk_seed, fors_tree_addr); /* fors_gen_sk */
        sig += p->n;
        treehash(pid, roots + i * p->n, sig, sk_seed, pub_seed, indices[i], idx_offset, p->a,
                 fors_gen_leaf, fors_tree_addr);
        sig += p->n * p->a;
    }
    spxo_thash(pid, pk, roots, p->k, pub_seed, fors_pk_addr);
}

void spxo_fors_pk_from_sig(int pid, uint8_t *pk, const uint8_t *sig, const uint8_t *m,
                           const uint8_t *pub_seed, const uint32_t fors_addr[8])
{
    const spxo_params *p = P(pid); /* fors.cu:596-633 */
    uint32_t indices[64];
    uint8_t roots[64 * 32], leaf[32];
    uint32_t fors_tree_addr[8] = {0}, fors_pk_addr[8] = {0};
    copy_keypair(p, fors_tree_addr, fors_addr);
    copy_keypair(p, fors_pk_addr, fors_addr);
    set_type(fors_tree_addr, T_FORSTREE);
    set_type(fors_pk_addr, T_FORSPK);
    spxo_message_to_indices(pid, indices, m);
    for (int i = 0; i < p->k; i++) {
        uint32_t idx_offset = (uint32_t)i * (1u << p->a);
        set_height(fors_tree_addr, 0);
        set_index(fors_tree_addr, indices[i] + idx_offset);
        spxo_thash(pid, leaf, sig, 1, pub_seed, fors_tree_addr); /* fors_sk_to_leaf */
        sig += p->n;
        spxo_compute_root(pid, roots + i * p->n, leaf, indices[i], idx_offset, sig, p->a,
                          pub_seed, fors_tree_addr);
        sig += p->n * p->a;
    }
    spxo_thash(pid, pk, roots, p->k, pub_seed, fors_pk_addr);
}

/* ------------------------------------------------------------------ */
/* api.h surface (sign.cu:124-352, 1924-1994, 3048-3121, 3302-3311,    */
/* 6058-6079)                                                           */
/* ------------------------------------------------------------------ */

/* The reference absorbs pub_seed once per operation (initialize_hash_function,
 * sign.cu:173,1942,3069).  One compression per operation is added to the
 * counter here so counts match the reference's work. */
static void count_midstate(const spxo_params *p) { if (p->hash == SPXO_SHA2) g_cnt_sha++; }

int spxo_seed_keypair(int pid, uint8_t *pk, uint8_t *sk, const uint8_t *seed)
{
    const spxo_params *p = P(pid); /* sign.cu:154-182 */
    uint8_t auth_path[16 * 32];
    uint32_t top_tree_addr[8] = {0};
    set_layer(top_tree_addr, p->d - 1);
    set_type(top_tree_addr, T_HASHTREE);
    memcpy(sk, seed, 3 * p->n);
    memcpy(pk, sk + 2 * p->n, p->n);
    count_midstate(p);
    treehash(pid, sk + 3 * p->n, auth_path, sk, sk + 2 * p->n, 0, 0, p->tree_h,
             spxo_wots_gen_leaf, top_tree_addr);
    memcpy(pk + p->n, sk + 3 * p->n, p->n);
    return 0;
}

int spxo_keypair(int pid, uint8_t *pk, uint8_t *sk)
{
    const spxo_params *p = P(pid); /* sign.cu:342-352 */
    uint8_t seed[96];
    spxo_randombytes(seed, 3 * p->n);
    if (g_debug_mode)
        for (int i = 0; i < 3 * p->n; i++) seed[i] = (uint8_t)i;
    return spxo_seed_keypair(pid, pk, sk, seed);
}

static int signature_core(int pid, uint8_t *sig, const uint8_t *m, size_t mlen, const uint8_t *sk,
                          const uint8_t *optrand, uint8_t *t_fors_pk, uint8_t *t_roots,
                          uint64_t *t_trees, uint32_t *t_leaves)
{
    const spxo_params *p = P(pid); /* sign.cu:1924-1994 */
    const uint8_t *sk_seed = sk, *sk_prf = sk + p->n, *pk = sk + 2 * p->n, *pub_seed = pk;
    uint8_t mhash[64], root[32];
    uint64_t tree;
    uint32_t idx_leaf;
    uint32_t wots_addr[8] = {0}, tree_addr[8] = {0};
    count_midstate(p);
    set_type(wots_addr, T_WOTS);
    set_type(tree_addr, T_HASHTREE);
    spxo_gen_message_random(pid, sig, sk_prf, optrand, m, mlen);
    spxo_hash_message(pid, mhash, &tree, &idx_leaf, sig, pk, m, mlen);
    sig += p->n;
    set_tree(wots_addr, tree);
    set_keypair(p, wots_addr, idx_leaf);
    spxo_fors_sign(pid, sig, root, mhash, sk_seed, pub_seed, wots_addr);
    if (t_fors_pk) memcpy(t_fors_pk, root, p->n);
    sig += p->fors_bytes;
    for (int i = 0; i < p->d; i++) {
        set_layer(tree_addr, i);
        set_tree(tree_addr, tree);
        copy_subtree(wots_addr, tree_addr);
        set_keypair(p, wots_addr, idx_leaf);
        if (t_trees) t_trees[i] = tree;
        if (t_leaves) t_leaves[i] = idx_leaf;
        spxo_wots_sign(pid, sig, root, sk_seed, pub_seed, wots_addr);
        sig += p->wots_bytes;
        treehash(pid, root, sig, sk_seed, pub_seed, idx_leaf, 0, p->tree_h, spxo_wots_gen_leaf,
                 tree_addr);
        if (t_roots) memcpy(t_roots + i * p->n, root, p->n);
        sig += p->tree_h * p->n;
        idx_leaf = (uint32_t)(tree & ((1u << p->tree_h) - 1));
        tree = tree >> p->tree_h;
    }
    return 0;
}

int spxo_signature_with_optrand(int pid, uint8_t *sig, const uint8_t *m, size_t mlen,
                                const uint8_t *sk, const uint8_t *optrand)
{
    return signature_core(pid, sig, m, mlen, sk, optrand, NULL, NULL, NULL, NULL);
}

int spxo_sign_trace(int pid, const uint8_t *m, size_t mlen, const uint8_t *sk,
                    const uint8_t *optrand, uint8_t *fors_pk, uint8_t *roots, uint64_t *trees,
                    uint32_t *leaves)
{
    const spxo_params *p = P(pid);
    uint8_t *sig = (uint8_t *)malloc(p->sig_bytes);
    int r = signature_core(pid, sig, m, mlen, sk, optrand, fors_pk, roots, trees, leaves);
    free(sig);
    return r;
}

int spxo_signature(int pid, uint8_t *sig, size_t *siglen, const uint8_t *m, size_t mlen,
                   const uint8_t *sk)
{
    const spxo_params *p = P(pid);
    uint8_t optrand[32];
    spxo_randombytes(optrand, p->n); /* sign.cu:1950 */
    if (g_debug_mode) memset(optrand, 0xad, p->n);
    signature_core(pid, sig, m, mlen, sk, optrand, NULL, NULL, NULL, NULL);
    *siglen = p->sig_bytes;
    return 0;
}

int spxo_verify(int pid, const uint8_t *sig, size_t siglen, const uint8_t *m, size_t mlen,
                const uint8_t *pk)
{
    const spxo_params *p = P(pid); /* sign.cu:3048-3121 */
    const uint8_t *pub_seed = pk, *pub_root = pk + p->n;
    uint8_t mhash[64], wots_pk[80 * 32], root[32], leaf[32];
    uint64_t tree;
    uint32_t idx_leaf;
    uint32_t wots_addr[8] = {0}, tree_addr[8] = {0}, wots_pk_addr[8] = {0};
    if (siglen != (size_t)p->sig_bytes) return -1;
    count_midstate(p);
    set_type(wots_addr, T_WOTS);
    set_type(tree_addr, T_HASHTREE);
    set_type(wots_pk_addr, T_WOTSPK);
    spxo_hash_message(pid, mhash, &tree, &idx_leaf, sig, pk, m, mlen);
    sig += p->n;
    set_tree(wots_addr, tree);
    set_keypair(p, wots_addr, idx_leaf);
    spxo_fors_pk_from_sig(pid, root, sig, mhash, pub_seed, wots_addr);
    sig += p->fors_bytes;
    for (int i = 0; i < p->d; i++) {
        set_layer(tree_addr, i);
        set_tree(tree_addr, tree);
        copy_subtree(wots_addr, tree_addr);
        set_keypair(p, wots_addr, idx_leaf);
        copy_keypair(p, wots_pk_addr, wots_addr);
        spxo_wots_pk_from_sig(pid, wots_pk, sig, root, pub_seed, wots_addr);
        sig += p->wots_bytes;
        spxo_thash(pid, leaf, wots_pk, p->len, pub_seed, wots_pk_addr);
        spxo_compute_root(pid, root, leaf, idx_leaf, 0, sig, p->tree_h, pub_seed, tree_addr);
        sig += p->tree_h * p->n;
        idx_leaf = (uint32_t)(tree & ((1u << p->tree_h) - 1));
        tree = tree >> p->tree_h;
    }
    return memcmp(root, pub_root, p->n) ? -1 : 0;
}

int spxo_sign(int pid, uint8_t *sm, uint64_t *smlen, const uint8_t *m, uint64_t mlen,
              const uint8_t *sk)
{
    const spxo_params *p = P(pid); /* sign.cu:3302-3311 */
    size_t siglen;
    /* the message may alias sm + sig_bytes: sign from a copy, then move it */
    uint8_t *mc = (uint8_t *)malloc(mlen ? mlen : 1);
    memcpy(mc, m, mlen);
    spxo_signature(pid, sm, &siglen, mc, mlen, sk);
    memmove(sm + p->sig_bytes, mc, mlen);
    *smlen = siglen + mlen;
    free(mc);
    return 0;
}

int spxo_open(int pid, uint8_t *m, uint64_t *mlen, const uint8_t *sm, uint64_t smlen,
              const uint8_t *pk)
{
    const spxo_params *p = P(pid); /* sign.cu:6058-6079 */
    if (smlen < (uint64_t)p->sig_bytes) {
        memset(m, 0, smlen);
        *mlen = 0;
        return -1;
    }
    *mlen = smlen - p->sig_bytes;
    if (spxo_verify(pid, sm, p->sig_bytes, sm + p->sig_bytes, *mlen, pk)) {
        memset(m, 0, smlen);
        *mlen = 0;
        return -1;
    }
    memmove(m, sm + p->sig_bytes, *mlen);
    return 0;
}
