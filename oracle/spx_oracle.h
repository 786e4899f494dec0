/*
 * spx_oracle.h -- CPU oracle for the SPHINCS+ round-3 "simple" host path of
 * the reference (CUSPX).  TEST INFRASTRUCTURE ONLY: this library is the
 * checker for the HIP engine in cuspx_amd/.  Only tests/, the smoke entry and
 * bench.py's cpu_baseline leg may load it; the product path never does.
 *
 * It restates, from scratch and in plain C, the reference's host functions
 * (sign.cu, wots.cu, fors.cu, utils.cu, address.cu, sha256.cu, fips202.cu,
 * hash_*.cu, thash_*.cu, rng.cu), with the parameter set chosen at run time
 * instead of at compile time.  Every function cites the reference file:line
 * it follows in spx_oracle.c.
 *
 * Pinning status per function (how each is checked by tests/):
 *   sha256 / inc / mgf1 / hmac ......... hashlib / hmac closed forms, FIPS 180-4 vectors
 *   keccak / shake256 .................. hashlib.shake_256, FIPS 202 vectors
 *   aes256 / DRBG ...................... FIPS-197 C.3 vector, NIST PQC KAT count-0 seed/msg
 *   thash / prf / prf_msg / h_msg ...... hashlib closed forms (both families)
 *   address layout ..................... byte table from sha256_offsets.h
 *   chain_lengths / base_w ............. closed form + exhaustive small cases
 *   gen_chain .......................... composition identity
 *   wots ............................... sign/pk_from_sig round trip (test/wots.cu invariant)
 *   treehash / compute_root ............ brute-force recursive Merkle, every leaf index
 *   fors ............................... sign/pk_from_sig round trip (test/fors.cu invariant)
 *   sign / verify / keypair ............ round trip, bit-flip negative, determinism;
 *                                         end-to-end bytes vs the reference binary:
 *                                         parity unpinned (no reference vector exists)
 */
#ifndef SPX_ORACLE_H
#define SPX_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { SPXO_SHA2 = 0, SPXO_SHAKE = 1 };

typedef struct {
    int hash;          /* SPXO_SHA2 / SPXO_SHAKE */
    int n, full_h, d, a, k, w;
    int logw, len1, len2, len, tree_h;
    int wots_bytes, fors_msg_bytes, fors_bytes, sig_bytes, pk_bytes, sk_bytes;
    int tree_bits, tree_bytes, leaf_bits, leaf_bytes, dgst_bytes;
    char name[32];
} spxo_params;

/* parameter registry: ids 0..11 = {sha2,shake} x {128s,128f,192s,192f,256s,256f} */
int spxo_num_params(void);
int spxo_get_params(int pid, spxo_params *out);
int spxo_param_id(const char *name);

/* debug_mode 1 = the reference as shipped (DEBUG_MODE: seed[i]=i, optrand=0xad),
 * 0 = DRBG-derived seed and optrand. */
void spxo_set_debug_mode(int on);
int spxo_get_debug_mode(void);

/* counters of primitive invocations since the last reset */
void spxo_reset_counts(void);
uint64_t spxo_count_sha256_blocks(void);
uint64_t spxo_count_keccak_perms(void);

/* primitives */
void spxo_sha256(uint8_t out[32], const uint8_t *in, size_t inlen);
void spxo_sha256_compress_words(uint32_t state[8], const uint32_t w[16]);
void spxo_hmac_sha256(uint8_t out[32], const uint8_t *key, size_t keylen,
                      const uint8_t *msg, size_t msglen);
void spxo_mgf1(uint8_t *out, size_t outlen, const uint8_t *in, size_t inlen);
void spxo_keccakf1600(uint64_t s[25]);
void spxo_shake256(uint8_t *out, size_t outlen, const uint8_t *in, size_t inlen);
void spxo_aes256_ecb(const uint8_t key[32], const uint8_t in[16], uint8_t out[16]);

/* AES-256-CTR DRBG (NIST PQC KAT generator), one global instance like the reference */
void spxo_randombytes_init(const uint8_t entropy[48], const uint8_t *pers, int strength);
int spxo_randombytes(uint8_t *x, uint64_t xlen);

/* address helpers: addr is 8 uint32 words viewed as 32 bytes */
void spxo_addr_set(uint32_t addr[8], int field, uint64_t value, int pid);

/* tweakable hash layer */
void spxo_seed_state(int pid, uint8_t state[40], const uint8_t *pub_seed);
void spxo_thash(int pid, uint8_t *out, const uint8_t *in, unsigned inblocks,
                const uint8_t *pub_seed, const uint32_t addr[8]);
void spxo_prf_addr(int pid, uint8_t *out, const uint8_t *key, const uint32_t addr[8]);
void spxo_gen_message_random(int pid, uint8_t *R, const uint8_t *sk_prf,
                             const uint8_t *optrand, const uint8_t *m, uint64_t mlen);
void spxo_hash_message(int pid, uint8_t *digest, uint64_t *tree, uint32_t *leaf_idx,
                       const uint8_t *R, const uint8_t *pk, const uint8_t *m, uint64_t mlen);

/* WOTS / FORS / Merkle */
void spxo_chain_lengths(int pid, unsigned *lengths, const uint8_t *msg);
void spxo_gen_chain(int pid, uint8_t *out, const uint8_t *in, unsigned start, unsigned steps,
                    const uint8_t *pub_seed, uint32_t addr[8]);
void spxo_wots_gen_pk(int pid, uint8_t *pk, const uint8_t *sk_seed, const uint8_t *pub_seed,
                      uint32_t addr[8]);
void spxo_wots_sign(int pid, uint8_t *sig, const uint8_t *msg, const uint8_t *sk_seed,
                    const uint8_t *pub_seed, uint32_t addr[8]);
void spxo_wots_pk_from_sig(int pid, uint8_t *pk, const uint8_t *sig, const uint8_t *msg,
                           const uint8_t *pub_seed, uint32_t addr[8]);
void spxo_wots_gen_leaf(int pid, uint8_t *leaf, const uint8_t *sk_seed, const uint8_t *pub_seed,
                        uint32_t addr_idx, const uint32_t tree_addr[8]);
void spxo_xmss_treehash(int pid, uint8_t *root, uint8_t *auth_path, const uint8_t *sk_seed,
                        const uint8_t *pub_seed, uint32_t leaf_idx, uint32_t tree_addr[8]);
void spxo_compute_root(int pid, uint8_t *root, const uint8_t *leaf, uint32_t leaf_idx,
                       uint32_t idx_offset, const uint8_t *auth_path, uint32_t tree_height,
                       const uint8_t *pub_seed, uint32_t addr[8]);
void spxo_message_to_indices(int pid, uint32_t *indices, const uint8_t *m);
void spxo_fors_sign(int pid, uint8_t *sig, uint8_t *pk, const uint8_t *m, const uint8_t *sk_seed,
                    const uint8_t *pub_seed, const uint32_t fors_addr[8]);
void spxo_fors_pk_from_sig(int pid, uint8_t *pk, const uint8_t *sig, const uint8_t *m,
                           const uint8_t *pub_seed, const uint32_t fors_addr[8]);

/* api.h surface (parameter set passed explicitly) */
int spxo_seed_keypair(int pid, uint8_t *pk, uint8_t *sk, const uint8_t *seed);
int spxo_keypair(int pid, uint8_t *pk, uint8_t *sk);
int spxo_signature_with_optrand(int pid, uint8_t *sig, const uint8_t *m, size_t mlen,
                                const uint8_t *sk, const uint8_t *optrand);
int spxo_signature(int pid, uint8_t *sig, size_t *siglen, const uint8_t *m, size_t mlen,
                   const uint8_t *sk);
int spxo_verify(int pid, const uint8_t *sig, size_t siglen, const uint8_t *m, size_t mlen,
                const uint8_t *pk);
int spxo_sign(int pid, uint8_t *sm, uint64_t *smlen, const uint8_t *m, uint64_t mlen,
              const uint8_t *sk);
int spxo_open(int pid, uint8_t *m, uint64_t *mlen, const uint8_t *sm, uint64_t smlen,
              const uint8_t *pk);

/* hypertree intermediates of one signature (for kernel-level parity tests):
 * fors_pk[n], roots[d*n] (root of each layer's subtree), tree[d], leaf[d] */
int spxo_sign_trace(int pid, const uint8_t *m, size_t mlen, const uint8_t *sk,
                    const uint8_t *optrand, uint8_t *fors_pk, uint8_t *roots,
                    uint64_t *trees, uint32_t *leaves);

#ifdef __cplusplus
}
#endif
#endif
