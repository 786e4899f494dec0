/*
 * cuspx_params.h -- compile-time sizes of one SPHINCS+ parameter set, for
 * callers written against the reference's params.h / api.h: the SPX_* size
 * macros of params.h:33-151 and the CRYPTO_* macros of api.h:9-14, with the
 * same names and values.
 *
 * The set is chosen the way the reference's Makefile chooses it
 * (Makefile:109-155): -DSHA256 or -DSHAKE256, plus one of -DSPX_128S,
 * -DSPX_128F, -DSPX_192S, -DSPX_192F, -DSPX_256S, -DSPX_256F.
 * CUSPX_PARAM_NAME is the library's name of that set ("sha2-128f", ...),
 * for cuspx_param_id().  cuspx_api.h includes this header and also makes the
 * chosen set the library's default, so the api.h calls act on it.
 */
#ifndef CUSPX_PARAMS_H
#define CUSPX_PARAMS_H

#if defined(SHA256) && defined(SHAKE256)
#error define only one of SHA256 and SHAKE256
#elif defined(SHA256)
#define CUSPX_HASH_NAME "sha2"
#elif defined(SHAKE256)
#define CUSPX_HASH_NAME "shake256"
#else
#error define SHA256 or SHAKE256 to select the hash family
#endif

/* n, full height h, layers d, FORS height a, FORS trees k */
#if defined(SPX_128S)
#define SPX_N 16
#define SPX_FULL_HEIGHT 63
#define SPX_D 7
#define SPX_FORS_HEIGHT 12
#define SPX_FORS_TREES 14
#define CUSPX_SIZE_NAME "128s"
#elif defined(SPX_128F)
#define SPX_N 16
#define SPX_FULL_HEIGHT 66
#define SPX_D 22
#define SPX_FORS_HEIGHT 6
#define SPX_FORS_TREES 33
#define CUSPX_SIZE_NAME "128f"
#elif defined(SPX_192S)
#define SPX_N 24
#define SPX_FULL_HEIGHT 63
#define SPX_D 7
#define SPX_FORS_HEIGHT 14
#define SPX_FORS_TREES 17
#define CUSPX_SIZE_NAME "192s"
#elif defined(SPX_192F)
#define SPX_N 24
#define SPX_FULL_HEIGHT 66
#define SPX_D 22
#define SPX_FORS_HEIGHT 8
#define SPX_FORS_TREES 33
#define CUSPX_SIZE_NAME "192f"
#elif defined(SPX_256S)
#define SPX_N 32
#define SPX_FULL_HEIGHT 64
#define SPX_D 8
#define SPX_FORS_HEIGHT 14
#define SPX_FORS_TREES 22
#define CUSPX_SIZE_NAME "256s"
#elif defined(SPX_256F)
#define SPX_N 32
#define SPX_FULL_HEIGHT 68
#define SPX_D 17
#define SPX_FORS_HEIGHT 9
#define SPX_FORS_TREES 35
#define CUSPX_SIZE_NAME "256f"
#else
#error define one of SPX_128S SPX_128F SPX_192S SPX_192F SPX_256S SPX_256F
#endif

#define CUSPX_PARAM_NAME CUSPX_HASH_NAME "-" CUSPX_SIZE_NAME

/* Winternitz w = 16: 4 bits per chain digit; the message takes 2n digits and
 * the checksum (at most 2n * 15 < 16^3 for every n here) three more */
#define SPX_WOTS_W 16
#define SPX_WOTS_LOGW 4
#define SPX_ADDR_BYTES 32
#define SPX_WOTS_LEN1 (2 * SPX_N)
#define SPX_WOTS_LEN2 3
#define SPX_WOTS_LEN (SPX_WOTS_LEN1 + SPX_WOTS_LEN2)
#define SPX_WOTS_BYTES (SPX_WOTS_LEN * SPX_N)
#define SPX_WOTS_PK_BYTES SPX_WOTS_BYTES
#define SPX_TREE_HEIGHT (SPX_FULL_HEIGHT / SPX_D)

/* FORS: k * a index bits, each tree's record is its leaf secret and a
 * auth-path nodes */
#define SPX_FORS_MSG_BYTES ((SPX_FORS_HEIGHT * SPX_FORS_TREES + 7) / 8)
#define SPX_FORS_BYTES ((SPX_FORS_HEIGHT + 1) * SPX_FORS_TREES * SPX_N)
#define SPX_FORS_PK_BYTES SPX_N

/* signature = R || FORS || d x (WOTS signature || auth path) */
#define SPX_BYTES (SPX_N + SPX_FORS_BYTES + SPX_D * SPX_WOTS_BYTES + SPX_FULL_HEIGHT * SPX_N)
#define SPX_PK_BYTES (2 * SPX_N)
#define SPX_SK_BYTES (2 * SPX_N + SPX_PK_BYTES)
#define SPX_OPTRAND_BYTES 32

/* the batched calls' fixed message stride and signed-message record */
#define SPX_MLEN 32
#define SPX_SM_BYTES (SPX_MLEN + SPX_BYTES)
#define SM_BYTES (SPX_MLEN + SPX_BYTES)

#define CRYPTO_ALGNAME "SPHINCS+"
#define CRYPTO_SECRETKEYBYTES SPX_SK_BYTES
#define CRYPTO_PUBLICKEYBYTES SPX_PK_BYTES
#define CRYPTO_BYTES SPX_BYTES
#define CRYPTO_SEEDBYTES (3 * SPX_N)

#endif
