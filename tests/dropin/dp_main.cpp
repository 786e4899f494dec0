// A caller written the way the reference's test/dp_spx.cu and
// test/dp_benchmark.cu drive the batched surface (face_mdp_* and the
// multi-stream face_ms_mdp_*), compiled against include/cuspx_api.h alone:
//
//   g++ -DSHA256 -DSPX_128F -I include dp_main.cpp -L cuspx_amd -lcuspx
//   ./dp_main NUM [OUT]
//
// or, with -DPLAIN_API_H, against the plain declarations and size macros
// only, linked to the per-set library alone (as spx_main.cpp):
//
//   g++ -DSHA256 -DSPX_128F -DPLAIN_API_H -I include dp_main.cpp -L cuspx_amd -lspx_sha2_128f
//
// The DRBG is seeded with the NIST KAT entropy 00..2f and draws the NUM
// 32-byte messages; keys and optrand follow the shipped DEBUG_MODE.  OUT (if
// given) receives m || pk || sk || sm of the face_mdp_* calls for the test to
// compare with the CPU oracle.  Prints "OK" and exits 0 if every check passed.
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <vector>

#ifdef PLAIN_API_H
#include "cuspx.h"
#include "cuspx_params.h"
#else
#include "cuspx_api.h"
#endif

static int failures = 0;
static void expect(bool ok, const char* what) {
    if (!ok) {
        printf("  X %s\n", what);
        failures++;
    }
}

int main(int argc, char** argv) {
    setbuf(stdout, NULL);
    const unsigned num = argc > 1 ? (unsigned)atoi(argv[1]) : 64;
    const char* out = argc > 2 ? argv[2] : NULL;
    unsigned char entropy[48];
    for (int i = 0; i < 48; i++) entropy[i] = (unsigned char)i;

    typedef std::vector<unsigned char> Buf;
    Buf pk((size_t)SPX_PK_BYTES * num), sk((size_t)SPX_SK_BYTES * num);
    Buf m((size_t)SPX_MLEN * num), sm((size_t)SM_BYTES * num), mout((size_t)SPX_MLEN * num);
    unsigned long long smlen = 0, mlen = 0;

    randombytes_init(entropy, NULL, 256);
    randombytes(m.data(), (unsigned long long)SPX_MLEN * num);
    expect(face_mdp_crypto_sign_keypair(pk.data(), sk.data(), num) == 0, "batched keygen");
    expect(face_mdp_crypto_sign(sm.data(), &smlen, m.data(), SPX_MLEN, sk.data(), num) == 0,
           "batched signing");
    expect(smlen == SPX_SM_BYTES, "smlen != SPX_SM_BYTES");

    // every record opens on its own with the single-item call (the first 16)
    for (unsigned i = 0; i < num && i < 16; i++) {
        Buf one(SM_BYTES);
        expect(crypto_sign_open(one.data(), &mlen, sm.data() + (size_t)i * SM_BYTES, smlen,
                                pk.data() + (size_t)i * SPX_PK_BYTES) == 0,
               "single-item verification of a batched signature");
        expect(mlen == SPX_MLEN && memcmp(one.data(), m.data() + (size_t)i * SPX_MLEN,
                                          SPX_MLEN) == 0,
               "single-item recovered message");
    }

    // batched verification recovers every message
    expect(face_mdp_crypto_sign_open(mout.data(), &mlen, sm.data(), smlen, pk.data(), num) == 0,
           "batched verification");
    expect(mlen == SPX_MLEN && mout == m, "batched recovered messages");

    // the multi-stream forms draw the same DRBG bytes and give the same bytes
    {
        Buf pk2(pk.size()), sk2(sk.size()), m2(m.size()), sm2(sm.size()), mo2(mout.size());
        unsigned long long smlen2 = 0, mlen2 = 0;
        randombytes_init(entropy, NULL, 256);
        randombytes(m2.data(), (unsigned long long)SPX_MLEN * num);
        expect(face_ms_mdp_crypto_sign_keypair(pk2.data(), sk2.data(), num) == 0,
               "multi-stream keygen");
        expect(face_ms_mdp_crypto_sign(sm2.data(), &smlen2, m2.data(), SPX_MLEN, sk2.data(),
                                       num) == 0,
               "multi-stream signing");
        expect(pk2 == pk && sk2 == sk && smlen2 == smlen && sm2 == sm,
               "multi-stream outputs differ from face_mdp_*");
        expect(face_ms_mdp_crypto_sign_open(mo2.data(), &mlen2, sm2.data(), smlen2, pk2.data(),
                                            num) == 0,
               "multi-stream verification");
        expect(mlen2 == SPX_MLEN && mo2 == m, "multi-stream recovered messages");
        // a flipped message bit of one record fails the call and clears that slot
        const unsigned bad = num / 2;
        sm2[(size_t)bad * SM_BYTES + smlen2 - 1] ^= 1;
        expect(face_ms_mdp_crypto_sign_open(mo2.data(), &mlen2, sm2.data(), smlen2, pk2.data(),
                                            num) != 0,
               "multi-stream verification accepted a flipped message bit");
        bool rest = true, cleared = true;
        for (unsigned i = 0; i < num; i++) {
            const bool same = memcmp(mo2.data() + (size_t)i * SPX_MLEN,
                                     m.data() + (size_t)i * SPX_MLEN, SPX_MLEN) == 0;
            if (i == bad) {
                for (int b = 0; b < SPX_MLEN; b++) cleared = cleared && mo2[(size_t)i * SPX_MLEN + b] == 0;
            } else {
                rest = rest && same;
            }
        }
        expect(rest && cleared, "multi-stream per-item results after one flipped bit");
    }

    // a flipped bit in the first record's signature fails the batch
    sm[100] ^= 1;
    expect(face_mdp_crypto_sign_open(mout.data(), &mlen, sm.data(), smlen, pk.data(), num) != 0,
           "batched verification accepted a flipped signature bit");
    sm[100] ^= 1;

    // the fixed 32-byte message stride: longer messages are refused (-4)
    expect(face_mdp_crypto_sign(sm.data(), &smlen, m.data(), SPX_MLEN + 1, sk.data(), num) == -4,
           "mlen > SPX_MLEN accepted");
    expect(face_mdp_crypto_sign_open(mout.data(), &mlen, sm.data(), SPX_SM_BYTES + 1, pk.data(),
                                     num) == -4,
           "smlen > SPX_SM_BYTES accepted");

    if (out) {
        FILE* f = fopen(out, "wb");
        expect(f != NULL, "open the output file");
        if (f) {
            fwrite(m.data(), 1, m.size(), f);
            fwrite(pk.data(), 1, pk.size(), f);
            fwrite(sk.data(), 1, sk.size(), f);
            fwrite(sm.data(), 1, sm.size(), f);
            fclose(f);
        }
    }
    if (failures) {
        printf("FAILED %d\n", failures);
        return 1;
    }
    printf("OK %s %u\n", CUSPX_PARAM_NAME, num);
    return 0;
}
