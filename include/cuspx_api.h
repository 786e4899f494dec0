/*
 * cuspx_api.h -- drop-in replacement for the reference's api.h + params.h +
 * rng.h includes.  A caller built against the reference switches by
 * including this header instead, keeping its -DSHA256 / -DSHAKE256 and
 * -DSPX_<size> flags, and linking libcuspx.so:
 *
 *   - the SPX_* and CRYPTO_* size macros (cuspx_params.h);
 *   - every api.h / rng.h function the library replaces (cuspx.h);
 *   - a load-time constructor in the including program that makes the
 *     compiled-in set the library's default, so crypto_sign_*,
 *     face_mdp_* and face_ms_mdp_* act on it as the reference's build does.
 */
#ifndef CUSPX_API_H
#define CUSPX_API_H

#include "cuspx.h"
#include "cuspx_params.h"

#ifdef __cplusplus
namespace {
struct CuspxSelectParams {
    CuspxSelectParams() { cuspx_set_default_param(cuspx_param_id(CUSPX_PARAM_NAME)); }
};
const CuspxSelectParams cuspx_select_params_;
}  // namespace
#else
__attribute__((constructor)) static void cuspx_select_params_(void) {
    cuspx_set_default_param(cuspx_param_id(CUSPX_PARAM_NAME));
}
#endif

#endif
