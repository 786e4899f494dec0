// shake256-256s: n=32 h=64 d=8 a=14 k=22 (params.h:33-79); FORS: 2^6 leaves per lane
#include "set_impl.hpp"

using SetParams = spx::Params<spx::kShake, 32, 64, 8, 14, 22>;
SPX_REGISTER_SET(cuspx_set_shake256_256s, "shake256-256s", 10, SetParams, 6, 6)
