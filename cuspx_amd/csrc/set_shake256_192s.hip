// shake256-192s: n=24 h=63 d=7 a=14 k=17 (params.h:33-79); FORS: 2^6 leaves per lane
#include "set_impl.hpp"

using SetParams = spx::Params<spx::kShake, 24, 63, 7, 14, 17>;
SPX_REGISTER_SET(cuspx_set_shake256_192s, "shake256-192s", 8, SetParams, 6, 6)
