// shake256-128f: n=16 h=66 d=22 a=6 k=33 (params.h:33-79); FORS: 2^2 leaves per lane
#include "set_impl.hpp"

using SetParams = spx::Params<spx::kShake, 16, 66, 22, 6, 33>;
SPX_REGISTER_SET(cuspx_set_shake256_128f, "shake256-128f", 7, SetParams, 2, 2)
