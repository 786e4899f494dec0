// shake256-256f: n=32 h=68 d=17 a=9 k=35 (params.h:33-79); FORS: 2^3 leaves
// per lane, 2^5 in the fused launch of large batches (16 lanes per tree)
#include "set_impl.hpp"

using SetParams = spx::Params<spx::kShake, 32, 68, 17, 9, 35>;
SPX_REGISTER_SET(cuspx_set_shake256_256f, "shake256-256f", 11, SetParams, 3, 5)
