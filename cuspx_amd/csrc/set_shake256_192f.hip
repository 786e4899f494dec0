// shake256-192f: n=24 h=66 d=22 a=8 k=33 (params.h:33-79); FORS: 2^2 leaves
// per lane, 2^5 in the fused launch of large batches (8 lanes per tree)
#include "set_impl.hpp"

using SetParams = spx::Params<spx::kShake, 24, 66, 22, 8, 33>;
SPX_REGISTER_SET(cuspx_set_shake256_192f, "shake256-192f", 9, SetParams, 2, 5)
