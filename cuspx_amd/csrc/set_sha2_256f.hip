// sha2-256f: n=32 h=68 d=17 a=9 k=35 (params.h:33-79); FORS: 2^3 leaves per
// lane, 2^6 in the fused launch of large batches (8 lanes per tree)
#include "set_impl.hpp"

using SetParams = spx::Params<spx::kSha2, 32, 68, 17, 9, 35>;
SPX_REGISTER_SET(cuspx_set_sha2_256f, "sha2-256f", 5, SetParams, 3, 6)
