// sha2-192f: n=24 h=66 d=22 a=8 k=33 (params.h:33-79); FORS: 2^2 leaves per
// lane, 2^6 in the fused launch of large batches (4 lanes per tree)
#include "set_impl.hpp"

using SetParams = spx::Params<spx::kSha2, 24, 66, 22, 8, 33>;
SPX_REGISTER_SET(cuspx_set_sha2_192f, "sha2-192f", 3, SetParams, 2, 6)
