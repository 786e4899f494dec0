// sha2-128f: n=16 h=66 d=22 a=6 k=33 (params.h:33-79); FORS: 2^2 leaves per
// lane, 2^4 in the fused launch of large batches (4 lanes per tree)
#include "set_impl.hpp"

using SetParams = spx::Params<spx::kSha2, 16, 66, 22, 6, 33>;
SPX_REGISTER_SET(cuspx_set_sha2_128f, "sha2-128f", 1, SetParams, 2, 4)
