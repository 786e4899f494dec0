// sha2-128f / sha2-192f / sha2-256f mixed batches in one persistent signing
// kernel; J (FORS leaves per lane): 2^3 for 128f (8 lanes per tree, timed
// best for a mixed 128f/192f/256f batch of 8190), else as in each set's own
// fused launch of large batches
#include "family_impl.hpp"

using F128 = spx::Params<spx::kSha2, 16, 66, 22, 6, 33>;
using F192 = spx::Params<spx::kSha2, 24, 66, 22, 8, 33>;
using F256 = spx::Params<spx::kSha2, 32, 68, 17, 9, 35>;
SPX_REGISTER_FAMILY(cuspx_family_sha2_f, F128, 3, 1, F192, 6, 3, F256, 6, 5)
