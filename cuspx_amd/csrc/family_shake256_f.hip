// shake256-128f / shake256-192f / shake256-256f mixed batches in one
// persistent signing kernel; J (FORS leaves per lane): 2^3 for 128f (8 lanes
// per tree, timed best for the mixed 128f/192f batch of 8192), else as in
// each set's own fused launch of large batches
#include "family_impl.hpp"

using F128 = spx::Params<spx::kShake, 16, 66, 22, 6, 33>;
using F192 = spx::Params<spx::kShake, 24, 66, 22, 8, 33>;
using F256 = spx::Params<spx::kShake, 32, 68, 17, 9, 35>;
SPX_REGISTER_FAMILY(cuspx_family_shake256_f, F128, 3, 7, F192, 5, 9, F256, 5, 11)
