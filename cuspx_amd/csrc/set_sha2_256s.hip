// sha2-256s: n=32 h=64 d=8 a=14 k=22 (params.h:33-79); FORS: 2^6 leaves per lane
#include "set_impl.hpp"

using SetParams = spx::Params<spx::kSha2, 32, 64, 8, 14, 22>;
SPX_REGISTER_SET(cuspx_set_sha2_256s, "sha2-256s", 4, SetParams, 6, 6)
