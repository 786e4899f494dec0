// sha2-128s: n=16 h=63 d=7 a=12 k=14 (params.h:33-79); FORS: 2^4 leaves per lane
#include "set_impl.hpp"

using SetParams = spx::Params<spx::kSha2, 16, 63, 7, 12, 14>;
SPX_REGISTER_SET(cuspx_set_sha2_128s, "sha2-128s", 0, SetParams, 4, 4)
