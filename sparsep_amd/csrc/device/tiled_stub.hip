// 2D tiled plans are built in tiled.hip once the host tilings exist.
#include "plan.hpp"

int build_tiled_plan(sp_plan *, const sp_coo *) { return SP_ERR_UNSUPPORTED; }
