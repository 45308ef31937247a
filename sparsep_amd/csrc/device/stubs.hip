// Temporary: formats whose kernels are not built yet report SP_ERR_UNSUPPORTED.
#include "plan.hpp"

int coo_setup(sp_plan *, const sp_coo *) { return SP_ERR_UNSUPPORTED; }
int coo_run(sp_plan *, const void *, void *, int, int, int) { return SP_ERR_UNSUPPORTED; }
int bcsr_setup(sp_plan *, const sp_bcsr *) { return SP_ERR_UNSUPPORTED; }
int bcsr_run(sp_plan *, const void *, void *, int, int, int) { return SP_ERR_UNSUPPORTED; }
int bcoo_setup(sp_plan *, const sp_bcoo *) { return SP_ERR_UNSUPPORTED; }
int bcoo_run(sp_plan *, const void *, void *, int, int, int) { return SP_ERR_UNSUPPORTED; }
int tiled_plan_create(sp_ctx *, int, const sp_coo *, sp_plan *) { return SP_ERR_UNSUPPORTED; }
int tiled_run(sp_plan *, const void *, void *, int, int, int) { return SP_ERR_UNSUPPORTED; }
