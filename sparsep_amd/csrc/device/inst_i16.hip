// Kernel instantiations for one value type (split per type so the build
// compiles them in parallel).
#include "launch.hpp"

template int run_work_t<int16_t>(const Work &, sp_ctx *, const void *, void *, const uintptr_t *, const uintptr_t *,
                            int, int, int);
template int run_tables_t<int16_t>(const Work &, sp_ctx *, const uintptr_t *, const uintptr_t *, int);
