// dispatch_detail.hpp -- dispatch-table calls on a device basis handle,
// shared by Engine and BatchEngine.
#pragma once

#include <libaccint/engine.hpp>

namespace libaccint::detail {

DispatchTable autotune_handle(lai_basis* b, int flags, const Real* d_D, hipStream_t s, int reps, const char* what);
DispatchTable dispatch_table_of(lai_basis* b, const char* what);
void set_dispatch_table_of(lai_basis* b, const DispatchTable& table, const char* what);
// applies the file named by LIBACCINT_DISPATCH_TABLE, if any
void apply_dispatch_table_env(lai_basis* b);

}  // namespace libaccint::detail
