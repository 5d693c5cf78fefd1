// api_internal.hpp -- what the C ABI translation units share besides
// lai_device.h: the thread's error message and a read-only view of a basis
// handle's shell data (both defined in api.hip).
#pragma once

#include <string>

#include "lai_device.h"

namespace lai {

// Records msg as the calling thread's last error and returns code.
int set_error(int code, const std::string& msg);

struct BasisView {
  lai_ctx* ctx;
  int ns, nbf, n_mol;
  const int* L;         // [ns]
  const int* nprim;     // [ns]
  const int* prim_off;  // [ns]
  const int* func_off;  // [ns], within the shell's molecule
  const double* exps;   // by prim_off
  const double* coefs;  // by prim_off, as uploaded (normalised)
};
BasisView basis_view(const lai_basis* b);

}  // namespace lai
