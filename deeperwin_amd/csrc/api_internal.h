// Internal interface between the C-ABI translation units of the dwmi library (api.cpp and noalloc.hip).
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>
#include <string>

#include "dwmi.h"

namespace dwmi {

// Arguments of dw_mcmc_steps (include/dwmi.h), shared by its allocating and non-allocating variants.
struct McmcArgs {
  float* r;
  int32_t n_walkers;
  const float* R;
  const float* Z;
  float* log_psi_sqr;
  int32_t* walker_age;
  const float* normals;
  const float* uniforms;
  int32_t n_steps;
  dw_chain_scalars* chain;
  int32_t max_age;
  int32_t stepsize_update_interval;
  float target_acceptance;
  float min_stepsize;
  float max_stepsize;
  float* out_acc;
  uint32_t* out_flags;
};

// A fixed device region that a call must carve all of its buffers from (no allocation).  With check_only
// the call validates its arguments, that they fit `bytes` and that the weights are on the device, and
// returns without touching the device (base is not used).
struct FixedScratch {
  char* base;
  size_t bytes;
  bool check_only;
};

// Run the chain.  fx == nullptr: the model's own buffers (grown with hipMalloc on demand, or carved from
// its caller-owned workspace).  Otherwise every device buffer of the call is carved from fx and nothing is
// allocated: the call fails with DW_ERR_INVALID when the walkers do not fit and with DW_ERR_STATE when the
// weights have not been uploaded yet.
dw_status mcmc_run(dw_model* m, const McmcArgs& a, const FixedScratch* fx, hipStream_t s);

// Set the thread-local message returned by dw_last_error and return `code`.
dw_status set_error(dw_status code, const std::string& msg);

}  // namespace dwmi
