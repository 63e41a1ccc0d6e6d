// dw_mcmc_steps_noalloc: the Metropolis-Hastings driver of dw_mcmc_steps with every device buffer of the
// call carved from one fixed scratch array, so that it calls no hipMalloc / hipFree (for callers under
// HIP graph capture or in loops that must not allocate).
//
// `scratch` is a file-scope device array; every GPU the library runs on holds its own copy, at its own
// address, until the process ends.  On the host its name is only a stand-in for the symbol, so the
// device address is looked up once per device with hipGetSymbolAddress (with that device current) and
// kept; only that address is handed to kernels.  All calls on one GPU share the array: calls that
// overlap in time on one GPU (two streams, two threads) race on it.  Every buffer a call reads is written
// earlier in the same call (the acceptance counters are cleared by a fill kernel), so nothing carries
// over between calls.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "api_internal.h"
#include "dwmi.h"

namespace {

static __device__ __align__(256) unsigned char scratch[DW_NOALLOC_SCRATCH_BYTES];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
char* g_scratch_ptr[kMaxDevices] = {};  // device address of `scratch` per device (nullptr: not looked up yet)

dw_status scratch_address(char** out) {
  int dev = -1;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return dwmi::set_error(DW_ERR_HIP, std::string("hipGetDevice failed: ") + hipGetErrorString(e));
  if (dev < 0 || dev >= kMaxDevices) return dwmi::set_error(DW_ERR_INVALID, "device ordinal out of range");
  std::lock_guard<std::mutex> lock(g_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return dwmi::set_error(DW_ERR_HIP, std::string("hipGetSymbolAddress(scratch) failed: ") + hipGetErrorString(e));
    g_scratch_ptr[dev] = (char*)p;
  }
  *out = g_scratch_ptr[dev];
  return DW_OK;
}

}  // namespace

extern "C" dw_status dw_mcmc_steps_noalloc(dw_model* m, float* r, int32_t n_walkers, const float* R, const float* Z,
                                           float* log_psi_sqr, int32_t* walker_age, const float* normals,
                                           const float* uniforms, int32_t n_steps, dw_chain_scalars* chain,
                                           int32_t max_age, int32_t stepsize_update_interval, float target_acceptance,
                                           float min_stepsize, float max_stepsize, float* out_acc, uint32_t* out_flags,
                                           dw_stream stream) {
  const dwmi::McmcArgs a{r, n_walkers, R, Z, log_psi_sqr, walker_age, normals, uniforms, n_steps, chain, max_age,
                         stepsize_update_interval, target_acceptance, min_stepsize, max_stepsize, out_acc, out_flags};
  // size and state checks need no device: a bad call fails before the symbol lookup
  const dwmi::FixedScratch check{nullptr, DW_NOALLOC_SCRATCH_BYTES, true};
  dw_status st = dwmi::mcmc_run(m, a, &check, nullptr);
  if (st) return st;
  char* base = nullptr;
  st = scratch_address(&base);
  if (st) return st;
  const dwmi::FixedScratch fx{base, DW_NOALLOC_SCRATCH_BYTES, false};
  return dwmi::mcmc_run(m, a, &fx, (hipStream_t)stream);
}
