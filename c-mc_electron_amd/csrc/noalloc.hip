// emc_poisson_solve_noalloc: the Poisson solve for callers that cannot allocate
// device memory (code under HIP graph capture, serving loops).  Its work vectors
// live in one file-scope device array, `scratch`, of a fixed size; nothing is
// allocated or freed per call.
//
// In host code `scratch` names a host-side stand-in, not device memory: its
// device address is looked up with hipGetSymbolAddress once per device (each
// device holds its own copy of the array) and only that address is handed to
// the kernels and copies.
//
// Every call on one device shares that device's array and finds the previous
// call's contents in it: the solve initialises every word it reads (the staged
// right-hand side and start, the CG vectors, partial sums, the done flag and
// the multigrid vectors are all written before they are read).  Calls that
// overlap on one device race on the array: the caller serialises them.
#include <hip/hip_runtime.h>

#include <mutex>

#include "noalloc.h"

namespace {

// 80 MiB: about 10.6 doubles per mesh node, meshes up to about 980 k nodes.
constexpr size_t SCRATCH_DOUBLES = (80ull << 20) / sizeof(double);
constexpr int MAX_DEVICES = 64;

__device__ __align__(256) double scratch[SCRATCH_DOUBLES];

std::mutex scratch_mu;
double* scratch_ptr[MAX_DEVICES] = {};

}  // namespace

int emc_noalloc_setup(int device) {
  if (device < 0 || device >= MAX_DEVICES) return EMC_ERR_ARG;
  std::lock_guard<std::mutex> lk(scratch_mu);
  if (scratch_ptr[device]) return EMC_OK;
  int prev = -1;
  if (hipGetDevice(&prev) != hipSuccess) return EMC_ERR_HIP;
  if (prev != device && hipSetDevice(device) != hipSuccess) return EMC_ERR_HIP;
  void* p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (prev != device) (void)hipSetDevice(prev);
  if (e != hipSuccess || !p) return EMC_ERR_HIP;
  scratch_ptr[device] = (double*)p;
  return EMC_OK;
}

extern "C" int emc_poisson_solve_noalloc(emc_ctx* ctx, const double* rhs, double* sol, double rtol, int maxit,
                                         int* iters, double* relres) {
  if (!ctx || !rhs || !sol) return EMC_ERR_ARG;
  const int dev = emc_ctx_device(ctx);
  if (dev < 0) return EMC_ERR_NOGPU;
  double* ws = nullptr;
  {
    std::lock_guard<std::mutex> lk(scratch_mu);
    ws = dev < MAX_DEVICES ? scratch_ptr[dev] : nullptr;
  }
  if (!ws) return EMC_ERR_STATE;  // emc_create looks the address up; it did not succeed
  return emc_poisson_solve_ws(ctx, rhs, sol, rtol, maxit, iters, relres, ws, SCRATCH_DOUBLES);
}
