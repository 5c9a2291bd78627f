// The Poisson solve on a caller-supplied device workspace, and the set-up of
// the static scratch array that emc_poisson_solve_noalloc hands it
// (csrc/noalloc.hip).
#pragma once
#include <stddef.h>

#include "../../include/emc.h"

// Doubles of workspace one solve on the context's mesh needs (0 without a GPU).
size_t emc_poisson_ws_doubles(emc_ctx* c);
// The device ordinal of the context (-1 without a GPU).
int emc_ctx_device(emc_ctx* c);
// emc_poisson_solve with every vector the kernels write carved out of ws
// (device memory, ws_doubles long): allocates nothing.
int emc_poisson_solve_ws(emc_ctx* c, const double* rhs, double* sol, double rtol, int maxit, int* iters,
                         double* relres, double* ws, size_t ws_doubles);
// Looks up, once per device, the device address of the scratch array (with that
// device current); EMC_OK or EMC_ERR_HIP.
int emc_noalloc_setup(int device);
