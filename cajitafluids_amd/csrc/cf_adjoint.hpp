// cf_adjoint.hpp -- launchers of the reverse-mode kernels (cf_adjoint.hip).
#pragma once

#include <hip/hip_runtime.h>

#include "cf_kernels.hpp"

namespace cfmi
{

// advect^T: bar (zeroed by the caller) += the transpose of k_advect applied
// to obar, linearised at cur (ghosts gathered).
void launch_adj_advect(const Geo& g, Fields cur, Fields obar, Fields bar,
                       hipStream_t s);
// inputs^T with pre = the values k_inputs saw; dvel[0..2] += dL/dU_d.
void launch_adj_inputs(const Geo& g, Fields pre, Fields bar, double* dvel,
                       hipStream_t s);
// projection^T pieces, in order: wall mask; y = -s G^T bar (owned cells) and
// ysum = sum(y); y -= diag(A) ysum / diag_sum; (PCG: z = L y);
// bar += -(1/dx) D^T z.
void launch_adj_mask(const Geo& g, Fields bar, hipStream_t s);
void launch_adj_gradT(const Geo& g, Fields bar, double* y, double* ysum,
                      RedBuf rb, hipStream_t s);
void launch_adj_shift(const Geo& g, double* y, const double* ysum,
                      double diag_sum, hipStream_t s);
void launch_adj_divT(const Geo& g, const double* z, Fields bar, hipStream_t s);

} // namespace cfmi
