/*
 * cf_oracle_adjoint.c -- reverse mode of the oracle's step (placeholder
 * until the adjoint is restated).  TEST INFRASTRUCTURE ONLY.
 */
#include "cf_oracle_internal.h"

int cfo_adjoint_run(cfo_state* s, int nsteps, int ckpt, const double* dq_T,
                    double* du0, double* dv0, double* dw0, double dvel[3],
                    double* dq0)
{
    return CFO_EINVAL;
}
