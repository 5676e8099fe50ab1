// Host-side interface of the adjoint kernels (lbm_adjoint.hip).
#pragma once

#include "lbm_common.hpp"

namespace plbm {

int adjoint_partials(const KArgs& a);
int launch_adjoint_step(int lattice, const KArgs& a, int coll, int storage, bool field,
                        const void* f, const float* lin, float* lout, float* dF, double* part,
                        hipStream_t s);
int launch_adjoint_seed(int lattice, const KArgs& a, int storage, bool field, const void* f,
                        const float* G, float* lam, float* dF, hipStream_t s);
int launch_adjoint_init(int lattice, int N, const double* rho0, const double* u0,
                        const float* lam, double* du, double* drho, hipStream_t s);

}  // namespace plbm
