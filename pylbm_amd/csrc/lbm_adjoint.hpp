// Host-side interface of the adjoint kernels (lbm_adjoint.hip).
#pragma once

#include "lbm_common.hpp"

namespace plbm {

int adjoint_partials(const KArgs& a);
// pattern: PYLBM_AB (lin -> lout), or PYLBM_AA_EVEN / PYLBM_AA_ODD in place (lin == lout);
// rec: f is the step's moment record (N x 4 fp32) instead of the stored state (part must be
// nullptr)
int launch_adjoint_step(int lattice, const KArgs& a, int coll, int storage, int lam_storage,
                        int pattern, bool field, bool rec, const void* f, const void* lin,
                        void* lout, float* dF, double* part, hipStream_t s);
// Two record-mode adjoint steps per launch (lam_in -> lam_out, two buffers): recA is the
// record of the later forward step, recB of the earlier one.  Applies to D3Q19 boxes
// periodic on every axis with rows of at most 128 nodes, fp32 adjoint states and no force
// field (adjoint_pair_applies); PYLBM_ENOTSUP otherwise.
bool adjoint_pair_applies(const KArgs& a, int lattice, int coll, int lam_storage, bool field);
int launch_adjoint_pair(const KArgs& a, int coll, const void* recA, const void* recB,
                        const void* lin, void* lout, hipStream_t s);
int launch_adjoint_seed(int lattice, const KArgs& a, int storage, int lam_storage, bool field,
                        const void* f, const float* G, void* lam, float* dF, hipStream_t s);
int launch_adjoint_init(int lattice, int lam_storage, int N, const double* rho0,
                        const double* u0, const void* lam, double* du, double* drho,
                        hipStream_t s);

}  // namespace plbm
