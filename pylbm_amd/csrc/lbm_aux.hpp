// Host-side interface of the auxiliary kernels (encode / decode / init / observables / copy).
#pragma once

#include "lbm_common.hpp"

namespace plbm {

enum AuxOp { kOpDecode = 0, kOpEncode = 1, kOpInit = 2, kOpMacro = 3, kOpReduce = 4 };

struct AuxArgs {
  int lattice, storage, aa;
  int N;
  int has_force;
  float F[3];
  const float* ff;
  int kind;                 // reduction kind
  const void* f;            // stored state (input)
  const void* f2;           // second stored state (MAXABS_DIFF)
  void* f_out;              // stored state (output)
  const double* d_in;       // canonical state / rho
  const double* d_in2;      // u
  double* d_out;            // canonical state / rho / reduction result
  double* d_out2;           // u
  double* partials;         // reduction: one double per workgroup (reduce_blocks)
};

// A reduction runs on min(ceil(N / kBlock), kReducePartials) workgroups, one float64
// partial each; the allocation-free entry keeps them in a static device array of this
// size (32 KiB per GPU).
constexpr int kReducePartials = 4096;
constexpr int kMaxDevices = 64;

int launch_aux(int op, const AuxArgs& x, hipStream_t s);
int reduce_blocks(int64_t n_nodes);
// Device address of the current GPU's static partials array, looked up with
// hipGetSymbolAddress on the first call on each device; 0 or an error code.
int reduce_scratch(double** out);
int launch_copy(const void* src, void* dst, int64_t nbytes, hipStream_t s);
// kind 0 copy (non-temporal), 1 copy (ordinary loads / stores), 2 read only, 3 write only;
// unroll 1, 2, 4 or 8 (16-byte accesses per lane)
int launch_plane_probe(int planes, int elem_bytes, const void* src, void* dst,
                       int64_t plane_bytes, hipStream_t s);
int launch_probe(int kind, int unroll, const void* src, void* dst, int64_t nbytes,
                 hipStream_t s);

}  // namespace plbm
