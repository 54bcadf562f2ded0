// Numeric multifrontal Cholesky factorisation on the device.
//
// The symbolic phase (ordering, supernodes, row structure, front plan) runs
// once on the host (cholesky_symbolic); the numeric phase runs here, level by
// level of the nested-dissection tree:
//   1. scatter the S blocks owned by the level's supernodes into their fronts;
//   2. extend-add each child's updated Schur complement into its parent's
//      front (two passes: first children, then second children);
//   3. per group of fronts with equal (ns, ms): batched dense Cholesky of the
//      leading block, triangular solve of the panel, Schur update of the
//      trailing block, inverse of the leading factor and W21 = L21 Linv
//      (rocSOLVER / rocBLAS batched calls on the context stream);
// then one pass writes the sweep layout of pack_factor (fwd / bwd) from the
// fronts.  The result equals cholesky_factor + pack_factor up to rounding.
#pragma once
#include <hip/hip_runtime.h>

#include "cholesky.hpp"

namespace pecs {

struct DeviceFactorStats {
  double seconds = 0.0;  // wall time of the numeric phase (synchronised)
  double flops = 0.0;    // dense flops of the batched calls
  double front_bytes = 0.0;
};

// One dense-library (rocBLAS / rocSOLVER) handle, created once per context
// and reused by all of its factorisations.
void* factor_handle_create();
void factor_handle_destroy(void* handle);

// fwd / bwd: device arrays of P.nvals doubles (P from pack_layout of F).
// handle: from factor_handle_create.  The fronts (plan.front_total doubles)
// are allocated for the call and freed before it returns; the call checks
// first that the device has room for them.
DeviceFactorStats factor_on_device(const FrontPlan& plan, const SupernodalFactor& F, const PackedFactor& P,
                                   double* fwd, double* bwd, hipStream_t stream, void* handle);

// The MFMA-blocked copies of pack_blocked, built on the device from fwd / bwd
// (P must hold blocked_layout's offsets; fblk / bblk have fblk_off.back() /
// bblk_off.back() doubles).
void blocked_on_device(const PackedFactor& P, const double* fwd, const double* bwd, double* fblk, double* bblk,
                       hipStream_t stream);

}  // namespace pecs
