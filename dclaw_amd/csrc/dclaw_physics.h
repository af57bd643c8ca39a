// dclaw_physics.h -- point and edge physics of the D-Claw step, FP64.
//
// Every routine is __host__ __device__ so the same code runs in the gfx950
// kernels and in the host-side lane emulator used by the tests.  The
// floating-point evaluation order follows the reference Fortran statement by
// statement (left-to-right products and sums, x**2 as x*x, unary minus
// outside the product); build with -ffp-contract=off so that no FMA
// contraction changes a rounding.  MAX/MIN are (a > b ? a : b) and
// (a < b ? a : b); SIGN(a, b) is copysign(|a|, b).
//
// Structurally zero f-wave components are not carried: with
// delta_huvm_revert = .true. (riemannsolvers_dclaw.f:66) the middle solver
// wave keeps only its h and normal-momentum parts (:513-516), so rpn2's
// waves 3 and 4 are identically zero (rpn2_dclaw.f:303-315) and wave 2 has
// only h and normal momentum (:296-301); no wave has a bdif component.
// Dropping +-0 terms from sums that start at +0 leaves every sum bit-exact.
#pragma once

#include <hip/hip_runtime.h>
#include <math.h>

#include "dclaw_hip.h"

#define DCM_FN __host__ __device__ __forceinline__
#include "dcmath.h"

#define DC_HD __host__ __device__ __forceinline__

namespace dclaw {

using Real = double;
using Params = dclaw_params;
// src2method 2 (mp_update_FE_4quad) is compiled into the step itself; the
// derivative instantiations leave it out (its substep logic over values
// with tangents would spill the shared src2 body of every method)
constexpr bool kFe4quad = true;

#include "dclaw_physics_body.inc"

}  // namespace dclaw
