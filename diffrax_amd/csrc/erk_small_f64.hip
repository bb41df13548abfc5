// erk_small_f64.hip -- float64 parity build of the small-system kernel.  Compiled with
// -ffp-contract=off: the compiler fuses nothing, so the only fused multiply-adds are the
// explicit ones (the stage, b, b_err and interpolant sums, written as fma chains in index
// order and restated the same way by the oracle); every other a*b+c is the rounded
// product then the rounded sum.
#include "erk_adjoint.hpp"
#include "erk_small_dispatch.hpp"

namespace dfx {
int erk_small_solve_f64(int solver, int vf, int d, const ErkKArgs& a, int block, int64_t lanes,
                        hipStream_t s) {
  return dispatch_small<double>(solver, vf, d, a, block, lanes, s);
}
int erk_backsolve_f64(int solver, int vf, int d, const ErkKArgs& a, const BacksolveArgs& b,
                      hipStream_t s) {
  return dispatch_backsolve<double>(solver, vf, d, a, b, s);
}
int erk_adjoint_f64(int solver, int vf, int d, const ErkKArgs& a, const ErkKArgs& ra,
                    const DiscreteAdjArgs& b, hipStream_t s) {
  return dispatch_adjoint<double>(solver, vf, d, a, ra, b, s);
}
}  // namespace dfx
