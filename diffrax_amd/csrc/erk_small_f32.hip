// erk_small_f32.hip -- float32-state / float64-time build of the small-system kernel.
// Stage algebra runs in float32; time, step size, the error norm and the controller stay
// in float64.  Built with -ffp-contract=on: a * b + c is fused only where one source
// expression says so, the same in every kernel the step loop is inlined into (the forward
// solve and the discrete adjoint's replay of it).
#include "erk_adjoint.hpp"
#include "erk_small_dispatch.hpp"

namespace dfx {
int erk_small_solve_f32(int solver, int vf, int d, const ErkKArgs& a, int block, int64_t lanes,
                        hipStream_t s) {
  return dispatch_small<float>(solver, vf, d, a, block, lanes, s);
}
int erk_backsolve_f32(int solver, int vf, int d, const ErkKArgs& a, const BacksolveArgs& b,
                      hipStream_t s) {
  return dispatch_backsolve<float>(solver, vf, d, a, b, s);
}
int erk_adjoint_f32(int solver, int vf, int d, const ErkKArgs& a, const ErkKArgs& ra,
                    const DiscreteAdjArgs& b, hipStream_t s) {
  return dispatch_adjoint<float>(solver, vf, d, a, ra, b, s);
}
}  // namespace dfx
