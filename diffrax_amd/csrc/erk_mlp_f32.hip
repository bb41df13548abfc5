// erk_mlp_f32.hip -- the neural-ODE solve kernels of the float32 MFMA form at four and eight
// waves per tile (erk_mlp.hpp).
#include "erk_mlp.hpp"

namespace dfx {

void launch_erk_mlp_f32(int waves, const MlpArgs& a, const MlpWeights<PolicyF32>& W, bool grad,
                       bool integral, unsigned nblocks, hipStream_t s) {
  if (waves == 8) launch_erk_mlp<PolicyF32, 8>(a, W, grad, integral, nblocks, s);
  else launch_erk_mlp<PolicyF32, 4>(a, W, grad, integral, nblocks, s);
}

}  // namespace dfx
