// erk_mlp_bf16.hip -- the neural-ODE solve kernels of the three-piece bfloat16 MFMA form at four and eight
// waves per tile (erk_mlp.hpp).
#include "erk_mlp.hpp"

namespace dfx {

void launch_erk_mlp_bf16(int waves, const MlpArgs& a, const MlpWeights<PolicyBF16x3>& W, bool grad,
                        bool integral, unsigned nblocks, hipStream_t s) {
  if (waves == 8) launch_erk_mlp<PolicyBF16x3, 8>(a, W, grad, integral, nblocks, s);
  else launch_erk_mlp<PolicyBF16x3, 4>(a, W, grad, integral, nblocks, s);
}

}  // namespace dfx
