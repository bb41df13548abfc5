// sde_heun_f64.hip -- double instantiations of the Heun SDE kernels (no FMA contraction: bit parity with the FP64 oracle).
#include "sde_heun.hpp"

namespace dfx {

int sde_launch_fwd_f64(const SdeArgs& a, hipStream_t s) {
  const int64_t nb = (a.P * (a.d / 2) + kSdeBlock - 1) / kSdeBlock;
  if (a.n_ts > 0 || a.save_steps > 0 || a.dense_ys)
    hipLaunchKernelGGL((sde_heun_fwd_kernel<double, 2, 16, true>), dim3((unsigned)nb), dim3(kSdeBlock), 0, s, a);
  else
    hipLaunchKernelGGL((sde_heun_fwd_kernel<double, 2, 16, false>), dim3((unsigned)nb), dim3(kSdeBlock), 0, s, a);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

int sde_launch_adj_f64(const SdeArgs& a, int64_t nb, size_t lds, hipStream_t s) {
  auto go = [&](auto kern) {
    hipLaunchKernelGGL(kern, dim3((unsigned)nb), dim3(kSdeBlock), lds, s, a);
  };
  const bool steps = a.save_steps > 0 || a.n_ts > 0;
  if (a.CK == 32)
    steps ? go(sde_heun_adj_kernel<double, 32, true>) : go(sde_heun_adj_kernel<double, 32, false>);
  else if (a.CK == 16)
    steps ? go(sde_heun_adj_kernel<double, 16, true>) : go(sde_heun_adj_kernel<double, 16, false>);
  else
    steps ? go(sde_heun_adj_kernel<double, 8, true>) : go(sde_heun_adj_kernel<double, 8, false>);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

int sde_launch_dense_f64(const SdeDenseArgs& a, hipStream_t s) {
  const int64_t nb = (a.P * a.n_q * a.d + 255) / 256;
  hipLaunchKernelGGL((sde_dense_eval_kernel<double>), dim3((unsigned)nb), dim3(256), 0, s, a);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

}  // namespace dfx
