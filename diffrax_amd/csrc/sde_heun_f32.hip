// sde_heun_f32.hip -- float instantiations of the Heun SDE kernels (FMA contraction allowed).
#include "sde_heun.hpp"

namespace dfx {

int sde_launch_fwd_f32(const SdeArgs& a, hipStream_t s) {
  const int64_t nb = (a.P * (a.d / 4) + kSdeBlock - 1) / kSdeBlock;
  if (a.n_ts > 0 || a.save_steps > 0 || a.dense_ys)
    hipLaunchKernelGGL((sde_heun_fwd_kernel<float, 4, 16, true>), dim3((unsigned)nb), dim3(kSdeBlock), 0, s, a);
  else
    hipLaunchKernelGGL((sde_heun_fwd_kernel<float, 4, 16, false>), dim3((unsigned)nb), dim3(kSdeBlock), 0, s, a);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

int sde_launch_adj_f32(const SdeArgs& a, int64_t nb, size_t lds, hipStream_t s) {
  auto go = [&](auto kern) {
    hipLaunchKernelGGL(kern, dim3((unsigned)nb), dim3(kSdeBlock), lds, s, a);
  };
  const bool steps = a.save_steps > 0 || a.n_ts > 0;
  if (a.CK == 32)
    steps ? go(sde_heun_adj_kernel<float, 32, true>) : go(sde_heun_adj_kernel<float, 32, false>);
  else if (a.CK == 16)
    steps ? go(sde_heun_adj_kernel<float, 16, true>) : go(sde_heun_adj_kernel<float, 16, false>);
  else
    steps ? go(sde_heun_adj_kernel<float, 8, true>) : go(sde_heun_adj_kernel<float, 8, false>);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

int sde_launch_dense_f32(const SdeDenseArgs& a, hipStream_t s) {
  const int64_t nb = (a.P * a.n_q * a.d + 255) / 256;
  hipLaunchKernelGGL((sde_dense_eval_kernel<float>), dim3((unsigned)nb), dim3(256), 0, s, a);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

}  // namespace dfx
