// sde_heun_f32.hip -- float instantiations of the Heun SDE kernels (FMA contraction allowed).
#include "sde_heun.hpp"

namespace dfx {

int sde_launch_fwd_f32(const SdeArgs& a, hipStream_t s) {
  const int64_t nb = (a.P * (a.d / 4) + kSdeBlock - 1) / kSdeBlock;
  hipLaunchKernelGGL((sde_heun_fwd_kernel<float, 4, 16>), dim3((unsigned)nb), dim3(kSdeBlock), 0, s, a);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

int sde_launch_adj_f32(const SdeArgs& a, int64_t nb, size_t lds, hipStream_t s) {
  if (a.CK == 32)
    hipLaunchKernelGGL((sde_heun_adj_kernel<float, 32>), dim3((unsigned)nb), dim3(kSdeBlock), lds, s, a);
  else if (a.CK == 16)
    hipLaunchKernelGGL((sde_heun_adj_kernel<float, 16>), dim3((unsigned)nb), dim3(kSdeBlock), lds, s, a);
  else
    hipLaunchKernelGGL((sde_heun_adj_kernel<float, 8>), dim3((unsigned)nb), dim3(kSdeBlock), lds, s, a);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

}  // namespace dfx
