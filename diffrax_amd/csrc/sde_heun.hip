// sde_heun.hip -- C ABI of the Heun SDE path (kernels in sde_heun.hpp; per-precision
// instantiations in sde_heun_f32.hip and sde_heun_f64.hip).
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string.h>

#include <mutex>

#include "../../include/diffrax_amd.h"
#include "sde_heun.hpp"

namespace dfx {

// Sum the per-block partials: one block per output q; thread t adds blocks t, t + 256, ...
// in order, then a fixed LDS tree combines the 256 sums (deterministic: the order depends on
// nothing but nblocks).  Writes theta/mu/sigma cotangents to their own buffers.
__global__ void __launch_bounds__(256) sde_reduce_kernel(const double* __restrict__ partial,
                                                         int nblocks, int d,
                                                         double* __restrict__ thb,
                                                         double* __restrict__ mub,
                                                         double* __restrict__ sgb) {
  __shared__ double sum[256];
  const int q = blockIdx.x;
  const int nout = 3 * d;
  double acc = 0.0;
  for (int b = threadIdx.x; b < nblocks; b += 256) acc += partial[(int64_t)b * nout + q];
  sum[threadIdx.x] = acc;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) sum[threadIdx.x] += sum[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    double* dst = (q < d) ? thb : (q < 2 * d ? mub : sgb);
    dst[q % d] = sum[0];
  }
}

}  // namespace dfx

using namespace dfx;

namespace {
constexpr int kBlock = kSdeBlock;

int sde_check(const dfx_sde_problem* p) {
  if (!p) return DFX_E_ARG;
  if (p->precision != DFX_F64 && p->precision != DFX_F32) return DFX_E_ARG;
  if (p->d <= 0 || p->d > 256 || p->paths < 0 || p->num_steps < 0) return DFX_E_ARG;
  const int V = (p->precision == DFX_F32) ? 4 : 2;
  if (p->d % V != 0) return DFX_E_ARG;
  if (p->save_steps < 0 || p->n_ts < 0) return DFX_E_ARG;
  if (kBlock % (p->d / V) != 0) return DFX_E_ARG;
  return DFX_OK;
}

int64_t sde_threads(const dfx_sde_problem* p, int V) { return p->paths * (p->d / V); }

int sde_adjoint_impl(const dfx_sde_problem* p, const dfx_sde_io* io, const dfx_sde_grad_io* g,
                     void* workspace, size_t workspace_bytes, void* stream);
}  // namespace

// Scratch of dfx_sde_heun_adjoint_static: the adjoint's per-block partial sums, one copy per
// GPU.  Host code uses the device address looked up per device (static_scratch), never the
// array's name.
static __device__ __align__(256) unsigned char scratch[DFX_SDE_STATIC_SCRATCH_BYTES];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

// The device address of `scratch` on the current device, looked up on first use.
int static_scratch(void** out) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) return DFX_E_HIP;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* ptr = nullptr;
    if (hipGetSymbolAddress((void**)&ptr, HIP_SYMBOL(scratch)) != hipSuccess || !ptr)
      return DFX_E_HIP;
    g_scratch_ptr[dev] = ptr;
  }
  *out = g_scratch_ptr[dev];
  return DFX_OK;
}
}  // namespace

extern "C" {

int dfx_sde_n_out(const dfx_sde_problem* p) {
  if (!p || p->save_steps < 0 || p->num_steps < 0) return DFX_E_ARG;
  if (p->n_ts < 0) return DFX_E_ARG;
  int n = (p->save_t0 ? 1 : 0) + p->n_ts;
  if (p->save_steps > 0) n += p->num_steps / p->save_steps;
  if (p->save_t1 && !(p->save_steps > 0 && p->num_steps % p->save_steps == 0)) n += 1;
  return n;
}

int dfx_sde_n_ckpt(const dfx_sde_problem* p) {
  if (!p || p->checkpoint_every <= 0) return 0;
  return (p->num_steps + p->checkpoint_every - 1) / p->checkpoint_every - 1;
}

int dfx_sde_workspace_bytes(const dfx_sde_problem* p, size_t* bytes) {
  if (sde_check(p) != DFX_OK || !bytes) return DFX_E_ARG;
  const int64_t nblocks = (sde_threads(p, 1) + kBlock - 1) / kBlock;
  *bytes = (size_t)(nblocks * 3 * p->d) * sizeof(double) + 256;
  return DFX_OK;
}

int dfx_sde_heun_solve(const dfx_sde_problem* p, const dfx_sde_io* io, void* stream) {
  int rc = sde_check(p);
  if (rc != DFX_OK) return rc;
  if (!io || !io->y0 || !io->theta || !io->mu || !io->sigma || !io->grid) return DFX_E_ARG;
  if (p->num_steps > 0 && !io->dW) return DFX_E_ARG;
  if (dfx_sde_n_out(p) > 0 && !io->ys) return DFX_E_ARG;
  if (p->n_ts > 0 && !io->ts) return DFX_E_ARG;
  if (io->ckpt && p->checkpoint_every <= 0) return DFX_E_ARG;
  if ((io->dense_ys == nullptr) != (io->dense_k == nullptr)) return DFX_E_ARG;
  if (io->dense_ys && (p->checkpoint_every > 0 || io->ckpt || p->num_steps < 1)) return DFX_E_ARG;
  if (p->paths == 0) return DFX_OK;
  SdeArgs a;
  memset(&a, 0, sizeof(a));
  a.y0 = io->y0; a.theta = io->theta; a.mu = io->mu; a.sigma = io->sigma;
  a.dW = io->dW; a.grid = io->grid; a.ys = io->ys; a.ckpt = io->ckpt;
  a.P = p->paths; a.d = p->d; a.N = p->num_steps; a.CK = p->checkpoint_every > 0 ? p->checkpoint_every : 1;
  a.n_out = dfx_sde_n_out(p);
  a.save_t0 = p->save_t0; a.save_t1 = p->save_t1; a.save_steps = p->save_steps;
  a.n_ts = p->n_ts; a.ts = io->ts;
  a.dense_ys = io->dense_ys; a.dense_k = io->dense_k;
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  if ((p->precision == DFX_F32 ? sde_launch_fwd_f32(a, s) : sde_launch_fwd_f64(a, s)) != 0)
    return DFX_E_HIP;
  return hipGetLastError() == hipSuccess ? DFX_OK : DFX_E_HIP;
}

int dfx_sde_dense_eval(const dfx_sde_problem* p, const dfx_sde_dense_io* io, void* stream) {
  int rc = sde_check(p);
  if (rc != DFX_OK) return rc;
  if (!io || !io->dense_ys || !io->dense_k || !io->grid || p->num_steps < 1 || io->n_q < 0)
    return DFX_E_ARG;
  if (io->tq_stride != 0 && io->tq_stride < io->n_q) return DFX_E_ARG;
  if (p->paths == 0 || io->n_q == 0) return DFX_OK;
  if (!io->tq || !io->out) return DFX_E_ARG;
  SdeDenseArgs a;
  a.dense_ys = io->dense_ys; a.dense_k = io->dense_k; a.grid = io->grid; a.tq = io->tq;
  a.out = io->out; a.P = p->paths; a.n_q = io->n_q; a.tq_stride = io->tq_stride;
  a.d = p->d; a.N = p->num_steps; a.left = io->left ? 1 : 0;
  a.derivative = io->derivative ? 1 : 0;
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  if ((p->precision == DFX_F32 ? sde_launch_dense_f32(a, s) : sde_launch_dense_f64(a, s)) != 0)
    return DFX_E_HIP;
  return hipGetLastError() == hipSuccess ? DFX_OK : DFX_E_HIP;
}

int dfx_sde_heun_adjoint(const dfx_sde_problem* p, const dfx_sde_io* io,
                         const dfx_sde_grad_io* g, void* workspace, size_t workspace_bytes,
                         void* stream) {
  return sde_adjoint_impl(p, io, g, workspace, workspace_bytes, stream);
}

int dfx_sde_heun_adjoint_static(const dfx_sde_problem* p, const dfx_sde_io* io,
                                const dfx_sde_grad_io* g, void* stream) {
  size_t need = 0;
  if (dfx_sde_workspace_bytes(p, &need) != DFX_OK) return DFX_E_ARG;
  if (need > DFX_SDE_STATIC_SCRATCH_BYTES) return DFX_E_TOO_LARGE;
  void* ptr = nullptr;
  const int rc = static_scratch(&ptr);
  if (rc != DFX_OK) return rc;
  return sde_adjoint_impl(p, io, g, ptr, DFX_SDE_STATIC_SCRATCH_BYTES, stream);
}

}  // extern "C"

namespace {
int sde_adjoint_impl(const dfx_sde_problem* p, const dfx_sde_io* io, const dfx_sde_grad_io* g,
                     void* workspace, size_t workspace_bytes, void* stream) {
  int rc = sde_check(p);
  if (rc != DFX_OK) return rc;
  if (!io || !g || !io->y0 || !io->theta || !io->mu || !io->sigma || !io->grid || !g->ys_bar)
    return DFX_E_ARG;
  if (!g->theta_bar || !g->mu_bar || !g->sigma_bar) return DFX_E_ARG;
  if (p->checkpoint_every != 32 && p->checkpoint_every != 16 && p->checkpoint_every != 8)
    return DFX_E_ARG;
  if (dfx_sde_n_ckpt(p) > 0 && !io->ckpt) return DFX_E_ARG;
  if (p->n_ts > 0 && !io->ts) return DFX_E_ARG;
  size_t need = 0;
  dfx_sde_workspace_bytes(p, &need);
  if (!workspace || workspace_bytes < need) return DFX_E_WORKSPACE;
  SdeArgs a;
  memset(&a, 0, sizeof(a));
  a.y0 = io->y0; a.theta = io->theta; a.mu = io->mu; a.sigma = io->sigma;
  a.dW = io->dW; a.grid = io->grid; a.ckpt = io->ckpt;
  a.ys_bar = g->ys_bar; a.y0_bar = g->y0_bar;
  a.P = p->paths; a.d = p->d; a.N = p->num_steps; a.CK = p->checkpoint_every;
  a.n_out = dfx_sde_n_out(p);
  a.save_t0 = p->save_t0; a.save_t1 = p->save_t1; a.save_steps = p->save_steps;
  a.n_ts = p->n_ts; a.ts = io->ts;
  const int64_t nb = (sde_threads(p, 1) + kBlock - 1) / kBlock;
  a.partial = reinterpret_cast<double*>(workspace);
  const size_t lds = (size_t)3 * kBlock * sizeof(double);
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  if (p->paths > 0) {
    if ((p->precision == DFX_F32 ? sde_launch_adj_f32(a, nb, lds, s)
                                 : sde_launch_adj_f64(a, nb, lds, s)) != 0)
      return DFX_E_HIP;
    const int nout = 3 * p->d;
    hipLaunchKernelGGL(sde_reduce_kernel, dim3(nout), dim3(256), 0, s, a.partial,
                       (int)nb, p->d, g->theta_bar, g->mu_bar, g->sigma_bar);
  }
  return hipGetLastError() == hipSuccess ? DFX_OK : DFX_E_HIP;
}
}  // namespace
