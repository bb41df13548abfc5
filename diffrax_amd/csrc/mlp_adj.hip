// mlp_adj.hip -- discrete adjoint of the neural-ODE solve (erk_mlp.hip): gradients of a
// loss on the saved states with respect to y0 and all MLP weights and biases.
//
// Semantics (RecursiveCheckpointAdjoint, _adjoint.py:264-312): reverse mode through the
// accepted step sequence with the step sizes held fixed -- the controller's factor is
// stop_gradient (_step_size_controller/pid.py:498-530) and rejected steps leave y and the
// FSAL state unchanged (_integrate.py:422-426), so they carry no gradient.  The forward
// pass (gradient mode) recorded the accepted (t_n, t_{n+1}) and the state at every C-th
// accepted step.  The backward walks segments of C steps from the last one; for each
// segment and each tile of 32 trajectories this kernel
//   1. recomputes the segment forward from its checkpoint (identical arithmetic to the
//      forward kernel), keeping every stage's f(z_i) in scratch and its layer inputs
//      (z_i, h1, h2) in blocked row buffers;
//   2. runs the exact reverse sweep: cotangents of the SaveAt(ts) values enter through the
//      Tsit5 interpolant (k_i gets b_i(theta) ybar_s), the SSAL output y_{n+1} = z_6, and
//      each stage k_i = dt f(z_i) is pulled back through the MLP on MFMA
//      (tanh' = 1 - f^2 from the kept f, softplus' = 1 - exp(-h) from the kept h), with no
//      forward re-evaluation;
//   3. stages the deltas of every stage in blocked row buffers beside the layer inputs.
// A split-K MFMA GEMM then forms dW_l = sum_rows delta_l^T input_l for the segment, and a
// fixed-order reduction adds it to the float32 accumulators (deterministic result).
#include <hip/hip_runtime.h>
#include <math.h>
#include <string.h>

#include "../../include/diffrax_amd.h"
#include "erk_small.hpp"
#include "mlp_tile.hpp"

namespace dfx {

constexpr int ADJ_CHUNK = 4096;  // rows per split-K chunk of the weight-gradient GEMM

// Six row buffers: D1 (width 256), Z (64), D2 (256), H1 (256), D3 (64), H2 (256).

template <class P>
struct RowBufs {
  typename P::act_t* b[6];
};

struct AdjArgs {
  const double* grid;
  const float* ckpt;
  const int32_t* stats;
  const double* ts;
  const float* ys_bar;
  float* ybar;
  float* zs;
  int64_t B, B_pad, ts_stride;
  int32_t max_steps, n_ckpt, C, Cs, seg, n_ts, n_out, save_t0;
  double t0, t1;
};

// Copy 32 tile rows of an LDS buffer [32][ld] into a blocked global row buffer of width
// Wd: element (row, col) at ((row / KA) * Wd + col) * KA + row % KA.
template <class P>
__device__ __forceinline__ void rows_out(const typename P::act_t* lds, int ld, int Wd,
                                         typename P::act_t* dst, int64_t row0) {
  constexpr int KA = P::KA;
  const int pairs = (MLP_TILE / KA) * Wd;
  for (int e = threadIdx.x; e < pairs; e += blockDim.x) {
    const int rb = e / Wd, col = e - rb * Wd;
    typename P::act_t* o = dst + ((row0 / KA + rb) * Wd + col) * KA;
#pragma unroll
    for (int j = 0; j < KA; ++j) o[j] = lds[(rb * KA + j) * ld + col];
  }
}

// Element (row, col) of a blocked row buffer of width Wd.
template <class P>
__device__ __forceinline__ float row_val(const typename P::act_t* buf, int Wd, int64_t row,
                                         int col) {
  constexpr int KA = P::KA;
  return (float)buf[((row / KA) * Wd + col) * KA + row % KA];
}

template <class P, int NW>
__global__ void __launch_bounds__(64 * NW, 1)
mlp_adj_segment_kernel(const AdjArgs a, const MlpWeights<P> W, const RowBufs<P> R) {
  extern __shared__ __align__(16) unsigned char smem[];
  MlpLds<P>& L = *reinterpret_cast<MlpLds<P>*>(smem);
  constexpr bool F = P::fast_math;
  using Tab = Tsit5Tab;
  constexpr int S = Tab::S;
  using WV = MlpWaves<NW>;
  constexpr int CPT = WV::CPT, GRP = MLP_D / CPT, NT = WV::NT12, NT3 = WV::NT3;
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const int orow = tid / GRP, c0 = (tid % GRP) * CPT;
  const int tile = blockIdx.x;
  const int64_t traj = (int64_t)tile * MLP_TILE + orow;
  const bool valid = traj < a.B;
  const int nacc = valid ? a.stats[traj * 3 + 1] : 0;
  const int n0 = a.seg * a.C;
  int steps_here = nacc - n0;
  steps_here = steps_here < 0 ? 0 : (steps_here > a.Cs ? a.Cs : steps_here);
  const int dir = (a.t0 < a.t1) ? 1 : -1;
  const double* ts_row = a.ts ? a.ts + (valid ? traj : 0) * a.ts_stride : nullptr;
  using TL = TabLit<Tab, float>;  // float32 stage algebra: the coefficients as literals
  float* fs_tile = a.zs + (int64_t)tile * a.Cs * S * MLP_TILE * MLP_D;
  auto fslot = [&](int q, int i) { return fs_tile + (((int64_t)q * S + i) * MLP_TILE + orow) * MLP_D + c0; };
  // Rows of stage (q, i) in the segment's row buffers.
  auto stage_row0 = [&](int q, int i) { return ((int64_t)q * S + i) * a.B_pad + (int64_t)tile * MLP_TILE; };
  // The forward factors of the weight gradients (layer inputs z, h1, h2) straight from LDS
  // after the stage's MLP evaluation.
  auto stash_rows = [&](int q, int i) {
    const int64_t r0 = stage_row0(q, i);
    rows_out<P>(L.h1, L.LDH, MLP_H, R.b[3], r0);
    rows_out<P>(L.h2, L.LDH, MLP_H, R.b[5], r0);
    rows_out<P>(L.x, L.LDX, MLP_D, R.b[1], r0);
  };

  // ---- 1. recompute the segment: keep every stage's f (scratch) and its layer inputs (row
  //      buffers), so that the reverse sweep needs no forward re-evaluation ----
  float y[CPT], fl[CPT];
#pragma unroll
  for (int j = 0; j < CPT; ++j) y[j] = (steps_here > 0) ? a.ckpt[(traj * a.n_ckpt + a.seg) * MLP_D + c0 + j] : 0.0f;
  for (int q = 0; q < a.Cs; ++q) {
    const bool act = q < steps_here;
    const int n = n0 + q;
    const double ta = act ? a.grid[(traj * a.max_steps + n) * 2 + 0] : 0.0;
    const double tb = act ? a.grid[(traj * a.max_steps + n) * 2 + 1] : 0.0;
    const float control = (float)((double)dir * (tb - ta));
    float k[S][CPT], z[CPT], f[CPT];
    if (q == 0) {
#pragma unroll
      for (int j = 0; j < CPT; ++j) store_input<P>(L, orow, c0 + j, y[j]);
      __syncthreads();
      mlp_forward<P, NW>(W, L, wave, lane, orow, c0, fl);
      stash_rows(0, 0);
      __syncthreads();
    }
    {
      // stage 0 is the FSAL evaluation: f(y_n), whose rows the previous step's last stage
      // (same input, z_6 = y_{n+1}) or the evaluation above wrote
      float* fp = fslot(q, 0);
#pragma unroll
      for (int j = 0; j < CPT; ++j) { fp[j] = fl[j]; k[0][j] = control * fl[j]; }
    }
    static_for<1, S>([&](auto ic) {
      constexpr int i = decltype(ic)::value;
#pragma unroll
      for (int j = 0; j < CPT; ++j) {
        float kc[S];
#pragma unroll
        for (int m = 0; m < S; ++m) kc[m] = (m < i) ? k[m][j] : 0.0f;
        z[j] = y[j] + comb<Tab, 0, i, 0, i, false>(TL::v.a[i], kc, 0.0f);
        store_input<P>(L, orow, c0 + j, z[j]);
      }
      __syncthreads();
      mlp_forward<P, NW>(W, L, wave, lane, orow, c0, f);
      stash_rows(q, i);
      if constexpr (i == S - 1) {
        if (q + 1 < a.Cs) stash_rows(q + 1, 0);
      }
      __syncthreads();
      float* fp = fslot(q, i);
#pragma unroll
      for (int j = 0; j < CPT; ++j) { fp[j] = f[j]; k[i][j] = control * f[j]; }
    });
#pragma unroll
    for (int j = 0; j < CPT; ++j) {
      if (act) y[j] = z[j];
      fl[j] = f[j];
    }
  }

  // ---- 2. reverse sweep ----
  float yb[CPT];
#pragma unroll
  for (int j = 0; j < CPT; ++j) yb[j] = valid ? a.ybar[traj * MLP_D + c0 + j] : 0.0f;
  for (int q = a.Cs - 1; q >= 0; --q) {
    const bool act = q < steps_here;
    const int n = n0 + q;
    const double ta = act ? a.grid[(traj * a.max_steps + n) * 2 + 0] : 0.0;
    const double tb = act ? a.grid[(traj * a.max_steps + n) * 2 + 1] : 0.0;
    const float control = (float)((double)dir * (tb - ta));
    float kb[S][CPT], ybn[CPT];
#pragma unroll
    for (int j = 0; j < CPT; ++j) {
      ybn[j] = yb[j];
#pragma unroll
      for (int m = 0; m < S; ++m) kb[m][j] = 0.0f;
    }
    if (act) {
      // y_{n+1} = z_6 = y_n + sum_j a_6j k_j
#pragma unroll
      for (int m = 0; m < S - 1; ++m) {
#pragma unroll
        for (int j = 0; j < CPT; ++j) kb[m][j] += TL::v.a[S - 1][m] * yb[j];
      }
      // SaveAt(ts) values interpolated inside this step
      int jlo = 0, jhi = 0;
      for (int s = 0; s < a.n_ts; ++s) {
        const double tsn = ts_row[s] * dir;
        if (n > 0 && tsn <= ta) jlo = s + 1;
        if (tsn <= tb) jhi = s + 1;
      }
      for (int s = jlo; s < jhi; ++s) {
        const double th = linear_rescale(ta, ts_row[s] * dir, tb);
        double bb[7];
        tsit5_weights(th, bb);
        const float* g = a.ys_bar + (traj * a.n_out + (a.save_t0 ? 1 : 0) + s) * MLP_D + c0;
#pragma unroll
        for (int j = 0; j < CPT; ++j) {
          const float gv = g[j];
          ybn[j] += gv;
#pragma unroll
          for (int m = 0; m < 7; ++m) kb[m][j] += (float)bb[m] * gv;
        }
      }
    }
    // stages in reverse
    static_for<0, S>([&](auto ic) {
      constexpr int i = S - 1 - decltype(ic)::value;
      const int64_t row0 = stage_row0(q, i);
      const float* fp = fslot(q, i);
      // delta3 = fbar * (1 - f^2), fbar = control * kbar_i
#pragma unroll
      for (int j = 0; j < CPT; ++j) {
        const float fb = act ? control * kb[i][j] : 0.0f;
        const float fv = fp[j];
        L.d3[orow * L.LDX + c0 + j] = P::to_act(fb * (1.0f - fv * fv));
      }
      __syncthreads();
      rows_out<P>(L.d3, L.LDX, MLP_D, R.b[4], row0);
      // h2bar = d3 . W3 ; delta2 = h2bar * softplus'(.) (h2 from its row buffer)
      floatx16_t acc[NT];
      wave_gemm<P, MLP_D / P::KSTEP, NT>(L.d3, L.d3, L.LDX, W.w3t, NT * wave, 0, MLP_D / P::KSTEP,
                                         acc, lane);
#pragma unroll
      for (int t = 0; t < NT; ++t) {
        const int col = 32 * (NT * wave + t) + (lane & 31);
#pragma unroll
        for (int v = 0; v < 16; ++v) {
          const int r = acc_row(lane, v);
          const float h = row_val<P>(R.b[5], MLP_H, row0 + r, col);
          L.d[r * L.LDH + col] = P::to_act(acc[t][v] * dsoftplus_from_value<F>(h));
        }
      }
      __syncthreads();
      // h1bar = d2 . W2 ; delta1 = h1bar * softplus'(.), kept in L.h2
      wave_gemm<P, MLP_H / P::KSTEP, NT>(L.d, L.d, L.LDH, W.w2t, NT * wave, 0, MLP_H / P::KSTEP,
                                         acc, lane);
#pragma unroll
      for (int t = 0; t < NT; ++t) {
        const int col = 32 * (NT * wave + t) + (lane & 31);
#pragma unroll
        for (int v = 0; v < 16; ++v) {
          const int r = acc_row(lane, v);
          const float h = row_val<P>(R.b[3], MLP_H, row0 + r, col);
          L.h2[r * L.LDH + col] = P::to_act(acc[t][v] * dsoftplus_from_value<F>(h));
        }
      }
      __syncthreads();
      rows_out<P>(L.d, L.LDH, MLP_H, R.b[2], row0);
      rows_out<P>(L.h2, L.LDH, MLP_H, R.b[0], row0);
      // zbar = d1 . W1 (split-K over the 256 hidden units)
      constexpr int KS1 = MLP_H / P::KSTEP;
      const int qk = WV::quarter(wave), n3 = WV::nt3(wave);
      floatx16_t acc3[NT3];
      wave_gemm<P, KS1, NT3>(L.h2, L.h2, L.LDH, W.w1t, n3, qk * (KS1 / 4), (qk + 1) * (KS1 / 4),
                             acc3, lane);
#pragma unroll
      for (int t = 0; t < NT3; ++t) {
        const int col = 32 * (n3 + t) + (lane & 31);
#pragma unroll
        for (int v = 0; v < 16; ++v) L.part[(qk * MLP_TILE + acc_row(lane, v)) * L.LDP + col] = acc3[t][v];
      }
      __syncthreads();
      float zb[CPT];
#pragma unroll
      for (int j = 0; j < CPT; ++j) {
        const int c = c0 + j;
        float s = L.part[(0 * MLP_TILE + orow) * L.LDP + c];
        s += L.part[(1 * MLP_TILE + orow) * L.LDP + c];
        s += L.part[(2 * MLP_TILE + orow) * L.LDP + c];
        s += L.part[(3 * MLP_TILE + orow) * L.LDP + c];
        zb[j] = act ? s : 0.0f;
      }
      // z_i = y_n + sum_{m<i} a_im k_m
#pragma unroll
      for (int j = 0; j < CPT; ++j) {
        ybn[j] += zb[j];
#pragma unroll
        for (int m = 0; m < i; ++m) kb[m][j] += TL::v.a[i][m] * zb[j];
      }
      __syncthreads();
    });
    if (act) {
#pragma unroll
      for (int j = 0; j < CPT; ++j) yb[j] = ybn[j];
    }
  }
  if (valid) {
#pragma unroll
    for (int j = 0; j < CPT; ++j) a.ybar[traj * MLP_D + c0 + j] = yb[j];
  }
}

// ---- the three-piece bfloat16 form's adjoint ----------------------------------------------
// The recompute and reverse sweep of mlp_adj_segment_kernel with every product on the
// six-pass three-piece MFMA (wave_gemm3): the deltas are split like the activations, so the
// gradients are those of the float32-weight model that the forward evaluates.  The row
// buffers hold float32 values (layer inputs and deltas written before their split into
// pieces) for the float32 weight-gradient GEMM.  LDS: the forward planes;
// in the sweep the x planes take delta3, the h1 planes delta2 and then the split-K partials,
// the h2 planes delta1 (the activations come back from the row buffers).  NW waves per tile
// share the products as in the forward (MlpWaves); the arithmetic of every element does not
// depend on NW.

template <int NW>
__global__ void __launch_bounds__(64 * NW, 1)
mlp_adj_segment3_kernel(const AdjArgs a, const MlpWeights<PolicyBF16x3> W,
                        const RowBufs<PolicyF32> R) {
  using P = PolicyBF16x3;
  extern __shared__ __align__(16) unsigned char smem[];
  MlpLds<P>& L = *reinterpret_cast<MlpLds<P>*>(smem);
  using Tab = Tsit5Tab;
  constexpr int S = Tab::S;
  using WV = MlpWaves<NW>;
  constexpr int CPT = WV::CPT, GRP = MLP_D / CPT, NT = WV::NT12, NT3 = WV::NT3;
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const int orow = tid / GRP, c0 = (tid % GRP) * CPT;
  const int tile = blockIdx.x;
  const int64_t traj = (int64_t)tile * MLP_TILE + orow;
  const bool valid = traj < a.B;
  const int nacc = valid ? a.stats[traj * 3 + 1] : 0;
  const int n0 = a.seg * a.C;
  int steps_here = nacc - n0;
  steps_here = steps_here < 0 ? 0 : (steps_here > a.Cs ? a.Cs : steps_here);
  const int dir = (a.t0 < a.t1) ? 1 : -1;
  const double* ts_row = a.ts ? a.ts + (valid ? traj : 0) * a.ts_stride : nullptr;
  using TL = TabLit<Tab, float>;
  float* fs_tile = a.zs + (int64_t)tile * a.Cs * S * MLP_TILE * MLP_D;
  auto fslot = [&](int q, int i) { return fs_tile + (((int64_t)q * S + i) * MLP_TILE + orow) * MLP_D + c0; };
  auto stage_row0 = [&](int q, int i) { return ((int64_t)q * S + i) * a.B_pad + (int64_t)tile * MLP_TILE; };
  // The layer inputs of stage (q, i) go to the row buffers as they are formed: the stage
  // value by its owner threads, the activations from the forward's epilogues.  The last
  // stage's rows are also the next step's stage-0 rows (FSAL: the same input).
  auto sink_of = [&](int q, int i) {
    MlpRowSink k;
    k.h1 = R.b[3];
    k.h2 = R.b[5];
    k.row0 = stage_row0(q, i);
    k.row0b = (i == S - 1 && q + 1 < a.Cs) ? stage_row0(q + 1, 0) : -1;
    return k;
  };
  auto put_x = [&](const MlpRowSink& k, int j, float v) {
    R.b[1][(k.row0 + orow) * MLP_D + c0 + j] = v;
    if (k.row0b >= 0) R.b[1][(k.row0b + orow) * MLP_D + c0 + j] = v;
  };

  // ---- 1. recompute the segment (as mlp_adj_segment_kernel) ----
  float y[CPT], fl[CPT];
#pragma unroll
  for (int j = 0; j < CPT; ++j) y[j] = (steps_here > 0) ? a.ckpt[(traj * a.n_ckpt + a.seg) * MLP_D + c0 + j] : 0.0f;
  for (int q = 0; q < a.Cs; ++q) {
    const bool act = q < steps_here;
    const int n = n0 + q;
    const double ta = act ? a.grid[(traj * a.max_steps + n) * 2 + 0] : 0.0;
    const double tb = act ? a.grid[(traj * a.max_steps + n) * 2 + 1] : 0.0;
    const float control = (float)((double)dir * (tb - ta));
    float k[S][CPT], z[CPT], f[CPT];
    if (q == 0) {
      const MlpRowSink k0 = sink_of(0, 0);
#pragma unroll
      for (int j = 0; j < CPT; ++j) {
        store_input<P>(L, orow, c0 + j, y[j]);
        put_x(k0, j, y[j]);
      }
      __syncthreads();
      mlp_forward3<NW>(W, L, wave, lane, orow, c0, fl, k0);
      __syncthreads();
    }
    {
      float* fp = fslot(q, 0);
#pragma unroll
      for (int j = 0; j < CPT; ++j) { fp[j] = fl[j]; k[0][j] = control * fl[j]; }
    }
    static_for<1, S>([&](auto ic) {
      constexpr int i = decltype(ic)::value;
      const MlpRowSink ki = sink_of(q, i);
#pragma unroll
      for (int j = 0; j < CPT; ++j) {
        float kc[S];
#pragma unroll
        for (int m = 0; m < S; ++m) kc[m] = (m < i) ? k[m][j] : 0.0f;
        z[j] = y[j] + comb<Tab, 0, i, 0, i, false>(TL::v.a[i], kc, 0.0f);
        store_input<P>(L, orow, c0 + j, z[j]);
        put_x(ki, j, z[j]);
      }
      __syncthreads();
      mlp_forward3<NW>(W, L, wave, lane, orow, c0, f, ki);
      __syncthreads();
      float* fp = fslot(q, i);
#pragma unroll
      for (int j = 0; j < CPT; ++j) { fp[j] = f[j]; k[i][j] = control * f[j]; }
    });
#pragma unroll
    for (int j = 0; j < CPT; ++j) {
      if (act) y[j] = z[j];
      fl[j] = f[j];
    }
  }

  // ---- 2. reverse sweep ----
  float yb[CPT];
#pragma unroll
  for (int j = 0; j < CPT; ++j) yb[j] = valid ? a.ybar[traj * MLP_D + c0 + j] : 0.0f;
  for (int q = a.Cs - 1; q >= 0; --q) {
    const bool act = q < steps_here;
    const int n = n0 + q;
    const double ta = act ? a.grid[(traj * a.max_steps + n) * 2 + 0] : 0.0;
    const double tb = act ? a.grid[(traj * a.max_steps + n) * 2 + 1] : 0.0;
    const float control = (float)((double)dir * (tb - ta));
    float kb[S][CPT], ybn[CPT];
#pragma unroll
    for (int j = 0; j < CPT; ++j) {
      ybn[j] = yb[j];
#pragma unroll
      for (int m = 0; m < S; ++m) kb[m][j] = 0.0f;
    }
    if (act) {
#pragma unroll
      for (int m = 0; m < S - 1; ++m) {
#pragma unroll
        for (int j = 0; j < CPT; ++j) kb[m][j] += TL::v.a[S - 1][m] * yb[j];
      }
      int jlo = 0, jhi = 0;
      for (int s = 0; s < a.n_ts; ++s) {
        const double tsn = ts_row[s] * dir;
        if (n > 0 && tsn <= ta) jlo = s + 1;
        if (tsn <= tb) jhi = s + 1;
      }
      for (int s = jlo; s < jhi; ++s) {
        const double th = linear_rescale(ta, ts_row[s] * dir, tb);
        double bb[7];
        tsit5_weights(th, bb);
        const float* g = a.ys_bar + (traj * a.n_out + (a.save_t0 ? 1 : 0) + s) * MLP_D + c0;
#pragma unroll
        for (int j = 0; j < CPT; ++j) {
          const float gv = g[j];
          ybn[j] += gv;
#pragma unroll
          for (int m = 0; m < 7; ++m) kb[m][j] += (float)bb[m] * gv;
        }
      }
    }
    static_for<0, S>([&](auto ic) {
      constexpr int i = S - 1 - decltype(ic)::value;
      const int64_t row0 = stage_row0(q, i);
      const float* fp = fslot(q, i);
      // delta3 = fbar (1 - f^2), fbar = control * kbar_i: its row, and its pieces in x
#pragma unroll
      for (int j = 0; j < CPT; ++j) {
        const float fb = act ? control * kb[i][j] : 0.0f;
        const float fv = fp[j];
        const float d3 = fb * (1.0f - fv * fv);
        R.b[4][(row0 + orow) * MLP_D + c0 + j] = d3;
        const int li = orow * L.LDX + c0 + j;
        split3(d3, L.x[0][li], L.x[1][li], L.x[2][li]);
      }
      __syncthreads();
      // delta2 = (delta3 . W3) softplus'(h2): its row, and its pieces in h1
      floatx16_t acc[NT];
      wave_gemm3<MLP_D / 16, NT>(L.x[0], L.x[1], L.x[2], L.LDX, W.w3t, W.plane3, NT * wave, 0,
                                 MLP_D / 16, acc, lane);
#pragma unroll
      for (int t = 0; t < NT; ++t) {
        const int col = 32 * (NT * wave + t) + (lane & 31);
#pragma unroll
        for (int v = 0; v < 16; ++v) {
          const int r = acc_row(lane, v);
          const float h = R.b[5][(row0 + r) * MLP_H + col];
          const float d2 = acc[t][v] * dsoftplus_from_value<true>(h);
          R.b[2][(row0 + r) * MLP_H + col] = d2;
          const int li = r * L.LDH + col;
          split3(d2, L.h1[0][li], L.h1[1][li], L.h1[2][li]);
        }
      }
      __syncthreads();
      // delta1 = (delta2 . W2) softplus'(h1): its row, and its pieces in h2
      wave_gemm3<MLP_H / 16, NT>(L.h1[0], L.h1[1], L.h1[2], L.LDH, W.w2t, W.plane2, NT * wave, 0,
                                 MLP_H / 16, acc, lane);
#pragma unroll
      for (int t = 0; t < NT; ++t) {
        const int col = 32 * (NT * wave + t) + (lane & 31);
#pragma unroll
        for (int v = 0; v < 16; ++v) {
          const int r = acc_row(lane, v);
          const float h = R.b[3][(row0 + r) * MLP_H + col];
          const float d1 = acc[t][v] * dsoftplus_from_value<true>(h);
          R.b[0][(row0 + r) * MLP_H + col] = d1;
          const int li = r * L.LDH + col;
          split3(d1, L.h2[0][li], L.h2[1][li], L.h2[2][li]);
        }
      }
      __syncthreads();
      // zbar = delta1 . W1 (split-K over the 256 hidden units); the partials take the h1
      // planes, whose last readers (the delta1 products) are behind the barrier above
      constexpr int KS1 = MLP_H / 16;
      const int qk = WV::quarter(wave), n3 = WV::nt3(wave);
      floatx16_t acc3[NT3];
      wave_gemm3<KS1, NT3>(L.h2[0], L.h2[1], L.h2[2], L.LDH, W.w1t, W.plane1, n3,
                           qk * (KS1 / 4), (qk + 1) * (KS1 / 4), acc3, lane);
      float* part = L.part();
#pragma unroll
      for (int t = 0; t < NT3; ++t) {
        const int col = 32 * (n3 + t) + (lane & 31);
#pragma unroll
        for (int v = 0; v < 16; ++v) part[(qk * MLP_TILE + acc_row(lane, v)) * L.LDP + col] = acc3[t][v];
      }
      __syncthreads();
      float zb[CPT];
#pragma unroll
      for (int j = 0; j < CPT; ++j) {
        const int c = c0 + j;
        float sv = part[(0 * MLP_TILE + orow) * L.LDP + c];
        sv += part[(1 * MLP_TILE + orow) * L.LDP + c];
        sv += part[(2 * MLP_TILE + orow) * L.LDP + c];
        sv += part[(3 * MLP_TILE + orow) * L.LDP + c];
        zb[j] = act ? sv : 0.0f;
      }
#pragma unroll
      for (int j = 0; j < CPT; ++j) {
        ybn[j] += zb[j];
#pragma unroll
        for (int m = 0; m < i; ++m) kb[m][j] += TL::v.a[i][m] * zb[j];
      }
      __syncthreads();
    });
    if (act) {
#pragma unroll
      for (int j = 0; j < CPT; ++j) yb[j] = ybn[j];
    }
  }
  if (valid) {
#pragma unroll
    for (int j = 0; j < CPT; ++j) a.ybar[traj * MLP_D + c0 + j] = yb[j];
  }
}

// ---- weight gradients of one segment ---------------------------------------------------
// Output tiles: matrix 0 = dW1 [256 x 64] (D1, Z), 1 = dW2 [256 x 256] (D2, H1),
// 2 = dW3 [64 x 256] (D3, H2); 32x32 tiles numbered matrix by matrix, row-major.
struct DwLayout {
  int n1[3], n2[3], first[4];
};
__host__ __device__ inline DwLayout dw_layout() {
  DwLayout l{{MLP_H, MLP_H, MLP_D}, {MLP_D, MLP_H, MLP_H}, {0, 0, 0, 0}};
  for (int m = 0; m < 3; ++m) l.first[m + 1] = l.first[m] + (l.n1[m] / 32) * (l.n2[m] / 32);
  return l;
}

// Launch order is XCD-aware: workgroups are dealt to the 8 XCDs round-robin by linear id,
// so block b runs on XCD b % 8; XCD x takes chunks x, x + 8, ... and runs the blocks of a
// chunk back to back, which then read the chunk's rows through one L2 instead of eight.
constexpr int kXcds = 8;

// partial[chunk][tile] = A^T B over one chunk of ADJ_CHUNK rows (A = deltas, B = layer
// inputs) for every 32 x 32 output tile, each wave on a 2 x 2 block of tiles (a 64 x 64
// piece of one matrix): per k-step it loads two delta and two input fragments for four
// MFMAs (a wave per tile, measured before, loaded twice as much per MFMA and took the
// backward of 8,192 trajectories from 2.84 to 2.91 ms, the backsolve from 19.0 to 19.6 ms,
// with the same bits).  The fragments come in groups of 4 k-steps held in two register
// buffers (the next group loads while the current one feeds the MFMAs).  The row buffers
// are allocated in whole chunks and rows past the segment's last are zero, so every chunk
// runs its full length.  The 96 tiles of a chunk form 24 blocks (dW1 4 x 1, dW2 4 x 4,
// dW3 1 x 4); blocks in the first block column of a matrix also produce the chunk's column
// sums of A (the bias gradient partials, partial_db: [chunk][db1 | db2 | db3]).
// SC: the delta rows are multiplied by rscale[row] as they load (the backsolve's per-row
// weights dt * b_i, zero for rejected steps and idle rows; a row of zero weight is dropped
// from both operands, so a non-finite value in it contributes nothing).  Float32 rows only.
__host__ __device__ inline int dw_blocks_of(int m) {
  const DwLayout l = dw_layout();
  return (l.n1[m] / 64) * (l.n2[m] / 64);
}

template <class P, bool SC = false>
__global__ void __launch_bounds__(64) mlp_dw_gemm_kernel(const RowBufs<P> R,
                                                           float* __restrict__ partial,
                                                           float* __restrict__ partial_db,
                                                           int nch,
                                                           const float* __restrict__ rscale = nullptr) {
  static_assert(!SC || P::KA == 1, "row scales need one row per fragment element");
  const DwLayout lay = dw_layout();
  const int nblk = dw_blocks_of(0) + dw_blocks_of(1) + dw_blocks_of(2);
  const int xcd = (int)(blockIdx.x % kXcds), slot = (int)(blockIdx.x / kXcds);
  const int chunk = (slot / nblk) * kXcds + xcd;
  int bg = slot - (slot / nblk) * nblk;
  if (chunk >= nch) return;
  int m = 0;
  while (bg >= dw_blocks_of(m)) { bg -= dw_blocks_of(m); ++m; }
  const int nbj = lay.n2[m] / 64;
  const int bi = bg / nbj, bj = bg - bi * nbj;
  const typename P::act_t* A = R.b[2 * m];       // deltas (width n1)
  const typename P::act_t* Bm = R.b[2 * m + 1];  // layer inputs (width n2)
  const int WA = lay.n1[m], WB = lay.n2[m];
  const int lane = threadIdx.x;
  constexpr int KA = P::KA, KS = P::KSTEP, U = 4;
  static_assert(ADJ_CHUNK % (KS * U) == 0, "chunk must hold whole k-step groups");
  const bool col_sums = (bj == 0);
  const int64_t r_begin = (int64_t)chunk * ADJ_CHUNK;
  floatx16_t acc[2][2];
#pragma unroll
  for (int x = 0; x < 2; ++x) { acc[x][0] = floatx16_t{}; acc[x][1] = floatx16_t{}; }
  float csum[2] = {0.0f, 0.0f};
  const int ca = bi * 64 + (lane & 31), cb = bj * 64 + (lane & 31);
  using frag_t = typename P::frag_t;
  auto load_group = [&](int64_t r0, frag_t (&fa)[U][2], frag_t (&fb)[U][2]) {
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int64_t rb = (r0 + u * KS) / KA + (lane >> 5);
#pragma unroll
      for (int x = 0; x < 2; ++x) {
        fa[u][x] = P::load_a(A + (rb * WA + ca + 32 * x) * KA);
        fb[u][x] = P::load_a(Bm + (rb * WB + cb + 32 * x) * KA);
      }
      if constexpr (SC) {
        const float w = rscale[rb];
#pragma unroll
        for (int x = 0; x < 2; ++x) {
          fa[u][x] = (w != 0.0f) ? fa[u][x] * w : 0.0f;
          fb[u][x] = (w != 0.0f) ? fb[u][x] : 0.0f;
        }
      }
    }
  };
  auto mma_group = [&](const frag_t (&fa)[U][2], const frag_t (&fb)[U][2]) {
#pragma unroll
    for (int u = 0; u < U; ++u) {
#pragma unroll
      for (int x = 0; x < 2; ++x) {
        acc[x][0] = P::mfma(fa[u][x], fb[u][0], acc[x][0]);
        acc[x][1] = P::mfma(fa[u][x], fb[u][1], acc[x][1]);
        if (col_sums) {
          if constexpr (KA == 1) {
            csum[x] += fa[u][x];
          } else {
#pragma unroll
            for (int j = 0; j < KA; ++j) csum[x] += (float)fa[u][x][j];
          }
        }
      }
    }
  };
  constexpr int G = ADJ_CHUNK / (KS * U);
  static_assert(G % 2 == 0, "groups are processed in pairs");
  frag_t a0[U][2], b0[U][2], a1[U][2], b1[U][2];
  load_group(r_begin, a0, b0);
  for (int g = 0; g < G; g += 2) {
    load_group(r_begin + (int64_t)(g + 1) * KS * U, a1, b1);
    mma_group(a0, b0);
    load_group(r_begin + (int64_t)min(g + 2, G - 1) * KS * U, a0, b0);
    mma_group(a1, b1);
  }
  const int ntj = lay.n2[m] / 32;
#pragma unroll
  for (int x = 0; x < 2; ++x) {
#pragma unroll
    for (int y = 0; y < 2; ++y) {
      const int tg = lay.first[m] + (2 * bi + x) * ntj + (2 * bj + y);
      float* o = partial + ((int64_t)chunk * lay.first[3] + tg) * 1024 + lane * 16;
#pragma unroll
      for (int v = 0; v < 16; ++v) o[v] = acc[x][y][v];
    }
  }
  if (col_sums) {
    const int off = (m == 0) ? 0 : (m == 1 ? MLP_H : 2 * MLP_H);
#pragma unroll
    for (int x = 0; x < 2; ++x) {
      const float c = csum[x] + __shfl_xor(csum[x], 32);
      if (lane < 32) partial_db[(int64_t)chunk * (2 * MLP_H + MLP_D) + off + ca + 32 * x] = c;
    }
  }
}

// dw_acc (dW1 | dW2 | dW3 | db1 | db2 | db3, row-major) += partials summed in chunk order.
__global__ void mlp_dw_reduce_kernel(const float* __restrict__ partial,
                                     const float* __restrict__ partial_db, int nchunks,
                                     float* __restrict__ dw_acc) {
  const DwLayout lay = dw_layout();
  const int nW = MLP_H * MLP_D + MLP_H * MLP_H + MLP_D * MLP_H;
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nW + 2 * MLP_H + MLP_D) return;
  float s = 0.0f;
  if (e < nW) {
    int m = 0, off = e;
    const int sz[3] = {MLP_H * MLP_D, MLP_H * MLP_H, MLP_D * MLP_H};
    while (off >= sz[m]) { off -= sz[m]; ++m; }
    const int n1 = off / lay.n2[m], n2 = off - (off / lay.n2[m]) * lay.n2[m];
    const int tg = lay.first[m] + (n1 / 32) * (lay.n2[m] / 32) + n2 / 32;
    const int r = n1 % 32, c = n2 % 32;
    const int lane = c + 32 * ((r % 8) / 4);
    const int v = 4 * (r / 8) + (r % 4);
    for (int ch = 0; ch < nchunks; ++ch) s += partial[((int64_t)ch * lay.first[3] + tg) * 1024 + lane * 16 + v];
  } else {
    const int col = e - nW;
    for (int ch = 0; ch < nchunks; ++ch) s += partial_db[(int64_t)ch * (2 * MLP_H + MLP_D) + col];
  }
  dw_acc[e] += s;
}

// ---- BacksolveAdjoint of the neural ODE ---------------------------------------------
// The continuous adjoint (_adjoint.py:539-873) with the adjoint seminorm
// (_misc.py:85-116): between consecutive save times, from the last one back to t0, the
// augmented state (y, a_y, a_theta) is integrated backwards by Tsit5 with the integral PID
// controller, after a_y takes that save's cotangent and y restarts from the forward's saved
// state, under the adjoint vector field (AdjointTerm, _term.py:767-917):
//   dy/dt = f(y),  da_y/dt = vjp(f, y)(-a_y),  da_theta/dt = vjp(f, theta)(-a_y).
// The seminorm leaves a_theta out of the error norm (RMS over the 128 components of y and
// a_y), so a_theta only accumulates: a_theta(t0) = sum over accepted steps and stages of
// dt b_i vjp_theta(z_i).  Each stage's vjp_theta is the set of outer products delta_l x
// input_l; the kernel writes those rows (unscaled) to the row buffers and their weights
// dt b_i (0 for rejected steps and idle rows) to rscale, and the split-K GEMM above forms
// the sum over all rows (trajectories, steps and stages) for the launch.
// One workgroup per tile of 32 rows; each launch runs at most Q lock-step iterations (one
// attempted step of every row) and leaves unfinished rows' state in BsRow / ya for the
// next launch.  Each iteration evaluates stage 0 afresh: f(y) at an unchanged y gives the
// same value as the FSAL stage, and fresh evaluation also covers the first step after
// each save (made_jump).
struct BsRow {
  double tprev, tnext, T1, floor_t1, inv, pinv;
  int32_t at_dtmin, s, nsteps, dir, result, res_total, done, pad_;
};

struct BsArgs {
  const float* ys;       // [B, n_out, 64] forward saved states
  const double* ts_out;  // [B, n_out] their times
  const float* ys_bar;   // [B, n_out, 64] cotangents
  float* ya;             // [B_pad, 128] y then a_y of each row between launches
  BsRow* rows;           // [B_pad]
  float* rscale;         // [Q * 7 * B_pad] weights of the weight-gradient rows
  float* ybar;           // out [B, 64] cotangent of y0
  int32_t* result;       // out [B]
  int32_t* counters;     // [2]: rows unfinished after the launch, iterations run (max)
  int64_t B, B_pad;
  int32_t n_out, save_t0, max_steps, Q, first;
  double t0, dt0;        // the forward t0; dt0 of the backward solves
  PidParams pid;         // the backward controller
};

template <class P, int NW>
__global__ void __launch_bounds__(64 * NW, 1)
mlp_backsolve_kernel(const BsArgs a, const MlpWeights<P> W, const RowBufs<P> R) {
  static_assert(P::KA == 1, "float32 rows");
  extern __shared__ __align__(16) unsigned char smem[];
  MlpLds<P>& L = *reinterpret_cast<MlpLds<P>*>(smem);
  constexpr bool F = P::fast_math;
  using Tab = Tsit5Tab;
  constexpr int S = Tab::S;
  using TL = TabLit<Tab, float>;
  using WV = MlpWaves<NW>;
  constexpr int CPT = WV::CPT, GRP = MLP_D / CPT, NT = WV::NT12, NT3 = WV::NT3;
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const int orow = tid / GRP, c0 = (tid % GRP) * CPT, gbase = lane & ~(GRP - 1);
  const bool lead = (tid % GRP) == 0;
  const int tile = blockIdx.x;
  const int64_t traj = (int64_t)tile * MLP_TILE + orow;
  const bool valid = traj < a.B;
  const int s_end = a.save_t0 ? 1 : 0;

  float y[CPT], av[CPT];
  BsRow st;
  // Start the backward solve over save interval s: y from the forward's row s, a_y plus
  // that row's cotangent, span [ts_out[s], ts_out[s - 1] or t0] (_scan_fun).
  auto begin = [&](int s) {
    const float* ysr = a.ys + (traj * a.n_out + s) * MLP_D + c0;
    const float* gr = a.ys_bar + (traj * a.n_out + s) * MLP_D + c0;
#pragma unroll
    for (int j = 0; j < CPT; ++j) { y[j] = ysr[j]; av[j] = av[j] + gr[j]; }
    const double t0r = a.ts_out[traj * a.n_out + s];
    const double t1r = s > 0 ? a.ts_out[traj * a.n_out + s - 1] : a.t0;
    st.s = s;
    st.dir = (t0r < t1r) ? 1 : -1;
    const double T0 = t0r * st.dir;
    st.T1 = t1r * st.dir;
    double fl = st.T1;
    for (int r = 0; r < 100; ++r) fl = prevbefore(fl);
    st.floor_t1 = fl;
    double dt0v = a.dt0 * st.dir;
    if (a.pid.has_dtmax) dt0v = jmin(dt0v, a.pid.dtmax);
    if (a.pid.has_dtmin) dt0v = jmax(dt0v, a.pid.dtmin);
    st.tprev = T0;
    st.tnext = jmin(T0 + dt0v, st.T1);
    st.inv = 1.0; st.pinv = 1.0; st.at_dtmin = 0;
    st.nsteps = 0;
    st.result = DFX_SUCCESSFUL;
  };
  // Move past finished intervals; after the last one the row's a_y (plus the t0 row's
  // cotangent when t0 was saved) is dL/dy0.
  auto settle = [&]() {
    while (!st.done) {
      if (st.tprev < st.T1 && st.result == DFX_SUCCESSFUL && st.nsteps < a.max_steps) return;
      if (st.tprev < st.T1 && st.result == DFX_SUCCESSFUL) st.result = DFX_MAX_STEPS_REACHED;
      if (st.res_total == DFX_SUCCESSFUL) st.res_total = st.result;
      if (st.s - 1 >= s_end) {
        begin(st.s - 1);
      } else {
        if (a.save_t0 && a.n_out > 0) {
          const float* g0 = a.ys_bar + traj * a.n_out * MLP_D + c0;
#pragma unroll
          for (int j = 0; j < CPT; ++j) av[j] = av[j] + g0[j];
        }
#pragma unroll
        for (int j = 0; j < CPT; ++j) a.ybar[traj * MLP_D + c0 + j] = av[j];
        if (lead) a.result[traj] = st.res_total;
        st.done = 1;
      }
    }
  };

  if (!valid) {
    st.done = 1;
  } else if (a.first) {
#pragma unroll
    for (int j = 0; j < CPT; ++j) av[j] = 0.0f;
    st.res_total = DFX_SUCCESSFUL;
    st.done = 0;
    if (a.n_out - 1 >= s_end) {
      begin(a.n_out - 1);
    } else {
      st.tprev = st.T1 = 0.0;  // nothing saved after t0: no interval
      st.s = s_end;
      st.result = DFX_SUCCESSFUL;
    }
    settle();
  } else {
    st = a.rows[traj];
#pragma unroll
    for (int j = 0; j < CPT; ++j) {
      y[j] = a.ya[traj * 128 + c0 + j];
      av[j] = a.ya[traj * 128 + 64 + c0 + j];
    }
  }

  const int64_t rb0 = (int64_t)tile * MLP_TILE + orow;
  auto stage_row0 = [&](int q, int i) { return ((int64_t)q * S + i) * a.B_pad + (int64_t)tile * MLP_TILE; };
  int q = 0;
  for (; q < a.Q; ++q) {
    const bool run = !st.done;
    if (!__syncthreads_or(run)) break;
    const double ts0 = st.tprev, ts1 = st.tnext;
    const float control = run ? (float)((double)st.dir * (ts1 - ts0)) : 0.0f;
    float ky[S][CPT], ka[S][CPT], zy[CPT], za[CPT];
    static_for<0, S>([&](auto ic) {
      constexpr int i = decltype(ic)::value;
#pragma unroll
      for (int j = 0; j < CPT; ++j) {
        if constexpr (i == 0) {
          zy[j] = y[j];
          za[j] = av[j];
        } else {
          float cy[S], ca[S];
#pragma unroll
          for (int m = 0; m < S; ++m) {
            cy[m] = (m < i) ? ky[m][j] : 0.0f;
            ca[m] = (m < i) ? ka[m][j] : 0.0f;
          }
          zy[j] = y[j] + comb<Tab, 0, i, 0, i, false>(TL::v.a[i], cy, 0.0f);
          za[j] = av[j] + comb<Tab, 0, i, 0, i, false>(TL::v.a[i], ca, 0.0f);
        }
        if (!run) { zy[j] = 0.0f; za[j] = 0.0f; }  // idle rows: zero deltas
        store_input<P>(L, orow, c0 + j, zy[j]);
      }
      __syncthreads();
      float f[CPT];
      mlp_forward<P, NW>(W, L, wave, lane, orow, c0, f);
      const int64_t row0 = stage_row0(q, i);
      rows_out<P>(L.x, L.LDX, MLP_D, R.b[1], row0);
      rows_out<P>(L.h1, L.LDH, MLP_H, R.b[3], row0);
      rows_out<P>(L.h2, L.LDH, MLP_H, R.b[5], row0);
      // vjp with cotangent w = -a_y of the stage: delta3 = w (1 - f^2)
#pragma unroll
      for (int j = 0; j < CPT; ++j) L.d3[orow * L.LDX + c0 + j] = P::to_act(-za[j] * (1.0f - f[j] * f[j]));
      __syncthreads();
      rows_out<P>(L.d3, L.LDX, MLP_D, R.b[4], row0);
      // delta2 = (delta3 . W3) softplus'(h2)
      floatx16_t acc[NT];
      wave_gemm<P, MLP_D / P::KSTEP, NT>(L.d3, L.d3, L.LDX, W.w3t, NT * wave, 0, MLP_D / P::KSTEP,
                                         acc, lane);
#pragma unroll
      for (int t = 0; t < NT; ++t) {
        const int col = 32 * (NT * wave + t) + (lane & 31);
#pragma unroll
        for (int v = 0; v < 16; ++v) {
          const int r = acc_row(lane, v);
          const float h = (float)L.h2[r * L.LDH + col];
          L.d[r * L.LDH + col] = P::to_act(acc[t][v] * dsoftplus_from_value<F>(h));
        }
      }
      __syncthreads();
      // delta1 = (delta2 . W2) softplus'(h1), into the h2 buffer (h2 is in its row buffer)
      wave_gemm<P, MLP_H / P::KSTEP, NT>(L.d, L.d, L.LDH, W.w2t, NT * wave, 0, MLP_H / P::KSTEP,
                                         acc, lane);
#pragma unroll
      for (int t = 0; t < NT; ++t) {
        const int col = 32 * (NT * wave + t) + (lane & 31);
#pragma unroll
        for (int v = 0; v < 16; ++v) {
          const int r = acc_row(lane, v);
          const float h = (float)L.h1[r * L.LDH + col];
          L.h2[r * L.LDH + col] = P::to_act(acc[t][v] * dsoftplus_from_value<F>(h));
        }
      }
      __syncthreads();
      rows_out<P>(L.d, L.LDH, MLP_H, R.b[2], row0);
      rows_out<P>(L.h2, L.LDH, MLP_H, R.b[0], row0);
      // vjp_y = delta1 . W1 (split-K over the 256 hidden units)
      constexpr int KS1 = MLP_H / P::KSTEP;
      const int qk = WV::quarter(wave), n3 = WV::nt3(wave);
      floatx16_t acc3[NT3];
      wave_gemm<P, KS1, NT3>(L.h2, L.h2, L.LDH, W.w1t, n3, qk * (KS1 / 4), (qk + 1) * (KS1 / 4),
                             acc3, lane);
#pragma unroll
      for (int t = 0; t < NT3; ++t) {
        const int col = 32 * (n3 + t) + (lane & 31);
#pragma unroll
        for (int v = 0; v < 16; ++v) L.part[(qk * MLP_TILE + acc_row(lane, v)) * L.LDP + col] = acc3[t][v];
      }
      __syncthreads();
#pragma unroll
      for (int j = 0; j < CPT; ++j) {
        const int c = c0 + j;
        float sv = L.part[(0 * MLP_TILE + orow) * L.LDP + c];
        sv += L.part[(1 * MLP_TILE + orow) * L.LDP + c];
        sv += L.part[(2 * MLP_TILE + orow) * L.LDP + c];
        sv += L.part[(3 * MLP_TILE + orow) * L.LDP + c];
        ky[i][j] = control * f[j];
        ka[i][j] = control * sv;
      }
      __syncthreads();
    });
    // y1, a1 = the last stage (SSAL); embedded error over the 128 components of (y, a_y)
    bool anynan = false;
    float ey[CPT], ea[CPT];
#pragma unroll
    for (int j = 0; j < CPT; ++j) {
      float cy[S], ca[S];
#pragma unroll
      for (int m = 0; m < S; ++m) { cy[m] = ky[m][j]; ca[m] = ka[m][j]; }
      ey[j] = comb<Tab, 2, 0, 0, S, false>(TL::v.berr, cy, 0.0f);
      ea[j] = comb<Tab, 2, 0, 0, S, false>(TL::v.berr, ca, 0.0f);
      anynan = anynan || (zy[j] != zy[j]) || (za[j] != za[j]);
    }
    const unsigned long long nanmask = __ballot(anynan);
    const bool group_nan = ((nanmask >> gbase) & ((1ull << GRP) - 1)) != 0;
    auto sq_err = [&](float e0, float v0, float v1) {
      const double v1c = group_nan ? (double)v0 : (double)v1;
      const double yy = jmax(fabs((double)v0), fabs(v1c));
      double e = (double)e0;
      if (e != e) e = __builtin_inf();
      e = e / (a.pid.atol + yy * a.pid.rtol);
      return e * e;
    };
    double ty[CPT], ta[CPT];
#pragma unroll
    for (int j = 0; j < CPT; ++j) {
      ty[j] = sq_err(ey[j], y[j], zy[j]);
      ta[j] = sq_err(ea[j], av[j], za[j]);
    }
    // Eight partials, one per eight-component block: the block's y terms then its a_y
    // terms, in component order; then the partials in order.  With four components per
    // owner, the pair's sums pass between its two owners so that the order is the same.
    double part = 0.0;
    int stride = 1, first = 0;
    if constexpr (CPT == 8) {
#pragma unroll
      for (int j = 0; j < CPT; ++j) part += ty[j];
#pragma unroll
      for (int j = 0; j < CPT; ++j) part += ta[j];
    } else {
#pragma unroll
      for (int j = 0; j < CPT; ++j) part += ty[j];      // even: its y terms
      part = __shfl(part, lane & ~1);
#pragma unroll
      for (int j = 0; j < CPT; ++j) part += ty[j];      // odd: + its y terms
      part = __shfl(part, lane | 1);
#pragma unroll
      for (int j = 0; j < CPT; ++j) part += ta[j];      // even: + its a_y terms
      part = __shfl(part, lane & ~1);
#pragma unroll
      for (int j = 0; j < CPT; ++j) part += ta[j];      // odd: + its a_y terms
      stride = 2;
      first = 1;
    }
    double ssum = 0.0;
#pragma unroll
    for (int k = 0; k < 8; ++k) ssum += __shfl(part, gbase + stride * k + first);
    const double scaled = sqrt(ssum / 128.0);
    PidState npid;
    npid.inv = st.inv; npid.pinv = st.pinv; npid.at_dtmin = st.at_dtmin;
    double ntprev, ntnext;
    int code = 0;
    const bool keep = pid_adapt<PID_INTEGRAL>(a.pid, scaled, ts0, ts1, npid, ntprev, ntnext, code);
    ntprev = jmin(ntprev, st.T1);
    if (ntnext > st.floor_t1) ntnext = keep ? st.T1 : ntprev + 0.5 * (st.T1 - ntprev);
    // weights of this iteration's weight-gradient rows: dt b_i for an accepted step
    if (lead) {
#pragma unroll
      for (int i = 0; i < S; ++i)
        a.rscale[((int64_t)q * S + i) * a.B_pad + rb0] = (run && keep) ? control * TL::v.b[i] : 0.0f;
    }
    if (run) {
      st.inv = npid.inv; st.pinv = npid.pinv; st.at_dtmin = npid.at_dtmin;
      ++st.nsteps;
      if (st.result == DFX_SUCCESSFUL) st.result = code;
      if (keep) {
#pragma unroll
        for (int j = 0; j < CPT; ++j) { y[j] = zy[j]; av[j] = za[j]; }
      }
      st.tprev = ntprev;
      st.tnext = ntnext;
      settle();
    }
  }
  // rows of the iterations this tile did not run carry no weight
  if (lead) {
    for (int qq = q; qq < a.Q; ++qq) {
#pragma unroll
      for (int i = 0; i < S; ++i) a.rscale[((int64_t)qq * S + i) * a.B_pad + rb0] = 0.0f;
    }
  }
  if (tid == 0 && q > 0) atomicMax(&a.counters[1], q);
  // Every row's state goes back, finished rows included: the next launch must see their
  // done flag (or it would run their intervals again from an older state).
  if (valid) {
    if (!st.done) {
#pragma unroll
      for (int j = 0; j < CPT; ++j) {
        a.ya[traj * 128 + c0 + j] = y[j];
        a.ya[traj * 128 + 64 + c0 + j] = av[j];
      }
    }
    if (lead) {
      a.rows[traj] = st;
      if (!st.done) atomicAdd(&a.counters[0], 1);
    }
  }
}

}  // namespace dfx

using namespace dfx;

namespace {

int64_t padded_batch(int64_t B) { return (B + MLP_TILE - 1) / MLP_TILE * MLP_TILE; }

int64_t seg_rows(int64_t B, int32_t seg_steps) { return (int64_t)seg_steps * 7 * padded_batch(B); }

int64_t seg_chunks(int64_t B, int32_t seg_steps) {
  return (seg_rows(B, seg_steps) + ADJ_CHUNK - 1) / ADJ_CHUNK;
}

// The backsolve's state buffer: ya [B_pad, 128] floats, rows [B_pad] BsRow, rscale [nrow]
// floats (every row of the allocated row buffers), counters [2] int32; 256-byte aligned.
struct BsStateLayout {
  size_t ya, rows, rscale, counters, total;
};
BsStateLayout bs_state_layout(int64_t Bp, int64_t nrow) {
  auto up = [](size_t x) { return (x + 255) / 256 * 256; };
  BsStateLayout l;
  l.ya = 0;
  l.rows = up(l.ya + (size_t)Bp * 128 * sizeof(float));
  l.rscale = up(l.rows + (size_t)Bp * sizeof(BsRow));
  l.counters = up(l.rscale + (size_t)nrow * sizeof(float));
  l.total = up(l.counters + 2 * sizeof(int32_t));
  return l;
}

// Zero the launch counters and the row weights past the kernel's rows (the padding of the
// last split-K chunk), which the kernel never writes.
__global__ void bs_reset_kernel(int32_t* counters, float* rscale, int64_t used, int64_t nrow) {
  if (threadIdx.x < 2) counters[threadIdx.x] = 0;
  for (int64_t r = used + threadIdx.x; r < nrow; r += blockDim.x) rscale[r] = 0.0f;
}

// The weight-gradient GEMM over nch chunks of float32 rows (SC: per-row weights rscale).
template <bool SC>
void launch_dw_gemm(const RowBufs<PolicyF32>& R, float* partial, float* partial_db, int nch,
                    const float* rscale, hipStream_t s) {
  const int groups = (nch + kXcds - 1) / kXcds;
  const int nblk = dw_blocks_of(0) + dw_blocks_of(1) + dw_blocks_of(2);
  hipLaunchKernelGGL((mlp_dw_gemm_kernel<PolicyF32, SC>), dim3(kXcds * nblk * groups), dim3(64),
                     0, s, R, partial, partial_db, nch, rscale);
}

// One segment: the recompute + sweep kernel `kern` with the weights of policy P (float32:
// mlp_adj_segment_kernel on a float32 buffer; the three-piece bfloat16 form:
// mlp_adj_segment3_kernel on the three split3 planes of a bf16 buffer), then the float32
// weight-gradient GEMM over its float32 row buffers and the fixed-order reduction.
template <class P, class Kern>
int run_segment(Kern kern, int threads, const dfx_mlp_problem* p, const void* packed,
                const float* b1, const float* b2, const float* b3, const dfx_mlp_adj_io* io,
                int32_t seg, int32_t seg_steps, hipStream_t s) {
  AdjArgs a;
  memset(&a, 0, sizeof(a));
  a.grid = io->grid; a.ckpt = io->ckpt; a.stats = io->stats; a.ts = io->ts; a.ys_bar = io->ys_bar;
  a.ybar = io->ybar; a.zs = io->zs;
  a.B = p->batch; a.B_pad = padded_batch(p->batch);
  a.max_steps = p->max_steps; a.n_ckpt = dfx_mlp_n_ckpt(p); a.C = p->ckpt_every;
  a.Cs = seg_steps; a.seg = seg; a.n_ts = p->n_ts; a.ts_stride = p->ts_stride;
  a.n_out = (p->save_t0 ? 1 : 0) + p->n_ts + (p->save_t1 ? 1 : 0);
  a.save_t0 = p->save_t0; a.t0 = p->t0; a.t1 = p->t1;
  const MlpWeights<P> W = mlp_weights_from<P>(packed, P::PIECES, b1, b2, b3);
  RowBufs<PolicyF32> R;
  const int64_t rows_alloc = seg_chunks(p->batch, seg_steps) * ADJ_CHUNK;
  float* base = reinterpret_cast<float*>(io->rows);
  int64_t off = 0;
  const int widths[6] = {MLP_H, MLP_D, MLP_H, MLP_H, MLP_D, MLP_H};
  for (int i = 0; i < 6; ++i) { R.b[i] = base + off; off += rows_alloc * widths[i]; }
  const unsigned ntiles = (unsigned)(a.B_pad / MLP_TILE);
  (void)hipFuncSetAttribute(reinterpret_cast<const void*>(kern),
                            hipFuncAttributeMaxDynamicSharedMemorySize, (int)sizeof(MlpLds<P>));
  hipLaunchKernelGGL(kern, dim3(ntiles), dim3(threads), sizeof(MlpLds<P>), s, a, W, R);
  const int nch = (int)seg_chunks(p->batch, seg_steps);
  const DwLayout lay = dw_layout();
  float* partial = io->partial;
  float* partial_db = io->partial + (int64_t)nch * lay.first[3] * 1024;
  launch_dw_gemm<false>(R, partial, partial_db, nch, nullptr, s);
  const int nout = MLP_H * MLP_D + MLP_H * MLP_H + MLP_D * MLP_H + 2 * MLP_H + MLP_D;
  hipLaunchKernelGGL(mlp_dw_reduce_kernel, dim3((nout + 255) / 256), dim3(256), 0, s, partial,
                     partial_db, nch, io->dw);
  return hipGetLastError() == hipSuccess ? DFX_OK : DFX_E_HIP;
}

}  // namespace

extern "C" {

int dfx_mlp_adj_workspace(const dfx_mlp_problem* p, int32_t seg_steps, size_t* zs_bytes,
                          size_t* rows_bytes, size_t* partial_bytes) {
  if (!p || seg_steps <= 0 || !zs_bytes || !rows_bytes || !partial_bytes) return DFX_E_ARG;
  const int64_t Bp = padded_batch(p->batch);
  *zs_bytes = (size_t)(Bp * seg_steps * 7 * MLP_D) * sizeof(float);
  const int64_t rows_alloc = seg_chunks(p->batch, seg_steps) * ADJ_CHUNK;
  int64_t wsum = 0;
  const int widths[6] = {MLP_H, MLP_D, MLP_H, MLP_H, MLP_D, MLP_H};
  for (int i = 0; i < 6; ++i) wsum += widths[i];
  *rows_bytes = (size_t)(rows_alloc * wsum) * sizeof(float);   // float32 rows for both forms
  const DwLayout lay = dw_layout();
  *partial_bytes = (size_t)(seg_chunks(p->batch, seg_steps) * ((int64_t)lay.first[3] * 1024 + 2 * MLP_H + MLP_D)) * sizeof(float);
  return DFX_OK;
}

int dfx_mlp_backsolve_workspace(const dfx_mlp_problem* p, int32_t chunk_iters,
                                size_t* state_bytes, size_t* rows_bytes, size_t* partial_bytes) {
  if (!p || chunk_iters <= 0 || !state_bytes || !rows_bytes || !partial_bytes) return DFX_E_ARG;
  const int64_t Bp = padded_batch(p->batch);
  const int64_t nrow = seg_chunks(p->batch, chunk_iters) * ADJ_CHUNK;
  *state_bytes = bs_state_layout(Bp, nrow).total;
  int64_t wsum = 0;
  const int widths[6] = {MLP_H, MLP_D, MLP_H, MLP_H, MLP_D, MLP_H};
  for (int i = 0; i < 6; ++i) wsum += widths[i];
  *rows_bytes = (size_t)(nrow * wsum) * sizeof(float);
  const DwLayout lay = dw_layout();
  *partial_bytes = (size_t)(seg_chunks(p->batch, chunk_iters) *
                            ((int64_t)lay.first[3] * 1024 + 2 * MLP_H + MLP_D)) * sizeof(float);
  return DFX_OK;
}

int dfx_mlp_backsolve(const dfx_mlp_problem* back, const void* packed, const float* b1,
                      const float* b2, const float* b3, const dfx_mlp_backsolve_io* io,
                      int32_t chunk_iters, int32_t first, void* stream) {
  if (!back || !packed || !b1 || !b2 || !b3 || !io) return DFX_E_ARG;
  if (back->d != MLP_D || back->hidden != MLP_H || back->solver != DFX_TSIT5) return DFX_E_ARG;
  if (back->precision != DFX_F32) return DFX_E_ARG;  // float32 products on a float32 packing
  if (chunk_iters <= 0 || isnan(back->dt0) || back->max_steps <= 0) return DFX_E_ARG;
  const int32_t n_out = (back->save_t0 ? 1 : 0) + back->n_ts + (back->save_t1 ? 1 : 0);
  if (!io->ys || !io->ts_out || !io->ys_bar || !io->state || !io->rows || !io->partial ||
      !io->dw || !io->ybar || !io->result)
    return DFX_E_ARG;
  if (back->batch == 0) return DFX_OK;
  BsArgs a;
  memset(&a, 0, sizeof(a));
  if (!mlp_pid_params(back, a.pid)) return DFX_E_ARG;
  const int64_t Bp = padded_batch(back->batch);
  const int nch = (int)seg_chunks(back->batch, chunk_iters);
  const BsStateLayout sl = bs_state_layout(Bp, (int64_t)nch * ADJ_CHUNK);
  unsigned char* sb = reinterpret_cast<unsigned char*>(io->state);
  a.ys = io->ys; a.ts_out = io->ts_out; a.ys_bar = io->ys_bar;
  a.ya = reinterpret_cast<float*>(sb + sl.ya);
  a.rows = reinterpret_cast<BsRow*>(sb + sl.rows);
  a.rscale = reinterpret_cast<float*>(sb + sl.rscale);
  a.counters = reinterpret_cast<int32_t*>(sb + sl.counters);
  a.ybar = io->ybar; a.result = io->result;
  a.B = back->batch; a.B_pad = Bp;
  a.n_out = n_out; a.save_t0 = back->save_t0; a.max_steps = back->max_steps;
  a.Q = chunk_iters; a.first = first ? 1 : 0;
  a.t0 = back->t0; a.dt0 = back->dt0;
  // The products run in float32 MFMA for both matmul forms (the caller packs the weights
  // for DFX_F32): the bfloat16 form's forward evaluates the float32-weight model too.
  using P = PolicyF32;
  const MlpWeights<P> W = mlp_weights_from<P>(packed, 1, b1, b2, b3);
  RowBufs<P> R;
  float* base = reinterpret_cast<float*>(io->rows);
  int64_t off = 0;
  const int widths[6] = {MLP_H, MLP_D, MLP_H, MLP_H, MLP_D, MLP_H};
  for (int i = 0; i < 6; ++i) { R.b[i] = base + off; off += (int64_t)nch * ADJ_CHUNK * widths[i]; }
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  hipLaunchKernelGGL(bs_reset_kernel, dim3(1), dim3(64), 0, s, a.counters, a.rscale,
                     (int64_t)chunk_iters * 7 * Bp, (int64_t)nch * ADJ_CHUNK);
  const unsigned ntiles = (unsigned)(Bp / MLP_TILE);
  // four waves per tile: at eight the kernel spills 193 VGPRs per lane (at four, 76) and
  // took 34.9 ms against 19.6 ms for the backsolve of 8,192 trajectories (bit-identical
  // gradients)
  constexpr int NWB = 4;
  (void)hipFuncSetAttribute(reinterpret_cast<const void*>(&mlp_backsolve_kernel<P, NWB>),
                            hipFuncAttributeMaxDynamicSharedMemorySize, (int)sizeof(MlpLds<P>));
  hipLaunchKernelGGL((mlp_backsolve_kernel<P, NWB>), dim3(ntiles), dim3(64 * NWB),
                     sizeof(MlpLds<P>), s, a, W, R);
  const DwLayout lay = dw_layout();
  float* partial = io->partial;
  float* partial_db = io->partial + (int64_t)nch * lay.first[3] * 1024;
  launch_dw_gemm<true>(R, partial, partial_db, nch, a.rscale, s);
  const int nout = MLP_H * MLP_D + MLP_H * MLP_H + MLP_D * MLP_H + 2 * MLP_H + MLP_D;
  hipLaunchKernelGGL(mlp_dw_reduce_kernel, dim3((nout + 255) / 256), dim3(256), 0, s, partial,
                     partial_db, nch, io->dw);
  return hipGetLastError() == hipSuccess ? DFX_OK : DFX_E_HIP;
}

int dfx_mlp_backsolve_status(const dfx_mlp_problem* back, int32_t chunk_iters,
                             const void* state, const int32_t** counters) {
  if (!back || chunk_iters <= 0 || !state || !counters) return DFX_E_ARG;
  const int64_t Bp = padded_batch(back->batch);
  const BsStateLayout sl = bs_state_layout(Bp, seg_chunks(back->batch, chunk_iters) * ADJ_CHUNK);
  *counters = reinterpret_cast<const int32_t*>(reinterpret_cast<const unsigned char*>(state) +
                                               sl.counters);
  return DFX_OK;
}

int dfx_mlp_adj_segment(const dfx_mlp_problem* p, const void* packed, const float* b1,
                        const float* b2, const float* b3, const dfx_mlp_adj_io* io,
                        int32_t segment, int32_t seg_steps, void* stream) {
  if (!p || !packed || !b1 || !b2 || !b3 || !io) return DFX_E_ARG;
  if (p->d != MLP_D || p->hidden != MLP_H || p->solver != DFX_TSIT5) return DFX_E_ARG;
  if (p->precision != DFX_F32 && p->precision != DFX_BF16_F32) return DFX_E_ARG;
  if (p->ckpt_every <= 0 || seg_steps <= 0 || seg_steps > p->ckpt_every || segment < 0 ||
      segment >= dfx_mlp_n_ckpt(p))
    return DFX_E_ARG;
  if (!io->grid || !io->ckpt || !io->stats || !io->ys_bar || !io->ybar || !io->zs || !io->rows ||
      !io->partial || !io->dw)
    return DFX_E_ARG;
  if (p->n_ts > 0 && !io->ts) return DFX_E_ARG;
  if (p->ts_stride != 0 && p->ts_stride != p->n_ts) return DFX_E_ARG;
  if (p->batch == 0) return DFX_OK;
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  // eight waves per tile, two per SIMD: at four the kernels held 415 (float32) and 512 + 93
  // spilled (three-piece) registers per lane, one wave per SIMD.  The backward of 8,192
  // trajectories went from 3.11 to 2.86 ms (float32) and from 2.51-2.55 to 2.16-2.17 ms
  // (three-piece), with bit-identical gradients.
  if (p->precision == DFX_BF16_F32)
    return run_segment<PolicyBF16x3>(&mlp_adj_segment3_kernel<8>, 64 * 8, p, packed, b1, b2, b3,
                                      io, segment, seg_steps, s);
  if (p->precision == DFX_F32)
    return run_segment<PolicyF32>(&mlp_adj_segment_kernel<PolicyF32, 8>, 64 * 8, p, packed, b1,
                                  b2, b3, io, segment, seg_steps, s);
  return DFX_E_ARG;
}

}  // extern "C"
