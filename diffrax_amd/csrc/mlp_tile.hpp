// mlp_tile.hpp -- the registered neural vector field f(y) = tanh(W3 sp(W2 sp(W1 y + b1) + b2) + b3)
// (sp = softplus; d = 64 -> 256 -> 256 -> 64) evaluated for a tile of 32 trajectories by one
// 256-thread workgroup on MFMA, plus its vector-Jacobian product.
//
// Layout of one tile (4 waves):
//  - owner threads: thread t owns trajectory t/8 of the tile and state components
//    (t%8)*8 .. +8, so a trajectory's 8 owners are adjacent lanes of one wave;
//  - activations live in LDS, row-major [32 rows][K + PAD]; every MFMA reads its A fragment
//    (32 rows x KSTEP) straight from there;
//  - weights are read as MFMA B fragments from a host-packed copy (fragment (n_tile, k_step)
//    is 64 lanes x KA contiguous elements, one coalesced load per lane);
//  - layers 1 and 2: wave w produces output columns [64w, 64w + 64) (two 32x32 tiles);
//    layer 3 (64 outputs, K = 256) is split over K: wave w reduces K in [64w, 64w + 64) and
//    the four partial sums meet in LDS, where the owner threads add them in wave order.
// Precision policies: BF16 operands with f32 accumulate (v_mfma_f32_32x32x16_bf16), or
// f32 throughout (v_mfma_f32_32x32x2f32).  Activation functions are evaluated in f32.
#pragma once
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include "../../include/diffrax_amd.h"
#include "erk_common.hpp"

namespace dfx {

typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef float floatx16_t __attribute__((ext_vector_type(16)));

constexpr int MLP_D = 64;
constexpr int MLP_H = 256;
constexpr int MLP_TILE = 32;
constexpr int MLP_THREADS = 256;

struct PolicyBF16 {
  using act_t = __bf16;
  using frag_t = bf16x8_t;
  static constexpr int KSTEP = 16;  // K per MFMA
  static constexpr int KA = 8;      // elements per lane per fragment
  static constexpr int PAD = 8;     // LDS row padding (elements); keeps 16-byte alignment
  static constexpr bool fast_math = true;
  // Every layer input (the RK stage value and both hidden activations) is fed as hi + lo
  // bfloat16 terms, ~16 significant bits.  A single bfloat16 (8 bits) turns f into a step
  // function at the scale of the O(dt f) differences between stages, and the embedded error
  // estimate then measures rounding noise instead of truncation error.  Each layer costs two
  // bfloat16 MFMA passes (float32 accumulate); the weights themselves are bfloat16.
  static constexpr bool split_input = true;
  static constexpr int PIECES = 1;  // weight planes read (plane 0 of a three-piece buffer)
  __device__ __forceinline__ static frag_t load_a(const act_t* p) {
    return *reinterpret_cast<const frag_t*>(p);
  }
  __device__ __forceinline__ static frag_t load_b(const act_t* packed, int frag) {
    return reinterpret_cast<const frag_t*>(packed)[frag];
  }
  __device__ __forceinline__ static floatx16_t mfma(frag_t a, frag_t b, floatx16_t c) {
    return __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, c, 0, 0, 0);
  }
  __device__ __forceinline__ static act_t to_act(float x) { return (act_t)x; }
};

struct PolicyF32 {
  using act_t = float;
  using frag_t = float;
  static constexpr int KSTEP = 2;
  static constexpr int KA = 1;
  static constexpr int PAD = 1;     // odd row stride: conflict-free column reads
  // Softplus / tanh through the hardware exp2, log2 and reciprocal (about 1 ulp each): the
  // vector field stays within 8.4e-7 of a float64 evaluation (the libm forms gave 8e-7)
  // and the forward runs 2x faster, its VALU work having been the larger part of a step.
  static constexpr bool fast_math = true;
  static constexpr bool split_input = false;
  static constexpr int PIECES = 1;
  __device__ __forceinline__ static frag_t load_a(const act_t* p) { return *p; }
  __device__ __forceinline__ static frag_t load_b(const act_t* packed, int frag) {
    return packed[frag];
  }
  __device__ __forceinline__ static floatx16_t mfma(frag_t a, frag_t b, floatx16_t c) {
    return __builtin_amdgcn_mfma_f32_32x32x2f32(a, b, c, 0, 0, 0);
  }
  __device__ __forceinline__ static act_t to_act(float x) { return x; }
};

// BF16 MFMA (float32 accumulate) with every operand carried as three bfloat16 pieces,
// v = p0 + p1 + p2 exactly for a float32 v (split3).  The six products p_i q_j with
// i + j <= 2 give each layer sum to ~2^-24 relative, float32 MFMA accuracy, from six bf16
// MFMA passes: the float32-weight model itself is evaluated (the weights are split, not
// rounded), and the products of the leading pieces accumulate apart from the five
// correction products.  bf16 MFMA has 16x the float32 MFMA rate on gfx950.
struct PolicyBF16x3 : PolicyBF16 {
  static constexpr bool split_input = false;
  static constexpr int PIECES = 3;
};

// v = p0 + p1 + p2 with bfloat16 pieces: each residual is exact in float32 and the last
// one has at most 8 significant bits.
__device__ __forceinline__ void split3(float v, __bf16& p0, __bf16& p1, __bf16& p2) {
  p0 = (__bf16)v;
  const float r1 = v - (float)p0;
  p1 = (__bf16)r1;
  p2 = (__bf16)(r1 - (float)p1);
}

// Packed weight buffer: six matrices (w1, w2, w3 forward; w1t, w2t, w3t transposed), each
// `pieces` planes of its fragment array (bf16: the three split3 pieces, plane 0 the
// rounded weight; float32: one plane).  Element offsets of plane 0 and plane sizes.
struct MlpPackLayout {
  int64_t off[6], size[6], total;
};
__host__ __device__ inline MlpPackLayout mlp_pack_layout(int pieces) {
  const int64_t D = MLP_D, H = MLP_H;
  const int64_t sz[6] = {H * D, H * H, D * H, D * H, H * H, H * D};
  MlpPackLayout l;
  int64_t o = 0;
  for (int i = 0; i < 6; ++i) {
    l.off[i] = o;
    l.size[i] = sz[i];
    o += sz[i] * pieces;
  }
  l.total = o;
  return l;
}

// Packed weights of one solve.  Forward fragments: B[k][n] = W[n][k].  Backward fragments
// (transposed products): B[k][n] = W[k][n].
template <class P>
struct MlpWeights {
  const typename P::act_t* w1;   // layer 1 forward fragments (N = 256, K = 64)
  const typename P::act_t* w2;   // layer 2 forward (N = 256, K = 256)
  const typename P::act_t* w3;   // layer 3 forward (N = 64, K = 256)
  const typename P::act_t* w1t;  // z-bar = d1 . W1   (N = 64, K = 256)
  const typename P::act_t* w2t;  // h1-bar = d2 . W2  (N = 256, K = 256)
  const typename P::act_t* w3t;  // h2-bar = d3 . W3  (N = 256, K = 64)
  const float* b1;
  const float* b2;
  const float* b3;
  int64_t plane1, plane2, plane3;  // elements between the pieces of w1, w2, w3 (PIECES > 1)
};

// The weights of a packed buffer (mlp_pack_layout) for a policy; a one-piece policy on a
// three-piece buffer reads plane 0.
template <class P>
__host__ __device__ inline MlpWeights<P> mlp_weights_from(const void* packed, int pieces,
                                                          const float* b1, const float* b2,
                                                          const float* b3) {
  const MlpPackLayout l = mlp_pack_layout(pieces);
  const typename P::act_t* base = reinterpret_cast<const typename P::act_t*>(packed);
  MlpWeights<P> W;
  W.w1 = base + l.off[0]; W.w2 = base + l.off[1]; W.w3 = base + l.off[2];
  W.w1t = base + l.off[3]; W.w2t = base + l.off[4]; W.w3t = base + l.off[5];
  W.b1 = b1; W.b2 = b2; W.b3 = b3;
  W.plane1 = l.size[0]; W.plane2 = l.size[1]; W.plane3 = l.size[2];
  return W;
}

template <class P>
struct MlpLds {
  static constexpr int LDX = MLP_D + P::PAD;
  static constexpr int LDH = MLP_H + P::PAD;
  static constexpr int LDP = MLP_D + 4;
  typename P::act_t x[MLP_TILE * LDX];    // layer-1 input (stage value), high part
  typename P::act_t xlo[P::split_input ? MLP_TILE * LDX : 1];  // layer-1 input, low part
  typename P::act_t h1[MLP_TILE * LDH];   // softplus(layer 1) (high part)
  typename P::act_t h2[MLP_TILE * LDH];   // softplus(layer 2) (high part)
  typename P::act_t h1lo[P::split_input ? MLP_TILE * LDH : 1];  // low parts
  typename P::act_t h2lo[P::split_input ? MLP_TILE * LDH : 1];
  typename P::act_t d[MLP_TILE * LDH];    // backward: delta of layer 2
  typename P::act_t d3[MLP_TILE * LDX];   // backward: delta of layer 3
  float part[4 * MLP_TILE * LDP];         // split-K partial sums of the 64-wide outputs
};

// Store one float value of the layer-1 input (row r, column c) in its policy's form.
template <class P>
__device__ __forceinline__ void store_input(MlpLds<P>& L, int r, int c, float v) {
  const typename P::act_t hi = P::to_act(v);
  L.x[r * L.LDX + c] = hi;
  if constexpr (P::split_input) L.xlo[r * L.LDX + c] = P::to_act(v - (float)hi);
}

// Store a hidden activation (hi, and lo for split policies).
template <class P>
__device__ __forceinline__ void store_hidden(typename P::act_t* hi_buf, typename P::act_t* lo_buf,
                                             int idx, float v) {
  const typename P::act_t hi = P::to_act(v);
  hi_buf[idx] = hi;
  if constexpr (P::split_input) lo_buf[idx] = P::to_act(v - (float)hi);
}

// ---- activations ------------------------------------------------------------------------
template <bool Fast>
__device__ __forceinline__ float softplus_f(float x) {
  // jax.nn.softplus(x) = logaddexp(x, 0) = max(x, 0) + log1p(exp(-|x|))
  if constexpr (Fast) {
    const float e = __builtin_amdgcn_exp2f(-fabsf(x) * 1.4426950408889634f);
    return fmaxf(x, 0.0f) + 0.6931471805599453f * __builtin_amdgcn_logf(1.0f + e);
  } else {
    return fmaxf(x, 0.0f) + log1pf(expf(-fabsf(x)));
  }
}
template <bool Fast>
__device__ __forceinline__ float tanh_f(float x) {
  if constexpr (Fast) {
    // tanh(x) = sign(x) (1 - 2 / (exp(2|x|) + 1))
    const float e = __builtin_amdgcn_exp2f(fminf(2.0f * fabsf(x), 60.0f) * 1.4426950408889634f);
    const float t = 1.0f - 2.0f * __builtin_amdgcn_rcpf(e + 1.0f);
    return copysignf(t, x);
  } else {
    return tanhf(x);
  }
}
// derivative of softplus from its value h = softplus(s): sigmoid(s) = 1 - exp(-h)
template <bool Fast>
__device__ __forceinline__ float dsoftplus_from_value(float h) {
  if constexpr (Fast) return 1.0f - __builtin_amdgcn_exp2f(-h * 1.4426950408889634f);
  else return -expm1f(-h);
}

// ---- one wave's share of a layer --------------------------------------------------------
// acc[nt] (32x32 f32) = sum over k-steps [ks0, ks1) of A(lds rows, k) * B(fragment (nt0+nt, ks)),
// where A = a_lds, or a_lds + a_lo (LO: the hi + lo halves of a split operand, two MFMAs per
// B fragment).  16-byte fragments (bf16) are loaded PF k-steps ahead of their MFMAs through a
// ring of registers (PF = 4; 8 for layer 2: 7 % faster forward, 12 % faster adjoint than one
// step ahead); 4-byte fragments (f32, 2 k per MFMA, PF = 0) are left to the compiler's
// unrolled schedule, where explicit prefetch measured slower.
template <class P, int NKS, int NT, bool LO = false, int PF = (P::KA == 1 ? 0 : 4)>
__device__ __forceinline__ void wave_gemm(const typename P::act_t* a_lds,
                                          const typename P::act_t* a_lo, int lda,
                                          const typename P::act_t* wp, int nt0, int ks0,
                                          int ks1, floatx16_t (&acc)[NT], int lane) {
  using frag_t = typename P::frag_t;
  const int row = lane & 31;
  const int kh = (lane >> 5) * P::KA;
#pragma unroll
  for (int t = 0; t < NT; ++t) acc[t] = floatx16_t{};
  if constexpr (P::KA == 1 && PF == 0) {
    static_assert(!LO, "split operands are bfloat16");
#pragma unroll 8
    for (int ks = ks0; ks < ks1; ++ks) {
      const frag_t a = P::load_a(a_lds + row * lda + ks * P::KSTEP + kh);
#pragma unroll
      for (int t = 0; t < NT; ++t) {
        const frag_t b = P::load_b(wp, ((nt0 + t) * NKS + ks) * 64 + lane);
        acc[t] = P::mfma(a, b, acc[t]);
      }
    }
  } else {
    // B fragments come from L2 (the weights of every layer, shared by all workgroups), whose
    // latency is several k-steps of MFMA work: PF k-steps are kept in flight.
    frag_t buf[PF][NT];
#pragma unroll
    for (int p = 0; p < PF; ++p) {
      const int kp = (ks0 + p < ks1) ? ks0 + p : ks1 - 1;
#pragma unroll
      for (int t = 0; t < NT; ++t) buf[p][t] = P::load_b(wp, ((nt0 + t) * NKS + kp) * 64 + lane);
    }
    for (int ks = ks0; ks < ks1; ks += PF) {
#pragma unroll
      for (int p = 0; p < PF; ++p) {
        const int k = ks + p;
        if (k < ks1) {
          frag_t b[NT];
          const int kn = (k + PF < ks1) ? k + PF : ks1 - 1;  // the tail re-reads the last one
#pragma unroll
          for (int t = 0; t < NT; ++t) {
            b[t] = buf[p][t];
            buf[p][t] = P::load_b(wp, ((nt0 + t) * NKS + kn) * 64 + lane);
          }
          const frag_t a = P::load_a(a_lds + row * lda + k * P::KSTEP + kh);
#pragma unroll
          for (int t = 0; t < NT; ++t) acc[t] = P::mfma(a, b[t], acc[t]);
          if constexpr (LO) {
            const frag_t al = P::load_a(a_lo + row * lda + k * P::KSTEP + kh);
#pragma unroll
            for (int t = 0; t < NT; ++t) acc[t] = P::mfma(al, b[t], acc[t]);
          }
        }
      }
    }
  }
}

// C-layout helpers: accumulator element v (= 4i + j) of lane l sits at
// row 8i + (l/32)*4 + j, column l%32 of the 32x32 tile.
__device__ __forceinline__ int acc_row(int lane, int v) { return 8 * (v >> 2) + (lane >> 5) * 4 + (v & 3); }

// ---- the three-piece bf16 form --------------------------------------------------------
// LDS: the three pieces of the layer-1 input and of both hidden activations (115 KB).  The
// split-K partials of layer 3 reuse the h1 planes, which nobody reads after layer 2.
template <>
struct MlpLds<PolicyBF16x3> {
  static constexpr int LDX = MLP_D + PolicyBF16x3::PAD;
  static constexpr int LDH = MLP_H + PolicyBF16x3::PAD;
  static constexpr int LDP = MLP_D + 4;
  __bf16 x[3][MLP_TILE * LDX];
  __bf16 h1[3][MLP_TILE * LDH];
  __bf16 h2[3][MLP_TILE * LDH];
  __device__ __forceinline__ float* part() { return reinterpret_cast<float*>(&h1[0][0]); }
};
static_assert(4 * MLP_TILE * MlpLds<PolicyBF16x3>::LDP * sizeof(float) <=
                  sizeof(MlpLds<PolicyBF16x3>::h1),
              "layer-3 partials fit in the h1 planes");

template <>
__device__ __forceinline__ void store_input<PolicyBF16x3>(MlpLds<PolicyBF16x3>& L, int r, int c,
                                                          float v) {
  const int i = r * L.LDX + c;
  split3(v, L.x[0][i], L.x[1][i], L.x[2][i]);
}

// acc[nt] (32x32 f32) = sum over k-steps [ks0, ks1) of A . B with A and B three-piece
// operands: A from LDS planes a0..a2, B fragment (nt0 + nt, ks) from planes wp + p * plane.
// The leading products a0 b0 accumulate in `big`, the five corrections (a0 b1, a1 b0, a0 b2,
// a1 b1, a2 b0) in `sml`, added at the end.  B fragments are loaded PF k-steps ahead.
template <int NKS, int NT, int PF = 1>
__device__ __forceinline__ void wave_gemm3(const __bf16* a0, const __bf16* a1, const __bf16* a2,
                                           int lda, const __bf16* wp, int64_t plane, int nt0,
                                           int ks0, int ks1, floatx16_t (&acc)[NT], int lane) {
  using P = PolicyBF16x3;
  using frag_t = P::frag_t;
  const int row = lane & 31;
  const int kh = (lane >> 5) * P::KA;
  floatx16_t big[NT], sml[NT];
#pragma unroll
  for (int t = 0; t < NT; ++t) { big[t] = floatx16_t{}; sml[t] = floatx16_t{}; }
  auto bfrag = [&](int pc, int t, int k) {
    return P::load_b(wp + pc * plane, ((nt0 + t) * NKS + k) * 64 + lane);
  };
  frag_t buf[PF][3][NT];
#pragma unroll
  for (int p = 0; p < PF; ++p) {
    const int kp = (ks0 + p < ks1) ? ks0 + p : ks1 - 1;
#pragma unroll
    for (int pc = 0; pc < 3; ++pc) {
#pragma unroll
      for (int t = 0; t < NT; ++t) buf[p][pc][t] = bfrag(pc, t, kp);
    }
  }
  // one PF-step group per iteration: a fully unrolled K loop (seven inlined evaluations per
  // step) lets the scheduler hoist loads until the registers spill
#pragma unroll 1
  for (int ks = ks0; ks < ks1; ks += PF) {
#pragma unroll
    for (int p = 0; p < PF; ++p) {
      const int k = ks + p;
      if (k < ks1) {
        frag_t b[3][NT];
#pragma unroll
        for (int pc = 0; pc < 3; ++pc) {
#pragma unroll
          for (int t = 0; t < NT; ++t) b[pc][t] = buf[p][pc][t];
        }
        if (k + PF < ks1) {
#pragma unroll
          for (int pc = 0; pc < 3; ++pc) {
#pragma unroll
            for (int t = 0; t < NT; ++t) buf[p][pc][t] = bfrag(pc, t, k + PF);
          }
        }
        const int ao = row * lda + k * P::KSTEP + kh;
        const frag_t x0 = P::load_a(a0 + ao), x1 = P::load_a(a1 + ao), x2 = P::load_a(a2 + ao);
#pragma unroll
        for (int t = 0; t < NT; ++t) big[t] = P::mfma(x0, b[0][t], big[t]);
#pragma unroll
        for (int t = 0; t < NT; ++t) {
          sml[t] = P::mfma(x0, b[1][t], sml[t]);
          sml[t] = P::mfma(x1, b[0][t], sml[t]);
          sml[t] = P::mfma(x0, b[2][t], sml[t]);
          sml[t] = P::mfma(x1, b[1][t], sml[t]);
          sml[t] = P::mfma(x2, b[0][t], sml[t]);
        }
      }
    }
  }
#pragma unroll
  for (int t = 0; t < NT; ++t) acc[t] = big[t] + sml[t];
}

// Float32 copies of the hidden activations for the adjoint's row buffers ([rows][256] each,
// the tile's rows from row0, and again from row0b when row0b >= 0): written in the layer
// epilogues, before layer 3's partials reuse the h1 planes.
struct MlpRowSink {
  float* h1 = nullptr;
  float* h2 = nullptr;
  int64_t row0 = 0, row0b = -1;
  __device__ __forceinline__ void put(float* dst, int r, int col, float v) const {
    dst[(row0 + r) * MLP_H + col] = v;
    if (row0b >= 0) dst[(row0b + r) * MLP_H + col] = v;
  }
};

// The tile's waves split the work: layers 1 and 2 by output columns (NT12 tiles of 32 per
// wave), layer 3 by K quarters (four waves: both 32-column tiles of one quarter; eight
// waves: one tile of one quarter), into the same four partial slabs.  The owners of a row
// hold 32 / NW of its components.
template <int NW>
struct MlpWaves {
  static_assert(NW == 4 || NW == 8, "four or eight waves per tile");
  static constexpr int NT12 = 8 / NW;
  static constexpr int NT3 = NW == 4 ? 2 : 1;
  static constexpr int CPT = 32 / NW;
  __device__ static int nt3(int wave) { return NW == 4 ? 0 : (wave & 1); }
  __device__ static int quarter(int wave) { return NW == 4 ? wave : (wave >> 1); }
};

template <int NW = 4>
__device__ __forceinline__ void mlp_forward3(const MlpWeights<PolicyBF16x3>& W,
                                             MlpLds<PolicyBF16x3>& L, int wave, int lane,
                                             int owner_row, int owner_c0,
                                             float (&fout)[MlpWaves<NW>::CPT],
                                             const MlpRowSink& sink) {
  using WV = MlpWaves<NW>;
  constexpr int NT = WV::NT12;
  constexpr int KS1 = MLP_D / 16, KS2 = MLP_H / 16;
  floatx16_t acc[NT];
  // layer 1: [32 x 64] . W1^T -> [32 x 256], wave's columns [64w, 64w + 64)
  wave_gemm3<KS1, NT>(L.x[0], L.x[1], L.x[2], L.LDX, W.w1, W.plane1, NT * wave, 0, KS1, acc, lane);
#pragma unroll
  for (int t = 0; t < NT; ++t) {
    const int col = 32 * (NT * wave + t) + (lane & 31);
    const float bias = W.b1[col];
#pragma unroll
    for (int v = 0; v < 16; ++v) {
      const int i = acc_row(lane, v) * L.LDH + col;
      const float h = softplus_f<true>(acc[t][v] + bias);
      if (sink.h1) sink.put(sink.h1, acc_row(lane, v), col, h);
      split3(h, L.h1[0][i], L.h1[1][i], L.h1[2][i]);
    }
  }
  __syncthreads();
  // layer 2: [32 x 256] . W2^T -> [32 x 256]
  wave_gemm3<KS2, NT>(L.h1[0], L.h1[1], L.h1[2], L.LDH, W.w2, W.plane2, NT * wave, 0, KS2, acc,
                      lane);
#pragma unroll
  for (int t = 0; t < NT; ++t) {
    const int col = 32 * (NT * wave + t) + (lane & 31);
    const float bias = W.b2[col];
#pragma unroll
    for (int v = 0; v < 16; ++v) {
      const int i = acc_row(lane, v) * L.LDH + col;
      const float h = softplus_f<true>(acc[t][v] + bias);
      if (sink.h2) sink.put(sink.h2, acc_row(lane, v), col, h);
      split3(h, L.h2[0][i], L.h2[1][i], L.h2[2][i]);
    }
  }
  __syncthreads();
  // layer 3 (split-K): wave w sums K in [64w, 64w + 64) for all 64 outputs; the partials go
  // to the h1 planes (layer 2, their last reader, finished before the barrier above)
  constexpr int NT3 = WV::NT3;
  const int qk = WV::quarter(wave), n3 = WV::nt3(wave);
  floatx16_t acc3[NT3];
  wave_gemm3<KS2, NT3>(L.h2[0], L.h2[1], L.h2[2], L.LDH, W.w3, W.plane3, n3, qk * (KS2 / 4),
                       (qk + 1) * (KS2 / 4), acc3, lane);
  float* part = L.part();
#pragma unroll
  for (int t = 0; t < NT3; ++t) {
    const int col = 32 * (n3 + t) + (lane & 31);
#pragma unroll
    for (int v = 0; v < 16; ++v) part[(qk * MLP_TILE + acc_row(lane, v)) * L.LDP + col] = acc3[t][v];
  }
  __syncthreads();
#pragma unroll
  for (int j = 0; j < WV::CPT; ++j) {
    const int c = owner_c0 + j;
    float s = part[(0 * MLP_TILE + owner_row) * L.LDP + c];
    s += part[(1 * MLP_TILE + owner_row) * L.LDP + c];
    s += part[(2 * MLP_TILE + owner_row) * L.LDP + c];
    s += part[(3 * MLP_TILE + owner_row) * L.LDP + c];
    fout[j] = tanh_f<true>(s + W.b3[c]);
  }
}

// Forward of the MLP on the tile: reads L.x (stage values), leaves softplus activations in
// L.h1 / L.h2 and returns f for the owner's 32 / NW components in fout (NW waves per tile;
// MlpWaves).  The three-piece bf16 form runs mlp_forward3.
template <class P, int NW = 4>
__device__ __forceinline__ void mlp_forward(const MlpWeights<P>& W, MlpLds<P>& L, int wave,
                                            int lane, int owner_row, int owner_c0,
                                            float (&fout)[MlpWaves<NW>::CPT]) {
  if constexpr (P::PIECES == 3) {
    mlp_forward3<NW>(W, L, wave, lane, owner_row, owner_c0, fout, MlpRowSink{});
  } else {
  using WV = MlpWaves<NW>;
  constexpr int NT = WV::NT12;
  constexpr bool F = P::fast_math;
  floatx16_t acc[NT];
  // layer 1: [32 x 64] . W1^T -> [32 x 256], wave's columns [64w, 64w + 64)
  constexpr bool S = P::split_input;
  wave_gemm<P, MLP_D / P::KSTEP, NT, S>(L.x, L.xlo, L.LDX, W.w1, NT * wave, 0, MLP_D / P::KSTEP,
                                        acc, lane);
#pragma unroll
  for (int t = 0; t < NT; ++t) {
    const int col = 32 * (NT * wave + t) + (lane & 31);
    const float bias = W.b1[col];
#pragma unroll
    for (int v = 0; v < 16; ++v)
      store_hidden<P>(L.h1, L.h1lo, acc_row(lane, v) * L.LDH + col, softplus_f<F>(acc[t][v] + bias));
  }
  __syncthreads();
  // layer 2: [32 x 256] . W2^T -> [32 x 256]
  wave_gemm<P, MLP_H / P::KSTEP, NT, S, (P::KA == 1 ? 0 : 8)>(L.h1, L.h1lo, L.LDH, W.w2,
                                                              NT * wave, 0, MLP_H / P::KSTEP,
                                                              acc, lane);
#pragma unroll
  for (int t = 0; t < NT; ++t) {
    const int col = 32 * (NT * wave + t) + (lane & 31);
    const float bias = W.b2[col];
#pragma unroll
    for (int v = 0; v < 16; ++v)
      store_hidden<P>(L.h2, L.h2lo, acc_row(lane, v) * L.LDH + col, softplus_f<F>(acc[t][v] + bias));
  }
  __syncthreads();
  // layer 3 (split-K): wave w sums K in [64w, 64w + 64) for all 64 outputs
  constexpr int KS3 = MLP_H / P::KSTEP;
  constexpr int NT3 = WV::NT3;
  const int qk = WV::quarter(wave), n3 = WV::nt3(wave);
  floatx16_t acc3[NT3];
  wave_gemm<P, KS3, NT3, S>(L.h2, L.h2lo, L.LDH, W.w3, n3, qk * (KS3 / 4), (qk + 1) * (KS3 / 4),
                            acc3, lane);
#pragma unroll
  for (int t = 0; t < NT3; ++t) {
    const int col = 32 * (n3 + t) + (lane & 31);
#pragma unroll
    for (int v = 0; v < 16; ++v) L.part[(qk * MLP_TILE + acc_row(lane, v)) * L.LDP + col] = acc3[t][v];
  }
  __syncthreads();
#pragma unroll
  for (int j = 0; j < WV::CPT; ++j) {
    const int c = owner_c0 + j;
    float s = L.part[(0 * MLP_TILE + owner_row) * L.LDP + c];
    s += L.part[(1 * MLP_TILE + owner_row) * L.LDP + c];
    s += L.part[(2 * MLP_TILE + owner_row) * L.LDP + c];
    s += L.part[(3 * MLP_TILE + owner_row) * L.LDP + c];
    fout[j] = tanh_f<F>(s + W.b3[c]);
  }
  }
}

// The PID parameters of a neural-ODE problem (PIDController.__init__, pid.py:299-311, with
// its coefficients divided by the error order); false unless the controller is the integral
// one the neural-ODE kernels implement.
inline bool mlp_pid_params(const dfx_mlp_problem* p, PidParams& q) {
  const double error_order = isnan(p->error_order) ? 5.0 : p->error_order;
  q.rtol = p->rtol; q.atol = p->atol;
  q.coeff1 = (p->icoeff + p->pcoeff + p->dcoeff) / error_order;
  q.coeff2 = -(p->pcoeff + 2.0 * p->dcoeff) / error_order;
  q.coeff3 = p->dcoeff / error_order;
  q.dtmin = p->dtmin; q.dtmax = p->dtmax;
  q.has_dtmin = !isnan(p->dtmin); q.has_dtmax = !isnan(p->dtmax);
  q.force_dtmin = p->force_dtmin; q.factormin = p->factormin; q.factormax = p->factormax;
  q.safety = p->safety;
  q.pad_ = 0;
  return pid_is_integral(q);
}

}  // namespace dfx
