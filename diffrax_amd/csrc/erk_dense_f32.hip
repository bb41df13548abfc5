// erk_dense_f32.hip -- float instantiations of the dense-output evaluation kernel
// (FMA contraction, as the float32-state solve kernel).
#include "erk_dense.hpp"

namespace dfx {

int erk_dense_eval_f32(int solver, int d, const DenseKArgs& a, hipStream_t s) {
  return dense_dispatch<float>(solver, d, a, s);
}

}  // namespace dfx
