// erk_dense_f64.hip -- double instantiations of the dense-output evaluation kernel
// (no FMA contraction (bit parity with the solve kernel's saves)).
#include "erk_dense.hpp"

namespace dfx {

int erk_dense_eval_f64(int solver, int d, const DenseKArgs& a, hipStream_t s) {
  return dense_dispatch<double>(solver, d, a, s);
}

}  // namespace dfx
