// FP32 instantiation of the engine (the mixed path: FP32 fields and
// transforms, FP64 reductions and host scalars).
#include "mlq_ops_core.hpp"
#include "mlq_rcg.hpp"
#include "mlq_fast.hpp"
#include "mlq_diag.hpp"

namespace mlq {

// the FP64 twin of the mixed line search is the FP64 engine, instantiated in
// its own unit
extern template class Engine<double>;

template class Engine<float>;

std::unique_ptr<IEngine> make_engine_fp32(int n, const mlq_params& p) {
  return std::unique_ptr<IEngine>(new Engine<float>(n, p));
}

}  // namespace mlq
