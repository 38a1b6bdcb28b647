// FP64 instantiation of the engine (the parity path).
#include "mlq_ops_core.hpp"
#include "mlq_rcg.hpp"
#include "mlq_fast.hpp"
#include "mlq_diag.hpp"

namespace mlq {

// out[k] = sum_i part[i*K + k]; one workgroup per k (grid K), fixed order
__global__ void k_reduce_final(const double* part, int nblk, int K, double* out) {
  __shared__ double sh[RED_THREADS];
  const int k = blockIdx.x;
  double s = 0.0;
  for (int i = threadIdx.x; i < nblk; i += blockDim.x) s += part[(long long)i * K + k];
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int o = blockDim.x / 2; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[k] = sh[0];
}

// out[s] = sum of the workgroup partials part[b * n + s]: one workgroup per
// shell (grid n), a fixed summation order
__global__ void k_shell_final(const double* part, int nblk, int n, double* out) {
  __shared__ double sh[256];
  const int s = blockIdx.x;
  double acc = 0.0;
  for (int b = threadIdx.x; b < nblk; b += 256) acc += part[(long long)b * n + s];
  sh[threadIdx.x] = acc;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[s] = sh[0];
}

template class Engine<double>;

std::unique_ptr<IEngine> make_engine_fp64(int n, const mlq_params& p) {
  return std::unique_ptr<IEngine>(new Engine<double>(n, p));
}

}  // namespace mlq
