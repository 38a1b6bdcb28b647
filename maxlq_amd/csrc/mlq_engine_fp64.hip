// FP64 instantiation of the engine (the parity path).
#include "mlq_ops_core.hpp"
#include "mlq_rcg.hpp"
#include "mlq_fast.hpp"

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

// out[s] = sum of the workgroup partials part[b * n + s] in workgroup order
__global__ void k_shell_final(const double* part, int nblk, int n, double* out) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  double acc = 0.0;
#pragma unroll 8
  for (int b = 0; b < nblk; ++b) acc += part[(long long)b * n + s];
  out[s] = acc;
}

template class Engine<double>;

std::unique_ptr<IEngine> make_engine_fp64(int n, const mlq_params& p) {
  return std::unique_ptr<IEngine>(new Engine<double>(n, p));
}

}  // namespace mlq
