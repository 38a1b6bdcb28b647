// Forward real 3-D transform (fftfwdr) on device arrays with no device
// allocation: the intermediate spectrum of the x, y and z passes lives in a
// file-scope device array (`scratch`) instead of the engine's allocated
// transform scratch.  For callers that cannot call hipMalloc (code under HIP
// graph capture, tight serving loops, harnesses that forbid it).
//
// The array holds one half spectrum of up to 128^3 FP64 values with the
// engine's row pitch: 128 x 128 x 72 complex doubles, 18.9 MB, held by every
// GPU the library runs on until the process ends.  Its device address is
// taken once per device by noalloc_bind() (mlq_init) with
// hipGetSymbolAddress; host code passes that address, never the array's
// name, to the kernels.  Every call on one GPU shares the array and finds the
// last call's contents there; nothing in it needs to start at zero (the x
// pass writes every entry the y and z passes and the unpacking read).  Calls
// that overlap on one GPU race on it.
#include <mutex>

#include "mlq_engine.hpp"
#include "mlq_kernels.hpp"

namespace mlq {

constexpr int NOALLOC_MAXN = 128;
constexpr long long NOALLOC_ELEMS =
    (long long)NOALLOC_MAXN * NOALLOC_MAXN * spec_pitch<double>(NOALLOC_MAXN);

static __device__ __align__(256) double2 scratch[NOALLOC_ELEMS];

namespace {
constexpr int MAXDEV = 64;
std::mutex bind_mu;
double2* bound[MAXDEV] = {};   // device address of `scratch` per device

// half spectrum with the engine's row pitch -> the reference's (n/2+1) rows
__global__ void k_unpad(const double2* __restrict__ s, double2* __restrict__ o, int n, int pp) {
  const long long px = n / 2 + 1, total = px * n * n;
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= total) return;
  const long long row = i / px, kx = i - row * px;
  o[i] = s[row * pp + kx];
}

template <int N>
void fwdr_n(hipStream_t st, const double2* tw, const double* u, double2* out, double2* scr) {
  constexpr int E = FftShape<N>::E, Tf = FftShape<N>::Tf;
  constexpr int PX = N / 2 + 1, PP = spec_pitch<double>(N);
  constexpr int NF = (256 / Tf > 0 ? 256 / Tf : 1) < N * N / 2 ? (256 / Tf > 0 ? 256 / Tf : 1)
                                                                 : N * N / 2;
  constexpr int C = 8 > 256 / Tf ? 8 : 256 / Tf;   // the strided passes' default tile
  static_assert((long long)N * N * PP <= NOALLOC_ELEMS, "one half spectrum fits the scratch");
  XArgs<double> xa{};
  xa.rsrc[0] = u;
  xa.spec = scr;
  xa.tw = tw;
  xa.sstride = (long long)N * N * PP;
  xa.nb = 1;
  xa.keep = -1;
  xa.scale = 1.0;
  hipLaunchKernelGGL((k_xfwd<double, N, NF, false>), dim3((unsigned)((long long)N * N / 2 / NF)),
                     dim3(NF * Tf), 0, st, xa);
  MLQ_HIP(hipGetLastError());
  SArgs<double> sa{};
  sa.spec = scr;
  sa.tw = tw;
  sa.sstride = (long long)N * N * PP;
  sa.keep = -1;
  hipLaunchKernelGGL((k_strided<double, N, E, C, -1, 0>), dim3((PX + C - 1) / C, N), dim3(C * Tf),
                     0, st, sa, PX);
  MLQ_HIP(hipGetLastError());
  sa.dst[0] = scr;   // in place: each column is read whole before it is written
  sa.op[0] = sp_none();
  hipLaunchKernelGGL((k_strided<double, N, E, C, -1, 2>), dim3((N * PP + C - 1) / C, 1),
                     dim3(C * Tf), 0, st, sa, N * PP);
  MLQ_HIP(hipGetLastError());
  const long long total = (long long)PX * N * N;
  hipLaunchKernelGGL(k_unpad, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, st, scr, out, N,
                     PP);
  MLQ_HIP(hipGetLastError());
}
}  // namespace

void noalloc_bind() {
  int dev = 0;
  MLQ_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= MAXDEV) throw Error(MLQ_EUNSUP, "device index beyond 63");
  std::lock_guard<std::mutex> lk(bind_mu);
  if (bound[dev]) return;
  void* p = nullptr;
  MLQ_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
  bound[dev] = static_cast<double2*>(p);
}

void fftfwdr_noalloc(int n, hipStream_t st, const double2* tw, const double* u, double* fu) {
  if (n > NOALLOC_MAXN)
    throw Error(MLQ_EINVAL, "mlq_fftfwdr_noalloc: n = " + std::to_string(n) +
                                " exceeds the no-allocation scratch (n <= 128)");
  if (!u || !fu) throw Error(MLQ_EINVAL, "mlq_fftfwdr_noalloc: null device array");
  int dev = 0;
  MLQ_HIP(hipGetDevice(&dev));
  double2* scr = nullptr;
  {
    std::lock_guard<std::mutex> lk(bind_mu);
    if (dev >= 0 && dev < MAXDEV) scr = bound[dev];
  }
  if (!scr)
    throw Error(MLQ_EINVAL, "mlq_fftfwdr_noalloc: no context was created on this device");
  double2* out = reinterpret_cast<double2*>(fu);
  switch (n) {
    case 8: fwdr_n<8>(st, tw, u, out, scr); break;
    case 16: fwdr_n<16>(st, tw, u, out, scr); break;
    case 32: fwdr_n<32>(st, tw, u, out, scr); break;
    case 64: fwdr_n<64>(st, tw, u, out, scr); break;
    case 128: fwdr_n<128>(st, tw, u, out, scr); break;
    default: throw Error(MLQ_EINVAL, "mlq_fftfwdr_noalloc: unsupported n");
  }
  MLQ_HIP(hipStreamSynchronize(st));
}

}  // namespace mlq
