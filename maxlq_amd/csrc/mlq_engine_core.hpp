// Engine<T> definitions: memory, transforms, primitives.
#pragma once
#include <cstdio>
#include <cstdlib>
#include "mlq_engine_impl.hpp"

namespace mlq {

// ---------------------------------------------------------------------------
// set-up
// ---------------------------------------------------------------------------
template <typename T> Engine<T>::Engine(int n, const mlq_params& p) : p_(p) {
  n_ = n;
  px_ = n / 2 + 1;
  pp_ = spec_pitch<T>(n);
  nr_ = (long long)n * n * n;
  ns_ = (long long)n * n * pp_;
  resolve_params();
  MLQ_HIP(hipStreamCreateWithFlags(&st_, hipStreamNonBlocking));
  {
    const char* f = std::getenv("MLQ_FUSED");
    fused_ = !(f && f[0] == '0');
  }
  // twiddles exp(-2 pi i k/N), computed in FP64 on the host
  std::vector<double2> tw(n);
  for (int k = 0; k < n; ++k) {
    const double a = -2.0 * M_PI * (double)k / (double)n;
    tw[k].x = std::cos(a);
    tw[k].y = std::sin(a);
  }
  tw_ = (double2*)dmalloc(sizeof(double2) * n);
  MLQ_HIP(hipMemcpy(tw_, tw.data(), sizeof(double2) * n, hipMemcpyHostToDevice));
  // transform scratch: up to MAXB spectra, fewer when one spectrum is a
  // sizeable share of the device (1024^3: 3 spectra instead of 12)
  {
    size_t fr = 0, tot = 0;
    MLQ_HIP(hipMemGetInfo(&fr, &tot));
    const double per = (double)sizeof(C) * (double)ns_;
    sb_ = std::max(3, std::min(MAXB, (int)(0.06 * (double)tot / per)));
  }
  scratch_ = (C*)dmalloc(sizeof(C) * ns_ * sb_);
  part_ = (double*)dmalloc(sizeof(double) * std::max(RED_BLOCKS, RED_MAXBLK) * RED_MAXK);
  red_dev_ = (double*)dmalloc(sizeof(double) * 64);
  MLQ_HIP(hipHostMalloc((void**)&red_host_, sizeof(double) * 64, hipHostMallocDefault));
  MLQ_HIP(hipEventCreateWithFlags(&done_ev_, hipEventDisableTiming));
}

template <typename T> Engine<T>::~Engine() {
  hipStreamSynchronize(st_);
  for (auto& f : fields_) {
    if (f.phys) hipFree(f.phys);
    if (f.spec) hipFree(f.spec);
  }
  slabs_.for_each_slab([](char* base, size_t) { hipFree(base); });
  for (auto& r : prof_recs_) { hipEventDestroy(r.a); hipEventDestroy(r.b); }
  for (auto e : ev_pool_) hipEventDestroy(e);
  hipFree(tw_);
  hipFree(scratch_);
  hipFree(part_);
  if (lspart_) hipFree(lspart_);
  if (lsrisky_) hipFree(lsrisky_);
  hipFree(red_dev_);
  hipHostFree(red_host_);
  hipEventDestroy(done_ev_);
  hipStreamDestroy(st_);
}

template <typename T> bool Engine<T>::fast_path_fits() const {
  size_t fr = 0, tot = 0;
  if (hipMemGetInfo(&fr, &tot) != hipSuccess) { (void)hipGetLastError(); return true; }
  return fast_path_bytes() <= 0.95 * ((double)fr + (double)pooled_);
}

template <typename T> void Engine<T>::resolve_params() {
  if (p_.use_e_u < 0) use_e_u_ = p_.q < 4.0 - 2.0e-16;   // initialize.f90:79-85
  else use_e_u_ = p_.use_e_u != 0;
}

template <typename T> void* Engine<T>::dmalloc(size_t bytes) {
  void* p = nullptr;
  const bool log = alloc_log_level() > 0;
  double t0 = log ? alloc_log_ms() : 0.0;
  hipError_t e = hipMalloc(&p, bytes);
  if (e != hipSuccess) {
    // give back the slabs that hold no temporary and retry
    (void)hipGetLastError();
    const size_t had = release_free_slabs();
    if (log) {
      const double t1 = alloc_log_ms();
      std::fprintf(stderr, "[mlq] at %.0f ms: %zu bytes of free slabs given back for %zu: %.1f ms\n",
                   t0, had, bytes, t1 - t0);
      t0 = t1;
    }
    e = hipMalloc(&p, bytes);
  }
  if (e != hipSuccess) {
    (void)hipGetLastError();
    throw Error(MLQ_EOOM, "hipMalloc of " + std::to_string(bytes) + " bytes failed");
  }
  used_ += bytes;
  peak_ = std::max(peak_, used_);
  note_live();
  if (log)
    std::fprintf(stderr, "[mlq] at %.0f ms: device allocation of %zu bytes (in use %zu): %.1f ms\n",
                 t0, bytes, used_, alloc_log_ms() - t0);
  return p;
}
template <typename T> void Engine<T>::dfree(void* p, size_t bytes) {
  if (!p) return;
  const double t0 = alloc_log_level() > 0 ? alloc_log_ms() : 0.0;
  MLQ_HIP(hipStreamSynchronize(st_));
  MLQ_HIP(hipFree(p));
  used_ -= bytes;
  if (alloc_log_level() > 0)
    std::fprintf(stderr, "[mlq] at %.0f ms: device free of %zu bytes: %.1f ms\n", t0, bytes,
                 alloc_log_ms() - t0);
}

template <typename T> size_t Engine<T>::release_free_slabs() {
  const size_t had = slabs_.release_free([&](char* base, size_t size) { dfree(base, size); });
  pooled_ -= had;
  return had;
}

template <typename T> void* Engine<T>::carve(size_t bytes) {
  bytes = SlabSet::round(bytes);
  char* p = slabs_.take(bytes);
  if (!p) {
    // a new slab: at least 9 real fields (the largest temporaries are
    // 9-component gradients), or just this block when that does not fit
    const size_t want = std::max(bytes, SlabSet::round(sizeof(T) * (size_t)nr_ * 9));
    void* base = nullptr;
    size_t got = want;
    if (want > bytes && hipMalloc(&base, want) == hipSuccess) {
      used_ += want;
      peak_ = std::max(peak_, used_);
      if (alloc_log_level() > 0)
        std::fprintf(stderr, "[mlq] at %.0f ms: slab of %zu bytes (in use %zu)\n", alloc_log_ms(),
                     want, used_);
    } else {
      (void)hipGetLastError();
      got = bytes;
      base = dmalloc(bytes);
    }
    slabs_.add((char*)base, got);
    pooled_ += got;
    p = slabs_.take(bytes);
  }
  pooled_ -= bytes;
  note_live();
  return p;
}

template <typename T> void Engine<T>::pool_put(void* p, size_t bytes) {
  if (alloc_log_level() > 1) std::fprintf(stderr, "[mlq] put %zu %p\n", bytes, p);
  const size_t len = slabs_.give(p);
  if (!len) {
    std::fprintf(stderr, "[mlq] temporary %p was not carved from a slab\n", p);
    std::abort();
  }
  pooled_ += len;
}

template <typename T> typename Engine<T>::RT Engine<T>::rtmp(int k) {
  const size_t bytes = sizeof(T) * nr_ * k;
  void* p = carve(bytes);
  if (alloc_log_level() > 1) std::fprintf(stderr, "[mlq] get %zu %p\n", bytes, p);
  return RT(this, (T*)p, bytes);
}
template <typename T> typename Engine<T>::ST Engine<T>::stmp(int k) {
  const size_t bytes = sizeof(C) * ns_ * k;
  void* p = carve(bytes);
  if (alloc_log_level() > 1) std::fprintf(stderr, "[mlq] get %zu %p\n", bytes, p);
  return ST(this, (C*)p, bytes);
}

template <typename T> DevField& Engine<T>::fld(int h) {
  if (h < 0 || h >= (int)fields_.size() || fields_[h].ncomp == 0)
    throw Error(MLQ_EINVAL, "bad field handle " + std::to_string(h));
  return fields_[h];
}
template <typename T> const DevField& Engine<T>::fld(int h) const {
  if (h < 0 || h >= (int)fields_.size() || fields_[h].ncomp == 0)
    throw Error(MLQ_EINVAL, "bad field handle " + std::to_string(h));
  return fields_[h];
}

template <typename T> int Engine<T>::field_alloc(int ncomp) {
  if (ncomp != 1 && ncomp != 3) throw Error(MLQ_EINVAL, "ncomp must be 1 or 3");
  DevField f;
  f.ncomp = ncomp;
  f.phys = dmalloc(sizeof(T) * nr_ * ncomp);
  MLQ_HIP(hipMemsetAsync(f.phys, 0, sizeof(T) * nr_ * ncomp, st_));
  for (size_t i = 0; i < fields_.size(); ++i)
    if (fields_[i].ncomp == 0) { fields_[i] = f; return (int)i; }
  fields_.push_back(f);
  return (int)fields_.size() - 1;
}
template <typename T> void Engine<T>::field_free(int h) {
  DevField& f = fld(h);
  dfree(f.phys, sizeof(T) * nr_ * f.ncomp);
  if (f.spec) dfree(f.spec, sizeof(C) * ns_ * f.ncomp);
  f = DevField();
}

template <typename T> void Engine<T>::upload(int h, const double* host) {
  DevField& f = fld(h);
  const long long cnt = nr_ * f.ncomp;
  if (sizeof(T) == sizeof(double)) {
    MLQ_HIP(hipMemcpyAsync(f.phys, host, sizeof(double) * cnt, hipMemcpyHostToDevice, st_));
  } else {
    std::vector<T> tmp(cnt);
    for (long long i = 0; i < cnt; ++i) tmp[i] = (T)host[i];
    MLQ_HIP(hipMemcpyAsync(f.phys, tmp.data(), sizeof(T) * cnt, hipMemcpyHostToDevice, st_));
    sync();
  }
  f.spec_valid = false;
  sync();
}
template <typename T> void Engine<T>::download(int h, double* host) {
  DevField& f = fld(h);
  const long long cnt = nr_ * f.ncomp;
  if (sizeof(T) == sizeof(double)) {
    MLQ_HIP(hipMemcpyAsync(host, f.phys, sizeof(double) * cnt, hipMemcpyDeviceToHost, st_));
    sync();
  } else {
    std::vector<T> tmp(cnt);
    MLQ_HIP(hipMemcpyAsync(tmp.data(), f.phys, sizeof(T) * cnt, hipMemcpyDeviceToHost, st_));
    sync();
    for (long long i = 0; i < cnt; ++i) host[i] = (double)tmp[i];
  }
}

// ---------------------------------------------------------------------------
// profiling (HIP events on the engine stream)
// ---------------------------------------------------------------------------
template <typename T> void Engine<T>::prof_begin(int cls) {
  if (!prof_) return;
  ProfRec r;
  r.cls = cls;
  r.bytes = 0;
  for (hipEvent_t* e : {&r.a, &r.b}) {
    if (!ev_pool_.empty()) { *e = ev_pool_.back(); ev_pool_.pop_back(); }
    else MLQ_HIP(hipEventCreate(e));
  }
  MLQ_HIP(hipEventRecord(r.a, st_));
  prof_recs_.push_back(r);
}
template <typename T> void Engine<T>::prof_end(int cls, double bytes) {
  if (!prof_) return;
  ProfRec& r = prof_recs_.back();
  r.bytes = bytes;
  MLQ_HIP(hipEventRecord(r.b, st_));
  if (prof_recs_.size() > 4096) prof_flush();
}
template <typename T> void Engine<T>::prof_flush() {
  if (prof_recs_.empty()) return;
  MLQ_HIP(hipStreamSynchronize(st_));
  for (auto& r : prof_recs_) {
    float ms = 0.f;
    MLQ_HIP(hipEventElapsedTime(&ms, r.a, r.b));
    prof_count_[r.cls] += 1;
    prof_ms_[r.cls] += ms;
    prof_bytes_[r.cls] += r.bytes;
    ev_pool_.push_back(r.a);
    ev_pool_.push_back(r.b);
  }
  prof_recs_.clear();
}
template <typename T> void Engine<T>::prof_enable(bool on) {
  prof_flush();
  prof_ = on;
  if (on) {
    for (int i = 0; i < PC_COUNT; ++i) { prof_count_[i] = 0; prof_ms_[i] = 0; prof_bytes_[i] = 0; }
  }
}
template <typename T> void Engine<T>::prof_get(int cls, long long* count, double* ms, double* bytes) {
  prof_flush();
  if (cls < 0 || cls >= PC_COUNT) throw Error(MLQ_EINVAL, "bad profile class");
  *count = prof_count_[cls];
  *ms = prof_ms_[cls];
  *bytes = prof_bytes_[cls];
}

// ---------------------------------------------------------------------------
// generic launches
// ---------------------------------------------------------------------------
template <typename T> template <class F> void Engine<T>::map(const F& f, long long n, int cls) {
  constexpr int vec = VecW<F>::v;
  if (n % vec) throw Error(MLQ_EINVAL, "map length not a multiple of its vector width");
  const double ab = (double)n * AlgB<F>::of(f);
  n /= vec;
  if (n <= 0) return;
  // one call per thread (the grid-stride loop only runs past 2^24
  // workgroups): a capped grid that strides over the fields measured 25%
  // below the streaming rate of a one-pass grid
  long long blocks = (n + MAP_THREADS - 1) / MAP_THREADS;
  blocks = std::min<long long>(blocks, MLQ_MAP_CAP);
  prof_begin(cls);
  hipLaunchKernelGGL((k_map<F>), dim3((unsigned)blocks), dim3(MAP_THREADS), 0, st_, f, n);
  MLQ_HIP(hipGetLastError());
  prof_end(cls, ab);
}

template <typename T> template <int K, class F>
void Engine<T>::reduce(const F& f, long long n, double* out) {
  constexpr int vec = VecW<F>::v;
  if (n % vec) throw Error(MLQ_EINVAL, "reduction length not a multiple of its vector width");
  const double ab = (double)n * AlgB<F>::of(f);
  n /= vec;
  prof_begin(PC_REDUCE);
  const int nblk = (int)red_blocks(n);
  hipLaunchKernelGGL((k_reduce<K, F>), dim3(nblk), dim3(RED_THREADS), 0, st_, f, n, part_);
  MLQ_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_reduce_final, dim3(K), dim3(RED_THREADS), 0, st_, part_, nblk, K,
                     red_dev_);
  MLQ_HIP(hipGetLastError());
  prof_end(PC_REDUCE, ab);
  MLQ_HIP(hipMemcpyAsync(red_host_, red_dev_, sizeof(double) * K, hipMemcpyDeviceToHost, st_));
  wait_spin();
  for (int k = 0; k < K; ++k) out[k] = red_host_[k];
}

template <typename T> void Engine<T>::fetch(int n, const C* const* p, std::complex<double>* out) {
  if (n < 1 || n > FETCH_MAX) throw Error(MLQ_EINVAL, "fetch: count out of range");
  FetchArgs<T> fa{};
  for (int i = 0; i < n; ++i) fa.p[i] = p[i];
  fa.n = n;
  hipLaunchKernelGGL((k_fetch<T>), dim3(1), dim3(64), 0, st_, fa, red_dev_);
  MLQ_HIP(hipGetLastError());
  MLQ_HIP(hipMemcpyAsync(red_host_, red_dev_, sizeof(double) * 2 * n, hipMemcpyDeviceToHost, st_));
  wait_spin();
  for (int i = 0; i < n; ++i) out[i] = std::complex<double>(red_host_[2 * i], red_host_[2 * i + 1]);
}

// Wait for the stream by polling an event: a result the host needs before
// its next launch (a reduction, a few scalars) comes back without the
// wake-up latency of a blocking synchronisation (measured ~25-40 us per
// read-back between two kernels of the iteration).
template <typename T> void Engine<T>::wait_spin() {
  MLQ_HIP(hipEventRecord(done_ev_, st_));
  for (;;) {
    const hipError_t e = hipEventQuery(done_ev_);
    if (e == hipSuccess) return;
    if (e != hipErrorNotReady) MLQ_HIP(e);
  }
}

template <typename T> double Engine<T>::dot(const T* a, const T* b) {
  double r;
  reduce<1>(R_dot<T>{a, b}, nr_, &r);
  return r;
}
template <typename T> double Engine<T>::sum(const T* a) {
  double r;
  reduce<1>(R_sum<T>{a}, nr_, &r);
  return r;
}
template <typename T> bool Engine<T>::has_nan(const T* a, long long n) {
  double r;
  reduce<1>(R_nan<T>{a}, n, &r);
  return r != 0.0;
}
template <typename T> double Engine<T>::spec_dot(const C* a, const C* b, int kind) {
  double r;
  const double s = 1.5 - 1.0 / p_.q;
  reduce<1>(R_specdot<T>{a, b, n_, pp_, kind, p_.lambda1, s}, ns_, &r);
  return r;
}

// ---------------------------------------------------------------------------
// batched 3-D transforms
// ---------------------------------------------------------------------------
template <typename T> struct StridedCfg {
  static constexpr int cmin = sizeof(T) == 4 ? 16 : 8;  // 128-byte row segments
};
template <int N, typename T> constexpr int strided_C() {
  constexpr int Tf = FftShape<N>::Tf;
  constexpr int c = StridedCfg<T>::cmin > 256 / Tf ? StridedCfg<T>::cmin : 256 / Tf;
  return c;
}
// Tile variants of the strided (y/z) passes: E elements per thread, C adjacent
// columns per workgroup.  Variant 0 is the default; 1-3 are the candidates
// the autotuner times for N >= 128 (invalid shapes fall back to variant 0).
template <typename T, int N, int V> struct SVar {
  static constexpr int cmin = sizeof(T) == 4 ? 16 : 8;  // 128-byte row segments
  static constexpr int E0 = FftShape<N>::E;
  static constexpr int C0 = cmin > 256 / (N / E0) ? cmin : 256 / (N / E0);
  static constexpr int Er = V == 0 ? E0 : (V == 3 ? 8 : 16);
  static constexpr int Cr = V == 0 ? C0 : (V == 1 ? cmin : (V == 2 ? 2 * cmin : cmin / 2));
  // the column tile and the twiddle table (N complex doubles) fit the LDS
  static constexpr bool ok = V == 0 || (N >= 128 && Er <= N && (N / Er) * Cr <= 1024 &&
                                        (long long)N * Cr * 2 * (long long)sizeof(T) + 16LL * N <=
                                            160 * 1024);
  static constexpr int E = ok ? Er : E0;
  static constexpr int C = ok ? Cr : C0;
};

#ifndef MLQ_XINV_ZB
#define MLQ_XINV_ZB 1   // band-limited inverse passes: zero-band first radix pass
#endif
#ifndef MLQ_FWD_ZB
#define MLQ_FWD_ZB 1    // band-limited forward passes: zero-band last radix pass
#endif
template <typename T, int N, int V, int DIR, int MODE>
void launch_strided_v(hipStream_t st, SArgs<T> sa, int ncols, int planes) {
  using S = SVar<T, N, V>;
  // pruned z passes that read and write kept columns only run the kept tiles
  // alone (the other tiles of a full grid exit at once, 57% of them at the
  // 2/3 mask)
  sa.kcomp = MODE != 0 && sa.keep >= 0 && 2 * sa.keep + 1 < N && (MODE == 1 || sa.kzero == 0);
  const unsigned gx = sa.kcomp ? (unsigned)(2 * sa.keep + 1) * (unsigned)((sa.keep + S::C) / S::C)
                               : (unsigned)((ncols + S::C - 1) / S::C);
  const dim3 grid(gx, planes), block(S::C * (N / S::E));
  // band-limited launches: zero-band first (inverse) or last (forward) radix pass
  if ((DIR > 0 ? MLQ_XINV_ZB : MLQ_FWD_ZB) && sa.keep >= 0 && 8 * sa.keep < 3 * N) {
    hipLaunchKernelGGL((k_strided<T, N, S::E, S::C, DIR, MODE, true>), grid, block, 0, st, sa,
                       ncols);
    return;
  }
  hipLaunchKernelGGL((k_strided<T, N, S::E, S::C, DIR, MODE>), grid, block, 0, st, sa, ncols);
}
// var: bits 0-1 the tile variant, bit 2 the XCD-contiguous tile order
template <typename T, int N, int DIR, int MODE>
void launch_strided(int var, hipStream_t st, SArgs<T> sa, int ncols, int planes) {
  sa.remap = (var >> 2) & 1;
  switch (var & 3) {
    case 1: launch_strided_v<T, N, 1, DIR, MODE>(st, sa, ncols, planes); break;
    case 2: launch_strided_v<T, N, 2, DIR, MODE>(st, sa, ncols, planes); break;
    case 3: launch_strided_v<T, N, 3, DIR, MODE>(st, sa, ncols, planes); break;
    default: launch_strided_v<T, N, 0, DIR, MODE>(st, sa, ncols, planes); break;
  }
}

// threads per workgroup of the x passes (row pairs of Tf threads each):
// 256 measured faster than 128 and 512 at 512^3, and equal to 64
#ifndef MLQ_XPASS_THREADS
#define MLQ_XPASS_THREADS 256
#endif
template <int N> constexpr int xpass_NF() {
  constexpr int Tf = FftShape<N>::Tf;
  constexpr int nf = MLQ_XPASS_THREADS / Tf > 0 ? MLQ_XPASS_THREADS / Tf : 1;
  return nf < N * N / 2 ? nf : N * N / 2;
}

// the band of an inverse batch: the effective input op[b] * src[b] is
// band-limited when the caller says so (keep >= 0) or when every op masks
template <typename T> int Engine<T>::inv_keep(int nb, const SpecOp* ops, int keep) const {
  int km = -1;
  for (int b = 0; b < nb; ++b) {
    if (!(ops[b].flags & SP_MASK) || ops[b].mcut >= n_ / 2) { km = -1; break; }
    km = std::max(km, ops[b].mcut);
  }
  if (km >= 0 && (keep < 0 || km < keep)) keep = km;
  return keep >= n_ / 2 ? -1 : keep;
}

// pruned when every output of the batch is masked to a cube smaller than the grid
template <typename T> int Engine<T>::fwd_keep(int nb, const SpecOp* ops) const {
  int keep = -1;
  for (int b = 0; b < nb; ++b) {
    if (!(ops[b].flags & SP_MASK) || ops[b].mcut >= n_ / 2) return -1;
    keep = std::max(keep, ops[b].mcut);
  }
  return keep;
}

template <typename T> template <int N>
void Engine<T>::fwd_n(int nb, const T* const* src, C* const* dst, const SpecOp* ops) {
  constexpr int NF = xpass_NF<N>();
  constexpr int Tf = FftShape<N>::Tf;
  XArgs<T> xa{};
  for (int b = 0; b < nb; ++b) xa.rsrc[b] = src[b];
  xa.spec = scratch_;
  xa.tw = tw_;
  xa.sstride = ns_;
  xa.nb = nb;
  xa.scale = (T)1;
  const int keep = fwd_keep(nb, ops);
  xa.keep = keep;
  nfwd_ += nb;
  const long long pairs = (long long)nb * N * N / 2;
  prof_begin(PC_XFWD);
  if (MLQ_FWD_ZB && keep >= 0 && 8 * keep < 3 * N)   // zero-band last radix pass
    hipLaunchKernelGGL((k_xfwd<T, N, NF, true, true>), dim3((unsigned)(pairs / NF)), dim3(NF * Tf), 0,
                       st_, xa);
  else if (keep >= 0)
    hipLaunchKernelGGL((k_xfwd<T, N, NF, true>), dim3((unsigned)(pairs / NF)), dim3(NF * Tf), 0, st_, xa);
  else
    hipLaunchKernelGGL((k_xfwd<T, N, NF, false>), dim3((unsigned)(pairs / NF)), dim3(NF * Tf), 0, st_, xa);
  MLQ_HIP(hipGetLastError());
  prof_end(PC_XFWD, nb * (real_bytes() + ab_xcols(keep)));
  fwd_yz_n<N>(nb, dst, ops, keep);
}

// y and z passes of a forward transform whose x pass left nb spectra in the
// scratch slots 0..nb-1
template <typename T> template <int N>
void Engine<T>::fwd_yz_n(int nb, C* const* dst, const SpecOp* ops, int keep, bool kept_only) {
  constexpr int PX = N / 2 + 1, PP = spec_pitch<T>(N);
  const double sb = (double)nb * spec_bytes();
  SArgs<T> sa{};
  sa.spec = scratch_;
  sa.tw = tw_;
  sa.sstride = ns_;
  sa.keep = keep;
  prof_begin(PC_YFWD);
  launch_strided<T, N, -1, 0>(svar_[keep >= 0 ? 4 : 0], st_, sa, keep >= 0 ? keep + 1 : PX, N * nb);
  MLQ_HIP(hipGetLastError());
  prof_end(PC_YFWD, nb * (ab_xcols(keep) + ab_kcols(keep)));
  for (int b = 0; b < nb; ++b) { sa.dst[b] = dst[b]; sa.op[b] = ops[b]; }
  sa.kzero = kept_only ? 0 : 1;
  prof_begin(PC_ZFWD);
  launch_strided<T, N, -1, 2>(svar_[keep >= 0 ? 5 : 1], st_, sa, N * PP, nb);
  MLQ_HIP(hipGetLastError());
  prof_end(PC_ZFWD, nb * (ab_kcols(keep) + (kept_only ? ab_cube(keep) : spec_bytes())));
}

template <typename T> template <int N, class P>
void Engine<T>::fwd_prod_n(const P& prod, C* const* dst, const SpecOp* ops, bool& done) {
  if constexpr (xfwd_p_fits<P, T, N>()) {
    constexpr int Tf = FftShape<N>::Tf;
    XArgs<T> xa{};
    xa.spec = scratch_;
    xa.tw = tw_;
    xa.sstride = ns_;
    xa.nb = P::NB;
    xa.scale = (T)1;
    const int keep = fwd_keep(P::NB, ops);
    xa.keep = keep;
    const unsigned pairs = (unsigned)((long long)N * N / 2);
    prof_begin(PC_XFWD);
    if (MLQ_FWD_ZB && keep >= 0 && 8 * keep < 3 * N)   // zero-band last radix pass
      hipLaunchKernelGGL((k_xfwd_p<T, N, true, P, true>), dim3(pairs), dim3(P::NB * Tf), 0, st_, prod,
                         xa);
    else if (keep >= 0)
      hipLaunchKernelGGL((k_xfwd_p<T, N, true, P>), dim3(pairs), dim3(P::NB * Tf), 0, st_, prod, xa);
    else
      hipLaunchKernelGGL((k_xfwd_p<T, N, false, P>), dim3(pairs), dim3(P::NB * Tf), 0, st_, prod, xa);
    MLQ_HIP(hipGetLastError());
    prof_end(PC_XFWD, (double)nr_ * ProdB<P>::of(prod) + P::NB * ab_xcols(keep));
    fwd_yz_n<N>(P::NB, dst, ops, keep, true);
    nfwd_ += P::NB;
    done = true;
  }
}

// inverse transforms of P::NI fields (stored to dst[b] where set) and forward
// transforms of the P::NB fields that P computes per point from them and its
// own loads, with one x pass for both (k_xinv_fwd_p); false (nothing
// launched) when the batch does not run that way: the caller then runs inv
// and fwd_prod.  Both sides band-limited with the zero band (keep < 3N/8).
template <typename T> template <int N, class P>
void Engine<T>::inv_fwd_prod_n(const InvSrc* src, const SpecOp* iops, T* const* idst, int ikeep,
                               const P& prod, C* const* dst, const SpecOp* ops, bool& done) {
  if constexpr (xinv_fwd_p_fits<P, T, N>()) {
    constexpr int Tf = FftShape<N>::Tf;
    const int okeep = fwd_keep(P::NB, ops);
    auto zb = [](int k) { return k >= 0 && 8 * k < 3 * N; };
    if (!MLQ_XINV_ZB || !MLQ_FWD_ZB || !zb(okeep) || !zb(inv_keep(P::NI, iops, ikeep))) return;
    const XArgs<T> ai = inv_zy_n<N>(P::NI, src, iops, idst, ikeep);
    XArgs<T> ao{};
    ao.spec = scratch_;
    ao.tw = tw_;
    ao.sstride = ns_;
    ao.nb = P::NB;
    ao.scale = (T)1;
    ao.keep = okeep;
    int nst = 0;
    for (int b = 0; b < P::NI; ++b) nst += idst[b] != nullptr;
    prof_begin(PC_XFWD);
    hipLaunchKernelGGL((k_xinv_fwd_p<T, N, P>), dim3((unsigned)((long long)N * N / 2)),
                       dim3(P::NB * Tf), 0, st_, prod, ai, ao);
    MLQ_HIP(hipGetLastError());
    prof_end(PC_XFWD, P::NI * ab_xcols(ikeep) + nst * real_bytes() +
                          (double)nr_ * ProdB<P>::of(prod) + P::NB * ab_xcols(okeep));
    fwd_yz_n<N>(P::NB, dst, ops, okeep, true);
    nfwd_ += P::NB;
    done = true;
  }
}

// z and y passes of an inverse transform; returns the arguments of its x
// pass (the spectra it reads from the scratch slots, keep, scale)
template <typename T> template <int N>
XArgs<T> Engine<T>::inv_zy_n(int nb, const InvSrc* src, const SpecOp* ops, T* const* dst, int keep) {
  constexpr int PX = N / 2 + 1, PP = spec_pitch<T>(N);
  keep = inv_keep(nb, ops, keep);
  ninv_ += nb;
  // Factored plan.  The i K_x and i K_y parts of a spectral factor commute
  // with the z pass (i K_x also with the y pass), so they are applied in the
  // x and y passes instead, and outputs that differ only in them share their
  // z (and y) transforms: d and its 3 derivatives need 2 z and 3 y
  // transforms instead of 4 each.
  SpecOp zop[MAXB];
  int fdim[MAXB];
  for (int b = 0; b < nb; ++b) {
    zop[b] = ops[b];
    fdim[b] = -1;
    if ((ops[b].flags & SP_IK) && ops[b].dim < 2) {
      fdim[b] = ops[b].dim;
      zop[b].flags &= ~SP_IK;
      zop[b].dim = 0;
    }
  }
  auto same_z = [&](int a, int b) {
    const SpecOp &p = zop[a], &q = zop[b];
    if (p.flags != q.flags || p.scale != q.scale) return false;
    if ((p.flags & SP_MASK) && p.mcut != q.mcut) return false;
    if ((p.flags & SP_IK) && p.dim != q.dim) return false;
    if (src[a].ns != src[b].ns) return false;
    for (int i = 0; i < src[a].ns; ++i)
      if (src[a].s[i] != src[b].s[i] || src[a].c[i] != src[b].c[i]) return false;
    return true;
  };
  int zrep[MAXB], zid[MAXB], nz = 0;
  for (int b = 0; b < nb; ++b) {
    int j = 0;
    while (j < nz && !same_z(zrep[j], b)) ++j;
    if (j == nz) zrep[nz++] = b;
    zid[b] = j;
  }
  // y transforms: one per (z transform, i K_y or not); the first of each z
  // transform runs in place, the others (launched first) to new spectra
  int yz[MAXB], yk[MAXB], yslot[MAXB], yid[MAXB], ny = 0, next = nz;
  for (int b = 0; b < nb; ++b) {
    const int k = fdim[b] == 1;
    int p = 0;
    while (p < ny && !(yz[p] == zid[b] && yk[p] == k)) ++p;
    if (p == ny) {
      bool first = true;
      for (int r = 0; r < ny; ++r) first = first && yz[r] != zid[b];
      yz[ny] = zid[b];
      yk[ny] = k;
      yslot[ny] = first ? zid[b] : next++;
      ++ny;
    }
    yid[b] = p;
  }
  SArgs<T> sa{};
  sa.spec = scratch_;
  sa.tw = tw_;
  sa.sstride = ns_;
  sa.keep = keep;
  double zin = 0;
  for (int j = 0; j < nz; ++j) {
    const int b = zrep[j];
    sa.nsrc[j] = src[b].ns;
    for (int i = 0; i < src[b].ns; ++i) { sa.src[j][i] = src[b].s[i]; sa.coef[j][i] = src[b].c[i]; }
    sa.op[j] = zop[b];
    zin += src[b].ns;
  }
  prof_begin(PC_ZINV);
  launch_strided<T, N, 1, 1>(svar_[keep >= 0 ? 6 : 2], st_, sa, N * PP, nz);
  MLQ_HIP(hipGetLastError());
  prof_end(PC_ZINV, zin * ab_cube(keep) + nz * ab_kcols(keep));
  prof_begin(PC_YINV);
  for (int pass = 0; pass < 2; ++pass) {   // out of place first, then in place
    SArgs<T> ya = sa;
    ya.ymap = 1;
    int nf = 0;
    for (int p = 0; p < ny; ++p) {
      if ((yslot[p] == yz[p]) != (pass == 1)) continue;
      ya.yin[nf] = yz[p];
      ya.yout[nf] = yslot[p];
      ya.yik[nf] = yk[p];
      ++nf;
    }
    if (nf == 0) continue;
    launch_strided<T, N, 1, 0>(svar_[keep >= 0 ? 7 : 3], st_, ya, keep >= 0 ? keep + 1 : PX,
                               N * nf);
    MLQ_HIP(hipGetLastError());
  }
  prof_end(PC_YINV, ny * (ab_kcols(keep) + ab_xcols(keep)));
  XArgs<T> xa{};
  for (int b = 0; b < nb; ++b) {
    xa.rdst[b] = dst[b];
    xa.xin[b] = yslot[yid[b]];
    xa.xik[b] = fdim[b] == 0;
  }
  xa.xmap = 1;
  xa.spec = scratch_;
  xa.tw = tw_;
  xa.sstride = ns_;
  xa.nb = nb;
  xa.keep = keep;
  xa.scale = (T)(1.0 / ((double)N * N * N));
  return xa;
}

template <typename T> template <int N>
void Engine<T>::inv_n(int nb, const InvSrc* src, const SpecOp* ops, T* const* dst, int keep_in) {
  constexpr int NF = xpass_NF<N>();
  constexpr int Tf = FftShape<N>::Tf;
  const XArgs<T> xa = inv_zy_n<N>(nb, src, ops, dst, keep_in);
  const int keep = xa.keep;
  const double rb = (double)nb * real_bytes();
  const long long pairs = (long long)nb * N * N / 2;
  prof_begin(PC_XINV);
  if (MLQ_XINV_ZB && keep >= 0 && 8 * keep < 3 * N)   // the first radix pass skips the zero band
    hipLaunchKernelGGL((k_xinv<T, N, NF, true, true>), dim3((unsigned)(pairs / NF)), dim3(NF * Tf), 0,
                       st_, xa);
  else if (keep >= 0)
    hipLaunchKernelGGL((k_xinv<T, N, NF, true>), dim3((unsigned)(pairs / NF)), dim3(NF * Tf), 0, st_, xa);
  else
    hipLaunchKernelGGL((k_xinv<T, N, NF, false>), dim3((unsigned)(pairs / NF)), dim3(NF * Tf), 0, st_, xa);
  MLQ_HIP(hipGetLastError());
  prof_end(PC_XINV, rb + nb * ab_xcols(keep));
}

// Slot plan of a round-trip group whose inputs sit in scratch slots
// 0..nin-1 (outputs' `in` index those slots).  z transforms: one per distinct
// (input, factor without its i K_x / i K_y part), in place when it is its
// input's only z transform, else into a fresh slot (the input slot is free
// after the z launch).  y transforms: one per (z transform, i K_y or not),
// the first in place, the others into free or fresh slots.  The i K_x part
// is applied in the inverse x pass.  Returns the slots needed.
template <typename T>
int Engine<T>::rtrip_plan(int nin, int nout, const RtOut* outs, RtPlan& pl) const {
  int zof[MAXB];
  pl.nz = 0;
  for (int o = 0; o < nout; ++o) {
    SpecOp z = outs[o].op;
    pl.oxik[o] = 0;
    int yik = 0;
    if ((z.flags & SP_IK) && z.dim < 2) {
      (z.dim == 0 ? pl.oxik[o] : yik) = 1;
      z.flags &= ~SP_IK;
      z.dim = 0;
    }
    int j = 0;
    for (; j < pl.nz; ++j) {
      const SpecOp& p = pl.zop[j];
      if (pl.zin[j] == outs[o].in && p.flags == z.flags && p.scale == z.scale &&
          (!(p.flags & SP_MASK) || p.mcut == z.mcut) && (!(p.flags & SP_IK) || p.dim == z.dim) &&
          (!(p.flags & SP_HSCALE) || p.hs == z.hs))
        break;
    }
    if (j == pl.nz) {
      if (pl.nz == MAXB) return 1 << 20;
      pl.zin[j] = outs[o].in;
      pl.zop[j] = z;
      ++pl.nz;
    }
    zof[o] = j;
    pl.oy[o] = yik;   // y key, resolved below
  }
  int next = nin;
  int freel[MAXB], nfree = 0;
  for (int i = 0; i < nin; ++i) {
    int cnt = 0, last = -1;
    for (int j = 0; j < pl.nz; ++j)
      if (pl.zin[j] == i) { ++cnt; last = j; }
    if (cnt == 1) {
      pl.zout[last] = i;
    } else {
      for (int j = 0; j < pl.nz; ++j)
        if (pl.zin[j] == i) pl.zout[j] = next++;
      if (cnt > 1) freel[nfree++] = i;
    }
  }
  pl.ny = 0;
  for (int o = 0; o < nout; ++o) {
    const int k = pl.oy[o], z = zof[o];
    int p = 0;
    while (p < pl.ny && !(pl.yz[p] == z && pl.yk[p] == k)) ++p;
    if (p == pl.ny) {
      bool first = true;
      for (int r = 0; r < pl.ny; ++r) first = first && pl.yz[r] != z;
      pl.yz[p] = z;
      pl.yk[p] = k;
      pl.yslot[p] = first ? pl.zout[z] : (nfree > 0 ? freel[--nfree] : next++);
      ++pl.ny;
    }
    pl.oy[o] = p;
  }
  pl.need = next;
  return next;
}

template <typename T> template <int N>
void Engine<T>::rtrip_n(int nin, const T* const* src, int nout, const RtOut* outs, int keep,
                        const T* const* mul) {
  constexpr int NF = xpass_NF<N>();
  constexpr int Tf = FftShape<N>::Tf;
  constexpr int PX = N / 2 + 1, PP = spec_pitch<T>(N);
  RtPlan pl;
  rtrip_plan(nin, nout, outs, pl);
  const double sb = (double)spec_bytes();
  // forward x and y passes of the inputs into slots 0..nin-1
  {
    XArgs<T> xa{};
    int nmul = 0;
    for (int b = 0; b < nin; ++b) {
      xa.rsrc[b] = src[b];
      xa.rmul[b] = mul ? mul[b] : nullptr;
      nmul += xa.rmul[b] != nullptr;
    }
    xa.spec = scratch_;
    xa.tw = tw_;
    xa.sstride = ns_;
    xa.nb = nin;
    xa.scale = (T)1;
    xa.keep = keep;
    const long long pairs = (long long)nin * N * N / 2;
    prof_begin(PC_XFWD);
    if (MLQ_FWD_ZB && keep >= 0 && 8 * keep < 3 * N)
      hipLaunchKernelGGL((k_xfwd<T, N, NF, true, true>), dim3((unsigned)(pairs / NF)), dim3(NF * Tf),
                         0, st_, xa);
    else if (keep >= 0)
      hipLaunchKernelGGL((k_xfwd<T, N, NF, true>), dim3((unsigned)(pairs / NF)), dim3(NF * Tf), 0,
                         st_, xa);
    else
      hipLaunchKernelGGL((k_xfwd<T, N, NF, false>), dim3((unsigned)(pairs / NF)), dim3(NF * Tf), 0,
                         st_, xa);
    MLQ_HIP(hipGetLastError());
    prof_end(PC_XFWD, (nin + nmul) * real_bytes() + nin * ab_xcols(keep));
    SArgs<T> sa{};
    sa.spec = scratch_;
    sa.tw = tw_;
    sa.sstride = ns_;
    sa.keep = keep;
    prof_begin(PC_YFWD);
    launch_strided<T, N, -1, 0>(svar_[keep >= 0 ? 4 : 0], st_, sa, keep >= 0 ? keep + 1 : PX,
                                N * nin);
    MLQ_HIP(hipGetLastError());
    prof_end(PC_YFWD, nin * (ab_xcols(keep) + ab_kcols(keep)));
  }
  nfwd_ += nin;
  // fused z round trips
  {
    ZArgs<T> za{};
    za.spec = scratch_;
    za.tw = tw_;
    za.sstride = ns_;
    za.keep = keep;
    for (int j = 0; j < pl.nz; ++j) {
      za.zin[j] = pl.zin[j];
      za.zout[j] = pl.zout[j];
      za.op[j] = pl.zop[j];
    }
    constexpr int E = FftShape<N>::E, CZ = zrt_C<T, N>();
    const int ncols = N * PP;
    // band-limited: only the tiles of kept columns (2 keep + 1 kept ky rows of
    // ceil((keep + 1) / CZ) tiles) are launched
    za.kcomp = keep >= 0 && 2 * keep + 1 < N;
    const unsigned gx = za.kcomp ? (unsigned)(2 * keep + 1) * (unsigned)((keep + CZ) / CZ)
                                 : (unsigned)((ncols + CZ - 1) / CZ);
    prof_begin(PC_ZRT);
    hipLaunchKernelGGL((k_zrt<T, N, E, CZ>), dim3(gx, pl.nz), dim3(CZ * (N / E)), 0, st_, za,
                       ncols);
    MLQ_HIP(hipGetLastError());
    prof_end(PC_ZRT, 2 * pl.nz * ab_kcols(keep));
  }
  // inverse y passes (out of place first: they read slots the in-place ones rewrite)
  {
    SArgs<T> ya{};
    ya.spec = scratch_;
    ya.tw = tw_;
    ya.sstride = ns_;
    ya.keep = keep;
    ya.ymap = 1;
    prof_begin(PC_YINV);
    for (int pass = 0; pass < 2; ++pass) {
      int nf = 0;
      for (int p = 0; p < pl.ny; ++p) {
        const int zs = pl.zout[pl.yz[p]];
        if ((pl.yslot[p] == zs) != (pass == 1)) continue;
        ya.yin[nf] = zs;
        ya.yout[nf] = pl.yslot[p];
        ya.yik[nf] = pl.yk[p];
        ++nf;
      }
      if (nf == 0) continue;
      launch_strided<T, N, 1, 0>(svar_[keep >= 0 ? 7 : 3], st_, ya, keep >= 0 ? keep + 1 : PX,
                                 N * nf);
      MLQ_HIP(hipGetLastError());
    }
    prof_end(PC_YINV, pl.ny * (ab_kcols(keep) + ab_xcols(keep)));
  }
  // inverse x pass (1/N^3 folded in)
  {
    XArgs<T> xa{};
    for (int o = 0; o < nout; ++o) {
      xa.rdst[o] = outs[o].dst;
      xa.xin[o] = pl.yslot[pl.oy[o]];
      xa.xik[o] = pl.oxik[o];
    }
    xa.xmap = 1;
    xa.spec = scratch_;
    xa.tw = tw_;
    xa.sstride = ns_;
    xa.nb = nout;
    xa.keep = keep;
    xa.scale = (T)(1.0 / ((double)N * N * N));
    const long long pairs = (long long)nout * N * N / 2;
    prof_begin(PC_XINV);
    if (MLQ_XINV_ZB && keep >= 0 && 8 * keep < 3 * N)
      hipLaunchKernelGGL((k_xinv<T, N, NF, true, true>), dim3((unsigned)(pairs / NF)), dim3(NF * Tf),
                         0, st_, xa);
    else if (keep >= 0)
      hipLaunchKernelGGL((k_xinv<T, N, NF, true>), dim3((unsigned)(pairs / NF)), dim3(NF * Tf), 0,
                         st_, xa);
    else
      hipLaunchKernelGGL((k_xinv<T, N, NF, false>), dim3((unsigned)(pairs / NF)), dim3(NF * Tf), 0,
                         st_, xa);
    MLQ_HIP(hipGetLastError());
    prof_end(PC_XINV, nout * (real_bytes() + ab_xcols(keep)));
  }
  ninv_ += nout;
}

#define MLQ_DISPATCH_N(FN, ...)                                              \
  switch (n_) {                                                              \
    case 8: FN<8>(__VA_ARGS__); break;                                       \
    case 16: FN<16>(__VA_ARGS__); break;                                     \
    case 32: FN<32>(__VA_ARGS__); break;                                     \
    case 64: FN<64>(__VA_ARGS__); break;                                     \
    case 128: FN<128>(__VA_ARGS__); break;                                   \
    case 256: FN<256>(__VA_ARGS__); break;                                   \
    case 512: FN<512>(__VA_ARGS__); break;                                   \
    case 1024: FN<1024>(__VA_ARGS__); break;                                 \
    default: throw Error(MLQ_EINVAL, "unsupported n " + std::to_string(n_)); \
  }

// forward transforms of the P::NB fields a producer computes pointwise
// (k_xfwd_p); false (nothing launched) when the scratch or the LDS cannot
// hold the batch, the caller then materialises the fields.  A band-limited
// batch stores the kept cube of its spectra only: every reader of these
// outputs (band-limited inverses, kept-cube spectral kernels) stays inside it.
template <typename T> template <class P>
bool Engine<T>::fwd_prod(const P& prod, C* const* dst, const SpecOp* ops) {
  if (!fused_ || P::NB > sb_) return false;
  bool done = false;
  MLQ_DISPATCH_N(fwd_prod_n, prod, dst, ops, done);
  return done;
}

template <typename T>
void Engine<T>::fwd(int nb, const T* const* src, C* const* dst, const SpecOp* ops) {
  for (int b0 = 0; b0 < nb; b0 += sb_) {
    const int k = std::min(sb_, nb - b0);
    MLQ_DISPATCH_N(fwd_n, k, src + b0, dst + b0, ops + b0);
  }
}
template <typename T> template <class P>
bool Engine<T>::inv_fwd_prod(const InvSrc* src, const SpecOp* iops, T* const* idst, int ikeep,
                             const P& prod, C* const* dst, const SpecOp* ops) {
  if (!fused_ || P::NI > sb_ || P::NB > sb_) return false;
  bool done = false;
  MLQ_DISPATCH_N(inv_fwd_prod_n, src, iops, idst, ikeep, prod, dst, ops, done);
  return done;
}
template <typename T>
void Engine<T>::inv(int nb, const InvSrc* src, const SpecOp* ops, T* const* dst, int keep) {
  for (int b0 = 0; b0 < nb; b0 += sb_) {
    const int k = std::min(sb_, nb - b0);
    MLQ_DISPATCH_N(inv_n, k, src + b0, ops + b0, dst + b0, keep);
  }
}

template <typename T>
void Engine<T>::mul_dealias(int nb, const T* const* a, const T* const* b, T* const* out,
                            bool dealias, double order) {
  const int mc = mcut_of(order);
  if (!dealias || mask_is_identity(mc)) {
    for (int i = 0; i < nb; ++i) map(F_mul<T>{a[i], b[i], out[i]}, nr_);
    return;
  }
  std::vector<RtOut> o(nb);
  for (int i = 0; i < nb; ++i) o[i] = RtOut{i, sp_mask(mc), out[i]};
  rtrip(nb, a, nb, o.data(), b);
}

template <typename T> void Engine<T>::dealias_dev(T* const* f, int nb, double order) {
  const int mc = mcut_of(order);
  if (mask_is_identity(mc)) return;   // identity mask: the round trip only adds round-off
  std::vector<RtOut> outs(nb);
  for (int b = 0; b < nb; ++b) outs[b] = RtOut{b, sp_mask(mc), f[b]};
  rtrip(nb, f, nb, outs.data());
}

// Round trips in launch groups that fit the transform scratch: inputs are
// added to a group while its slot plan fits; an input whose outputs alone do
// not fit runs one output at a time.  A group's outputs are written by its
// last pass, after every read of its inputs: an output may alias an input of
// its own group (dealias in place), never an input of a later group.
template <typename T>
void Engine<T>::rtrip(int nin, const T* const* src, int nout, const RtOut* outs,
                      const T* const* mul) {
  int keep = -1;
  for (int o = 0; o < nout; ++o) {
    const SpecOp& op = outs[o].op;
    if (!(op.flags & SP_MASK) || op.mcut >= n_ / 2) { keep = -1; break; }
    keep = std::max(keep, op.mcut);
  }
  for (int o = 0; o < nout; ++o)
    if (outs[o].in < 0 || outs[o].in >= nin) throw Error(MLQ_EINVAL, "rtrip: bad input index");
  int i0 = 0;
  while (i0 < nin) {
    // grow the group [i0, i1) while it fits
    int i1 = i0;
    std::vector<RtOut> g;
    while (i1 < nin) {
      std::vector<RtOut> t = g;
      for (int o = 0; o < nout; ++o)
        if (outs[o].in == i1) { RtOut r = outs[o]; r.in = i1 - i0; t.push_back(r); }
      RtPlan pl;
      if ((int)t.size() > MAXB || rtrip_plan(i1 - i0 + 1, (int)t.size(), t.data(), pl) > sb_) break;
      g = std::move(t);
      ++i1;
    }
    if (i1 == i0) {
      // one input whose outputs do not fit together: one output per launch
      for (int o = 0; o < nout; ++o) {
        if (outs[o].in != i0) continue;
        RtOut r = outs[o];
        r.in = 0;
        const T* s = src[i0];
        const T* m = mul ? mul[i0] : nullptr;
        MLQ_DISPATCH_N(rtrip_n, 1, &s, 1, &r, keep, &m);
      }
      i0 += 1;
      continue;
    }
    if (!g.empty())
      MLQ_DISPATCH_N(rtrip_n, i1 - i0, src + i0, (int)g.size(), g.data(), keep,
                     mul ? mul + i0 : nullptr);
    i0 = i1;
  }
}

template <typename T> void Engine<T>::time_fft(int batch, int reps, double* t_ms) {
  batch = std::max(1, std::min(batch, sb_));
  std::vector<RT> r;
  std::vector<ST> s;
  std::vector<T*> rp;
  std::vector<C*> sp;
  for (int b = 0; b < batch; ++b) {
    r.push_back(rtmp()); s.push_back(stmp());
    rp.push_back(r.back().get()); sp.push_back(s.back().get());
    map(F_noise<T>{rp[b], 12345ULL, nr_, b}, nr_);
  }
  std::vector<SpecOp> ops(batch, sp_none());
  std::vector<InvSrc> is(batch);
  for (int b = 0; b < batch; ++b) is[b] = InvSrc{{sp[b], nullptr, nullptr}, {1.0, 0, 0}, 1};
  const bool was = prof_;
  prof_enable(true);
  for (int i = 0; i < reps; ++i) {
    fwd(batch, rp.data(), sp.data(), ops.data());
    inv(batch, is.data(), ops.data(), rp.data());
  }
  prof_flush();
  for (int c = 0; c < 6; ++c) t_ms[c] = prof_count_[c] ? prof_ms_[c] / prof_count_[c] : 0.0;
  t_ms[6] = batch;
  prof_enable(was);
}

// One strided pass alone, in place on the scratch spectra, timed with HIP
// events on the engine stream (ms per launch).
template <typename T> template <int N>
double Engine<T>::time_pass_n(int pass, int var, int batch, int reps, int keep, C* ext) {
  constexpr int PX = N / 2 + 1, PP = spec_pitch<T>(N);
  SArgs<T> sa{};
  sa.spec = scratch_;
  sa.tw = tw_;
  sa.sstride = ns_;
  sa.keep = keep;
  const int ycols = keep >= 0 ? keep + 1 : PX;
  // as in a transform: the forward z pass writes, and the inverse z pass
  // reads, spectra outside the scratch (ext, when given)
  for (int b = 0; b < batch; ++b) {
    sa.src[b][0] = (ext ? ext : scratch_) + b * ns_;
    sa.coef[b][0] = 1.0;
    sa.nsrc[b] = 1;
    sa.dst[b] = (ext ? ext : scratch_) + b * ns_;
    sa.op[b] = sp_none();
  }
  hipEvent_t a, z;
  MLQ_HIP(hipEventCreate(&a));
  MLQ_HIP(hipEventCreate(&z));
  auto run = [&]() {
    switch (pass) {
      case 0: launch_strided<T, N, -1, 0>(var, st_, sa, ycols, N * batch); break;
      case 1: launch_strided<T, N, -1, 2>(var, st_, sa, N * PP, batch); break;
      case 2: launch_strided<T, N, 1, 1>(var, st_, sa, N * PP, batch); break;
      default: launch_strided<T, N, 1, 0>(var, st_, sa, ycols, N * batch); break;
    }
  };
  run();  // warm-up
  MLQ_HIP(hipEventRecord(a, st_));
  for (int i = 0; i < reps; ++i) run();
  MLQ_HIP(hipEventRecord(z, st_));
  MLQ_HIP(hipEventSynchronize(z));
  float ms = 0.f;
  MLQ_HIP(hipEventElapsedTime(&ms, a, z));
  MLQ_HIP(hipEventDestroy(a));
  MLQ_HIP(hipEventDestroy(z));
  return (double)ms / reps;
}

// candidates per pass: 4 tile shapes x {dispatch order, XCD-contiguous order};
// passes 0-3 full transforms, 4-7 band-limited ones (the 2/3 mask's cube)
constexpr int NTILE = 8;
template <typename T> void Engine<T>::autotune(int batch, double* times, int* choice) {
  batch = std::max(1, std::min(batch, sb_));
  // nonzero spectra in the scratch and in the outside spectra the z passes
  // write or read (an in-place z pass measured up to 15% off the real one)
  ST ext = stmp(batch);
  for (int b = 0; b < batch; ++b) {
    map(F_noise<T>{reinterpret_cast<T*>(scratch_ + b * ns_), 777ULL, 2 * ns_, b}, 2 * ns_);
    map(F_noise<T>{reinterpret_cast<T*>(ext.get() + b * ns_), 778ULL, 2 * ns_, b}, 2 * ns_);
  }
  const int mc = mcut_of(2.0);
  for (int slot = 0; slot < 8; ++slot) {
    const int pass = slot & 3;
    const int keep = slot >= 4 && !mask_is_identity(mc) ? mc : -1;
    // two rounds over the candidates, the faster of each candidate's two
    // timings: one short timing per candidate picked tiles 15% apart from
    // run to run
    double tmin[NTILE];
    for (int v = 0; v < NTILE; ++v) tmin[v] = 1e300;
    for (int round = 0; round < 2; ++round) {
      for (int v = 0; v < NTILE; ++v) {
        double t = 0.0;
        MLQ_DISPATCH_N(t = time_pass_n, pass, v, batch, 3, keep, ext.get());
        tmin[v] = std::min(tmin[v], t);
      }
    }
    double best = 1e300;
    int bv = 0;
    for (int v = 0; v < NTILE; ++v) {
      if (times) times[slot * NTILE + v] = tmin[v];
      if (tmin[v] < best) { best = tmin[v]; bv = v; }
    }
    svar_[slot] = bv;
    if (choice) choice[slot] = bv;
  }
}

// column copies in the y or z pattern of tile shape `var` (bits 0-1) and
// order (bit 2), nfields spectra: ms per launch and algorithmic bytes
template <typename T> template <int N>
double Engine<T>::time_colcopy_n(int axis, int var, const C* src, C* dst, int nfields, int reps) {
  constexpr int PX = N / 2 + 1, PP = spec_pitch<T>(N);
  const int ncols = axis == 1 ? PX : N * PP;
  const int planes = axis == 1 ? N * nfields : nfields;
  const int remap = (var >> 2) & 1;
  auto launch = [&](auto s) {
    using S = decltype(s);
    const dim3 grid((ncols + S::C - 1) / S::C, planes), block(S::C * (N / S::E));
    if (axis == 1)
      hipLaunchKernelGGL((k_colcopy<T, N, S::E, S::C, 1>), grid, block, 0, st_, src, dst, ncols, remap);
    else
      hipLaunchKernelGGL((k_colcopy<T, N, S::E, S::C, 2>), grid, block, 0, st_, src, dst, ncols, remap);
  };
  auto run = [&]() {
    switch (var & 3) {
      case 1: launch(SVar<T, N, 1>{}); break;
      case 2: launch(SVar<T, N, 2>{}); break;
      case 3: launch(SVar<T, N, 3>{}); break;
      default: launch(SVar<T, N, 0>{}); break;
    }
    MLQ_HIP(hipGetLastError());
  };
  hipEvent_t a, z;
  MLQ_HIP(hipEventCreate(&a));
  MLQ_HIP(hipEventCreate(&z));
  run();  // warm-up
  MLQ_HIP(hipEventRecord(a, st_));
  for (int i = 0; i < reps; ++i) run();
  MLQ_HIP(hipEventRecord(z, st_));
  MLQ_HIP(hipEventSynchronize(z));
  float ms = 0.f;
  MLQ_HIP(hipEventElapsedTime(&ms, a, z));
  MLQ_HIP(hipEventDestroy(a));
  MLQ_HIP(hipEventDestroy(z));
  return (double)ms / reps;
}

template <typename T>
void Engine<T>::time_column_copy(int axis, int var, int nfields, int reps, double* t_ms,
                                 double* bytes) {
  if ((axis != 1 && axis != 2) || var < 0 || var > 7 || nfields < 1 || reps < 1)
    throw Error(MLQ_EINVAL, "time_column_copy: axis 1 or 2, var 0..7, nfields and reps >= 1");
  const size_t sz = sizeof(C) * (size_t)ns_ * (size_t)nfields;
  C* a = (C*)dmalloc(sz);
  C* b = (C*)dmalloc(sz);
  MLQ_HIP(hipMemsetAsync(a, 0, sz, st_));
  double t = 0.0;
  MLQ_DISPATCH_N(t = time_colcopy_n, axis, var, a, b, nfields, reps);
  dfree(a, sz);
  dfree(b, sz);
  *t_ms = t;
  *bytes = 2.0 * nfields * 2.0 * sizeof(T) * (double)(n_ / 2 + 1) * (double)n_ * (double)n_;
}

template <typename T> void Engine<T>::time_copy(int nfields, int reps, double* t_ms) {
  nfields = std::max(1, nfields);
  RT a = rtmp(nfields), b = rtmp(nfields);
  map(F_fill<T>{a, (T)1}, nr_ * nfields);
  const bool was = prof_;
  prof_enable(true);
  for (int i = 0; i < reps; ++i) map(F_copy<T>{a, b}, nr_ * nfields);
  prof_flush();
  *t_ms = prof_count_[PC_POINTWISE] ? prof_ms_[PC_POINTWISE] / prof_count_[PC_POINTWISE] : 0.0;
  prof_enable(was);
}

}  // namespace mlq
