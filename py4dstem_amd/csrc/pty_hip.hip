// HIP kernels and C-ABI for single-slice GD ptychography on MI355X (gfx950).
//
// Work decomposition (one mini-batch of B positions, frames S x S):
//   k_fused          one work group per position: ifft2(probe_hat * ramp),
//                    periodic object gather, exit wave, fft2, amplitude
//                    projection, ifft2, adjoint products.  The frame lives in
//                    LDS (S=128: 136 KiB of the 160 KiB per CU); a row is
//                    owned by 8 lanes of one wave, so passes along rows are
//                    separated by wave barriers, everything else by
//                    work-group barriers (nine per position).
//   k_gather         one work group per 4 x 64 object tile: float64 sums of
//                    the staged patches in batch order (deterministic).
//   k_object_update  normalised object step.
//   k_probe_partial / k_probe_update   probe step from staged terms.
//   k_constraint_*   per-iteration object constraints; k_probe_com probe
//                    recentring; k_error_reduce per-iteration error.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "pty_core.h"
#include "../../include/pty_hip.h"

using namespace pty;

namespace {

thread_local std::string g_err;

int fail(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}

int hip_check(hipError_t e, const char* what) {
  return e == hipSuccess ? PTY_OK : fail(PTY_E_LAUNCH, "%s: %s", what, hipGetErrorString(e));
}

int check_launch(const char* what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(PTY_E_LAUNCH, "%s: %s", what, hipGetErrorString(e));
  return PTY_OK;
}

__host__ __device__ inline size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

// the kernels index the object and the staged frames with 32-bit element
// offsets, and read the object through a buffer resource of 32-bit byte
// range: the object must stay below 2^28 elements (16384^2; 2 GiB of
// complex64), the staged frames below 2^31 elements
int check_index_range(long long Px, long long Py, long long frames, long long S) {
  const long long lim = 0x7fffffffll;
  if (Px * Py >= (1ll << 28)) return fail(PTY_E_UNSUPPORTED, "object %lldx%lld beyond 2^28 elements", Px, Py);
  if (frames * S * S > lim)
    return fail(PTY_E_UNSUPPORTED, "%lld staged %lldx%lld frames beyond 2^31 elements", frames, S, S);
  return PTY_OK;
}

bool aligned8(const void* p) { return (reinterpret_cast<uintptr_t>(p) & 7u) == 0; }

inline hipStream_t st(pty_stream_t s) { return reinterpret_cast<hipStream_t>(s); }

// LDS operations of one wave execute in order; the fences keep the compiler
// from moving this wave's LDS accesses across the boundary.  Enough between
// two passes along the same lines, whose lanes all belong to one wave.
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// ---------------------------------------------------------------------------
// fused per-position kernel
// ---------------------------------------------------------------------------
template <int S> struct GpuRunner {
  using K = Fused<S>;
  static constexpr int T = Frame<S>::T;
  typename K::Regs& r;
  LdsPtr lds;
  const cf* tw;
  int tid;
  double* red;
  const FusedArgs& a;
  // The empty asm makes the thread id opaque at every phase boundary, so the
  // per-phase LDS/global address arithmetic is recomputed instead of being
  // shared (and kept live) across phases.
  template <class Fn> __device__ __forceinline__ void phase(Fn fn) {
    int t = tid;
    asm volatile("" : "+v"(t));
    __builtin_assume(t >= 0 && t < T);  // lets index math use unsigned shifts
    fn(t, r, lds, tw);
  }
  __device__ __forceinline__ void barrier() { __syncthreads(); }
  __device__ __forceinline__ void wave_barrier() { wave_sync(); }
  __device__ __forceinline__ void reduce_err(int b) {
    double v = r.err;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    if ((tid & 63) == 0) red[tid >> 6] = v;
    __syncthreads();
    if (tid == 0) {
      double s = 0.0;
      for (int w = 0; w < T / 64; ++w) s += red[w];
      a.err[b] = s;
    }
  }
};

template <int S, bool MASK = false, bool POT = false>
__device__ __forceinline__ void fused_body(const FusedArgs& a, int b) {
  __shared__ cf lds_mem[Fused<S>::LDS_TOTAL];
  __shared__ cf tw[Fused<S>::TW_WORDS];
  __shared__ double red[Frame<S>::T / 64];
  const int tid = threadIdx.x;
  // the probe-spectrum loads (independent of the slot) go out first, ahead
  // of the twiddle work and the set-up's slot/position reads
  typename Fused<S>::Regs r;
  Fused<S>::p_load_probe_coalesced(tid, a, r);
  // threads [2S, 3S) are idle in the setup phase; its closing barrier
  // publishes the table
  Fused<S>::twiddles(tid - 2 * S, LdsPtr{tw}, Frame<S>::T - 2 * S);
  GpuRunner<S> run{r, LdsPtr{lds_mem}, tw, tid, red, a};
  fused_position<S, GpuRunner<S>, MASK, POT, true>(run, a, b);
}

template <int S, bool MASK, bool POT = false>
__global__ __launch_bounds__(Frame<S>::T) void k_fused(FusedArgs a) {
  fused_body<S, MASK, POT>(a, a.b0 + (int)blockIdx.x);  // grid covers slots b0..B-1
}

// mixed-state forward/adjoint terms, one work group per position
template <int S>
__global__ __launch_bounds__(Frame<S>::T) void k_mixed(FusedArgs a) {
  __shared__ cf lds_mem[Fused<S>::LDS_TOTAL];
  __shared__ cf tw[Fused<S>::TW_WORDS];
  __shared__ double red[Frame<S>::T / 64];
  const int tid = threadIdx.x;
  Fused<S>::twiddles(tid - 2 * S, LdsPtr{tw}, Frame<S>::T - 2 * S);  // published by the first barrier
  typename Fused<S>::Regs r;
  GpuRunner<S> run{r, LdsPtr{lds_mem}, tw, tid, red, a};
  mixed_position<S>(run, a, a.b0 + (int)blockIdx.x);
}

// projection-set exit-wave update, one work group per position
template <int S>
__global__ __launch_bounds__(Frame<S>::T) void k_projection(FusedArgs a) {
  __shared__ cf lds_mem[Fused<S>::LDS_TOTAL];
  __shared__ cf tw[Fused<S>::TW_WORDS];
  __shared__ double red[Frame<S>::T / 64];
  const int tid = threadIdx.x;
  Fused<S>::twiddles(tid - 2 * S, LdsPtr{tw}, Frame<S>::T - 2 * S);  // published by the setup barrier
  typename Fused<S>::Regs r;
  GpuRunner<S> run{r, LdsPtr{lds_mem}, tw, tid, red, a};
  projection_position<S>(run, a, a.b0 + (int)blockIdx.x);
}

// backward of the probe path, one work group per position
template <int S>
__global__ __launch_bounds__(Frame<S>::T) void k_probe_grad(FusedArgs a) {
  __shared__ cf lds_mem[Fused<S>::LDS_TOTAL];
  __shared__ cf tw[Fused<S>::TW_WORDS];
  const int tid = threadIdx.x;
  Fused<S>::twiddles(tid - 2 * S, LdsPtr{tw}, Frame<S>::T - 2 * S);  // published by the setup barrier
  typename Fused<S>::Regs r;
  GpuRunner<S> run{r, LdsPtr{lds_mem}, tw, tid, nullptr, a};
  probe_grad_position<S>(run, a, a.b0 + (int)blockIdx.x);
}

// forward model only (synthetic data), one work group per position
template <int S>
__global__ __launch_bounds__(Frame<S>::T) void k_simulate(FusedArgs a) {
  __shared__ cf lds_mem[Fused<S>::LDS_TOTAL];
  __shared__ cf tw[Fused<S>::TW_WORDS];
  const int tid = threadIdx.x;
  Fused<S>::twiddles(tid - 2 * S, LdsPtr{tw}, Frame<S>::T - 2 * S);  // published by the setup barrier
  typename Fused<S>::Regs r;
  GpuRunner<S> run{r, LdsPtr{lds_mem}, tw, tid, nullptr, a};
  simulate_position<S>(run, a, a.b0 + (int)blockIdx.x);
}

// ---------------------------------------------------------------------------
// position correction, one work group per position of the mini-batch
// ---------------------------------------------------------------------------
template <int S> struct GpuPosRunner {
  using K = PosCorr<S>;
  static constexpr int T = Frame<S>::T;
  typename K::Regs& r;
  LdsPtr lds;
  const cf* tw;
  int tid;
  double* red;  // [T / 64][POS_SUMS]
  template <class Fn> __device__ __forceinline__ void phase(Fn fn) {
    int t = tid;
    asm volatile("" : "+v"(t));
    __builtin_assume(t >= 0 && t < T);  // lets index math use unsigned shifts
    fn(t, r, lds, tw);
  }
  __device__ __forceinline__ void barrier() { __syncthreads(); }
  __device__ __forceinline__ void wave_barrier() { wave_sync(); }
  // wave sums by butterfly, then thread 0 adds the waves in order
  __device__ __forceinline__ void reduce_sums(double* tot) {
#pragma unroll
    for (int k = 0; k < POS_SUMS; ++k) {
      double v = r.s[k];
#pragma unroll
      for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
      if ((tid & 63) == 0) red[(tid >> 6) * POS_SUMS + k] = v;
    }
    __syncthreads();
    if (tid == 0) {
#pragma unroll
      for (int k = 0; k < POS_SUMS; ++k) {
        double s = 0.0;
        for (int w = 0; w < T / 64; ++w) s += red[w * POS_SUMS + k];
        tot[k] = s;
      }
    }
  }
};

template <int S, bool ACC = false>
__global__ __launch_bounds__(Frame<S>::T) void k_position_correction(PosArgs a) {
  __shared__ cf lds_mem[PosCorr<S>::LDS_TOTAL];
  __shared__ cf tw[Fused<S>::TW_WORDS];
  __shared__ double red[Frame<S>::T / 64 * POS_SUMS];
  const int tid = threadIdx.x;
  Fused<S>::twiddles(tid, LdsPtr{tw}, Frame<S>::T);  // published by the setup barrier
  typename PosCorr<S>::Regs r;
  GpuPosRunner<S> run{r, LdsPtr{lds_mem}, tw, tid, red};
  position_correction_position<S, GpuPosRunner<S>, ACC>(run, a, (int)blockIdx.x);
}

// probe modes: the least-squares sums of slot b from the mode-summed
// accumulators (ptychographic_methods.py:2128-2196 with the mixed-state
// intensity sqrt(sum_k |F_k|^2)^2 of :2736-2740), then the shared update
template <int S>
__global__ __launch_bounds__(Frame<S>::T) void k_position_solve(PosArgs a) {
#pragma clang fp contract(off)
  using F = Frame<S>;
  constexpr int T = F::T, N1 = F::N1;
  __shared__ double red[T / 64 * POS_SUMS];
  const int b = (int)blockIdx.x, tid = threadIdx.x;
  int idx = a.order ? a.order[b] : b;
  idx = ((unsigned)idx < (unsigned)a.N) ? idx : 0;
  const float* A = a.amps + (size_t)idx * S * S + F::row_line(tid) * S + N1 * F::row_q(tid);
  const float* acc = a.acc + (size_t)b * 3 * S * S + tid;
  double s[POS_SUMS] = {0.0, 0.0, 0.0, 0.0, 0.0};
#pragma unroll
  for (int j = 0; j < N1; ++j) {
    const double px = acc[(size_t)j * T];
    const double py = acc[(size_t)S * S + (size_t)j * T];
    const float m = sqrt_rn(acc[(size_t)2 * S * S + (size_t)j * T]);
    const double d = A[j] * A[j] - m * m;  // float32 difference
    s[0] += px * px;
    s[1] += px * py;
    s[2] += py * py;
    s[3] += px * d;
    s[4] += py * d;
  }
#pragma unroll
  for (int k = 0; k < POS_SUMS; ++k) {
    double v = s[k];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    if ((tid & 63) == 0) red[(tid >> 6) * POS_SUMS + k] = v;
  }
  __syncthreads();
  if (tid == 0) {
    double tot[POS_SUMS];
    for (int k = 0; k < POS_SUMS; ++k) {
      double v = 0.0;
      for (int w = 0; w < T / 64; ++w) v += red[w * POS_SUMS + k];
      tot[k] = v;
    }
    position_update(a, b, idx, tot);
  }
}

// positions -= mean(positions) - com (phase/ptychographic_constraints.py:1292-1311)
// with the reference's arithmetic: NumPy's mean over axis 0 of a float32
// [N][2] array adds the rows one after another in float32, then divides by N.
// One thread adds them in that order from LDS-staged chunks (a dependent
// chain of N adds per axis, tens of microseconds for a 128 x 128 scan, once
// per iteration); every thread then applies the float32 shift.
constexpr int COM_CHUNK = 4096;

__global__ __launch_bounds__(1024) void k_positions_com(float* __restrict__ pos, int N, float com_x, float com_y) {
#pragma clang fp contract(off)
  __shared__ float2 stage[COM_CHUNK];
  __shared__ float shift[2];
  const int tid = threadIdx.x;
  float sx = 0.f, sy = 0.f;
  for (int base = 0; base < N; base += COM_CHUNK) {
    const int n = (N - base) < COM_CHUNK ? (N - base) : COM_CHUNK;
    for (int i = tid; i < n; i += 1024) stage[i] = make_float2(pos[2 * (base + i)], pos[2 * (base + i) + 1]);
    __syncthreads();
    if (tid == 0) {
      for (int i = 0; i < n; ++i) {
        const float2 v = stage[i];
        sx += v.x;
        sy += v.y;
      }
    }
    __syncthreads();
  }
  if (tid == 0) {
    // float32 / N, correctly rounded (the float64 quotient rounds once more
    // to the same float32: 53 >= 2 * 24 + 2)
    shift[0] = (float)((double)sx / (double)N) - com_x;
    shift[1] = (float)((double)sy / (double)N) - com_y;
  }
  __syncthreads();
  const float dx = shift[0], dy = shift[1];
  for (int i = tid; i < N; i += 1024) {
    pos[2 * i] -= dx;
    pos[2 * i + 1] -= dy;
  }
}

// sum over b of complex frames [B][n], float64, in 64-slot chunks
__global__ __launch_bounds__(256) void k_sum_frames_partial(const cf* __restrict__ frames, int B, int n,
                                                            double* __restrict__ part) {
  const int pix = blockIdx.x * blockDim.x + threadIdx.x;
  const int chunk = blockIdx.y;
  if (pix >= n) return;
  const int b1 = min(B, (chunk + 1) * 64);
  double re = 0.0, im = 0.0;
  for (int b = chunk * 64; b < b1; ++b) {
    const cf v = frames[(size_t)b * n + pix];
    re += (double)v.x;
    im += (double)v.y;
  }
  part[((size_t)chunk * n + pix) * 2] = re;
  part[((size_t)chunk * n + pix) * 2 + 1] = im;
}

__global__ __launch_bounds__(256) void k_sum_frames_finish(const double* __restrict__ part, int nchunk, int n,
                                                           double* __restrict__ acc) {
  const int pix = blockIdx.x * blockDim.x + threadIdx.x;
  if (pix >= n) return;
  double re = 0.0, im = 0.0;
  for (int c = 0; c < nchunk; ++c) {
    re += part[((size_t)c * n + pix) * 2];
    im += part[((size_t)c * n + pix) * 2 + 1];
  }
  acc[2 * pix] += re;
  acc[2 * pix + 1] += im;
}

__global__ __launch_bounds__(256) void k_fill_f64(double* __restrict__ x, long long n, double v) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    x[i] = v;
}

__global__ __launch_bounds__(256) void k_acc_to_c64(const double* __restrict__ acc, long long n, double scale,
                                                    cf* __restrict__ out) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    out[i] = mkcf((float)(scale * acc[2 * i]), (float)(scale * acc[2 * i + 1]));
}

// ---------------------------------------------------------------------------
// standalone 2-D FFT of one frame per work group
//   forward: natural in -> natural (or kernel-layout) out
//   inverse: natural (or kernel-layout) in -> natural out, scaled 1/S^2
// ---------------------------------------------------------------------------
template <int S, bool INV>
__device__ __forceinline__ void fft2_body(const cf* __restrict__ in, cf* __restrict__ out, int natural_io,
                                          int blk) {
  using F = Frame<S>;
  using K = Fused<S>;
  __shared__ cf lds_mem[F::LDS_WORDS];
  __shared__ cf tw[Fused<S>::TW_WORDS];
  const int tid = threadIdx.x;
  K::twiddles(tid, LdsPtr{tw}, Frame<S>::T);
  __syncthreads();
  const cf* src = in + (size_t)blk * S * S;
  cf* dst = out + (size_t)blk * S * S;
  LdsPtr lds{lds_mem};
  typename K::Regs r;
  if (!INV) {
    const int cc = F::col_line(tid), n2 = F::col_q(tid);
#pragma unroll
    for (int m = 0; m < F::N1; ++m) r.v[m] = src[(n2 + F::N2 * m) * S + cc];
    Pass<S>::template A<false>(n2, r.v, tw);
#pragma unroll
    for (int m = 0; m < F::N1; ++m) lds.st(F::at(n2 + F::N2 * m, cc), r.v[m]);
    __syncthreads();
    K::template p_colB<false>(tid, r, lds, tw);
    __syncthreads();
    K::template p_rowA<false>(tid, r, lds, tw);
    wave_sync();
    const int rr = F::row_line(tid), t = F::row_q(tid);
#pragma unroll
    for (int j = 0; j < F::N1; ++j) r.v[j] = lds.ld(F::at(rr, F::N1 * t + j));
    Pass<S>::template B<false>(t, r.v, tw);
    if (natural_io) {
      const int kr = F::kidx(rr);
#pragma unroll
      for (int j = 0; j < F::N1; ++j) dst[kr * S + F::kidx(F::N1 * t + j)] = r.v[j];
    } else {
      // kernel layout out through the frame: a thread's own 16 elements are
      // 128 contiguous bytes, which would make every store instruction touch
      // 64 cache lines; pieces of two elements per lane are contiguous
#pragma unroll
      for (int j = 0; j < F::N1; ++j) lds.st(F::at(rr, F::N1 * t + j), r.v[j]);
      __syncthreads();
#pragma unroll
      for (int k = 0; k < F::N1 / 2; ++k) {
        const int e = 2 * (k * F::T + tid);
        dst[e] = lds.ld(F::at(e / S, e % S));
        dst[e + 1] = lds.ld(F::at(e / S, e % S) + 1);
      }
    }
  } else {
    const int rr = F::row_line(tid), t = F::row_q(tid);
    if (natural_io) {
      const int kr = F::kidx(rr);
#pragma unroll
      for (int j = 0; j < F::N1; ++j) r.v[j] = src[kr * S + F::kidx(F::N1 * t + j)];
    } else {
      // kernel layout in through the frame, in coalesced two-element pieces
#pragma unroll
      for (int k = 0; k < F::N1 / 2; ++k) {
        const int e = 2 * (k * F::T + tid);
        lds.st(F::at(e / S, e % S), src[e]);
        lds.st(F::at(e / S, e % S) + 1, src[e + 1]);
      }
      __syncthreads();
#pragma unroll
      for (int j = 0; j < F::N1; ++j) r.v[j] = lds.ld(F::at(rr, F::N1 * t + j));
    }
    Pass<S>::template B<true>(t, r.v, tw);
#pragma unroll
    for (int j = 0; j < F::N1; ++j) lds.st(F::at(rr, F::N1 * t + j), r.v[j]);
    wave_sync();
    K::template p_rowA<true>(tid, r, lds, tw);
    __syncthreads();
    K::template p_colB<true>(tid, r, lds, tw);
    __syncthreads();
    K::p_colAinv_to_regs(tid, r, lds, tw);
    const int cc = F::col_line(tid), n2 = F::col_q(tid);
#pragma unroll
    for (int m = 0; m < F::N1; ++m) dst[(n2 + F::N2 * m) * S + cc] = r.v[m];
  }
}

template <int S, bool INV>
__global__ __launch_bounds__(Frame<S>::T) void k_fft2(const cf* __restrict__ in, cf* __restrict__ out,
                                                      int natural_io) {
  fft2_body<S, INV>(in, out, natural_io, blockIdx.x);
}

template <int S>
__global__ void k_permute(const float* __restrict__ in, float* __restrict__ out, long long total,
                          int to_kernel) {
  using F = Frame<S>;
  for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < total;
       e += (long long)gridDim.x * blockDim.x) {
    const long long f = e / (S * S);
    const int rem = (int)(e - f * S * S);
    const int p = rem / S, q = rem % S;
    const long long nat = f * S * S + (long long)F::kidx(p) * S + F::kidx(q);
    if (to_kernel) out[e] = in[nat];
    else out[nat] = in[e];
  }
}

// ---------------------------------------------------------------------------
// scatter-add by tile
// ---------------------------------------------------------------------------
struct GatherGpuRunner {
  Gather::Regs& r;
  GatherSharedView sh;
  int tid;
  template <class Fn> __device__ __forceinline__ void phase(Fn fn) {
    int t = tid;
    asm volatile("" : "+v"(t));
    fn(t, r, sh);
  }
  __device__ __forceinline__ void barrier() { __syncthreads(); }
};

__device__ __forceinline__ void gather_body(const GatherArgs& a, int tile) {
  __shared__ GatherShared sh;
  Gather::Regs r;
  GatherGpuRunner run{r, GatherSharedView{&sh}, (int)threadIdx.x};
  gather_tile(run, a, tile);
}

__global__ __launch_bounds__(GATHER_THREADS) void k_gather(GatherArgs a) { gather_body(a, (int)blockIdx.x); }

__device__ float block_max(float v, float* sh) {
  sh[threadIdx.x] = v;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) sh[threadIdx.x] = fmaxf(sh[threadIdx.x], sh[threadIdx.x + s]);
    __syncthreads();
  }
  float m = sh[0];
  __syncthreads();
  return m;
}

// Correctly rounded float32 square root of x >= 0 (IEEE / NumPy semantics).
// The hardware estimate r is within one ulp.  The midpoints between r and its
// float neighbours have at most 25 significant bits, so their squares are
// exact in float64; comparing x with them selects the correctly rounded
// result (x, a float, can never equal such a square).  x is 0 or normal here.

// 1 / sqrt(1e-16 + (c1 s)^2 + (c2 m)^2) in float32, every operation rounded
// (contraction off, correctly rounded sqrt and division)
__device__ __forceinline__ float overlap_norm(float s, float m, float c1, float c2) {
#pragma clang fp contract(off)
  const float a = c1 * s;
  float t = 1e-16f + a * a;
  const float b = c2 * m;
  t = t + b * b;
  return 1.0f / sqrt_rn(t);
}

__device__ __forceinline__ cf step_add(cf o, cf g, float nrm, float step) {
#pragma clang fp contract(off)
  const float ux = step * (g.x * nrm);
  const float uy = step * (g.y * nrm);
  return mkcf(o.x + ux, o.y + uy);
}

// o + step * g * nrm (gradient descent), or g * nrm when `replace` (the
// projection-set adjoint, ptychographic_methods.py:1910-1998)
__device__ __forceinline__ cf update_value(cf o, cf g, float nrm, float step, int replace) {
#pragma clang fp contract(off)
  return replace ? mkcf(g.x * nrm, g.y * nrm) : step_add(o, g, nrm, step);
}

__device__ __forceinline__ void object_update_body(cf* __restrict__ obj, const cf* __restrict__ grad,
                                                   const float* __restrict__ spsum,
                                                   const float* __restrict__ tilemax, int ntiles,
                                                   long long npix, float step, float c1, float c2, int blk,
                                                   int nblk, int replace = 0) {
  __shared__ float sh[256];
  float m = 0.f;
  for (int i = threadIdx.x; i < ntiles; i += blockDim.x) m = fmaxf(m, tilemax[i]);
  m = block_max(m, sh);
  for (long long p = (long long)blk * blockDim.x + threadIdx.x; p < npix; p += (long long)nblk * blockDim.x) {
    obj[p] = update_value(obj[p], grad[p], overlap_norm(spsum[p], m, c1, c2), step, replace);
  }
}

// potential object (ptychographic_methods.py:1868-1880): V += step * (R * n),
// R = Re(-i conj(exp(i V)) G) = cos V Im G - sin V Re G with G the float64
// scatter of conj(P_shift) chi (conj(exp(i V)) is common to a pixel's terms;
// the reference rounds each term's real part to float32 before its float64
// sum, a difference at float32 rounding level)
__global__ __launch_bounds__(256) void k_object_update_potential(float* __restrict__ V, const cf* __restrict__ grad,
                                                                 const float* __restrict__ spsum,
                                                                 const float* __restrict__ tilemax, int ntiles,
                                                                 long long npix, float step, float c1, float c2) {
#pragma clang fp contract(off)
  __shared__ float sh[256];
  float m = 0.f;
  for (int i = threadIdx.x; i < ntiles; i += blockDim.x) m = fmaxf(m, tilemax[i]);
  m = block_max(m, sh);
  for (long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x; p < npix;
       p += (long long)gridDim.x * blockDim.x) {
    const float v = V[p];
    const cf o = expi(v);
    const cf g = grad[p];
    const float R = (float)((double)o.x * (double)g.y - (double)o.y * (double)g.x);
    const float n = overlap_norm(spsum[p], m, c1, c2);
    V[p] = v + step * (R * n);
  }
}

__global__ __launch_bounds__(256) void k_object_update(cf* __restrict__ obj, const cf* __restrict__ grad,
                                                       const float* __restrict__ spsum,
                                                       const float* __restrict__ tilemax, int ntiles,
                                                       long long npix, float step, float c1, float c2,
                                                       int replace) {
  object_update_body(obj, grad, spsum, tilemax, ntiles, npix, step, c1, c2, blockIdx.x, gridDim.x, replace);
}

// the object step of a mini-batch and, in work group 0 of the same launch,
// fft2 of the just-updated probe for the next mini-batch's fused pass (the
// two are independent; the same arithmetic as k_object_update and k_fft2)
template <int S>
__global__ __launch_bounds__(Frame<S>::T) void k_object_update_fft(cf* __restrict__ obj, const cf* __restrict__ grad,
                                                                  const float* __restrict__ spsum,
                                                                  const float* __restrict__ tilemax, int ntiles,
                                                                  long long npix, float step, float c1, float c2,
                                                                  const cf* __restrict__ probe,
                                                                  cf* __restrict__ phat) {
  if (blockIdx.x == 0) {
    fft2_body<S, false>(probe, phat, 0, 0);
    return;
  }
  constexpr int T = Frame<S>::T;
  __shared__ float sh[T];
  const int tid = threadIdx.x;
  float m = 0.f;
  for (int i = tid; i < ntiles; i += T) m = fmaxf(m, tilemax[i]);
  sh[tid] = m;
  __syncthreads();
  for (int s = T / 2; s > 0; s >>= 1) {
    if (tid < s) sh[tid] = fmaxf(sh[tid], sh[tid + s]);
    __syncthreads();
  }
  m = sh[0];
  const long long nblk = gridDim.x - 1;
  for (long long p = (long long)(blockIdx.x - 1) * T + tid; p < npix; p += nblk * T)
    obj[p] = update_value(obj[p], grad[p], overlap_norm(spsum[p], m, c1, c2), step, 0);
}

// mixed state (ptychographic_methods.py:2945-2975): the mode overlaps summed
// in float32, modes in order, with a per-block maximum
__global__ __launch_bounds__(256) void k_mixed_overlap(const float* __restrict__ spsum, int K, long long npix,
                                                       float* __restrict__ total, float* __restrict__ bmax) {
  __shared__ float sh[256];
  float m = 0.f;
  for (long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x; p < npix;
       p += (long long)gridDim.x * blockDim.x) {
    float s = spsum[p];
    for (int k = 1; k < K; ++k) s = s + spsum[(size_t)k * npix + p];
    total[p] = s;
    m = fmaxf(m, s);
  }
  m = block_max(m, sh);
  if (threadIdx.x == 0) bmax[blockIdx.x] = m;
}

// object += (sum_k step * grad_k) * norm(total), the reference's float32 order
__global__ __launch_bounds__(256) void k_mixed_object_update(cf* __restrict__ obj, const cf* __restrict__ grad,
                                                             int K, const float* __restrict__ total,
                                                             const float* __restrict__ bmax, int nblk,
                                                             long long npix, float step, float c1, float c2) {
#pragma clang fp contract(off)
  __shared__ float sh[256];
  float m = 0.f;
  for (int i = threadIdx.x; i < nblk; i += blockDim.x) m = fmaxf(m, bmax[i]);
  m = block_max(m, sh);
  for (long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x; p < npix;
       p += (long long)gridDim.x * blockDim.x) {
    float ux = 0.f, uy = 0.f;
    for (int k = 0; k < K; ++k) {
      const cf g = grad[(size_t)k * npix + p];
      ux = ux + step * g.x;
      uy = uy + step * g.y;
    }
    const float n = overlap_norm(total[p], m, c1, c2);
    const cf o = obj[p];
    obj[p] = mkcf(o.x + ux * n, o.y + uy * n);
  }
}

// ---------------------------------------------------------------------------
// probe update: partial sums over chunks of the batch, then one work group
// ---------------------------------------------------------------------------
#ifndef PTY_PROBE_CHUNK
#define PTY_PROBE_CHUNK 64
#endif
constexpr int PROBE_CHUNK = PTY_PROBE_CHUNK;  // batch slots per partial sum (a multiple of 16)
// staged launches start at multiples of 64 slots (stage_size) and index the
// partial sums by s0 / PROBE_CHUNK: a chunk must not straddle a stage
static_assert(PROBE_CHUNK % 16 == 0 && 64 % PROBE_CHUNK == 0, "probe chunks must tile the 64-slot stages");

struct ProbePartialArgs {
  const cf* chi;
  const cf* object;        // complex object, or null with object_pot
  const float* positions;
  const int* order;
  int N, B, S, Px, Py;
  double* part;  // [nchunk][S*S][3] (re, im, |O|^2)
  const float* object_pot = nullptr;  // potential V: the patch is exp(i V)
};

__device__ __forceinline__ void probe_partial_body(const ProbePartialArgs& a, int blk, int chunk) {
#if defined(__clang__)
#pragma clang fp contract(off)
#endif
  __shared__ int anchor[2 * PROBE_CHUNK];
  const int S = a.S;
  const int b0 = chunk * PROBE_CHUNK;
  if (b0 >= a.B) return;  // uniform over the work group
  const int nb = min(a.B - b0, PROBE_CHUNK);
  // the chunk's rounded patch anchors, once per work group: the loop below
  // then issues only independent object and chi loads
  for (int t = threadIdx.x; t < nb; t += blockDim.x) {
    int idx = a.order[b0 + t];
    idx = ((unsigned)idx < (unsigned)a.N) ? idx : 0;
    anchor[2 * t] = wrapi_fast((int)rintf(a.positions[2 * idx]), a.Px);
    anchor[2 * t + 1] = wrapi_fast((int)rintf(a.positions[2 * idx + 1]), a.Py);
  }
  __syncthreads();
  const int pix = blk * blockDim.x + threadIdx.x;
  if (pix >= S * S) return;
  const int i = pix / S, j = pix % S;
  const int fi = i < S / 2 ? i : i - S;
  const int fj = j < S / 2 ? j : j - S;
  const cf* chi = a.chi + (size_t)b0 * S * S + pix;
  double re = 0.0, im = 0.0, so = 0.0;
  // anchors in the object and |fi|, |fj| <= S/2 <= Px/2, Py/2: one
  // conditional wrap per side; offsets fit 32 bits (object < 2^31 elements)
#pragma unroll 8
  for (int k = 0; k < nb; ++k) {  // batch order
    const int x0 = anchor[2 * k], y0 = anchor[2 * k + 1];
    const unsigned off = (unsigned)(wrap1(x0 + fi, a.Px) * a.Py + wrap1(y0 + fj, a.Py));
    const cf o = a.object_pot ? expi(elem(a.object_pot, off)) : elem(a.object, off);
    const float ao = hypot_f(o.x, o.y);
    so += (double)(ao * ao);
    const cf g = cconjmul_rn(o, elem(chi, (unsigned)(k * S * S)));
    re += (double)g.x;
    im += (double)g.y;
  }
  double* p = a.part + ((size_t)chunk * S * S + pix) * 3;
  p[0] = re;
  p[1] = im;
  p[2] = so;
}

__global__ __launch_bounds__(256) void k_probe_partial(ProbePartialArgs a) {
  probe_partial_body(a, blockIdx.x, blockIdx.y);
}

// the scatter of a stage (work groups [0, ntiles)) and its probe partial
// sums (the rest, pixel block fastest) in one launch: both only read the
// fused pass's staged terms
__global__ __launch_bounds__(256) void k_gather_probe(GatherArgs g, ProbePartialArgs pa, int ntiles,
                                                      int pix_blocks) {
  const int b = (int)blockIdx.x;
  if (b < ntiles) gather_body(g, b);
  else probe_partial_body(pa, (b - ntiles) % pix_blocks, (b - ntiles) / pix_blocks);
}

// per-pixel sum of the chunk partials: PROBE_REDUCE_GROUPS contiguous chunk
// ranges summed side by side, then added in range order; max of sum |O|^2
// per 256 pixels.  Work group: 256 pixels x PROBE_REDUCE_GROUPS threads.
constexpr int PROBE_REDUCE_GROUPS = 4;
constexpr int PROBE_REDUCE_THREADS = 256 * PROBE_REDUCE_GROUPS;

__device__ __forceinline__ void probe_reduce_body(const double* __restrict__ part, int nchunk, int npix,
                                                  cf* __restrict__ gsum, float* __restrict__ osum,
                                                  float* __restrict__ bmax, int blk) {
  constexpr int G = PROBE_REDUCE_GROUPS;
  __shared__ double acc[G - 1][3][256];
  __shared__ float sh[256];
  const int lane = threadIdx.x & 255, g = threadIdx.x >> 8;
  const int pix = blk * 256 + lane;
  const int c0 = nchunk * g / G, c1 = nchunk * (g + 1) / G;
  double re = 0.0, im = 0.0, s = 0.0;
  if (pix < npix) {
#pragma unroll 4
    for (int c = c0; c < c1; ++c) {
      const double* p = part + ((size_t)c * npix + pix) * 3;
      re += p[0];
      im += p[1];
      s += p[2];
    }
  }
  if (g > 0) {
    acc[g - 1][0][lane] = re;
    acc[g - 1][1][lane] = im;
    acc[g - 1][2][lane] = s;
  }
  __syncthreads();
  if (g == 0) {
    for (int h = 0; h < G - 1; ++h) {
      re += acc[h][0][lane];
      im += acc[h][1][lane];
      s += acc[h][2][lane];
    }
    float so = 0.f;
    if (pix < npix) {
      gsum[pix] = mkcf((float)re, (float)im);
      so = (float)s;
      osum[pix] = so;
    }
    sh[lane] = so;
  }
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if ((int)threadIdx.x < st) sh[threadIdx.x] = fmaxf(sh[threadIdx.x], sh[threadIdx.x + st]);
    __syncthreads();
  }
  if (threadIdx.x == 0) bmax[blk] = sh[0];
}

__global__ __launch_bounds__(PROBE_REDUCE_THREADS) void k_probe_reduce(const double* __restrict__ part, int nchunk,
                                                      int npix, cf* __restrict__ gsum,
                                                      float* __restrict__ osum,
                                                      float* __restrict__ bmax) {
  probe_reduce_body(part, nchunk, npix, gsum, osum, bmax, blockIdx.x);
}

__device__ __forceinline__ void probe_apply_body(cf* __restrict__ probe, const cf* __restrict__ gsum,
                                                 const float* __restrict__ osum, const float* __restrict__ bmax,
                                                 int nblk, int npix, float step, float c1, float c2, int blk,
                                                 int replace = 0) {
  __shared__ float sh[256];
  float m = 0.f;
  for (int i = threadIdx.x; i < nblk; i += blockDim.x) m = fmaxf(m, bmax[i]);
  m = block_max(m, sh);
  const int pix = blk * blockDim.x + threadIdx.x;
  if (pix < npix) probe[pix] = update_value(probe[pix], gsum[pix], overlap_norm(osum[pix], m, c1, c2), step, replace);
}

__global__ __launch_bounds__(256) void k_probe_apply(cf* __restrict__ probe, const cf* __restrict__ gsum,
                                                     const float* __restrict__ osum,
                                                     const float* __restrict__ bmax, int nblk, int npix,
                                                     float step, float c1, float c2, int replace) {
  probe_apply_body(probe, gsum, osum, bmax, nblk, npix, step, c1, c2, blockIdx.x, replace);
}

// ---------------------------------------------------------------------------
// constraints
// ---------------------------------------------------------------------------
constexpr int CONSTRAINT_BLOCKS = 256;

// masked sum of the phase (complex object) or of the potential (pot != null)
__device__ __forceinline__ void constraint_partial_body(const cf* __restrict__ obj,
                                                        const unsigned char* __restrict__ mask, long long npix,
                                                        double* part, int blk, int nblk,
                                                        const float* __restrict__ pot = nullptr) {
  __shared__ double ssum[256];
  __shared__ double scnt[256];
  double s = 0.0, c = 0.0;
  for (long long p = (long long)blk * blockDim.x + threadIdx.x; p < npix; p += (long long)nblk * blockDim.x) {
    if (mask[p]) {
      if (pot) {
        s += (double)pot[p];
      } else {
        const cf o = obj[p];
        s += (double)atan2f(o.y, o.x);
      }
      c += 1.0;
    }
  }
  ssum[threadIdx.x] = s;
  scnt[threadIdx.x] = c;
  __syncthreads();
  for (int k = blockDim.x / 2; k > 0; k >>= 1) {
    if ((int)threadIdx.x < k) {
      ssum[threadIdx.x] += ssum[threadIdx.x + k];
      scnt[threadIdx.x] += scnt[threadIdx.x + k];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    part[2 * blk] = ssum[0];
    part[2 * blk + 1] = scnt[0];
  }
}

__global__ __launch_bounds__(256) void k_constraint_partial(const cf* __restrict__ obj,
                                                            const unsigned char* __restrict__ mask,
                                                            long long npix, double* part,
                                                            const float* __restrict__ pot) {
  constraint_partial_body(obj, mask, npix, part, blockIdx.x, gridDim.x, pot);
}

__device__ __forceinline__ void constraint_apply_body(cf* __restrict__ obj, long long npix,
                                                      const double* __restrict__ part, int nparts, int use_mask,
                                                      float shrinkage_rad, int pure_phase, int blk, int nblk) {
  __shared__ float shrink_sh;
  if (threadIdx.x == 0) {
    float shrink = shrinkage_rad;
    if (use_mask) {
      double s = 0.0, c = 0.0;
      for (int k = 0; k < nparts; ++k) {
        s += part[2 * k];
        c += part[2 * k + 1];
      }
      shrink = shrink + (float)(s / c);
    }
    shrink_sh = shrink;
  }
  __syncthreads();
  const float shrink = shrink_sh;
  const bool do_shrink = use_mask || shrinkage_rad > 0.f;
  for (long long p = (long long)blk * blockDim.x + threadIdx.x; p < npix; p += (long long)nblk * blockDim.x) {
    cf o = obj[p];
    if (do_shrink) {
      const float ph = atan2f(o.y, o.x) - shrink;
      const float am = hypot_f(o.x, o.y);
      float sn, cs;
      sincosf(ph, &sn, &cs);
      o = mkcf(am * cs, am * sn);
    }
    const float ph = atan2f(o.y, o.x);
    const float am = pure_phase ? 1.0f : fminf(hypot_f(o.x, o.y), 1.0f);
    float sn, cs;
    sincosf(ph, &sn, &cs);
    obj[p] = mkcf(am * cs, am * sn);
  }
}

__global__ __launch_bounds__(256) void k_constraint_apply(cf* __restrict__ obj, long long npix,
                                                          const double* __restrict__ part, int nparts,
                                                          int use_mask, float shrinkage_rad,
                                                          int pure_phase) {
  constraint_apply_body(obj, npix, part, nparts, use_mask, shrinkage_rad, pure_phase, blockIdx.x, gridDim.x);
}

// potential object (ptychographic_constraints.py:60-118, 447-462): V minus
// (shrinkage_rad + masked mean of V) when shrinking, then clip at zero when
// object_positivity; NaN stays NaN as under numpy clip
__global__ __launch_bounds__(256) void k_constraint_apply_potential(float* __restrict__ V, long long npix,
                                                                    const double* __restrict__ part, int nparts,
                                                                    int use_mask, float shrinkage_rad,
                                                                    int positivity) {
  __shared__ float shrink_sh;
  if (threadIdx.x == 0) {
    float shrink = shrinkage_rad;
    if (use_mask) {
      double s = 0.0, c = 0.0;
      for (int k = 0; k < nparts; ++k) {
        s += part[2 * k];
        c += part[2 * k + 1];
      }
      shrink = shrink + (float)(s / c);
    }
    shrink_sh = shrink;
  }
  __syncthreads();
  const float shrink = shrink_sh;
  const bool do_shrink = use_mask || shrinkage_rad > 0.f;
  for (long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x; p < npix;
       p += (long long)gridDim.x * blockDim.x) {
    float v = V[p];
    if (do_shrink) v = v - shrink;
    if (positivity) v = v < 0.f ? 0.f : v;
    V[p] = v;
  }
}

// probe centre of mass (corner-centred coordinates) and Fourier recentring
template <int S> __device__ __forceinline__ void probe_com_body(cf* __restrict__ probe) {
  using F = Frame<S>;
  using K = Fused<S>;
  constexpr int T = F::T;
  __shared__ cf lds_mem[F::LDS_WORDS];
  __shared__ cf tw[Fused<S>::TW_WORDS];
  __shared__ double red[3][T / 64];
  __shared__ float shift[2];
  const int tid = threadIdx.x;
  K::twiddles(tid, LdsPtr{tw}, Frame<S>::T);
  LdsPtr lds{lds_mem};
  typename K::Regs r;
  const int cc = F::col_line(tid), n2 = F::col_q(tid);
  double sI = 0.0, sx = 0.0, sy = 0.0;
#pragma unroll
  for (int m = 0; m < F::N1; ++m) {
    const int i = n2 + F::N2 * m;
    r.v[m] = probe[i * S + cc];
    const float a = hypot_f(r.v[m].x, r.v[m].y);
    const double I = (double)(a * a);
    sI += I;
    sx += (double)F::fidx(i) * I;
    sy += (double)F::fidx(cc) * I;
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    sI += __shfl_xor(sI, off);
    sx += __shfl_xor(sx, off);
    sy += __shfl_xor(sy, off);
  }
  if ((tid & 63) == 0) {
    red[0][tid >> 6] = sI;
    red[1][tid >> 6] = sx;
    red[2][tid >> 6] = sy;
  }
  __syncthreads();
  if (tid == 0) {
    double a = 0, b = 0, c = 0;
    for (int w = 0; w < T / 64; ++w) { a += red[0][w]; b += red[1][w]; c += red[2][w]; }
    shift[0] = (float)(-(b / a));
    shift[1] = (float)(-(c / a));
  }
  // forward fft2 (columns then rows) with the ramp applied in kernel layout
  Pass<S>::template A<false>(n2, r.v, tw);
#pragma unroll
  for (int m = 0; m < F::N1; ++m) lds.st(F::at(n2 + F::N2 * m, cc), r.v[m]);
  __syncthreads();
  K::template p_colB<false>(tid, r, lds, tw);
  __syncthreads();
  K::template p_rowA<false>(tid, r, lds, tw);
  __syncthreads();
  {
    const int rr = F::row_line(tid), t = F::row_q(tid);
#pragma unroll
    for (int j = 0; j < F::N1; ++j) r.v[j] = lds.ld(F::at(rr, F::N1 * t + j));
    Pass<S>::template B<false>(t, r.v, tw);
    const cf ex = K::ramp1(F::kidx(rr), shift[0]);
#pragma unroll
    for (int j = 0; j < F::N1; ++j) {
      const cf e = cmul(ex, K::ramp1(F::kidx(F::N1 * t + j), shift[1]));
      r.v[j] = cmul(r.v[j], e);
    }
    Pass<S>::template B<true>(t, r.v, tw);
#pragma unroll
    for (int j = 0; j < F::N1; ++j) lds.st(F::at(rr, F::N1 * t + j), r.v[j]);
  }
  __syncthreads();
  K::template p_rowA<true>(tid, r, lds, tw);
  __syncthreads();
  K::template p_colB<true>(tid, r, lds, tw);
  __syncthreads();
  K::p_colAinv_to_regs(tid, r, lds, tw);
#pragma unroll
  for (int m = 0; m < F::N1; ++m) probe[(n2 + F::N2 * m) * S + cc] = r.v[m];
}

template <int S>
__global__ __launch_bounds__(Frame<S>::T) void k_probe_com(cf* __restrict__ probe) {
  probe_com_body<S>(probe);
}

// mixed state: the centre of mass of mode 0's intensity (one work group;
// the same float64 sums as probe_com_body), then every mode shifted by it
template <int S>
__global__ __launch_bounds__(Frame<S>::T) void k_probe_com_measure(const cf* __restrict__ probe,
                                                                   float* __restrict__ shift) {
  using F = Frame<S>;
  constexpr int T = F::T;
  __shared__ double red[3][T / 64];
  const int tid = threadIdx.x;
  const int cc = F::col_line(tid), n2 = F::col_q(tid);
  double sI = 0.0, sx = 0.0, sy = 0.0;
#pragma unroll
  for (int m = 0; m < F::N1; ++m) {
    const int i = n2 + F::N2 * m;
    const cf v = probe[i * S + cc];
    const float a = hypot_f(v.x, v.y);
    const double I = (double)(a * a);
    sI += I;
    sx += (double)F::fidx(i) * I;
    sy += (double)F::fidx(cc) * I;
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    sI += __shfl_xor(sI, off);
    sx += __shfl_xor(sx, off);
    sy += __shfl_xor(sy, off);
  }
  if ((tid & 63) == 0) {
    red[0][tid >> 6] = sI;
    red[1][tid >> 6] = sx;
    red[2][tid >> 6] = sy;
  }
  __syncthreads();
  if (tid == 0) {
    double a = 0, b = 0, c = 0;
    for (int w = 0; w < T / 64; ++w) { a += red[0][w]; b += red[1][w]; c += red[2][w]; }
    shift[0] = (float)(-(b / a));
    shift[1] = (float)(-(c / a));
  }
}

template <int S>
__global__ __launch_bounds__(Frame<S>::T) void k_probe_shift_modes(cf* __restrict__ probe,
                                                                   const float* __restrict__ shift) {
  using F = Frame<S>;
  using K = Fused<S>;
  __shared__ cf lds_mem[F::LDS_WORDS];
  __shared__ cf tw[Fused<S>::TW_WORDS];
  const int tid = threadIdx.x;
  K::twiddles(tid, LdsPtr{tw}, Frame<S>::T);
  LdsPtr lds{lds_mem};
  typename K::Regs r;
  cf* pk = probe + (size_t)blockIdx.x * S * S;
  const float s0 = shift[0], s1 = shift[1];
  const int cc = F::col_line(tid), n2 = F::col_q(tid);
#pragma unroll
  for (int m = 0; m < F::N1; ++m) r.v[m] = pk[(n2 + F::N2 * m) * S + cc];
  __syncthreads();  // twiddle table
  Pass<S>::template A<false>(n2, r.v, tw);
#pragma unroll
  for (int m = 0; m < F::N1; ++m) lds.st(F::at(n2 + F::N2 * m, cc), r.v[m]);
  __syncthreads();
  K::template p_colB<false>(tid, r, lds, tw);
  __syncthreads();
  K::template p_rowA<false>(tid, r, lds, tw);
  __syncthreads();
  {
    const int rr = F::row_line(tid), t = F::row_q(tid);
#pragma unroll
    for (int j = 0; j < F::N1; ++j) r.v[j] = lds.ld(F::at(rr, F::N1 * t + j));
    Pass<S>::template B<false>(t, r.v, tw);
    const cf ex = K::ramp1(F::kidx(rr), s0);
#pragma unroll
    for (int j = 0; j < F::N1; ++j) {
      const cf e = cmul(ex, K::ramp1(F::kidx(F::N1 * t + j), s1));
      r.v[j] = cmul(r.v[j], e);
    }
    Pass<S>::template B<true>(t, r.v, tw);
#pragma unroll
    for (int j = 0; j < F::N1; ++j) lds.st(F::at(rr, F::N1 * t + j), r.v[j]);
  }
  __syncthreads();
  K::template p_rowA<true>(tid, r, lds, tw);
  __syncthreads();
  K::template p_colB<true>(tid, r, lds, tw);
  __syncthreads();
  K::p_colAinv_to_regs(tid, r, lds, tw);
#pragma unroll
  for (int m = 0; m < F::N1; ++m) pk[(n2 + F::N2 * m) * S + cc] = r.v[m];
}

// per-iteration error: fixed-order float64 sum, normalised
__device__ __forceinline__ void error_reduce_body(const double* __restrict__ err, int n, double norm,
                                                  double* __restrict__ out) {
  __shared__ double sh[1024];
  double s = 0.0;
  for (int i = threadIdx.x; i < n; i += blockDim.x) s += err[i];
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int k = blockDim.x / 2; k > 0; k >>= 1) {
    if ((int)threadIdx.x < k) sh[threadIdx.x] += sh[threadIdx.x + k];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] = sh[0] / norm;
}

__global__ __launch_bounds__(1024) void k_error_reduce(const double* __restrict__ err, int n,
                                                       double norm, double* __restrict__ out) {
  error_reduce_body(err, n, norm, out);
}

// sum over the work group of a per-thread float64 value (thread 0 holds it
// after the call; every thread must call it)
template <int T> __device__ __forceinline__ double block_sum_f64(double v, double* red) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  double s = 0.0;
  if (threadIdx.x == 0)
    for (int w = 0; w < T / 64; ++w) s += red[w];
  __syncthreads();
  return s;
}

// fix_probe_aperture (ptychographic_constraints.py:1004-1033): the probe's
// Fourier amplitude replaced by the initial aperture (kernel layout), its
// Fourier phase kept, the real-space intensity restored; one work group
template <int S>
__global__ __launch_bounds__(Frame<S>::T) void k_probe_aperture(cf* __restrict__ probe,
                                                                const float* __restrict__ aperture) {
#pragma clang fp contract(off)
  using F = Frame<S>;
  using K = Fused<S>;
  constexpr int T = F::T;
  __shared__ cf lds_mem[F::LDS_WORDS];
  __shared__ cf tw[Fused<S>::TW_WORDS];
  __shared__ double red[T / 64];
  __shared__ float norm;
  const int tid = threadIdx.x;
  K::twiddles(tid, LdsPtr{tw}, Frame<S>::T);
  LdsPtr lds{lds_mem};
  typename K::Regs r;
  const int cc = F::col_line(tid), n2 = F::col_q(tid);
  double s = 0.0;
#pragma unroll
  for (int m = 0; m < F::N1; ++m) {
    r.v[m] = probe[(n2 + F::N2 * m) * S + cc];
    const float a = hypot_f(r.v[m].x, r.v[m].y);
    s += (double)(a * a);
  }
  const double s0 = block_sum_f64<T>(s, red);  // also publishes the twiddles
  Pass<S>::template A<false>(n2, r.v, tw);
#pragma unroll
  for (int m = 0; m < F::N1; ++m) lds.st(F::at(n2 + F::N2 * m, cc), r.v[m]);
  __syncthreads();
  K::template p_colB<false>(tid, r, lds, tw);
  __syncthreads();
  K::template p_rowA<false>(tid, r, lds, tw);
  __syncthreads();
  {
    const int rr = F::row_line(tid), t = F::row_q(tid);
#pragma unroll
    for (int j = 0; j < F::N1; ++j) r.v[j] = lds.ld(F::at(rr, F::N1 * t + j));
    Pass<S>::template B<false>(t, r.v, tw);
    const float* ap = aperture + rr * S + F::N1 * t;
#pragma unroll
    for (int j = 0; j < F::N1; ++j) {
      const cf f = r.v[j];
      const float m2 = cabs2(f);
      const cf u = m2 != 0.f ? cscale(f, frsq(m2)) : mkcf(1.f, 0.f);  // exp(i angle F); NaN stays NaN
      r.v[j] = cscale(u, ap[j]);
    }
    Pass<S>::template B<true>(t, r.v, tw);
#pragma unroll
    for (int j = 0; j < F::N1; ++j) lds.st(F::at(rr, F::N1 * t + j), r.v[j]);
  }
  __syncthreads();
  K::template p_rowA<true>(tid, r, lds, tw);
  __syncthreads();
  K::template p_colB<true>(tid, r, lds, tw);
  __syncthreads();
  K::p_colAinv_to_regs(tid, r, lds, tw);
  s = 0.0;
#pragma unroll
  for (int m = 0; m < F::N1; ++m) {
    const float a = hypot_f(r.v[m].x, r.v[m].y);
    s += (double)(a * a);
  }
  const double s1 = block_sum_f64<T>(s, red);
  if (tid == 0) norm = sqrt_rn((float)s0 / (float)s1);
  __syncthreads();
  const float n = norm;
#pragma unroll
  for (int m = 0; m < F::N1; ++m) probe[(n2 + F::N2 * m) * S + cc] = cscale(r.v[m], n);
}

// constrain_probe_amplitude (ptychographic_constraints.py:920-957): real-space
// top-hat 0.5 (1 - erf(sigma r / (1 - r^2))), r = |q| - radius, in float64 as
// the reference evaluates it, intensity restored; one work group per mode
template <int S> __device__ __forceinline__ double tophat_mask(int i, int c, double radius, double sigma) {
  using F = Frame<S>;
  const double x = (double)F::fidx(i) / S, y = (double)F::fidx(c) / S;
  const double r = sqrt(x * x + y * y) - radius;
  return 0.5 * (1.0 - erf(sigma * r / (1.0 - r * r)));
}

template <int S>
__global__ __launch_bounds__(Frame<S>::T) void k_probe_amplitude(cf* __restrict__ probe, double radius,
                                                                 double width) {
  using F = Frame<S>;
  constexpr int T = F::T;
  __shared__ double red[T / 64];
  __shared__ double norm;
  const int tid = threadIdx.x;
  cf* pk = probe + (size_t)blockIdx.x * S * S;
  const int cc = F::col_line(tid), n2 = F::col_q(tid);
  const double sigma = 1.7724538509055160273 / width;  // sqrt(pi) / width
  double s = 0.0, s1 = 0.0;
  for (int m = 0; m < F::N1; ++m) {  // sums first; the mask is recomputed below
    const int i = n2 + F::N2 * m;
    const cf v = pk[i * S + cc];
    const float a = hypot_f(v.x, v.y);
    s += (double)(a * a);
    const double mk = tophat_mask<S>(i, cc, radius, sigma);
    const double ux = (double)v.x * mk, uy = (double)v.y * mk;
    s1 += ux * ux + uy * uy;
  }
  const double s0 = block_sum_f64<T>(s, red);
  const double t1 = block_sum_f64<T>(s1, red);
  if (tid == 0) norm = sqrt((double)(float)s0 / t1);
  __syncthreads();
  const double n = norm;
  for (int m = 0; m < F::N1; ++m) {
    const int i = n2 + F::N2 * m;
    const cf v = pk[i * S + cc];
    const double mk = tophat_mask<S>(i, cc, radius, sigma) * n;
    pk[i * S + cc] = mkcf((float)((double)v.x * mk), (float)((double)v.y * mk));
  }
}

// ---------------------------------------------------------------------------
// launch helpers
// ---------------------------------------------------------------------------
#include "pty_256.inc"

// frames whose transforms run in one work group with the frame in LDS
bool frame_supported(int S) { return S == 32 || S == 64 || S == 128; }
// frames of the gradient-descent path: also 256 (line passes, pty_256.inc)
bool frame_supported_gd(int S) { return frame_supported(S) || S == 256; }

int launch_fused(const FusedArgs& a, int S, hipStream_t stream) {
  const int grid = a.B - a.b0;
  if (grid <= 0) return PTY_OK;
  if (S == 256) return launch_fused256(a, stream);
  if (a.object_pot) {  // potential object (no detector mask: the caller rejects the pair)
    if (a.fmask) return fail(PTY_E_UNSUPPORTED, "detector masks with potential objects are not supported");
    if (S == 32) hipLaunchKernelGGL((k_fused<32, false, true>), dim3(grid), dim3(Frame<32>::T), 0, stream, a);
    else if (S == 64) hipLaunchKernelGGL((k_fused<64, false, true>), dim3(grid), dim3(Frame<64>::T), 0, stream, a);
    else hipLaunchKernelGGL((k_fused<128, false, true>), dim3(grid), dim3(Frame<128>::T), 0, stream, a);
  } else if (a.fmask) {
    if (S == 32) hipLaunchKernelGGL((k_fused<32, true>), dim3(grid), dim3(Frame<32>::T), 0, stream, a);
    else if (S == 64) hipLaunchKernelGGL((k_fused<64, true>), dim3(grid), dim3(Frame<64>::T), 0, stream, a);
    else hipLaunchKernelGGL((k_fused<128, true>), dim3(grid), dim3(Frame<128>::T), 0, stream, a);
  } else {
    if (S == 32) hipLaunchKernelGGL((k_fused<32, false>), dim3(grid), dim3(Frame<32>::T), 0, stream, a);
    else if (S == 64) hipLaunchKernelGGL((k_fused<64, false>), dim3(grid), dim3(Frame<64>::T), 0, stream, a);
    else hipLaunchKernelGGL((k_fused<128, false>), dim3(grid), dim3(Frame<128>::T), 0, stream, a);
  }
  return check_launch("k_fused");
}

int launch_mixed(const FusedArgs& a, int S, hipStream_t stream) {
  const int grid = a.B - a.b0;
  if (grid <= 0) return PTY_OK;
  if (S == 32) hipLaunchKernelGGL(k_mixed<32>, dim3(grid), dim3(Frame<32>::T), 0, stream, a);
  else if (S == 64) hipLaunchKernelGGL(k_mixed<64>, dim3(grid), dim3(Frame<64>::T), 0, stream, a);
  else hipLaunchKernelGGL(k_mixed<128>, dim3(grid), dim3(Frame<128>::T), 0, stream, a);
  return check_launch("k_mixed");
}

int launch_projection(const FusedArgs& a, int S, hipStream_t stream) {
  const int grid = a.B - a.b0;
  if (grid <= 0) return PTY_OK;
  if (S == 32) hipLaunchKernelGGL(k_projection<32>, dim3(grid), dim3(Frame<32>::T), 0, stream, a);
  else if (S == 64) hipLaunchKernelGGL(k_projection<64>, dim3(grid), dim3(Frame<64>::T), 0, stream, a);
  else hipLaunchKernelGGL(k_projection<128>, dim3(grid), dim3(Frame<128>::T), 0, stream, a);
  return check_launch("k_projection");
}

int launch_fft(const cf* in, cf* out, int batch, int S, bool inv, int natural, hipStream_t stream) {
  if (batch <= 0) return PTY_OK;
  if (S == 256) return launch_fft256(in, out, batch, inv, stream);  // natural order in and out
  if (S == 32) {
    if (inv) hipLaunchKernelGGL((k_fft2<32, true>), dim3(batch), dim3(Frame<32>::T), 0, stream, in, out, natural);
    else hipLaunchKernelGGL((k_fft2<32, false>), dim3(batch), dim3(Frame<32>::T), 0, stream, in, out, natural);
  } else if (S == 64) {
    if (inv) hipLaunchKernelGGL((k_fft2<64, true>), dim3(batch), dim3(Frame<64>::T), 0, stream, in, out, natural);
    else hipLaunchKernelGGL((k_fft2<64, false>), dim3(batch), dim3(Frame<64>::T), 0, stream, in, out, natural);
  } else {
    if (inv) hipLaunchKernelGGL((k_fft2<128, true>), dim3(batch), dim3(Frame<128>::T), 0, stream, in, out, natural);
    else hipLaunchKernelGGL((k_fft2<128, false>), dim3(batch), dim3(Frame<128>::T), 0, stream, in, out, natural);
  }
  return check_launch("k_fft2");
}

int num_tiles(int Px, int Py) { return gather_num_tiles(Px, Py); }

int launch_gather(const GatherArgs& a, int Px, int Py, hipStream_t stream) {
  hipLaunchKernelGGL(k_gather, dim3(num_tiles(Px, Py)), dim3(GATHER_THREADS), 0, stream, a);
  return check_launch("k_gather");
}

int grid_for(long long n, int threads, int cap = 4096) {
  long long g = (n + threads - 1) / threads;
  return (int)(g < 1 ? 1 : (g > cap ? cap : g));
}

int launch_object_update(cf* obj, const cf* grad, const float* spsum, const float* tilemax, int ntiles,
                         int Px, int Py, float step, float nmin, hipStream_t stream, int replace = 0) {
  const long long npix = (long long)Px * Py;
  const float c1 = (float)(1.0 - (double)nmin), c2 = nmin;
  hipLaunchKernelGGL(k_object_update, dim3(grid_for(npix, 256)), dim3(256), 0, stream, obj, grad, spsum,
                     tilemax, ntiles, npix, step, c1, c2, replace);
  return check_launch("k_object_update");
}

int launch_object_update_fft(cf* obj, const cf* grad, const float* spsum, const float* tilemax, int ntiles, int Px,
                             int Py, float step, float nmin, const cf* probe, cf* phat, int S, hipStream_t stream) {
  const long long npix = (long long)Px * Py;
  const float c1 = (float)(1.0 - (double)nmin), c2 = nmin;
  if (S == 32)
    hipLaunchKernelGGL(k_object_update_fft<32>, dim3(1 + grid_for(npix, Frame<32>::T)), dim3(Frame<32>::T), 0, stream,
                       obj, grad, spsum, tilemax, ntiles, npix, step, c1, c2, probe, phat);
  else if (S == 64)
    hipLaunchKernelGGL(k_object_update_fft<64>, dim3(1 + grid_for(npix, Frame<64>::T)), dim3(Frame<64>::T), 0, stream,
                       obj, grad, spsum, tilemax, ntiles, npix, step, c1, c2, probe, phat);
  else
    hipLaunchKernelGGL(k_object_update_fft<128>, dim3(1 + grid_for(npix, Frame<128>::T)), dim3(Frame<128>::T), 0,
                       stream, obj, grad, spsum, tilemax, ntiles, npix, step, c1, c2, probe, phat);
  return check_launch("k_object_update_fft");
}

int launch_object_update_potential(float* V, const cf* grad, const float* spsum, const float* tilemax, int ntiles,
                                   int Px, int Py, float step, float nmin, hipStream_t stream) {
  const long long npix = (long long)Px * Py;
  const float c1 = (float)(1.0 - (double)nmin), c2 = nmin;
  hipLaunchKernelGGL(k_object_update_potential, dim3(grid_for(npix, 256)), dim3(256), 0, stream, V, grad, spsum,
                     tilemax, ntiles, npix, step, c1, c2);
  return check_launch("k_object_update_potential");
}

__host__ __device__ inline int probe_nchunk(int B) { return (B + PROBE_CHUNK - 1) / PROBE_CHUNK; }

// partials [nchunk][S*S][3] doubles, then gsum [S*S] cf, osum [S*S] f32, bmax
size_t probe_ws_bytes(int S, int B) {
  const size_t S2 = (size_t)S * S;
  return align256((size_t)probe_nchunk(B) * S2 * 24) + align256(S2 * 8) + align256(S2 * 4) +
         align256((S2 / 256 + 1) * 4);
}

int launch_probe_apply(cf* probe, double* ws, int S, int B, float step, float nmin, hipStream_t stream,
                       int replace = 0) {
  const size_t S2 = (size_t)S * S;
  char* base = (char*)ws;
  double* part = ws;
  cf* gsum = (cf*)(base + align256((size_t)probe_nchunk(B) * S2 * 24));
  float* osum = (float*)((char*)gsum + align256(S2 * 8));
  float* bmax = (float*)((char*)osum + align256(S2 * 4));
  const int npix = (int)S2, nblk = (npix + 255) / 256;
  hipLaunchKernelGGL(k_probe_reduce, dim3(nblk), dim3(PROBE_REDUCE_THREADS), 0, stream, part, probe_nchunk(B), npix, gsum, osum,
                     bmax);
  int rc = check_launch("k_probe_reduce");
  if (rc) return rc;
  const float c1 = (float)(1.0 - (double)nmin), c2 = nmin;
  hipLaunchKernelGGL(k_probe_apply, dim3(nblk), dim3(256), 0, stream, probe, gsum, osum, bmax, nblk, npix, step,
                     c1, c2, replace);
  return check_launch("k_probe_apply");
}

int launch_probe_update(cf* probe, const cf* chi, const cf* obj, int Px, int Py, const float* pos, int N,
                        const int* order, int B, int S, float step, float nmin, double* part,
                        hipStream_t stream) {
  ProbePartialArgs pa{chi, obj, pos, order, N, B, S, Px, Py, part};
  const int nchunk = probe_nchunk(B);
  hipLaunchKernelGGL(k_probe_partial, dim3((S * S + 255) / 256, nchunk), dim3(256), 0, stream, pa);
  int rc = check_launch("k_probe_partial");
  if (rc) return rc;
  return launch_probe_apply(probe, part, S, B, step, nmin, stream);
}

// ---------------------------------------------------------------------------
// gaussian smoothing of the object (ptychographic_constraints.py:119-151,
// which calls scipy.ndimage.gaussian_filter with its defaults: reflect
// boundaries, truncate 4).  One pass per axis, rows (axis 0) first; each
// output is scipy's correlate1d sum for a symmetric kernel in float64,
// w[r] x[i] + sum over j = r .. 1 of (x[i-j] + x[i+j]) w[r-j], rounded to
// float32 after the axis.  The weights are computed on the host as scipy
// computes them.  Pure-phase objects smooth the phase and rebuild exp(i phi).
// ---------------------------------------------------------------------------
__device__ __forceinline__ int reflect_index(int i, int n) {  // half-sample symmetric, period 2n
  const int p = 2 * n;
  int m = i % p;
  m += m < 0 ? p : 0;
  return m < n ? m : p - 1 - m;
}

template <class T> struct Lanes;
template <> struct Lanes<float> {
  static constexpr int N = 1;
  static __device__ double get(float v, int) { return (double)v; }
  static __device__ void set(float& v, int, double d) { v = (float)d; }
};
template <> struct Lanes<cf> {
  static constexpr int N = 2;
  static __device__ double get(cf v, int c) { return (double)(c ? v.y : v.x); }
  static __device__ void set(cf& v, int c, double d) { (c ? v.y : v.x) = (float)d; }
};

template <class T, int AXIS>
__global__ __launch_bounds__(256) void k_gaussian_axis(const T* __restrict__ in, T* __restrict__ out, int Px,
                                                       int Py, const double* __restrict__ w, int r) {
#pragma clang fp contract(off)
  const long long npix = (long long)Px * Py;
  const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= npix) return;
  const int x = (int)(p / Py), y = (int)(p % Py);
  const int n = AXIS == 0 ? Px : Py;
  const int i = AXIS == 0 ? x : y;
  const long long base = AXIS == 0 ? y : (long long)x * Py;
  const long long step = AXIS == 0 ? Py : 1;
  T res;
#pragma unroll
  for (int c = 0; c < Lanes<T>::N; ++c) {
    double acc = Lanes<T>::get(in[base + (long long)i * step], c) * w[r];
    for (int j = r; j >= 1; --j) {
      const double lo = Lanes<T>::get(in[base + (long long)reflect_index(i - j, n) * step], c);
      const double hi = Lanes<T>::get(in[base + (long long)reflect_index(i + j, n) * step], c);
      acc += (lo + hi) * w[r - j];
    }
    Lanes<T>::set(res, c, acc);
  }
  out[p] = res;
}

__global__ __launch_bounds__(256) void k_object_phase(const cf* __restrict__ obj, float* __restrict__ phase,
                                                      long long npix) {
  const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p < npix) phase[p] = atan2f(obj[p].y, obj[p].x);
}

__global__ __launch_bounds__(256) void k_object_expi(const float* __restrict__ phase, cf* __restrict__ obj,
                                                     long long npix) {
  const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p < npix) {
    float s, c;
    sincosf(phase[p], &s, &c);
    obj[p] = mkcf(c, s);
  }
}

// tmp: Px * Py * 8 bytes
int launch_object_gaussian(cf* obj, int Px, int Py, const double* w, int r, int pure, void* tmp,
                           hipStream_t stream) {
  const long long npix = (long long)Px * Py;
  const dim3 grid(grid_for(npix, 256)), blk(256);
  if (pure) {
    float* ph = (float*)tmp;
    float* ph2 = ph + npix;
    hipLaunchKernelGGL(k_object_phase, grid, blk, 0, stream, obj, ph, npix);
    hipLaunchKernelGGL((k_gaussian_axis<float, 0>), grid, blk, 0, stream, ph, ph2, Px, Py, w, r);
    hipLaunchKernelGGL((k_gaussian_axis<float, 1>), grid, blk, 0, stream, ph2, ph, Px, Py, w, r);
    hipLaunchKernelGGL(k_object_expi, grid, blk, 0, stream, ph, obj, npix);
  } else {
    cf* t = (cf*)tmp;
    hipLaunchKernelGGL((k_gaussian_axis<cf, 0>), grid, blk, 0, stream, obj, t, Px, Py, w, r);
    hipLaunchKernelGGL((k_gaussian_axis<cf, 1>), grid, blk, 0, stream, t, obj, Px, Py, w, r);
  }
  return check_launch("k_gaussian_axis");
}

// a real array (the potential object): both axes on the float32 values; tmp:
// Px * Py * 4 bytes
int launch_gaussian_real(float* V, int Px, int Py, const double* w, int r, void* tmp, hipStream_t stream) {
  const long long npix = (long long)Px * Py;
  const dim3 grid(grid_for(npix, 256)), blk(256);
  float* t = (float*)tmp;
  hipLaunchKernelGGL((k_gaussian_axis<float, 0>), grid, blk, 0, stream, V, t, Px, Py, w, r);
  hipLaunchKernelGGL((k_gaussian_axis<float, 1>), grid, blk, 0, stream, t, V, Px, Py, w, r);
  return check_launch("k_gaussian_axis");
}

// ---------------------------------------------------------------------------
// Butterworth filter of the object (ptychographic_constraints.py:153-201):
// O - mean(O), fft2, times env(|q|) on the fftfreq grid, ifft2, + mean.  The
// object is any size, so each 2-D transform is two complex products with the
// DFT matrices (rows, then columns), W[a][b] = exp(-+2 pi i ((a b) mod n) / n)
// generated in float64 while the tiles are loaded; float32 products, float32
// accumulation over n terms.
// ---------------------------------------------------------------------------
constexpr int DFT_T = 32;  // output tile DFT_T x DFT_T, k-step DFT_T

__device__ __forceinline__ cf dft_twiddle(int a, int b, int n, int inverse) {
  const int m = (int)(((long long)a * b) % n);
  double s, c;
  sincospi((inverse ? 2.0 : -2.0) * (double)m / (double)n, &s, &c);
  return mkcf((float)c, (float)s);
}

// LEFT:  Y[i][j] = sum_k W_M[i][k] X[k][j]   (transform along axis 0)
// RIGHT: Y[i][j] = sum_k X[i][k] W_N[k][j]   (transform along axis 1)
template <bool LEFT>
__global__ __launch_bounds__(256) void k_dft(const cf* __restrict__ X, cf* __restrict__ Y, int M, int N,
                                             int inverse) {
  __shared__ cf As[DFT_T][DFT_T + 1];
  __shared__ cf Bs[DFT_T][DFT_T + 1];
  const int K = LEFT ? M : N;
  const int i0 = blockIdx.y * DFT_T, j0 = blockIdx.x * DFT_T;
  const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
  float ar[2][2] = {{0.f, 0.f}, {0.f, 0.f}}, ai[2][2] = {{0.f, 0.f}, {0.f, 0.f}};
  for (int k0 = 0; k0 < K; k0 += DFT_T) {
    for (int e = threadIdx.x; e < DFT_T * DFT_T; e += 256) {
      const int r = e / DFT_T, c = e % DFT_T;
      const int gi = i0 + r, gk = k0 + c;
      cf v = mkcf(0.f, 0.f);
      if (gi < M && gk < K) v = LEFT ? dft_twiddle(gi, gk, K, inverse) : X[(size_t)gi * N + gk];
      As[r][c] = v;
      const int hk = k0 + r, gj = j0 + c;
      cf u = mkcf(0.f, 0.f);
      if (hk < K && gj < N) u = LEFT ? X[(size_t)hk * N + gj] : dft_twiddle(hk, gj, K, inverse);
      Bs[r][c] = u;
    }
    __syncthreads();
#pragma unroll 8
    for (int k = 0; k < DFT_T; ++k) {
      const cf a0 = As[2 * ty][k], a1 = As[2 * ty + 1][k];
      const cf b0 = Bs[k][2 * tx], b1 = Bs[k][2 * tx + 1];
      const cf a[2] = {a0, a1}, b[2] = {b0, b1};
#pragma unroll
      for (int u = 0; u < 2; ++u)
#pragma unroll
        for (int v = 0; v < 2; ++v) {
          ar[u][v] += a[u].x * b[v].x - a[u].y * b[v].y;
          ai[u][v] += a[u].x * b[v].y + a[u].y * b[v].x;
        }
    }
    __syncthreads();
  }
#pragma unroll
  for (int u = 0; u < 2; ++u)
#pragma unroll
    for (int v = 0; v < 2; ++v) {
      const int gi = i0 + 2 * ty + u, gj = j0 + 2 * tx + v;
      if (gi < M && gj < N) Y[(size_t)gi * N + gj] = mkcf(ar[u][v], ai[u][v]);
    }
}

// ---------------------------------------------------------------------------
// resampling / padding of the patterns (phase_base_class.py:223-420): every
// reshaping method there (Fourier crop or embed, order-1 zoom, binning, zero
// padding) is linear along each axis, so a pattern X (Qx x Qy, real) becomes
// Re(Mx X My^T) with complex Mx (Sx x Qx), My (Sy x Qy) built on the host.
// Two batched products (blockIdx.z = pattern), 32 x 32 output tiles.
// ---------------------------------------------------------------------------
// T[b] = Mx (m x k) X[b] (k x n): complex m x n
__global__ __launch_bounds__(256) void k_lin_left(const cf* __restrict__ Mx, const float* __restrict__ X,
                                                  cf* __restrict__ T, int m, int k, int n) {
  __shared__ cf As[DFT_T][DFT_T + 1];
  __shared__ float Bs[DFT_T][DFT_T + 1];
  const float* Xb = X + (size_t)blockIdx.z * k * n;
  cf* Tb = T + (size_t)blockIdx.z * m * n;
  const int i0 = blockIdx.y * DFT_T, j0 = blockIdx.x * DFT_T;
  const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
  float ar[2][2] = {{0.f, 0.f}, {0.f, 0.f}}, ai[2][2] = {{0.f, 0.f}, {0.f, 0.f}};
  for (int k0 = 0; k0 < k; k0 += DFT_T) {
    for (int e = threadIdx.x; e < DFT_T * DFT_T; e += 256) {
      const int r = e / DFT_T, c = e % DFT_T;
      const int gi = i0 + r, gk = k0 + c;
      As[r][c] = (gi < m && gk < k) ? Mx[(size_t)gi * k + gk] : mkcf(0.f, 0.f);
      const int hk = k0 + r, gj = j0 + c;
      Bs[r][c] = (hk < k && gj < n) ? Xb[(size_t)hk * n + gj] : 0.f;
    }
    __syncthreads();
#pragma unroll 8
    for (int q = 0; q < DFT_T; ++q) {
      const cf a[2] = {As[2 * ty][q], As[2 * ty + 1][q]};
      const float b[2] = {Bs[q][2 * tx], Bs[q][2 * tx + 1]};
#pragma unroll
      for (int u = 0; u < 2; ++u)
#pragma unroll
        for (int v = 0; v < 2; ++v) {
          ar[u][v] += a[u].x * b[v];
          ai[u][v] += a[u].y * b[v];
        }
    }
    __syncthreads();
  }
#pragma unroll
  for (int u = 0; u < 2; ++u)
#pragma unroll
    for (int v = 0; v < 2; ++v) {
      const int gi = i0 + 2 * ty + u, gj = j0 + 2 * tx + v;
      if (gi < m && gj < n) Tb[(size_t)gi * n + gj] = mkcf(ar[u][v], ai[u][v]);
    }
}

// Y[b] = Re(T[b] (m x k) My^T), My n x k: real m x n
__global__ __launch_bounds__(256) void k_lin_right(const cf* __restrict__ T, const cf* __restrict__ My,
                                                   float* __restrict__ Y, int m, int k, int n) {
  __shared__ cf As[DFT_T][DFT_T + 1];
  __shared__ cf Bs[DFT_T][DFT_T + 1];
  const cf* Tb = T + (size_t)blockIdx.z * m * k;
  float* Yb = Y + (size_t)blockIdx.z * m * n;
  const int i0 = blockIdx.y * DFT_T, j0 = blockIdx.x * DFT_T;
  const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
  float acc[2][2] = {{0.f, 0.f}, {0.f, 0.f}};
  for (int k0 = 0; k0 < k; k0 += DFT_T) {
    for (int e = threadIdx.x; e < DFT_T * DFT_T; e += 256) {
      const int r = e / DFT_T, c = e % DFT_T;
      const int gi = i0 + r, gk = k0 + c;
      As[r][c] = (gi < m && gk < k) ? Tb[(size_t)gi * k + gk] : mkcf(0.f, 0.f);
      const int hk = k0 + r, gj = j0 + c;  // Bs[q][j] = My[j][q]
      Bs[r][c] = (hk < k && gj < n) ? My[(size_t)gj * k + hk] : mkcf(0.f, 0.f);
    }
    __syncthreads();
#pragma unroll 8
    for (int q = 0; q < DFT_T; ++q) {
      const cf a[2] = {As[2 * ty][q], As[2 * ty + 1][q]};
      const cf b[2] = {Bs[q][2 * tx], Bs[q][2 * tx + 1]};
#pragma unroll
      for (int u = 0; u < 2; ++u)
#pragma unroll
        for (int v = 0; v < 2; ++v) acc[u][v] += a[u].x * b[v].x - a[u].y * b[v].y;
    }
    __syncthreads();
  }
#pragma unroll
  for (int u = 0; u < 2; ++u)
#pragma unroll
    for (int v = 0; v < 2; ++v) {
      const int gi = i0 + 2 * ty + u, gj = j0 + 2 * tx + v;
      if (gi < m && gj < n) Yb[(size_t)gi * n + gj] = acc[u][v];
    }
}

// float64 partial sums of the object (re, im) per block, fixed grid
constexpr int MEAN_BLOCKS = 256;
static_assert(MEAN_BLOCKS <= CONSTRAINT_BLOCKS, "mean partials share the constraint partial buffer");
__global__ __launch_bounds__(256) void k_complex_sum_partial(const cf* __restrict__ obj, long long npix,
                                                             double* __restrict__ part) {
  __shared__ double sr[256], si[256];
  double a = 0.0, b = 0.0;
  for (long long p = (long long)blockIdx.x * 256 + threadIdx.x; p < npix; p += (long long)MEAN_BLOCKS * 256) {
    a += (double)obj[p].x;
    b += (double)obj[p].y;
  }
  sr[threadIdx.x] = a;
  si[threadIdx.x] = b;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) {
      sr[threadIdx.x] += sr[threadIdx.x + s];
      si[threadIdx.x] += si[threadIdx.x + s];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = sr[0];
    part[2 * blockIdx.x + 1] = si[0];
  }
}

__device__ __forceinline__ cf mean_from_partials(const double* part, long long npix) {
  double a = 0.0, b = 0.0;
  for (int k = 0; k < MEAN_BLOCKS; ++k) {
    a += part[2 * k];
    b += part[2 * k + 1];
  }
  return mkcf((float)(a / (double)npix), (float)(b / (double)npix));
}

// sign = -1: O -= mean; sign = +1: O += mean
__global__ __launch_bounds__(256) void k_shift_by_mean(cf* __restrict__ obj, long long npix,
                                                       const double* __restrict__ part, float sign) {
  __shared__ cf m_sh;
  if (threadIdx.x == 0) m_sh = mean_from_partials(part, npix);
  __syncthreads();
  const cf m = m_sh;
  for (long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x; p < npix;
       p += (long long)gridDim.x * blockDim.x) {
    const cf o = obj[p];
    obj[p] = mkcf(o.x + sign * m.x, o.y + sign * m.y);
  }
}

// spectrum times env(|q|) / (Px Py) (the inverse transform's normalisation),
// env evaluated in float64 as the reference evaluates it
__global__ __launch_bounds__(256) void k_butterworth_env(cf* __restrict__ spec, int Px, int Py, double dx,
                                                         double dy, double q_lowpass, double q_highpass,
                                                         double order) {
  const long long npix = (long long)Px * Py;
  const double vx = 1.0 / (Px * dx), vy = 1.0 / (Py * dy);
  const double norm = 1.0 / ((double)Px * (double)Py);
  for (long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x; p < npix;
       p += (long long)gridDim.x * blockDim.x) {
    const int x = (int)(p / Py), y = (int)(p % Py);
    const int kx = x < (Px - 1) / 2 + 1 ? x : x - Px;  // numpy fftfreq order
    const int ky = y < (Py - 1) / 2 + 1 ? y : y - Py;
    const double qx = kx * vx, qy = ky * vy;
    const double qr = sqrt(qx * qx + qy * qy);
    double env = 1.0;
    if (q_highpass != 0.0) env *= 1.0 - 1.0 / (1.0 + pow(qr / q_highpass, 2.0 * order));
    if (q_lowpass != 0.0) env *= 1.0 / (1.0 + pow(qr / q_lowpass, 2.0 * order));
    const double f = env * norm;
    const cf v = spec[p];
    spec[p] = mkcf((float)((double)v.x * f), (float)((double)v.y * f));
  }
}

// tmp: Px * Py * 8 bytes; part: MEAN_BLOCKS * 16 bytes
int launch_object_butterworth(cf* obj, int Px, int Py, double q_lowpass, double q_highpass, double order,
                              double dx, double dy, void* tmp, double* part, hipStream_t stream) {
  const long long npix = (long long)Px * Py;
  cf* t = (cf*)tmp;
  const dim3 g2((Py + DFT_T - 1) / DFT_T, (Px + DFT_T - 1) / DFT_T), b2(256);
  const dim3 g1(grid_for(npix, 256)), b1(256);
  hipLaunchKernelGGL(k_complex_sum_partial, dim3(MEAN_BLOCKS), dim3(256), 0, stream, obj, npix, part);
  hipLaunchKernelGGL(k_shift_by_mean, g1, b1, 0, stream, obj, npix, part, -1.f);
  hipLaunchKernelGGL(k_dft<true>, g2, b2, 0, stream, obj, t, Px, Py, 0);
  hipLaunchKernelGGL(k_dft<false>, g2, b2, 0, stream, t, obj, Px, Py, 0);
  hipLaunchKernelGGL(k_butterworth_env, g1, b1, 0, stream, obj, Px, Py, dx, dy, q_lowpass, q_highpass, order);
  hipLaunchKernelGGL(k_dft<true>, g2, b2, 0, stream, obj, t, Px, Py, 1);
  hipLaunchKernelGGL(k_dft<false>, g2, b2, 0, stream, t, obj, Px, Py, 1);
  hipLaunchKernelGGL(k_shift_by_mean, g1, b1, 0, stream, obj, npix, part, 1.f);
  return check_launch("k_dft");
}

// smoothing filters ahead of the shrinkage and threshold
// (ptychographic_constraints.py:431-446); tmp: Px * Py * 8 bytes, part:
// MEAN_BLOCKS * 16 bytes
int launch_object_filters(cf* obj, int Px, int Py, const pty_gd_params& q, void* tmp, double* part,
                          hipStream_t stream) {
  int rc = PTY_OK;
  if (q.gaussian_weights && q.gaussian_radius >= 0 &&
      (rc = launch_object_gaussian(obj, Px, Py, q.gaussian_weights, q.gaussian_radius, q.pure_phase_object, tmp,
                                   stream)))
    return rc;
  if (q.butterworth &&
      (rc = launch_object_butterworth(obj, Px, Py, q.q_lowpass, q.q_highpass, q.butterworth_order,
                                      q.sampling[0], q.sampling[1], tmp, part, stream)))
    return rc;
  return rc;
}

int launch_constraints(cf* obj, int Px, int Py, const unsigned char* mask, float shrink, int pure,
                       double* part, hipStream_t stream) {
  const long long npix = (long long)Px * Py;
  if (mask) {
    hipLaunchKernelGGL(k_constraint_partial, dim3(CONSTRAINT_BLOCKS), dim3(256), 0, stream, obj, mask, npix,
                       part, (const float*)nullptr);
    int rc = check_launch("k_constraint_partial");
    if (rc) return rc;
  }
  hipLaunchKernelGGL(k_constraint_apply, dim3(grid_for(npix, 256)), dim3(256), 0, stream, obj, npix, part,
                     CONSTRAINT_BLOCKS, mask ? 1 : 0, shrink, pure);
  return check_launch("k_constraint_apply");
}

int launch_constraints_potential(float* V, int Px, int Py, const unsigned char* mask, float shrink, int positivity,
                                 double* part, hipStream_t stream) {
  const long long npix = (long long)Px * Py;
  if (mask) {
    hipLaunchKernelGGL(k_constraint_partial, dim3(CONSTRAINT_BLOCKS), dim3(256), 0, stream, (const cf*)nullptr, mask,
                       npix, part, (const float*)V);
    int rc = check_launch("k_constraint_partial");
    if (rc) return rc;
  }
  hipLaunchKernelGGL(k_constraint_apply_potential, dim3(grid_for(npix, 256)), dim3(256), 0, stream, V, npix, part,
                     CONSTRAINT_BLOCKS, mask ? 1 : 0, shrink, positivity);
  return check_launch("k_constraint_apply_potential");
}

// probe recentring; 256 px frames need `tmp` (one frame) and `shift` (2
// floats) of device scratch
int launch_probe_com(cf* probe, int S, hipStream_t stream, cf* tmp = nullptr, float* shift = nullptr) {
  if (S == 256) {
    if (!tmp || !shift) return fail(PTY_E_UNSUPPORTED, "the 256 px probe recentring needs a workspace (pty_gd_run)");
    return launch_probe_com256(probe, tmp, shift, stream);
  }
  if (S == 32) hipLaunchKernelGGL(k_probe_com<32>, dim3(1), dim3(Frame<32>::T), 0, stream, probe);
  else if (S == 64) hipLaunchKernelGGL(k_probe_com<64>, dim3(1), dim3(Frame<64>::T), 0, stream, probe);
  else hipLaunchKernelGGL(k_probe_com<128>, dim3(1), dim3(Frame<128>::T), 0, stream, probe);
  return check_launch("k_probe_com");
}

// the probe constraints after the recentring (ptychographic_constraints.py:
// 1091-1147, 1218-1286): the aperture on mode 0, the amplitude top-hat on
// every mode
int launch_probe_extra(cf* probe, int K, int S, const pty_gd_params& q, hipStream_t stream) {
  int rc = PTY_OK;
  if (S == 256 && (q.fix_probe_aperture || q.constrain_probe_amplitude))
    return fail(PTY_E_UNSUPPORTED, "probe aperture/amplitude constraints take frames of 32, 64 or 128 px");
  if (q.fix_probe_aperture) {
    if (!q.probe_aperture) return fail(PTY_E_ARG, "fix_probe_aperture needs the initial aperture");
    if (S == 32) hipLaunchKernelGGL(k_probe_aperture<32>, dim3(1), dim3(Frame<32>::T), 0, stream, probe,
                                    q.probe_aperture);
    else if (S == 64) hipLaunchKernelGGL(k_probe_aperture<64>, dim3(1), dim3(Frame<64>::T), 0, stream, probe,
                                    q.probe_aperture);
    else hipLaunchKernelGGL(k_probe_aperture<128>, dim3(1), dim3(Frame<128>::T), 0, stream, probe,
                            q.probe_aperture);
    if ((rc = check_launch("k_probe_aperture"))) return rc;
  }
  if (q.constrain_probe_amplitude) {
    const double rad = q.probe_amplitude_radius, wid = q.probe_amplitude_width;
    if (S == 32) hipLaunchKernelGGL(k_probe_amplitude<32>, dim3(K), dim3(Frame<32>::T), 0, stream, probe, rad, wid);
    else if (S == 64) hipLaunchKernelGGL(k_probe_amplitude<64>, dim3(K), dim3(Frame<64>::T), 0, stream, probe, rad, wid);
    else hipLaunchKernelGGL(k_probe_amplitude<128>, dim3(K), dim3(Frame<128>::T), 0, stream, probe, rad, wid);
    if ((rc = check_launch("k_probe_amplitude"))) return rc;
  }
  return rc;
}

// every one of K modes shifted by mode 0's centre of mass (shift: 2 floats of scratch)
int launch_probe_com_modes(cf* probe, int K, int S, float* shift, hipStream_t stream) {
  if (S == 32) hipLaunchKernelGGL(k_probe_com_measure<32>, dim3(1), dim3(Frame<32>::T), 0, stream, probe, shift);
  else if (S == 64) hipLaunchKernelGGL(k_probe_com_measure<64>, dim3(1), dim3(Frame<64>::T), 0, stream, probe, shift);
  else hipLaunchKernelGGL(k_probe_com_measure<128>, dim3(1), dim3(Frame<128>::T), 0, stream, probe, shift);
  int rc = check_launch("k_probe_com_measure");
  if (rc) return rc;
  if (S == 32) hipLaunchKernelGGL(k_probe_shift_modes<32>, dim3(K), dim3(Frame<32>::T), 0, stream, probe, shift);
  else if (S == 64) hipLaunchKernelGGL(k_probe_shift_modes<64>, dim3(K), dim3(Frame<64>::T), 0, stream, probe, shift);
  else hipLaunchKernelGGL(k_probe_shift_modes<128>, dim3(K), dim3(Frame<128>::T), 0, stream, probe, shift);
  return check_launch("k_probe_shift_modes");
}

int launch_position_correction(const PosArgs& a, int S, hipStream_t stream) {
  if (a.B <= 0) return PTY_OK;
  if (S == 32) hipLaunchKernelGGL(k_position_correction<32>, dim3(a.B), dim3(Frame<32>::T), 0, stream, a);
  else if (S == 64) hipLaunchKernelGGL(k_position_correction<64>, dim3(a.B), dim3(Frame<64>::T), 0, stream, a);
  else hipLaunchKernelGGL(k_position_correction<128>, dim3(a.B), dim3(Frame<128>::T), 0, stream, a);
  return check_launch("k_position_correction");
}

int launch_positions_com(float* pos, int N, const float* com, hipStream_t stream) {
  hipLaunchKernelGGL(k_positions_com, dim3(1), dim3(1024), 0, stream, pos, N, com[0], com[1]);
  return check_launch("k_positions_com");
}

int validate_position_params(const pty_position_params* pp, const float* positions) {
  if (!pp) return PTY_OK;
  if (!pp->positions || pp->positions != positions)
    return fail(PTY_E_ARG, "position params must update the problem's positions buffer");
  if (!pp->positions_initial) return fail(PTY_E_ARG, "positions_initial is NULL");
  if (!std::isfinite(pp->step_size)) return fail(PTY_E_ARG, "positions step size is not finite");
  return PTY_OK;
}

PosArgs position_args(const pty_position_params& pp, const cf* ps, const cf* far, const cf* obj, int Px, int Py,
                      const float* amps, int N, const int* order, int B) {
  PosArgs a{};
  a.ps = ps; a.far = far; a.object = obj; a.amps = amps;
  a.positions = pp.positions; a.positions_initial = pp.positions_initial; a.order = order;
  a.N = N; a.B = B; a.Px = Px; a.Py = Py;
  a.step = pp.step_size;
  a.clip_update = pp.clip_update; a.max_update = pp.max_update;
  a.clip_total = pp.clip_total; a.max_total = pp.max_total;
  a.sums = nullptr; a.singular = pp.singular;
  a.far_modes = 1; a.acc = nullptr; a.acc_add = 0;
  return a;
}

// near-equal batch sizes, larger ones first (phase/utils.py:539-579)
std::vector<int> subdivide(int n, int max_batch) {
  const int nb = (n + ((max_batch - n % max_batch) % max_batch)) / max_batch;
  std::vector<int> v(nb, n / nb);
  const int extra = n % nb;
  for (int i = 0; i < extra; ++i) v[i] += 1;
  return v;
}

struct WsLayout {
  size_t phat, ps, chi, err, grad, spsum, tilemax, ppart, cpart, run, far, total;
};

// positions per fused launch: a multiple of 64 (probe-partial chunk) or the
// whole mini-batch
int stage_size(const pty_gd_params& q) {
  if (q.stage <= 0 || q.stage >= q.max_batch) return q.max_batch;
  const int s = ((q.stage + 63) / 64) * 64;
  return s < q.max_batch ? s : q.max_batch;
}

// With position correction every launch covers the whole mini-batch: the
// correction reads the batch's shifted probes and far fields after the update.
int stage_for(const pty_gd_params& q, bool positions) { return positions ? q.max_batch : stage_size(q); }

// stages pipelined over two streams: P_shift and chi double-buffered
bool pipelined(const pty_gd_params& q, bool positions) {
  return q.pipeline != 0 && !positions && stage_size(q) < q.max_batch;
}

// the second stream and the stage events of pipelined runs, per device;
// `mu` is held while a pipelined run enqueues its launches
struct SidePipe {
  std::mutex mu;
  hipStream_t s = nullptr;
  hipEvent_t fused_done[2] = {nullptr, nullptr};
  hipEvent_t scatter_done[2] = {nullptr, nullptr};
};
constexpr int PIPE_MAX_DEVICES = 64;
std::mutex g_pipe_mu;
SidePipe g_pipe[PIPE_MAX_DEVICES];

int side_pipe(SidePipe*& out) {
  int dev = -1;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return fail(PTY_E_LAUNCH, "hipGetDevice: %s", hipGetErrorString(e));
  if (dev < 0 || dev >= PIPE_MAX_DEVICES)
    return fail(PTY_E_ARG, "device %d beyond the %d the pipeline table covers", dev, PIPE_MAX_DEVICES);
  std::lock_guard<std::mutex> lock(g_pipe_mu);
  SidePipe& sp = g_pipe[dev];
  if (!sp.s) {
    // default priority, not blocking on the null stream
    if ((e = hipStreamCreateWithFlags(&sp.s, hipStreamNonBlocking)) != hipSuccess) {
      sp.s = nullptr;
      return fail(PTY_E_LAUNCH, "hipStreamCreateWithFlags: %s", hipGetErrorString(e));
    }
    for (int i = 0; i < 2; ++i) {
      if ((e = hipEventCreateWithFlags(&sp.fused_done[i], hipEventDisableTiming)) != hipSuccess ||
          (e = hipEventCreateWithFlags(&sp.scatter_done[i], hipEventDisableTiming)) != hipSuccess)
        return fail(PTY_E_LAUNCH, "hipEventCreateWithFlags: %s", hipGetErrorString(e));
    }
  }
  out = &sp;
  return PTY_OK;
}

// projection sets stage no exit-wave differences (chi): the exit waves
// themselves are the caller's persistent buffer
WsLayout layout(const pty_problem& p, const pty_gd_params& q, bool positions = false, bool no_chi = false) {
  WsLayout L{};
  const size_t S2 = (size_t)p.S * p.S;
  const size_t B = (size_t)stage_for(q, positions);
  const size_t nbuf = pipelined(q, positions) && !no_chi ? 2 : 1;
  const size_t npix = (size_t)p.Px * p.Py;
  size_t o = 0;
  L.phat = o; o = align256(o + S2 * 8);
  L.ps = o; o = align256(o + nbuf * B * S2 * 8);
  L.chi = o; o = align256(o + (no_chi ? 0 : nbuf * B * S2 * 8));
  L.err = o; o = align256(o + (size_t)p.N * 8);
  L.grad = o; o = align256(o + npix * 8);
  L.spsum = o; o = align256(o + npix * 4);
  L.tilemax = o; o = align256(o + (size_t)num_tiles(p.Px, p.Py) * 4);
  L.ppart = o; o = align256(o + (q.fix_probe ? 0 : probe_ws_bytes(p.S, q.max_batch)));
  L.cpart = o; o = align256(o + (size_t)CONSTRAINT_BLOCKS * 16);
  L.run = o; o = align256(o + ((int)B < q.max_batch ? npix * 24 : 0));
  L.far = o; o = align256(o + (positions ? B * S2 * 8 : 0));
  L.total = o;
  return L;
}

int validate_problem(const pty_problem* p) {
  if (!p) return fail(PTY_E_ARG, "problem is NULL");
  if (!frame_supported_gd(p->S))
    return fail(PTY_E_UNSUPPORTED, "frame size S=%d not supported (32, 64, 128 or 256)", p->S);
  if (p->Px < p->S || p->Py < p->S)
    return fail(PTY_E_ARG, "object %dx%d smaller than the %dx%d frame", p->Px, p->Py, p->S, p->S);
  if (p->N <= 0) return fail(PTY_E_ARG, "no scan positions");
  if (!p->object || !p->probe || !p->positions || !p->amplitudes)
    return fail(PTY_E_ARG, "null device buffer in problem");
  if (p->object_type != PTY_OBJECT_COMPLEX && p->object_type != PTY_OBJECT_POTENTIAL)
    return fail(PTY_E_ARG, "object_type %d is neither PTY_OBJECT_COMPLEX nor PTY_OBJECT_POTENTIAL", p->object_type);
  if (!aligned8(p->probe) || (p->object_type == PTY_OBJECT_COMPLEX && !aligned8(p->object)))
    return fail(PTY_E_ARG, "complex buffers must be 8-byte aligned");
  return check_index_range(p->Px, p->Py, p->N, p->S);
}

#include "pty_multi.inc"

}  // namespace

// ---------------------------------------------------------------------------
// Static workspace of the allocation-free gradient-descent entry point: one
// copy per GPU (each device's code object holds its own), kept until the
// process ends.  Host code must use the per-device address from
// hipGetSymbolAddress, never the array's name.
// ---------------------------------------------------------------------------
static __device__ __align__(256) unsigned char scratch[PTY_NOALLOC_SCRATCH_BYTES];

namespace {

constexpr int SCRATCH_MAX_DEVICES = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[SCRATCH_MAX_DEVICES];

// device address of the current GPU's scratch array, looked up once per device
int scratch_pointer(void** out) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return fail(PTY_E_LAUNCH, "hipGetDevice: %s", hipGetErrorString(e));
  if (dev < 0 || dev >= SCRATCH_MAX_DEVICES)
    return fail(PTY_E_ARG, "device %d beyond the %d the scratch table covers", dev, SCRATCH_MAX_DEVICES);
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* ptr = nullptr;
    e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr)
      return fail(PTY_E_LAUNCH, "hipGetSymbolAddress(scratch) on device %d: %s", dev, hipGetErrorString(e));
    g_scratch_ptr[dev] = ptr;
  }
  *out = g_scratch_ptr[dev];
  return PTY_OK;
}

// the largest positions-per-launch setting (the whole mini-batch, else a
// multiple of 64 not above the caller's) whose workspace fits the scratch
int fit_stage_to_scratch(const pty_problem& p, pty_gd_params& q) {
  pty_gd_params t = q;
  int s = (q.stage > 0 && q.stage < q.max_batch) ? stage_size(q) : q.max_batch;
  size_t need = 0;
  while (true) {
    t.stage = s >= q.max_batch ? 0 : s;
    need = layout(p, t).total;
    if (need <= (size_t)PTY_NOALLOC_SCRATCH_BYTES) {
      q.stage = t.stage;
      return PTY_OK;
    }
    const int lower = ((s - 1) / 64) * 64;
    if (lower < 64) break;
    s = lower;
  }
  return fail(PTY_E_WORKSPACE,
              "workspace of %zu bytes at 64 positions per launch exceeds the %zu-byte static scratch "
              "(S=%d, object %dx%d); use pty_gd_run with a caller workspace",
              need, (size_t)PTY_NOALLOC_SCRATCH_BYTES, p.S, p.Px, p.Py);
}

}  // namespace

// ===========================================================================
// C ABI
// ===========================================================================
namespace {

// num_iter iterations of the GD loop; pp (optional) adds the per-batch
// position correction and the per-iteration centre-of-mass step
int gd_run(const pty_problem* problem, const pty_gd_params* params, const pty_position_params* pp,
           const int* orders, double* errors, void* workspace, size_t workspace_bytes, hipStream_t stream) {
  int rc = validate_problem(problem);
  if (rc) return rc;
  if (!params) return fail(PTY_E_ARG, "params is NULL");
  const pty_problem& p = *problem;
  const pty_gd_params& q = *params;
  if (q.num_iter < 0 || q.max_batch <= 0 || q.max_batch > p.N)
    return fail(PTY_E_ARG, "bad num_iter=%d / max_batch=%d for N=%d", q.num_iter, q.max_batch, p.N);
  if (!orders || !errors || !workspace) return fail(PTY_E_ARG, "null orders/errors/workspace");
  if ((rc = validate_position_params(pp, p.positions))) return rc;
  const bool correct = pp != nullptr;
  const bool pot = p.object_type == PTY_OBJECT_POTENTIAL;
  if (pot && (correct || p.fourier_mask || q.butterworth))
    return fail(PTY_E_UNSUPPORTED, "potential objects: no position correction, detector mask or Butterworth filter");
  if (p.S == 256 && (correct || (!q.fix_probe && (q.fix_probe_aperture || q.constrain_probe_amplitude))))
    return fail(PTY_E_UNSUPPORTED, "256 px frames: no position correction or probe aperture/amplitude constraint");
  const WsLayout L = layout(p, q, correct);
  if (workspace_bytes < L.total)
    return fail(PTY_E_WORKSPACE, "workspace %zu bytes < required %zu", workspace_bytes, L.total);
  char* ws = (char*)workspace;
  cf* phat = (cf*)(ws + L.phat);
  cf* ps = (cf*)(ws + L.ps);
  cf* chi = (cf*)(ws + L.chi);
  double* err = (double*)(ws + L.err);
  cf* grad = (cf*)(ws + L.grad);
  float* spsum = (float*)(ws + L.spsum);
  float* tilemax = (float*)(ws + L.tilemax);
  double* ppart = (double*)(ws + L.ppart);
  double* cpart = (double*)(ws + L.cpart);
  double* runbuf = (double*)(ws + L.run);
  cf* far = correct ? (cf*)(ws + L.far) : nullptr;
  const int stage = stage_for(q, correct);
  cf* obj = pot ? nullptr : (cf*)p.object;
  float* V = pot ? (float*)p.object : nullptr;
  cf* probe = (cf*)p.probe;
  const std::vector<int> batches = subdivide(p.N, q.max_batch);
  // pipelined stages: stage k's scatter and probe sums on the side stream
  // beside stage k + 1's fused pass; P_shift and chi alternate between two
  // buffers, so the fused pass of stage k + 2 waits for the scatter of
  // stage k, and the mini-batch's updates wait for its last scatter
  SidePipe* pipe = nullptr;
  std::unique_lock<std::mutex> pipe_lock;
  if (pipelined(q, correct)) {
    if ((rc = side_pipe(pipe))) return rc;
    pipe_lock = std::unique_lock<std::mutex>(pipe->mu);
  }
  const size_t stage_elems = (size_t)stage * p.S * p.S;
  const int ntiles = num_tiles(p.Px, p.Py);
  const double norm = (double)(float)(p.mean_intensity * (float)p.N);
  bool phat_valid = false;
  // complex object with probe update: the object step carries the fft2 of
  // the updated probe (k_object_update_fft), so phat stays current across
  // the mini-batches of an iteration
  const bool fft_with_update = !q.fix_probe && !pot && p.S != 256;
  for (int it = 0; it < q.num_iter; ++it) {
    const int* order_it = orders + (size_t)it * p.N;
    int start = 0;
    for (int nb : batches) {
      const int* ob = order_it + start;
      if (!phat_valid) {
        if ((rc = launch_fft(probe, phat, 1, p.S, false, 0, stream))) return rc;
        phat_valid = q.fix_probe != 0 || fft_with_update;
      }
      int k = 0;
      for (int s0 = 0; s0 < nb; s0 += stage, ++k) {
        const int ns = (nb - s0) < stage ? (nb - s0) : stage;
        cf* ps_k = pipe ? ps + (size_t)(k & 1) * stage_elems : ps;
        cf* chi_k = pipe ? chi + (size_t)(k & 1) * stage_elems : chi;
        hipStream_t gs = pipe ? pipe->s : stream;
        if (pipe && k >= 2 && (rc = hip_check(hipStreamWaitEvent(stream, pipe->scatter_done[k & 1], 0),
                                              "hipStreamWaitEvent")))
          return rc;
        FusedArgs a{};
        a.probe_hat = phat; a.object = obj; a.object_pot = V; a.positions = p.positions; a.amps = p.amplitudes;
        a.order = ob + s0; a.N = p.N; a.B = ns; a.Px = p.Px; a.Py = p.Py; a.b0 = 0;
        a.ps_out = ps_k; a.chi_out = chi_k; a.err = err + start + s0; a.fwd_out = far;
        a.fmask = p.fourier_mask;
        if ((rc = launch_fused(a, p.S, stream))) return rc;
        if (pipe && ((rc = hip_check(hipEventRecord(pipe->fused_done[k & 1], stream), "hipEventRecord")) ||
                     (rc = hip_check(hipStreamWaitEvent(gs, pipe->fused_done[k & 1], 0), "hipStreamWaitEvent"))))
          return rc;
        GatherArgs g{};
        g.ps = ps_k; g.chi = chi_k; g.positions = p.positions; g.order = ob + s0; g.N = p.N; g.B = ns; g.S = p.S;
        g.Px = p.Px; g.Py = p.Py; g.tiles_y = gather_tiles_y(p.Py); g.grad = grad; g.spsum = spsum;
        g.tilemax = tilemax;
        if (ns < nb) {
          g.run = runbuf;
          g.run_first = s0 == 0;
          g.run_last = s0 + ns == nb;
        }
        if (q.fix_probe) {
          if ((rc = launch_gather(g, p.Px, p.Py, gs))) return rc;
        } else {
          ProbePartialArgs pa{chi_k, obj, p.positions, ob + s0, p.N, ns, p.S, p.Px, p.Py,
                              ppart + (size_t)(s0 / PROBE_CHUNK) * p.S * p.S * 3};
          pa.object_pot = V;
          const int pix_blocks = (p.S * p.S + 255) / 256;
          hipLaunchKernelGGL(k_gather_probe, dim3(ntiles + pix_blocks * probe_nchunk(ns)), dim3(GATHER_THREADS), 0,
                             gs, g, pa, ntiles, pix_blocks);
          if ((rc = check_launch("k_gather_probe"))) return rc;
        }
        if (pipe && (rc = hip_check(hipEventRecord(pipe->scatter_done[k & 1], gs), "hipEventRecord"))) return rc;
      }
      if (pipe && (rc = hip_check(hipStreamWaitEvent(stream, pipe->scatter_done[(k - 1) & 1], 0),
                                  "hipStreamWaitEvent")))
        return rc;
      // the probe step first: the object step may carry the probe's fft2
      if (!q.fix_probe) {
        if ((rc = launch_probe_apply(probe, ppart, p.S, nb, q.step_size, q.normalization_min, stream))) return rc;
      }
      if (fft_with_update) {
        if ((rc = launch_object_update_fft(obj, grad, spsum, tilemax, ntiles, p.Px, p.Py, q.step_size,
                                           q.normalization_min, probe, phat, p.S, stream)))
          return rc;
      } else if (pot) {
        if ((rc = launch_object_update_potential(V, grad, spsum, tilemax, ntiles, p.Px, p.Py, q.step_size,
                                                 q.normalization_min, stream)))
          return rc;
      } else if ((rc = launch_object_update(obj, grad, spsum, tilemax, ntiles, p.Px, p.Py, q.step_size,
                                     q.normalization_min, stream)))
        return rc;
      if (!q.fix_probe && !fft_with_update) phat_valid = false;
      if (correct) {  // singleslice_ptychography.py:913-926, after the adjoint
        const PosArgs pa = position_args(*pp, ps, far, obj, p.Px, p.Py, p.amplitudes, p.N, ob, nb);
        if ((rc = launch_position_correction(pa, p.S, stream))) return rc;
      }
      start += nb;
    }
    hipLaunchKernelGGL(k_error_reduce, dim3(1), dim3(1024), 0, stream, err, p.N, norm, errors + it);
    if ((rc = check_launch("k_error_reduce"))) return rc;
    if (pot) {
      if (q.gaussian_weights && q.gaussian_radius >= 0 &&
          (rc = launch_gaussian_real(V, p.Px, p.Py, q.gaussian_weights, q.gaussian_radius, grad, stream)))
        return rc;
      if ((rc = launch_constraints_potential(V, p.Px, p.Py, p.fov_mask_inverse, q.shrinkage_rad, q.object_positivity,
                                             cpart, stream)))
        return rc;
    } else if ((rc = launch_object_filters(obj, p.Px, p.Py, q, grad, cpart, stream))) {
      return rc;
    } else if ((rc = launch_constraints(obj, p.Px, p.Py, p.fov_mask_inverse, q.shrinkage_rad, q.pure_phase_object, cpart,
                                 stream)))
      return rc;
    if (!q.fix_probe && q.fix_probe_com) {
      if ((rc = launch_probe_com(probe, p.S, stream, grad, (float*)cpart))) return rc;
    }
    if (!q.fix_probe && (rc = launch_probe_extra(probe, 1, p.S, q, stream))) return rc;
    if (!q.fix_probe) phat_valid = false;  // the recentring and constraints moved the probe
    if (correct && pp->center_of_mass) {
      if ((rc = launch_positions_com(pp->positions, p.N, pp->initial_com, stream))) return rc;
    }
  }
  return PTY_OK;
}

// num_iter projection-set iterations (singleslice_ptychography.py:847-983
// with use_projection_scheme): one batch of all N positions in `order`, the
// exit waves advanced in place, object and probe replaced by their normalised
// adjoints, then the same constraints as gradient descent
int proj_run(const pty_problem* problem, const pty_gd_params* params, const pty_projection_params* jp,
             const int* order, double* errors, void* workspace, size_t workspace_bytes, hipStream_t stream,
             const pty_position_params* pp = nullptr) {
  int rc = validate_problem(problem);
  if (rc) return rc;
  if (!params || !jp) return fail(PTY_E_ARG, "params are NULL");
  const pty_problem& p = *problem;
  const pty_gd_params& q = *params;
  if (!frame_supported(p.S))
    return fail(PTY_E_UNSUPPORTED, "projection sets and probe modes take frames of 32, 64 or 128 px, not %d", p.S);
  if (p.object_type != PTY_OBJECT_COMPLEX)
    return fail(PTY_E_UNSUPPORTED, "potential objects run through pty_gd_run (gradient descent)");
  if (q.num_iter < 0 || q.max_batch != p.N)
    return fail(PTY_E_ARG, "projection sets update from all %d positions at once (max_batch=%d)", p.N, q.max_batch);
  if (!order || !errors || !workspace || !jp->exit_waves) return fail(PTY_E_ARG, "null order/errors/workspace/exit waves");
  if (p.fourier_mask) return fail(PTY_E_UNSUPPORTED, "a detector Fourier mask is not defined for projection sets");
  if (!aligned8(jp->exit_waves)) return fail(PTY_E_ARG, "exit waves must be 8-byte aligned");
  if ((rc = validate_position_params(pp, p.positions))) return rc;
  const bool correct = pp != nullptr;
  // position correction needs P_shift and fft2(P_shift * O) of every position
  // from the object before the update: whole-batch ps, chi and far buffers
  const WsLayout L = correct ? layout(p, q, true, false) : layout(p, q, false, true);
  if (workspace_bytes < L.total)
    return fail(PTY_E_WORKSPACE, "workspace %zu bytes < required %zu", workspace_bytes, L.total);
  char* ws = (char*)workspace;
  cf* phat = (cf*)(ws + L.phat);
  cf* ps = (cf*)(ws + L.ps);
  double* err = (double*)(ws + L.err);
  cf* grad = (cf*)(ws + L.grad);
  float* spsum = (float*)(ws + L.spsum);
  float* tilemax = (float*)(ws + L.tilemax);
  double* ppart = (double*)(ws + L.ppart);
  double* cpart = (double*)(ws + L.cpart);
  double* runbuf = (double*)(ws + L.run);
  cf* chi = (cf*)(ws + L.chi);
  cf* far = (cf*)(ws + L.far);
  cf* exits = (cf*)jp->exit_waves;
  // with position correction the layout holds whole-batch staging and no
  // running sums, so every launch covers all N positions
  const int stage = correct ? p.N : stage_size(q);
  cf* obj = (cf*)p.object;
  cf* probe = (cf*)p.probe;
  const int ntiles = num_tiles(p.Px, p.Py);
  const double norm = (double)(float)(p.mean_intensity * (float)p.N);
  const size_t S2 = (size_t)p.S * p.S;
  // the reference's Python-float coefficients, rounded to float32 where they
  // multiply complex64 arrays
  const double x = 1.0 - jp->a - jp->b, y = 1.0 - jp->c;
  int valid = jp->exit_valid != 0;
  bool phat_valid = false;
  for (int it = 0; it < q.num_iter; ++it) {
    if (!phat_valid) {
      if ((rc = launch_fft(probe, phat, 1, p.S, false, 0, stream))) return rc;
      phat_valid = q.fix_probe != 0;
    }
    for (int s0 = 0; s0 < p.N; s0 += stage) {
      const int ns = (p.N - s0) < stage ? (p.N - s0) : stage;
      cf* ex = exits + (size_t)s0 * S2;
      FusedArgs a{};
      a.probe_hat = phat; a.object = obj; a.positions = p.positions; a.amps = p.amplitudes;
      a.order = order + s0; a.N = p.N; a.B = ns; a.Px = p.Px; a.Py = p.Py; a.b0 = 0;
      a.ps_out = ps; a.chi_in = ex; a.chi_out = ex; a.err = err + s0; a.fwd_out = nullptr;
      a.proj_a = (float)jp->a; a.proj_b = (float)jp->b; a.proj_c = (float)jp->c;
      a.proj_x = (float)x; a.proj_y = (float)y; a.exit_valid = valid;
      if ((rc = launch_projection(a, p.S, stream))) return rc;
      GatherArgs g{};
      g.ps = ps; g.chi = ex; g.positions = p.positions; g.order = order + s0; g.N = p.N; g.B = ns; g.S = p.S;
      g.Px = p.Px; g.Py = p.Py; g.tiles_y = gather_tiles_y(p.Py); g.grad = grad; g.spsum = spsum;
      g.tilemax = tilemax;
      if (ns < p.N) {
        g.run = runbuf;
        g.run_first = s0 == 0;
        g.run_last = s0 + ns == p.N;
      }
      if (q.fix_probe) {
        if ((rc = launch_gather(g, p.Px, p.Py, stream))) return rc;
      } else {  // scatter and probe partial sums in one launch
        ProbePartialArgs pa{ex, obj, p.positions, order + s0, p.N, ns, p.S, p.Px, p.Py,
                            ppart + (size_t)(s0 / PROBE_CHUNK) * S2 * 3};
        const int pix_blocks = (p.S * p.S + 255) / 256;
        hipLaunchKernelGGL(k_gather_probe, dim3(ntiles + pix_blocks * probe_nchunk(ns)), dim3(GATHER_THREADS), 0,
                           stream, g, pa, ntiles, pix_blocks);
        if ((rc = check_launch("k_gather_probe"))) return rc;
      }
    }
    valid = 1;
    hipLaunchKernelGGL(k_error_reduce, dim3(1), dim3(1024), 0, stream, err, p.N, norm, errors + it);
    if ((rc = check_launch("k_error_reduce"))) return rc;
    if (correct) {
      // the overlap's far field with the object and probe of this iteration's
      // forward pass (singleslice_ptychography.py:913-926 passes `overlap`);
      // the gradient-descent pass forms it (its chi and err are not used)
      FusedArgs a{};
      a.probe_hat = phat; a.object = obj; a.positions = p.positions; a.amps = p.amplitudes;
      a.order = order; a.N = p.N; a.B = p.N; a.Px = p.Px; a.Py = p.Py; a.b0 = 0;
      a.ps_out = ps; a.chi_out = chi; a.err = err; a.fwd_out = far;
      if ((rc = launch_fused(a, p.S, stream))) return rc;
    }
    if ((rc = launch_object_update(obj, grad, spsum, tilemax, ntiles, p.Px, p.Py, 0.f, q.normalization_min, stream,
                                   1)))
      return rc;
    if (!q.fix_probe) {
      if ((rc = launch_probe_apply(probe, ppart, p.S, p.N, 0.f, q.normalization_min, stream, 1))) return rc;
    }
    if (correct) {  // after the adjoint, with the updated object
      const PosArgs pa = position_args(*pp, ps, far, obj, p.Px, p.Py, p.amplitudes, p.N, order, p.N);
      if ((rc = launch_position_correction(pa, p.S, stream))) return rc;
    }
    if ((rc = launch_object_filters(obj, p.Px, p.Py, q, grad, cpart, stream))) return rc;
    if ((rc = launch_constraints(obj, p.Px, p.Py, p.fov_mask_inverse, q.shrinkage_rad, q.pure_phase_object, cpart,
                                 stream)))
      return rc;
    if (!q.fix_probe && q.fix_probe_com) {
      if ((rc = launch_probe_com(probe, p.S, stream))) return rc;
    }
    if (!q.fix_probe && (rc = launch_probe_extra(probe, 1, p.S, q, stream))) return rc;
    if (correct && pp->center_of_mass) {
      if ((rc = launch_positions_com(pp->positions, p.N, pp->initial_com, stream))) return rc;
    }
  }
  return PTY_OK;
}

// workspace of the mixed-state loop: per-mode staging, far fields, scatter
// outputs and probe partials for whole-batch launches
struct MixedLayout {
  size_t phat, ps, chi, far, err, grad, spsum, tilemax, total, bmax, ppart, pstride, cpart, shift, acc, all;
};
constexpr int MIXED_BLOCKS = 1024;

MixedLayout mixed_layout(const pty_problem& p, const pty_gd_params& q, int K, bool positions = false) {
  MixedLayout L{};
  const size_t S2 = (size_t)p.S * p.S, B = (size_t)q.max_batch, npix = (size_t)p.Px * p.Py;
  size_t o = 0;
  L.phat = o; o = align256(o + K * S2 * 8);
  L.ps = o; o = align256(o + K * B * S2 * 8);
  L.chi = o; o = align256(o + K * B * S2 * 8);
  L.far = o; o = align256(o + K * B * S2 * 8);
  L.err = o; o = align256(o + (size_t)p.N * 8);
  L.grad = o; o = align256(o + K * npix * 8);
  L.spsum = o; o = align256(o + K * npix * 4);
  L.tilemax = o; o = align256(o + (size_t)K * num_tiles(p.Px, p.Py) * 4);
  L.total = o; o = align256(o + npix * 4);
  L.bmax = o; o = align256(o + MIXED_BLOCKS * 4);
  L.pstride = q.fix_probe ? 0 : align256(probe_ws_bytes(p.S, q.max_batch));
  L.ppart = o; o = align256(o + K * L.pstride);
  L.cpart = o; o = align256(o + (size_t)CONSTRAINT_BLOCKS * 16);
  L.shift = o; o = align256(o + 8);
  L.acc = o; o = align256(o + (positions ? 3 * B * S2 * 4 : 0));
  L.all = o;
  return L;
}

// num_iter mixed-state GD iterations (mixedstate_ptychography.py:846-1027 with
// ObjectNDProbeMixedMethodsMixin, ptychographic_methods.py:2696-3000); probe
// [K][S][S]; the per-iteration probe constraint is the recentring by mode 0's
// centre of mass (the orthogonalisation runs on the host between calls)
int mixed_run(const pty_problem* problem, const pty_gd_params* params, int K, const int* orders, double* errors,
              void* workspace, size_t workspace_bytes, hipStream_t stream,
              const pty_position_params* pp = nullptr) {
  int rc = validate_problem(problem);
  if (rc) return rc;
  if (!params) return fail(PTY_E_ARG, "params is NULL");
  const pty_problem& p = *problem;
  const pty_gd_params& q = *params;
  if (!frame_supported(p.S))
    return fail(PTY_E_UNSUPPORTED, "projection sets and probe modes take frames of 32, 64 or 128 px, not %d", p.S);
  if (p.object_type != PTY_OBJECT_COMPLEX)
    return fail(PTY_E_UNSUPPORTED, "potential objects run through pty_gd_run (gradient descent)");
  if (K < 1 || K > 64) return fail(PTY_E_ARG, "num_probes=%d out of range 1..64", K);
  if ((rc = check_index_range(p.Px, p.Py, (long long)q.max_batch * K, p.S))) return rc;
  if (q.num_iter < 0 || q.max_batch <= 0 || q.max_batch > p.N)
    return fail(PTY_E_ARG, "bad num_iter=%d / max_batch=%d for N=%d", q.num_iter, q.max_batch, p.N);
  if (!orders || !errors || !workspace) return fail(PTY_E_ARG, "null orders/errors/workspace");
  if ((rc = validate_position_params(pp, p.positions))) return rc;
  const MixedLayout L = mixed_layout(p, q, K, pp != nullptr);
  if (workspace_bytes < L.all)
    return fail(PTY_E_WORKSPACE, "workspace %zu bytes < required %zu", workspace_bytes, L.all);
  char* ws = (char*)workspace;
  cf* phat = (cf*)(ws + L.phat);
  cf* ps = (cf*)(ws + L.ps);
  cf* chi = (cf*)(ws + L.chi);
  cf* far = (cf*)(ws + L.far);
  double* err = (double*)(ws + L.err);
  cf* grad = (cf*)(ws + L.grad);
  float* spsum = (float*)(ws + L.spsum);
  float* tilemax = (float*)(ws + L.tilemax);
  float* total = (float*)(ws + L.total);
  float* bmax = (float*)(ws + L.bmax);
  double* cpart = (double*)(ws + L.cpart);
  cf* obj = (cf*)p.object;
  cf* probe = (cf*)p.probe;
  const size_t S2 = (size_t)p.S * p.S;
  const long long npix = (long long)p.Px * p.Py;
  const int ntiles = num_tiles(p.Px, p.Py);
  const int nblk = grid_for(npix, 256, MIXED_BLOCKS);
  const std::vector<int> batches = subdivide(p.N, q.max_batch);
  const double norm = (double)(float)(p.mean_intensity * (float)p.N);
  const float c1 = (float)(1.0 - (double)q.normalization_min), c2 = q.normalization_min;
  bool phat_valid = false;
  for (int it = 0; it < q.num_iter; ++it) {
    const int* order_it = orders + (size_t)it * p.N;
    int start = 0;
    for (int nb : batches) {
      const int* ob = order_it + start;
      if (!phat_valid) {
        if ((rc = launch_fft(probe, phat, K, p.S, false, 0, stream))) return rc;
        phat_valid = q.fix_probe != 0;
      }
      FusedArgs a{};
      a.probe_hat = phat; a.object = obj; a.positions = p.positions; a.amps = p.amplitudes;
      a.order = ob; a.N = p.N; a.B = nb; a.Px = p.Px; a.Py = p.Py; a.b0 = 0;
      a.ps_out = ps; a.chi_out = chi; a.err = err + start; a.fwd_out = far; a.K = K;
      a.fmask = p.fourier_mask;
      if ((rc = launch_mixed(a, p.S, stream))) return rc;
      for (int k = 0; k < K; ++k) {
        const cf* ps_k = ps + (size_t)k * nb * S2;
        const cf* chi_k = chi + (size_t)k * nb * S2;
        GatherArgs g{};
        g.ps = ps_k; g.chi = chi_k; g.positions = p.positions; g.order = ob; g.N = p.N; g.B = nb; g.S = p.S;
        g.Px = p.Px; g.Py = p.Py; g.tiles_y = gather_tiles_y(p.Py); g.grad = grad + (size_t)k * npix;
        g.spsum = spsum + (size_t)k * npix; g.tilemax = tilemax + (size_t)k * ntiles;
        if (q.fix_probe) {
          if ((rc = launch_gather(g, p.Px, p.Py, stream))) return rc;
        } else {  // the mode's scatter and probe partial sums in one launch
          ProbePartialArgs pa{chi_k, obj, p.positions, ob, p.N, nb, p.S, p.Px, p.Py,
                              (double*)(ws + L.ppart + k * L.pstride)};
          const int pix_blocks = (p.S * p.S + 255) / 256;
          hipLaunchKernelGGL(k_gather_probe, dim3(ntiles + pix_blocks * probe_nchunk(nb)), dim3(GATHER_THREADS), 0,
                             stream, g, pa, ntiles, pix_blocks);
          if ((rc = check_launch("k_gather_probe"))) return rc;
        }
      }
      hipLaunchKernelGGL(k_mixed_overlap, dim3(nblk), dim3(256), 0, stream, spsum, K, npix, total, bmax);
      if ((rc = check_launch("k_mixed_overlap"))) return rc;
      hipLaunchKernelGGL(k_mixed_object_update, dim3(grid_for(npix, 256)), dim3(256), 0, stream, obj, grad, K,
                         total, bmax, nblk, npix, q.step_size, c1, c2);
      if ((rc = check_launch("k_mixed_object_update"))) return rc;
      if (!q.fix_probe) {
        for (int k = 0; k < K; ++k) {
          if ((rc = launch_probe_apply(probe + (size_t)k * S2, (double*)(ws + L.ppart + k * L.pstride), p.S, nb,
                                       q.step_size, q.normalization_min, stream)))
            return rc;
        }
      }
      if (pp) {  // position correction with the updated object (mixedstate_ptychography.py:942-956)
        PosArgs base = position_args(*pp, ps, far, obj, p.Px, p.Py, p.amplitudes, p.N, ob, nb);
        base.far_modes = K;
        base.acc = (float*)(ws + L.acc);
        for (int k = 0; k < K; ++k) {
          PosArgs ak = base;
          ak.ps = ps + (size_t)k * nb * S2;
          ak.far = far + (size_t)k * S2;
          ak.acc_add = k > 0;
          if (p.S == 32) hipLaunchKernelGGL((k_position_correction<32, true>), dim3(nb), dim3(Frame<32>::T), 0, stream, ak);
          else if (p.S == 64)
            hipLaunchKernelGGL((k_position_correction<64, true>), dim3(nb), dim3(Frame<64>::T), 0, stream, ak);
          else
            hipLaunchKernelGGL((k_position_correction<128, true>), dim3(nb), dim3(Frame<128>::T), 0, stream, ak);
          if ((rc = check_launch("k_position_correction"))) return rc;
        }
        if (p.S == 32) hipLaunchKernelGGL(k_position_solve<32>, dim3(nb), dim3(Frame<32>::T), 0, stream, base);
        else if (p.S == 64) hipLaunchKernelGGL(k_position_solve<64>, dim3(nb), dim3(Frame<64>::T), 0, stream, base);
        else hipLaunchKernelGGL(k_position_solve<128>, dim3(nb), dim3(Frame<128>::T), 0, stream, base);
        if ((rc = check_launch("k_position_solve"))) return rc;
      }
      start += nb;
    }
    hipLaunchKernelGGL(k_error_reduce, dim3(1), dim3(1024), 0, stream, err, p.N, norm, errors + it);
    if ((rc = check_launch("k_error_reduce"))) return rc;
    if ((rc = launch_object_filters(obj, p.Px, p.Py, q, grad, cpart, stream))) return rc;
    if ((rc = launch_constraints(obj, p.Px, p.Py, p.fov_mask_inverse, q.shrinkage_rad, q.pure_phase_object, cpart,
                                 stream)))
      return rc;
    if (!q.fix_probe && q.fix_probe_com) {
      if ((rc = launch_probe_com_modes(probe, K, p.S, (float*)(ws + L.shift), stream))) return rc;
    }
    if (!q.fix_probe && (rc = launch_probe_extra(probe, K, p.S, q, stream))) return rc;
    if (pp && pp->center_of_mass) {
      if ((rc = launch_positions_com(pp->positions, p.N, pp->initial_com, stream))) return rc;
    }
  }
  return PTY_OK;
}

}  // namespace

extern "C" {

int pty_version(void) { return 10000; }

const char* pty_last_error(void) { return g_err.c_str(); }

int pty_fft2_c64(const void* in, void* out, int batch, int S, int inverse, pty_stream_t stream) {
  if (!frame_supported_gd(S)) return fail(PTY_E_UNSUPPORTED, "frame size S=%d not supported", S);
  if (batch < 0 || !in || !out) return fail(PTY_E_ARG, "bad fft2 arguments");
  if (in == out) return fail(PTY_E_ARG, "fft2 must not run in place");
  return launch_fft((const cf*)in, (cf*)out, batch, S, inverse != 0, 1, st(stream));
}

int pty_permute_frames(const float* in, float* out, int count, int S, int direction, pty_stream_t stream) {
  if (!frame_supported_gd(S)) return fail(PTY_E_UNSUPPORTED, "frame size S=%d not supported", S);
  if (count < 0 || !in || !out || in == out) return fail(PTY_E_ARG, "bad permute arguments");
  const long long total = (long long)count * S * S;
  if (total == 0) return PTY_OK;
  const int to_kernel = direction == 0;
  if (S == 256) hipLaunchKernelGGL(k_permute<256>, dim3(grid_for(total, 256, 65536)), dim3(256), 0, st(stream), in, out, total, to_kernel);
  else if (S == 32) hipLaunchKernelGGL(k_permute<32>, dim3(grid_for(total, 256, 65536)), dim3(256), 0, st(stream), in, out, total, to_kernel);
  else if (S == 64) hipLaunchKernelGGL(k_permute<64>, dim3(grid_for(total, 256, 65536)), dim3(256), 0, st(stream), in, out, total, to_kernel);
  else hipLaunchKernelGGL(k_permute<128>, dim3(grid_for(total, 256, 65536)), dim3(256), 0, st(stream), in, out, total, to_kernel);
  return check_launch("k_permute");
}

int pty_probe_spectrum(const void* probe, void* probe_hat, int S, pty_stream_t stream) {
  if (!frame_supported_gd(S)) return fail(PTY_E_UNSUPPORTED, "frame size S=%d not supported", S);
  if (!probe || !probe_hat || probe == probe_hat) return fail(PTY_E_ARG, "bad probe spectrum arguments");
  return launch_fft((const cf*)probe, (cf*)probe_hat, 1, S, false, 0, st(stream));
}

int pty_forward_adjoint_batch(const void* probe_hat, const void* object, int Px, int Py, const float* positions,
                              int N, const float* amplitudes, const int* order, int B, int S, void* ps, void* chi,
                              double* err, void* far, pty_stream_t stream) {
  if (!frame_supported_gd(S)) return fail(PTY_E_UNSUPPORTED, "frame size S=%d not supported", S);
  if (Px < S || Py < S) return fail(PTY_E_ARG, "object %dx%d smaller than frame %d", Px, Py, S);
  if (N <= 0 || B < 0) return fail(PTY_E_ARG, "bad N=%d or B=%d", N, B);
  if (!probe_hat || !object || !positions || !amplitudes || !order || !ps || !chi || !err)
    return fail(PTY_E_ARG, "null buffer in forward_adjoint_batch");
  if (int rc = check_index_range(Px, Py, B, S)) return rc;
  FusedArgs a{};
  a.probe_hat = (const cf*)probe_hat;
  a.object = (const cf*)object;
  a.positions = positions;
  a.amps = amplitudes;
  a.order = order;
  a.N = N; a.B = B; a.Px = Px; a.Py = Py; a.b0 = 0;
  a.ps_out = (cf*)ps; a.chi_out = (cf*)chi; a.err = err; a.fwd_out = (cf*)far;
  return launch_fused(a, S, st(stream));
}

int pty_scatter_patches(const void* ps, const void* chi, const float* positions, int N, const int* order, int B,
                        int S, int Px, int Py, void* grad, float* spsum, float* tilemax, pty_stream_t stream) {
  // 32 x 32 windows are the learned initialiser's stitching patches
  if (!frame_supported_gd(S)) return fail(PTY_E_UNSUPPORTED, "frame size S=%d not supported", S);
  if (Px < S || Py < S || N <= 0 || B < 0) return fail(PTY_E_ARG, "bad scatter sizes");
  if (!ps || !chi || !positions || !order || !grad || !spsum || !tilemax) return fail(PTY_E_ARG, "null buffer in scatter");
  if (int rc = check_index_range(Px, Py, B, S)) return rc;
  GatherArgs a{};
  a.ps = (const cf*)ps; a.chi = (const cf*)chi; a.positions = positions; a.order = order;
  a.N = N; a.B = B; a.S = S; a.Px = Px; a.Py = Py; a.tiles_y = gather_tiles_y(Py);
  a.grad = (cf*)grad; a.spsum = spsum; a.tilemax = tilemax;
  return launch_gather(a, Px, Py, st(stream));
}

int pty_object_update(void* object, const void* grad, const float* spsum, const float* tilemax, int ntiles, int Px,
                      int Py, float step_size, float normalization_min, pty_stream_t stream) {
  if (!object || !grad || !spsum || !tilemax || ntiles <= 0 || Px <= 0 || Py <= 0)
    return fail(PTY_E_ARG, "bad object update arguments");
  return launch_object_update((cf*)object, (const cf*)grad, spsum, tilemax, ntiles, Px, Py, step_size,
                              normalization_min, st(stream));
}

size_t pty_probe_partials_bytes(int S, int B) { return probe_ws_bytes(S, B); }

int pty_scatter_num_tiles(int Px, int Py) { return gather_num_tiles(Px, Py); }

int pty_probe_update(void* probe, const void* chi, const void* object, int Px, int Py, const float* positions, int N,
                     const int* order, int B, int S, float step_size, float normalization_min, void* partials,
                     pty_stream_t stream) {
  if (!frame_supported_gd(S)) return fail(PTY_E_UNSUPPORTED, "frame size S=%d not supported", S);
  if (!probe || !chi || !object || !positions || !order || !partials || B <= 0 || N <= 0)
    return fail(PTY_E_ARG, "bad probe update arguments");
  if (int rc = check_index_range(Px, Py, B, S)) return rc;
  return launch_probe_update((cf*)probe, (const cf*)chi, (const cf*)object, Px, Py, positions, N, order, B, S,
                             step_size, normalization_min, (double*)partials, st(stream));
}

size_t pty_constraint_partials_bytes(void) { return (size_t)CONSTRAINT_BLOCKS * 16; }

int pty_object_constraints(void* object, int Px, int Py, const unsigned char* mask, float shrinkage_rad,
                           int pure_phase_object, void* partials, pty_stream_t stream) {
  if (!object || Px <= 0 || Py <= 0 || (mask && !partials)) return fail(PTY_E_ARG, "bad constraint arguments");
  return launch_constraints((cf*)object, Px, Py, mask, shrinkage_rad, pure_phase_object, (double*)partials,
                            st(stream));
}

int pty_object_gaussian(void* object, int Px, int Py, const double* weights, int radius, int pure_phase_object,
                        void* scratch, pty_stream_t stream) {
  if (!object || !weights || !scratch || Px <= 0 || Py <= 0 || radius < 0)
    return fail(PTY_E_ARG, "bad gaussian filter arguments");
  return launch_object_gaussian((cf*)object, Px, Py, weights, radius, pure_phase_object, scratch, st(stream));
}

int pty_object_butterworth(void* object, int Px, int Py, double q_lowpass, double q_highpass,
                           double butterworth_order, double sampling_x, double sampling_y, void* scratch,
                           void* partials, pty_stream_t stream) {
  if (!object || !scratch || !partials || Px <= 0 || Py <= 0 || !(sampling_x > 0) || !(sampling_y > 0))
    return fail(PTY_E_ARG, "bad butterworth filter arguments");
  return launch_object_butterworth((cf*)object, Px, Py, q_lowpass, q_highpass, butterworth_order, sampling_x,
                                   sampling_y, scratch, (double*)partials, st(stream));
}

size_t pty_resample_workspace_bytes(int batch, int Qy, int Sx) {
  return batch > 0 && Qy > 0 && Sx > 0 ? (size_t)batch * Sx * Qy * 8 : 0;
}

int pty_resample_frames(const float* in, int batch, int Qx, int Qy, const void* Mx, const void* My, int Sx, int Sy,
                        float* out, void* workspace, pty_stream_t stream) {
  if (!in || !out || !Mx || !My || !workspace || batch < 0 || Qx <= 0 || Qy <= 0 || Sx <= 0 || Sy <= 0 ||
      batch > 65535)
    return fail(PTY_E_ARG, "bad resampling arguments");
  if (batch == 0) return PTY_OK;
  cf* T = (cf*)workspace;
  hipLaunchKernelGGL(k_lin_left, dim3((Qy + DFT_T - 1) / DFT_T, (Sx + DFT_T - 1) / DFT_T, batch), dim3(256), 0,
                     st(stream), (const cf*)Mx, in, T, Sx, Qx, Qy);
  int rc = check_launch("k_lin_left");
  if (rc) return rc;
  hipLaunchKernelGGL(k_lin_right, dim3((Sy + DFT_T - 1) / DFT_T, (Sx + DFT_T - 1) / DFT_T, batch), dim3(256), 0,
                     st(stream), (const cf*)T, (const cf*)My, out, Sx, Qy, Sy);
  return check_launch("k_lin_right");
}

int pty_probe_com_constraint(void* probe, int S, pty_stream_t stream) {
  if (!frame_supported(S)) return fail(PTY_E_UNSUPPORTED, "frame size S=%d not supported", S);
  if (!probe) return fail(PTY_E_ARG, "probe is NULL");
  return launch_probe_com((cf*)probe, S, st(stream));
}

// ---- loss and its gradients (autograd backward) ----
namespace {
struct LossLayout {
  size_t phat, ps, chi, pg, err, oacc, ppart, pacc, tmp, total;
};
LossLayout loss_layout(int S, int Px, int Py, int N, int Bc) {
  LossLayout L{};
  const size_t S2 = (size_t)S * S;
  size_t o = 0;
  L.phat = o; o = align256(o + S2 * 8);
  L.ps = o; o = align256(o + (size_t)Bc * S2 * 8);
  L.chi = o; o = align256(o + (size_t)Bc * S2 * 8);
  L.pg = o; o = align256(o + (size_t)Bc * S2 * 8);
  L.err = o; o = align256(o + (size_t)N * 8);
  L.oacc = o; o = align256(o + (size_t)Px * Py * 16);
  L.ppart = o; o = align256(o + (size_t)probe_nchunk(Bc) * S2 * 16);
  L.pacc = o; o = align256(o + S2 * 16);
  L.tmp = o; o = align256(o + S2 * 8);
  L.total = o;
  return L;
}
}  // namespace

size_t pty_loss_workspace_bytes(int S, int Px, int Py, int num_slots, int chunk) {
  if (!frame_supported(S) || num_slots <= 0 || chunk <= 0) return 0;
  return loss_layout(S, Px, Py, num_slots, chunk < num_slots ? chunk : num_slots).total;
}

int pty_loss_gradients(const void* object, const void* probe, int Px, int Py, const float* positions, int N,
                       const float* amplitudes, const int* order, int num_slots, int chunk, int S, double scale,
                       double* loss,
                       void* grad_object, void* grad_probe, void* workspace, size_t workspace_bytes,
                       pty_stream_t stream_) {
  if (!frame_supported(S)) return fail(PTY_E_UNSUPPORTED, "frame size S=%d not supported", S);
  if (Px < S || Py < S || N <= 0 || num_slots <= 0 || chunk <= 0)
    return fail(PTY_E_ARG, "bad loss_gradients sizes");
  if (!object || !probe || !positions || !amplitudes || !order || !loss || !workspace)
    return fail(PTY_E_ARG, "null buffer in loss_gradients");
  const int M = num_slots;
  const int Bc = chunk < M ? chunk : M;
  const LossLayout L = loss_layout(S, Px, Py, M, Bc);
  if (workspace_bytes < L.total) return fail(PTY_E_WORKSPACE, "workspace %zu < %zu", workspace_bytes, L.total);
  hipStream_t stream = st(stream_);
  char* ws = (char*)workspace;
  cf* phat = (cf*)(ws + L.phat);
  cf* ps = (cf*)(ws + L.ps);
  cf* chi = (cf*)(ws + L.chi);
  cf* pg = (cf*)(ws + L.pg);
  double* err = (double*)(ws + L.err);
  double* oacc = (double*)(ws + L.oacc);
  double* ppart = (double*)(ws + L.ppart);
  double* pacc = (double*)(ws + L.pacc);
  cf* tmp = (cf*)(ws + L.tmp);
  const int S2 = S * S;
  int rc;
  if (grad_object) {
    hipLaunchKernelGGL(k_fill_f64, dim3(grid_for((long long)Px * Py * 2, 256)), dim3(256), 0, stream, oacc,
                       (long long)Px * Py * 2, 0.0);
    if ((rc = check_launch("k_fill_f64"))) return rc;
  }
  if (grad_probe) {
    hipLaunchKernelGGL(k_fill_f64, dim3(grid_for((long long)S2 * 2, 256)), dim3(256), 0, stream, pacc,
                       (long long)S2 * 2, 0.0);
    if ((rc = check_launch("k_fill_f64"))) return rc;
  }
  if ((rc = launch_fft((const cf*)probe, phat, 1, S, false, 0, stream))) return rc;
  for (int start = 0; start < M; start += Bc) {
    const int nb = (M - start) < Bc ? (M - start) : Bc;
    FusedArgs a{};
    a.probe_hat = phat; a.object = (const cf*)object; a.positions = positions; a.amps = amplitudes;
    a.order = order + start; a.N = N; a.B = nb; a.Px = Px; a.Py = Py; a.b0 = 0;
    a.ps_out = ps; a.chi_out = chi; a.err = err + start; a.fwd_out = nullptr;
    if ((rc = launch_fused(a, S, stream))) return rc;
    if (grad_object) {
      GatherArgs g{};
      g.ps = ps; g.chi = chi; g.positions = positions; g.order = order + start; g.N = N; g.B = nb; g.S = S;
      g.Px = Px; g.Py = Py; g.tiles_y = gather_tiles_y(Py); g.grad = nullptr; g.spsum = nullptr;
      g.tilemax = nullptr; g.acc = oacc;
      if ((rc = launch_gather(g, Px, Py, stream))) return rc;
    }
    if (grad_probe) {
      a.chi_in = chi;
      a.pgrad_out = pg;
      if (S == 32) hipLaunchKernelGGL(k_probe_grad<32>, dim3(nb), dim3(Frame<32>::T), 0, stream, a);
      else if (S == 64) hipLaunchKernelGGL(k_probe_grad<64>, dim3(nb), dim3(Frame<64>::T), 0, stream, a);
      else hipLaunchKernelGGL(k_probe_grad<128>, dim3(nb), dim3(Frame<128>::T), 0, stream, a);
      if ((rc = check_launch("k_probe_grad"))) return rc;
      const int nchunk = probe_nchunk(nb);
      hipLaunchKernelGGL(k_sum_frames_partial, dim3((S2 + 255) / 256, nchunk), dim3(256), 0, stream, pg, nb, S2,
                         ppart);
      if ((rc = check_launch("k_sum_frames_partial"))) return rc;
      hipLaunchKernelGGL(k_sum_frames_finish, dim3((S2 + 255) / 256), dim3(256), 0, stream, ppart, nchunk, S2,
                         pacc);
      if ((rc = check_launch("k_sum_frames_finish"))) return rc;
    }
  }
  hipLaunchKernelGGL(k_error_reduce, dim3(1), dim3(1024), 0, stream, err, M, 1.0, loss);
  if ((rc = check_launch("k_error_reduce"))) return rc;
  const double g = -2.0 * (double)S2 * scale;
  if (grad_object) {
    hipLaunchKernelGGL(k_acc_to_c64, dim3(grid_for((long long)Px * Py, 256)), dim3(256), 0, stream, oacc,
                       (long long)Px * Py, g, (cf*)grad_object);
    if ((rc = check_launch("k_acc_to_c64"))) return rc;
  }
  if (grad_probe) {
    hipLaunchKernelGGL(k_acc_to_c64, dim3(grid_for(S2, 256)), dim3(256), 0, stream, pacc, (long long)S2, g, tmp);
    if ((rc = check_launch("k_acc_to_c64"))) return rc;
    if ((rc = launch_fft(tmp, (cf*)grad_probe, 1, S, true, 0, stream))) return rc;
  }
  return PTY_OK;
}

int pty_simulate_intensities(const void* probe_hat, const void* object, int Px, int Py, const float* positions,
                             int N, int S, float* out, pty_stream_t stream) {
  if (!frame_supported(S)) return fail(PTY_E_UNSUPPORTED, "frame size S=%d not supported", S);
  if (Px < S || Py < S || N < 0) return fail(PTY_E_ARG, "bad simulate sizes");
  if (!probe_hat || !object || !positions || !out) return fail(PTY_E_ARG, "null buffer in simulate");
  if (N == 0) return PTY_OK;
  FusedArgs a{};
  a.probe_hat = (const cf*)probe_hat; a.object = (const cf*)object; a.positions = positions;
  a.order = nullptr; a.N = N; a.B = N; a.Px = Px; a.Py = Py; a.b0 = 0; a.sim_out = out;
  if (S == 32) hipLaunchKernelGGL(k_simulate<32>, dim3(N), dim3(Frame<32>::T), 0, st(stream), a);
  else if (S == 64) hipLaunchKernelGGL(k_simulate<64>, dim3(N), dim3(Frame<64>::T), 0, st(stream), a);
  else hipLaunchKernelGGL(k_simulate<128>, dim3(N), dim3(Frame<128>::T), 0, st(stream), a);
  return check_launch("k_simulate");
}

size_t pty_gd_workspace_bytes(const pty_problem* problem, const pty_gd_params* params) {
  if (!problem || !params || params->max_batch <= 0) return 0;
  return layout(*problem, *params).total;
}

int pty_gd_run(const pty_problem* problem, const pty_gd_params* params, const int* orders, double* errors,
               void* workspace, size_t workspace_bytes, pty_stream_t stream) {
  return gd_run(problem, params, nullptr, orders, errors, workspace, workspace_bytes, st(stream));
}

int pty_gd_run_noalloc(const pty_problem* problem, const pty_gd_params* params, const int* orders, double* errors,
                       pty_stream_t stream) {
  int rc = validate_problem(problem);
  if (rc) return rc;
  if (!params) return fail(PTY_E_ARG, "params is NULL");
  if (params->max_batch <= 0 || params->max_batch > problem->N)
    return fail(PTY_E_ARG, "bad max_batch=%d for N=%d", params->max_batch, problem->N);
  pty_gd_params q = *params;
  q.pipeline = 0;  // the side stream would be created on first use
  if ((rc = fit_stage_to_scratch(*problem, q))) return rc;
  void* ws = nullptr;
  if ((rc = scratch_pointer(&ws))) return rc;
  return gd_run(problem, &q, nullptr, orders, errors, ws, (size_t)PTY_NOALLOC_SCRATCH_BYTES, st(stream));
}

size_t pty_gd_positions_workspace_bytes(const pty_problem* problem, const pty_gd_params* params) {
  if (!problem || !params || params->max_batch <= 0) return 0;
  return layout(*problem, *params, true).total;
}

int pty_gd_run_positions(const pty_problem* problem, const pty_gd_params* params,
                         const pty_position_params* positions, const int* orders, double* errors,
                         void* workspace, size_t workspace_bytes, pty_stream_t stream) {
  if (!positions) return fail(PTY_E_ARG, "position params are NULL");
  return gd_run(problem, params, positions, orders, errors, workspace, workspace_bytes, st(stream));
}

int pty_position_correction_batch(const pty_position_params* pp, const void* ps, const void* far,
                                  const void* object, int Px, int Py, const float* amplitudes, int N,
                                  const int* order, int B, int S, double* sums, pty_stream_t stream) {
  if (!frame_supported(S)) return fail(PTY_E_UNSUPPORTED, "frame size S=%d not supported", S);
  if (Px < S || Py < S) return fail(PTY_E_ARG, "object %dx%d smaller than frame %d", Px, Py, S);
  if (N <= 0 || B < 0 || B > N) return fail(PTY_E_ARG, "bad N=%d or B=%d", N, B);
  if (!pp || !ps || !far || !object || !amplitudes || !order)
    return fail(PTY_E_ARG, "null buffer in position_correction_batch");
  int rc = validate_position_params(pp, pp->positions);
  if (rc) return rc;
  PosArgs a = position_args(*pp, (const cf*)ps, (const cf*)far, (const cf*)object, Px, Py, amplitudes, N, order, B);
  a.sums = sums;
  return launch_position_correction(a, S, st(stream));
}

size_t pty_projection_workspace_bytes(const pty_problem* problem, const pty_gd_params* params) {
  if (!problem || !params || params->max_batch <= 0) return 0;
  return layout(*problem, *params, false, true).total;
}

int pty_projection_run(const pty_problem* problem, const pty_gd_params* params,
                       const pty_projection_params* projection, const int* order, double* errors,
                       void* workspace, size_t workspace_bytes, pty_stream_t stream) {
  return proj_run(problem, params, projection, order, errors, workspace, workspace_bytes, st(stream));
}

size_t pty_projection_positions_workspace_bytes(const pty_problem* problem, const pty_gd_params* params) {
  if (!problem || !params || params->max_batch <= 0) return 0;
  return layout(*problem, *params, true, false).total;
}

int pty_projection_run_positions(const pty_problem* problem, const pty_gd_params* params,
                                 const pty_projection_params* projection, const pty_position_params* positions,
                                 const int* order, double* errors, void* workspace, size_t workspace_bytes,
                                 pty_stream_t stream) {
  if (!positions) return fail(PTY_E_ARG, "position params are NULL");
  return proj_run(problem, params, projection, order, errors, workspace, workspace_bytes, st(stream), positions);
}

int pty_projection_batch(const void* probe_hat, const void* object, int Px, int Py, const float* positions,
                         int N, const float* amplitudes, const int* order, int B, int S,
                         const pty_projection_params* projection, void* ps, double* err, pty_stream_t stream) {
  if (!frame_supported(S)) return fail(PTY_E_UNSUPPORTED, "frame size S=%d not supported", S);
  if (Px < S || Py < S) return fail(PTY_E_ARG, "object %dx%d smaller than frame %d", Px, Py, S);
  if (N <= 0 || B < 0) return fail(PTY_E_ARG, "bad N=%d or B=%d", N, B);
  if (!probe_hat || !object || !positions || !amplitudes || !order || !projection || !projection->exit_waves ||
      !err)
    return fail(PTY_E_ARG, "null buffer in projection_batch");
  if (int rc = check_index_range(Px, Py, B, S)) return rc;
  const double x = 1.0 - projection->a - projection->b, y = 1.0 - projection->c;
  FusedArgs a{};
  a.probe_hat = (const cf*)probe_hat; a.object = (const cf*)object; a.positions = positions;
  a.amps = amplitudes; a.order = order; a.N = N; a.B = B; a.Px = Px; a.Py = Py; a.b0 = 0;
  a.ps_out = (cf*)ps; a.chi_in = (const cf*)projection->exit_waves; a.chi_out = (cf*)projection->exit_waves;
  a.err = err;
  a.proj_a = (float)projection->a; a.proj_b = (float)projection->b; a.proj_c = (float)projection->c;
  a.proj_x = (float)x; a.proj_y = (float)y; a.exit_valid = projection->exit_valid != 0;
  return launch_projection(a, S, st(stream));
}

size_t pty_mixed_workspace_bytes(const pty_problem* problem, const pty_gd_params* params, int num_probes) {
  if (!problem || !params || params->max_batch <= 0 || num_probes < 1) return 0;
  return mixed_layout(*problem, *params, num_probes).all;
}

int pty_mixed_run(const pty_problem* problem, const pty_gd_params* params, int num_probes, const int* orders,
                  double* errors, void* workspace, size_t workspace_bytes, pty_stream_t stream) {
  return mixed_run(problem, params, num_probes, orders, errors, workspace, workspace_bytes, st(stream));
}

size_t pty_mixed_positions_workspace_bytes(const pty_problem* problem, const pty_gd_params* params,
                                           int num_probes) {
  if (!problem || !params || params->max_batch <= 0 || num_probes < 1) return 0;
  return mixed_layout(*problem, *params, num_probes, true).all;
}

int pty_mixed_run_positions(const pty_problem* problem, const pty_gd_params* params, int num_probes,
                            const pty_position_params* positions, const int* orders, double* errors,
                            void* workspace, size_t workspace_bytes, pty_stream_t stream) {
  if (!positions) return fail(PTY_E_ARG, "position parameters are NULL");
  return mixed_run(problem, params, num_probes, orders, errors, workspace, workspace_bytes, st(stream), positions);
}

int pty_mixed_forward_adjoint_batch(const void* probe_hat, int num_probes, const void* object, int Px, int Py,
                                    const float* positions, int N, const float* amplitudes, const int* order,
                                    int B, int S, void* ps, void* chi, double* err, void* far,
                                    pty_stream_t stream) {
  if (!frame_supported(S)) return fail(PTY_E_UNSUPPORTED, "frame size S=%d not supported", S);
  if (Px < S || Py < S) return fail(PTY_E_ARG, "object %dx%d smaller than frame %d", Px, Py, S);
  if (N <= 0 || B < 0 || num_probes < 1 || num_probes > 64) return fail(PTY_E_ARG, "bad N, B or num_probes");
  if (!probe_hat || !object || !positions || !amplitudes || !order || !ps || !chi || !err || !far)
    return fail(PTY_E_ARG, "null buffer in mixed_forward_adjoint_batch");
  if (int rc = check_index_range(Px, Py, (long long)B * num_probes, S)) return rc;
  FusedArgs a{};
  a.probe_hat = (const cf*)probe_hat; a.object = (const cf*)object; a.positions = positions;
  a.amps = amplitudes; a.order = order; a.N = N; a.B = B; a.Px = Px; a.Py = Py; a.b0 = 0;
  a.ps_out = (cf*)ps; a.chi_out = (cf*)chi; a.err = err; a.fwd_out = (cf*)far; a.K = num_probes;
  return launch_mixed(a, S, st(stream));
}

int pty_probe_com_modes(void* probe, int num_probes, int S, float* shift, pty_stream_t stream) {
  if (!frame_supported(S)) return fail(PTY_E_UNSUPPORTED, "frame size S=%d not supported", S);
  if (!probe || !shift || num_probes < 1) return fail(PTY_E_ARG, "bad probe modes");
  return launch_probe_com_modes((cf*)probe, num_probes, S, shift, st(stream));
}

int pty_positions_com(float* positions, int N, const float* initial_com, pty_stream_t stream) {
  if (!positions || !initial_com || N <= 0) return fail(PTY_E_ARG, "bad positions_com arguments");
  return launch_positions_com(positions, N, initial_com, st(stream));
}

}  // extern "C"

// ---- many reconstructions in one launch sequence per iteration -------------
extern "C" {

size_t pty_multi_workspace_bytes(const pty_multi_problem* problems, int count) {
  MultiPlan plan;
  if (!problems || count <= 0 || multi_make_plan(problems, count, plan)) return 0;
  return multi_layout(problems, plan).total;
}

int pty_multi_run(const pty_multi_problem* problems, int count, int num_iter, int use_graph, void* workspace,
                  size_t workspace_bytes, pty_stream_t stream_) {
  if (!problems || count <= 0 || num_iter < 0 || !workspace) return fail(PTY_E_ARG, "bad multi-run arguments");
  MultiPlan plan;
  int rc = multi_make_plan(problems, count, plan);
  if (rc) return rc;
  const MultiLayout L = multi_layout(problems, plan);
  if (workspace_bytes < L.total) return fail(PTY_E_WORKSPACE, "workspace %zu < %zu", workspace_bytes, L.total);
  hipStream_t stream = st(stream_);
  char* ws = (char*)workspace;
  // descriptors and batch tables
  std::vector<ProblemDev> host(count);
  std::vector<int> rounds;
  std::vector<size_t> rstart(count);
  for (int i = 0; i < count; ++i) {
    rstart[i] = rounds.size();
    rounds.insert(rounds.end(), plan.bstart[i].begin(), plan.bstart[i].end());
    rounds.insert(rounds.end(), plan.bsize[i].begin(), plan.bsize[i].end());
  }
  int* dRounds = (int*)(ws + L.rounds);
  for (int i = 0; i < count; ++i) {
    const pty_problem& p = problems[i].problem;
    const pty_gd_params& q = problems[i].params;
    ProblemDev& d = host[i];
    d.S = p.S; d.Px = p.Px; d.Py = p.Py; d.N = p.N; d.nbatch = (int)plan.bsize[i].size(); d.stage = plan.stage[i];
    d.fix_probe = q.fix_probe; d.fix_probe_com = q.fix_probe_com; d.pure_phase = q.pure_phase_object;
    d.use_mask = p.fov_mask_inverse != nullptr;
    d.step = q.step_size; d.c1 = (float)(1.0 - (double)q.normalization_min); d.c2 = q.normalization_min;
    d.shrink = q.shrinkage_rad;
    d.norm = (double)(float)(p.mean_intensity * (float)p.N);
    d.object = (cf*)p.object; d.probe = (cf*)p.probe; d.phat = (cf*)(ws + L.phat[i]);
    d.positions = p.positions; d.amps = p.amplitudes; d.mask = p.fov_mask_inverse;
    d.orders = problems[i].orders; d.errors = problems[i].errors;
    d.bstart = dRounds + rstart[i];
    d.bsize = dRounds + rstart[i] + plan.bsize[i].size();
    d.ps = (cf*)(ws + L.ps[i]); d.chi = (cf*)(ws + L.chi[i]); d.err = (double*)(ws + L.err[i]);
    d.grad = (cf*)(ws + L.grad[i]); d.spsum = (float*)(ws + L.spsum[i]); d.tilemax = (float*)(ws + L.tilemax[i]);
    d.run = (double*)(ws + L.run[i]); d.ppart = (double*)(ws + L.ppart[i]); d.cpart = (double*)(ws + L.cpart[i]);
  }
  ProblemDev* dP = (ProblemDev*)(ws + L.desc);
  int* dIter = (int*)(ws + L.iter);
  int* dPl64 = (int*)(ws + L.plist64);
  int* dPl128 = (int*)(ws + L.plist128);
  hipError_t e = hipMemcpyAsync(dP, host.data(), sizeof(ProblemDev) * count, hipMemcpyHostToDevice, stream);
  if (e == hipSuccess && !rounds.empty())
    e = hipMemcpyAsync(dRounds, rounds.data(), sizeof(int) * rounds.size(), hipMemcpyHostToDevice, stream);
  if (e == hipSuccess && !plan.plist[0].empty())
    e = hipMemcpyAsync(dPl64, plan.plist[0].data(), sizeof(int) * plan.plist[0].size(), hipMemcpyHostToDevice,
                       stream);
  if (e == hipSuccess && !plan.plist[1].empty())
    e = hipMemcpyAsync(dPl128, plan.plist[1].data(), sizeof(int) * plan.plist[1].size(), hipMemcpyHostToDevice,
                       stream);
  if (e == hipSuccess) e = hipStreamSynchronize(stream);  // host tables are released on return
  if (e != hipSuccess) return fail(PTY_E_LAUNCH, "descriptor upload: %s", hipGetErrorString(e));
  hipLaunchKernelGGL(km_iter_set, dim3(1), dim3(64), 0, stream, dIter, 0);
  if ((rc = check_launch("km_iter_set"))) return rc;
  if (!use_graph) {
    for (int it = 0; it < num_iter; ++it)
      if ((rc = multi_enqueue_iteration(plan, dP, dIter, dPl64, dPl128, stream))) return rc;
    return PTY_OK;
  }
  // one iteration captured once, replayed num_iter times on a private stream
  hipStream_t cs = nullptr;
  hipEvent_t ready = nullptr, done = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  e = hipStreamCreateWithFlags(&cs, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&ready, hipEventDisableTiming);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&done, hipEventDisableTiming);
  if (e == hipSuccess) e = hipEventRecord(ready, stream);
  if (e == hipSuccess) e = hipStreamWaitEvent(cs, ready, 0);
  if (e == hipSuccess) e = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal);
  if (e == hipSuccess) {
    rc = multi_enqueue_iteration(plan, dP, dIter, dPl64, dPl128, cs);
    hipError_t e2 = hipStreamEndCapture(cs, &graph);
    if (!rc && e2 != hipSuccess) rc = fail(PTY_E_LAUNCH, "graph capture: %s", hipGetErrorString(e2));
    if (!rc) {
      e2 = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
      if (e2 != hipSuccess) rc = fail(PTY_E_LAUNCH, "graph instantiate: %s", hipGetErrorString(e2));
    }
    for (int it = 0; !rc && it < num_iter; ++it) {
      e2 = hipGraphLaunch(exec, cs);
      if (e2 != hipSuccess) rc = fail(PTY_E_LAUNCH, "graph launch: %s", hipGetErrorString(e2));
    }
    if (!rc) {
      e2 = hipEventRecord(done, cs);
      if (e2 == hipSuccess) e2 = hipStreamWaitEvent(stream, done, 0);
      if (e2 != hipSuccess) rc = fail(PTY_E_LAUNCH, "graph join: %s", hipGetErrorString(e2));
    }
    (void)hipStreamSynchronize(cs);
  } else {
    rc = fail(PTY_E_LAUNCH, "graph setup: %s", hipGetErrorString(e));
  }
  if (exec) (void)hipGraphExecDestroy(exec);
  if (graph) (void)hipGraphDestroy(graph);
  if (ready) (void)hipEventDestroy(ready);
  if (done) (void)hipEventDestroy(done);
  if (cs) (void)hipStreamDestroy(cs);
  return rc;
}

}  // extern "C"

#include "pty_net.inc"
#include "pty_prep.inc"
