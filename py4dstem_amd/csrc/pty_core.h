// Shared host/device core of the ptychography kernels.
//
// Every kernel body is written as a sequence of "phases": per-thread
// functions separated by work-group barriers.  The HIP kernels in
// pty_hip.hip call the phases with threadIdx.x and __syncthreads(); the CPU
// emulator used by the tests calls the same functions for every thread id in
// turn, with an LDS view that flags cross-thread races inside a phase and
// out-of-range indices.  No phase may read a word another thread writes in
// the same phase, or write a word another thread reads or writes in it.
//
// FFT scheme (frames of S x S complex64, S = N1 * N2):
//   forward  = pass A (N1-point DFTs at stride N2, then twiddles)
//              + pass B (N2-point DFTs of contiguous runs)
//            natural order in -> "permuted" order out: X[k1 + N1*k2] is
//            stored at position k2 + N2*k1.
//   inverse  = pass B' + pass A' (the transposed algorithm):
//            permuted order in -> natural order out.
// Pointwise work between a forward and an inverse transform (the far-field
// amplitude projection, the probe phase ramp) therefore never needs a
// reordering pass; the measured amplitudes and the probe spectrum are simply
// stored in the permuted order once.
#pragma once

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define PTY_HD __host__ __device__ __forceinline__
#else
#include <cmath>
#include <cstdint>
#include <cstring>
#define PTY_HD inline
#endif

#ifndef PTY_CHECK
#define PTY_CHECK(cond, what) ((void)0)
#endif

namespace pty {

struct __attribute__((aligned(8))) cf {
  float x, y;
};

PTY_HD cf mkcf(float x, float y) { cf r; r.x = x; r.y = y; return r; }
PTY_HD cf cadd(cf a, cf b) { return mkcf(a.x + b.x, a.y + b.y); }
PTY_HD cf csub(cf a, cf b) { return mkcf(a.x - b.x, a.y - b.y); }
PTY_HD cf cmul(cf a, cf b) { return mkcf(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x); }
// conj(a) * b
PTY_HD cf cconjmul(cf a, cf b) { return mkcf(a.x * b.x + a.y * b.y, a.x * b.y - a.y * b.x); }
PTY_HD cf cscale(cf a, float s) { return mkcf(a.x * s, a.y * s); }
PTY_HD cf cconj(cf a) { return mkcf(a.x, -a.y); }
PTY_HD float cabs2(cf a) { return a.x * a.x + a.y * a.y; }
// conj(a) * b and |a|^2 with every product rounded to float32 before the
// sum (no fused multiply-add), as NumPy's complex64 arithmetic forms them
PTY_HD cf cconjmul_rn(cf a, cf b) {
#if defined(__clang__)
#pragma clang fp contract(off)
#endif
  return mkcf(a.x * b.x + a.y * b.y, a.x * b.y - a.y * b.x);
}
PTY_HD float cabs2_rn(cf a) {
#if defined(__clang__)
#pragma clang fp contract(off)
#endif
  return a.x * a.x + a.y * a.y;
}

// reciprocal square root (hardware v_rsq_f32 on the GPU)
PTY_HD float frsq(float x) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __builtin_amdgcn_rsqf(x);
#else
  return 1.0f / std::sqrt(x);
#endif
}

// correctly rounded float32 square root: the hardware estimate r is moved to
// a neighbour when x lies beyond the midpoint between them (midpoints squared
// exactly in float64).  NaN stays NaN (as NumPy's sqrt); negative inputs,
// which only rounding could produce here, give 0.
PTY_HD float sqrt_rn(float x) {
#if defined(__HIP_DEVICE_COMPILE__)
  if (!(x > 0.f)) return x < 0.f ? 0.f : x;
  const float r = __builtin_amdgcn_sqrtf(x);
  const float lo = __uint_as_float(__float_as_uint(r) - 1u);
  const float hi = __uint_as_float(__float_as_uint(r) + 1u);
  const double m_lo = ((double)r + (double)lo) * 0.5, m_hi = ((double)r + (double)hi) * 0.5;
  const double xd = (double)x;
  if (xd < m_lo * m_lo) return lo;
  if (xd > m_hi * m_hi) return hi;
  return r;
#else
  return x < 0.f ? 0.f : std::sqrt(x);
#endif
}

// |x + iy| in float32 as sqrt of the float64 sum of the exact squares,
// rounded once more to float32 (the libm hypotf construction)
PTY_HD float hypot_f(float x, float y) {
  const double s = (double)x * (double)x + (double)y * (double)y;
#if defined(__HIP_DEVICE_COMPILE__)
  return (float)__builtin_sqrt(s);
#else
  return (float)std::sqrt(s);
#endif
}

// multiply by the quarter-turn of the transform direction: -i (forward), +i (inverse)
template <bool INV> PTY_HD cf rot90(cf a) { return INV ? mkcf(-a.y, a.x) : mkcf(a.y, -a.x); }

// ---------------------------------------------------------------------------
// in-register DFTs, natural order in and out; INV selects exp(+2 pi i nk/N)
// ---------------------------------------------------------------------------
template <bool INV> PTY_HD void dft4(cf& x0, cf& x1, cf& x2, cf& x3) {
  cf t0 = cadd(x0, x2), t1 = csub(x0, x2);
  cf t2 = cadd(x1, x3), t3 = rot90<INV>(csub(x1, x3));
  x0 = cadd(t0, t2);
  x2 = csub(t0, t2);
  x1 = cadd(t1, t3);
  x3 = csub(t1, t3);
}

template <bool INV> PTY_HD void dft8(cf* x) {
  const float c = 0.70710678118654752f;
  cf e0 = x[0], e1 = x[2], e2 = x[4], e3 = x[6];
  cf o0 = x[1], o1 = x[3], o2 = x[5], o3 = x[7];
  dft4<INV>(e0, e1, e2, e3);
  dft4<INV>(o0, o1, o2, o3);
  // twiddles W8^k (forward: exp(-i pi k / 4))
  cf w1 = INV ? mkcf(c * (o1.x - o1.y), c * (o1.x + o1.y)) : mkcf(c * (o1.x + o1.y), c * (o1.y - o1.x));
  cf w2 = rot90<INV>(o2);
  cf w3 = INV ? mkcf(-c * (o3.x + o3.y), c * (o3.x - o3.y)) : mkcf(c * (o3.y - o3.x), -c * (o3.x + o3.y));
  x[0] = cadd(e0, o0); x[4] = csub(e0, o0);
  x[1] = cadd(e1, w1); x[5] = csub(e1, w1);
  x[2] = cadd(e2, w2); x[6] = csub(e2, w2);
  x[3] = cadd(e3, w3); x[7] = csub(e3, w3);
}

// 16-point DFT as 4 x 4: inner DFT4 over n1 (stride 4), twiddle W16^(n2*k1),
// outer DFT4 over n2.
template <bool INV> PTY_HD void dft16(cf* x) {
  const float c1 = 0.92387953251128676f;  // cos(pi/8)
  const float s1 = 0.38268343236508977f;  // sin(pi/8)
  const float c2 = 0.70710678118654752f;
  const float sg = INV ? 1.f : -1.f;      // sign of the twiddle angle
  cf y[16];
#pragma unroll
  for (int n2 = 0; n2 < 4; ++n2) {
    cf a = x[n2], b = x[n2 + 4], cc = x[n2 + 8], d = x[n2 + 12];
    dft4<INV>(a, b, cc, d);
    y[n2 * 4 + 0] = a; y[n2 * 4 + 1] = b; y[n2 * 4 + 2] = cc; y[n2 * 4 + 3] = d;
  }
  // W16^m for m = n2*k1 in {1,2,3,2,4,6,3,6,9}
  const float wc[10] = {1.f, c1, c2, s1, 0.f, -s1, -c2, -c1, -1.f, -c1};
  const float ws[10] = {0.f, s1, c2, c1, 1.f, c1, c2, s1, 0.f, -s1};
#pragma unroll
  for (int n2 = 1; n2 < 4; ++n2) {
#pragma unroll
    for (int k1 = 1; k1 < 4; ++k1) {
      const int m = n2 * k1;
      if (m == 4) {  // W16^4 is the quarter turn: a swap and a sign
        y[n2 * 4 + k1] = rot90<INV>(y[n2 * 4 + k1]);
      } else {
        cf w = mkcf(wc[m], sg * ws[m]);
        y[n2 * 4 + k1] = cmul(y[n2 * 4 + k1], w);
      }
    }
  }
#pragma unroll
  for (int k1 = 0; k1 < 4; ++k1) {
    cf a = y[k1], b = y[4 + k1], cc = y[8 + k1], d = y[12 + k1];
    dft4<INV>(a, b, cc, d);
    x[k1] = a; x[k1 + 4] = b; x[k1 + 8] = cc; x[k1 + 12] = d;
  }
}

template <int N, bool INV> PTY_HD void dft(cf* x);
template <> PTY_HD void dft<8, false>(cf* x) { dft8<false>(x); }
template <> PTY_HD void dft<8, true>(cf* x) { dft8<true>(x); }
template <> PTY_HD void dft<16, false>(cf* x) { dft16<false>(x); }
template <> PTY_HD void dft<16, true>(cf* x) { dft16<true>(x); }
template <> PTY_HD void dft<4, false>(cf* x) { dft4<false>(x[0], x[1], x[2], x[3]); }
template <> PTY_HD void dft<4, true>(cf* x) { dft4<true>(x[0], x[1], x[2], x[3]); }

// ---------------------------------------------------------------------------
// frame geometry
// ---------------------------------------------------------------------------
template <int S> struct Geo;
template <> struct Geo<32> { static constexpr int N1 = 8, N2 = 4; };
template <> struct Geo<64> { static constexpr int N1 = 8, N2 = 8; };
template <> struct Geo<128> { static constexpr int N1 = 16, N2 = 8; };
// 256 px frames are transformed by line passes through global memory (no
// LDS-resident frame); their "kernel layout" is the natural order, which
// N1 = S, N2 = 1 expresses (kidx(p) = p).  Only the layout helpers are used
// at this size.
template <> struct Geo<256> { static constexpr int N1 = 256, N2 = 1; };

template <int S> struct Frame {
  static constexpr int N1 = Geo<S>::N1;
  static constexpr int N2 = Geo<S>::N2;
  static constexpr int R = N1 / N2;   // pass-B groups per thread
  static constexpr int T = S * N2;    // threads per frame
  static constexpr int PITCH = S + S / N1; // LDS row pitch in complex words
  static constexpr int LDS_WORDS = S * PITCH;
  // One pad word after every N1-word block of a row.  Pass A (stride N2)
  // and pass B (contiguous runs of N1 starting at N1*t) then hit distinct
  // banks, the row pitch S + S/N1 = S + 8 words staggers neighbouring rows,
  // and every per-thread access is base + compile-time offset.
  static PTY_HD int phys(int p) { return p + p / N1; }
  static PTY_HD int at(int row, int col) { return row * PITCH + phys(col); }
  // natural frequency index held at permuted position p
  static PTY_HD int kidx(int p) { return p / N2 + N1 * (p % N2); }
  // fftfreq integer of natural index k
  static PTY_HD int fidx(int k) { return k < S / 2 ? k : k - S; }
  // thread -> (line, q) for row passes and column passes.  A row's N2 lanes
  // are consecutive lanes of one wave, so the exchange between the two
  // passes along rows stays inside the wave (a wave barrier suffices).  A
  // column's lanes are S apart (one wave covers 64 consecutive columns, which
  // keeps global accesses of column passes row-contiguous); exchanges between
  // column passes cross waves and need a work-group barrier.
  static PTY_HD int row_line(int tid) { return tid / N2; }
  static PTY_HD int row_q(int tid) { return tid % N2; }
  static PTY_HD int col_line(int tid) { return tid % S; }
  static PTY_HD int col_q(int tid) { return tid / S; }
  // far-field staging layout: element j of row-pass thread t sits at j*T + t
  // (consecutive lanes write consecutive words); it holds the kernel-layout
  // position (row_line(t), N1 * row_q(t) + j)
  static PTY_HD int far_index(int t, int j) { return j * T + t; }
};

// element i of an array whose base address is the same for the whole work
// group (a frame of the batch, the object): the byte offset stays a 32-bit
// value, so the GPU access is scalar base + 32-bit vector offset instead of
// 64-bit address arithmetic per element.  Every such array is below 4 GiB
// from its base (frames are 512 KiB at most; objects below 23,000^2 pixels).
template <class T> PTY_HD T& elem(T* base, unsigned i) {
  return *reinterpret_cast<T*>(reinterpret_cast<char*>(base) + i * (unsigned)sizeof(T));
}
template <class T> PTY_HD const T& elem(const T* base, unsigned i) {
  return *reinterpret_cast<const T*>(reinterpret_cast<const char*>(base) + i * (unsigned)sizeof(T));
}

// A complex64 array in global memory addressed through a buffer resource on
// the GPU: element offset = a per-thread part (vi) plus a part the same for
// the whole work group (si, kept in a scalar register or the instruction's
// immediate), so a run of loads or stores costs no vector address
// arithmetic.  The resource covers `elems` elements from `p` (elems * 8 below
// 2^31 bytes); the hardware drops stores and zero-fills loads beyond it.
// The host version checks the range (an emulator run fails on an access the
// GPU would drop).
struct CfBuf {
#if defined(__HIP_DEVICE_COMPILE__)
  __amdgpu_buffer_rsrc_t r;
#else
  cf* p;
  unsigned n;
#endif
};

PTY_HD CfBuf cf_buf(const cf* p, unsigned elems) {
  CfBuf b;
#if defined(__HIP_DEVICE_COMPILE__)
  // raw buffer, stride 0; 0x00020000 = the CDNA (gfx9 family) data format word
  b.r = __builtin_amdgcn_make_buffer_rsrc(const_cast<cf*>(p), (short)0, (int)(elems * 8u), 0x00020000);
#else
  b.p = const_cast<cf*>(p);
  b.n = elems;
#endif
  return b;
}

PTY_HD void cf_st(const CfBuf& b, unsigned vi, unsigned si, cf v) {
#if defined(__HIP_DEVICE_COMPILE__)
  typedef unsigned u32x2 __attribute__((ext_vector_type(2)));
  // cache policy 2: non-temporal (the frames are written once per
  // position and read by another launch, so they should not evict the
  // probe spectrum and the object from L2)
  __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2, v), b.r, vi * 8u, si * 8u, 2);
#else
  PTY_CHECK(vi + si < b.n, "buffer store out of range");
  b.p[vi + si] = v;
#endif
}

PTY_HD cf cf_ld(const CfBuf& b, unsigned vi, unsigned si) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __builtin_bit_cast(cf, __builtin_amdgcn_raw_buffer_load_b64(b.r, vi * 8u, si * 8u, 0));
#else
  PTY_CHECK(vi + si < b.n, "buffer load out of range");
  return b.p[vi + si];
#endif
}

// bit patterns between float32 and uint32 (LDS tables of offsets)
PTY_HD float bits_f32(unsigned u) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __uint_as_float(u);
#else
  float f;
  std::memcpy(&f, &u, 4);
  return f;
#endif
}
PTY_HD unsigned f32_bits(float f) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __float_as_uint(f);
#else
  unsigned u;
  std::memcpy(&u, &f, 4);
  return u;
#endif
}

PTY_HD int wrapi(int a, int n) {
  int r = a % n;
  return r < 0 ? r + n : r;
}

PTY_HD unsigned umin(unsigned a, unsigned b) { return a < b ? a : b; }

// p[i] where the address is the same for the whole work group and the data
// do not change during the launch: on the GPU a scalar load (constant
// address space), which does not queue behind the work group's vector loads
template <class T> PTY_HD T uniform_load(const T* p, unsigned i) {
#if defined(__HIP_DEVICE_COMPILE__)
  typedef const __attribute__((address_space(4))) T* constant_ptr;
  return ((constant_ptr)p)[i];
#else
  return p[i];
#endif
}

// wrapi for a value known to lie in [-n, 2n): one conditional correction
// per side instead of an integer division
PTY_HD int wrap1(int a, int n) {
  a += a < 0 ? n : 0;
  return a - (a >= n ? n : 0);
}

// wrapi with the division only off the common path (an index already in
// [0, n) is returned as is)
PTY_HD int wrapi_fast(int a, int n) { return (unsigned)a < (unsigned)n ? a : wrapi(a, n); }

// Passes on one line (a row or a column) of a frame.  v[] is the thread's
// register block; for A passes v[m] sits at line position n2 + N2*m, for B
// passes v[j] sits at position N1*t + j.
template <int S> struct Pass {
  using F = Frame<S>;
  static constexpr int N1 = F::N1, N2 = F::N2, R = F::R;

  // tw: the twiddle table of Fused<S>::twiddles, W^(n2 k1) = exp(-2 pi i n2 k1 / S)
  // stored twice: at n2 (N1 + 1) + k1 (the A pass reads a thread's N1 values
  // at one base address) and at TW_B + k1 (N2 + 1) + n2 (the B' pass reads N2
  // values at one base), so every twiddle load is base + immediate offset.
  // The one-word pad per table row keeps the lanes of a row pass (which
  // differ in n2, or k1) on distinct LDS banks.  A third copy at TW_BS holds
  // the B' layout times 1/S^2, for the last inverse pass of a transform whose
  // output carries the ifft2 normalisation (a power of two: scaling the
  // twiddles is exact, and every later step is a sum or a product by a
  // constant, so the result equals the scaled-after one bit for bit).
  static constexpr int TW_A = N2 * (N1 + 1);
  static constexpr int TW_BS = TW_A + N1 * (N2 + 1);
  static constexpr int TW_WORDS = TW_BS + N1 * (N2 + 1);
  template <bool INV> static PTY_HD void A(int n2, cf* v, const cf* tw) {
    if (!INV) {
      dft<N1, false>(v);
      const cf* w = tw + n2 * (N1 + 1);
#pragma unroll
      for (int k1 = 1; k1 < N1; ++k1) v[k1] = cmul(v[k1], w[k1]);
    } else {
      dft<N1, true>(v);
    }
  }
  // SCALED (inverse only): the outputs carry the factor 1/S^2
  template <bool INV, bool SCALED = false> static PTY_HD void B(int t, cf* v, const cf* tw) {
    static_assert(INV || !SCALED, "only the inverse pass is scaled");
#pragma unroll
    for (int g = 0; g < R; ++g) {
      cf* u = v + g * N2;
      dft<N2, INV>(u);
      if (INV) {
        const cf* w = tw + (SCALED ? TW_BS : TW_A) + (R * t + g) * (N2 + 1);  // k1 = R t + g
        if (SCALED) u[0] = cscale(u[0], 1.0f / (float)(S * S));
#pragma unroll
        for (int n2 = 1; n2 < N2; ++n2) u[n2] = cmul(u[n2], cconj(w[n2]));
      }
    }
  }
  // twiddles for the inverse A' pass are applied in B'; the forward A pass
  // applies its own.  (Inverse: B' multiplies V[k1][n2] by conj W^(n2 k1).)
};

// LDS access policy used by the GPU kernels (the emulator supplies a checking one)
struct LdsPtr {
  cf* p;
  PTY_HD cf ld(int i) const { return p[i]; }
  PTY_HD void st(int i, cf v) const { p[i] = v; }
};

// ---------------------------------------------------------------------------
// Fused per-position kernel: shifted probe, exit wave, far-field projection
// and the adjoint terms, one work group per scan position.
// ---------------------------------------------------------------------------
struct FusedArgs {
  const cf* probe_hat;     // [S][S] fft2(probe), permuted layout
  const cf* object;        // [Px][Py]
  const float* positions;  // [N][2] float32 pixel positions
  const float* amps;       // [N][S][S] measured amplitudes, permuted layout
  const int* order;        // [B] position index of each batch slot
  int N, B, Px, Py;
  int b0;                  // first batch slot handled by this launch
  cf* ps_out;              // [B][S][S] shifted probe P_shift      (natural)
  cf* chi_out;             // [B][S][S] chi = ifft2(A e^{i angle F} - F) (natural)
  double* err;             // [B] sum (A - |F|)^2 per position
  cf* fwd_out;             // optional [B][N1][T] far field F, element j of thread t of the
                           // row pass at [j][t] (coalesced; see far_index), or null
  const cf* chi_in;        // probe-gradient pass: [B][S][S] chi (natural)
  cf* pgrad_out;           // probe-gradient pass: [B][S][S] (kernel layout)
  float* sim_out;          // simulation pass: [B][S][S] |F|^2 (natural)
  // optional detector mask [S][S] in kernel layout: multiplies F before the
  // projection and the far-field update after it (ptychographic_methods.py:
  // 1603-1616, 2773-2792); the staged F of fwd_out stays unmasked
  const float* fmask;
  // mixed state: K probe modes; probe_hat, ps_out and chi_out hold K slices
  // (mode-major: mode k of slot b at (k * B + b) * S * S), fwd_out the
  // [B][K] far fields in the staging layout
  int K;
  // projection sets: coefficients (float32 roundings of the host's float64
  // a, b, c, x = 1 - a - b, y = 1 - c); the exit waves of the previous
  // iteration are read from chi_in and the new ones written to chi_out
  // (slot-indexed, may alias); exit_valid = 0 takes the overlap instead
  float proj_a, proj_b, proj_c, proj_x, proj_y;
  int exit_valid;
  // potential object (object_type 'potential', ptychographic_methods.py:
  // 1522-1523): float32 V [Px][Py]; the patch is exp(i V).  Read by the
  // gradient-descent pass instantiated for it, in place of `object`
  const float* object_pot;
};

// exp(i v) as NumPy's complex64 exp(1j * v): (cos v, sin v)
PTY_HD cf expi(float v) {
  float s, c;
  sincosf(v, &s, &c);
  return mkcf(c, s);
}

template <int S> struct Fused {
  using F = Frame<S>;
  using P = Pass<S>;
  static constexpr int N1 = F::N1, N2 = F::N2, R = F::R, T = F::T;

  struct Regs {
    cf v[N1];
    float amp[N1]; // this thread's measured amplitudes (row pass B layout),
                   // requested phases before the projection reads them
    double err;
    int x0, y0;   // patch anchor, wrapped into [0, Px) x [0, Py)
    int idx;      // position index of this batch slot
  };

  // extra LDS words after the frame: exp(-2 pi i k f / S) ramp factors for
  // rows (k in [0,S)) then columns
  static constexpr int RAMP = F::LDS_WORDS;
  // then S words holding the object row offsets row(i) * Py of the window's
  // rows i (as float bit patterns in .x), written once per position
  static constexpr int ROWS = F::LDS_WORDS + 2 * S;
  static constexpr int LDS_TOTAL = F::LDS_WORDS + 3 * S;

  // twiddle table into LDS: words tid, tid + nthreads, ... < TW_WORDS of the
  // two layouts of Pass<S> (pad words included), each W^(n2 k1) the float32
  // rounding of the float64 value (n2 < N2, k1 < N1, n2 k1 < S)
  static constexpr int TW_WORDS = Pass<S>::TW_WORDS;
  template <class L> static PTY_HD void twiddles(int tid, L tw, int nthreads) {
    if (tid < 0) return;
    for (int i = tid; i < TW_WORDS; i += nthreads) {
      const bool a = i < Pass<S>::TW_A, scaled = i >= Pass<S>::TW_BS;
      const int u = a ? i : i - (scaled ? Pass<S>::TW_BS : Pass<S>::TW_A);
      const int m = a ? (u / (N1 + 1)) * (u % (N1 + 1)) : (u / (N2 + 1)) * (u % (N2 + 1));
#if defined(__HIP_DEVICE_COMPILE__)
      double sn, cs;
      sincospi(-2.0 * m / S, &sn, &cs);
#else
      double cs = std::cos(-2.0 * M_PI * m / S), sn = std::sin(-2.0 * M_PI * m / S);
#endif
      const float k = scaled ? 1.0f / (float)(S * S) : 1.0f;
      tw.st(i, mkcf((float)cs * k, (float)sn * k));
    }
  }

  // ramp factor exp(-2 pi i f k / S) evaluated in float32 the way the
  // reference's fourier_translation_operator does: theta = (-2 pi kf) * frac
  static PTY_HD cf ramp1(int k, float frac) {
    const float kf = (float)F::fidx(k) * (1.0f / S);
    const float th = (-6.28318530717958648f * kf) * frac;
    float s, c;
#if defined(__HIP_DEVICE_COMPILE__)
    sincosf(th, &s, &c);
#else
    s = std::sin(th); c = std::cos(th);
#endif
    return mkcf(c, s);
  }

  // P0: anchor of slot b and the ramp table (threads tid < 2S)
  template <class L> static PTY_HD void p_setup(int tid, const FusedArgs& a, int b, Regs& r, L lds,
                                                bool conj_ramp = false) {
    int idx = a.order ? uniform_load(a.order, (unsigned)b) : b;  // no order: slot b is position b
    PTY_CHECK(idx >= 0 && idx < a.N, "order index out of range");
    idx = ((unsigned)idx < (unsigned)a.N) ? idx : 0;  // memory-safe on bad input
    const float px = uniform_load(a.positions, 2u * idx), py = uniform_load(a.positions, 2u * idx + 1u);
    const float rx = rintf(px), ry = rintf(py);  // round half to even, like np.round
    r.idx = idx;
    r.x0 = wrapi_fast((int)rx, a.Px);
    r.y0 = wrapi_fast((int)ry, a.Py);
    r.err = 0.0;
    if (tid < 2 * S) {
      const int k = tid < S ? tid : tid - S;
      const float frac = tid < S ? px - rx : py - ry;
      // (-2 pi kf) * (-frac) is the exact negation: conj of the forward ramp.
      // The shifted probe's 1/S^2 (ifft2 normalisation) rides on the row
      // factors: a power of two, so scaling before the transform is exact
      // and the result is bit for bit the scaled-after one.
      const cf e = ramp1(k, conj_ramp ? -frac : frac);
      lds.st(RAMP + tid, (tid < S && !conj_ramp) ? cscale(e, 1.0f / (float)(S * S)) : e);
    }
    if (tid < S) {  // periodic row of window row `tid`, times the row length
      int row = r.x0 + F::fidx(tid);
      row += row < 0 ? a.Px : 0;
      row -= row >= a.Px ? a.Px : 0;
      lds.st(ROWS + tid, mkcf(bits_f32((unsigned)row * (unsigned)a.Py), 0.f));
    }
  }

  // P1: row pass B' of ifft2(probe_hat * ramp), the probe spectrum staged in
  // the frame by p_stage_probe
  template <class L> static PTY_HD void p_probe_rowBinv(int tid, const FusedArgs& a, Regs& r, L lds, const cf* tw) {
    const int rr = F::row_line(tid), t = F::row_q(tid);
    const cf ex = lds.ld(RAMP + F::kidx(rr));
#pragma unroll
    for (int j = 0; j < N1; ++j) {
      const cf e = cmul(ex, lds.ld(RAMP + S + F::kidx(N1 * t + j)));
      r.v[j] = cmul(lds.ld(F::at(rr, N1 * t + j)), e);
    }
    P::template B<true>(t, r.v, tw);
#pragma unroll
    for (int j = 0; j < N1; ++j) lds.st(F::at(rr, N1 * t + j), r.v[j]);
  }

  // generic passes reading and writing LDS
  template <bool INV, class L, class RG> static PTY_HD void p_rowA(int tid, RG& r, L lds, const cf* tw) {
    const int rr = F::row_line(tid), n2 = F::row_q(tid);
#pragma unroll
    for (int m = 0; m < N1; ++m) r.v[m] = lds.ld(F::at(rr, n2 + N2 * m));
    P::template A<INV>(n2, r.v, tw);
#pragma unroll
    for (int m = 0; m < N1; ++m) lds.st(F::at(rr, n2 + N2 * m), r.v[m]);
  }
  template <bool INV, class L, class RG> static PTY_HD void p_rowB(int tid, RG& r, L lds, const cf* tw) {
    const int rr = F::row_line(tid), t = F::row_q(tid);
#pragma unroll
    for (int j = 0; j < N1; ++j) r.v[j] = lds.ld(F::at(rr, N1 * t + j));
    P::template B<INV>(t, r.v, tw);
#pragma unroll
    for (int j = 0; j < N1; ++j) lds.st(F::at(rr, N1 * t + j), r.v[j]);
  }
  template <bool INV, class L, class RG, bool SCALED = false>
  static PTY_HD void p_colB(int tid, RG& r, L lds, const cf* tw) {
    const int cc = F::col_line(tid), t = F::col_q(tid);
#pragma unroll
    for (int j = 0; j < N1; ++j) r.v[j] = lds.ld(F::at(N1 * t + j, cc));
    P::template B<INV, SCALED>(t, r.v, tw);
#pragma unroll
    for (int j = 0; j < N1; ++j) lds.st(F::at(N1 * t + j, cc), r.v[j]);
  }
  // column pass A' leaving the natural-order result in registers, times
  // 1/S^2 unless the input already carries it (SCALE = false: the shifted
  // probe, whose row ramp factors hold the 1/S^2)
  template <class L, bool SCALE = true>
  static PTY_HD void p_colAinv_to_regs(int tid, Regs& r, L lds, const cf* tw) {
    const int cc = F::col_line(tid), n2 = F::col_q(tid);
#pragma unroll
    for (int m = 0; m < N1; ++m) r.v[m] = lds.ld(F::at(n2 + N2 * m, cc));
    P::template A<true>(n2, r.v, tw);
    if (SCALE) {
      const float s = 1.0f / (float)(S * S);
#pragma unroll
      for (int m = 0; m < N1; ++m) r.v[m] = cscale(r.v[m], s);
    }
  }

  // periodic object index of window element (i, j); x0/y0 already wrapped
  // and Px, Py >= S, so one conditional correction per axis suffices
  static PTY_HD unsigned object_offset(const FusedArgs& a, const Regs& r, int i, int j) {
    int row = r.x0 + F::fidx(i);
    row += row < 0 ? a.Px : 0;
    row -= row >= a.Px ? a.Px : 0;
    int col = r.y0 + F::fidx(j);
    col += col < 0 ? a.Py : 0;
    col -= col >= a.Py ? a.Py : 0;
    PTY_CHECK(row >= 0 && row < a.Px && col >= 0 && col < a.Py, "object index out of range");
    return (unsigned)row * (unsigned)a.Py + (unsigned)col;
  }

  // P4+P5: shifted probe out, exit wave psi = P * O_patch, then the forward
  // column pass A on psi straight from registers
  // POT: the object is a potential V and the patch exp(i V)
  template <class L, bool POT = false>
  static PTY_HD void p_exitwave_colA(int tid, const FusedArgs& a, int b, Regs& r, L lds, const cf* tw) {
    const int cc = F::col_line(tid), n2 = F::col_q(tid);
    int col = r.y0 + F::fidx(cc);
    col += col < 0 ? a.Py : 0;
    col -= col >= a.Py ? a.Py : 0;
    // window row n2 + N2 m: row offset from the set-up table
    auto offset = [&](int m) {
      const unsigned o = f32_bits(lds.ld(ROWS + n2 + N2 * m).x) + (unsigned)col;
      PTY_CHECK(o < (unsigned)a.Px * (unsigned)a.Py && o == object_offset(a, r, n2 + N2 * m, cc),
                "object index out of range");
      return o;
    };
    auto store_ps = [&]() {
      if (a.ps_out) {
        const CfBuf psb = cf_buf(a.ps_out + (size_t)b * S * S, S * S);
#pragma unroll
        for (int m = 0; m < N1; ++m) cf_st(psb, n2 * S + cc, N2 * m * S, r.v[m]);
      }
    };
    if constexpr (!POT) {
      // a complex patch is requested first, so its loads are in flight
      // during the inverse column pass
      cf op[N1];
      const CfBuf objb = cf_buf(a.object, (unsigned)a.Px * (unsigned)a.Py);
#pragma unroll
      for (int m = 0; m < N1; ++m) op[m] = cf_ld(objb, offset(m), 0u);
#if defined(__HIP_DEVICE_COMPILE__)
      __builtin_amdgcn_sched_barrier(0);
#endif
      p_colAinv_to_regs<L, false>(tid, r, lds, tw);
      store_ps();
#pragma unroll
      for (int m = 0; m < N1; ++m) r.v[m] = cmul(r.v[m], op[m]);
    } else {
      p_colAinv_to_regs<L, false>(tid, r, lds, tw);
      store_ps();
#pragma unroll
      for (int m = 0; m < N1; ++m) r.v[m] = cmul(r.v[m], expi(elem(a.object_pot, offset(m))));
    }
    P::template A<false>(n2, r.v, tw);
#pragma unroll
    for (int m = 0; m < N1; ++m) lds.st(F::at(n2 + N2 * m, cc), r.v[m]);
  }

  // P8: row pass B (completing fft2), amplitude projection, row pass B'
  // PROJ = false: g = A e^{i angle F} - F (gradient descent);
  // PROJ = true:  g = A e^{i angle F}     (projection sets);
  // MASK: a.fmask multiplies F before and g after the projection (a
  // compile-time choice, so the unmasked kernels carry no mask code)
  template <class L, bool PROJ = false, bool MASK = false>
  static PTY_HD void p_project(int tid, const FusedArgs& a, int b, Regs& r, L lds, const cf* tw) {
    const int rr = F::row_line(tid), t = F::row_q(tid);
#pragma unroll
    for (int j = 0; j < N1; ++j) r.v[j] = lds.ld(F::at(rr, N1 * t + j));
    P::template B<false>(t, r.v, tw);
    if (a.fwd_out) {  // one uniform branch for the thread's whole row of F
      const CfBuf farb = cf_buf(a.fwd_out + (size_t)b * S * S, S * S);
#pragma unroll
      for (int j = 0; j < N1; ++j) cf_st(farb, tid, j * T, r.v[j]);
    }
    const float* fm = MASK ? a.fmask + rr * S + N1 * t : nullptr;
    double acc = 0.0;
#pragma unroll
    for (int j = 0; j < N1; ++j) {
      const float mj = MASK ? fm[j] : 1.f;
      const cf f = MASK ? cscale(r.v[j], mj) : r.v[j];
      const float amp = r.amp[j];
      const float m2 = cabs2(f);
      // |F| = 0 takes angle 0 (NumPy's angle(0)): the unit phasor (1, 0) in
      // place of F/|F|; a NaN far field stays NaN.  Selects, not a branch.
      const bool nz = m2 != 0.f;
      const float inv = nz ? frsq(m2) : 0.f;
      const float mag = m2 * inv;
      const float ux = nz ? f.x * inv : 1.f, uy = f.y * inv;
      const cf g = PROJ ? mkcf(amp * ux, amp * uy) : mkcf(amp * ux - f.x, amp * uy - f.y);
      const float d = amp - mag;
      acc += (double)(d * d);
      r.v[j] = MASK ? cscale(g, mj) : g;
    }
    r.err += acc;
    P::template B<true>(t, r.v, tw);
#pragma unroll
    for (int j = 0; j < N1; ++j) lds.st(F::at(rr, N1 * t + j), r.v[j]);
  }

  // projection sets (ptychographic_methods.py:1637-1734): shifted probe out,
  // overlap psi = P * O_patch, previous exit wave e (psi on the first
  // iteration), c psi + y e into the forward column pass, and the partial
  // new exit wave w = x e + a psi (the reference's left-to-right float32
  // evaluation order) written over e: the same thread reads it back in the
  // last phase (keeping it in registers through the transforms would spill)
  template <class L> static PTY_HD void p_proj_colA(int tid, const FusedArgs& a, int b, Regs& r,
                                                    L lds, const cf* tw) {
#if defined(__clang__)
#pragma clang fp contract(off)
#endif
    p_colAinv_to_regs<L, false>(tid, r, lds, tw);
    const int cc = F::col_line(tid), n2 = F::col_q(tid);
    cf* ps_b = a.ps_out ? a.ps_out + (size_t)b * S * S : nullptr;
    const cf* ex_b = a.chi_in + (size_t)b * S * S;
    cf* w_b = a.chi_out + (size_t)b * S * S;
#pragma unroll
    for (int m = 0; m < N1; ++m) {
      const int i = n2 + N2 * m;
      if (ps_b) ps_b[i * S + cc] = r.v[m];
      const cf psi = cmul(r.v[m], a.object[object_offset(a, r, i, cc)]);
      const cf e = a.exit_valid ? ex_b[i * S + cc] : psi;
      r.v[m] = cadd(cscale(psi, a.proj_c), cscale(e, a.proj_y));
      w_b[i * S + cc] = cadd(cscale(e, a.proj_x), cscale(psi, a.proj_a));
    }
    P::template A<false>(n2, r.v, tw);
#pragma unroll
    for (int m = 0; m < N1; ++m) lds.st(F::at(n2 + N2 * m, cc), r.v[m]);
  }

  // projection sets, last phase: new exit wave w + b * ifft2(A e^{i angle F})
  // (w from the first phase, same thread and elements)
  template <class L> static PTY_HD void p_proj_out(int tid, const FusedArgs& a, int b, Regs& r,
                                                   L lds, const cf* tw) {
#if defined(__clang__)
#pragma clang fp contract(off)
#endif
    p_colAinv_to_regs(tid, r, lds, tw);
    const int cc = F::col_line(tid), n2 = F::col_q(tid);
    cf* ex_b = a.chi_out + (size_t)b * S * S;
#pragma unroll
    for (int m = 0; m < N1; ++m) {
      const int o = (n2 + N2 * m) * S + cc;
      ex_b[o] = cadd(ex_b[o], cscale(r.v[m], a.proj_b));
    }
  }

  // The probe spectrum in coalesced pieces: piece k (k < N1/2) of thread tid
  // is the element pair 2 (k T + tid), 2 (k T + tid) + 1 of the [S][S]
  // kernel-layout array, into r.v[2k], r.v[2k+1].  Consecutive lanes read
  // consecutive 16-byte words (a thread's own row segment, 128 bytes at
  // S = 128, would make every load instruction touch 64 cache lines).
  static PTY_HD void p_load_probe_coalesced(int tid, const FusedArgs& a, Regs& r) {
#pragma unroll
    for (int k = 0; k < N1 / 2; ++k) {
      const int e = 2 * (k * T + tid);
      r.v[2 * k] = a.probe_hat[e];
      r.v[2 * k + 1] = a.probe_hat[e + 1];
    }
  }
  // the pieces of p_load_probe_coalesced into the frame's LDS layout (the
  // pair shares a row and a 16-word block: e is even)
  template <class L> static PTY_HD void p_stage_probe(int tid, Regs& r, L lds) {
#pragma unroll
    for (int k = 0; k < N1 / 2; ++k) {
      const int e = 2 * (k * T + tid);
      const int row = e / S, col = e % S;
      lds.st(F::at(row, col), r.v[2 * k]);
      lds.st(F::at(row, col) + 1, r.v[2 * k + 1]);
    }
  }

  // this thread's measured amplitudes (row pass B layout) into r.amp; r.idx
  // set by p_setup
  static PTY_HD void p_load_amps(int tid, const FusedArgs& a, Regs& r) {
    const float* A = a.amps + (size_t)r.idx * S * S + F::row_line(tid) * S + N1 * F::row_q(tid);
#pragma unroll
    for (int j = 0; j < N1; ++j) r.amp[j] = A[j];
  }

  // ---- mixed state (ptychographic_methods.py:2696-2805) -------------------

  // row pass B completes F_k; F_k to the staging buffer; sum_k |F_k|^2 in
  // r.amp (float32, modes in order, |F| as the reference's abs then square)
  template <class L> static PTY_HD void p_mixed_far(int tid, const FusedArgs& a, int b, int k,
                                                    Regs& r, L lds, const cf* tw) {
#if defined(__clang__)
#pragma clang fp contract(off)
#endif
    const int rr = F::row_line(tid), t = F::row_q(tid);
#pragma unroll
    for (int j = 0; j < N1; ++j) r.v[j] = lds.ld(F::at(rr, N1 * t + j));
    P::template B<false>(t, r.v, tw);
    cf* far = a.fwd_out + ((size_t)b * a.K + k) * S * S + tid;
    const float* fm = a.fmask ? a.fmask + rr * S + N1 * t : nullptr;
#pragma unroll
    for (int j = 0; j < N1; ++j) {
      if (fm) r.v[j] = cscale(r.v[j], fm[j]);
      far[(size_t)j * T] = r.v[j];
      const float m = hypot_f(r.v[j].x, r.v[j].y);
      r.amp[j] = (k == 0 ? 0.f : r.amp[j]) + m * m;
    }
  }

  // far = sqrt(sum_k |F_k|^2), err += (A - far)^2, factor A / far (0 where
  // far = 0) kept in r.amp
  static PTY_HD void p_mixed_factor(int tid, const FusedArgs& a, Regs& r) {
#if defined(__clang__)
#pragma clang fp contract(off)
#endif
    const float* A = a.amps + (size_t)r.idx * S * S + F::row_line(tid) * S + N1 * F::row_q(tid);
    double acc = 0.0;
#pragma unroll
    for (int j = 0; j < N1; ++j) {
      const float amp = A[j];
      const float far = sqrt_rn(r.amp[j]);
      const float d = amp - far;
      acc += (double)(d * d);
      // float32 quotient, correctly rounded through float64 (53 >= 2 * 24 + 2)
      r.amp[j] = far == 0.f ? 0.f : (float)((double)amp / (double)far);
    }
    r.err += acc;
  }

  // mode k: F_k from the staging buffer, factor F_k - F_k, row pass B'
  template <class L> static PTY_HD void p_mixed_back(int tid, const FusedArgs& a, int b, int k,
                                                     Regs& r, L lds, const cf* tw) {
#if defined(__clang__)
#pragma clang fp contract(off)
#endif
    const int rr = F::row_line(tid), t = F::row_q(tid);
    const cf* far = a.fwd_out + ((size_t)b * a.K + k) * S * S + tid;
    const float* fm = a.fmask ? a.fmask + rr * S + N1 * t : nullptr;
#pragma unroll
    for (int j = 0; j < N1; ++j) {
      const cf f = far[(size_t)j * T];
      const cf g = csub(cscale(f, r.amp[j]), f);
      r.v[j] = fm ? cscale(g, fm[j]) : g;
    }
    P::template B<true>(t, r.v, tw);
#pragma unroll
    for (int j = 0; j < N1; ++j) lds.st(F::at(rr, N1 * t + j), r.v[j]);
  }

  // probe gradient, first phase: g = conj(O_patch) * chi from global memory,
  // forward column pass A from registers
  template <class L> static PTY_HD void p_pgrad_colA(int tid, const FusedArgs& a, int b, Regs& r,
                                                     L lds, const cf* tw) {
    const int cc = F::col_line(tid), n2 = F::col_q(tid);
    const cf* chi_b = a.chi_in + (size_t)b * S * S;
#pragma unroll
    for (int m = 0; m < N1; ++m) {
      const int i = n2 + N2 * m;
      r.v[m] = cconjmul(a.object[object_offset(a, r, i, cc)], chi_b[i * S + cc]);
    }
    P::template A<false>(n2, r.v, tw);
#pragma unroll
    for (int m = 0; m < N1; ++m) lds.st(F::at(n2 + N2 * m, cc), r.v[m]);
  }

  // probe gradient, last phase: row pass B completes fft2(g); multiply by the
  // conjugate ramp held in the LDS table; store in kernel layout
  template <class L> static PTY_HD void p_pgrad_out(int tid, const FusedArgs& a, int b, Regs& r,
                                                    L lds, const cf* tw) {
    const int rr = F::row_line(tid), t = F::row_q(tid);
#pragma unroll
    for (int j = 0; j < N1; ++j) r.v[j] = lds.ld(F::at(rr, N1 * t + j));
    P::template B<false>(t, r.v, tw);
    const cf ex = lds.ld(RAMP + F::kidx(rr));
    cf* out = a.pgrad_out + (size_t)b * S * S + rr * S + N1 * t;
#pragma unroll
    for (int j = 0; j < N1; ++j) {
      const cf e = cmul(ex, lds.ld(RAMP + S + F::kidx(N1 * t + j)));
      out[j] = cmul(r.v[j], e);
    }
  }

  // simulation: row pass B completes F = fft2(psi); |F|^2 to natural order
  template <class L> static PTY_HD void p_sim_out(int tid, const FusedArgs& a, int b, Regs& r, L lds,
                                                  const cf* tw) {
    const int rr = F::row_line(tid), t = F::row_q(tid);
#pragma unroll
    for (int j = 0; j < N1; ++j) r.v[j] = lds.ld(F::at(rr, N1 * t + j));
    P::template B<false>(t, r.v, tw);
    float* out = a.sim_out + (size_t)b * S * S + F::kidx(rr) * S;
#pragma unroll
    for (int j = 0; j < N1; ++j) out[F::kidx(N1 * t + j)] = cabs2(r.v[j]);
  }

  // P11: chi (natural order) out; PRESCALED: the column pass B' before it
  // already applied the 1/S^2
  template <class L, bool PRESCALED = false>
  static PTY_HD void p_outputs(int tid, const FusedArgs& a, int b, Regs& r, L lds, const cf* tw) {
    p_colAinv_to_regs<L, !PRESCALED>(tid, r, lds, tw);
    const int cc = F::col_line(tid), n2 = F::col_q(tid);
    const CfBuf chib = cf_buf(a.chi_out + (size_t)b * S * S, S * S);
#pragma unroll
    for (int m = 0; m < N1; ++m) cf_st(chib, n2 * S + cc, N2 * m * S, r.v[m]);
  }
};

// The phase sequence for one scan position.  ``run`` executes a phase for
// every thread (GPU: this thread; emulator: all thread ids in turn).
// ``run.barrier()`` is a work-group barrier, needed where data cross waves
// (the ramp table, column passes, row <-> column); ``run.wave_barrier()``
// only orders the LDS accesses of one wave, enough between two passes along
// rows.  Nine work-group barriers per position.
//
// The measured amplitudes are requested in the second phase, not at set-up:
// every work group of a launch starts at about the same time, and set-up
// loads beside the probe spectrum were issue-blocked behind the whole chip's
// simultaneous reads (they are first used by the projection, five phases
// later).  The probe spectrum is read in coalesced pieces
// (p_load_probe_coalesced) and staged through the frame's LDS by the set-up
// phase.  PROBE_LOADED: the caller has already loaded the pieces into r.v
// (issued before its twiddle-table work), else the set-up phase loads them.
template <int S, class Run, bool MASK = false, bool POT = false, bool PROBE_LOADED = false>
PTY_HD void fused_position(Run& run, const FusedArgs& a, int b) {
  using K = Fused<S>;
  using Regs = typename K::Regs;
  // ifft2(probe_hat * ramp): B', A' along rows, then B', A' along columns
  run.phase([&](int tid, Regs& r, auto lds, const cf*) {
    if (!PROBE_LOADED) K::p_load_probe_coalesced(tid, a, r);
    K::p_setup(tid, a, b, r, lds);
    K::p_stage_probe(tid, r, lds);
  });
  run.barrier();
  run.phase([&](int tid, Regs& r, auto lds, const cf* tw) {
    K::p_probe_rowBinv(tid, a, r, lds, tw);
  });
  run.wave_barrier();
  run.phase([&](int tid, Regs& r, auto lds, const cf* tw) {
    K::p_load_amps(tid, a, r);
    K::template p_rowA<true>(tid, r, lds, tw);
  });
  run.barrier();
  run.phase([&](int tid, Regs& r, auto lds, const cf* tw) { K::template p_colB<true>(tid, r, lds, tw); });
  run.barrier();
  // shifted probe in registers; exit wave; fft2 starts with columns
  run.phase([&](int tid, Regs& r, auto lds, const cf* tw) {
    K::template p_exitwave_colA<decltype(lds), POT>(tid, a, b, r, lds, tw);
  });
  run.barrier();
  run.phase([&](int tid, Regs& r, auto lds, const cf* tw) { K::template p_colB<false>(tid, r, lds, tw); });
  run.barrier();
  run.phase([&](int tid, Regs& r, auto lds, const cf* tw) { K::template p_rowA<false>(tid, r, lds, tw); });
  run.wave_barrier();
  // far field -> projection -> start of ifft2
  run.phase([&](int tid, Regs& r, auto lds, const cf* tw) {
    K::template p_project<decltype(lds), false, MASK>(tid, a, b, r, lds, tw);
  });
  run.wave_barrier();
  run.phase([&](int tid, Regs& r, auto lds, const cf* tw) { K::template p_rowA<true>(tid, r, lds, tw); });
  run.barrier();
  run.phase([&](int tid, Regs& r, auto lds, const cf* tw) {
    K::template p_colB<true, decltype(lds), Regs, true>(tid, r, lds, tw);
  });
  run.barrier();
  run.phase([&](int tid, Regs& r, auto lds, const cf* tw) {
    K::template p_outputs<decltype(lds), true>(tid, a, b, r, lds, tw);
  });
  run.reduce_err(b);  // ends with a barrier
}

// Projection-set version of fused_position (same barriers): the exit waves
// of chi_in are advanced in place into chi_out.
template <int S, class Run> PTY_HD void projection_position(Run& run, const FusedArgs& a, int b) {
  using K = Fused<S>;
  using Regs = typename K::Regs;
  run.phase([&](int tid, Regs& r, auto lds, const cf*) {
    K::p_load_probe_coalesced(tid, a, r);
    K::p_setup(tid, a, b, r, lds);
    K::p_stage_probe(tid, r, lds);
  });
  run.barrier();
  run.phase([&](int tid, Regs& r, auto lds, const cf* tw) {
    K::p_probe_rowBinv(tid, a, r, lds, tw);
  });
  run.wave_barrier();
  run.phase([&](int tid, Regs& r, auto lds, const cf* tw) {
    K::p_load_amps(tid, a, r);
    K::template p_rowA<true>(tid, r, lds, tw);
  });
  run.barrier();
  run.phase([&](int tid, Regs& r, auto lds, const cf* tw) { K::template p_colB<true>(tid, r, lds, tw); });
  run.barrier();
  run.phase([&](int tid, Regs& r, auto lds, const cf* tw) { K::p_proj_colA(tid, a, b, r, lds, tw); });
  run.barrier();
  run.phase([&](int tid, Regs& r, auto lds, const cf* tw) { K::template p_colB<false>(tid, r, lds, tw); });
  run.barrier();
  run.phase([&](int tid, Regs& r, auto lds, const cf* tw) { K::template p_rowA<false>(tid, r, lds, tw); });
  run.wave_barrier();
  run.phase([&](int tid, Regs& r, auto lds, const cf* tw) {
    K::template p_project<decltype(lds), true>(tid, a, b, r, lds, tw);
  });
  run.wave_barrier();
  run.phase([&](int tid, Regs& r, auto lds, const cf* tw) { K::template p_rowA<true>(tid, r, lds, tw); });
  run.barrier();
  run.phase([&](int tid, Regs& r, auto lds, const cf* tw) { K::template p_colB<true>(tid, r, lds, tw); });
  run.barrier();
  run.phase([&](int tid, Regs& r, auto lds, const cf* tw) { K::p_proj_out(tid, a, b, r, lds, tw); });
  run.reduce_err(b);  // ends with a barrier
}

// Mixed-state version of fused_position: per mode the shifted probe, exit
// wave and far field (staged), then the common amplitude factor, then per
// mode the inverse transform of factor * F_k - F_k.  ps/chi slices of mode k
// start at k * B * S * S.
template <int S, class Run> PTY_HD void mixed_position(Run& run, const FusedArgs& a, int b) {
  using K = Fused<S>;
  using Regs = typename K::Regs;
  for (int k = 0; k < a.K; ++k) {
    FusedArgs ak = a;
    ak.probe_hat = a.probe_hat + (size_t)k * S * S;
    ak.ps_out = a.ps_out ? a.ps_out + (size_t)k * a.B * S * S : nullptr;
    // mode k's probe spectrum in coalesced pieces staged through the frame
    // (free: the previous mode's far field has left it)
    run.phase([&](int tid, Regs& r, auto lds, const cf*) {
      K::p_load_probe_coalesced(tid, ak, r);
      if (k == 0) K::p_setup(tid, ak, b, r, lds);
      K::p_stage_probe(tid, r, lds);
    });
    run.barrier();
    run.phase([&](int tid, Regs& r, auto lds, const cf* tw) {
      K::p_probe_rowBinv(tid, ak, r, lds, tw);
    });
    run.wave_barrier();
    run.phase([&](int tid, Regs& r, auto lds, const cf* tw) { K::template p_rowA<true>(tid, r, lds, tw); });
    run.barrier();
    run.phase([&](int tid, Regs& r, auto lds, const cf* tw) { K::template p_colB<true>(tid, r, lds, tw); });
    run.barrier();
    run.phase([&](int tid, Regs& r, auto lds, const cf* tw) { K::p_exitwave_colA(tid, ak, b, r, lds, tw); });
    run.barrier();
    run.phase([&](int tid, Regs& r, auto lds, const cf* tw) { K::template p_colB<false>(tid, r, lds, tw); });
    run.barrier();
    run.phase([&](int tid, Regs& r, auto lds, const cf* tw) { K::template p_rowA<false>(tid, r, lds, tw); });
    run.wave_barrier();
    run.phase([&](int tid, Regs& r, auto lds, const cf* tw) { K::p_mixed_far(tid, a, b, k, r, lds, tw); });
    run.barrier();
  }
  run.phase([&](int tid, Regs& r, auto, const cf*) { K::p_mixed_factor(tid, a, r); });
  for (int k = 0; k < a.K; ++k) {
    FusedArgs ak = a;
    ak.chi_out = a.chi_out + (size_t)k * a.B * S * S;
    run.phase([&](int tid, Regs& r, auto lds, const cf* tw) { K::p_mixed_back(tid, a, b, k, r, lds, tw); });
    run.wave_barrier();
    run.phase([&](int tid, Regs& r, auto lds, const cf* tw) { K::template p_rowA<true>(tid, r, lds, tw); });
    run.barrier();
    run.phase([&](int tid, Regs& r, auto lds, const cf* tw) { K::template p_colB<true>(tid, r, lds, tw); });
    run.barrier();
    run.phase([&](int tid, Regs& r, auto lds, const cf* tw) { K::p_outputs(tid, ak, b, r, lds, tw); });
    run.barrier();
  }
  run.reduce_err(b);  // ends with a barrier
}

// Forward model only: intensities |fft2(P_shift * O_patch)|^2 of slot b.
template <int S, class Run> PTY_HD void simulate_position(Run& run, const FusedArgs& a, int b) {
  using K = Fused<S>;
  using Regs = typename K::Regs;
  run.phase([&](int tid, Regs& r, auto lds, const cf*) {
    K::p_load_probe_coalesced(tid, a, r);
    K::p_setup(tid, a, b, r, lds);
    K::p_stage_probe(tid, r, lds);
  });
  run.barrier();
  run.phase([&](int tid, Regs& r, auto lds, const cf* tw) {
    K::p_probe_rowBinv(tid, a, r, lds, tw);
  });
  run.wave_barrier();
  run.phase([&](int tid, Regs& r, auto lds, const cf* tw) { K::template p_rowA<true>(tid, r, lds, tw); });
  run.barrier();
  run.phase([&](int tid, Regs& r, auto lds, const cf* tw) { K::template p_colB<true>(tid, r, lds, tw); });
  run.barrier();
  run.phase([&](int tid, Regs& r, auto lds, const cf* tw) { K::p_exitwave_colA(tid, a, b, r, lds, tw); });
  run.barrier();
  run.phase([&](int tid, Regs& r, auto lds, const cf* tw) { K::template p_colB<false>(tid, r, lds, tw); });
  run.barrier();
  run.phase([&](int tid, Regs& r, auto lds, const cf* tw) { K::template p_rowA<false>(tid, r, lds, tw); });
  run.wave_barrier();
  run.phase([&](int tid, Regs& r, auto lds, const cf* tw) { K::p_sim_out(tid, a, b, r, lds, tw); });
  run.barrier();
}

// Backward of the probe path for one position:
//   pgrad_out[b] = fft2(conj(O_patch) * chi) * conj(T_b)   (kernel layout)
template <int S, class Run> PTY_HD void probe_grad_position(Run& run, const FusedArgs& a, int b) {
  using K = Fused<S>;
  using Regs = typename K::Regs;
  run.phase([&](int tid, Regs& r, auto lds, const cf*) { K::p_setup(tid, a, b, r, lds, true); });
  run.barrier();
  run.phase([&](int tid, Regs& r, auto lds, const cf* tw) { K::p_pgrad_colA(tid, a, b, r, lds, tw); });
  run.barrier();
  run.phase([&](int tid, Regs& r, auto lds, const cf* tw) { K::template p_colB<false>(tid, r, lds, tw); });
  run.barrier();
  run.phase([&](int tid, Regs& r, auto lds, const cf* tw) { K::template p_rowA<false>(tid, r, lds, tw); });
  run.wave_barrier();
  run.phase([&](int tid, Regs& r, auto lds, const cf* tw) { K::p_pgrad_out(tid, a, b, r, lds, tw); });
  run.barrier();  // the ramp table and frame are rewritten by the next position
}

// ---------------------------------------------------------------------------
// Position correction (phase/ptychographic_methods.py:2071-2225), one work
// group per scan position of the mini-batch, after the batch's object update.
// With F = fft2(P_shift * O_old patch) kept by the fused pass and
// G_x, G_y = fft2(P_shift * O_new patch moved one pixel along rows / columns):
//   p_x = 2 Re((F - G_x) conj F),  p_y = 2 Re((F - G_y) conj F),
//   d = A^2 - |F|^2                                  (float32, per pixel)
//   u = [[sum p_x^2, sum p_x p_y], [., sum p_y^2]]^-1 [sum p_x d, sum p_y d]
// (float64 sums), then positions -= step * u with the optional clips.  All
// sums run over every frequency, so the permuted order of F, G and the
// amplitudes needs no reordering.
// ---------------------------------------------------------------------------
struct PosArgs {
  const cf* ps;                    // [B][S][S] shifted probes of the batch (natural)
  const cf* far;                   // [B][N1][T] F of the fused pass (staging layout)
  const cf* object;                // [Px][Py] object after the batch's update
  const float* amps;               // [N][S][S] measured amplitudes (permuted)
  float* positions;                // [N][2] slot b updates its own entry only
  const float* positions_initial;  // [N][2]
  const int* order;                // [B] position index of each slot
  int N, B, Px, Py;
  float step;                      // positions_step_size
  int clip_update, clip_total;     // max_position_update/total_distance given
  double max_update, max_total;    // those limits in pixels
  double* sums;                    // optional [B][5]: p_x^2, p_x p_y, p_y^2, p_x d, p_y d
  int* singular;                   // incremented per singular 2x2 system
  // probe modes (ptychographic_methods.py:2187-2189, 2736-2740): `far`
  // points at mode k of slot 0 and slot b's is far_modes slots further; the
  // per-mode pass adds p_x, p_y and |F_k|^2 of every pixel into acc
  // ([B][3][N1][T], staging order; acc_add = 0 for the first mode stores);
  // far_modes 0 means 1
  int far_modes;
  float* acc;
  int acc_add;
};

constexpr int POS_SUMS = 5;

// 2x2 solve, step, clips and the position write of one slot (one thread).
// Float32 where the reference's arrays are float32 (step product, norms,
// distance test, the subtraction), float64 where it compares against or
// divides by the float64 limits.
PTY_HD void position_update(const PosArgs& a, int b, int idx, const double* s) {
#if defined(__clang__)
#pragma clang fp contract(off)
#endif
  if (a.sums) {
    for (int k = 0; k < POS_SUMS; ++k) a.sums[(size_t)b * POS_SUMS + k] = s[k];
  }
  const double det = s[0] * s[2] - s[1] * s[1];
  float fx = 0.f, fy = 0.f;
  if (det != 0.0) {
    const double ux = (s[2] * s[3] - s[1] * s[4]) / det;
    const double uy = (s[0] * s[4] - s[1] * s[3]) / det;
    fx = (float)ux * a.step;
    fy = (float)uy * a.step;
  } else if (a.singular) {
#if defined(__HIP_DEVICE_COMPILE__)
    atomicAdd(a.singular, 1);
#else
    ++*a.singular;
#endif
  }
  if (a.clip_update) {
    const float n = sqrtf(fx * fx + fy * fy);
    if ((double)n > a.max_update) {
      const double q = (double)n / a.max_update;
      fx = (float)((double)fx / q);
      fy = (float)((double)fy / q);
    }
  }
  const float cx = a.positions[2 * idx], cy = a.positions[2 * idx + 1];
  if (a.clip_total) {
    const float dx = (cx - a.positions_initial[2 * idx]) - fx;
    const float dy = (cy - a.positions_initial[2 * idx + 1]) - fy;
    if ((double)sqrtf(dx * dx + dy * dy) > a.max_total) {
      fx = 0.f;
      fy = 0.f;
    }
  }
  a.positions[2 * idx] = cx - fx;
  a.positions[2 * idx + 1] = cy - fy;
}

template <int S> struct PosCorr {
  using F = Frame<S>;
  using P = Pass<S>;
  static constexpr int N1 = F::N1, N2 = F::N2, T = F::T;
  static constexpr int LDS_TOTAL = F::LDS_WORDS;

  // Register budget (1024 threads: 128 VGPRs): F is loaded at set-up and
  // stays live; the amplitudes are loaded only for the second transform's
  // last row passes and the sums exist only in the final phase.
  struct Regs {
    cf v[N1];
    cf f[N1];      // F at this thread's row-pass-B positions
    float amp[N1]; // measured amplitudes at the same positions
    float px[N1];  // p_x at the same positions
    double s[POS_SUMS];
    int x0, y0, idx;
  };

  // anchor of slot b; the loads of F stay in flight across the first transform
  static PTY_HD void p_setup(int tid, const PosArgs& a, int b, Regs& r) {
    int idx = a.order ? a.order[b] : b;
    PTY_CHECK(idx >= 0 && idx < a.N, "order index out of range");
    idx = ((unsigned)idx < (unsigned)a.N) ? idx : 0;
    r.idx = idx;
    r.x0 = wrapi((int)rintf(a.positions[2 * idx]), a.Px);
    r.y0 = wrapi((int)rintf(a.positions[2 * idx + 1]), a.Py);
    const size_t modes = a.far_modes > 0 ? (size_t)a.far_modes : 1;  // 0: single mode
    const cf* Fb = a.far + (size_t)b * modes * S * S + tid;
#pragma unroll
    for (int j = 0; j < N1; ++j) r.f[j] = Fb[(size_t)j * T];
  }

  // row pass A of the second transform; the amplitude loads issued first
  // complete behind it
  template <class L> static PTY_HD void p_rowA_amps(int tid, const PosArgs& a, Regs& r, L lds,
                                                    const cf* tw) {
    const float* A = a.amps + (size_t)r.idx * S * S + F::row_line(tid) * S + N1 * F::row_q(tid);
#pragma unroll
    for (int j = 0; j < N1; ++j) r.amp[j] = A[j];
    Fused<S>::template p_rowA<false>(tid, r, lds, tw);
  }

  // column pass A of fft2(P_shift * O[rows + DX, cols + DY]) from global memory
  template <int DX, int DY, class L>
  static PTY_HD void p_colA(int tid, const PosArgs& a, int b, Regs& r, L lds, const cf* tw) {
    const int cc = F::col_line(tid), n2 = F::col_q(tid);
    const cf* ps_b = a.ps + (size_t)b * S * S;
    int col = r.y0 + F::fidx(cc) + DY;
    col += col < 0 ? a.Py : 0;
    col -= col >= a.Py ? a.Py : 0;
#pragma unroll
    for (int m = 0; m < N1; ++m) {
      const int i = n2 + N2 * m;
      int row = r.x0 + F::fidx(i) + DX;
      row += row < 0 ? a.Px : 0;
      row -= row >= a.Px ? a.Px : 0;
      PTY_CHECK(row >= 0 && row < a.Px && col >= 0 && col < a.Py, "object index out of range");
      r.v[m] = cmul(ps_b[i * S + cc], a.object[(unsigned)row * (unsigned)a.Py + (unsigned)col]);
    }
    P::template A<false>(n2, r.v, tw);
#pragma unroll
    for (int m = 0; m < N1; ++m) lds.st(F::at(n2 + N2 * m, cc), r.v[m]);
  }

  // 2 Re((F - G) conj F), the reference's complex64 product (real part)
  static PTY_HD float partial(cf f, cf g) {
#if defined(__clang__)
#pragma clang fp contract(off)
#endif
    const cf d = csub(f, g);
    return 2.f * (d.x * f.x + d.y * f.y);
  }

  // row pass B completes G_x; p_x kept in registers
  template <class L> static PTY_HD void p_partial_x(int tid, Regs& r, L lds, const cf* tw) {
    const int rr = F::row_line(tid), t = F::row_q(tid);
#pragma unroll
    for (int j = 0; j < N1; ++j) r.v[j] = lds.ld(F::at(rr, N1 * t + j));
    P::template B<false>(t, r.v, tw);
#pragma unroll
    for (int j = 0; j < N1; ++j) r.px[j] = partial(r.f[j], r.v[j]);
  }

  // row pass B completes G_y; the five float64 sums of this thread
  template <class L> static PTY_HD void p_partial_y_sums(int tid, Regs& r, L lds, const cf* tw) {
#if defined(__clang__)
#pragma clang fp contract(off)
#endif
    const int rr = F::row_line(tid), t = F::row_q(tid);
#pragma unroll
    for (int j = 0; j < N1; ++j) r.v[j] = lds.ld(F::at(rr, N1 * t + j));
    P::template B<false>(t, r.v, tw);
    double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0, s4 = 0.0;
#pragma unroll
    for (int j = 0; j < N1; ++j) {
      const double px = r.px[j];
      const double py = partial(r.f[j], r.v[j]);
      const double d = r.amp[j] * r.amp[j] - cabs2(r.f[j]);  // float32 difference
      s0 += px * px;
      s1 += px * py;
      s2 += py * py;
      s3 += px * d;
      s4 += py * d;
    }
    r.s[0] = s0; r.s[1] = s1; r.s[2] = s2; r.s[3] = s3; r.s[4] = s4;
  }
  // probe modes: row pass B completes G_y; p_x, p_y and |F|^2 (abs, then
  // squared, as the reference's mode sum) added into the accumulators in
  // mode order, float32 as the reference's sum over the mode axis
  template <class L> static PTY_HD void p_partial_y_acc(int tid, const PosArgs& a, int b, Regs& r, L lds,
                                                        const cf* tw) {
#if defined(__clang__)
#pragma clang fp contract(off)
#endif
    const int rr = F::row_line(tid), t = F::row_q(tid);
#pragma unroll
    for (int j = 0; j < N1; ++j) r.v[j] = lds.ld(F::at(rr, N1 * t + j));
    P::template B<false>(t, r.v, tw);
    float* acc = a.acc + (size_t)b * 3 * S * S + tid;
#pragma unroll
    for (int j = 0; j < N1; ++j) {
      const float py = partial(r.f[j], r.v[j]);
      const float m = hypot_f(r.f[j].x, r.f[j].y);
      const float i2 = m * m;
      float* e = acc + (size_t)j * T;
      if (a.acc_add) {
        e[0] = e[0] + r.px[j];
        e[S * S] = e[S * S] + py;
        e[2 * S * S] = e[2 * S * S] + i2;
      } else {
        e[0] = r.px[j];
        e[S * S] = py;
        e[2 * S * S] = i2;
      }
    }
  }
};

// The phase sequence of one position's correction (runner as for
// fused_position; ``run.reduce_sums`` adds r.s over the work group and hands
// the totals to thread 0).  Nine work-group barriers.
template <int S, class Run, bool ACC = false>
PTY_HD void position_correction_position(Run& run, const PosArgs& a, int b) {
  using K = PosCorr<S>;
  using Regs = typename K::Regs;
  using FK = Fused<S>;
  run.phase([&](int tid, Regs& r, auto, const cf*) { K::p_setup(tid, a, b, r); });
  run.barrier();  // twiddle table
  run.phase([&](int tid, Regs& r, auto lds, const cf* tw) { K::template p_colA<1, 0>(tid, a, b, r, lds, tw); });
  run.barrier();
  run.phase([&](int tid, Regs& r, auto lds, const cf* tw) { FK::template p_colB<false>(tid, r, lds, tw); });
  run.barrier();
  run.phase([&](int tid, Regs& r, auto lds, const cf* tw) { FK::template p_rowA<false>(tid, r, lds, tw); });
  run.wave_barrier();
  run.phase([&](int tid, Regs& r, auto lds, const cf* tw) { K::p_partial_x(tid, r, lds, tw); });
  run.barrier();
  run.phase([&](int tid, Regs& r, auto lds, const cf* tw) { K::template p_colA<0, 1>(tid, a, b, r, lds, tw); });
  run.barrier();
  run.phase([&](int tid, Regs& r, auto lds, const cf* tw) { FK::template p_colB<false>(tid, r, lds, tw); });
  run.barrier();
  run.phase([&](int tid, Regs& r, auto lds, const cf* tw) { K::p_rowA_amps(tid, a, r, lds, tw); });
  run.wave_barrier();
  if (ACC) {  // one probe mode: the sums follow once every mode is in
    run.phase([&](int tid, Regs& r, auto lds, const cf* tw) { K::p_partial_y_acc(tid, a, b, r, lds, tw); });
    return;
  }
  run.phase([&](int tid, Regs& r, auto lds, const cf* tw) { K::p_partial_y_sums(tid, r, lds, tw); });
  double tot[POS_SUMS];
  run.reduce_sums(tot);  // ends with a barrier; totals valid in thread 0
  run.phase([&](int tid, Regs& r, auto, const cf*) {
    if (tid == 0) position_update(a, b, r.idx, tot);
  });
}

// ---------------------------------------------------------------------------
// Deterministic scatter-add by object pixel ("sort by pixel").  One work group
// owns a TR x TC object tile, GT_RPT pixels of one column per thread (the
// column test of a hit is shared by them; the 64 lanes of a wave read 512
// contiguous bytes of a patch row).  For each segment of the
// batch it (1) tests which positions' windows touch the tile, (2) compacts
// the hits into an in-order list in LDS with a parallel prefix scan, and
// (3) every thread sums, in float64 and in list (= batch) order,
// conj(P_shift) * chi and |P_shift|^2 of the staged patch element covering its
// pixel.  Per pixel the sum order is the batch order, as in the reference's
// np.bincount; the float32 products are formed exactly as the reference does.
// ---------------------------------------------------------------------------
#ifndef PTY_GT_R
#define PTY_GT_R 8
#define PTY_GT_C 64
#endif
#ifndef PTY_GATHER_ROWS
#define PTY_GATHER_ROWS 2
#endif
#ifndef PTY_GATHER_UNROLL
#define PTY_GATHER_UNROLL 4
#endif
constexpr int GT_R = PTY_GT_R;           // tile rows
constexpr int GT_C = PTY_GT_C;           // tile columns (512-byte patch-row reads)
constexpr int GT_RPT = PTY_GATHER_ROWS;  // tile rows per thread, GT_R / GT_RPT apart
constexpr int GT_RSTEP = GT_R / GT_RPT;
static_assert(GT_R % GT_RPT == 0, "tile rows must split evenly over a thread's rows");
constexpr int GATHER_THREADS = GT_RSTEP * GT_C;
constexpr int GATHER_CAND = 4;           // candidates tested per thread per segment
constexpr int GATHER_SEG = GATHER_THREADS * GATHER_CAND;

struct GatherArgs {
  const cf* ps;            // [B][S][S] shifted probes
  const cf* chi;           // [B][S][S] exit-wave updates
  const float* positions;  // [N][2]
  const int* order;        // [B]
  int N, B, S, Px, Py, tiles_y;
  cf* grad;                // [Px][Py] float32 of the float64 sums
  float* spsum;            // [Px][Py]
  float* tilemax;          // [num_tiles] max of spsum over the tile
  double* acc;             // optional [Px][Py][2]: running float64 sums (loss gradient)
  double* run;             // optional [Px][Py][3]: running float64 sums carried across
                           // stages of one mini-batch (re, im, |P|^2)
  int run_first;           // this stage starts the running sums at zero
  int run_last;            // this stage writes grad/spsum/tilemax from the sums
};

// a window that meets the tile: its first staged element, and the window
// coordinates (row, column) of the tile origin, in [0, Px) x [0, Py)
struct GatherHit {
  unsigned base;
  int dr, dc;
};

// LDS of one gather work group, and the view the GPU runner hands to phases
struct GatherShared {
  int scan[2][GATHER_THREADS];
  GatherHit list[GATHER_SEG];
  float red[GATHER_THREADS];
};

template <class T> struct LdsArr {
  T* p;
  PTY_HD T ld(int i) const { return p[i]; }
  PTY_HD void st(int i, T v) const { p[i] = v; }
};

struct GatherSharedView {
  GatherShared* s;
  PTY_HD LdsArr<int> scan(int i) const { return LdsArr<int>{s->scan[i]}; }
  PTY_HD LdsArr<GatherHit> list() const { return LdsArr<GatherHit>{s->list}; }
  PTY_HD LdsArr<float> red() const { return LdsArr<float>{s->red}; }
};

PTY_HD int gather_tiles_y(int Py) { return (Py + GT_C - 1) / GT_C; }
PTY_HD int gather_num_tiles(int Px, int Py) { return ((Px + GT_R - 1) / GT_R) * gather_tiles_y(Py); }

struct Gather {
  struct Regs {
    int cnt;
    int total;
    double re[GT_RPT], im[GT_RPT], s[GT_RPT];  // rows lr + g * GT_RSTEP of the tile
    int hdr[GATHER_CAND], hdc[GATHER_CAND];  // tile origin in this thread's candidate windows
    unsigned hits;                           // bit k: candidate k hits the tile
  };

  static PTY_HD void origin(const GatherArgs& a, int tile, int& R0, int& C0) {
    R0 = (tile / a.tiles_y) * GT_R;
    C0 = (tile % a.tiles_y) * GT_C;
  }

  // hit test of slot b's window against the tile; (dr, dc): the tile origin
  // in window coordinates, wrapped into the object (R0, C0 in the object, the
  // anchor wrapped into it and S <= Px, Py keep them within one wrap)
  static PTY_HD bool test(const GatherArgs& a, int R0, int C0, int b, int& dr, int& dc) {
    int idx = a.order[b];
    idx = ((unsigned)idx < (unsigned)a.N) ? idx : 0;
    const int x0 = wrapi_fast((int)rintf(a.positions[2 * idx]), a.Px);
    const int y0 = wrapi_fast((int)rintf(a.positions[2 * idx + 1]), a.Py);
    const int S = a.S;
    const int tr = (a.Px - R0) < GT_R ? (a.Px - R0) : GT_R;
    const int tc = (a.Py - C0) < GT_C ? (a.Py - C0) : GT_C;
    dr = wrap1(R0 - x0 + S / 2, a.Px);
    dc = wrap1(C0 - y0 + S / 2, a.Py);
    return (dr < S || dr + tr > a.Px) && (dc < S || dc + tc > a.Py);
  }

  // running sums: zero, or carried over from the previous stage of the batch
  static PTY_HD void p_init(int tid, const GatherArgs& a, int tile, Regs& r) {
    int R0, C0;
    origin(a, tile, R0, C0);
    const int col = C0 + tid % GT_C;
#pragma unroll
    for (int g = 0; g < GT_RPT; ++g) {
      r.re[g] = 0.0; r.im[g] = 0.0; r.s[g] = 0.0;
      const int row = R0 + tid / GT_C + g * GT_RSTEP;
      if (!a.run || a.run_first || row >= a.Px || col >= a.Py) continue;
      const double* q = a.run + ((size_t)row * a.Py + col) * 3;
      r.re[g] = q[0]; r.im[g] = q[1]; r.s[g] = q[2];
    }
  }

  // phase 1: count hits among this thread's candidates of the segment
  template <class L> static PTY_HD void p_count(int tid, const GatherArgs& a, int tile, int seg0,
                                                Regs& r, L scan0) {
    int R0, C0, dr, dc;
    origin(a, tile, R0, C0);
    int c = 0;
    r.hits = 0u;
#pragma unroll
    for (int k = 0; k < GATHER_CAND; ++k) {
      const int b = seg0 + tid * GATHER_CAND + k;
      if (b < a.B && test(a, R0, C0, b, dr, dc)) {
        ++c;
        r.hits |= 1u << k;
        r.hdr[k] = dr;
        r.hdc[k] = dc;
      }
    }
    r.cnt = c;
    scan0.st(tid, c);
  }

  // Hillis-Steele inclusive scan step over the per-thread counts
  template <class L> static PTY_HD void p_scan(int tid, int d, L src, L dst) {
    int v = src.ld(tid);
    if (tid >= d) v += src.ld(tid - d);
    dst.st(tid, v);
  }

  // phase 3: write this thread's hits at its offset of the in-order list
  // (the hit tests of p_count, kept in registers)
  template <class L, class H> static PTY_HD void p_fill(int tid, const GatherArgs& a, int tile,
                                                       int seg0, Regs& r, L incl, H list) {
    (void)tile;
    const unsigned SS = (unsigned)(a.S * a.S);
    int o = incl.ld(tid) - r.cnt;
    r.total = incl.ld(GATHER_THREADS - 1);
#pragma unroll
    for (int k = 0; k < GATHER_CAND; ++k) {
      if (r.hits & (1u << k)) {
        GatherHit h;
        h.base = (unsigned)(seg0 + tid * GATHER_CAND + k) * SS;
        h.dr = r.hdr[k];
        h.dc = r.hdc[k];
        list.st(o++, h);
      }
    }
  }

  // one pixel of the window: conj(P_shift) * chi and |P_shift|^2 added in
  // float64
  static PTY_HD void add_term(const GatherArgs& a, unsigned o, double& re, double& im, double& s) {
#if defined(__clang__)
#pragma clang fp contract(off)
#endif
    const cf p = elem(a.ps, o);
    const cf g = cconjmul_rn(p, elem(a.chi, o));
    re += (double)g.x;
    im += (double)g.y;
    s += (double)cabs2_rn(p);
  }

  // phase 4: accumulate this thread's pixels over the segment's hit list,
  // PTY_GATHER_UNROLL hits per round (the loads first, then the float64 adds
  // in list order); the column test is shared by the thread's rows
  template <class H> static PTY_HD void p_accumulate(int tid, const GatherArgs& a, int tile,
                                                     Regs& r, H list) {
    // separately rounded float32 products (no fused multiply-add), in every
    // loop alike: the sums do not depend on how the hits fall into rounds
#if defined(__clang__)
#pragma clang fp contract(off)
#endif
    int R0, C0;
    origin(a, tile, R0, C0);
    const int lr = tid / GT_C, lc = tid % GT_C;
    if (R0 + lr >= a.Px || C0 + lc >= a.Py) return;
    const int S = a.S;
    const unsigned uS = (unsigned)S, half = uS / 2;
    // rows of this thread inside the object (the first one is)
    bool row_ok[GT_RPT];
#pragma unroll
    for (int g = 0; g < GT_RPT; ++g) row_ok[g] = R0 + lr + g * GT_RSTEP < a.Px;
    constexpr int U = PTY_GATHER_UNROLL;
    int h = 0;
    for (; h + U <= r.total; h += U) {
      cf p[U][GT_RPT], c[U][GT_RPT];
      bool in[U][GT_RPT];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const GatherHit e = list.ld(h + u);
        unsigned dc = (unsigned)(e.dc + lc);
        dc = umin(dc, dc - (unsigned)a.Py);
        const bool col_in = dc < uS;
        const unsigned wc = dc ^ half;
#pragma unroll
        for (int g = 0; g < GT_RPT; ++g) {
          unsigned dr = (unsigned)(e.dr + lr + g * GT_RSTEP);
          dr = umin(dr, dr - (unsigned)a.Px);
          in[u][g] = col_in && dr < uS && row_ok[g];
          const unsigned o = e.base + (in[u][g] ? ((dr ^ half) * uS | wc) : 0u);
          p[u][g] = elem(a.ps, o);
          c[u][g] = elem(a.chi, o);
        }
      }
#pragma unroll
      for (int u = 0; u < U; ++u) {
#pragma unroll
        for (int g = 0; g < GT_RPT; ++g) {
          if (in[u][g]) {
            const cf gg = cconjmul_rn(p[u][g], c[u][g]);
            r.re[g] += (double)gg.x;
            r.im[g] += (double)gg.y;
            r.s[g] += (double)cabs2_rn(p[u][g]);
          }
        }
      }
    }
    for (; h < r.total; ++h) {
      const GatherHit e = list.ld(h);
      unsigned dc = (unsigned)(e.dc + lc);
      dc = umin(dc, dc - (unsigned)a.Py);
      if (dc >= uS) continue;
      const unsigned wc = dc ^ half;
#pragma unroll
      for (int g = 0; g < GT_RPT; ++g) {
        unsigned dr = (unsigned)(e.dr + lr + g * GT_RSTEP);
        dr = umin(dr, dr - (unsigned)a.Px);
        if (dr < uS && row_ok[g]) add_term(a, e.base + ((dr ^ half) * uS | wc), r.re[g], r.im[g], r.s[g]);
      }
    }
  }

  // phase 5: store the pixel, stage its spsum for the tile maximum
  template <class L> static PTY_HD void p_store(int tid, const GatherArgs& a, int tile, Regs& r,
                                                L red) {
    int R0, C0;
    origin(a, tile, R0, C0);
    const int col = C0 + tid % GT_C;
    float m = 0.f;
#pragma unroll
    for (int g = 0; g < GT_RPT; ++g) {
      const int row = R0 + tid / GT_C + g * GT_RSTEP;
      if (row >= a.Px || col >= a.Py) continue;
      const size_t o = (size_t)row * a.Py + col;
      if (a.acc) {
        a.acc[2 * o] += r.re[g];
        a.acc[2 * o + 1] += r.im[g];
      } else if (a.run && !a.run_last) {
        double* q = a.run + o * 3;
        q[0] = r.re[g]; q[1] = r.im[g]; q[2] = r.s[g];
      } else {
        a.grad[o] = mkcf((float)r.re[g], (float)r.im[g]);
        const float v = (float)r.s[g];
        a.spsum[o] = v;
        m = v > m ? v : m;
      }
    }
    red.st(tid, m);
  }
};

// Whole gather sequence for one tile (see fused_position for the runner).
template <class Run> PTY_HD void gather_tile(Run& run, const GatherArgs& a, int tile) {
  using G = Gather;
  using Regs = G::Regs;
  run.phase([&](int tid, Regs& r, auto& sh) { G::p_init(tid, a, tile, r); (void)sh; });
  for (int seg0 = 0; seg0 < a.B; seg0 += GATHER_SEG) {
    run.phase([&](int tid, Regs& r, auto& sh) { G::p_count(tid, a, tile, seg0, r, sh.scan(0)); });
    run.barrier();
    int src = 0;
    for (int d = 1; d < GATHER_THREADS; d <<= 1) {
      run.phase([&](int tid, Regs&, auto& sh) { G::p_scan(tid, d, sh.scan(src), sh.scan(1 - src)); });
      run.barrier();
      src = 1 - src;
    }
    run.phase([&](int tid, Regs& r, auto& sh) { G::p_fill(tid, a, tile, seg0, r, sh.scan(src), sh.list()); });
    run.barrier();
    run.phase([&](int tid, Regs& r, auto& sh) { G::p_accumulate(tid, a, tile, r, sh.list()); });
    run.barrier();
  }
  run.phase([&](int tid, Regs& r, auto& sh) { G::p_store(tid, a, tile, r, sh.red()); });
  run.barrier();
  for (int d = GATHER_THREADS / 2; d > 0; d >>= 1) {
    run.phase([&](int tid, Regs&, auto& sh) {
      if (tid < d) {
        auto red = sh.red();
        const float x = red.ld(tid), y = red.ld(tid + d);
        red.st(tid, x > y ? x : y);
      }
    });
    run.barrier();
  }
  run.phase([&](int tid, Regs&, auto& sh) {
    if (tid == 0 && a.tilemax && (!a.run || a.run_last)) a.tilemax[tile] = sh.red().ld(0);
  });
}

}  // namespace pty
