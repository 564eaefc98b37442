// CPU emulator of the HIP kernels' phase sequences (test infrastructure).
//
// Runs the exact phase functions of py4dstem_amd/csrc/pty_core.h for every
// thread id in turn.  Between barriers it records, per LDS word, which thread
// wrote and which read it, and counts a violation for any cross-thread
// read-after-write, write-after-read or write-after-write inside one phase,
// and for any out-of-range LDS or global index it can see.  Built with g++
// by tests/conftest.py; the product library never links it.
#include <cstdio>
#include <cstring>
#include <vector>

static long g_violations = 0;
static char g_first[256] = {0};
static void note_violation(const char* what) {
  if (g_violations++ == 0) std::snprintf(g_first, sizeof g_first, "%s", what);
}
#define PTY_CHECK(cond, what) ((cond) ? (void)0 : note_violation(what))

#include "../../py4dstem_amd/csrc/pty_core.h"

using namespace pty;

// Per LDS word: the writer / reader thread of the current phase (wt, rd; -2 =
// several readers) and the wave that wrote / read it since the last
// work-group barrier (ww, rw; -1 none, -2 several waves).  Across a wave
// barrier only the same wave may touch a word again; anything else is a race.
constexpr int WAVE = 64;

struct LdsCheck {
  cf* mem;
  int n;
  int* wt;
  int* rd;
  int* ww;
  int* rw;
  const int* cur;
  cf ld(int i) const {
    if (i < 0 || i >= n) { note_violation("lds load out of range"); return mkcf(0, 0); }
    if (wt[i] >= 0 && wt[i] != *cur) note_violation("lds read of a word another thread wrote");
    if (ww[i] != -1 && ww[i] != *cur / WAVE) note_violation("lds read of a word another wave wrote (wave barrier only)");
    if (rd[i] == -1) rd[i] = *cur; else if (rd[i] != *cur) rd[i] = -2;
    return mem[i];
  }
  void st(int i, cf v) const {
    if (i < 0 || i >= n) { note_violation("lds store out of range"); return; }
    if (wt[i] >= 0 && wt[i] != *cur) note_violation("lds write-write race");
    if (rd[i] == -2 || (rd[i] >= 0 && rd[i] != *cur)) note_violation("lds write of a word another thread read");
    if (ww[i] != -1 && ww[i] != *cur / WAVE) note_violation("lds write of a word another wave wrote (wave barrier only)");
    if (rw[i] != -1 && rw[i] != *cur / WAVE) note_violation("lds write of a word another wave read (wave barrier only)");
    wt[i] = *cur;
    mem[i] = v;
  }
};

template <int S, class RegsT = typename Fused<S>::Regs> struct EmuRunner {
  using K = Fused<S>;
  static constexpr int T = Frame<S>::T;
  std::vector<cf> mem;
  std::vector<int> wt, rd, ww, rw;
  std::vector<RegsT> regs;
  std::vector<cf> tw;
  int cur = -1;
  const FusedArgs* a;
  EmuRunner(const FusedArgs* args)
      : mem(K::LDS_TOTAL, mkcf(0, 0)), wt(K::LDS_TOTAL, -1), rd(K::LDS_TOTAL, -1),
        ww(K::LDS_TOTAL, -1), rw(K::LDS_TOTAL, -1), regs(T), tw(K::TW_WORDS), a(args) {
    struct TwSt { cf* p; void st(int i, cf v) const { p[i] = v; } } view{tw.data()};
    K::twiddles(0, view, 1);
  }
  template <class Fn> void phase(Fn fn) {
    LdsCheck view{mem.data(), (int)mem.size(), wt.data(), rd.data(), ww.data(), rw.data(), &cur};
    for (int t = 0; t < T; ++t) {
      cur = t;
      fn(t, regs[t], view, tw.data());
    }
    cur = -1;
  }
  void barrier() {
    std::fill(wt.begin(), wt.end(), -1);
    std::fill(rd.begin(), rd.end(), -1);
    std::fill(ww.begin(), ww.end(), -1);
    std::fill(rw.begin(), rw.end(), -1);
  }
  // fold this phase's accesses into the per-wave record, then open a new phase
  void wave_barrier() {
    auto merge = [](int have, int w) { return have == -1 ? w : (have == w ? w : -2); };
    for (size_t i = 0; i < wt.size(); ++i) {
      if (wt[i] >= 0) ww[i] = merge(ww[i], wt[i] / WAVE);
      if (rd[i] >= 0) rw[i] = merge(rw[i], rd[i] / WAVE);
      if (rd[i] == -2) rw[i] = -2;  // several readers: conservatively several waves
      wt[i] = -1;
      rd[i] = -1;
    }
  }
  void reduce_err(int b) {
    double s = 0.0;
    for (int t = 0; t < T; ++t) s += regs[t].err;
    a->err[b] = s;
    barrier();
  }
  void reduce_sums(double* tot) {
    for (int k = 0; k < POS_SUMS; ++k) {
      double s = 0.0;
      for (int t = 0; t < T; ++t) s += regs[t].s[k];
      tot[k] = s;
    }
    barrier();
  }
};

// typed LDS array with the same race/bounds bookkeeping as LdsCheck
template <class T> struct LdsArrCheck {
  T* mem;
  int n;
  int* wt;
  int* rd;
  const int* cur;
  T ld(int i) const {
    if (i < 0 || i >= n) { note_violation("lds load out of range"); return T{}; }
    if (wt[i] >= 0 && wt[i] != *cur) note_violation("lds read of a word another thread wrote");
    if (rd[i] == -1) rd[i] = *cur; else if (rd[i] != *cur) rd[i] = -2;
    return mem[i];
  }
  void st(int i, T v) const {
    if (i < 0 || i >= n) { note_violation("lds store out of range"); return; }
    if (wt[i] >= 0 && wt[i] != *cur) note_violation("lds write-write race");
    if (rd[i] == -2 || (rd[i] >= 0 && rd[i] != *cur)) note_violation("lds write of a word another thread read");
    wt[i] = *cur;
    mem[i] = v;
  }
};

template <class T> struct TrackedArr {
  std::vector<T> mem;
  std::vector<int> wt, rd;
  explicit TrackedArr(int n) : mem(n), wt(n, -1), rd(n, -1) {}
  LdsArrCheck<T> view(const int* cur) { return LdsArrCheck<T>{mem.data(), (int)mem.size(), wt.data(), rd.data(), cur}; }
  void clear() { std::fill(wt.begin(), wt.end(), -1); std::fill(rd.begin(), rd.end(), -1); }
};

struct GatherEmu {
  int cur = -1;
  TrackedArr<int> scan0{GATHER_THREADS}, scan1{GATHER_THREADS};
  TrackedArr<GatherHit> list{GATHER_SEG};
  TrackedArr<float> red{GATHER_THREADS};
  std::vector<Gather::Regs> regs = std::vector<Gather::Regs>(GATHER_THREADS);
  struct View {
    GatherEmu* e;
    LdsArrCheck<int> scan(int i) const { return (i ? e->scan1 : e->scan0).view(&e->cur); }
    LdsArrCheck<GatherHit> list() const { return e->list.view(&e->cur); }
    LdsArrCheck<float> red() const { return e->red.view(&e->cur); }
  };
  template <class Fn> void phase(Fn fn) {
    View v{this};
    for (int t = 0; t < GATHER_THREADS; ++t) {
      cur = t;
      fn(t, regs[t], v);
    }
    cur = -1;
  }
  void barrier() { scan0.clear(); scan1.clear(); list.clear(); red.clear(); }
};

template <int S> static void run_fused(const FusedArgs& a) {
  EmuRunner<S> run(&a);
  for (int b = a.b0; b < a.B; ++b) fused_position<S>(run, a, b);
}

extern "C" {

long emu_violations(void) { return g_violations; }
const char* emu_first_violation(void) { return g_first; }
void emu_reset(void) { g_violations = 0; g_first[0] = 0; }

// Same argument meaning as pty_forward_adjoint_batch (complex arrays are
// interleaved float32 pairs).
int emu_fused(int S, const float* probe_hat, const float* object, int Px, int Py,
              const float* positions, int N, const float* amps, const int* order, int B,
              float* ps, float* chi, double* err, float* fwd_out) {
  FusedArgs a{};
  a.probe_hat = (const cf*)probe_hat;
  a.object = (const cf*)object;
  a.positions = positions;
  a.amps = amps;
  a.order = order;
  a.N = N; a.B = B; a.Px = Px; a.Py = Py; a.b0 = 0;
  a.ps_out = (cf*)ps; a.chi_out = (cf*)chi; a.err = err; a.fwd_out = (cf*)fwd_out;
  if (S == 32) run_fused<32>(a);
  else if (S == 64) run_fused<64>(a);
  else if (S == 128) run_fused<128>(a);
  else return -1;
  return 0;
}

}  // extern "C"

// forward 2-D FFT natural -> permuted and inverse permuted -> natural, with
// the same passes (and phase/barrier structure) the kernels use
template <int S> static void fft_fwd(const cf* in, cf* out) {
  using K = Fused<S>;
  FusedArgs dummy{};
  EmuRunner<S> run(&dummy);
  using Regs = typename K::Regs;
  using F = Frame<S>;
  run.phase([&](int tid, Regs& r, auto lds, const cf* tw) {
    const int cc = F::col_line(tid), n2 = F::col_q(tid);
    for (int m = 0; m < F::N1; ++m) r.v[m] = in[(n2 + F::N2 * m) * S + cc];
    Pass<S>::template A<false>(n2, r.v, tw);
    for (int m = 0; m < F::N1; ++m) lds.st(F::at(n2 + F::N2 * m, cc), r.v[m]);
  });
  run.barrier();
  run.phase([&](int tid, Regs& r, auto lds, const cf* tw) { K::template p_colB<false>(tid, r, lds, tw); });
  run.barrier();
  run.phase([&](int tid, Regs& r, auto lds, const cf* tw) { K::template p_rowA<false>(tid, r, lds, tw); });
  run.wave_barrier();
  run.phase([&](int tid, Regs& r, auto lds, const cf* tw) {
    const int rr = F::row_line(tid), t = F::row_q(tid);
    for (int j = 0; j < F::N1; ++j) r.v[j] = lds.ld(F::at(rr, F::N1 * t + j));
    Pass<S>::template B<false>(t, r.v, tw);
    for (int j = 0; j < F::N1; ++j) out[rr * S + F::N1 * t + j] = r.v[j];
  });
}

template <int S> static void fft_inv(const cf* in, cf* out) {
  using K = Fused<S>;
  FusedArgs dummy{};
  EmuRunner<S> run(&dummy);
  using Regs = typename K::Regs;
  using F = Frame<S>;
  run.phase([&](int tid, Regs& r, auto lds, const cf* tw) {
    const int rr = F::row_line(tid), t = F::row_q(tid);
    for (int j = 0; j < F::N1; ++j) r.v[j] = in[rr * S + F::N1 * t + j];
    Pass<S>::template B<true>(t, r.v, tw);
    for (int j = 0; j < F::N1; ++j) lds.st(F::at(rr, F::N1 * t + j), r.v[j]);
  });
  run.wave_barrier();
  run.phase([&](int tid, Regs& r, auto lds, const cf* tw) { K::template p_rowA<true>(tid, r, lds, tw); });
  run.barrier();
  run.phase([&](int tid, Regs& r, auto lds, const cf* tw) { K::template p_colB<true>(tid, r, lds, tw); });
  run.barrier();
  run.phase([&](int tid, Regs& r, auto lds, const cf* tw) {
    K::p_colAinv_to_regs(tid, r, lds, tw);
    const int cc = F::col_line(tid), n2 = F::col_q(tid);
    for (int m = 0; m < F::N1; ++m) out[(n2 + F::N2 * m) * S + cc] = r.v[m];
  });
}

extern "C" {

int emu_fft2_perm(int S, int inverse, const float* in, float* out) {
  const cf* i = (const cf*)in;
  cf* o = (cf*)out;
  if (S == 32) { if (inverse) fft_inv<32>(i, o); else fft_fwd<32>(i, o); }
  else if (S == 64) { if (inverse) fft_inv<64>(i, o); else fft_fwd<64>(i, o); }
  else if (S == 128) { if (inverse) fft_inv<128>(i, o); else fft_fwd<128>(i, o); }
  else return -1;
  return 0;
}

int emu_kidx(int S, int p) {
  if (S == 32) return Frame<32>::kidx(p);
  else if (S == 64) return Frame<64>::kidx(p);
  if (S == 128) return Frame<128>::kidx(p);
  return -1;
}

int emu_gather(int S, const float* ps, const float* chi, const float* positions, int N,
               const int* order, int B, int Px, int Py, float* grad, float* spsum, float* tilemax) {
  GatherArgs a{};
  a.ps = (const cf*)ps; a.chi = (const cf*)chi; a.positions = positions; a.order = order;
  a.N = N; a.B = B; a.S = S; a.Px = Px; a.Py = Py;
  a.tiles_y = gather_tiles_y(Py);
  a.grad = (cf*)grad; a.spsum = spsum; a.tilemax = tilemax;
  for (int tile = 0; tile < gather_num_tiles(Px, Py); ++tile) {
    GatherEmu run;
    gather_tile(run, a, tile);
  }
  return 0;
}

int emu_gather_num_tiles(int Px, int Py) { return gather_num_tiles(Px, Py); }

// the same scatter in stages of `stage` slots with float64 sums carried over
int emu_gather_staged(int S, const float* ps, const float* chi, const float* positions, int N,
                      const int* order, int B, int stage, int Px, int Py, float* grad, float* spsum,
                      float* tilemax, double* runbuf) {
  for (int s0 = 0; s0 < B; s0 += stage) {
    const int ns = (B - s0) < stage ? (B - s0) : stage;
    GatherArgs a{};
    a.ps = (const cf*)ps + (size_t)s0 * S * S; a.chi = (const cf*)chi + (size_t)s0 * S * S;
    a.positions = positions; a.order = order + s0;
    a.N = N; a.B = ns; a.S = S; a.Px = Px; a.Py = Py; a.tiles_y = gather_tiles_y(Py);
    a.grad = (cf*)grad; a.spsum = spsum; a.tilemax = tilemax;
    a.run = runbuf; a.run_first = s0 == 0; a.run_last = s0 + ns == B;
    for (int tile = 0; tile < gather_num_tiles(Px, Py); ++tile) {
      GatherEmu run;
      gather_tile(run, a, tile);
    }
  }
  return 0;
}

// forward model: out[b] = |fft2(P_shift_b O_b)|^2 (natural order)
int emu_simulate(int S, const float* probe_hat, const float* object, int Px, int Py, const float* positions,
                 int N, float* out) {
  FusedArgs a{};
  a.probe_hat = (const cf*)probe_hat; a.object = (const cf*)object; a.positions = positions;
  a.N = N; a.B = N; a.Px = Px; a.Py = Py; a.sim_out = out;
  if (S == 32) { EmuRunner<32> run(&a); for (int b = 0; b < N; ++b) simulate_position<32>(run, a, b); }
  else if (S == 64) { EmuRunner<64> run(&a); for (int b = 0; b < N; ++b) simulate_position<64>(run, a, b); }
  else if (S == 128) { EmuRunner<128> run(&a); for (int b = 0; b < N; ++b) simulate_position<128>(run, a, b); }
  else return -1;
  return 0;
}

// mixed-state forward/adjoint terms of one batch: same arguments as
// pty_mixed_forward_adjoint_batch
int emu_mixed(int S, const float* probe_hat, int K, const float* object, int Px, int Py, const float* positions,
              int N, const float* amps, const int* order, int B, float* ps, float* chi, double* err, float* far) {
  FusedArgs a{};
  a.probe_hat = (const cf*)probe_hat; a.object = (const cf*)object; a.positions = positions; a.amps = amps;
  a.order = order; a.N = N; a.B = B; a.Px = Px; a.Py = Py; a.b0 = 0; a.K = K;
  a.ps_out = (cf*)ps; a.chi_out = (cf*)chi; a.err = err; a.fwd_out = (cf*)far;
  if (S == 32) { EmuRunner<32> run(&a); for (int b = 0; b < B; ++b) mixed_position<32>(run, a, b); }
  else if (S == 64) { EmuRunner<64> run(&a); for (int b = 0; b < B; ++b) mixed_position<64>(run, a, b); }
  else if (S == 128) { EmuRunner<128> run(&a); for (int b = 0; b < B; ++b) mixed_position<128>(run, a, b); }
  else return -1;
  return 0;
}

// projection-set exit-wave step of one batch: same arguments as
// pty_projection_batch (exit waves advanced in place)
int emu_projection(int S, const float* probe_hat, const float* object, int Px, int Py, const float* positions,
                   int N, const float* amps, const int* order, int B, float* exits, int exit_valid, double pa,
                   double pb, double pc, float* ps, double* err) {
  FusedArgs a{};
  a.probe_hat = (const cf*)probe_hat; a.object = (const cf*)object; a.positions = positions; a.amps = amps;
  a.order = order; a.N = N; a.B = B; a.Px = Px; a.Py = Py; a.b0 = 0;
  a.ps_out = (cf*)ps; a.chi_in = (const cf*)exits; a.chi_out = (cf*)exits; a.err = err;
  a.proj_a = (float)pa; a.proj_b = (float)pb; a.proj_c = (float)pc;
  a.proj_x = (float)(1.0 - pa - pb); a.proj_y = (float)(1.0 - pc); a.exit_valid = exit_valid;
  if (S == 32) { EmuRunner<32> run(&a); for (int b = 0; b < B; ++b) projection_position<32>(run, a, b); }
  else if (S == 64) { EmuRunner<64> run(&a); for (int b = 0; b < B; ++b) projection_position<64>(run, a, b); }
  else if (S == 128) { EmuRunner<128> run(&a); for (int b = 0; b < B; ++b) projection_position<128>(run, a, b); }
  else return -1;
  return 0;
}

// position correction of one batch: same arguments as
// pty_position_correction_batch (positions updated in place)
int emu_position_correction(int S, const float* ps, const float* far, const float* object, int Px, int Py,
                            const float* amps, float* positions, const float* positions_initial, int N,
                            const int* order, int B, float step, int clip_update, double max_update,
                            int clip_total, double max_total, double* sums, int* singular) {
  PosArgs a{};
  a.ps = (const cf*)ps; a.far = (const cf*)far; a.object = (const cf*)object; a.amps = amps;
  a.positions = positions; a.positions_initial = positions_initial; a.order = order;
  a.N = N; a.B = B; a.Px = Px; a.Py = Py; a.step = step;
  a.clip_update = clip_update; a.max_update = max_update;
  a.clip_total = clip_total; a.max_total = max_total;
  a.sums = sums; a.singular = singular;
  if (S == 32) {
    EmuRunner<32, PosCorr<32>::Regs> run(nullptr);
    for (int b = 0; b < B; ++b) position_correction_position<32>(run, a, b);
  }
  else if (S == 64) {
    EmuRunner<64, PosCorr<64>::Regs> run(nullptr);
    for (int b = 0; b < B; ++b) position_correction_position<64>(run, a, b);
  } else if (S == 128) {
    EmuRunner<128, PosCorr<128>::Regs> run(nullptr);
    for (int b = 0; b < B; ++b) position_correction_position<128>(run, a, b);
  } else {
    return -1;
  }
  return 0;
}

// probe-gradient pass: out[b] = fft2(conj(O_b) chi_b) conj(T_b), kernel layout
int emu_probe_grad(int S, const float* chi, const float* object, int Px, int Py, const float* positions,
                   int N, const int* order, int B, float* out) {
  FusedArgs a{};
  a.object = (const cf*)object; a.positions = positions; a.order = order;
  a.N = N; a.B = B; a.Px = Px; a.Py = Py;
  a.chi_in = (const cf*)chi; a.pgrad_out = (cf*)out;
  if (S == 32) { EmuRunner<32> run(&a); for (int b = 0; b < B; ++b) probe_grad_position<32>(run, a, b); }
  else if (S == 64) { EmuRunner<64> run(&a); for (int b = 0; b < B; ++b) probe_grad_position<64>(run, a, b); }
  else if (S == 128) { EmuRunner<128> run(&a); for (int b = 0; b < B; ++b) probe_grad_position<128>(run, a, b); }
  else return -1;
  return 0;
}

}  // extern "C"
