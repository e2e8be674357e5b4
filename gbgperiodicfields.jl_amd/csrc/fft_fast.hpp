// Register-resident pencil FFT for power-of-two lengths (compile-time N).
//
// A pencil of length N is spread over TP = N/E threads; thread t holds the E
// elements x[t + TP*m], m = 0..E-1, in registers.  The transform is the
// radix-8 Stockham autosort factorisation (radix 8 while N/Ls allows, then
// one closing radix 2 or 4 stage):
//   stage (R, Ls), butterfly j in [0, N/R), k = j mod Ls:
//     v_r = x[j + r N/R] * w_{Ls R}^{r k};  V = DFT_R(v);  y[(j-k) R + k + r Ls] = V_r
// With j = t + TP*b the first stage reads exactly the registers a thread
// already holds and the last stage writes back to the same pattern, so a
// pencil is loaded from and stored to global memory directly in that
// pattern; only the stages in between exchange through LDS (one write, one
// barrier, one read, one barrier per exchange).  Twiddles come from an
// N-entry table e^{-2 pi i m/N} computed in extended precision on the host:
// w^1, w^2, w^4 are read and w^3, w^5, w^6, w^7 formed by one product each.
#pragma once

#include "fft_core.hpp"

namespace gbg {
namespace fast {

// wave-scope barrier: orders the wave's LDS writes before its later reads
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// SYNC = 0: the pencil's threads span several waves (workgroup barrier);
// SYNC = 1: all of a pencil's threads sit in one wave.
template <int SYNC>
__device__ __forceinline__ void lds_sync() {
  if (SYNC)
    wave_sync();
  else
    __syncthreads();
}

template <int N, int LS>
struct StageR {
  static constexpr int rem = N / LS;
  static constexpr int R = (rem % 8 == 0) ? 8 : rem;
};

template <int DIR, typename T>
__device__ __forceinline__ cpx<T> tw_load(const cpx<T>* __restrict__ tw, int m) {
  cpx<T> w = tw[m];
  if (DIR > 0) w.y = -w.y;
  return w;
}

template <typename T, int R, int DIR>
__device__ __forceinline__ void twiddle(cpx<T>* a, int k, int step, const cpx<T>* __restrict__ tw) {
  if constexpr (R == 8) {
    const int m = k * step;
    const cpx<T> w1 = tw_load<DIR>(tw, m), w2 = tw_load<DIR>(tw, 2 * m), w4 = tw_load<DIR>(tw, 4 * m);
    const cpx<T> w3 = cmul(w2, w1);
    a[1] = cmul(a[1], w1);
    a[2] = cmul(a[2], w2);
    a[3] = cmul(a[3], w3);
    a[4] = cmul(a[4], w4);
    a[5] = cmul(a[5], cmul(w4, w1));
    a[6] = cmul(a[6], cmul(w4, w2));
    a[7] = cmul(a[7], cmul(w4, w3));
  } else {
#pragma unroll
    for (int r = 1; r < R; ++r) a[r] = cmul(a[r], tw_load<DIR>(tw, r * k * step));
  }
}

constexpr int num_stages(int n) {
  int s = 0, ls = 1;
  while (ls < n) {
    const int rem = n / ls;
    ls *= (rem % 8 == 0) ? 8 : rem;
    ++s;
  }
  return s;
}

// Per-stage twiddle layout (TS).  The N-entry table read at m = k N/(LS R),
// 2m, 4m puts the butterflies k of consecutive lanes on a few LDS banks (up
// to eight lanes on one bank).  Stage (R, LS > 1) instead owns the entries
// ts_base(LS) + p LS + k, k in [0, LS), with w^{q_p k N/(LS R)} for q = 1, 2, 4
// (R = 8) or q = 1 .. R - 1 (a closing radix 2 or 4): consecutive k sit in
// consecutive entries.  The values are the N-entry table's own entries.
constexpr int ts_npow(int r) { return r == 8 ? 3 : r - 1; }
template <int N>
constexpr int ts_base(int ls_want) {
  int off = 0, ls = 1;
  while (ls < N) {
    const int rem = N / ls;
    const int r = (rem % 8 == 0) ? 8 : rem;
    if (ls == ls_want) return off;
    if (ls > 1) off += ts_npow(r) * ls;
    ls *= r;
  }
  return off;
}
template <int N>
constexpr int ts_count() { return ts_base<N>(N); }

// exponent e of entry i of the per-stage table (its value is tw[e]); the
// stage loop unrolls, so ls is a constant in each test
template <int N>
__device__ __forceinline__ int ts_exponent(int i) {
  constexpr int NS = num_stages(N);
  int e = 0, off = 0, ls = 1;
#pragma unroll
  for (int st = 0; st < NS; ++st) {
    const int rem = N / ls;
    const int r = (rem % 8 == 0) ? 8 : rem;
    if (ls > 1) {
      const int cnt = ts_npow(r) * ls;
      const int d = i - off;
      if (d >= 0 && d < cnt) {
        const int p = d / ls, k = d - p * ls;
        e = (r == 8 ? (1 << p) : p + 1) * k * (N / (ls * r));
      }
      off += cnt;
    }
    ls *= r;
  }
  return e;
}

// the per-stage table from the N-entry table tw, by the NT threads of a
// workgroup (thread tid): every load is issued before the first store
template <typename T, int N, int NT>
__device__ __forceinline__ void ts_fill(cpx<T>* dst, const cpx<T>* __restrict__ tw, int tid) {
  constexpr int CNT = ts_count<N>();
  constexpr int PER = (CNT + NT - 1) / NT;
  cpx<T> v[PER];
#pragma unroll
  for (int q = 0; q < PER; ++q) {
    const int i = tid + q * NT;
    v[q] = tw[ts_exponent<N>(i < CNT ? i : 0)];
  }
#pragma unroll
  for (int q = 0; q < PER; ++q) {
    const int i = tid + q * NT;
    if (i < CNT) dst[i] = v[q];
  }
}

template <typename T, int R, int DIR>
__device__ __forceinline__ void twiddle_ts(cpx<T>* a, int k, int ls, const cpx<T>* __restrict__ ts) {
  if constexpr (R == 8) {
    const cpx<T> w1 = tw_load<DIR>(ts, k), w2 = tw_load<DIR>(ts, ls + k),
                 w4 = tw_load<DIR>(ts, 2 * ls + k);
    const cpx<T> w3 = cmul(w2, w1);
    a[1] = cmul(a[1], w1);
    a[2] = cmul(a[2], w2);
    a[3] = cmul(a[3], w3);
    a[4] = cmul(a[4], w4);
    a[5] = cmul(a[5], cmul(w4, w1));
    a[6] = cmul(a[6], cmul(w4, w2));
    a[7] = cmul(a[7], cmul(w4, w3));
  } else {
#pragma unroll
    for (int r = 1; r < R; ++r) a[r] = cmul(a[r], tw_load<DIR>(ts, (r - 1) * ls + k));
  }
}

// LDS slot of pencil element e: PD = 0 stores e at e; PD = 1 leaves one
// empty slot after every 8 (slot e + e/8, a pencil spans N + N/8 slots).  With
// the pencil strides of Cfg this spreads the stride-8 writes of the first
// stage and the unit-stride accesses of all stages over distinct banks (32
// banks of 4 bytes; 16-byte accesses are served 8 lanes per cycle, 8-byte
// accesses 16).  Offsets that are multiples of 8 map linearly, so every
// access is a per-thread base plus a compile-time offset.
//
// PD = 2 leaves one empty slot after every 16 (slot e + e/16), for 8-byte
// entries in the x passes, where consecutive lanes hold consecutive elements:
// an 8-byte access serves 16 lanes per cycle, and with PD = 1 sixteen
// consecutive elements span 17 slots, so every unit-stride stage read (and
// the writes of the stage with LS = 64) puts two lanes on one bank pair.
// With PD = 2 those are conflict-free and so are the first stage's stride-8
// writes; the writes of the stage with LS = 8 keep a two-way conflict on a
// quarter of their lanes.  Offsets that are multiples of 16 map linearly; the
// LS = 8 stage's offsets 8 r (added to a base below 8 mod 16) map to
// 8 r + r / 2 (wr_off).
template <int PD>
__device__ __forceinline__ int slot(int e) {
  return PD == 2 ? e + (e >> 4) : (PD ? e + (e >> 3) : e);
}
template <int PD>
constexpr int slot_step(int c) {
  return PD == 2 ? c + c / 16 : (PD ? c + c / 8 : c);
}
// slot offset of output r of a stage with LS > 1, from slot((j - k) R + k)
template <int PD, int LS>
constexpr int wr_off(int r) {
  return PD == 2 && LS == 8 ? 8 * r + r / 2 : r * slot_step<PD>(LS);
}

// Stage (R, LS) of the ping-pong schedule: exchange k (k = 0, 1, ...) writes
// buffer k mod 2 and the next stage reads it, so one barrier per exchange
// suffices -- a buffer is rewritten only two exchanges after it was read.
// ONE (p0 == p1, one buffer where two do not fit in LDS): a block barrier also
// separates each stage's reads from its writes.
// TS: tw is the per-stage table (ts_count<N>() entries) instead of the N-entry one.
template <typename T, int N, int E, int LS, int DIR, int SYNC, int W, int PD, bool ONE, bool TS>
__device__ __forceinline__ void stages(cpx<T> (&v)[E], cpx<T>* p0, cpx<T>* p1, int t,
                                       const cpx<T>* __restrict__ tw) {
  constexpr int R = StageR<N, LS>::R;
  static_assert(E % R == 0, "elements per thread must be a multiple of each radix");
  constexpr int NB = E / R;
  constexpr int NR = N / R;
  constexpr int TP = N / E;
  constexpr bool FIRST = (LS == 1);
  constexpr bool LAST = (LS * R == N);
  static_assert(!PD || ((FIRST || NR % 8 == 0) && (LAST || LS % 8 == 0 || (LS == 1 && R == 8))),
                "slot offsets must be multiples of 8");
  static_assert(PD != 2 || ((FIRST || NR % 16 == 0) && (LAST || LS == 8 || LS % 16 == 0 ||
                                                       (LS == 1 && R == 8))),
                "PD = 2: read offsets multiples of 16, write offsets as wr_off");
  cpx<T>* const rd = W == 0 ? p1 : p0;   // written by the previous stage
  cpx<T>* const pen = W == 0 ? p0 : p1;  // written by this stage
  cpx<T> a[NB][R];
  if constexpr (FIRST) {
#pragma unroll
    for (int b = 0; b < NB; ++b)
#pragma unroll
      for (int r = 0; r < R; ++r) a[b][r] = v[b + r * NB];
  } else {
#pragma unroll
    for (int b = 0; b < NB; ++b) {
      const int j = t + TP * b;
      const cpx<T>* q = rd + slot<PD>(j);
#pragma unroll
      for (int r = 0; r < R; ++r) a[b][r] = q[r * slot_step<PD>(NR)];
    }
    if constexpr (ONE && !SYNC && !LAST) __syncthreads();
  }
#pragma unroll
  for (int b = 0; b < NB; ++b) {
    if constexpr (!FIRST) {
      const int j = t + TP * b;
      constexpr int TB = ts_base<N>(LS);   // a constant, not a loop in the kernel
      if constexpr (TS)
        twiddle_ts<T, R, DIR>(a[b], j % LS, LS, tw + TB);
      else
        twiddle<T, R, DIR>(a[b], j % LS, N / (LS * R), tw);
    }
    Dft<R, DIR, T>::run(a[b], tw, N);
  }
  if constexpr (LAST) {
#pragma unroll
    for (int b = 0; b < NB; ++b)
#pragma unroll
      for (int r = 0; r < R; ++r) v[b + r * NB] = a[b][r];
  } else {
#pragma unroll
    for (int b = 0; b < NB; ++b) {
      const int j = t + TP * b;
      const int k = j % LS;
      // LS = 1: slot(8 j + r) = slot(8 j) + r; LS >= 8: slot(o + r LS) = slot(o) + wr_off(r)
      cpx<T>* o = pen + slot<PD>((j - k) * R + k);
#pragma unroll
      for (int r = 0; r < R; ++r) o[LS == 1 ? r : wr_off<PD, LS>(r)] = a[b][r];
    }
    lds_sync<SYNC>();
    stages<T, N, E, LS * R, DIR, SYNC, 1 - W, PD, ONE, TS>(v, p0, p1, t, tw);
  }
}

// In-register FFT of one pencil: v[m] = x[t + TP m] in and out.  p0, p1 are
// this pencil's two LDS regions (slot<PD> layout, at least N elements; the
// same region with ONE).  On return p0 may be written at once (its last read
// was followed by a barrier); p1 may not.  TS: tw is the per-stage table.
template <typename T, int N, int E, int DIR, int SYNC, int PD, bool ONE = false, bool TS = false>
__device__ __forceinline__ void fft(cpx<T> (&v)[E], cpx<T>* p0, cpx<T>* p1, int t,
                                    const cpx<T>* __restrict__ tw) {
  stages<T, N, E, 1, DIR, SYNC, 0, PD, ONE, TS>(v, p0, p1, t, tw);
  // with an odd number of exchanges the last one read p0; with one buffer
  // the last exchange always read it
  if constexpr (ONE || (num_stages(N) - 1) % 2 == 1) lds_sync<SYNC>();
}


// Workgroups are dealt round-robin to the 8 XCDs, each with its own L2.
// Giving each XCD a contiguous run of logical tiles makes neighbouring tiles,
// which share cache lines at their edges, meet in the same L2.  The grid is
// launched with xcd_grid(nb) >= nb workgroups; logical ids >= nb are idle.
constexpr int kXcd = 8;
inline unsigned xcd_grid(long long nb) { return (unsigned)(((nb + kXcd - 1) / kXcd) * kXcd); }
__device__ __forceinline__ long long xcd_tile() {
  const long long per = gridDim.x / kXcd;
  return (long long)(blockIdx.x % kXcd) * per + blockIdx.x / kXcd;
}

// per-length launch shape: E elements per thread, TP threads per pencil,
// B pencils per workgroup of the strided passes
template <typename T, int N>
struct Cfg {
  // strided passes: E elements per thread, TP threads per pencil
  static constexpr int E = 8;
  static constexpr int TP = N / E;
  // workgroup threads: at most 64 pencils, so at short lengths a smaller
  // workgroup means more, smaller tiles (still >= 128-byte segments); the
  // persistent FP64 z passes at N >= 128 expect one 512-thread workgroup per CU
  static constexpr int WG = sizeof(T) == 8 ? (N <= 64 ? 256 : 512)
                                           : (N <= 128 ? 256 : (N <= 256 ? 512 : 1024));
  static constexpr int B0 = WG / TP;
  static constexpr int B = B0 < 1 ? 1 : (B0 > 64 ? 64 : B0);
  static constexpr int NT = B * TP;
  // strided passes: consecutive lanes hold different pencils of a tile; an
  // odd pencil stride puts them on distinct banks
  static constexpr int PD = N >= 1024 ? 1 : 0;
  static constexpr int LD = PD ? N + N / 8 + 2 : N + 1;
  // one exchange buffer (and a barrier between each stage's reads and
  // writes) where two do not fit in the 160 KiB of LDS
  static constexpr bool ONE = 2 * B * LD * (int)sizeof(cpx<T>) > 152 * 1024;
  static constexpr int NBUF = ONE ? 1 : 2;
  // x pass: XE elements per thread (16 for FP32 at N = 1024, so a row sits
  // in one wave), XTP threads per row, U row-pair units per workgroup
  static constexpr int XE = (sizeof(T) == 4 && N >= 1024) ? 16 : 8;
  static constexpr int XTP = N / XE;
  static constexpr int U = XTP >= 256 ? 1 : 256 / XTP;
  static constexpr int XNT = U * XTP;
  static constexpr int XSYNC = XTP <= 64 ? 1 : 0;
  // x pass: consecutive lanes hold consecutive elements of one pencil (slot
  // padding per 16 entries for FP32 where every stage read spans 16 entries)
  static constexpr int XPD = (sizeof(T) == 4 && N >= 128) ? 2 : 1;
  static constexpr int XLD = XPD == 2 ? N + N / 16 : N + N / 8;
};


// Row-pair helpers of the x passes, called with N = Nx.  G provides nhk (kx
// indices actually held, <= Nx/2 + 1).  The stored count Nx/2 + 1 and the
// Nyquist index Nx/2 follow from N at compile time, so with the element
// index k = t + XTP m most of the completion's cases (mirrored or not, self-
// conjugate or not) are fixed per m and cost no instruction.
template <int N>
__device__ __forceinline__ constexpr int half_count() { return N / 2 + 1; }
template <int N>
__device__ __forceinline__ constexpr int nyquist_x() { return N % 2 == 0 ? N / 2 : -1; }
//
// c2r_pair_load: Z_k = A_k + i B_k over the Hermitian completion of the two
// half-spectrum rows ra, rb (entries kx >= nhk are zero), with the
// self-conjugate entries taken real as a c2r transform treats them.  Split in
// c2r_pair_fetch (the raw entries a[m], b[m] this thread needs, issued as
// loads) and c2r_pair_combine (the completion), so a caller can issue the
// loads of one row pair before transforming another.
template <typename T, int N, int E, class G>
__device__ __forceinline__ void c2r_pair_fetch(cpx<T> (&a)[E], cpx<T> (&b)[E],
                                               const cpx<T>* __restrict__ ra,
                                               const cpx<T>* __restrict__ rb, bool oka, bool okb,
                                               int t, const G& g) {
  using C = Cfg<T, N>;
#pragma unroll
  for (int m = 0; m < E; ++m) {
    const int k = t + C::XTP * m;
    const int kk = k >= half_count<N>() ? N - k : k;
    a[m] = mk<T>(T(0), T(0));
    b[m] = mk<T>(T(0), T(0));
    if (kk < g.nhk) {
      if (oka) a[m] = ld_s(ra + kk);
      if (okb) b[m] = ld_s(rb + kk);
    }
  }
}

template <typename T, int N, int E, class G>
__device__ __forceinline__ void c2r_pair_combine(cpx<T> (&v)[E], const cpx<T> (&a)[E],
                                                 const cpx<T> (&b)[E], int t, const G& g) {
  using C = Cfg<T, N>;
#pragma unroll
  for (int m = 0; m < E; ++m) {
    const int k = t + C::XTP * m;
    const bool up = k >= half_count<N>();
    const int kk = up ? N - k : k;
    cpx<T> x = a[m], y = b[m];
    if (kk == 0 || kk == nyquist_x<N>()) {
      x.y = T(0);
      y.y = T(0);
    }
    if (up) {
      x.y = -x.y;
      y.y = -y.y;
    }
    v[m] = mk<T>(x.x - y.y, x.y + y.x);
  }
}

template <typename T, int N, int E, class G>
__device__ __forceinline__ void c2r_pair_load(cpx<T> (&v)[E], const cpx<T>* __restrict__ ra,
                                              const cpx<T>* __restrict__ rb, bool oka, bool okb,
                                              int t, const G& g) {
  cpx<T> a[E], b[E];
  c2r_pair_fetch<T, N, E>(a, b, ra, rb, oka, okb, t, g);
  c2r_pair_combine<T, N, E>(v, a, b, t, g);
}

// forward-transformed Z = a + i b -> A_k, B_k for k < nhk, stored to ra / rb
// (NTS: with the nontemporal hint; the fused x passes take it, the standalone
// forward transform measured faster without)
template <typename T, int N, int E, class G, bool NTS = (GBG_NT & 2) != 0>
__device__ __forceinline__ void r2c_pair_store(cpx<T> (&v)[E], cpx<T>* pen,
                                               cpx<T>* __restrict__ ra, cpx<T>* __restrict__ rb,
                                               bool oka, bool okb, int t, const G& g) {
  using C = Cfg<T, N>;
#pragma unroll
  for (int m = 0; m < E; ++m) pen[t + C::XTP * m] = v[m];
  lds_sync<C::XSYNC>();
  for (int k = t; k < g.nhk; k += C::XTP) {
    const cpx<T> z = pen[k];
    const cpx<T> zc = pen[k == 0 ? 0 : N - k];
    if (oka) st_s<NTS>(ra + k, mk<T>(T(0.5) * (z.x + zc.x), T(0.5) * (z.y - zc.y)));
    if (okb) st_s<NTS>(rb + k, mk<T>(T(0.5) * (z.y + zc.y), T(0.5) * (zc.x - z.x)));
  }
  lds_sync<C::XSYNC>();
}

}  // namespace fast
}  // namespace gbg
