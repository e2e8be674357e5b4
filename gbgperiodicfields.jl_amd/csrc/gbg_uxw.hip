// Pipelines: the dealiased rotational nonlinear term and the dealiased
// scalar product.
//
// Composed path (GBG_UXW_UNFUSED, and grids the fused passes do not cover):
// the standalone operators in the order of examples/HIST_CPU.jl:229-247,
//   u^t = M u^;  w^ = curl u^t;  E, Z;  u = F^-1 u^t;  w = F^-1 w^;
//   n = u x w;  N^ = F n;  N^ = P N^;  N^ = M N^;  N^(0) = 0.
//
// Fused path: five pencil passes that never touch a mode the 2/3 mask M
// removes.  With kc the cutoff, only kx <= kx_max, |ky| <= ky_max and the
// (kx, ky) disk kx^2 + ky^2 <= kc^2 can carry data before the z transforms,
// so the intermediates are stored compactly over those index ranges:
//   K1  z-inverse:   u^ (ball of M) -> 5 fields (u and the kz terms of
//                    w = curl u^) on the disk pencils; energy and enstrophy
//                    partial sums in FP64
//   K2  y-inverse:   5 fields -> 6 (u, w: the kx and ky terms of w formed in
//                    registers), (z, kx <= kx_max) pencils, loads on the disk only
//   K3  x-pass:      6 C2R rows -> u x w in registers -> 3 R2C rows (kx <= kx_max)
//   K4  y-forward:   3 fields, stores the disk ky range only
//   K5  z-forward:   3 fields -> Leray projection, mask, mean-free -> N^ (full
//                    layout, zeros outside the mask), on the output tiles that
//                    hold disk pencils; a write-only kernel zeros the other tiles
// Buffers: W1 = [6][Nz][nyc][nhp] (K1 -> K2, five used), reused as [3][Nz][nyc][nhp]
// (K4 -> K5); W2 = [6][Nz][Ny][nhp] (K2 -> K3), reused as [3][Nz][Ny][nhp]
// (K3 -> K4).  nhk = kx_max + 1 entries of each row are used, nhp = nhk
// rounded up to 16; nyc = number of kept ky indices.
// The z passes (K1, K5's transform kernel) have one 131-160 KB-LDS workgroup
// per CU; where their tile loop fits the register budget without spilling
// they run persistent (one workgroup per CU walking its XCD's tiles), so the
// CU does not sit empty between one tile's workgroup and the next.
#include <type_traits>
#include <cstdlib>
#include <vector>

#include "fft_fast.hpp"

namespace gbg {

// The composed path (the standalone transforms and operators in sequence) is
// defined after the fused passes, whose combination kernel it shares.
static gbg_status uxw_composed(gbg_plan* p, const void* u_hat, void* n_hat, double kc, char* w,
                               hipStream_t s, const UxwComb* cb = nullptr);

// ============================================================================
// fused path
// ============================================================================
struct Geo {
  int nx, ny, nz, nh;
  int nhk;          // kept kx indices [0, nhk)
  int nhp;          // row pitch of the compact buffers (nhk rounded up to 16)
  int nyc;          // kept ky indices, compact
  int ya;           // kyc < ya -> ky = kyc; else ky = ny - nyc + kyc
  int nyq_x;        // x Nyquist index (nh - 1) for even nx, else -1
  long long M;      // full spectral component
  long long s1;     // compact-ky component stride  nz * nyc * nhp
  long long s2;     // full-y component stride      nz * ny * nhp
  double hi2;       // kc^2 (FP64)
};

using fast::Cfg;

// Tile shapes of the x pass (K3, the product's P3), chosen per plan
// (gbg_plan::x_var, default x_default_shape): V = 0 the row-pair units of Cfg
// per workgroup; V = 1 half as many, so more, smaller workgroups share a CU.
// Measured per pass on the MI355X (u x omega, 2/3 cutoff): at Nx = 1024 in
// FP64 (two waves per row, no staging) V = 1 is 5 % faster (18.1 -> 17.1 ms);
// in FP32 at Nx = 1024, staged, V = 0 is 6 % faster (8.85 -> 8.34 ms); at
// Nx = 512 V = 1 is 9 % slower in FP64 and within run-to-run spread in FP32.
// (For the z-inverse pass half the pencils per workgroup or sixteen elements
// per thread measured slower or spilled, and are not built.)
template <typename T, int N, int V>
struct XCfg {
  using D = Cfg<T, N>;
  static constexpr int XE = D::XE, XTP = D::XTP, XSYNC = D::XSYNC, XPD = D::XPD, XLD = D::XLD;
  static constexpr int U = V == 1 ? (D::U > 1 ? D::U / 2 : 1) : D::U;
  static constexpr int XNT = U * XTP;
};

// With OPAQUE, the same pointer made opaque to the optimiser: table loads
// behind it stay where they are written instead of being hoisted (and held in
// registers) across the whole kernel.
template <bool OPAQUE, typename P>
__device__ __forceinline__ P* fresh(P* p) {
  if (OPAQUE) asm volatile("" : "+s"(p));
  return p;
}

// x where keep holds, else +0, as a bitwise AND with a mask the optimiser
// cannot trace back to keep.  A select on keep would let it move the load of
// x under a branch on keep; the load then completes where the branch rejoins
// instead of staying in flight behind later work.
template <typename T>
__device__ __forceinline__ cpx<T> keep_or_zero(cpx<T> x, bool keep) {
  using U = typename std::conditional<sizeof(T) == 8, unsigned long long, unsigned>::type;
  U msk = keep ? ~U(0) : U(0);
  asm("" : "+v"(msk));
  return mk<T>(__builtin_bit_cast(T, __builtin_bit_cast(U, x.x) & msk),
               __builtin_bit_cast(T, __builtin_bit_cast(U, x.y) & msk));
}

__device__ __forceinline__ int ky_of(int kyc, const Geo& g) {
  return kyc < g.ya ? kyc : g.ny - g.nyc + kyc;
}
// compact index of a full ky index, or -1 when the mask drops it
__host__ __device__ __forceinline__ int kyc_of(int ky, const Geo& g) {
  if (ky < g.ya) return ky;
  const int c = ky - (g.ny - g.nyc);
  return (c >= g.ya && c < g.nyc) ? c : -1;
}

// ---- K1: z-inverse of u^ and curl u^ on the disk pencils --------------------
// Tiles: B consecutive kx of one compact (kyc) row, pencil index
// p1 = kyc * nhp + kx.  Energy and enstrophy partials land in part[2 b],
// part[2 b + 1] for the launched workgroup b (zeros for idle ones).
// Persistent form of K1 (see k_uxw_z_inverse): where the tile loop holds in
// the register budget without spilling (256 VGPRs in FP64; 128 in FP32, where
// only the 32-bit-offset form at Nz = 512 fits).
template <typename T, int N, bool O32>
constexpr bool z_inverse_persist() {
  return sizeof(T) == 8 ? (N >= 128 && N <= 512) : (O32 && N == 512);
}

// Nontemporal stores in the z-inverse pass only where a tile's piece of a z
// plane (B pencils) fills 128-byte lines: a 64-byte piece (B = 4 in FP64, 8 in
// FP32, at Nz = 1024) shares its line with the next tile's, and with the hint
// the two halves can reach memory as partial lines instead of meeting in L2 (a
// probe of that pattern: 64-byte pieces 3.3-3.8 TB/s written with the hint,
// 128-byte pieces 5.0-6.0 TB/s).  Measured at 1024^3: K1 with plain stores
// 13.36 -> 13.22 ms FP64 (zero tiles beside it), 7.40 -> 7.23 ms FP32; the
// last pass's transform kernel, whose output tiles follow the tile list, was
// slower with plain stores (6.07 -> 6.49 ms FP64), so it keeps the hint.
template <typename T, int B>
constexpr bool z_store_nt() {
  return piece_store_nt(B * (int)sizeof(cpx<T>));
}

// Where the y-inverse pass forms the kx and ky terms of w = curl u^ (the
// z-inverse then writes five fields: u and the two kz terms of w): Ny <= 512,
// where the y passes stream.  At Ny = 1024 the forms of that pass that fit the
// registers (a loop over its six steps) spilled or, written out, measured a
// seventh slower than the six-field y pass (1024^3 FP64 K2 10.8 -> 12.4 ms
// against K1 13.4 -> 12.2 ms), so there the z-inverse writes all six fields.
template <int NY>
__host__ __device__ constexpr bool curl_in_y() {
  return NY <= 512;
}

// O32: every element offset inside one component of uh and w1 fits 32 bits
// (M and s1 below 2^32); the per-entry offsets then take one register each.
template <typename T, int N, bool O32, bool FIVE>
__global__ __launch_bounds__((Cfg<T, N>::NT)) void k_uxw_z_inverse(
    const cpx<T>* __restrict__ uh, cpx<T>* __restrict__ w1, Geo g, Tables tb,
    const cpx<T>* __restrict__ tw, double* __restrict__ part, long long ntiles) {
  using C = Cfg<T, N>;
  using Off = typename std::conditional<O32, unsigned, long long>::type;
  constexpr int E = C::E, TP = C::TP, B = C::B, LD = C::LD;
  constexpr bool NTS = z_store_nt<T, B>();
  __shared__ cpx<T> lds[C::NBUF * B * LD];
  __shared__ double red[2][C::NT / 64];
  const int pl = threadIdx.x % B, t = threadIdx.x / B;
  // The transforms' twiddles and the kz derivative factors are read from LDS
  // copies: a global load consumed inside the loop below would wait for every
  // load issued before it (the counter of outstanding vector-memory loads
  // drains in order), including the prefetched components.  The copies are
  // complete before their first use, which follows the first exchange barrier
  // of the first transform.  Where LDS cannot hold the twiddle table as well
  // (FP64, Nz = 1024), it stays in global memory.
  constexpr bool TWL = (size_t)C::NBUF * B * LD * sizeof(cpx<T>) + 2 * (C::NT / 64) * sizeof(double) +
                           (size_t)N * (sizeof(cpx<T>) + sizeof(T)) <=
                       160 * 1024;
  __shared__ cpx<T> twl[TWL ? N : 1];
  __shared__ T a3l[N];
  // the kz squared family as well where LDS holds it, in FP64 (the keep bits
  // of a tile then need no global read; in FP32 the LDS reads spilled at the
  // 128-VGPR limit)
  constexpr bool Q3L = sizeof(T) == 8 && (size_t)C::NBUF * B * LD * sizeof(cpx<T>) + 2 * (C::NT / 64) * sizeof(double) +
                           (TWL ? N * sizeof(cpx<T>) : 0) + (size_t)N * (sizeof(T) + sizeof(double)) <=
                       160 * 1024;
  __shared__ double q3l[Q3L ? N : 1];
  for (int i = threadIdx.x; i < N; i += C::NT) {
    if constexpr (TWL) twl[i] = tw[i];
    a3l[i] = T(tb.kd3[i]);
    if constexpr (Q3L) q3l[i] = tb.ks3[i];
  }
  auto twp = [&]() {
    if constexpr (TWL)
      return (const cpx<T>*)twl;
    else
      return tw;
  };
  auto q3 = [&](int kz) {
    if constexpr (Q3L)
      return q3l[kz];
    else
      return tb.ks3[kz];
  };
  double eacc = 0.0, zacc = 0.0;
  // FP64: the next tile's table values are read while a tile runs (below);
  // in FP32 the values held across the tile spill at the 128-VGPR limit
  constexpr bool PREF = sizeof(T) == 8 && z_inverse_persist<T, N, O32>();
  // a tile's per-pencil table values (read from global memory; with PREF
  // only, else run_tile reads them itself)
  struct TileGeo {
    int kx, ky;
    double qxy;
    T a1, a2;
    bool disk;
  };
  auto geo_of = [&](long long tile) {
    const long long p1 = tile * B + pl;
    TileGeo q;
    q.kx = 0;
    q.ky = 0;
    q.qxy = 0.0;
    q.disk = false;
    if (tile < ntiles && p1 < (long long)g.nyc * g.nhp) {
      const int kyc = (int)(p1 / g.nhp);
      q.kx = (int)(p1 - (long long)kyc * g.nhp);
      if (q.kx < g.nhk) {
        q.ky = ky_of(kyc, g);
        q.qxy = tb.ks1[q.kx] + tb.ks2[q.ky];
        q.disk = q.qxy <= g.hi2;
      } else {
        q.kx = 0;
      }
    }
    q.a1 = T(tb.kd1[q.kx]);
    q.a2 = T(tb.kd2[q.ky]);
    return q;
  };
  // one tile: B consecutive kx of one compact (kyc) row
  auto run_tile = [&](long long tile, const TileGeo& q) {
    const long long p1 = tile * B + pl;
    int kx = 0, ky = 0;
    double qxy = 0.0;
    bool disk = false;
    if constexpr (PREF) {
      kx = q.kx;
      ky = q.ky;
      qxy = q.qxy;
      disk = q.disk;
    } else if (tile < ntiles && p1 < (long long)g.nyc * g.nhp) {
      const int kyc = (int)(p1 / g.nhp);
      kx = (int)(p1 - (long long)kyc * g.nhp);
      if (kx < g.nhk) {
        ky = ky_of(kyc, g);
        qxy = tb.ks1[kx] + tb.ks2[ky];
        disk = qxy <= g.hi2;
      } else {
        kx = 0;
      }
    }
    if (!__syncthreads_or(disk)) return;
    // the sums accumulate straight into the workgroup's (a per-tile copy cost
    // the FP32 K1 at Nz = 512 four registers it spilled inside the transforms)
    double& ea = eacc;
    double& za = zacc;
    const T a1 = PREF ? q.a1 : T(tb.kd1[kx]), a2 = PREF ? q.a2 : T(tb.kd2[ky]);
    const double wx = (kx == 0 || kx == g.nyq_x) ? 1.0 : 2.0;
    const long long plane = (long long)g.ny * g.nh;
    const long long off = (long long)ky * g.nh + kx;
    bool keep[E];
#pragma unroll
    for (int m = 0; m < E; ++m) keep[m] = disk & (qxy + q3(t + TP * m) <= g.hi2);  // no branch
    cpx<T>* pen = lds + pl * LD;
    cpx<T>* pen1 = C::ONE ? pen : lds + (B + pl) * LD;
    const long long zs = (long long)g.nyc * g.nhp;
    // FIVE (curl_in_y): five output fields from four component loads, each
    // issued at least one transform before it is needed, with two raw
    // components X, Y held and no register copies between them.  The x and y
    // parts of w = curl u^ are constant along a z pencil, so the z-inverse
    // needs only u and the two kz terms: fields 0..4 = u1, u2, u3,
    // d1 = F(i a3 u1), d2 = F(i a3 u2), and the y pass forms w1 = i a2 U3 - d2,
    // w2 = d1 - i a1 U3, w3 = i (a1 U2 - a2 U1).  The enstrophy partials take
    // w^ itself (a pair of raw components per term):
    //   u1 | w2 = i(a3 X - a1 Y) | d1, X <- u2 | u3 | w1 = i(a2 Y - a3 X) | d2, Y <- u1 | u2 | w3 = i(a1 X - a2 Y)
    // starting from X = u1, Y = u3 (the last load only serves the enstrophy).
    // Otherwise six fields 0..5 = u1, u2, u3, w1, w2, w3 from the same loads:
    //   u1 | w2 = i(a3 X - a1 Y), X <- u2 | u3 | w1 = i(a2 Y - a3 X), Y <- u1 | u2 | w3 = i(a1 X - a2 Y)
    // Entries outside the ball load the pencil's kz = 0 entry instead (a valid
    // address; on the disk a line the kept lanes fetch anyway) and are zeroed
    // where they are used.  No branch surrounds a load or a store, so the
    // compiler's wait for one load never has to cover another issued after it
    // and each load stays in flight across the transform that follows it.
    // Every lane of a tile that holds a disk pencil stores (zeros off the disk):
    // with B dividing the row pitch nhp, all of them address the buffer.
    static_assert(16 % B == 0 || B > 16, "tile width");
    constexpr bool ALL_STORE = 16 % B == 0;
    const bool st_ok = ALL_STORE || p1 < (long long)g.nyc * g.nhp;
    cpx<T> X[E], Y[E];
    // The offsets pass through an opaque (volatile) copy at each use: the
    // optimiser neither splits a load by cases nor keeps eight 64-bit
    // addresses live across the tile, which spilled in FP32.
    auto loadc = [&](cpx<T> (&r)[E], int c) {
      const cpx<T>* u = uh + c * g.M;
#pragma unroll
      for (int m = 0; m < E; ++m) {
        Off zi = (Off)off + (keep[m] ? (Off)(t + TP * m) * (Off)plane : (Off)0);
        asm volatile("" : "+v"(zi));
        r[m] = ld_s(u + zi);
      }
    };
    auto emit = [&](cpx<T> (&v)[E], int field) {
      fast::fft<T, N, E, 1, 0, C::PD, C::ONE>(v, pen, pen1, t, twp());
      if (st_ok) {
        cpx<T>* o = w1 + field * g.s1;
#pragma unroll
        for (int m = 0; m < E; ++m) {
          Off zo = (Off)p1 + (Off)(t + TP * m) * (Off)zs;
          asm volatile("" : "+v"(zo));
          st_s<NTS>(o + zo, v[m]);
        }
      }
      asm volatile("" ::: "memory");
      __builtin_amdgcn_sched_barrier(0);
    };
    auto velocity = [&](const cpx<T> (&A)[E], int field) {
      cpx<T> v[E];
#pragma unroll
      for (int m = 0; m < E; ++m) {
        v[m] = keep_or_zero(A[m], keep[m]);
        ea += wx * ((double)v[m].x * v[m].x + (double)v[m].y * v[m].y);
      }
      emit(v, field);
    };
    // enstrophy of w = i(ca A - cb B), ia / ib = 1, 2, 3 naming which of a1,
    // a2, a3 ca and cb are
    auto enstrophy = [&](const cpx<T> (&A)[E], const cpx<T> (&Bv)[E], int ia, int ib) {
#pragma unroll
      for (int m = 0; m < E; ++m) {
        const T a3 = a3l[t + TP * m];
        const T ca = ia == 1 ? a1 : (ia == 2 ? a2 : a3);
        const T cb = ib == 1 ? a1 : (ib == 2 ? a2 : a3);
        const cpx<T> w = keep_or_zero(cmuli(csub(cscale(A[m], ca), cscale(Bv[m], cb))), keep[m]);
        za += wx * ((double)w.x * w.x + (double)w.y * w.y);
      }
    };
    // d = i a3 A into field, with the enstrophy of w = i(cp P - cq Q) taken in
    // the same pass over the elements; then component nc (>= 0) loads into L
    auto dz = [&](const cpx<T> (&A)[E], int field, cpx<T> (&L)[E], int nc, const cpx<T> (&P)[E],
                  const cpx<T> (&Q)[E], int ip, int iq) {
      cpx<T> v[E];
#pragma unroll
      for (int m = 0; m < E; ++m) {
        const T a3 = a3l[t + TP * m];
        const T cp = ip == 1 ? a1 : (ip == 2 ? a2 : a3);
        const T cq = iq == 1 ? a1 : (iq == 2 ? a2 : a3);
        const cpx<T> w = keep_or_zero(cmuli(csub(cscale(P[m], cp), cscale(Q[m], cq))), keep[m]);
        za += wx * ((double)w.x * w.x + (double)w.y * w.y);
        v[m] = keep_or_zero(cmuli(cscale(A[m], a3)), keep[m]);
      }
      if (nc >= 0) loadc(L, nc);
      emit(v, field);
    };
    // (six fields) w = i(ca A - cb B) into field, with its enstrophy; then the
    // next component (nc >= 0) loads into A
    auto vorticity = [&](cpx<T> (&A)[E], const cpx<T> (&Bv)[E], int ia, int ib, int field, int nc) {
      cpx<T> v[E];
#pragma unroll
      for (int m = 0; m < E; ++m) {
        const T a3 = a3l[t + TP * m];
        const T ca = ia == 1 ? a1 : (ia == 2 ? a2 : a3);
        const T cb = ib == 1 ? a1 : (ib == 2 ? a2 : a3);
        v[m] = keep_or_zero(cmuli(csub(cscale(A[m], ca), cscale(Bv[m], cb))), keep[m]);
        za += wx * ((double)v[m].x * v[m].x + (double)v[m].y * v[m].y);
      }
      if (nc >= 0) loadc(A, nc);
      emit(v, field);
    };
    // The transforms written out one after another hold more registers
    // than a loop over them; where that spills (short pencils, where more
    // pencils share a workgroup), a loop over the fields loads the components
    // of field f + 1 (one or two, eight in all) while field f transforms.
    constexpr bool UNROLLED = sizeof(T) == 8 ? N >= 256 : N >= 512;
    if constexpr (UNROLLED && FIVE) {
      loadc(X, 0);
      loadc(Y, 2);
      velocity(X, 0);
      dz(X, 3, X, 1, X, Y, 3, 1);
      velocity(Y, 2);
      dz(X, 4, Y, 0, Y, X, 2, 3);
      velocity(X, 1);
      enstrophy(X, Y, 1, 2);
    } else if constexpr (UNROLLED) {
      loadc(X, 0);
      loadc(Y, 2);
      velocity(X, 0);
      vorticity(X, Y, 3, 1, 4, 1);
      velocity(Y, 2);
      vorticity(Y, X, 2, 3, 3, 0);
      velocity(X, 1);
      vorticity(X, Y, 1, 2, 5, -1);
    } else if constexpr (FIVE) {
      // field f: f = 0, 1, 2 -> u_f (X);  f = 3, 4 -> d = i a3 X with X = u1, u2;
      // Y pairs with X for the enstrophy terms w1 (f = 1), w2 (f = 3), w3 (f = 4)
      auto load = [&](int f) {
        loadc(X, f < 3 ? f : f - 3);
        if (f == 1 || f >= 3) loadc(Y, f == 4 ? 0 : 2);
      };
      load(0);
#pragma unroll 1
      for (int f = 0; f < 5; ++f) {
        if (f == 1) enstrophy(Y, X, 2, 3);
        if (f == 3) enstrophy(X, Y, 3, 1);
        if (f == 4) enstrophy(X, Y, 1, 2);
        cpx<T> v[E];
#pragma unroll
        for (int m = 0; m < E; ++m) {
          if (f < 3) {
            v[m] = keep_or_zero(X[m], keep[m]);
            ea += wx * ((double)v[m].x * v[m].x + (double)v[m].y * v[m].y);
          } else {
            v[m] = keep_or_zero(cmuli(cscale(X[m], a3l[t + TP * m])), keep[m]);
          }
        }
        if (f + 1 < 5) load(f + 1);
        emit(v, f);
      }
    } else {
      // field f: f = 0, 1, 2 -> u_f;  f = 3, 4, 5 -> w = i (ca u_ia - cb u_ib)
      auto load = [&](int f) {
        loadc(X, f < 3 ? f : (f == 3 ? 2 : (f == 4 ? 0 : 1)));
        if (f >= 3) loadc(Y, f == 3 ? 1 : (f == 4 ? 2 : 0));
      };
      load(0);
#pragma unroll 1
      for (int f = 0; f < 6; ++f) {
        cpx<T> v[E];
#pragma unroll
        for (int m = 0; m < E; ++m) {
          if (f < 3) {
            v[m] = keep_or_zero(X[m], keep[m]);
            ea += wx * ((double)v[m].x * v[m].x + (double)v[m].y * v[m].y);
          } else {
            const T a3 = a3l[t + TP * m];
            const T ca = f == 3 ? a2 : (f == 4 ? a3 : a1);
            const T cb = f == 3 ? a3 : (f == 4 ? a1 : a2);
            v[m] = keep_or_zero(cmuli(csub(cscale(X[m], ca), cscale(Y[m], cb))), keep[m]);
            za += wx * ((double)v[m].x * v[m].x + (double)v[m].y * v[m].y);
          }
        }
        if (f + 1 < 6) load(f + 1);
        emit(v, f);
      }
    }
  };
  constexpr bool PERSIST = z_inverse_persist<T, N, O32>();
  if constexpr (PERSIST) {
    const long long per = (ntiles + fast::kXcd - 1) / fast::kXcd;
    const int xcd = blockIdx.x % fast::kXcd;
    const long long L = gridDim.x / fast::kXcd;
    const long long tend = (xcd + 1) * per < ntiles ? (xcd + 1) * per : ntiles;
    long long tile = xcd * per + blockIdx.x / fast::kXcd;
    if constexpr (PREF) {
      // the next tile's table values are read while this tile runs: read at
      // the next tile's start, their wait would also cover this tile's last
      // stores (loads and stores retire through one in-order counter)
      TileGeo cur = geo_of(tile < tend ? tile : ntiles);
#pragma unroll 1
      for (; tile < tend; tile += L) {
        const TileGeo nxt = geo_of(tile + L < tend ? tile + L : ntiles);
        run_tile(tile, cur);
        cur = nxt;
      }
    } else {
#pragma unroll 1
      for (; tile < tend; tile += L) run_tile(tile, TileGeo{});
    }
  } else {
    run_tile(fast::xcd_tile(), TileGeo{});
  }
  // FP64 block reduction of the energy / enstrophy partial sums, fixed order
  for (int s = 32; s > 0; s >>= 1) {
    eacc += __shfl_down(eacc, s, 64);
    zacc += __shfl_down(zacc, s, 64);
  }
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) {
    red[0][wid] = eacc;
    red[1][wid] = zacc;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double e = 0.0, zz = 0.0;
    for (int i = 0; i < C::NT / 64; ++i) {
      e += red[0][i];
      zz += red[1][i];
    }
    part[2 * blockIdx.x] = e;
    part[2 * blockIdx.x + 1] = zz;
  }
}

__global__ __launch_bounds__(256) void k_uxw_reduce(const double* __restrict__ part, int nblocks,
                                                    double norm, double* __restrict__ out) {
  __shared__ double red[2][256];
  double e = 0.0, z = 0.0;
  for (int i = threadIdx.x; i < nblocks; i += 256) {
    e += part[2 * i];
    z += part[2 * i + 1];
  }
  red[0][threadIdx.x] = e;
  red[1][threadIdx.x] = z;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) {
      red[0][threadIdx.x] += red[0][threadIdx.x + s];
      red[1][threadIdx.x] += red[1][threadIdx.x + s];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    out[0] = red[0][0] / norm;
    out[1] = red[1][0] / norm;
  }
}

// ---- P1: z-inverse of two scalar spectra (the dealiased product) -------------
// The tiles of K1; fields 0, 1 of the compact buffer receive a^ and b^ on the
// disk pencils, with the modes outside the 2/3 ball zeroed on load.  b's
// loads are in flight while a transforms; as in K1 no load or store sits
// under a branch and the twiddles are read from an LDS copy where it fits.
template <typename T, int N>
__global__ __launch_bounds__((Cfg<T, N>::NT)) void k_dp_z_inverse(
    const cpx<T>* __restrict__ ah, const cpx<T>* __restrict__ bh, cpx<T>* __restrict__ w1, Geo g,
    Tables tb, const cpx<T>* __restrict__ tw, long long ntiles) {
  using C = Cfg<T, N>;
  constexpr int E = C::E, TP = C::TP, B = C::B, LD = C::LD;
  __shared__ cpx<T> lds[C::NBUF * B * LD];
  const int pl = threadIdx.x % B, t = threadIdx.x / B;
  const long long tile = fast::xcd_tile();
  const long long p1 = tile * B + pl;
  int kx = 0, ky = 0;
  double qxy = 0.0;
  bool disk = false;
  if (tile < ntiles && p1 < (long long)g.nyc * g.nhp) {
    const int kyc = (int)(p1 / g.nhp);
    kx = (int)(p1 - (long long)kyc * g.nhp);
    if (kx < g.nhk) {
      ky = ky_of(kyc, g);
      qxy = tb.ks1[kx] + tb.ks2[ky];
      disk = qxy <= g.hi2;
    } else {
      kx = 0;
    }
  }
  if (!__syncthreads_or(disk)) return;
  constexpr bool TWL =
      (size_t)C::NBUF * B * LD * sizeof(cpx<T>) + (size_t)N * sizeof(cpx<T>) <= 160 * 1024;
  __shared__ cpx<T> twl[TWL ? N : 1];
  if constexpr (TWL)
    for (int i = threadIdx.x; i < N; i += C::NT) twl[i] = tw[i];
  auto twp = [&]() {
    if constexpr (TWL)
      return (const cpx<T>*)twl;
    else
      return tw;
  };
  const long long plane = (long long)g.ny * g.nh;
  const long long off = (long long)ky * g.nh + kx;
  const long long zs = (long long)g.nyc * g.nhp;
  bool keep[E];
#pragma unroll
  for (int m = 0; m < E; ++m) keep[m] = disk & (qxy + tb.ks3[t + TP * m] <= g.hi2);  // no branch
  cpx<T>* pen = lds + pl * LD;
  cpx<T>* pen1 = C::ONE ? pen : lds + (B + pl) * LD;
  constexpr bool ALL_STORE = 16 % B == 0;   // every lane of a launched tile addresses w1
  const bool st_ok = ALL_STORE || p1 < (long long)g.nyc * g.nhp;
  cpx<T> A[E], Bv[E];
#pragma unroll
  for (int m = 0; m < E; ++m) {
    long long zi = keep[m] ? (long long)(t + TP * m) * plane : 0;
    asm("" : "+v"(zi));  // keeps the optimiser from splitting the loads by cases
    A[m] = ah[off + zi];
    Bv[m] = bh[off + zi];
  }
  auto emit = [&](cpx<T> (&v)[E], int field) {
#pragma unroll
    for (int m = 0; m < E; ++m) v[m] = keep_or_zero(v[m], keep[m]);
    fast::fft<T, N, E, 1, 0, C::PD, C::ONE>(v, pen, pen1, t, twp());
    if (st_ok) {
      cpx<T>* o = w1 + field * g.s1 + p1;
#pragma unroll
      for (int m = 0; m < E; ++m) o[(long long)(t + TP * m) * zs] = v[m];
    }
  };
  emit(A, 0);
  asm volatile("" ::: "memory");
  __builtin_amdgcn_sched_barrier(0);
  emit(Bv, 1);
}

// ---- K2 / K4: y passes over (z, kx < nhk) pencils ------------------------------
// DIR = +1 (inverse; NF = 6 for u x omega, 2 for the scalar product): loads
//   the disk ky from the compact buffer, stores all y to the full-y buffer;
//   for u x omega the five z-pass fields give the six (w formed here).
// DIR = -1 (forward; NF = 3, or 1): loads all y from the full-y buffer,
//   stores the disk ky to the compact buffer.
// The next field's loads are issued before the current field's transform.
template <typename T, int N, int DIR, int NF>
__global__ __launch_bounds__((Cfg<T, N>::NT)) void k_uxw_y(const cpx<T>* __restrict__ src,
                                                         cpx<T>* __restrict__ dst, Geo g,
                                                         Tables tb, const cpx<T>* __restrict__ tw,
                                                         long long ntiles) {
  using C = Cfg<T, N>;
  constexpr int E = C::E, TP = C::TP, B = C::B, LD = C::LD;
  // Two schedules, chosen by measurement (MI355X, 512^3 and 1024^3, FP64 and
  // FP32):
  //  STREAM (inverse pass, N <= 512): no branch around any load or store and
  //    the field loop written out, so the compiler's waits for the field in
  //    hand never include the prefetched next field or earlier stores
  //    (outstanding loads and stores retire from one in-order counter); the
  //    twiddles come from an LDS copy for the same reason.
  //  otherwise: a loop over the fields, loads and stores under their masks,
  //    twiddles from global memory (at N = 1024 the written-out form and the
  //    LDS or opaque twiddle reads each cost about a quarter more time).
  constexpr bool STREAM = DIR > 0 && N <= 512;
  // nontemporal stores up to N = 512; at 1024 they made both y passes slower
  // (1024^3 FP64 K2 10.8 -> 11.5 ms, K4 5.95 -> 6.23 ms)
  constexpr bool NTS = (GBG_NT & 2) != 0 && N <= 512;
  __shared__ cpx<T> lds[C::NBUF * B * LD];
  const int pl = threadIdx.x % B, t = threadIdx.x / B;
  const long long tile = fast::xcd_tile();
  if (tile >= ntiles) return;
  const long long p2 = tile * B + pl;
  constexpr bool TWL =
      STREAM && (size_t)C::NBUF * B * LD * sizeof(cpx<T>) + (size_t)N * sizeof(cpx<T>) <= 160 * 1024;
  __shared__ cpx<T> twl[TWL ? N : 1];
  // complete before its first use, which follows the first exchange barrier
  if constexpr (TWL)
    for (int i = threadIdx.x; i < N; i += C::NT) twl[i] = tw[i];
  auto twp = [&]() {
    if constexpr (TWL)
      return (const cpx<T>*)twl;
    else
      return tw;
  };
  // col: the lane's column in the row pitch nhp (its padding column when
  // col >= nhk), used for addresses; kx: the same, 0 in padding, for tables
  int z = 0, col = 0, kx = 0;
  bool valid = false;
  if (p2 < (long long)g.nz * g.nhp) {
    z = (int)(p2 / g.nhp);
    col = (int)(p2 - (long long)z * g.nhp);
    valid = col < g.nhk;
    kx = valid ? col : 0;
  }
  const double q1 = tb.ks1[kx];
  // compact (disk) offsets of this thread's elements; -1 = not held
  int coff[E];
#pragma unroll
  for (int m = 0; m < E; ++m) {
    const int ky = t + TP * m;
    const int kyc = kyc_of(ky, g);
    coff[m] = (valid && kyc >= 0 && q1 + tb.ks2[ky] <= g.hi2) ? kyc * g.nhp : -1;
  }
  const long long cbase = (long long)z * g.nyc * g.nhp + col;   // compact ky
  const long long fbase = (long long)z * g.ny * g.nhp + col;    // all y
  cpx<T>* pen = lds + pl * LD;
  cpx<T>* pen1 = C::ONE ? pen : lds + (B + pl) * LD;
  auto held = [&](int m) { return DIR > 0 ? coff[m] >= 0 : valid; };
  auto idx = [&](int m) {
    return DIR > 0 ? (long long)coff[m] : (long long)(t + TP * m) * g.nhp;
  };
  // STREAM: entries not held read the pencil's first entry (a valid address)
  // and are zeroed where used.  With B dividing the row pitch every lane of
  // a launched tile addresses the buffer, and all of them store (zeros from
  // the padding columns kx >= nhk).
  constexpr bool ALL_STORE = 16 % B == 0;
  const bool st_ok = STREAM ? (ALL_STORE || p2 < (long long)g.nz * g.nhp) : valid;
  cpx<T> nxt[E];
  auto load = [&](int f, cpx<T>(&v)[E]) {
    const cpx<T>* in = DIR > 0 ? src + f * g.s1 + cbase : src + f * g.s2 + fbase;
#pragma unroll
    for (int m = 0; m < E; ++m) {
      if constexpr (STREAM) {
        long long zi = held(m) ? idx(m) : 0;
        asm("" : "+v"(zi));  // keeps the optimiser from splitting the load by cases
        v[m] = ld_s(in + zi);
      } else {
        v[m] = held(m) ? ld_s(in + idx(m)) : mk<T>(T(0), T(0));
      }
    }
  };
  if constexpr (DIR > 0 && NF == 6 && curl_in_y<N>()) {
    // u x omega: the z pass wrote 0..4 = U1, U2, U3, D1 = F(i a3 u1), D2 =
    // F(i a3 u2); the six fields out (0..5 = u1, u2, u3, w1, w2, w3) take w
    // from them in registers:
    //   w1 = i a2 U3 - D2,  w2 = D1 - i a1 U3,  w3 = i (a1 U2 - a2 U1)
    // with two raw fields R1, R2 held, each load issued at least one transform
    // before its use:  R1 = U3, R2 = D2 | w1, R2 <- D1 | w2, R2 <- U1 | u3,
    // R1 <- U2 | w3 | u1 | u2.
    // the ky factors from an LDS copy (held in registers they spilled in
    // FP32 at Ny = 1024), complete after the barrier below
    __shared__ T a2l[N];
    for (int i = threadIdx.x; i < N; i += C::NT) a2l[i] = T(tb.kd2[i]);
    __syncthreads();
    const T a1 = T(tb.kd1[kx]);
    cpx<T> R1[E], R2[E];
    auto out = [&](cpx<T>(&v)[E], int f) {
      fast::fft<T, N, E, DIR, 0, C::PD, C::ONE>(v, pen, pen1, t, twp());
      if (st_ok) {
        cpx<T>* o = dst + f * g.s2 + fbase;
#pragma unroll
        for (int m = 0; m < E; ++m) st_s<NTS>(o + (long long)(t + TP * m) * g.nhp, v[m]);
      }
      if constexpr (STREAM) {
        asm volatile("" ::: "memory");
        __builtin_amdgcn_sched_barrier(0);
      }
    };
    // STREAM: entries not held were loaded from the pencil's first entry
    auto held_only = [&](cpx<T> x, int m) { return STREAM ? keep_or_zero(x, held(m)) : x; };
    load(2, R1);
    load(4, R2);
    // step f: the field formed, the load issued before its transform, the
    // output field (the steps written out)
    auto curl_step = [&](int f) {
      cpx<T> v[E];
      int of;
      if (f == 0) {
#pragma unroll
        for (int m = 0; m < E; ++m) v[m] = held_only(csub(cmuli(cscale(R1[m], a2l[t + TP * m])), R2[m]), m);
        load(3, R2);
        of = 3;
      } else if (f == 1) {
#pragma unroll
        for (int m = 0; m < E; ++m) v[m] = held_only(csub(R2[m], cmuli(cscale(R1[m], a1))), m);
        load(0, R2);
        of = 4;
      } else if (f == 2) {
#pragma unroll
        for (int m = 0; m < E; ++m) v[m] = held_only(R1[m], m);
        load(1, R1);
        of = 2;
      } else if (f == 3) {
#pragma unroll
        for (int m = 0; m < E; ++m)
          v[m] = held_only(cmuli(csub(cscale(R1[m], a1), cscale(R2[m], a2l[t + TP * m]))), m);
        of = 5;
      } else if (f == 4) {
#pragma unroll
        for (int m = 0; m < E; ++m) v[m] = held_only(R2[m], m);
        of = 0;
      } else {
#pragma unroll
        for (int m = 0; m < E; ++m) v[m] = held_only(R1[m], m);
        of = 1;
      }
      out(v, of);
    };
#pragma unroll
    for (int f = 0; f < 6; ++f) curl_step(f);
    return;
  }
  auto step = [&](int f) {
    cpx<T> v[E];
#pragma unroll
    for (int m = 0; m < E; ++m) v[m] = STREAM ? keep_or_zero(nxt[m], held(m)) : nxt[m];
    if (f + 1 < NF) load(f + 1, nxt);
    fast::fft<T, N, E, DIR, 0, C::PD, C::ONE>(v, pen, pen1, t, twp());
    if (DIR > 0) {
      if (st_ok) {
        cpx<T>* o = dst + f * g.s2 + fbase;
#pragma unroll
        for (int m = 0; m < E; ++m) st_s<NTS>(o + (long long)(t + TP * m) * g.nhp, v[m]);
      }
    } else {
      cpx<T>* o = dst + f * g.s1 + cbase;
#pragma unroll
      for (int m = 0; m < E; ++m)
        if (coff[m] >= 0) st_s<NTS>(o + coff[m], v[m]);
    }
  };
  load(0, nxt);
  if constexpr (STREAM) {
#pragma unroll
    for (int f = 0; f < NF; ++f) {
      step(f);
      asm volatile("" ::: "memory");
      __builtin_amdgcn_sched_barrier(0);
    }
  } else {
#pragma unroll 1
    for (int f = 0; f < NF; ++f) step(f);
  }
}

// ---- K3: x pass -- C2R of 6 rows, u x w, R2C of 3 rows ------------------------
// One unit = a pair of rows (r0, r1) of the flattened (y, z) index; rows have
// pitch nhp.  Per row the six fields go through three complex transforms in
// the pairs (u1, w1), (u2, w2), (u3, w3), and the cross product is built as
// they arrive, so at most five real rows are held in registers:
//   after (u2, w2):  n3 = u1 w2 - u2 w1
//   after (u3, w3):  n1 = u2 w3 - u3 w2,  n2 = u3 w1 - u1 w3
// (n1, n2) of a row share one forward transform, n3 of the two rows another;
// n3 of the first row waits in LDS.  When a pencil sits in one wave (S) the
// exchanges need no second LDS buffer (a wave's LDS operations execute in
// order), the twiddle table is read from LDS, and with STG the half-spectrum
// entries of the next field pair are copied from HBM straight into a per-wave
// LDS staging area (global_load_lds) while the current pair is transformed.
// Without S, FP64 table loads go through fresh() so the optimiser does not
// hoist the twiddles of every stage and hold them in registers.
// PROD (the dealiased scalar product): the two fields a, b of a row share one
// complex transform, p = a b is formed in registers, and p of the two rows of
// the unit share one forward transform (one output field).
// cache-policy bits of the staging copies (2: nontemporal)
constexpr int kLdsAux = (GBG_NT & 4) != 0 ? 2 : 0;

template <typename T, int N, int V>
struct XStage {
  static constexpr int EPL = 16 / (int)sizeof(cpx<T>);        // entries per lane per copy
  // copies per field row: three in FP64 and in FP32 at N >= 1024, so the kept
  // kx range of the 2/3 cutoff (N/3 + 1 entries) fits the staging area
  static constexpr int CH = (sizeof(T) == 8 || N >= 1024) ? 3 : 2;
  static constexpr int TPE = Cfg<T, N>::XTP * EPL;             // entries per copy per unit
  static constexpr int CAP = CH * TPE;                         // staged entries per row
  static constexpr int WAVE_ELEMS = 2 * CH * 64 * EPL;         // staging elements per wave
};

template <typename T, int N, bool STG, bool PROD, int V>
__global__ __launch_bounds__((XCfg<T, N, V>::XNT), 2) void k_uxw_x(const cpx<T>* __restrict__ w2,
                                                                 cpx<T>* __restrict__ w3, Geo g,
                                                                 const cpx<T>* __restrict__ tw,
                                                                 T scale, long long ntiles) {
  using C = XCfg<T, N, V>;
  using X = XStage<T, N, V>;
  constexpr int E = C::XE, TP = C::XTP, LD = C::XLD, S = C::XSYNC;
  constexpr int QOFF = PROD ? 1 : 3;   // field pairs (q, q + QOFF)
  static_assert(!STG || S, "staging needs a pencil per wave");
  constexpr int NBUF = S ? 1 : 2;
  constexpr int NW = C::XNT / 64;
  __shared__ cpx<T> lds[NBUF * C::U * LD];
  __shared__ T held[S ? C::U * E * TP : 1];
  // with S the stages read the per-stage twiddle table (bank-conflict free
  // for the consecutive butterflies of a wave's lanes)
  __shared__ cpx<T> twl[S ? fast::ts_count<N>() : 1];
  __shared__ cpx<T> stg[STG ? NW * X::WAVE_ELEMS : 1];
  if (S) {
    fast::ts_fill<T, N, C::XNT>(twl, tw, threadIdx.x);
    __syncthreads();
  }
  auto twp = [&]() { return S ? (const cpx<T>*)twl : fresh<sizeof(T) == 8>(tw); };
  const int un = threadIdx.x / TP, t = threadIdx.x % TP;
  const long long tile = fast::xcd_tile();
  const long long nrows = (long long)g.ny * g.nz;
  const long long r0 = 2 * (tile * C::U + un), r1 = r0 + 1;
  const bool ok0 = tile < ntiles && r0 < nrows, ok1 = tile < ntiles && r1 < nrows;
  const long long ro0 = ok0 ? r0 * g.nhp : 0, ro1 = ok1 ? r1 * g.nhp : 0;
  cpx<T>* pen = lds + un * LD;
  cpx<T>* pen1 = S ? pen : lds + (C::U + un) * LD;
  T* hold = held + (S ? un * E * TP + t : 0);
  // staging: this wave's area holds field a then field b, CH copies each; copy
  // c of unit u' (lanes u' TP .. u' TP + TP - 1) lands at c 64 EPL + u' TPE
  cpx<T>* sw = stg + (STG ? (threadIdx.x / 64) * X::WAVE_ELEMS : 0);
  const int uw = (threadIdx.x % 64) / TP;
  auto issue = [&](long long ro, int q) {
    if constexpr (STG) {
#pragma unroll
      for (int c = 0; c < X::CH; ++c) {
        const int e0 = c * X::TPE + t * X::EPL;
        const int e = e0 < g.nhk ? e0 : 0;   // entries past nhk are never read back
        __builtin_amdgcn_global_load_lds(
            (const void __attribute__((address_space(1)))*)(w2 + q * g.s2 + ro + e),
            (void __attribute__((address_space(3)))*)(sw + c * 64 * X::EPL), 16, 0, kLdsAux);
        __builtin_amdgcn_global_load_lds(
            (const void __attribute__((address_space(1)))*)(w2 + (q + QOFF) * g.s2 + ro + e),
            (void __attribute__((address_space(3)))*)(sw + (X::CH + c) * 64 * X::EPL), 16, 0, kLdsAux);
      }
    }
  };
  // the raw entries of pair q of a row: from the staging area or from HBM
  auto fetch = [&](cpx<T> (&a)[E], cpx<T> (&b)[E], long long ro, bool ok, int q) {
    if constexpr (STG) {
      __builtin_amdgcn_s_waitcnt(0x0F70);   // vmcnt(0): this wave's copies have landed
#pragma unroll
      for (int m = 0; m < E; ++m) {
        const int k = t + TP * m;
        const int kk = k >= fast::half_count<N>() ? N - k : k;
        a[m] = mk<T>(T(0), T(0));
        b[m] = mk<T>(T(0), T(0));
        if (kk < g.nhk) {
          const int idx = (kk / X::TPE) * 64 * X::EPL + uw * X::TPE + kk % X::TPE;
          a[m] = sw[idx];
          b[m] = sw[X::CH * 64 * X::EPL + idx];
        }
      }
      __builtin_amdgcn_s_waitcnt(0xC07F);   // lgkmcnt(0): staging reads done before reuse
    } else {
      fast::c2r_pair_fetch<T, N, E>(a, b, w2 + q * g.s2 + ro, w2 + (q + QOFF) * g.s2 + ro, ok, ok,
                                    t, g);
    }
  };
  if constexpr (PROD) {
    // row 0: fetch (a, b), stage row 1, transform, p0 = a b; row 1 the same;
    // then p0 + i p1 forward and split into the two rows' half spectra
    T P0[E];
    cpx<T> z[E];
    issue(ro0, 0);
    {
      cpx<T> a[E], b[E];
      fetch(a, b, ro0, ok0, 0);
      issue(ro1, 0);
      fast::c2r_pair_combine<T, N, E>(z, a, b, t, g);
    }
    fast::fft<T, N, E, 1, S, C::XPD, false, S != 0>(z, pen, pen1, t, twp());
#pragma unroll
    for (int m = 0; m < E; ++m) P0[m] = (z[m].x * scale) * (z[m].y * scale);
    {
      cpx<T> a[E], b[E];
      fetch(a, b, ro1, ok1, 0);
      fast::c2r_pair_combine<T, N, E>(z, a, b, t, g);
    }
    fast::fft<T, N, E, 1, S, C::XPD, false, S != 0>(z, pen, pen1, t, twp());
#pragma unroll
    for (int m = 0; m < E; ++m) z[m] = mk<T>(P0[m], (z[m].x * scale) * (z[m].y * scale));
    fast::fft<T, N, E, -1, S, C::XPD, false, S != 0>(z, pen, pen1, t, twp());
    fast::r2c_pair_store<T, N, E>(z, pen, w3 + ro0, w3 + ro1, ok0, ok1, t, g);
    return;
  }
  T N3[E], n3a[E];
  // EARLY: the raw entries na, nb of the second row's first field pair are
  // fetched before the first row's output stores are issued -- the staging
  // wait covers every outstanding load and store, so fetching them after the
  // stores would also wait for the stores to complete -- and the two rows are
  // written out.  Elsewhere the rows loop and each fetches its first pair: at
  // Nx = 1024 without staging the written-out form spills, and where it fits
  // otherwise (FP64 at 32 and 256, FP32 at 1024) it measured no faster.
  constexpr bool EARLY = sizeof(T) == 4 ? N <= 512 : (N == 64 || N == 128 || N == 512);
  cpx<T> na[E], nb[E];
  issue(ro0, 0);
  if constexpr (EARLY) {
    fetch(na, nb, ro0, ok0, 0);
    issue(ro0, 1);
  }
  auto row = [&](int h) {
    const long long ro = h ? ro1 : ro0;
    const bool ok = h ? ok1 : ok0;
    T U1[E], W1[E], U2[E], W2[E];
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      cpx<T> z[E];
      if (EARLY && q == 0) {
        fast::c2r_pair_combine<T, N, E>(z, na, nb, t, g);
      } else {
        cpx<T> a[E], b[E];
        fetch(a, b, ro, ok, q);
        if (q < 2)
          issue(ro, q + 1);
        else if (h == 0)
          issue(ro1, 0);
        fast::c2r_pair_combine<T, N, E>(z, a, b, t, g);
      }
      fast::fft<T, N, E, 1, S, C::XPD, false, S != 0>(z, pen, pen1, t, twp());
      if (q == 0) {
#pragma unroll
        for (int m = 0; m < E; ++m) {
          U1[m] = z[m].x * scale;
          W1[m] = z[m].y * scale;
        }
      } else if (q == 1) {
#pragma unroll
        for (int m = 0; m < E; ++m) {
          U2[m] = z[m].x * scale;
          W2[m] = z[m].y * scale;
          N3[m] = U1[m] * W2[m] - U2[m] * W1[m];
        }
      } else {
#pragma unroll
        for (int m = 0; m < E; ++m) {
          const T u3 = z[m].x * scale, w3v = z[m].y * scale;
          z[m] = mk<T>(U2[m] * w3v - u3 * W2[m], u3 * W1[m] - U1[m] * w3v);
        }
        fast::fft<T, N, E, -1, S, C::XPD, false, S != 0>(z, pen, pen1, t, twp());
        if (EARLY && h == 0) {
          fetch(na, nb, ro1, ok1, 0);
          issue(ro1, 1);
        }
        fast::r2c_pair_store<T, N, E>(z, pen, w3 + ro, w3 + g.s2 + ro, ok, ok, t, g);
      }
    }
    if (h == 0) {
#pragma unroll
      for (int m = 0; m < E; ++m) {
        if (S)
          hold[m * TP] = N3[m];
        else
          n3a[m] = N3[m];
      }
    }
  };
  if constexpr (EARLY) {
    row(0);
    row(1);
  } else {
#pragma unroll 1
    for (int h = 0; h < 2; ++h) row(h);
  }
  cpx<T> v[E];
#pragma unroll
  for (int m = 0; m < E; ++m) v[m] = mk<T>(S ? hold[m * TP] : n3a[m], N3[m]);
  fast::fft<T, N, E, -1, S, C::XPD, false, S != 0>(v, pen, pen1, t, twp());
  fast::r2c_pair_store<T, N, E>(v, pen, w3 + 2 * g.s2 + ro0, w3 + 2 * g.s2 + ro1, ok0, ok1, t, g);
}

// ---- K5: z-forward, Leray projection, mask, mean-free ------------------------
// Tiles: B consecutive pencils of the output's flattened (ky, kx) index, so
// the (largest) output stream is written in aligned segments.  With a tile
// list, workgroup i takes tile list[i] (the tiles that hold kept pencils) and
// k_uxw_zero_tiles writes the others; without one, workgroup i takes tile i
// and tiles without a kept pencil are written as zeros here.
// NC = 3: u x omega (Leray projection, mask, mean-free); NC = 1: the scalar
// product (mask only: its mean mode is kept).
// Persistent form (one workgroup per CU walking its XCD's share of the
// tiles): FP64 at the lengths where the tile loop holds in 256 VGPRs without
// spilling.  In FP32 the tile loop spills at the 128-VGPR limit (also with
// 32-bit element offsets as in K1).
template <typename T, int N>
constexpr bool z_forward_persist() {
  return sizeof(T) == 8 && (N == 128 || N == 512);
}

// The integrating-factor combination K5 can apply to N^ before it leaves the
// registers (the step either side of N^ in an integrating-factor RK2 update,
// examples/HIST_CPU.jl:267-282):
//   dst = e (alpha a + beta b + delta N^) + gamma N^,  e = exp(c |k|^2),
// |k|^2 = (k1^2 + k2^2) + k3^2 from the squared family in FP64 and e rounded to
// T, as the standalone exponential kernel forms it (_exponential_exponent,
// src/Convolution.jl:1289-1294).  a, b may be null (their terms dropped); with
// store_n = 0 the output spectrum N^ itself is not written.
template <typename T>
struct Epi {
  cpx<T>* dst;
  const cpx<T>* a;
  const cpx<T>* b;
  T alpha, beta, delta, gamma;
  double coef;
  int store_n;
};

// the combination at entry o of component c (o < M) for the mode of squared
// wavenumber q
template <typename T>
__device__ __forceinline__ void epi_store(const Epi<T>& ep, long long co, double q, cpx<T> n) {
  const T e = T(exp(ep.coef * q));
  cpx<T> acc = cscale(n, ep.delta);
  if (ep.a) acc = cadd(cscale(ep.a[co], ep.alpha), acc);
  if (ep.b) acc = cadd(acc, cscale(ep.b[co], ep.beta));
  ep.dst[co] = cadd(cscale(acc, e), cscale(n, ep.gamma));
}

template <typename T, int N, int NC, bool EPI = false>
__global__ __launch_bounds__((Cfg<T, N>::NT)) void k_uxw_z_forward(
    const cpx<T>* __restrict__ w1, cpx<T>* __restrict__ nh, Geo g, Tables tb,
    const cpx<T>* __restrict__ tw, const int* __restrict__ list, long long ntiles, Epi<T> ep) {
  using C = Cfg<T, N>;
  constexpr int E = C::E, TP = C::TP, B = C::B, LD = C::LD;
  constexpr bool NTS = (GBG_NT & 2) != 0;
  __shared__ cpx<T> lds[C::NBUF * B * LD];
  const int pl = threadIdx.x % B, t = threadIdx.x / B;
  // A global load consumed between the component loads and their last use
  // would wait for all of them (the counter of outstanding loads drains in
  // order): for N <= 512 (STREAM) the twiddles come from an LDS copy
  // (complete before the first exchange barrier) and the component loads
  // carry no branch; at N = 1024 the plain global reads measured faster.  The
  // mask bits are computed before the components load.
  constexpr bool STREAM = N <= 512;
  constexpr bool TWL =
      STREAM && (size_t)C::NBUF * B * LD * sizeof(cpx<T>) + (size_t)N * sizeof(cpx<T>) <= 160 * 1024;
  __shared__ cpx<T> twl[TWL ? N : 1];
  // The kz derivative factors (and, where LDS holds them too, the kz squared
  // family) are read from LDS copies: a global table read between the stores
  // of one kz and the next waits for every store issued before it.  Both
  // copies are complete before their first use, which follows the barrier of
  // the first tile's disk vote.
  __shared__ T a3l[N];
  constexpr bool Q3L = (size_t)C::NBUF * B * LD * sizeof(cpx<T>) + (TWL ? N * sizeof(cpx<T>) : 0) +
                           (size_t)N * (sizeof(T) + sizeof(double)) <=
                       160 * 1024;
  __shared__ double q3l[Q3L ? N : 1];
  for (int i = threadIdx.x; i < N; i += C::NT) {
    if constexpr (TWL) twl[i] = tw[i];
    a3l[i] = T(tb.kd3[i]);
    if constexpr (Q3L) q3l[i] = tb.ks3[i];
  }
  auto q3 = [&](int kz) {
    if constexpr (Q3L)
      return q3l[kz];
    else
      return tb.ks3[kz];
  };
  auto twp = [&]() {
    if constexpr (TWL)
      return (const cpx<T>*)twl;
    else
      return tw;
  };
  // one tile; the barrier at its start also orders the previous tile's last
  // LDS reads before this tile's first writes
  auto run_tile = [&](long long lt) {
    const long long tile = (list && lt < ntiles) ? (long long)list[lt] : lt;
    const long long p5l = tile * B + pl;
    const bool valid = lt < ntiles && p5l < (long long)g.ny * g.nh;
    const int p5 = valid ? (int)p5l : 0;
    int kx = 0, ky = 0, kyc = -1;
    double qxy = 0.0;
    T a1 = T(0), a2 = T(0);
    bool disk = false;
    if (valid) {
      ky = p5 / g.nh;
      kx = p5 - ky * g.nh;
      kyc = kyc_of(ky, g);
      qxy = tb.ks1[kx] + tb.ks2[ky];
      a1 = T(tb.kd1[kx]);
      a2 = T(tb.kd2[ky]);
      disk = kx < g.nhk && kyc >= 0 && qxy <= g.hi2;
    }
    const long long plane = (long long)g.ny * g.nh;
    if (!__syncthreads_or(disk)) {
      if (valid) {
#pragma unroll
        for (int c = 0; c < NC; ++c)
#pragma unroll
          for (int m = 0; m < E; ++m) {
            const long long co = c * g.M + (long long)(t + TP * m) * plane + p5;
            if constexpr (EPI) {
              if (ep.store_n) nh[co] = mk<T>(T(0), T(0));
              epi_store(ep, co, qxy + q3(t + TP * m), mk<T>(T(0), T(0)));
            } else {
              st_s<NTS>(nh + co, mk<T>(T(0), T(0)));
            }
          }
      }
      return;
    }
    const long long zs = (long long)g.nyc * g.nhp;
    const long long cb = (long long)(disk ? kyc : 0) * g.nhp + (disk ? kx : 0);
    cpx<T>* pen = lds + pl * LD;
    cpx<T>* pen1 = C::ONE ? pen : lds + (B + pl) * LD;
    bool keep[E];
#pragma unroll
    for (int m = 0; m < E; ++m) {
      const int kz = t + TP * m;
      keep[m] = disk & (qxy + q3(kz) <= g.hi2) &  // no branch
                (NC == 1 || !(kx == 0 && ky == 0 && kz == 0));
    }
    // STREAM: components off the disk read the compact buffer's first entry (a
    // valid address) and are zeroed just before their transform
    cpx<T> n[NC][E];
#pragma unroll
    for (int c = 0; c < NC; ++c) {
      const cpx<T>* in = w1 + c * g.s1 + cb;
#pragma unroll
      for (int m = 0; m < E; ++m) {
        if constexpr (STREAM) {
          long long zi = disk ? (long long)(t + TP * m) * zs : 0;
          asm("" : "+v"(zi));  // keeps the optimiser from splitting the load by cases
          n[c][m] = ld_s(in + zi);
        } else {
          n[c][m] = disk ? ld_s(in + (long long)(t + TP * m) * zs) : mk<T>(T(0), T(0));
        }
      }
    }
#pragma unroll
    for (int c = 0; c < NC; ++c) {
      if constexpr (STREAM) {
#pragma unroll
        for (int m = 0; m < E; ++m) n[c][m] = keep_or_zero(n[c][m], disk);
      }
      fast::fft<T, N, E, -1, 0, C::PD, C::ONE>(n[c], pen, pen1, t, twp());
      if constexpr (STREAM) __builtin_amdgcn_sched_barrier(0);
    }
    if (!valid) return;
    if constexpr (NC == 1) {
#pragma unroll
      for (int m = 0; m < E; ++m) {
        const long long o = (long long)(t + TP * m) * plane + p5;
        st_s<NTS>(nh + o, mk<T>(keep[m] ? n[0][m].x : T(0), keep[m] ? n[0][m].y : T(0)));
      }
    } else {
#pragma unroll
      for (int m = 0; m < E; ++m) {
        const int kz = t + TP * m;
        const T a3 = a3l[kz];
        // (_helmholtz_projection_kernel!, src/DifferentialOperators.jl:191-230), alpha = 1
        const T sq = a1 * a1 + a2 * a2 + a3 * a3;
        const T den = sq == T(0) ? T(1) : sq;
        const cpx<T> dot = cadd(cadd(cscale(n[0][m], a1), cscale(n[1][m], a2)), cscale(n[2][m], a3));
        const cpx<T> lon = mk<T>(dot.x / den, dot.y / den);
        const cpx<T> p1 = csub(n[0][m], cscale(lon, a1));
        const cpx<T> p2 = csub(n[1][m], cscale(lon, a2));
        const cpx<T> p3 = csub(n[2][m], cscale(lon, a3));
        const long long o = (long long)kz * plane + p5;
        // component-wise selects (a select of whole complex values is lowered
        // through a private array in FP64)
        const cpx<T> r1 = mk<T>(keep[m] ? p1.x : T(0), keep[m] ? p1.y : T(0));
        const cpx<T> r2 = mk<T>(keep[m] ? p2.x : T(0), keep[m] ? p2.y : T(0));
        const cpx<T> r3 = mk<T>(keep[m] ? p3.x : T(0), keep[m] ? p3.y : T(0));
        if constexpr (EPI) {
          if (ep.store_n) {
            nh[o] = r1;
            nh[g.M + o] = r2;
            nh[2 * g.M + o] = r3;
          }
          const double q = qxy + q3(kz);
          epi_store(ep, o, q, r1);
          epi_store(ep, g.M + o, q, r2);
          epi_store(ep, 2 * g.M + o, q, r3);
        } else {
          st_s<NTS>(nh + o, r1);
          st_s<NTS>(nh + g.M + o, r2);
          st_s<NTS>(nh + 2 * g.M + o, r3);
        }
      }
    }
  };
  if constexpr (z_forward_persist<T, N>()) {
    const long long per = (ntiles + fast::kXcd - 1) / fast::kXcd;
    const int xcd = blockIdx.x % fast::kXcd;
    const long long L = gridDim.x / fast::kXcd;
    const long long tend = (xcd + 1) * per < ntiles ? (xcd + 1) * per : ntiles;
#pragma unroll 1
    for (long long lt = xcd * per + blockIdx.x / fast::kXcd; lt < tend; lt += L) run_tile(lt);
  } else {
    run_tile(fast::xcd_tile());
  }
}

// The output tiles that hold no kept (kx, ky) pencil: zeros in all kz planes
// and components.  Workgroup (x, y) covers kZeroRun consecutive pencil
// indices of kZeroPlanes consecutive (component, kz) planes: the tile flags of
// its stretch are read once, then each plane's stretch is streamed out with no
// load between the stores.
constexpr int kZeroRun = 2048;
constexpr int kZeroPlanes = 16;
template <typename T>
__global__ __launch_bounds__(256) void k_uxw_zero_tiles(cpx<T>* __restrict__ nh,
                                                        const int* __restrict__ flag, int b,
                                                        long long plane, int nplanes) {
  constexpr int R = kZeroRun / 256;
  const cpx<T> zero = mk<T>(T(0), T(0));
  const long long base = (long long)blockIdx.x * kZeroRun + threadIdx.x;
  bool z[R];
#pragma unroll
  for (int i = 0; i < R; ++i) {
    const long long p5 = base + i * 256;
    z[i] = p5 < plane && !flag[p5 / b];
  }
  const int p0 = blockIdx.y * kZeroPlanes;
  const int p1 = p0 + kZeroPlanes < nplanes ? p0 + kZeroPlanes : nplanes;
  for (int pl = p0; pl < p1; ++pl) {
    cpx<T>* out = nh + (long long)pl * plane;
#pragma unroll
    for (int i = 0; i < R; ++i)
      if (z[i]) st_s(out + base + i * 256, zero);
  }
}

// The same output tiles under the integrating-factor combination: N^ is zero
// there, so dst = e (alpha a + beta b) (and N^ = 0 when it is stored).  Plane
// pl is (component pl / nz, kz = pl mod nz); entry p5 of a plane is (kx, ky) =
// (p5 mod nh, p5 / nh).
template <typename T>
__global__ __launch_bounds__(256) void k_uxw_zero_tiles_epi(cpx<T>* __restrict__ nh, Epi<T> ep,
                                                            const int* __restrict__ flag, int b,
                                                            long long plane, int nplanes, int nz,
                                                            int nhx, Tables tb) {
  constexpr int R = kZeroRun / 256;
  const long long base = (long long)blockIdx.x * kZeroRun + threadIdx.x;
  bool z[R];
  double qxy[R];
#pragma unroll
  for (int i = 0; i < R; ++i) {
    const long long p5 = base + i * 256;
    z[i] = p5 < plane && !flag[p5 / b];
    const int ky = z[i] ? (int)(p5 / nhx) : 0, kx = z[i] ? (int)(p5 - (long long)ky * nhx) : 0;
    qxy[i] = tb.ks1[kx] + tb.ks2[ky];
  }
  const int p0 = blockIdx.y * kZeroPlanes;
  const int p1 = p0 + kZeroPlanes < nplanes ? p0 + kZeroPlanes : nplanes;
  for (int pl = p0; pl < p1; ++pl) {
    const double q3 = tb.ks3[pl % nz];
#pragma unroll
    for (int i = 0; i < R; ++i) {
      if (!z[i]) continue;
      const long long co = (long long)pl * plane + base + i * 256;
      if (ep.store_n) nh[co] = mk<T>(T(0), T(0));
      epi_store(ep, co, qxy[i] + q3, mk<T>(T(0), T(0)));
    }
  }
}

// The combination over every mode from a stored N^ (the composed path):
// one thread per mode, the three components.
template <typename T>
__global__ __launch_bounds__(256) void k_uxw_epilogue(const cpx<T>* __restrict__ n, Epi<T> ep,
                                                      long long M, int s0, int s1, Tables tb) {
  const long long idx = (long long)blockIdx.x * 256 + threadIdx.x;
  if (idx >= M) return;
  const long long r = idx / s0;
  const int kx = (int)(idx - r * s0), ky = (int)(r % s1), kz = (int)(r / s1);
  const double q = (tb.ks1[kx] + tb.ks2[ky]) + tb.ks3[kz];
#pragma unroll
  for (int c = 0; c < 3; ++c) epi_store(ep, c * M + idx, q, n[c * M + idx]);
}

// ---- host side ----------------------------------------------------------------
// grid of a pass over nt tiles: one workgroup per tile (XCD-padded), or for a
// persistent pass at most one per CU (a multiple of the XCD count)
static unsigned pass_grid(const gbg_plan* p, long long nt, bool persist) {
  unsigned grid = fast::xcd_grid(nt);
  const unsigned cap = (unsigned)(p->ncu / fast::kXcd * fast::kXcd);
  if (persist && cap >= (unsigned)fast::kXcd && grid > cap) grid = cap;
  return grid;
}

// 32-bit element offsets inside one component of the u x omega buffers (the
// input and output spectra, M modes each, and the compact intermediates, s1
// each); GBG_WIDE_OFFSETS=1 at plan creation forces the 64-bit form
static bool narrow_offsets(const gbg_plan* p, const Geo& g) {
  return !p->wide_offsets && g.M < (1LL << 32) && g.s1 < (1LL << 32);
}

// K1; returns the launched grid (the partial-sum count)
template <typename T, int N>
static unsigned launch_z_inverse(const gbg_plan* p, const Geo& g, const cpx<T>* uh, cpx<T>* W1,
                                 const Tables& tb, const cpx<T>* twz, double* part, hipStream_t s) {
  using C = Cfg<T, N>;
  const long long nt = ((long long)g.nyc * g.nhp + C::B - 1) / C::B;
  // five fields where the y pass forms the kx and ky terms of w (curl_in_y
  // of Ny, a power of two here); six otherwise
  const bool five = g.ny <= 512;
  static_assert(curl_in_y<512>() && !curl_in_y<1024>(), "five-field threshold");
  unsigned grid;
  if (narrow_offsets(p, g)) {
    grid = pass_grid(p, nt, z_inverse_persist<T, N, true>());
    if (five)
      hipLaunchKernelGGL((k_uxw_z_inverse<T, N, true, true>), dim3(grid), dim3(C::NT), 0, s, uh, W1,
                         g, tb, twz, part, nt);
    else
      hipLaunchKernelGGL((k_uxw_z_inverse<T, N, true, false>), dim3(grid), dim3(C::NT), 0, s, uh, W1,
                         g, tb, twz, part, nt);
  } else {
    grid = pass_grid(p, nt, z_inverse_persist<T, N, false>());
    if (five)
      hipLaunchKernelGGL((k_uxw_z_inverse<T, N, false, true>), dim3(grid), dim3(C::NT), 0, s, uh,
                         W1, g, tb, twz, part, nt);
    else
      hipLaunchKernelGGL((k_uxw_z_inverse<T, N, false, false>), dim3(grid), dim3(C::NT), 0, s, uh,
                         W1, g, tb, twz, part, nt);
  }
  return grid;
}

// a y pass (K2, K4, P2, P4)
template <typename T, int N, int DIR, int NF>
static void launch_y(const Geo& g, const cpx<T>* src, cpx<T>* dst, const Tables& tb,
                     const cpx<T>* twy, hipStream_t s) {
  using C = Cfg<T, N>;
  const long long nt = ((long long)g.nz * g.nhp + C::B - 1) / C::B;
  hipLaunchKernelGGL((k_uxw_y<T, N, DIR, NF>), dim3(fast::xcd_grid(nt)), dim3(C::NT), 0, s, src,
                     dst, g, tb, twy, nt);
}

// K3 / P3 in tile-shape variant V: the next field pair is staged through LDS
// when a row sits in one wave and the kept kx range fits the staging area
template <typename T, int N, bool PROD, int V>
static void launch_x_v(const Geo& g, const cpx<T>* W2, cpx<T>* W1, const cpx<T>* twx, T scale,
                       hipStream_t s) {
  using C = XCfg<T, N, V>;
  const long long pairs = ((long long)g.ny * g.nz + 1) / 2;
  const long long nt = (pairs + C::U - 1) / C::U;
  if (C::XSYNC && g.nhk <= XStage<T, N, V>::CAP)
    hipLaunchKernelGGL((k_uxw_x<T, N, C::XSYNC != 0, PROD, V>), dim3(fast::xcd_grid(nt)),
                       dim3(C::XNT), 0, s, W2, W1, g, twx, scale, nt);
  else
    hipLaunchKernelGGL((k_uxw_x<T, N, false, PROD, V>), dim3(fast::xcd_grid(nt)), dim3(C::XNT), 0,
                       s, W2, W1, g, twx, scale, nt);
}

template <typename T, int N, bool PROD>
static void launch_x(int var, const Geo& g, const cpx<T>* W2, cpx<T>* W1, const cpx<T>* twx, T scale,
                     hipStream_t s) {
  if constexpr (N >= 256)
    if (var == 1) return launch_x_v<T, N, PROD, 1>(g, W2, W1, twx, scale, s);
  launch_x_v<T, N, PROD, 0>(g, W2, W1, twx, scale, s);
}

// K5's transform kernel over nt tiles (tile list `list`, or all tiles)
template <typename T, int N, int NC>
static void launch_z_forward(const gbg_plan* p, const Geo& g, long long nt, const cpx<T>* w,
                             cpx<T>* out, const Tables& tb, const cpx<T>* twz, const int* list,
                             hipStream_t s, const Epi<T>* ep = nullptr) {
  using C = Cfg<T, N>;
  const dim3 grid(pass_grid(p, nt, z_forward_persist<T, N>()));
  if constexpr (NC == 3) {
    if (ep) {
      hipLaunchKernelGGL((k_uxw_z_forward<T, N, NC, true>), grid, dim3(C::NT), 0, s, w, out, g, tb,
                         twz, list, nt, *ep);
      return;
    }
  }
  hipLaunchKernelGGL((k_uxw_z_forward<T, N, NC>), grid, dim3(C::NT), 0, s, w, out, g, tb, twz, list,
                     nt, Epi<T>{});
}

static bool fast_len(int n) {
  return n == 32 || n == 64 || n == 128 || n == 256 || n == 512 || n == 1024;
}

// call f(std::integral_constant<int, n>) for a supported power-of-two length
template <typename F>
static bool with_len(int n, F&& f) {
  switch (n) {
    case 32: f(std::integral_constant<int, 32>{}); return true;
    case 64: f(std::integral_constant<int, 64>{}); return true;
    case 128: f(std::integral_constant<int, 128>{}); return true;
    case 256: f(std::integral_constant<int, 256>{}); return true;
    case 512: f(std::integral_constant<int, 512>{}); return true;
    case 1024: f(std::integral_constant<int, 1024>{}); return true;
  }
  return false;
}

static void axis_host(int n, double L, bool half, std::vector<double>& ks) {
  const double unit = (2.0 * M_PI) / L;
  const int stored = half ? n / 2 + 1 : n;
  ks.resize(stored);
  for (int i = 0; i < stored; ++i) {
    const int f = (half || i < (n + 1) / 2) ? i : i - n;
    const double k = unit * (double)f;
    ks[i] = k * k;
  }
}

static Geo make_geo(gbg_plan* p, double kc) {
  Geo g;
  g.nx = p->n[0];
  g.ny = p->n[1];
  g.nz = p->n[2];
  g.nh = p->nh;
  g.nyq_x = (g.nx % 2 == 0) ? g.nh - 1 : -1;
  g.M = (long long)p->M;
  g.hi2 = kc * kc;
  std::vector<double> s1, s2;
  axis_host(g.nx, p->L[0], true, s1);
  axis_host(g.ny, p->L[1], false, s2);
  int nhk = 0;
  while (nhk < g.nh && s1[nhk] <= g.hi2) ++nhk;
  int ya = 0;
  while (ya < (g.ny + 1) / 2 && s2[ya] <= g.hi2) ++ya;
  int b = 0;
  while (b < g.ny - (g.ny + 1) / 2 && s2[g.ny - 1 - b] <= g.hi2) ++b;
  g.nhk = nhk < 1 ? 1 : nhk;
  g.nhp = (g.nhk + 15) / 16 * 16;
  g.ya = ya;
  g.nyc = ya + b;
  g.s1 = (long long)g.nz * g.nyc * g.nhp;
  g.s2 = (long long)g.nz * g.ny * g.nhp;
  return g;
}

// The plan's tile set for cutoff g.hi2 and tile width b, built on first use
// when `build` holds.  Returns nullptr (the last pass then classifies tiles
// itself) when the set is not cached and is not built now: `build` is false
// (a caller that must not allocate), the stream is being captured, or the
// plan already holds kTileSets sets.
static const TileSet* tile_set(gbg_plan* p, const Geo& g, int b, double kc, hipStream_t s,
                               bool build = true) {
  for (int i = 0; i < p->ntset; ++i)
    if (p->tset[i].kc == kc && p->tset[i].b == b) return &p->tset[i];
  if (!build || p->ntset >= kTileSets) return nullptr;
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  if (hipStreamIsCapturing(s, &cs) != hipSuccess || cs != hipStreamCaptureStatusNone) return nullptr;
  std::vector<double> s1, s2;
  axis_host(g.nx, p->L[0], true, s1);
  axis_host(g.ny, p->L[1], false, s2);
  const long long plane = (long long)g.ny * g.nh;
  const int nt = (int)((plane + b - 1) / b);
  // a tile is listed when one of its pencils may be kept; the slack only
  // lists extra tiles, which the last pass then writes as zeros itself
  const double hi = g.hi2 * (1.0 + 1e-9);
  std::vector<int> list, flag(nt, 0);
  for (int t = 0; t < nt; ++t) {
    for (long long q = (long long)t * b; q < (long long)(t + 1) * b && q < plane; ++q) {
      const int ky = (int)(q / g.nh), kx = (int)(q - (long long)ky * g.nh);
      if (kx < g.nhk && kyc_of(ky, g) >= 0 && s1[kx] + s2[ky] <= hi) {
        flag[t] = 1;
        break;
      }
    }
    if (flag[t]) list.push_back(t);
  }
  const int nd = (int)list.size();
  list.insert(list.end(), flag.begin(), flag.end());
  const size_t bytes = sizeof(int) * list.size();
  int* dev = nullptr;
  if (hipMalloc((void**)&dev, bytes) != hipSuccess) {
    (void)hipGetLastError();
    return nullptr;
  }
  if (hipMemcpy(dev, list.data(), bytes, hipMemcpyHostToDevice) != hipSuccess) {
    (void)hipGetLastError();
    (void)hipFree(dev);
    return nullptr;
  }
  TileSet& ts = p->tset[p->ntset++];
  ts.kc = kc;
  ts.b = b;
  ts.nd = nd;
  ts.nt = nt;
  ts.dev = dev;
  p->device_bytes += bytes;
  return &ts;
}

// The zero-tile writes of the last pass (the output tiles that hold no kept
// pencil) depend on nothing the other passes compute, and the transform
// kernel writes only the other tiles.  When the output overlaps no input
// they run on the plan's second stream beside one of the first passes
// (zero_fork_pass: K1 by default), forked from s before that pass
// and joined back into s after the last (zero_tiles_join); whenever the fork
// is not taken the last pass launches them on s itself.  Nor is the fork
// taken while s is being captured into a graph: the batched form's graph of
// many small items ran 1.6x longer in FP32 with a fork and join per item.
// A failed query or set-up call only declines the fork; its error is cleared
// so that the next launch check does not report it.
static bool aux_stream(gbg_plan* p, hipStream_t s) {
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  if (hipStreamIsCapturing(s, &cs) != hipSuccess) {
    (void)hipGetLastError();
    return false;
  }
  if (cs != hipStreamCaptureStatusNone) return false;
  if (p->aux) return true;
  hipStream_t a = nullptr;
  hipEvent_t f = nullptr, j = nullptr;
  if (hipStreamCreateWithFlags(&a, hipStreamNonBlocking) == hipSuccess &&
      hipEventCreateWithFlags(&f, hipEventDisableTiming) == hipSuccess &&
      hipEventCreateWithFlags(&j, hipEventDisableTiming) == hipSuccess) {
    p->aux = a;
    p->aux_fork = f;
    p->aux_join = j;
    return true;
  }
  (void)hipGetLastError();
  if (f) (void)hipEventDestroy(f);
  if (a) (void)hipStreamDestroy(a);
  return false;
}

static bool overlaps(const void* a, size_t abytes, const void* b, size_t bbytes) {
  const char *x = (const char*)a, *y = (const char*)b;
  return x < y + bbytes && y < x + abytes;
}

template <typename T>
static void launch_zero_tiles(const Geo& g, const TileSet* ts, int b, int nc, cpx<T>* out,
                              hipStream_t s, const Epi<T>* ep = nullptr, const Tables* tb = nullptr) {
  // planes (component c, kz) are consecutive: c M + kz plane = (c nz + kz) plane
  const long long plane = (long long)g.ny * g.nh;
  const int np = nc * g.nz;
  const dim3 zg((unsigned)((plane + kZeroRun - 1) / kZeroRun),
                (unsigned)((np + kZeroPlanes - 1) / kZeroPlanes));
  if (ep)
    hipLaunchKernelGGL((k_uxw_zero_tiles_epi<T>), zg, dim3(256), 0, s, out, *ep, ts->dev + ts->nd, b,
                       plane, np, g.nz, g.nh, *tb);
  else
    hipLaunchKernelGGL((k_uxw_zero_tiles<T>), zg, dim3(256), 0, s, out, ts->dev + ts->nd, b, plane,
                       np);
}

// The pass of u x omega beside which the zero tiles run on the second stream
// (0: none, in stream order in the last pass).  With nontemporal stores in the
// passes, measured per term on the MI355X: 512^3 FP64 none 5.15, beside K1
// 5.09, beside K3 5.14 ms; 1024^3 FP64 none 53.44, K1 53.43, K3 53.03 ms;
// 512^3 FP32 none 2.77, K1 2.73 ms; the integrating-factor RK2 step at 512^3
// FP64 beside K3 13.78, beside K1 13.55 ms.  K1 (one persistent workgroup per
// CU, bound by latency) leaves the memory system room; its time then includes
// the zero tiles' writes, and the last pass is the transform kernel alone.
// GBG_ZERO_FORK_F64 / _F32 = 0..4 override the choice (the results are the
// same bits).
template <typename T>
static int zero_fork_pass(bool combination) {
  (void)combination;
  const char* e = std::getenv(sizeof(T) == 8 ? "GBG_ZERO_FORK_F64" : "GBG_ZERO_FORK_F32");
  if (e && e[0] >= '0' && e[0] <= '4' && e[1] == 0) return e[0] - '0';
  return 1;
}

// fork point: recorded on s before the pass the zero tiles run beside
template <typename T>
static const TileSet* zero_tiles_fork(gbg_plan* p, const Geo& g, double kc, int nc, cpx<T>* out,
                                      const void* in0, size_t in0_bytes, const void* in1,
                                      size_t in1_bytes, hipStream_t s) {
  const size_t ob = (size_t)nc * g.M * sizeof(cpx<T>);
  if (overlaps(out, ob, in0, in0_bytes) || (in1 && overlaps(out, ob, in1, in1_bytes))) return nullptr;
  const TileSet* fork = nullptr;
  with_len(g.nz, [&](auto L) {
    constexpr int B = Cfg<T, decltype(L)::value>::B;
    const TileSet* ts = tile_set(p, g, B, kc, s);
    if (!ts || ts->nd == ts->nt || !aux_stream(p, s)) return;
    if (hipEventRecord(p->aux_fork, s) == hipSuccess)
      fork = ts;
    else
      (void)hipGetLastError();
  });
  return fork;
}

// the zero-tile kernel on the second stream, enqueued after the first pass's
// launch so that pass's workgroups are dispatched first and the zero writes
// fill the room they leave
template <typename T>
static bool zero_tiles_launch(gbg_plan* p, const Geo& g, const TileSet* ts, int nc, cpx<T>* out,
                              const Epi<T>* ep = nullptr, const Tables* tb = nullptr) {
  if (!ts) return false;
  if (hipStreamWaitEvent(p->aux, p->aux_fork, 0) != hipSuccess) {
    (void)hipGetLastError();
    return false;
  }
  launch_zero_tiles<T>(g, ts, ts->b, nc, out, p->aux, ep, tb);
  if (hipEventRecord(p->aux_join, p->aux) == hipSuccess) return true;
  // no join point: let the second stream drain, and the last pass writes the
  // zero tiles on s as well
  (void)hipGetLastError();
  (void)hipStreamSynchronize(p->aux);
  (void)hipGetLastError();
  return false;
}

static gbg_status zero_tiles_join(gbg_plan* p, hipStream_t s) {
  return check_hip(hipStreamWaitEvent(s, p->aux_join, 0), "join of the zero-tile stream");
}

// ---- the z-slab chain of the middle passes ------------------------------------
// K2, K3 and K4 (P2, P3, P4 of the scalar product) each work inside z planes:
// a y pencil (z, kx) and an x row (z, y) lie in one plane.  They can therefore
// run slab by slab, Z planes at a time, with a slab's full-y fields (NF after
// the y-inverse, NO after the x pass) in a ring of R slots of (NF + NO) Z Ny nhp
// entries instead of two full-size buffers.  A slot is read back soon after it
// is written, while the device's memory-side last-level cache (256 MB on the
// MI355X) still holds it, so the schedule's largest intermediates mostly stay
// off HBM.  The y-forward of slab j writes its compact output into the first
// NO fields of the compact buffer at slab j's planes, whose input the
// y-inverse of slab j has already read (both in stream order); the last pass
// reads it there.  Slab j runs on stream j mod R (s itself for j mod R = 0),
// forked from s after the first pass and joined back before the last; for a
// caller that must not create resources, under a graph capture, or when a
// stream cannot be made, every slab runs on s in order.
struct SlabCfg {
  int z;   // planes per slab (0: no slab chain, the two full-size buffers)
  int r;   // ring slots, the slabs in flight
};

// The chain runs when the caller asks for it (GBG_UXW_SLABS): Z planes per
// slab such that a ring slot holds at most kSlabSlotBytes (two slots in
// flight stay within the 256 MB cache), two slots.  Measured on the MI355X
// (u x omega, medians of interleaved runs): 512^3 FP64 5.26 -> 5.21 ms (8
// planes per slab), 1024^3 FP64 54.06 -> 53.58 ms (2 planes), 512^3 FP32
// 2.86 -> 2.82 ms (16 planes), 1024^3 FP32 30.0 -> 30.2 ms (8 planes; 4 and
// 2 planes slower still); on smaller grids the short per-slab launches cannot
// fill the device.  Three slots in flight, and one stream per pass with the
// slabs flowing through them under events, measured slower.  Since the passes
// store with the nontemporal hint (and the z pass writes five fields) the chain
// is slower at 512^3 too: 4.93 -> 5.06 ms FP64, 2.65 -> 2.71 ms FP32 (medians
// of interleaved runs).  What the chain saves is memory: the six full-y fields (6 Nz Ny nhp entries, 4.4 GB at
// 512^3 FP64, 35 GB at 1024^3) become two slots.  GBG_UXW_SLAB=Z,R at plan
// creation sets the planes per slab and the slots for such calls (tests and
// tuning); the results are the same bits for every setting.
constexpr size_t kSlabSlotBytes = (size_t)104 << 20;

static SlabCfg slab_cfg(const gbg_plan* p, const Geo& g, unsigned flags) {
  if (!(flags & GBG_UXW_SLABS)) return SlabCfg{0, 0};
  if (p->slab_set) return SlabCfg{p->slab_z, p->slab_r};
  const size_t plane = (size_t)9 * g.ny * g.nhp * 2 * p->elem;
  const size_t z = kSlabSlotBytes / plane;
  return SlabCfg{z < 1 ? 1 : (z < (size_t)g.nz ? (int)z : g.nz), 2};
}

// the chain's nst streams besides s and their events (false: run on s, as
// while s is being captured into a graph)
static bool slab_streams(gbg_plan* p, hipStream_t s, int nst) {
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  if (hipStreamIsCapturing(s, &cs) != hipSuccess) {
    (void)hipGetLastError();
    return false;
  }
  if (cs != hipStreamCaptureStatusNone) return false;
  if (!p->slab_fork && hipEventCreateWithFlags(&p->slab_fork, hipEventDisableTiming) != hipSuccess) {
    p->slab_fork = nullptr;
    (void)hipGetLastError();
    return false;
  }
  for (int i = 0; i < nst; ++i) {
    if (!p->slab_st[i] && hipStreamCreateWithFlags(&p->slab_st[i], hipStreamNonBlocking) != hipSuccess) {
      p->slab_st[i] = nullptr;
      (void)hipGetLastError();
      return false;
    }
    if (!p->slab_join[i] &&
        hipEventCreateWithFlags(&p->slab_join[i], hipEventDisableTiming) != hipSuccess) {
      p->slab_join[i] = nullptr;
      (void)hipGetLastError();
      return false;
    }
  }
  return true;
}

// Entries of one ring slot (NF + NO full-y fields of Z planes)
static long long slab_slot(const Geo& g, const SlabCfg& sc, int nf, int no) {
  return (long long)(nf + no) * sc.z * g.ny * g.nhp;
}

// The middle passes slab by slab: wc the compact buffer (NF fields in, the
// first NO fields out, field stride g.s1), ring the R slots.
template <typename T, int NF, int NO>
static gbg_status slab_chain(gbg_plan* p, const Geo& g, const SlabCfg& sc, cpx<T>* wc, cpx<T>* ring,
                             const Tables& tb, const cpx<T>* twx, const cpx<T>* twy, T scale,
                             bool noalloc, hipStream_t s) {
  constexpr bool PROD = NF == 2;
  const long long slot = slab_slot(g, sc, NF, NO);
  const int nsl = (g.nz + sc.z - 1) / sc.z;
  const int r = sc.r < nsl ? sc.r : nsl;
  gbg_status st = GBG_OK;
  bool multi = r > 1 && !noalloc && slab_streams(p, s, r - 1);
  if (multi) {
    if ((st = check_hip(hipEventRecord(p->slab_fork, s), "slab chain fork"))) return st;
    for (int i = 0; i + 1 < r; ++i)
      if ((st = check_hip(hipStreamWaitEvent(p->slab_st[i], p->slab_fork, 0), "slab chain fork")))
        return st;
  }
  for (int j = 0; j < nsl; ++j) {
    const int k = j % r;
    const hipStream_t q = (multi && k > 0) ? p->slab_st[k - 1] : s;
    const int z0 = j * sc.z, zn = sc.z < g.nz - z0 ? sc.z : g.nz - z0;
    Geo gs = g;
    gs.nz = zn;
    gs.s2 = (long long)zn * g.ny * g.nhp;
    cpx<T>* fin = ring + k * slot;      // NF full-y fields of the slab
    cpx<T>* fout = fin + NF * gs.s2;    // NO full-y fields of the slab
    cpx<T>* cz = wc + (long long)z0 * g.nyc * g.nhp;
    with_len(g.ny, [&](auto L) { launch_y<T, decltype(L)::value, 1, NF>(gs, cz, fin, tb, twy, q); });
    with_len(g.nx, [&](auto L) {
      launch_x<T, decltype(L)::value, PROD>(p->x_var, gs, fin, fout, twx, scale, q);
    });
    with_len(g.ny, [&](auto L) { launch_y<T, decltype(L)::value, -1, NO>(gs, fout, cz, tb, twy, q); });
    if ((st = check_launch("slab chain"))) return st;
  }
  if (multi) {
    for (int i = 0; i + 1 < r; ++i) {
      if ((st = check_hip(hipEventRecord(p->slab_join[i], p->slab_st[i]), "slab chain join")))
        return st;
      if ((st = check_hip(hipStreamWaitEvent(s, p->slab_join[i], 0), "slab chain join"))) return st;
    }
  }
  return GBG_OK;
}

// bytes of u x omega's two intermediate buffers (element size cb)
static void uxw_buffers(const Geo& g, const SlabCfg& sc, size_t cb, size_t* b1, size_t* b2) {
  if (sc.z > 0) {
    *b1 = (size_t)6 * g.s1 * cb;
    const int nsl = (g.nz + sc.z - 1) / sc.z;
    *b2 = (size_t)(sc.r < nsl ? sc.r : nsl) * slab_slot(g, sc, 6, 3) * cb;
    return;
  }
  *b1 = (size_t)(6 * g.s1 > 3 * g.s2 ? 6 * g.s1 : 3 * g.s2) * cb;
  *b2 = (size_t)6 * g.s2 * cb;
}

// base: uxw_work_bytes of work area.  With noalloc the pass schedule creates
// nothing (no tile set, no second stream): the last pass uses a tile set only
// when the plan already caches one for this cutoff.
template <typename T>
static gbg_status uxw_fused(gbg_plan* p, const cpx<T>* uh, cpx<T>* out, const Geo& g,
                            const SlabCfg& sc, double kc, char* base, bool noalloc, hipStream_t s,
                            const Epi<T>* ep = nullptr) {
  const size_t cb = sizeof(cpx<T>);
  // W1: K1's 6 compact-ky fields, later K3's 3 full-y fields;
  // W2: K2's 6 full-y fields, later K4's 3 compact-ky fields.
  // With the z-slab chain: W1 the compact fields (K4 writes its 3 into the
  // first three), W2 the ring of slots.
  size_t b1, b2;
  uxw_buffers(g, sc, cb, &b1, &b2);
  gbg_status st = GBG_OK;
  cpx<T>* W1 = (cpx<T>*)base;
  cpx<T>* W2 = (cpx<T>*)(base + b1);
  double* part = (double*)(base + b1 + b2);
  const Tables tb = tables_of(p);
  const cpx<T>* twx = (const cpx<T>*)p->tw[0];
  const cpx<T>* twy = (const cpx<T>*)p->tw[1];
  const cpx<T>* twz = (const cpx<T>*)p->tw[2];
  const double P = (double)g.nx * g.ny * g.nz;
  // with the combination the zero tiles write dst (the caller keeps dst off
  // u^); N^ itself may be absent
  cpx<T>* zout = ep ? ep->dst : out;
  const int fpass = noalloc ? 0 : zero_fork_pass<T>(ep != nullptr);
  const TileSet* fork = nullptr;
  bool forked = false;
  // fork before pass k is launched, start the zero writes after it
  auto fork_at = [&](int k) {
    if (fpass == k)
      fork = zero_tiles_fork<T>(p, g, kc, 3, zout, uh, 3 * g.M * sizeof(cpx<T>), nullptr, 0, s);
  };
  auto start_at = [&](int k) {
    if (fpass == k) forked = zero_tiles_launch<T>(p, g, fork, 3, out, ep, &tb);
  };
  prof_begin(p, s);
  fork_at(1);
  with_len(g.nz, [&](auto L) {
    constexpr int N = decltype(L)::value;
    const unsigned grid = launch_z_inverse<T, N>(p, g, uh, W1, tb, twz, part, s);
    hipLaunchKernelGGL(k_uxw_reduce, dim3(1), dim3(256), 0, s, part, (int)grid, P * P, p->red_out);
  });
  if ((st = check_launch("uxw K1"))) return st;
  start_at(1);
  prof_mark(p, s);
  // K4's compact output: W2 with the two full-size buffers, W1 with the chain
  cpx<T>* W4 = W2;
  if (sc.z > 0) {
    // one span for K2..K4 (a fork point in 2..4 forks before the chain)
    if (fpass >= 2 && fpass <= 4) {
      fork_at(fpass);
      start_at(fpass);
    }
    if ((st = slab_chain<T, 6, 3>(p, g, sc, W1, W2, tb, twx, twy, T(1.0 / P), noalloc, s)))
      return st;
    W4 = W1;
  } else {
    fork_at(2);
    with_len(g.ny, [&](auto L) {
      launch_y<T, decltype(L)::value, 1, 6>(g, W1, W2, tb, twy, s);
    });
    if ((st = check_launch("uxw K2"))) return st;
    start_at(2);
    prof_mark(p, s);
    fork_at(3);
    with_len(g.nx, [&](auto L) {
      launch_x<T, decltype(L)::value, false>(p->x_var, g, W2, W1, twx, T(1.0 / P), s);
    });
    if ((st = check_launch("uxw K3"))) return st;
    start_at(3);
    prof_mark(p, s);
    fork_at(4);
    with_len(g.ny, [&](auto L) {
      launch_y<T, decltype(L)::value, -1, 3>(g, W1, W2, tb, twy, s);
    });
    if ((st = check_launch("uxw K4"))) return st;
    start_at(4);
  }
  prof_mark(p, s);
  with_len(g.nz, [&](auto L) {
    constexpr int N = decltype(L)::value;
    using C = Cfg<T, N>;
    const long long plane = (long long)g.ny * g.nh;
    const TileSet* ts = tile_set(p, g, C::B, kc, s, !noalloc);
    if (!ts) {
      const long long nt = (plane + C::B - 1) / C::B;
      launch_z_forward<T, N, 3>(p, g, nt, W4, out, tb, twz, nullptr, s, ep);
      return;
    }
    if (!forked && ts->nd < ts->nt) launch_zero_tiles<T>(g, ts, C::B, 3, out, s, ep, &tb);
    if (ts->nd > 0)
      launch_z_forward<T, N, 3>(p, g, ts->nd, W4, out, tb, twz, ts->dev, s, ep);
  });
  if ((st = check_launch("uxw K5"))) return st;
  if (forked && (st = zero_tiles_join(p, s))) return st;
  prof_mark(p, s);
  return GBG_OK;
}

// the fused passes take x as the folded axis (the default layout)
static bool uxw_fusable(const gbg_plan* p, unsigned flags) {
  return !(flags & GBG_UXW_UNFUSED) && !p->generic_only && p->fold == 0 && fast_len(p->n[0]) &&
         fast_len(p->n[1]) && fast_len(p->n[2]);
}

size_t uxw_work_bytes(gbg_plan* p, double kc, unsigned flags) {
  const size_t cs = (size_t)p->elem * 2, rs = (size_t)p->elem;
  if (!uxw_fusable(p, flags)) return 2 * (3 * p->M * cs) + 2 * (3 * p->P * rs);
  const Geo g = make_geo(p, kc);
  // W1, W2 and the K1 partials (2 doubles per launched workgroup, at most
  // one per tile plus the XCD padding)
  size_t b1, b2;
  uxw_buffers(g, slab_cfg(p, g, flags), cs, &b1, &b2);
  return b1 + b2 + sizeof(double) * 2 * ((size_t)g.nyc * g.nhp + fast::kXcd);
}

template <typename T>
static Epi<T> make_epi(const UxwComb& c, bool store_n) {
  Epi<T> e;
  e.dst = (cpx<T>*)c.dst;
  e.a = (const cpx<T>*)c.a;
  e.b = (const cpx<T>*)c.b;
  e.alpha = T(c.alpha);
  e.beta = T(c.beta);
  e.delta = T(c.delta);
  e.gamma = T(c.gamma);
  e.coef = c.coef;
  e.store_n = store_n ? 1 : 0;
  return e;
}

// ============================================================================
// composed path
// ============================================================================
// w: 2 (3 M s_c) + 2 (3 P s) bytes of work area.  With a combination (cb) and
// no n_hat, N^ is formed in the first 3 M s_c of w (free by then) and the
// combination is applied to it in one more streaming pass.
static gbg_status uxw_composed(gbg_plan* p, const void* u_hat, void* n_hat, double kc, char* w,
                               hipStream_t s, const UxwComb* cb) {
  const size_t cs = (size_t)p->elem * 2;  // complex element bytes
  const size_t rs = (size_t)p->elem;      // real element bytes
  const size_t spec3 = 3 * p->M * cs, phys3 = 3 * p->P * rs;
  gbg_status st = GBG_OK;
  void* ut = w;
  void* wt = w + spec3;
  void* u = w + 2 * spec3;
  void* om = w + 2 * spec3 + phys3;
  prof_begin(p, s);
  if ((st = op_convolve_band(p, ut, u_hat, 3, 0.0, kc, 1.0, 0.0, s))) return st;
  if ((st = op_curl(p, wt, ut, 0, s))) return st;
  if ((st = op_inner(p, ut, ut, 3, 0, s))) return st;
  if ((st = op_inner(p, wt, wt, 3, 1, s))) return st;
  prof_mark(p, s);
  if ((st = inv_transform(p, u, ut, ut, 3, s))) return st;
  if ((st = inv_transform(p, om, wt, wt, 3, s))) return st;
  prof_mark(p, s);
  if ((st = op_cross(p, u, u, om, s))) return st;
  prof_mark(p, s);
  void* nt = n_hat ? n_hat : ut;
  if ((st = fwd_transform(p, nt, u, 3, s))) return st;
  prof_mark(p, s);
  if ((st = op_helmholtz(p, nt, nt, 0, 0.0, 0.0, 1.0, 0.0, s))) return st;
  if ((st = op_convolve_band(p, nt, nt, 3, 0.0, kc, 1.0, 0.0, s))) return st;
  if ((st = op_mean_free(p, nt, 3, s))) return st;
  if (cb) {
    const unsigned grid = (unsigned)((p->M + 255) / 256);
    if (p->elem == 8)
      hipLaunchKernelGGL(k_uxw_epilogue<double>, dim3(grid), dim3(256), 0, s,
                         (const cpx<double>*)nt, make_epi<double>(*cb, false), (long long)p->M,
                         p->s[0], p->s[1], tables_of(p));
    else
      hipLaunchKernelGGL(k_uxw_epilogue<float>, dim3(grid), dim3(256), 0, s,
                         (const cpx<float>*)nt, make_epi<float>(*cb, false), (long long)p->M,
                         p->s[0], p->s[1], tables_of(p));
    if ((st = check_launch("uxw combination"))) return st;
  }
  prof_mark(p, s);
  return GBG_OK;
}

static gbg_status uxw_run(gbg_plan* p, const void* u_hat, void* n_hat, double kc, unsigned flags,
                          char* work, bool noalloc, hipStream_t s, const UxwComb* cb = nullptr) {
  if (!uxw_fusable(p, flags)) return uxw_composed(p, u_hat, n_hat, kc, work, s, cb);
  const Geo g = make_geo(p, kc);
  const SlabCfg sc = slab_cfg(p, g, flags);
  if (p->elem == 8) {
    const Epi<double> e = cb ? make_epi<double>(*cb, n_hat != nullptr) : Epi<double>{};
    return uxw_fused<double>(p, (const cpx<double>*)u_hat, (cpx<double>*)n_hat, g, sc, kc, work,
                             noalloc, s, cb ? &e : nullptr);
  }
  const Epi<float> e = cb ? make_epi<float>(*cb, n_hat != nullptr) : Epi<float>{};
  return uxw_fused<float>(p, (const cpx<float>*)u_hat, (cpx<float>*)n_hat, g, sc, kc, work,
                          noalloc, s, cb ? &e : nullptr);
}

gbg_status pipe_uxw_comb(gbg_plan* p, const void* u_hat, void* n_hat, double kc, const UxwComb& cb,
                         unsigned flags, hipStream_t s) {
  gbg_status st = ensure_work(p, uxw_work_bytes(p, kc, flags));
  if (st) return st;
  return uxw_run(p, u_hat, n_hat, kc, flags, (char*)p->work, false, s, &cb);
}

gbg_status pipe_uxw(gbg_plan* p, const void* u_hat, void* n_hat, double kc, unsigned flags,
                    hipStream_t s) {
  gbg_status st = ensure_work(p, uxw_work_bytes(p, kc, flags));
  if (st) return st;
  return uxw_run(p, u_hat, n_hat, kc, flags, (char*)p->work, false, s);
}

gbg_status pipe_uxw_in(gbg_plan* p, const void* u_hat, void* n_hat, double kc, unsigned flags,
                       void* work, size_t work_bytes, hipStream_t s) {
  const size_t need = uxw_work_bytes(p, kc, flags);
  if (need > work_bytes)
    return fail(GBG_E_UNSUPPORTED, "u x omega on this grid and cutoff needs " + std::to_string(need) +
                                       " bytes of work area; the caller's area holds " +
                                       std::to_string(work_bytes));
  return uxw_run(p, u_hat, n_hat, kc, flags, (char*)work, true, s);
}

// Fused scalar product p^ = M F(F^-1(M a^) F^-1(M b^)), the five passes of
// u x omega with two fields in and one out (15 M s_c at the fused ideal):
//   P1 z-inverse of a^, b^ on the disk pencils -> W1 [2][Nz][nyc][nhp]
//   P2 y-inverse of 2 fields                    -> W2 [2][Nz][Ny][nhp]
//   P3 x: (a, b) of a row in one complex C2R, p = a b, p of two rows in one
//      R2C                                      -> W1 [1][Nz][Ny][nhp]
//   P4 y-forward of 1 field, disk ky            -> W2 [1][Nz][nyc][nhp]
//   P5 zero tiles + z-forward of 1 field, mask  -> p^ (full layout)
static size_t dp_work_bytes(const Geo& g, size_t cb) {
  const size_t b1 = (size_t)(2 * g.s1 > g.s2 ? 2 * g.s1 : g.s2) * cb;
  return b1 + (size_t)2 * g.s2 * cb;
}

template <typename T>
static gbg_status dp_fused(gbg_plan* p, const cpx<T>* ah, const cpx<T>* bh, cpx<T>* out,
                           const Geo& g, double kc, hipStream_t s) {
  const size_t cb = sizeof(cpx<T>);
  gbg_status st = ensure_work(p, dp_work_bytes(g, cb));
  if (st) return st;
  cpx<T>* W1 = (cpx<T>*)p->work;
  cpx<T>* W2 = (cpx<T>*)((char*)p->work + (size_t)(2 * g.s1 > g.s2 ? 2 * g.s1 : g.s2) * cb);
  const Tables tb = tables_of(p);
  const cpx<T>* twx = (const cpx<T>*)p->tw[0];
  const cpx<T>* twy = (const cpx<T>*)p->tw[1];
  const cpx<T>* twz = (const cpx<T>*)p->tw[2];
  const double P = (double)g.nx * g.ny * g.nz;
  const size_t cb1 = g.M * sizeof(cpx<T>);
  const TileSet* fork =
      sizeof(T) == 4 ? zero_tiles_fork<T>(p, g, kc, 1, out, ah, cb1, bh, cb1, s) : nullptr;
  prof_begin(p, s);
  with_len(g.nz, [&](auto L) {
    constexpr int N = decltype(L)::value;
    using C = Cfg<T, N>;
    const long long nt = ((long long)g.nyc * g.nhp + C::B - 1) / C::B;
    hipLaunchKernelGGL((k_dp_z_inverse<T, N>), dim3(fast::xcd_grid(nt)), dim3(C::NT), 0, s, ah, bh,
                       W1, g, tb, twz, nt);
  });
  if ((st = check_launch("product P1"))) return st;
  const bool forked = zero_tiles_launch<T>(p, g, fork, 1, out);
  prof_mark(p, s);
  with_len(g.ny, [&](auto L) {
    launch_y<T, decltype(L)::value, 1, 2>(g, W1, W2, tb, twy, s);
  });
  if ((st = check_launch("product P2"))) return st;
  prof_mark(p, s);
  with_len(g.nx, [&](auto L) {
    launch_x<T, decltype(L)::value, true>(p->x_var, g, W2, W1, twx, T(1.0 / P), s);
  });
  if ((st = check_launch("product P3"))) return st;
  prof_mark(p, s);
  with_len(g.ny, [&](auto L) {
    launch_y<T, decltype(L)::value, -1, 1>(g, W1, W2, tb, twy, s);
  });
  if ((st = check_launch("product P4"))) return st;
  prof_mark(p, s);
  with_len(g.nz, [&](auto L) {
    constexpr int N = decltype(L)::value;
    using C = Cfg<T, N>;
    const long long plane = (long long)g.ny * g.nh;
    const TileSet* ts = tile_set(p, g, C::B, kc, s);
    if (!ts) {
      const long long nt = (plane + C::B - 1) / C::B;
      launch_z_forward<T, N, 1>(p, g, nt, W2, out, tb, twz, nullptr, s);
      return;
    }
    if (!forked && ts->nd < ts->nt) launch_zero_tiles<T>(g, ts, C::B, 1, out, s);
    if (ts->nd > 0)
      launch_z_forward<T, N, 1>(p, g, ts->nd, W2, out, tb, twz, ts->dev, s);
  });
  if ((st = check_launch("product P5"))) return st;
  if (forked && (st = zero_tiles_join(p, s))) return st;
  prof_mark(p, s);
  return GBG_OK;
}

static bool fusable_grid(const gbg_plan* p) {
  return !p->generic_only && p->fold == 0 && fast_len(p->n[0]) && fast_len(p->n[1]) &&
         fast_len(p->n[2]);
}

size_t dealiased_product_work_bytes(gbg_plan* p, double kc) {
  const size_t cs = (size_t)p->elem * 2, rs = (size_t)p->elem;
  if (!fusable_grid(p)) return 2 * p->M * cs + 2 * p->P * rs;
  return dp_work_bytes(make_geo(p, kc), cs);
}

gbg_status pipe_dealiased_product(gbg_plan* p, void* p_hat, const void* a_hat, const void* b_hat,
                                  double kc, hipStream_t s) {
  if (fusable_grid(p)) {
    const Geo g = make_geo(p, kc);
    if (p->elem == 8)
      return dp_fused<double>(p, (const cpx<double>*)a_hat, (const cpx<double>*)b_hat,
                              (cpx<double>*)p_hat, g, kc, s);
    return dp_fused<float>(p, (const cpx<float>*)a_hat, (const cpx<float>*)b_hat,
                           (cpx<float>*)p_hat, g, kc, s);
  }
  // composed: the standalone operators and transforms (non-power-of-two axes)
  const size_t cs = (size_t)p->elem * 2, rs = (size_t)p->elem;
  const size_t spec1 = p->M * cs, phys1 = p->P * rs;
  gbg_status st = ensure_work(p, 2 * spec1 + 2 * phys1);
  if (st) return st;
  char* w = (char*)p->work;
  void* at = w;
  void* bt = w + spec1;
  void* a = w + 2 * spec1;
  void* b = w + 2 * spec1 + phys1;
  prof_begin(p, s);
  if ((st = op_convolve_band(p, at, a_hat, 1, 0.0, kc, 1.0, 0.0, s))) return st;
  if ((st = op_convolve_band(p, bt, b_hat, 1, 0.0, kc, 1.0, 0.0, s))) return st;
  if ((st = inv_transform(p, a, at, at, 1, s))) return st;
  if ((st = inv_transform(p, b, bt, bt, 1, s))) return st;
  if ((st = op_multiply(p, a, a, b, s))) return st;
  if ((st = fwd_transform(p, p_hat, a, 1, s))) return st;
  if ((st = op_convolve_band(p, p_hat, p_hat, 1, 0.0, kc, 1.0, 0.0, s))) return st;
  prof_mark(p, s);
  return GBG_OK;
}

}  // namespace gbg
