// Pencil FFT passes: the 3-D real-to-complex transform of periodic fields.
//
// forward (to_spectral!):  x-pass R2C (phys -> spec), y-pass C2C, z-pass C2C (in place)
// inverse (to_physical!):  z-pass C2C, y-pass C2C, x-pass C2R (1/N applied here)
//
// Strided passes load B consecutive pencils of the fast index into LDS, so
// every global row segment is B contiguous complex elements; the x passes
// pair two real rows into one complex transform (z = a + i b).  Tiles map to
// workgroups XCD-aware (fast::xcd_tile): with the odd row pitch N/2 + 1 a
// tile's segments straddle cache lines, and the neighbouring tile that shares
// them runs on the same XCD, behind the same L2.
#include "fft_fast.hpp"

namespace gbg {

// Strided pencil addressing: pencil p, element e along the axis, component c
//   offset = c*cstride + (p / inner)*ostride + (p % inner) + e*se
struct Strided {
  long long npencil, inner, ostride, se, cstride;
  int B, bshift, ld;
};

template <typename T, int DIR>
__global__ __launch_bounds__(kNT) void k_c2c_strided(const cpx<T>* in, cpx<T>* out, Strided d,
                                                     AxisPlan ap, const cpx<T>* __restrict__ tw,
                                                     T scale) {
  __shared__ cpx<T> s[2 * Lds<T>::kElems];
  const long long coff = (long long)blockIdx.y * d.cstride;
  const int pl = threadIdx.x & (d.B - 1);
  const int e0 = threadIdx.x >> d.bshift;
  const int estep = kNT >> d.bshift;
  const long long p = fast::xcd_tile() * d.B + pl;
  const bool valid = p < d.npencil;
  const long long base = coff + (valid ? (p / d.inner) * d.ostride + (p % d.inner) : 0);
  const int n = ap.n;
  for (int e = e0; e < n; e += estep) {
    cpx<T> v = mk<T>(0, 0);
    if (valid) v = in[base + (long long)e * d.se];
    s[pl * d.ld + e] = v;
  }
  __syncthreads();
  const cpx<T>* r = fft_any<T, DIR>(s, s + Lds<T>::kElems, d.ld, ap, d.B, tw);
  if (valid) {
    for (int e = e0; e < n; e += estep) out[base + (long long)e * d.se] = cscale(r[pl * d.ld + e], scale);
  }
}

// x-pass R2C: B row pairs per workgroup.  Row r of component data is at
// r*nx (physical) / r*nh (spectral); rows of all components are contiguous.
template <typename T>
__global__ __launch_bounds__(kNT) void k_r2c_x(const T* __restrict__ in, cpx<T>* __restrict__ out,
                                               long long nrows, int nx, int nh, int B, int ld,
                                               AxisPlan ap, const cpx<T>* __restrict__ tw) {
  __shared__ cpx<T> s[2 * Lds<T>::kElems];
  const long long pair0 = fast::xcd_tile() * B;
  for (int idx = threadIdx.x; idx < B * nx; idx += kNT) {
    const int q = idx / nx, x = idx - q * nx;
    const long long r0 = 2 * (pair0 + q), r1 = r0 + 1;
    const T a = r0 < nrows ? in[r0 * nx + x] : T(0);
    const T b = r1 < nrows ? in[r1 * nx + x] : T(0);
    s[q * ld + x] = mk<T>(a, b);
  }
  __syncthreads();
  const cpx<T>* rs = fft_any<T, -1>(s, s + Lds<T>::kElems, ld, ap, B, tw);
  for (int idx = threadIdx.x; idx < B * nh; idx += kNT) {
    const int q = idx / nh, k = idx - q * nh;
    const long long r0 = 2 * (pair0 + q), r1 = r0 + 1;
    if (r0 >= nrows) continue;
    const cpx<T> z = rs[q * ld + k];
    const cpx<T> zc = rs[q * ld + (k == 0 ? 0 : nx - k)];
    // A = (Z_k + conj Z_{n-k}) / 2,  B = (Z_k - conj Z_{n-k}) / (2i)
    out[r0 * nh + k] = mk<T>(T(0.5) * (z.x + zc.x), T(0.5) * (z.y - zc.y));
    if (r1 < nrows) out[r1 * nh + k] = mk<T>(T(0.5) * (z.y + zc.y), T(0.5) * (zc.x - z.x));
  }
}

// x-pass C2R: Z_k = A_k + i B_k over the Hermitian completion, with the
// self-conjugate entries (k = 0 and the even-axis Nyquist) taken real, as a
// c2r transform treats them.  Output scaled by `scale` (1/N of the 3-D
// inverse).
template <typename T>
__global__ __launch_bounds__(kNT) void k_c2r_x(const cpx<T>* __restrict__ in, T* __restrict__ out,
                                               long long nrows, int nx, int nh, int B, int ld,
                                               AxisPlan ap, const cpx<T>* __restrict__ tw, T scale) {
  __shared__ cpx<T> s[2 * Lds<T>::kElems];
  const long long pair0 = fast::xcd_tile() * B;
  const int nyq = (nx % 2 == 0) ? nx / 2 : -1;
  for (int idx = threadIdx.x; idx < B * nx; idx += kNT) {
    const int q = idx / nx, k = idx - q * nx;
    const long long r0 = 2 * (pair0 + q), r1 = r0 + 1;
    const bool upper = k >= nh;
    const int kk = upper ? nx - k : k;
    cpx<T> a = r0 < nrows ? in[r0 * nh + kk] : mk<T>(0, 0);
    cpx<T> b = r1 < nrows ? in[r1 * nh + kk] : mk<T>(0, 0);
    if (kk == 0 || kk == nyq) {
      a.y = T(0);
      b.y = T(0);
    }
    if (upper) {
      a.y = -a.y;
      b.y = -b.y;
    }
    s[q * ld + k] = mk<T>(a.x - b.y, a.y + b.x);
  }
  __syncthreads();
  const cpx<T>* rs = fft_any<T, 1>(s, s + Lds<T>::kElems, ld, ap, B, tw);
  for (int idx = threadIdx.x; idx < B * nx; idx += kNT) {
    const int q = idx / nx, x = idx - q * nx;
    const long long r0 = 2 * (pair0 + q), r1 = r0 + 1;
    if (r0 >= nrows) continue;
    const cpx<T> z = rs[q * ld + x];
    out[r0 * nx + x] = z.x * scale;
    if (r1 < nrows) out[r1 * nx + x] = z.y * scale;
  }
}

// R2C along a strided folded axis (y or z of a permuted layout): pencil pairs
// (2q, 2q + 1) of the pencil enumeration are transformed as z = a + i b, then
// split into the two half spectra as in k_r2c_x.  `pi` addresses the real
// input, `po` the complex output (same pencil enumeration, their own strides).
template <typename T>
__global__ __launch_bounds__(kNT) void k_r2c_strided(const T* __restrict__ in, cpx<T>* __restrict__ out,
                                                     Strided pi, Strided po, int n, int nh, int B,
                                                     int ld, AxisPlan ap,
                                                     const cpx<T>* __restrict__ tw) {
  __shared__ cpx<T> s[2 * Lds<T>::kElems];
  const long long pair0 = fast::xcd_tile() * B;
  const long long ci = (long long)blockIdx.y * pi.cstride, co = (long long)blockIdx.y * po.cstride;
  auto off = [](const Strided& d, long long p) { return (p / d.inner) * d.ostride + (p % d.inner); };
  for (int idx = threadIdx.x; idx < B * n; idx += kNT) {
    const int q = idx / n, e = idx - q * n;
    const long long p0 = 2 * (pair0 + q), p1 = p0 + 1;
    const T a = p0 < pi.npencil ? in[ci + off(pi, p0) + e * pi.se] : T(0);
    const T b = p1 < pi.npencil ? in[ci + off(pi, p1) + e * pi.se] : T(0);
    s[q * ld + e] = mk<T>(a, b);
  }
  __syncthreads();
  const cpx<T>* rs = fft_any<T, -1>(s, s + Lds<T>::kElems, ld, ap, B, tw);
  for (int idx = threadIdx.x; idx < B * nh; idx += kNT) {
    const int q = idx / nh, k = idx - q * nh;
    const long long p0 = 2 * (pair0 + q), p1 = p0 + 1;
    if (p0 >= po.npencil) continue;
    const cpx<T> z = rs[q * ld + k];
    const cpx<T> zc = rs[q * ld + (k == 0 ? 0 : n - k)];
    out[co + off(po, p0) + k * po.se] = mk<T>(T(0.5) * (z.x + zc.x), T(0.5) * (z.y - zc.y));
    if (p1 < po.npencil)
      out[co + off(po, p1) + k * po.se] = mk<T>(T(0.5) * (z.y + zc.y), T(0.5) * (zc.x - z.x));
  }
}

// C2R along a strided folded axis: the Hermitian completion of two pencils'
// half spectra (self-conjugate entries taken real) as Z = A + i B, inverse
// transform, real part to the first pencil and imaginary part to the second,
// scaled by `scale`.
template <typename T>
__global__ __launch_bounds__(kNT) void k_c2r_strided(const cpx<T>* __restrict__ in, T* __restrict__ out,
                                                     Strided pi, Strided po, int n, int nh, int B,
                                                     int ld, AxisPlan ap,
                                                     const cpx<T>* __restrict__ tw, T scale) {
  __shared__ cpx<T> s[2 * Lds<T>::kElems];
  const long long pair0 = fast::xcd_tile() * B;
  const long long ci = (long long)blockIdx.y * pi.cstride, co = (long long)blockIdx.y * po.cstride;
  auto off = [](const Strided& d, long long p) { return (p / d.inner) * d.ostride + (p % d.inner); };
  const int nyq = (n % 2 == 0) ? n / 2 : -1;
  for (int idx = threadIdx.x; idx < B * n; idx += kNT) {
    const int q = idx / n, k = idx - q * n;
    const long long p0 = 2 * (pair0 + q), p1 = p0 + 1;
    const bool upper = k >= nh;
    const int kk = upper ? n - k : k;
    cpx<T> a = p0 < pi.npencil ? in[ci + off(pi, p0) + kk * pi.se] : mk<T>(0, 0);
    cpx<T> b = p1 < pi.npencil ? in[ci + off(pi, p1) + kk * pi.se] : mk<T>(0, 0);
    if (kk == 0 || kk == nyq) {
      a.y = T(0);
      b.y = T(0);
    }
    if (upper) {
      a.y = -a.y;
      b.y = -b.y;
    }
    s[q * ld + k] = mk<T>(a.x - b.y, a.y + b.x);
  }
  __syncthreads();
  const cpx<T>* rs = fft_any<T, 1>(s, s + Lds<T>::kElems, ld, ap, B, tw);
  for (int idx = threadIdx.x; idx < B * n; idx += kNT) {
    const int q = idx / n, e = idx - q * n;
    const long long p0 = 2 * (pair0 + q), p1 = p0 + 1;
    if (p0 >= po.npencil) continue;
    const cpx<T> z = rs[q * ld + e];
    out[co + off(po, p0) + e * po.se] = z.x * scale;
    if (p1 < po.npencil) out[co + off(po, p1) + e * po.se] = z.y * scale;
  }
}

// ----------------------------------------------------------------------------
// register-resident passes for power-of-two lengths (fft_fast.hpp)
// ----------------------------------------------------------------------------
using fast::Cfg;

// strided C2C pass: B pencils per workgroup, thread (pl, t) holds elements
// t + TP*m of pencil pl; consecutive threads read consecutive pencils, so
// every global access is a B-element contiguous segment.  In place safe:
// every load of the tile precedes the first barrier, every store follows
// the last one.
template <typename T, int N, int DIR>
__global__ __launch_bounds__((Cfg<T, N>::NT)) void k_fast_c2c(const cpx<T>* in, cpx<T>* out,
                                                            Strided d, const cpx<T>* __restrict__ tw,
                                                            T scale) {
  using C = Cfg<T, N>;
  constexpr int E = C::E, TP = C::TP, B = C::B, LD = C::LD;
  __shared__ cpx<T> lds[C::NBUF * B * LD];
  const int pl = threadIdx.x % B, t = threadIdx.x / B;
  const long long p = fast::xcd_tile() * B + pl;
  const bool valid = p < d.npencil;
  const long long base =
      (long long)blockIdx.y * d.cstride + (valid ? (p / d.inner) * d.ostride + (p % d.inner) : 0);
  cpx<T> v[E];
#pragma unroll
  for (int m = 0; m < E; ++m)
    v[m] = valid ? in[base + (long long)(t + TP * m) * d.se] : mk<T>(T(0), T(0));
  fast::fft<T, N, E, DIR, 0, C::PD, C::ONE>(v, lds + pl * LD,
                                            lds + (C::ONE ? pl : B + pl) * LD, t, tw);
  if (valid) {
#pragma unroll
    for (int m = 0; m < E; ++m) out[base + (long long)(t + TP * m) * d.se] = cscale(v[m], scale);
  }
}

// Persistent form of the strided pass, in FP32 for pencil counts that are a
// multiple of B (every lane of every tile holds a pencil) and 128 <= N <= 512
// (512^3 vector transform 2.07 -> 2.01 ms; in FP64 it measured 3 % slower): one
// workgroup per CU walks its XCD's contiguous share of the (component, tile)
// items, and the next item's pencils are loaded while the current item is
// transformed.  The two register sets X and Y take alternate items in a loop
// written out twice, so no register copy waits for a load in flight; every
// load and store is unconditional (the item past the end reloads the current
// one); the twiddles are read from an LDS copy, since a global load inside the
// transform would wait for the prefetched pencils too (the counter of
// outstanding vector-memory operations drains in order).  In place safe: the
// items are disjoint, and a tile's loads precede its stores.
template <typename T, int N>
constexpr bool c2c_persist() {
  return sizeof(T) == 4 && N >= 128 && N <= 512;
}

template <typename T, int N, int DIR>
__global__ __launch_bounds__((Cfg<T, N>::NT)) void k_fast_c2c_walk(const cpx<T>* in, cpx<T>* out,
                                                                 Strided d,
                                                                 const cpx<T>* __restrict__ tw,
                                                                 T scale, long long ntiles,
                                                                 long long nitems) {
  using C = Cfg<T, N>;
  constexpr int E = C::E, TP = C::TP, B = C::B, LD = C::LD;
  __shared__ cpx<T> lds[C::NBUF * B * LD];
  __shared__ cpx<T> twl[N];
  for (int i = threadIdx.x; i < N; i += C::NT) twl[i] = tw[i];
  const int pl = threadIdx.x % B, t = threadIdx.x / B;
  cpx<T>* pen = lds + pl * LD;
  cpx<T>* pen1 = lds + (C::ONE ? pl : B + pl) * LD;
  auto base_of = [&](long long item) {
    const long long c = item / ntiles;
    const long long p = (item - c * ntiles) * B + pl;
    return c * d.cstride + (p / d.inner) * d.ostride + (p % d.inner);
  };
  auto load = [&](cpx<T> (&v)[E], long long b) {
#pragma unroll
    for (int m = 0; m < E; ++m) v[m] = in[b + (long long)(t + TP * m) * d.se];
  };
  auto run = [&](cpx<T> (&v)[E], long long b) {
    fast::fft<T, N, E, DIR, 0, C::PD, C::ONE>(v, pen, pen1, t, (const cpx<T>*)twl);
#pragma unroll
    for (int m = 0; m < E; ++m) out[b + (long long)(t + TP * m) * d.se] = cscale(v[m], scale);
    asm volatile("" ::: "memory");
    __builtin_amdgcn_sched_barrier(0);
  };
  const long long per = (nitems + fast::kXcd - 1) / fast::kXcd;
  const int xcd = blockIdx.x % fast::kXcd;
  const long long L = gridDim.x / fast::kXcd;
  const long long end = (xcd + 1) * per < nitems ? (xcd + 1) * per : nitems;
  long long item = xcd * per + blockIdx.x / fast::kXcd;
  if (item >= end) return;
  cpx<T> X[E], Y[E];
  long long bx = base_of(item);
  load(X, bx);
#pragma unroll 1
  while (true) {
    const long long n1 = item + L;
    const long long by = base_of(n1 < end ? n1 : item);
    load(Y, by);
    run(X, bx);
    if (n1 >= end) break;
    item = n1;
    const long long n2 = item + L;
    bx = base_of(n2 < end ? n2 : item);
    load(X, bx);
    run(Y, by);
    if (n2 >= end) break;
    item = n2;
  }
}

struct RowG {
  int nh, nhk, nyq_x;
};

// x-pass R2C over row pairs (z = a + i b), one pair per TP threads
template <typename T, int N>
__global__ __launch_bounds__((Cfg<T, N>::XNT)) void k_fast_r2c_x(const T* __restrict__ in,
                                                               cpx<T>* __restrict__ out,
                                                               long long nrows, RowG g,
                                                               const cpx<T>* __restrict__ tw) {
  using C = Cfg<T, N>;
  constexpr int E = C::XE, TP = C::XTP, LD = C::XLD;
  __shared__ cpx<T> lds[2 * C::U * LD];
  const int un = threadIdx.x / TP, t = threadIdx.x % TP;
  const long long r0 = 2 * (fast::xcd_tile() * C::U + un), r1 = r0 + 1;
  const bool ok0 = r0 < nrows, ok1 = r1 < nrows;
  cpx<T> v[E];
#pragma unroll
  for (int m = 0; m < E; ++m) {
    const int x = t + TP * m;
    v[m] = mk<T>(ok0 ? in[r0 * N + x] : T(0), ok1 ? in[r1 * N + x] : T(0));
  }
  cpx<T>* pen = lds + un * LD;
  cpx<T>* pen1 = lds + (C::U + un) * LD;
  fast::fft<T, N, E, -1, C::XSYNC, C::XPD>(v, pen, pen1, t, tw);
  fast::r2c_pair_store<T, N, E, RowG, false>(v, pen, out + (ok0 ? r0 * g.nh : 0), out + (ok1 ? r1 * g.nh : 0),
                                ok0, ok1, t, g);
}

// x-pass C2R over row pairs; output scaled by `scale`
template <typename T, int N>
__global__ __launch_bounds__((Cfg<T, N>::XNT)) void k_fast_c2r_x(const cpx<T>* __restrict__ in,
                                                               T* __restrict__ out, long long nrows,
                                                               RowG g, const cpx<T>* __restrict__ tw,
                                                               T scale) {
  using C = Cfg<T, N>;
  constexpr int E = C::XE, TP = C::XTP, LD = C::XLD;
  __shared__ cpx<T> lds[2 * C::U * LD];
  const int un = threadIdx.x / TP, t = threadIdx.x % TP;
  const long long r0 = 2 * (fast::xcd_tile() * C::U + un), r1 = r0 + 1;
  const bool ok0 = r0 < nrows, ok1 = r1 < nrows;
  cpx<T> v[E];
  fast::c2r_pair_load<T, N, E>(v, in + (ok0 ? r0 * g.nh : 0), in + (ok1 ? r1 * g.nh : 0), ok0, ok1,
                               t, g);
  fast::fft<T, N, E, 1, C::XSYNC, C::XPD>(v, lds + un * LD, lds + (C::U + un) * LD, t, tw);
#pragma unroll
  for (int m = 0; m < E; ++m) {
    const int x = t + TP * m;
    if (ok0) out[r0 * N + x] = v[m].x * scale;
    if (ok1) out[r1 * N + x] = v[m].y * scale;
  }
}

static bool fast_len(int n) {
  return n == 32 || n == 64 || n == 128 || n == 256 || n == 512 || n == 1024;
}

template <typename T, int N>
static void launch_fast_c2c(int dir, const cpx<T>* in, cpx<T>* out, const Strided& d, int ncomp,
                            const cpx<T>* tw, T scale, int ncu, hipStream_t s) {
  using C = Cfg<T, N>;
  if constexpr (c2c_persist<T, N>()) {
    const unsigned cap = (unsigned)(ncu / fast::kXcd * fast::kXcd);
    if (d.npencil % C::B == 0 && cap >= (unsigned)fast::kXcd) {
      const long long ntiles = d.npencil / C::B, nitems = ntiles * ncomp;
      unsigned g = fast::xcd_grid(nitems);
      if (g > cap) g = cap;
      if (dir < 0)
        hipLaunchKernelGGL((k_fast_c2c_walk<T, N, -1>), dim3(g), dim3(C::NT), 0, s, in, out, d, tw,
                           scale, ntiles, nitems);
      else
        hipLaunchKernelGGL((k_fast_c2c_walk<T, N, 1>), dim3(g), dim3(C::NT), 0, s, in, out, d, tw,
                           scale, ntiles, nitems);
      return;
    }
  }
  dim3 grid(fast::xcd_grid((d.npencil + C::B - 1) / C::B), (unsigned)ncomp);
  if (dir < 0)
    hipLaunchKernelGGL((k_fast_c2c<T, N, -1>), grid, dim3(C::NT), 0, s, in, out, d, tw, scale);
  else
    hipLaunchKernelGGL((k_fast_c2c<T, N, 1>), grid, dim3(C::NT), 0, s, in, out, d, tw, scale);
}

template <typename T>
static bool try_fast_c2c(int n, int dir, const cpx<T>* in, cpx<T>* out, const Strided& d, int ncomp,
                         const cpx<T>* tw, T scale, int ncu, hipStream_t s) {
  switch (n) {
    case 32: launch_fast_c2c<T, 32>(dir, in, out, d, ncomp, tw, scale, ncu, s); return true;
    case 64: launch_fast_c2c<T, 64>(dir, in, out, d, ncomp, tw, scale, ncu, s); return true;
    case 128: launch_fast_c2c<T, 128>(dir, in, out, d, ncomp, tw, scale, ncu, s); return true;
    case 256: launch_fast_c2c<T, 256>(dir, in, out, d, ncomp, tw, scale, ncu, s); return true;
    case 512: launch_fast_c2c<T, 512>(dir, in, out, d, ncomp, tw, scale, ncu, s); return true;
    case 1024: launch_fast_c2c<T, 1024>(dir, in, out, d, ncomp, tw, scale, ncu, s); return true;
  }
  return false;
}

template <typename T, int N>
static void launch_fast_x(int dir, const void* in, void* out, long long nrows, RowG g,
                          const cpx<T>* tw, T scale, hipStream_t s) {
  using C = Cfg<T, N>;
  const long long pairs = (nrows + 1) / 2;
  dim3 grid(fast::xcd_grid((pairs + C::U - 1) / C::U));
  if (dir < 0)
    hipLaunchKernelGGL((k_fast_r2c_x<T, N>), grid, dim3(C::XNT), 0, s, (const T*)in, (cpx<T>*)out,
                       nrows, g, tw);
  else
    hipLaunchKernelGGL((k_fast_c2r_x<T, N>), grid, dim3(C::XNT), 0, s, (const cpx<T>*)in, (T*)out,
                       nrows, g, tw, scale);
}

template <typename T>
static bool try_fast_x(int n, int dir, const void* in, void* out, long long nrows, RowG g,
                       const cpx<T>* tw, T scale, hipStream_t s) {
  switch (n) {
    case 32: launch_fast_x<T, 32>(dir, in, out, nrows, g, tw, scale, s); return true;
    case 64: launch_fast_x<T, 64>(dir, in, out, nrows, g, tw, scale, s); return true;
    case 128: launch_fast_x<T, 128>(dir, in, out, nrows, g, tw, scale, s); return true;
    case 256: launch_fast_x<T, 256>(dir, in, out, nrows, g, tw, scale, s); return true;
    case 512: launch_fast_x<T, 512>(dir, in, out, nrows, g, tw, scale, s); return true;
    case 1024: launch_fast_x<T, 1024>(dir, in, out, nrows, g, tw, scale, s); return true;
  }
  return false;
}

// ----------------------------------------------------------------------------
// host launchers
// ----------------------------------------------------------------------------
static int ilog2(int b) {
  int s = 0;
  while ((1 << s) < b) ++s;
  return s;
}

// pencils along `axis` of a column-major array of extents e[3]: pencil p,
// element k at (p / inner) ostride + (p mod inner) + k se
static Strided pencils(const long long e[3], int axis, long long cstride) {
  Strided d;
  if (axis == 0) {        // x: pencils over (y, z)
    d.npencil = e[1] * e[2];
    d.inner = 1;
    d.ostride = e[0];
    d.se = 1;
  } else if (axis == 1) { // y: pencils over (z, x)
    d.npencil = e[2] * e[0];
    d.inner = e[0];
    d.ostride = e[0] * e[1];
    d.se = e[0];
  } else {                // z: pencils over (y, x)
    d.npencil = e[1] * e[0];
    d.inner = e[1] * e[0];
    d.ostride = 0;
    d.se = e[1] * e[0];
  }
  d.cstride = cstride;
  d.B = 1;
  d.bshift = 0;
  d.ld = 0;
  return d;
}

template <typename T>
static gbg_status launch_strided(gbg_plan* p, int axis, int dir, const cpx<T>* in, cpx<T>* out,
                                 int ncomp, T scale, hipStream_t s) {
  const AxisPlan& ap = p->ax[axis];
  if (ap.n == 1 && in == out) return GBG_OK;
  const long long ext[3] = {p->s[0], p->s[1], p->s[2]};
  Strided d = pencils(ext, axis, (long long)p->M);
  d.B = p->bsel[axis] ? p->bsel[axis] : pick_b(pencil_len(ap), Lds<T>::kCap);
  d.bshift = ilog2(d.B);
  d.ld = pencil_len(ap) + 1;
  const cpx<T>* tw = reinterpret_cast<const cpx<T>*>(p->tw[axis]);
  if (!p->generic_only && p->use_fast[axis] &&
      try_fast_c2c<T>(ap.n, dir, in, out, d, ncomp, tw, scale, p->ncu, s))
    return check_launch("c2c pass");
  dim3 grid(fast::xcd_grid((d.npencil + d.B - 1) / d.B), (unsigned)ncomp);
  if (dir < 0)
    hipLaunchKernelGGL((k_c2c_strided<T, -1>), grid, dim3(kNT), 0, s, in, out, d, ap, tw, scale);
  else
    hipLaunchKernelGGL((k_c2c_strided<T, 1>), grid, dim3(kNT), 0, s, in, out, d, ap, tw, scale);
  return check_launch("c2c pass");
}

// the real <-> half-spectrum pass along a strided folded axis (y or z)
template <typename T>
static gbg_status launch_fold_pass(gbg_plan* p, int dir, const void* in, void* out, int ncomp,
                                   T scale, hipStream_t s) {
  const int f = p->fold;
  const AxisPlan& ap = p->ax[f];
  const long long pe[3] = {p->n[0], p->n[1], p->n[2]};
  const long long se[3] = {p->s[0], p->s[1], p->s[2]};
  const Strided dp = pencils(pe, f, (long long)p->P), ds = pencils(se, f, (long long)p->M);
  const int len = pencil_len(ap);
  const int B = p->bsel[f] ? p->bsel[f] : pick_b(len, Lds<T>::kCap);
  const long long pairs = (dp.npencil + 1) / 2;
  dim3 grid(fast::xcd_grid((pairs + B - 1) / B), (unsigned)ncomp);
  const cpx<T>* tw = reinterpret_cast<const cpx<T>*>(p->tw[f]);
  if (dir < 0)
    hipLaunchKernelGGL((k_r2c_strided<T>), grid, dim3(kNT), 0, s, (const T*)in, (cpx<T>*)out, dp, ds,
                       p->n[f], p->s[f], B, len + 1, ap, tw);
  else
    hipLaunchKernelGGL((k_c2r_strided<T>), grid, dim3(kNT), 0, s, (const cpx<T>*)in, (T*)out, ds, dp,
                       p->n[f], p->s[f], B, len + 1, ap, tw, scale);
  return check_launch(dir < 0 ? "r2c folded pass" : "c2r folded pass");
}

// a permuted layout: R2C along the folded axis, then C2C along the other two
template <typename T>
static gbg_status fwd_folded(gbg_plan* p, cpx<T>* spec, const T* phys, int ncomp, hipStream_t s) {
  gbg_status st = launch_fold_pass<T>(p, -1, phys, spec, ncomp, T(1), s);
  for (int a = 0; a < 3 && !st; ++a)
    if (a != p->fold) st = launch_strided<T>(p, a, -1, spec, spec, ncomp, T(1), s);
  return st;
}

// its inverse: C2C along the two full axes (the first out of place into the
// work buffer), then C2R along the folded axis carrying 1/N
template <typename T>
static gbg_status inv_folded(gbg_plan* p, T* phys, cpx<T>* spec_in, cpx<T>* work, int ncomp,
                             hipStream_t s) {
  gbg_status st = GBG_OK;
  bool first = true;
  for (int a = 2; a >= 0 && !st; --a) {
    if (a == p->fold) continue;
    st = launch_strided<T>(p, a, 1, first ? spec_in : work, work, ncomp, T(1), s);
    first = false;
  }
  if (st) return st;
  const T scale = T(1.0 / ((double)p->n[0] * (double)p->n[1] * (double)p->n[2]));
  return launch_fold_pass<T>(p, 1, work, phys, ncomp, scale, s);
}

template <typename T>
static gbg_status fwd_impl(gbg_plan* p, cpx<T>* spec, const T* phys, int ncomp, hipStream_t s) {
  if (p->fold != 0) return fwd_folded<T>(p, spec, phys, ncomp, s);
  const int nx = p->n[0], nh = p->nh;
  const long long nrows = (long long)ncomp * p->n[1] * p->n[2];
  const cpx<T>* twx = reinterpret_cast<const cpx<T>*>(p->tw[0]);
  const RowG rg{nh, nh, nx % 2 == 0 ? nh - 1 : -1};
  if (p->generic_only || !p->use_fast[0] ||
      !try_fast_x<T>(nx, -1, phys, spec, nrows, rg, twx, T(1), s)) {
    const int len = pencil_len(p->ax[0]);
    const int B = p->bsel[0] ? p->bsel[0] : pick_b(len, Lds<T>::kCap);
    const long long pairs = (nrows + 1) / 2;
    dim3 grid(fast::xcd_grid((pairs + B - 1) / B));
    hipLaunchKernelGGL((k_r2c_x<T>), grid, dim3(kNT), 0, s, phys, spec, nrows, nx, nh, B, len + 1,
                       p->ax[0], twx);
  }
  gbg_status st = check_launch("r2c x pass");
  if (st) return st;
  if ((st = launch_strided<T>(p, 1, -1, spec, spec, ncomp, T(1), s))) return st;
  return launch_strided<T>(p, 2, -1, spec, spec, ncomp, T(1), s);
}

template <typename T>
static gbg_status inv_impl(gbg_plan* p, T* phys, cpx<T>* spec_in, cpx<T>* work, int ncomp,
                           hipStream_t s) {
  if (p->fold != 0) return inv_folded<T>(p, phys, spec_in, work, ncomp, s);
  gbg_status st;
  if ((st = launch_strided<T>(p, 2, 1, spec_in, work, ncomp, T(1), s))) return st;
  if ((st = launch_strided<T>(p, 1, 1, work, work, ncomp, T(1), s))) return st;
  const int nx = p->n[0], nh = p->nh;
  const long long nrows = (long long)ncomp * p->n[1] * p->n[2];
  const T scale = T(1.0 / ((double)p->n[0] * (double)p->n[1] * (double)p->n[2]));
  const cpx<T>* twx = reinterpret_cast<const cpx<T>*>(p->tw[0]);
  const RowG rg{nh, nh, nx % 2 == 0 ? nh - 1 : -1};
  if (p->generic_only || !p->use_fast[0] ||
      !try_fast_x<T>(nx, 1, work, phys, nrows, rg, twx, scale, s)) {
    const int len = pencil_len(p->ax[0]);
    const int B = p->bsel[0] ? p->bsel[0] : pick_b(len, Lds<T>::kCap);
    const long long pairs = (nrows + 1) / 2;
    dim3 grid(fast::xcd_grid((pairs + B - 1) / B));
    hipLaunchKernelGGL((k_c2r_x<T>), grid, dim3(kNT), 0, s, work, phys, nrows, nx, nh, B, len + 1,
                       p->ax[0], twx, scale);
  }
  return check_launch("c2r x pass");
}

gbg_status fwd_transform(gbg_plan* p, void* spec, const void* phys, int ncomp, hipStream_t s) {
  if (p->elem == 8)
    return fwd_impl<double>(p, (cpx<double>*)spec, (const double*)phys, ncomp, s);
  return fwd_impl<float>(p, (cpx<float>*)spec, (const float*)phys, ncomp, s);
}

gbg_status inv_transform(gbg_plan* p, void* phys, void* spec_in, void* spec_work, int ncomp,
                         hipStream_t s) {
  if (p->elem == 8)
    return inv_impl<double>(p, (double*)phys, (cpx<double>*)spec_in, (cpx<double>*)spec_work,
                            ncomp, s);
  return inv_impl<float>(p, (float*)phys, (cpx<float>*)spec_in, (cpx<float>*)spec_work, ncomp, s);
}

}  // namespace gbg
