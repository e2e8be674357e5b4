// Device-side mixed-radix Stockham FFT over pencils held in LDS.
//
// A workgroup holds B pencils of length n in LDS (pencil stride `ld`).  Each
// stage of radix R performs the autosort (Stockham, decimation in time) step
//   for butterfly j in [0, n/R):  k = j mod Ls
//     v_r = x[j + r n/R] * w_{Ls R}^{r k}      (twiddle, skipped when Ls == 1)
//     V = DFT_R(v)
//     y[(j - k) R + k + r Ls] = V_r
// with Ls the product of the radices of the earlier stages; after the last
// stage the pencil is in natural order.  Values go LDS -> registers ->
// barrier -> LDS, so one buffer suffices.  Twiddles w_{Ls R}^{r k} are read
// from an n-entry table e^{-2 pi i m / n} (FP64-accurate, rounded to T) at
// m = r k n / (Ls R); the inverse uses the conjugate.
#pragma once

#include "gbg_internal.hpp"

namespace gbg {

// sqrt(1/2), cos/sin of 2pi/3, 2pi/5, 4pi/5 to double precision
constexpr double kSqrtHalf = 0.70710678118654752440084436210484903928;
constexpr double kSin60 = 0.86602540378443864676372317075293618347;
constexpr double kC1_5 = 0.30901699437494742410229341718281905886;   // cos(2pi/5)
constexpr double kC2_5 = -0.80901699437494742410229341718281905886;  // cos(4pi/5)
constexpr double kS1_5 = 0.95105651629515357211643933337938214340;   // sin(2pi/5)
constexpr double kS2_5 = 0.58778525229247312916870595463907276860;   // sin(4pi/5)

// multiply by DIR*i (forward DIR = -1 gives -i, inverse +i)
template <int DIR, typename T>
__device__ __forceinline__ cpx<T> rot(cpx<T> a) {
  if (DIR < 0) return mk<T>(a.y, -a.x);
  return mk<T>(-a.y, a.x);
}

// multiply by w8 = e^{DIR 2 pi i / 8}
template <int DIR, typename T>
__device__ __forceinline__ cpx<T> w8(cpx<T> a) {
  const T c = T(kSqrtHalf);
  if (DIR < 0) return mk<T>(c * (a.x + a.y), c * (a.y - a.x));
  return mk<T>(c * (a.x - a.y), c * (a.x + a.y));
}

template <int R, int DIR, typename T>
struct Dft;

template <int DIR, typename T>
struct Dft<2, DIR, T> {
  __device__ static __forceinline__ void run(cpx<T>* v, const cpx<T>*, int) {
    const cpx<T> a = v[0], b = v[1];
    v[0] = cadd(a, b);
    v[1] = csub(a, b);
  }
};

template <int DIR, typename T>
struct Dft<4, DIR, T> {
  __device__ static __forceinline__ void run(cpx<T>* v, const cpx<T>*, int) {
    const cpx<T> t0 = cadd(v[0], v[2]), t1 = csub(v[0], v[2]);
    const cpx<T> t2 = cadd(v[1], v[3]), t3 = rot<DIR>(csub(v[1], v[3]));
    v[0] = cadd(t0, t2);
    v[2] = csub(t0, t2);
    v[1] = cadd(t1, t3);
    v[3] = csub(t1, t3);
  }
};

template <int DIR, typename T>
struct Dft<8, DIR, T> {
  __device__ static __forceinline__ void run(cpx<T>* v, const cpx<T>*, int) {
    cpx<T> e[4] = {v[0], v[2], v[4], v[6]};
    cpx<T> o[4] = {v[1], v[3], v[5], v[7]};
    Dft<4, DIR, T>::run(e, nullptr, 0);
    Dft<4, DIR, T>::run(o, nullptr, 0);
    const cpx<T> o1 = w8<DIR>(o[1]);
    const cpx<T> o2 = rot<DIR>(o[2]);
    const cpx<T> o3 = rot<DIR>(w8<DIR>(o[3]));
    v[0] = cadd(e[0], o[0]);
    v[4] = csub(e[0], o[0]);
    v[1] = cadd(e[1], o1);
    v[5] = csub(e[1], o1);
    v[2] = cadd(e[2], o2);
    v[6] = csub(e[2], o2);
    v[3] = cadd(e[3], o3);
    v[7] = csub(e[3], o3);
  }
};

template <int DIR, typename T>
struct Dft<3, DIR, T> {
  __device__ static __forceinline__ void run(cpx<T>* v, const cpx<T>*, int) {
    const cpx<T> s = cadd(v[1], v[2]);
    const cpx<T> d = cscale(rot<DIR>(csub(v[1], v[2])), T(kSin60));
    const cpx<T> t = csub(v[0], cscale(s, T(0.5)));
    v[0] = cadd(v[0], s);
    v[1] = cadd(t, d);
    v[2] = csub(t, d);
  }
};

template <int DIR, typename T>
struct Dft<5, DIR, T> {
  __device__ static __forceinline__ void run(cpx<T>* v, const cpx<T>*, int) {
    const cpx<T> s14 = cadd(v[1], v[4]), d14 = csub(v[1], v[4]);
    const cpx<T> s23 = cadd(v[2], v[3]), d23 = csub(v[2], v[3]);
    const T c1 = T(kC1_5), c2 = T(kC2_5), s1 = T(kS1_5), s2 = T(kS2_5);
    const cpx<T> a1 = cadd(v[0], cadd(cscale(s14, c1), cscale(s23, c2)));
    const cpx<T> a2 = cadd(v[0], cadd(cscale(s14, c2), cscale(s23, c1)));
    const cpx<T> b1 = rot<DIR>(cadd(cscale(d14, s1), cscale(d23, s2)));
    const cpx<T> b2 = rot<DIR>(csub(cscale(d14, s2), cscale(d23, s1)));
    v[0] = cadd(v[0], cadd(s14, s23));
    v[1] = cadd(a1, b1);
    v[4] = csub(a1, b1);
    v[2] = cadd(a2, b2);
    v[3] = csub(a2, b2);
  }
};

// Generic odd prime radix: direct O(R^2) sum with roots read from the
// n-entry table (w_R^m = tw[(m mod R) n/R]).
template <int R, int DIR, typename T>
struct Dft {
  __device__ static __forceinline__ void run(cpx<T>* v, const cpx<T>* tw, int n) {
    cpx<T> out[R];
    const int step = n / R;
#pragma unroll
    for (int q = 0; q < R; ++q) {
      cpx<T> acc = v[0];
#pragma unroll
      for (int r = 1; r < R; ++r) {
        cpx<T> w = tw[((r * q) % R) * step];
        if (DIR > 0) w.y = -w.y;
        acc = cadd(acc, cmul(v[r], w));
      }
      out[q] = acc;
    }
#pragma unroll
    for (int q = 0; q < R; ++q) v[q] = out[q];
  }
};

// One Stockham stage over B pencils, LDS ping-pong: butterflies are read
// from `src`, transformed in registers and written to `dst` at once, so a
// thread keeps only R values live.
template <typename T, int R, int DIR>
__device__ __forceinline__ void stage(const cpx<T>* src, cpx<T>* dst, int ld, int n, int Ls, int B,
                                      const cpx<T>* __restrict__ tw) {
  const int nr = n / R;
  const int nbf = B * nr;
  const int twstep = n / (Ls * R);
  for (int idx = threadIdx.x; idx < nbf; idx += blockDim.x) {
    const int p = idx / nr, j = idx - p * nr;
    const int k = j % Ls;
    cpx<T> v[R];
    const cpx<T>* in = src + p * ld + j;
#pragma unroll
    for (int r = 0; r < R; ++r) v[r] = in[r * nr];
    if (Ls > 1) {
#pragma unroll
      for (int r = 1; r < R; ++r) {
        cpx<T> w = tw[r * k * twstep];
        if (DIR > 0) w.y = -w.y;
        v[r] = cmul(v[r], w);
      }
    }
    Dft<R, DIR, T>::run(v, tw, n);
    cpx<T>* out = dst + p * ld + (j - k) * R + k;
#pragma unroll
    for (int r = 0; r < R; ++r) out[r * Ls] = v[r];
  }
  __syncthreads();
}

// Whole-pencil transform of the B pencils in buffer s0 (natural order in
// and out); s1 is the ping-pong partner.  Returns the buffer that holds the
// result.
template <typename T, int DIR>
__device__ cpx<T>* fft_lds(cpx<T>* s0, cpx<T>* s1, int ld, const AxisPlan& ap, int B,
                           const cpx<T>* __restrict__ tw) {
  int Ls = 1;
  cpx<T>* a = s0;
  cpx<T>* b = s1;
  for (int st = 0; st < ap.nst; ++st) {
    const int R = ap.radix[st];
    switch (R) {
      case 8: stage<T, 8, DIR>(a, b, ld, ap.n, Ls, B, tw); break;
      case 4: stage<T, 4, DIR>(a, b, ld, ap.n, Ls, B, tw); break;
      case 2: stage<T, 2, DIR>(a, b, ld, ap.n, Ls, B, tw); break;
      case 3: stage<T, 3, DIR>(a, b, ld, ap.n, Ls, B, tw); break;
      case 5: stage<T, 5, DIR>(a, b, ld, ap.n, Ls, B, tw); break;
      case 7: stage<T, 7, DIR>(a, b, ld, ap.n, Ls, B, tw); break;
      case 11: stage<T, 11, DIR>(a, b, ld, ap.n, Ls, B, tw); break;
      case 13: stage<T, 13, DIR>(a, b, ld, ap.n, Ls, B, tw); break;
      default: break;
    }
    Ls *= R;
    cpx<T>* t = a;
    a = b;
    b = t;
  }
  return a;
}

// Whole-pencil transform of any axis plan: the mixed-radix form above, or for
// a chirp-z axis (ap.m > 0) Bluestein's convolution
//   X_k = c_k sum_j (x_j c_j) conj(c_{k-j}),  c_j = e^{DIR i pi j^2 / n},
// carried out in LDS: the n entries of each pencil times the chirps, zeros up
// to m, a forward m-point transform, the filter spectrum (1/m included), an
// inverse m-point transform, and the chirps again on the first n entries.  The
// pencils in s0 hold n entries (stride ld >= m + 1) and a barrier separates
// their writes from this call; the result (n entries) is in the returned buffer.
template <typename T, int DIR>
__device__ cpx<T>* fft_any(cpx<T>* s0, cpx<T>* s1, int ld, const AxisPlan& ap, int B,
                           const cpx<T>* __restrict__ tw) {
  if (ap.m == 0) return fft_lds<T, DIR>(s0, s1, ld, ap, B, tw);
  const int n = ap.n, m = ap.m;
  const cpx<T>* chirp = tw + m;
  const cpx<T>* filt = tw + m + n + (DIR > 0 ? m : 0);
  auto c_at = [&](int j) {
    cpx<T> c = chirp[j];
    if (DIR > 0) c.y = -c.y;
    return c;
  };
  for (int idx = threadIdx.x; idx < B * m; idx += blockDim.x) {
    const int p = idx / m, e = idx - p * m;
    cpx<T>* v = s0 + p * ld + e;
    *v = e < n ? cmul(*v, c_at(e)) : mk<T>(T(0), T(0));
  }
  __syncthreads();
  AxisPlan am = ap;
  am.n = m;
  cpx<T>* r = fft_lds<T, -1>(s0, s1, ld, am, B, tw);
  for (int idx = threadIdx.x; idx < B * m; idx += blockDim.x) {
    const int p = idx / m, k = idx - p * m;
    r[p * ld + k] = cmul(r[p * ld + k], filt[k]);
  }
  __syncthreads();
  cpx<T>* y = fft_lds<T, 1>(r, r == s0 ? s1 : s0, ld, am, B, tw);
  for (int idx = threadIdx.x; idx < B * n; idx += blockDim.x) {
    const int p = idx / n, k = idx - p * n;
    y[p * ld + k] = cmul(y[p * ld + k], c_at(k));
  }
  __syncthreads();
  return y;
}

// LDS of the generic pass kernels: two buffers (ping-pong) of 64 KiB payload
// each plus one padding element per pencil (B <= 64).
template <typename T>
struct Lds {
  static constexpr int kCap = 65536 / (2 * sizeof(T));
  static constexpr int kElems = kCap + 64;
};

constexpr int kNT = 256;      // threads per workgroup of the generic passes
constexpr int kMaxB = 64;     // pencils per workgroup, at most

// pencils per workgroup for a transform of length n holding `per` LDS
// elements per pencil: the largest power of two with B * (per + 1) <= cap.
inline int pick_b(int per, int cap) {
  int b = 1;
  while (b * 2 <= kMaxB && (b * 2) * (per + 1) <= cap && b * 2 <= kNT) b *= 2;
  return b;
}

}  // namespace gbg
