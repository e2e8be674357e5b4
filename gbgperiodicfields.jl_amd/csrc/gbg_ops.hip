// Spectral operators, physical products and the Parseval reduction as
// standalone streaming kernels, one work item per mode (or grid point).
// Per-mode arithmetic follows the reference kernels expression for
// expression (src/DifferentialOperators.jl, src/Convolution.jl,
// src/FieldAlgebra.jl, src/Reductions.jl); wavenumbers come from the plan's
// FP64 tables and are rounded to T where the field is Float32, while the
// band mask is always evaluated in FP64.
#include "gbg_internal.hpp"

namespace gbg {

constexpr int kOpNT = 256;

// decode a linear mode index into (kx, ky, kz).  The two divisions by the
// run-time nh and ny go through an FP64 reciprocal: n rh (n < 2^32, rh = 1/nh
// rounded) is within 2^-21 of n / nh, so its truncation is the quotient or,
// when n / nh is an integer, possibly one less, which the remainder test
// corrects.  Seven instructions instead of an integer division's thirty.
struct ModeIdx {
  unsigned nh, ny;
  double rh, ry;
  __device__ __forceinline__ static unsigned div(unsigned n, unsigned d, double rd) {
    unsigned q = (unsigned)((double)n * rd);
    q += (n - q * d) >= d ? 1u : 0u;
    return q;
  }
  __device__ __forceinline__ void operator()(unsigned idx, unsigned& kx, unsigned& ky,
                                             unsigned& kz) const {
    const unsigned t = div(idx, nh, rh);
    kx = idx - t * nh;
    kz = div(t, ny, ry);
    ky = t - kz * ny;
  }
};

// spectral_band membership in FP64: lo2 <= (k1^2 + k2^2) + k3^2 <= hi2
__device__ __forceinline__ bool in_band(const Tables& t, unsigned kx, unsigned ky, unsigned kz,
                                        double lo2, double hi2) {
  const double q = (t.ks1[kx] + t.ks2[ky]) + t.ks3[kz];
  return lo2 <= q && q <= hi2;
}

static unsigned nblocks(size_t n) { return (unsigned)((n + kOpNT - 1) / kOpNT); }

// Kernels that move few bytes per entry (the Laplacians of a scalar, the scalar
// product) take kSpan entries per thread, every load issued before the first
// store: with one entry per thread too few bytes were in flight to cover the
// HBM latency.  Entries past the end load the last entry (a valid address)
// and are not stored.
// Only FP32 takes the spans: in FP64 (16-byte modes) one entry per thread
// measured faster for the vector Laplacian.
template <typename T>
constexpr int span() { return sizeof(T) == 4 ? 4 : 1; }
template <typename T>
static unsigned nblocks_span(size_t n) {
  return (unsigned)((n + (size_t)kOpNT * span<T>() - 1) / ((size_t)kOpNT * span<T>()));
}

// ---- gradient ----------------------------------------------------------------
// scalar -> vector: d_j = (i a_j) s      (_scalar_gradient_kernel!, :357-368)
// vector -> tensor: d_ij = (i a_j) u_i   (_vector_gradient_kernel!, :414-441)
//   tensor component (i, j) is stored at block i + 3 j (Julia column-major).
template <typename T>
__global__ __launch_bounds__(kOpNT) void k_gradient(cpx<T>* __restrict__ dst,
                                                    const cpx<T>* __restrict__ src, unsigned M,
                                                    int nsrc, ModeIdx mi, Tables t) {
  const unsigned idx = blockIdx.x * kOpNT + threadIdx.x;
  if (idx >= M) return;
  unsigned kx, ky, kz;
  mi(idx, kx, ky, kz);
  const T a[3] = {T(t.kd1[kx]), T(t.kd2[ky]), T(t.kd3[kz])};
  for (int i = 0; i < nsrc; ++i) {
    const cpx<T> v = src[(size_t)i * M + idx];
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      // (i a) * v with the factor formed first: (0 + i a)(x + i y)
      const cpx<T> f = mk<T>(T(0), a[j]);
      dst[(size_t)(i + nsrc * j) * M + idx] = cmul(f, v);
    }
  }
}

// FP32 scalar -> vector with an even mode count and 16-byte aligned buffers:
// two consecutive modes per lane, one 16-byte load and three 16-byte stores,
// the same per-mode arithmetic as k_gradient.
__global__ __launch_bounds__(kOpNT) void k_gradient_pair(float4* __restrict__ dst,
                                                         const float4* __restrict__ src,
                                                         unsigned M2, ModeIdx mi, Tables t) {
  const unsigned idx = blockIdx.x * kOpNT + threadIdx.x;
  if (idx >= M2) return;
  float a[2][3];
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    unsigned kx, ky, kz;
    mi(2 * idx + h, kx, ky, kz);
    a[h][0] = float(t.kd1[kx]);
    a[h][1] = float(t.kd2[ky]);
    a[h][2] = float(t.kd3[kz]);
  }
  const float4 v = src[idx];
  const cpx<float> v0 = mk<float>(v.x, v.y), v1 = mk<float>(v.z, v.w);
#pragma unroll
  for (int j = 0; j < 3; ++j) {
    const cpx<float> c0 = cmul(mk<float>(0.0f, a[0][j]), v0);
    const cpx<float> c1 = cmul(mk<float>(0.0f, a[1][j]), v1);
    dst[(size_t)j * M2 + idx] = make_float4(c0.x, c0.y, c1.x, c1.y);
  }
}

// ---- symmetric gradient --------------------------------------------------------
// vector -> symmetric tensor (_symmetric_gradient_kernel!, :666-699):
//   S_ii = (i a_i) u_i,  S_ij = S_ji = (i/2) (a_j u_i + a_i u_j)
// with (i/2) the complex (0, 1/2) multiplying the sum.  P consecutive modes
// per lane (P = 2 in FP32: 16-byte accesses), the three sources read once and
// the six distinct values written to the nine blocks (block i + 3 j).
template <typename T, int P>
struct alignas(P * sizeof(cpx<T>)) ModePack {
  cpx<T> c[P];
};

template <typename T, int P>
__global__ __launch_bounds__(kOpNT) void k_symmetric_gradient(ModePack<T, P>* __restrict__ dst,
                                                              const ModePack<T, P>* __restrict__ src,
                                                              unsigned MP, ModeIdx mi, Tables t) {
  const unsigned idx = blockIdx.x * kOpNT + threadIdx.x;
  if (idx >= MP) return;
  ModePack<T, P> u[3], d[6];
#pragma unroll
  for (int i = 0; i < 3; ++i) u[i] = src[(size_t)i * MP + idx];
  const cpx<T> half_i = mk<T>(T(0), T(0.5));
#pragma unroll
  for (int h = 0; h < P; ++h) {
    unsigned kx, ky, kz;
    mi(P * idx + h, kx, ky, kz);
    const T a[3] = {T(t.kd1[kx]), T(t.kd2[ky]), T(t.kd3[kz])};
#pragma unroll
    for (int i = 0; i < 3; ++i) d[i].c[h] = cmul(mk<T>(T(0), a[i]), u[i].c[h]);
    // d[3] = S_12, d[4] = S_13, d[5] = S_23
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      const int i = q == 2 ? 1 : 0, j = q == 0 ? 1 : 2;
      const cpx<T> sum = cadd(cscale(u[i].c[h], a[j]), cscale(u[j].c[h], a[i]));
      d[3 + q].c[h] = cmul(half_i, sum);
    }
  }
  // blocks: (1,1) 0, (2,1) 1, (3,1) 2, (1,2) 3, (2,2) 4, (3,2) 5, (1,3) 6, (2,3) 7, (3,3) 8
  constexpr int kBlock[9] = {0, 3, 4, 3, 1, 5, 4, 5, 2};
#pragma unroll
  for (int b = 0; b < 9; ++b) dst[(size_t)b * MP + idx] = d[kBlock[b]];
}

// ---- divergence --------------------------------------------------------------
// i ((a1 s1 + a2 s2) + a3 s3), optionally alpha * g * (.) + beta * dst
//   (_apply_spectral_divergence!, :120-139)
template <typename T>
__global__ __launch_bounds__(kOpNT) void k_divergence(cpx<T>* dst, const cpx<T>* src, unsigned M,
                                                      int nout, ModeIdx mi, Tables t, int has_band,
                                                      double lo2, double hi2, T alpha, T beta) {
  const unsigned idx = blockIdx.x * kOpNT + threadIdx.x;
  if (idx >= M) return;
  unsigned kx, ky, kz;
  mi(idx, kx, ky, kz);
  const T a1 = T(t.kd1[kx]), a2 = T(t.kd2[ky]), a3 = T(t.kd3[kz]);
  const T g = has_band ? (in_band(t, kx, ky, kz, lo2, hi2) ? T(1) : T(0)) : T(1);
  for (int i = 0; i < nout; ++i) {
    // source block (i, j) of a tensor is i + nout*j; a vector has nout = 1
    const cpx<T> s1 = src[(size_t)(i + nout * 0) * M + idx];
    const cpx<T> s2 = src[(size_t)(i + nout * 1) * M + idx];
    const cpx<T> s3 = src[(size_t)(i + nout * 2) * M + idx];
    const cpx<T> sum = cadd(cadd(cscale(s1, a1), cscale(s2, a2)), cscale(s3, a3));
    cpx<T> r = cmul(mk<T>(T(0), T(1)), sum);
    if (has_band || alpha != T(1) || beta != T(0)) {
      r = cscale(r, alpha * g);
      if (beta != T(0)) r = cadd(r, cscale(dst[(size_t)i * M + idx], beta));
    }
    dst[(size_t)i * M + idx] = r;
  }
}

// ---- curl / inverse curl -----------------------------------------------------
// i (a2 u3 - a3 u2, a3 u1 - a1 u3, a1 u2 - a2 u1)          (_curl_kernel!, :502-523)
// the same times (|a|^2 == 0 ? 0 : 1/|a|^2)                (_inverse_curl_kernel!, :580-605)
// All reads of a mode precede its writes: dst may alias src.
template <typename T, bool INV>
__global__ __launch_bounds__(kOpNT) void k_curl(cpx<T>* dst, const cpx<T>* src, unsigned M,
                                                ModeIdx mi, Tables t) {
  const unsigned idx = blockIdx.x * kOpNT + threadIdx.x;
  if (idx >= M) return;
  unsigned kx, ky, kz;
  mi(idx, kx, ky, kz);
  const T a1 = T(t.kd1[kx]), a2 = T(t.kd2[ky]), a3 = T(t.kd3[kz]);
  const cpx<T> u1 = src[idx], u2 = src[(size_t)M + idx], u3 = src[2 * (size_t)M + idx];
  cpx<T> c1 = cmuli(csub(cscale(u3, a2), cscale(u2, a3)));
  cpx<T> c2 = cmuli(csub(cscale(u1, a3), cscale(u3, a1)));
  cpx<T> c3 = cmuli(csub(cscale(u2, a1), cscale(u1, a2)));
  if (INV) {
    const T sq = a1 * a1 + a2 * a2 + a3 * a3;
    const T rec = sq == T(0) ? T(0) : T(1) / sq;
    c1 = cscale(c1, rec);
    c2 = cscale(c2, rec);
    c3 = cscale(c3, rec);
  }
  dst[idx] = c1;
  dst[(size_t)M + idx] = c2;
  dst[2 * (size_t)M + idx] = c3;
}

// ---- Laplacian / inverse Laplacian --------------------------------------------
// -|k|^2 f   and   (-f) * (|k|^2 == 0 ? 0 : inv(|k|^2))   (:147-164)
// |k|^2 = (k1^2 + k2^2) + k3^2 from the squared family, in T.
template <typename T, bool INV>
__global__ __launch_bounds__(kOpNT) void k_laplacian(cpx<T>* dst, const cpx<T>* src, unsigned M,
                                                     int ncomp, ModeIdx mi, Tables t) {
  constexpr int kSpan = span<T>();
  const unsigned base = blockIdx.x * (kOpNT * kSpan) + threadIdx.x;
  unsigned at[kSpan];
  T f[kSpan];
#pragma unroll
  for (int j = 0; j < kSpan; ++j) {
    const unsigned idx = base + j * kOpNT;
    at[j] = idx < M ? idx : M - 1;
    unsigned kx, ky, kz;
    mi(at[j], kx, ky, kz);
    const T q = (T(t.ks1[kx]) + T(t.ks2[ky])) + T(t.ks3[kz]);
    f[j] = INV ? (q == T(0) ? T(0) : T(1) / q) : -q;
  }
  for (int c = 0; c < ncomp; ++c) {
    cpx<T> v[kSpan];
#pragma unroll
    for (int j = 0; j < kSpan; ++j) v[j] = src[(size_t)c * M + at[j]];
#pragma unroll
    for (int j = 0; j < kSpan; ++j) {
      const unsigned idx = base + j * kOpNT;
      if (idx < M)
        dst[(size_t)c * M + idx] = INV ? cscale(mk<T>(-v[j].x, -v[j].y), f[j]) : cscale(v[j], f[j]);
    }
  }
}

// FP32 scalar field with an even mode count and 16-byte aligned buffers: two
// consecutive modes per lane, one 16-byte load and store each, two such pairs
// per thread (8-byte lanes left the FP32 scalar Laplacians at about 4.6 TB/s;
// 512^3 inverse Laplacian 0.23 -> 0.21 ms).  For a vector field it measured
// slower (0.57 -> 0.59 ms), so three components keep k_laplacian.  The same
// per-mode arithmetic as k_laplacian.
template <bool INV>
__global__ __launch_bounds__(kOpNT) void k_laplacian_pair(float4* dst, const float4* src,
                                                          unsigned M2, int ncomp, ModeIdx mi,
                                                          Tables t) {
  constexpr int kPairs = 2;
  const unsigned base = blockIdx.x * (kOpNT * kPairs) + threadIdx.x;
  unsigned at[kPairs];
  float f[kPairs][2];
#pragma unroll
  for (int j = 0; j < kPairs; ++j) {
    const unsigned idx = base + j * kOpNT;
    at[j] = idx < M2 ? idx : M2 - 1;
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      unsigned kx, ky, kz;
      mi(2 * at[j] + h, kx, ky, kz);
      const float q = (float(t.ks1[kx]) + float(t.ks2[ky])) + float(t.ks3[kz]);
      f[j][h] = INV ? (q == 0.0f ? 0.0f : 1.0f / q) : -q;
    }
  }
  for (int c = 0; c < ncomp; ++c) {
    float4 v[kPairs];
#pragma unroll
    for (int j = 0; j < kPairs; ++j) v[j] = src[(size_t)c * M2 + at[j]];
#pragma unroll
    for (int j = 0; j < kPairs; ++j) {
      const unsigned idx = base + j * kOpNT;
      if (idx < M2) {
        float4 o;
        if (INV) {
          o.x = -v[j].x * f[j][0];
          o.y = -v[j].y * f[j][0];
          o.z = -v[j].z * f[j][1];
          o.w = -v[j].w * f[j][1];
        } else {
          o.x = v[j].x * f[j][0];
          o.y = v[j].y * f[j][0];
          o.z = v[j].z * f[j][1];
          o.w = v[j].w * f[j][1];
        }
        dst[(size_t)c * M2 + idx] = o;
      }
    }
  }
}

// ---- Helmholtz (Leray) projection ----------------------------------------------
// L = (a.u) / (a.a == 0 ? 1 : a.a);  u_i <- alpha g (u_i - a_i L) (+ beta d_i)
//   (_helmholtz_projection_kernel!, :191-230)
template <typename T>
__global__ __launch_bounds__(kOpNT) void k_helmholtz(cpx<T>* dst, const cpx<T>* src, unsigned M,
                                                     ModeIdx mi, Tables t, int has_band, double lo2,
                                                     double hi2, T alpha, T beta) {
  const unsigned idx = blockIdx.x * kOpNT + threadIdx.x;
  if (idx >= M) return;
  unsigned kx, ky, kz;
  mi(idx, kx, ky, kz);
  const T a1 = T(t.kd1[kx]), a2 = T(t.kd2[ky]), a3 = T(t.kd3[kz]);
  const T sq = a1 * a1 + a2 * a2 + a3 * a3;
  const cpx<T> u1 = src[idx], u2 = src[(size_t)M + idx], u3 = src[2 * (size_t)M + idx];
  const cpx<T> dot = cadd(cadd(cscale(u1, a1), cscale(u2, a2)), cscale(u3, a3));
  const T den = sq == T(0) ? T(1) : sq;
  const cpx<T> lon = mk<T>(dot.x / den, dot.y / den);
  const T g = has_band ? (in_band(t, kx, ky, kz, lo2, hi2) ? T(1) : T(0)) : T(1);
  const T sc = alpha * g;
  cpx<T> p1 = cscale(csub(u1, cscale(lon, a1)), sc);
  cpx<T> p2 = cscale(csub(u2, cscale(lon, a2)), sc);
  cpx<T> p3 = cscale(csub(u3, cscale(lon, a3)), sc);
  if (beta != T(0)) {
    p1 = cadd(p1, cscale(dst[idx], beta));
    p2 = cadd(p2, cscale(dst[(size_t)M + idx], beta));
    p3 = cadd(p3, cscale(dst[2 * (size_t)M + idx], beta));
  }
  dst[idx] = p1;
  dst[(size_t)M + idx] = p2;
  dst[2 * (size_t)M + idx] = p3;
}

// ---- radial band convolution (the 2/3 mask) ---------------------------------------
// alpha * sigma(|k|^2) * f (+ beta dst)   (_radial_convolution_pass!, src/Convolution.jl:756-787)
template <typename T>
__global__ __launch_bounds__(kOpNT) void k_band(cpx<T>* dst, const cpx<T>* src, unsigned M,
                                                int ncomp, ModeIdx mi, Tables t, double lo2,
                                                double hi2, T alpha, T beta) {
  const unsigned idx = blockIdx.x * kOpNT + threadIdx.x;
  if (idx >= M) return;
  unsigned kx, ky, kz;
  mi(idx, kx, ky, kz);
  const T g = in_band(t, kx, ky, kz, lo2, hi2) ? T(1) : T(0);
  const T f = alpha * g;
  for (int c = 0; c < ncomp; ++c) {
    cpx<T> r = cscale(src[(size_t)c * M + idx], f);
    if (beta != T(0)) r = cadd(r, cscale(dst[(size_t)c * M + idx], beta));
    dst[(size_t)c * M + idx] = r;
  }
}

// ---- mean-free projection: zero the k = 0 coefficient of each component -------------
template <typename T>
__global__ void k_mean_free(cpx<T>* spec, size_t M, int ncomp) {
  const int c = threadIdx.x;
  if (c < ncomp) spec[(size_t)c * M] = mk<T>(T(0), T(0));
}

// ---- physical products --------------------------------------------------------------
// (u x v)_i = u_j v_k - u_k v_j   (_cross_formula, src/FiberSections.jl:641-652)
template <typename T>
__global__ __launch_bounds__(kOpNT) void k_cross(T* dst, const T* u, const T* v, size_t P) {
  const size_t i = (size_t)blockIdx.x * kOpNT + threadIdx.x;
  if (i >= P) return;
  const T u1 = u[i], u2 = u[P + i], u3 = u[2 * P + i];
  const T v1 = v[i], v2 = v[P + i], v3 = v[2 * P + i];
  dst[i] = u2 * v3 - u3 * v2;
  dst[P + i] = u3 * v1 - u1 * v3;
  dst[2 * P + i] = u1 * v2 - u2 * v1;
}

template <typename T>
__global__ __launch_bounds__(kOpNT) void k_multiply(T* dst, const T* a, const T* b, size_t P) {
  constexpr int kSpan = span<T>();
  const size_t base = (size_t)blockIdx.x * (kOpNT * kSpan) + threadIdx.x;
  T x[kSpan], y[kSpan];
#pragma unroll
  for (int j = 0; j < kSpan; ++j) {
    const size_t i = base + (size_t)j * kOpNT;
    const size_t ic = i < P ? i : P - 1;
    x[j] = a[ic];
    y[j] = b[ic];
  }
#pragma unroll
  for (int j = 0; j < kSpan; ++j) {
    const size_t i = base + (size_t)j * kOpNT;
    if (i < P) dst[i] = x[j] * y[j];
  }
}

// ---- Parseval inner, FP64 accumulation in a fixed order --------------------------------
// sum_c sum_modes w(k_f) Re(f conj g), w = 1 on the folded axis's k_f = 0 plane and its
// even Nyquist plane, else 2; result / N^2   (src/Reductions.jl:883-925, weight :956-965)
// Each block sums a fixed contiguous chunk; a thread takes every kOpNT-th entry
// of it, kRedU loads in flight at a time into four accumulators (a single
// accumulator made each step wait for the previous add and left the loads
// unoverlapped).  With st the storage stride of the folded axis and period =
// st x its extent (a divisor of M), q = i mod period over the flattened
// (component, mode) index advances by kOpNT mod period, and k_f = 0 is
// q < st, the Nyquist plane nlo <= q < nhi (x folded: st = 1, q = kx).
constexpr int kRedU = 8;

struct FoldW {
  unsigned period, st, nlo, nhi;
};

static FoldW fold_weights(const gbg_plan* p) {
  unsigned st = 1;
  for (int a = 0; a < p->fold; ++a) st *= (unsigned)p->s[a];
  const unsigned ext = (unsigned)p->s[p->fold];
  const bool even = p->n[p->fold] % 2 == 0 && ext > 1;
  return FoldW{st * ext, st, even ? (ext - 1) * st : 0u, even ? ext * st : 0u};
}

template <typename T>
__global__ __launch_bounds__(kOpNT) void k_inner_part(const cpx<T>* f, const cpx<T>* g, size_t M,
                                                      int ncomp, FoldW fw, double* part) {
  __shared__ double red[kOpNT];
  const size_t total = M * (size_t)ncomp;
  const size_t chunk = (total + gridDim.x - 1) / gridDim.x;
  const size_t lo = chunk * blockIdx.x;
  const size_t hi = lo + chunk < total ? lo + chunk : total;
  const unsigned step = kOpNT % fw.period;
  size_t i = lo + threadIdx.x;
  unsigned q = (unsigned)(i % fw.period);
  double acc[4] = {0.0, 0.0, 0.0, 0.0};
  auto add = [&](int u, const cpx<T>& a, const cpx<T>& b) {
    const double w = (q < fw.st || (q >= fw.nlo && q < fw.nhi)) ? 1.0 : 2.0;
    acc[u & 3] += w * ((double)a.x * (double)b.x + (double)a.y * (double)b.y);
    q += step;
    if (q >= fw.period) q -= fw.period;
  };
  for (; i + (size_t)(kRedU - 1) * kOpNT < hi; i += (size_t)kRedU * kOpNT) {
    cpx<T> a[kRedU], b[kRedU];
#pragma unroll
    for (int u = 0; u < kRedU; ++u) {
      a[u] = f[i + (size_t)u * kOpNT];
      b[u] = g[i + (size_t)u * kOpNT];
    }
#pragma unroll
    for (int u = 0; u < kRedU; ++u) add(u, a[u], b[u]);
  }
  for (; i < hi; i += kOpNT) add(0, f[i], g[i]);
  red[threadIdx.x] = (acc[0] + acc[1]) + (acc[2] + acc[3]);
  __syncthreads();
  for (int s = kOpNT / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

// physical inner: sum of a[i] b[i] over every stored real (all components),
// FP64 partials in the same fixed grid (inner, src/Reductions.jl:838-853)
template <typename T>
__global__ __launch_bounds__(kOpNT) void k_dot_part(const T* a, const T* b, size_t total,
                                                    double* part) {
  __shared__ double red[kOpNT];
  const size_t chunk = (total + gridDim.x - 1) / gridDim.x;
  const size_t lo = chunk * blockIdx.x;
  const size_t hi = lo + chunk < total ? lo + chunk : total;
  size_t i = lo + threadIdx.x;
  double acc[4] = {0.0, 0.0, 0.0, 0.0};
  for (; i + (size_t)(kRedU - 1) * kOpNT < hi; i += (size_t)kRedU * kOpNT) {
    T x[kRedU], y[kRedU];
#pragma unroll
    for (int u = 0; u < kRedU; ++u) {
      x[u] = a[i + (size_t)u * kOpNT];
      y[u] = b[i + (size_t)u * kOpNT];
    }
#pragma unroll
    for (int u = 0; u < kRedU; ++u) acc[u & 3] += (double)x[u] * (double)y[u];
  }
  for (; i < hi; i += kOpNT) acc[0] += (double)a[i] * (double)b[i];
  red[threadIdx.x] = (acc[0] + acc[1]) + (acc[2] + acc[3]);
  __syncthreads();
  for (int s = kOpNT / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

__global__ __launch_bounds__(kOpNT) void k_inner_final(const double* part, int nparts, double norm,
                                                       double* out) {
  __shared__ double red[kOpNT];
  double acc = 0.0;
  for (int i = threadIdx.x; i < nparts; i += kOpNT) acc += part[i];
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int s = kOpNT / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = red[0] / norm;
}

// ----------------------------------------------------------------------------------------
// host launchers
// ----------------------------------------------------------------------------------------
static ModeIdx modes_of(gbg_plan* p) {
  return ModeIdx{(unsigned)p->s[0], (unsigned)p->s[1], 1.0 / (double)p->s[0], 1.0 / (double)p->s[1]};
}

gbg_status op_gradient(gbg_plan* p, void* dst, const void* src, int src_order, hipStream_t s) {
  const unsigned M = (unsigned)p->M;
  const int nsrc = src_order == 0 ? 1 : 3;
  if (p->elem == 8)
    hipLaunchKernelGGL((k_gradient<double>), nblocks(M), kOpNT, 0, s, (cpx<double>*)dst,
                       (const cpx<double>*)src, M, nsrc, modes_of(p), tables_of(p));
  else if (nsrc == 1 && M % 2 == 0 && ((uintptr_t)dst | (uintptr_t)src) % 16 == 0)
    hipLaunchKernelGGL(k_gradient_pair, nblocks(M / 2), kOpNT, 0, s, (float4*)dst,
                       (const float4*)src, M / 2, modes_of(p), tables_of(p));
  else
    hipLaunchKernelGGL((k_gradient<float>), nblocks(M), kOpNT, 0, s, (cpx<float>*)dst,
                       (const cpx<float>*)src, M, nsrc, modes_of(p), tables_of(p));
  return check_launch("gradient");
}

gbg_status op_symmetric_gradient(gbg_plan* p, void* dst, const void* src, hipStream_t s) {
  const unsigned M = (unsigned)p->M;
  if (p->elem == 8)
    hipLaunchKernelGGL((k_symmetric_gradient<double, 1>), nblocks(M), kOpNT, 0, s,
                       (ModePack<double, 1>*)dst, (const ModePack<double, 1>*)src, M, modes_of(p),
                       tables_of(p));
  else if (M % 2 == 0 && ((uintptr_t)dst | (uintptr_t)src) % 16 == 0)
    hipLaunchKernelGGL((k_symmetric_gradient<float, 2>), nblocks(M / 2), kOpNT, 0, s,
                       (ModePack<float, 2>*)dst, (const ModePack<float, 2>*)src, M / 2,
                       modes_of(p), tables_of(p));
  else
    hipLaunchKernelGGL((k_symmetric_gradient<float, 1>), nblocks(M), kOpNT, 0, s,
                       (ModePack<float, 1>*)dst, (const ModePack<float, 1>*)src, M, modes_of(p),
                       tables_of(p));
  return check_launch("symmetric gradient");
}

gbg_status op_divergence(gbg_plan* p, void* dst, const void* src, int src_order, int has_band,
                         double lo, double hi, double alpha, double beta, hipStream_t s) {
  const unsigned M = (unsigned)p->M;
  const int nout = src_order == 1 ? 1 : 3;
  const double lo2 = lo * lo, hi2 = hi * hi;
  if (p->elem == 8)
    hipLaunchKernelGGL((k_divergence<double>), nblocks(M), kOpNT, 0, s, (cpx<double>*)dst,
                       (const cpx<double>*)src, M, nout, modes_of(p), tables_of(p), has_band, lo2,
                       hi2, alpha, beta);
  else
    hipLaunchKernelGGL((k_divergence<float>), nblocks(M), kOpNT, 0, s, (cpx<float>*)dst,
                       (const cpx<float>*)src, M, nout, modes_of(p), tables_of(p), has_band, lo2,
                       hi2, (float)alpha, (float)beta);
  return check_launch("divergence");
}

gbg_status op_curl(gbg_plan* p, void* dst, const void* src, int inverse, hipStream_t s) {
  const unsigned M = (unsigned)p->M;
  if (p->elem == 8) {
    if (inverse)
      hipLaunchKernelGGL((k_curl<double, true>), nblocks(M), kOpNT, 0, s, (cpx<double>*)dst,
                         (const cpx<double>*)src, M, modes_of(p), tables_of(p));
    else
      hipLaunchKernelGGL((k_curl<double, false>), nblocks(M), kOpNT, 0, s, (cpx<double>*)dst,
                         (const cpx<double>*)src, M, modes_of(p), tables_of(p));
  } else {
    if (inverse)
      hipLaunchKernelGGL((k_curl<float, true>), nblocks(M), kOpNT, 0, s, (cpx<float>*)dst,
                         (const cpx<float>*)src, M, modes_of(p), tables_of(p));
    else
      hipLaunchKernelGGL((k_curl<float, false>), nblocks(M), kOpNT, 0, s, (cpx<float>*)dst,
                         (const cpx<float>*)src, M, modes_of(p), tables_of(p));
  }
  return check_launch("curl");
}

gbg_status op_laplacian(gbg_plan* p, void* dst, const void* src, int ncomp, int inverse,
                        hipStream_t s) {
  const unsigned M = (unsigned)p->M;
  if (p->elem == 8) {
    if (inverse)
      hipLaunchKernelGGL((k_laplacian<double, true>), nblocks_span<double>(M), kOpNT, 0, s, (cpx<double>*)dst,
                         (const cpx<double>*)src, M, ncomp, modes_of(p), tables_of(p));
    else
      hipLaunchKernelGGL((k_laplacian<double, false>), nblocks_span<double>(M), kOpNT, 0, s, (cpx<double>*)dst,
                         (const cpx<double>*)src, M, ncomp, modes_of(p), tables_of(p));
  } else if (ncomp == 1 && M % 2 == 0 && ((uintptr_t)dst | (uintptr_t)src) % 16 == 0) {
    const unsigned M2 = M / 2;
    const unsigned nb = (unsigned)((M2 + 2 * kOpNT - 1) / (2 * kOpNT));
    if (inverse)
      hipLaunchKernelGGL((k_laplacian_pair<true>), nb, kOpNT, 0, s, (float4*)dst, (const float4*)src,
                         M2, ncomp, modes_of(p), tables_of(p));
    else
      hipLaunchKernelGGL((k_laplacian_pair<false>), nb, kOpNT, 0, s, (float4*)dst, (const float4*)src,
                         M2, ncomp, modes_of(p), tables_of(p));
  } else {
    if (inverse)
      hipLaunchKernelGGL((k_laplacian<float, true>), nblocks_span<float>(M), kOpNT, 0, s, (cpx<float>*)dst,
                         (const cpx<float>*)src, M, ncomp, modes_of(p), tables_of(p));
    else
      hipLaunchKernelGGL((k_laplacian<float, false>), nblocks_span<float>(M), kOpNT, 0, s, (cpx<float>*)dst,
                         (const cpx<float>*)src, M, ncomp, modes_of(p), tables_of(p));
  }
  return check_launch("laplacian");
}

gbg_status op_helmholtz(gbg_plan* p, void* dst, const void* src, int has_band, double lo,
                        double hi, double alpha, double beta, hipStream_t s) {
  const unsigned M = (unsigned)p->M;
  const double lo2 = lo * lo, hi2 = hi * hi;
  if (p->elem == 8)
    hipLaunchKernelGGL((k_helmholtz<double>), nblocks(M), kOpNT, 0, s, (cpx<double>*)dst,
                       (const cpx<double>*)src, M, modes_of(p), tables_of(p), has_band, lo2, hi2,
                       alpha, beta);
  else
    hipLaunchKernelGGL((k_helmholtz<float>), nblocks(M), kOpNT, 0, s, (cpx<float>*)dst,
                       (const cpx<float>*)src, M, modes_of(p), tables_of(p), has_band, lo2, hi2,
                       (float)alpha, (float)beta);
  return check_launch("helmholtz");
}

gbg_status op_convolve_band(gbg_plan* p, void* dst, const void* src, int ncomp, double lo,
                            double hi, double alpha, double beta, hipStream_t s) {
  const unsigned M = (unsigned)p->M;
  const double lo2 = lo * lo, hi2 = hi * hi;
  if (p->elem == 8)
    hipLaunchKernelGGL((k_band<double>), nblocks(M), kOpNT, 0, s, (cpx<double>*)dst,
                       (const cpx<double>*)src, M, ncomp, modes_of(p), tables_of(p), lo2, hi2,
                       alpha, beta);
  else
    hipLaunchKernelGGL((k_band<float>), nblocks(M), kOpNT, 0, s, (cpx<float>*)dst,
                       (const cpx<float>*)src, M, ncomp, modes_of(p), tables_of(p), lo2, hi2,
                       (float)alpha, (float)beta);
  return check_launch("convolve band");
}

gbg_status op_mean_free(gbg_plan* p, void* spec, int ncomp, hipStream_t s) {
  if (p->elem == 8)
    hipLaunchKernelGGL((k_mean_free<double>), 1, 64, 0, s, (cpx<double>*)spec, p->M, ncomp);
  else
    hipLaunchKernelGGL((k_mean_free<float>), 1, 64, 0, s, (cpx<float>*)spec, p->M, ncomp);
  return check_launch("mean free");
}

gbg_status op_cross(gbg_plan* p, void* dst, const void* u, const void* v, hipStream_t s) {
  if (p->elem == 8)
    hipLaunchKernelGGL((k_cross<double>), nblocks(p->P), kOpNT, 0, s, (double*)dst,
                       (const double*)u, (const double*)v, p->P);
  else
    hipLaunchKernelGGL((k_cross<float>), nblocks(p->P), kOpNT, 0, s, (float*)dst, (const float*)u,
                       (const float*)v, p->P);
  return check_launch("cross");
}

gbg_status op_multiply(gbg_plan* p, void* dst, const void* a, const void* b, hipStream_t s) {
  if (p->elem == 8)
    hipLaunchKernelGGL((k_multiply<double>), nblocks_span<double>(p->P), kOpNT, 0, s, (double*)dst,
                       (const double*)a, (const double*)b, p->P);
  else
    hipLaunchKernelGGL((k_multiply<float>), nblocks_span<float>(p->P), kOpNT, 0, s, (float*)dst,
                       (const float*)a, (const float*)b, p->P);
  return check_launch("multiply");
}

gbg_status op_inner(gbg_plan* p, const void* f, const void* g, int ncomp, int slot, hipStream_t s) {
  const FoldW fw = fold_weights(p);
  if (p->elem == 8)
    hipLaunchKernelGGL((k_inner_part<double>), kRedBlocks, kOpNT, 0, s, (const cpx<double>*)f,
                       (const cpx<double>*)g, p->M, ncomp, fw,
                       p->red_part + (size_t)slot * kRedBlocks);
  else
    hipLaunchKernelGGL((k_inner_part<float>), kRedBlocks, kOpNT, 0, s, (const cpx<float>*)f,
                       (const cpx<float>*)g, p->M, ncomp, fw,
                       p->red_part + (size_t)slot * kRedBlocks);
  gbg_status st = check_launch("inner partials");
  if (st) return st;
  const double N = (double)p->n[0] * (double)p->n[1] * (double)p->n[2];
  hipLaunchKernelGGL(k_inner_final, 1, kOpNT, 0, s, p->red_part + (size_t)slot * kRedBlocks,
                     kRedBlocks, N * N, p->red_out + slot);
  return check_launch("inner final");
}

gbg_status op_inner_physical(gbg_plan* p, const void* f, const void* g, int ncomp, int slot,
                             hipStream_t s) {
  const size_t total = p->P * (size_t)ncomp;
  if (p->elem == 8)
    hipLaunchKernelGGL((k_dot_part<double>), kRedBlocks, kOpNT, 0, s, (const double*)f,
                       (const double*)g, total, p->red_part + (size_t)slot * kRedBlocks);
  else
    hipLaunchKernelGGL((k_dot_part<float>), kRedBlocks, kOpNT, 0, s, (const float*)f,
                       (const float*)g, total, p->red_part + (size_t)slot * kRedBlocks);
  gbg_status st = check_launch("physical inner partials");
  if (st) return st;
  hipLaunchKernelGGL(k_inner_final, 1, kOpNT, 0, s, p->red_part + (size_t)slot * kRedBlocks,
                     kRedBlocks, (double)p->P, p->red_out + slot);
  return check_launch("physical inner final");
}

}  // namespace gbg
