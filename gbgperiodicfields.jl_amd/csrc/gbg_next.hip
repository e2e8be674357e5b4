// The operators either side of the nonlinear term: the exponential
// (integrating-factor) kernel, the shell-binned radial spectrum, and the
// reality condition on the self-conjugate planes.  One streaming kernel per
// operation; wavenumbers from the plan's FP64 tables.
#include <cmath>

#include "gbg_internal.hpp"

namespace gbg {

constexpr int kNT = 256;

static unsigned grid_of(size_t n) { return (unsigned)((n + kNT - 1) / kNT); }

// ---- exponential kernel ---------------------------------------------------------------
// factor = exp(c |k|^2 + s w) per mode, |k|^2 = (k1^2 + k2^2) + k3^2 from the squared
// family (_exponential_exponent, src/Convolution.jl:1289-1294), evaluated in FP64; a
// complex offset w gives exp(re) (cos im + i sin im).  The factor is formed once per mode
// and applied to every component (_exponential_convolution_kernel!, :1326-1346).
struct ExpArgs {
  double coefficient, offset_scale, alpha, beta;
};

template <typename T>
__device__ __forceinline__ void exp_factor(const Tables& t, unsigned idx, unsigned nh, unsigned ny,
                                           const cpx<T>* __restrict__ offset, const ExpArgs& e,
                                           double& fr, double& fi) {
  const unsigned kx = idx % nh, r = idx / nh, ky = r % ny, kz = r / ny;
  const double k2 = (t.ks1[kx] + t.ks2[ky]) + t.ks3[kz];
  double re = e.coefficient * k2, im = 0.0;
  if (offset) {
    const cpx<T> w = offset[idx];
    re += e.offset_scale * (double)w.x;
    im = e.offset_scale * (double)w.y;
  }
  const double mag = exp(re);
  if (im == 0.0) {
    fr = mag;
    fi = 0.0;
  } else {
    double sn, cs;
    sincos(im, &sn, &cs);
    fr = mag * cs;
    fi = mag * sn;
  }
}

// dst = (alpha factor) f (+ beta dst when beta != 0)
template <typename T>
__global__ __launch_bounds__(kNT) void k_exponential(cpx<T>* dst, const cpx<T>* src, unsigned M,
                                                     int ncomp, unsigned nh, unsigned ny, Tables t,
                                                     const cpx<T>* offset, ExpArgs e) {
  const unsigned idx = blockIdx.x * kNT + threadIdx.x;
  if (idx >= M) return;
  double fr, fi;
  exp_factor<T>(t, idx, nh, ny, offset, e, fr, fi);
  const cpx<T> sf = mk<T>(T(e.alpha * fr), T(e.alpha * fi));
  const T beta = T(e.beta);
  for (int c = 0; c < ncomp; ++c) {
    const size_t i = (size_t)c * M + idx;
    const cpx<T> v = src[i];
    cpx<T> r = fi == 0.0 ? cscale(v, sf.x) : cmul(sf, v);
    if (e.beta != 0.0) r = cadd(r, cscale(dst[i], beta));
    dst[i] = r;
  }
}

// dst = factor (alpha a + beta b)   (_exponential_combination_kernel!, :1349-1362)
template <typename T>
__global__ __launch_bounds__(kNT) void k_exponential_combine(cpx<T>* dst, const cpx<T>* a,
                                                             const cpx<T>* b, unsigned M, int ncomp,
                                                             unsigned nh, unsigned ny, Tables t,
                                                             const cpx<T>* offset, ExpArgs e) {
  const unsigned idx = blockIdx.x * kNT + threadIdx.x;
  if (idx >= M) return;
  double fr, fi;
  exp_factor<T>(t, idx, nh, ny, offset, e, fr, fi);
  const cpx<T> f = mk<T>(T(fr), T(fi));
  const T al = T(e.alpha), be = T(e.beta);
  for (int c = 0; c < ncomp; ++c) {
    const size_t i = (size_t)c * M + idx;
    const cpx<T> comb = cadd(cscale(a[i], al), cscale(b[i], be));
    dst[i] = fi == 0.0 ? cscale(comb, f.x) : cmul(f, comb);
  }
}

// ---- radial spectrum -----------------------------------------------------------------
// Shell m = round(|k| / dk) (ties to even), Parseval weight 1 on the kx = 0 and even-Nx
// Nyquist planes, 2 elsewhere, density sum_c |f_c|^2 or Re(f_c conj g_c)
// (_spectral_shell_bin_kernel!, src/Reductions.jl:1309-1327).  Deterministic FP64 sums
// instead of the reference's atomics: along an x row |k| grows with kx, so a wave reads 64
// consecutive kx of a row, adds up equal-shell runs with a segmented scan, and the run ends
// add into the wave's own LDS histogram (distinct shells, rows in a fixed order).  When y
// or z is the folded axis the x row holds every frequency; lane j then reads kx index
// 0, 1, n-1, 2, n-2, ... (frequencies 0, 1, -1, 2, -2, ...), along which |k| grows too.  The
// workgroup writes the sum of its four histograms to part[shell][workgroup], a second
// kernel adds each shell's row in a fixed order.
constexpr int kRadialBlocks = 4096;
constexpr int kRadialMaxShells = 2048;   // 4 waves x 2048 x 8 B = 64 KiB of LDS

template <typename T>
__global__ __launch_bounds__(kNT) void k_radial_bin(const cpx<T>* __restrict__ f,
                                                    const cpx<T>* __restrict__ g, size_t M,
                                                    int ncomp, unsigned nh, unsigned ny,
                                                    unsigned nrows, int fold, int nyq, Tables t,
                                                    double dk, int nsh, double* __restrict__ part) {
  extern __shared__ double hist[];   // [4][nsh]
  for (int i = threadIdx.x; i < 4 * nsh; i += kNT) hist[i] = 0.0;
  __syncthreads();
  const int w = threadIdx.x / 64, lane = threadIdx.x % 64;
  double* h = hist + w * nsh;
  const unsigned nw = gridDim.x * 4;
  for (unsigned r = blockIdx.x * 4 + w; r < nrows; r += nw) {
    const unsigned ky = r % ny, kz = r / ny;
    const double s2 = t.ks2[ky], s3 = t.ks3[kz];
    // the Parseval weight of a row when y or z is the folded axis
    const int kf = fold == 1 ? (int)ky : (int)kz;
    const double wrow = (kf == 0 || kf == nyq) ? 1.0 : 2.0;
    const size_t base = (size_t)r * nh;
    for (unsigned c0 = 0; c0 < nh; c0 += 64) {
      const unsigned j = c0 + lane;
      const unsigned kx = fold == 0 || j == 0 ? j : ((j & 1u) ? (j + 1) / 2 : nh - j / 2);
      const bool valid = j < nh;
      int sh = -1;
      double v = 0.0;
      if (valid) {
        const double k2 = (t.ks1[kx] + s2) + s3;
        sh = (int)rint(sqrt(k2) / dk);
        double acc = 0.0;
        for (int c = 0; c < ncomp; ++c) {
          const cpx<T> a = f[(size_t)c * M + base + kx];
          if (g) {
            const cpx<T> b = g[(size_t)c * M + base + kx];
            acc += (double)a.x * (double)b.x + (double)a.y * (double)b.y;
          } else {
            acc += (double)a.x * (double)a.x + (double)a.y * (double)a.y;
          }
        }
        v = (fold == 0 ? (((int)kx == 0 || (int)kx == nyq) ? 1.0 : 2.0) : wrow) * acc;
      }
      // segmented inclusive scan over runs of equal shell (runs are contiguous)
#pragma unroll
      for (int d = 1; d < 64; d <<= 1) {
        const double ov = __shfl_up(v, d, 64);
        const int os = __shfl_up(sh, d, 64);
        if (lane >= d && os == sh) v += ov;
      }
      const int next = __shfl_down(sh, 1, 64);
      if (sh >= 0 && (lane == 63 || next != sh)) h[sh] += v;
    }
  }
  __syncthreads();
  for (int s = threadIdx.x; s < nsh; s += kNT)
    part[(size_t)s * gridDim.x + blockIdx.x] =
        ((hist[s] + hist[nsh + s]) + hist[2 * nsh + s]) + hist[3 * nsh + s];
}

// one workgroup per shell: strided partial sums, then a fixed tree
__global__ __launch_bounds__(kNT) void k_radial_final(const double* __restrict__ part, int nparts,
                                                      double norm, double* __restrict__ out) {
  __shared__ double red[kNT];
  const double* row = part + (size_t)blockIdx.x * nparts;
  double acc = 0.0;
  for (int b = threadIdx.x; b < nparts; b += kNT) acc += row[b];
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int k = kNT / 2; k > 0; k >>= 1) {
    if (threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[blockIdx.x] = red[0] / norm;
}

// ---- reality condition ---------------------------------------------------------------
// Planes of the folded axis f: k_f = 0 and, for even N_f, its Nyquist plane.  Within a
// plane, spanned by the other two axes a and b, the entry (j, l) pairs with
// ((-j) mod na, (-l) mod nb).  Work item = (j, l, component, plane) as in
// _reality_projection_kernel! (src/Fields.jl:2974-2990); the owner of a pair (j < j', or
// equal and l <= l') reads both members and writes both (_project_pair!, :2949-2958).
// Strides are in entries of one component (x folded: a = y, b = z, sf = 1).
struct Planes {
  unsigned na, nb;
  size_t sa, sb, sf;
  unsigned last;   // index of the second plane along the folded axis
  int nplanes;
};

template <typename T>
__global__ __launch_bounds__(kNT) void k_reality_project(cpx<T>* data, size_t M, Planes pl,
                                                         int ncomp) {
  const size_t item = (size_t)blockIdx.x * kNT + threadIdx.x;
  const size_t per_plane = (size_t)pl.na * pl.nb * ncomp;
  if (item >= per_plane * pl.nplanes) return;
  size_t rest = item;
  const unsigned j = rest % pl.na;
  rest /= pl.na;
  const unsigned l = rest % pl.nb;
  rest /= pl.nb;
  const unsigned c = rest % ncomp;
  const size_t p = (rest / ncomp) == 0 ? 0 : (size_t)pl.last * pl.sf;
  const unsigned jp = (pl.na - j) % pl.na, lp = (pl.nb - l) % pl.nb;
  if (!(j < jp || (j == jp && l <= lp))) return;
  const size_t off = (size_t)c * M + p;
  const size_t ix = off + pl.sa * j + pl.sb * l;
  const size_t iy = off + pl.sa * jp + pl.sb * lp;
  const cpx<T> x = data[ix];
  const cpx<T> y = ix == iy ? x : data[iy];
  const cpx<T> m = mk<T>((x.x + y.x) / T(2), (x.y - y.y) / T(2));
  data[ix] = m;
  if (ix != iy) data[iy] = mk<T>(m.x, -m.y);
}

// largest |f(k) - conj f(-k)| over the planes (reality_residual, src/Fields.jl:2858-2869)
template <typename T>
__global__ __launch_bounds__(kNT) void k_reality_residual(const cpx<T>* __restrict__ data, size_t M,
                                                          Planes pl, int ncomp,
                                                          double* __restrict__ part) {
  __shared__ double red[kNT];
  const size_t total = (size_t)pl.na * pl.nb * ncomp * pl.nplanes;
  double mx = 0.0;
  for (size_t item = (size_t)blockIdx.x * kNT + threadIdx.x; item < total;
       item += (size_t)gridDim.x * kNT) {
    size_t rest = item;
    const unsigned j = rest % pl.na;
    rest /= pl.na;
    const unsigned l = rest % pl.nb;
    rest /= pl.nb;
    const unsigned c = rest % ncomp;
    const size_t p = (rest / ncomp) == 0 ? 0 : (size_t)pl.last * pl.sf;
    const unsigned jp = (pl.na - j) % pl.na, lp = (pl.nb - l) % pl.nb;
    const size_t off = (size_t)c * M + p;
    const cpx<T> x = data[off + pl.sa * j + pl.sb * l];
    const cpx<T> y = data[off + pl.sa * jp + pl.sb * lp];
    const T dx = x.x - y.x, dy = x.y + y.y;   // f(k) - conj f(-k), in T
    const double a = sqrt((double)dx * (double)dx + (double)dy * (double)dy);
    mx = a > mx ? a : mx;
  }
  red[threadIdx.x] = mx;
  __syncthreads();
  for (int s = kNT / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s && red[threadIdx.x + s] > red[threadIdx.x]) red[threadIdx.x] = red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

__global__ void k_max_final(const double* __restrict__ part, int n, double* __restrict__ out) {
  __shared__ double red[kNT];
  double mx = 0.0;
  for (int i = threadIdx.x; i < n; i += kNT) mx = part[i] > mx ? part[i] : mx;
  red[threadIdx.x] = mx;
  __syncthreads();
  for (int s = kNT / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s && red[threadIdx.x + s] > red[threadIdx.x]) red[threadIdx.x] = red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = red[0];
}

// ---- host launchers ------------------------------------------------------------------
gbg_status op_exponential(gbg_plan* p, void* dst, const void* src, int ncomp, double coefficient,
                          const void* offset, double offset_scale, double alpha, double beta,
                          hipStream_t s) {
  const unsigned M = (unsigned)p->M;
  const ExpArgs e{coefficient, offset_scale, alpha, beta};
  if (p->elem == 8)
    hipLaunchKernelGGL((k_exponential<double>), grid_of(M), kNT, 0, s, (cpx<double>*)dst,
                       (const cpx<double>*)src, M, ncomp, (unsigned)p->s[0], (unsigned)p->s[1],
                       tables_of(p), (const cpx<double>*)offset, e);
  else
    hipLaunchKernelGGL((k_exponential<float>), grid_of(M), kNT, 0, s, (cpx<float>*)dst,
                       (const cpx<float>*)src, M, ncomp, (unsigned)p->s[0], (unsigned)p->s[1],
                       tables_of(p), (const cpx<float>*)offset, e);
  return check_launch("exponential convolve");
}

gbg_status op_exponential_combine(gbg_plan* p, void* dst, double alpha, const void* a, double beta,
                                  const void* b, int ncomp, double coefficient, const void* offset,
                                  double offset_scale, hipStream_t s) {
  const unsigned M = (unsigned)p->M;
  const ExpArgs e{coefficient, offset_scale, alpha, beta};
  if (p->elem == 8)
    hipLaunchKernelGGL((k_exponential_combine<double>), grid_of(M), kNT, 0, s, (cpx<double>*)dst,
                       (const cpx<double>*)a, (const cpx<double>*)b, M, ncomp, (unsigned)p->s[0],
                       (unsigned)p->s[1], tables_of(p), (const cpx<double>*)offset, e);
  else
    hipLaunchKernelGGL((k_exponential_combine<float>), grid_of(M), kNT, 0, s, (cpx<float>*)dst,
                       (const cpx<float>*)a, (const cpx<float>*)b, M, ncomp, (unsigned)p->s[0],
                       (unsigned)p->s[1], tables_of(p), (const cpx<float>*)offset, e);
  return check_launch("exponential combine");
}

void radial_shells(const gbg_plan* p, int* nshells, double* dk) {
  // the largest squared wavenumber per axis: folded axis at its last stored index,
  // full axes at index (n - 1) / 2 or, for even n, n / 2 (the Nyquist, k^2 true there)
  double corner = 0.0;
  for (int a = 0; a < 3; ++a) {
    const int n = p->n[a];
    const int f = n / 2;   // |frequency| of the largest |k| on either family
    const double k = (2.0 * M_PI) / p->L[a] * (double)f;
    corner += k * k;
  }
  const double Lmax = p->L[0] > p->L[1] ? (p->L[0] > p->L[2] ? p->L[0] : p->L[2])
                                        : (p->L[1] > p->L[2] ? p->L[1] : p->L[2]);
  const double d = (2.0 * M_PI) / Lmax;
  *dk = d;
  *nshells = (int)std::nearbyint(std::sqrt(corner) / d) + 1;
}

gbg_status op_radial_spectrum(gbg_plan* p, const void* f, const void* g, int ncomp, int nsh,
                              double* out_dev, double* part, hipStream_t s) {
  int n = 0;
  double dk = 0.0;
  radial_shells(p, &n, &dk);
  if (nsh != n) return fail(GBG_E_DIM, "nshells does not match the plan's shell count");
  if (nsh > kRadialMaxShells)
    return fail(GBG_E_UNSUPPORTED, "more than 2048 spectral shells (grid too anisotropic)");
  const unsigned nrows = (unsigned)(p->s[1] * p->s[2]);
  const unsigned nb = nrows / 4 + 1 < (unsigned)kRadialBlocks ? nrows / 4 + 1 : kRadialBlocks;
  const int nyq = (p->n[p->fold] % 2 == 0 && p->s[p->fold] > 1) ? p->s[p->fold] - 1 : -1;
  const size_t lds = sizeof(double) * 4 * (size_t)nsh;
  if (p->elem == 8)
    hipLaunchKernelGGL((k_radial_bin<double>), nb, kNT, lds, s, (const cpx<double>*)f,
                       (const cpx<double>*)g, p->M, ncomp, (unsigned)p->s[0], (unsigned)p->s[1],
                       nrows, p->fold, nyq, tables_of(p), dk, nsh, part);
  else
    hipLaunchKernelGGL((k_radial_bin<float>), nb, kNT, lds, s, (const cpx<float>*)f,
                       (const cpx<float>*)g, p->M, ncomp, (unsigned)p->s[0], (unsigned)p->s[1],
                       nrows, p->fold, nyq, tables_of(p), dk, nsh, part);
  gbg_status st = check_launch("radial spectrum bins");
  if (st) return st;
  const double N = (double)p->n[0] * (double)p->n[1] * (double)p->n[2];
  hipLaunchKernelGGL(k_radial_final, nsh, kNT, 0, s, part, (int)nb, N * N, out_dev);
  return check_launch("radial spectrum final");
}

size_t radial_work_bytes(const gbg_plan* p) {
  int n = 0;
  double dk = 0.0;
  radial_shells(p, &n, &dk);
  return sizeof(double) * ((size_t)kRadialBlocks * n + n);
}

// the self-conjugate planes of the plan's folded axis
static Planes reality_planes(const gbg_plan* p) {
  const int f = p->fold, a = f == 0 ? 1 : 0, b = f == 2 ? 1 : 2;
  size_t stride[3];
  stride[0] = 1;
  stride[1] = (size_t)p->s[0];
  stride[2] = (size_t)p->s[0] * p->s[1];
  Planes pl;
  pl.na = (unsigned)p->s[a];
  pl.nb = (unsigned)p->s[b];
  pl.sa = stride[a];
  pl.sb = stride[b];
  pl.sf = stride[f];
  pl.last = (unsigned)p->s[f] - 1;
  pl.nplanes = (p->n[f] % 2 == 0 && p->s[f] > 1) ? 2 : 1;
  return pl;
}

gbg_status op_reality_project(gbg_plan* p, void* spec, int ncomp, hipStream_t s) {
  const Planes pl = reality_planes(p);
  const size_t items = (size_t)pl.na * pl.nb * ncomp * pl.nplanes;
  if (p->elem == 8)
    hipLaunchKernelGGL((k_reality_project<double>), grid_of(items), kNT, 0, s, (cpx<double>*)spec,
                       p->M, pl, ncomp);
  else
    hipLaunchKernelGGL((k_reality_project<float>), grid_of(items), kNT, 0, s, (cpx<float>*)spec,
                       p->M, pl, ncomp);
  return check_launch("reality projection");
}

gbg_status op_reality_residual(gbg_plan* p, const void* spec, int ncomp, double* out_dev,
                               hipStream_t s) {
  const Planes pl = reality_planes(p);
  const size_t items = (size_t)pl.na * pl.nb * ncomp * pl.nplanes;
  const unsigned nb = grid_of(items) < (unsigned)kRedBlocks ? grid_of(items) : kRedBlocks;
  if (p->elem == 8)
    hipLaunchKernelGGL((k_reality_residual<double>), nb, kNT, 0, s, (const cpx<double>*)spec, p->M,
                       pl, ncomp, p->red_part);
  else
    hipLaunchKernelGGL((k_reality_residual<float>), nb, kNT, 0, s, (const cpx<float>*)spec, p->M,
                       pl, ncomp, p->red_part);
  gbg_status st = check_launch("reality residual");
  if (st) return st;
  hipLaunchKernelGGL(k_max_final, 1, kNT, 0, s, p->red_part, (int)nb, out_dev);
  return check_launch("reality residual final");
}

}  // namespace gbg
