// Setup, conversion and diagnostic kernels of libptf_hip (host runtime only).
#pragma once
#include "ptf_kernels.hpp"

namespace ptf {

// --------------------------------------------------------------------------
// setup, conversion and diagnostic kernels
// --------------------------------------------------------------------------

// ETDRK4 coefficients ([FF] getetdcoeffs(dt, L; ncirc=32, rcirc=1), getexpLs): FP64
__global__ void etd_coeff_kernel(const double* __restrict__ L, long long n, double dt, int ncirc,
                                 double rcirc, double* E, double* E2, double* zeta, double* alpha,
                                 double* beta, double* gam) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double zr = dt * L[i];
  double sz = 0, sa = 0, sb = 0, sg = 0;  // real parts of the contour sums
  for (int jj = 0; jj < ncirc; ++jj) {
    double sn, cs;
    sincospi(2.0 * (jj + 0.5) / ncirc, &sn, &cs);
    // z = zr + rcirc e^{i theta}
    const double xr = zr + rcirc * cs, xi = rcirc * sn;
    // e^z, e^{z/2}
    const double ea = exp(xr);
    double s1, c1;
    sincos(xi, &s1, &c1);
    const double ezr = ea * c1, ezi = ea * s1;
    const double eh = exp(xr / 2);
    double s2, c2;
    sincos(xi / 2, &s2, &c2);
    const double ehr = eh * c2, ehi = eh * s2;
    // 1/z, 1/z^3
    const double den = xr * xr + xi * xi;
    const double izr = xr / den, izi = -xi / den;
    const double z2r = xr * xr - xi * xi, z2i = 2 * xr * xi;
    const double iz2r = izr * izr - izi * izi, iz2i = 2 * izr * izi;
    const double iz3r = iz2r * izr - iz2i * izi, iz3i = iz2r * izi + iz2i * izr;
    // zeta: (e^{z/2} - 1)/z
    {
      const double nr = ehr - 1, ni = ehi;
      sz += nr * izr - ni * izi;
    }
    // alpha: (-4 - z + e^z (4 - 3z + z^2))/z^3
    {
      const double pr = 4 - 3 * xr + z2r, pi = -3 * xi + z2i;
      const double nr = -4 - xr + (ezr * pr - ezi * pi);
      const double ni = -xi + (ezr * pi + ezi * pr);
      sa += nr * iz3r - ni * iz3i;
    }
    // beta: (2 + z + e^z (-2 + z))/z^3
    {
      const double pr = -2 + xr, pi = xi;
      const double nr = 2 + xr + (ezr * pr - ezi * pi);
      const double ni = xi + (ezr * pi + ezi * pr);
      sb += nr * iz3r - ni * iz3i;
    }
    // Gamma: (-4 - 3z - z^2 + e^z (4 - z))/z^3
    {
      const double pr = 4 - xr, pi = -xi;
      const double nr = -4 - 3 * xr - z2r + (ezr * pr - ezi * pi);
      const double ni = -3 * xi - z2i + (ezr * pi + ezi * pr);
      sg += nr * iz3r - ni * iz3i;
    }
  }
  zeta[i] = dt * (sz / ncirc);
  alpha[i] = dt * (sa / ncirc);
  beta[i] = dt * (sb / ncirc);
  gam[i] = dt * (sg / ncirc);
  E[i] = exp(dt * L[i]);
  E2[i] = exp(dt * L[i] / 2);
}

template <typename TO, typename TI>
__global__ void convert_kernel(TO* __restrict__ out, const TI* __restrict__ in, long long n) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = (TO)in[i];
}

template <typename T>
__global__ void copy_kernel(T* __restrict__ out, const T* __restrict__ in, long long n) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = in[i];
}

// Neumaier-compensated FP64 partial sums of w_k |sol_k|^2 (Hermitian weights
// 1 at kx = 0 and nx/2, 2 elsewhere), one (sum, compensation) pair per block;
// the field may be in the brick layout (kx of element i is (i >> lb) mod nxh).
template <typename T>
__global__ void variance_partial_kernel(const cpx<T>* __restrict__ sol, const cpx<T>* __restrict__ sol_lo,
                                        long long n, int nxh, int lb, double* __restrict__ partial) {
  __shared__ double ssum[256];
  __shared__ double scomp[256];
  double sum = 0.0, comp = 0.0;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x) {
    const int kx = (int)((i >> lb) % nxh);   // brick layout: 2^lb planes innermost
    const double w = (kx == 0 || kx == nxh - 1) ? 1.0 : 2.0;
    // the mixed form's compensated state is sol + sol_lo, summed in FP64
    const cpx<T> v = sol[i];
    double vr = (double)v.re, vi = (double)v.im;
    if (sol_lo) {
      vr += (double)sol_lo[i].re;
      vi += (double)sol_lo[i].im;
    }
    const double x = w * (vr * vr + vi * vi);
    const double tt = sum + x;
    if (fabs(sum) >= fabs(x)) comp += (sum - tt) + x;
    else comp += (x - tt) + sum;
    sum = tt;
  }
  ssum[threadIdx.x] = sum;
  scomp[threadIdx.x] = comp;
  __syncthreads();
  for (int off = blockDim.x / 2; off > 0; off >>= 1) {
    if ((int)threadIdx.x < off) {
      const double a = ssum[threadIdx.x], b = ssum[threadIdx.x + off];
      const double tt = a + b;
      double e = (fabs(a) >= fabs(b)) ? (a - tt) + b : (b - tt) + a;
      ssum[threadIdx.x] = tt;
      scomp[threadIdx.x] += scomp[threadIdx.x + off] + e;
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    partial[2 * blockIdx.x] = ssum[0];
    partial[2 * blockIdx.x + 1] = scomp[0];
  }
}

// Element (row r, entry k) of a field in the working layout: rows are the (y, z)
// lines of x (r = y + ny z); with the last axis in bricks of 2^lb planes the rows
// are regrouped so that 2^lb consecutive planes sit innermost.
__device__ __forceinline__ long long work_index(long long r, long long k, long long rowlen, int lb,
                                                int ndim, int ny) {
  if (lb == 0) return r * rowlen + k;
  const long long B = 1ll << lb, mask = B - 1;
  const long long mid = ndim == 3 ? ny : 1;
  const long long l = r / mid, m = r % mid;
  const long long b = ((l >> lb) * mid + m) * B + (l & mask);
  return ((b >> lb) * rowlen + k) * B + (b & mask);
}

// Public FP64 layout (rows x rowlen entries of w doubles, x fastest) -> working
// layout in T; with dst_lo the FP64 value is split into hi + lo (compensated state).
template <typename T>
__global__ void layout_in_kernel(const double* __restrict__ src, T* __restrict__ dst,
                                 T* __restrict__ dst_lo, long long rows, int rowlen, int w, int lb,
                                 int ndim, int ny) {
  const long long n = rows * rowlen * w;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x) {
    const int e = (int)(i % w);
    const long long q = i / w;
    const long long k = q % rowlen, r = q / rowlen;
    const long long d = work_index(r, k, rowlen, lb, ndim, ny) * w + e;
    const double x = src[i];
    const T hi = (T)x;
    dst[d] = hi;
    if (dst_lo) dst_lo[d] = (T)(x - (double)hi);
  }
}

// Working layout (T, plus the low part when src_lo) -> public FP64 layout.
template <typename T>
__global__ void layout_out_kernel(const T* __restrict__ src, const T* __restrict__ src_lo,
                                  double* __restrict__ dst, long long rows, int rowlen, int w,
                                  int lb, int ndim, int ny) {
  const long long n = rows * rowlen * w;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x) {
    const int e = (int)(i % w);
    const long long q = i / w;
    const long long k = q % rowlen, r = q / rowlen;
    const long long d = work_index(r, k, rowlen, lb, ndim, ny) * w + e;
    double x = (double)src[d];
    if (src_lo) x += (double)src_lo[d];
    dst[i] = x;
  }
}

// Analytic separable flow at one time: out[c] at physical point (ix, iy, iz) is
// sum over the terms k of component c of ((amp[k] tx[k][ix]) ty[k][iy]) tz[k][iz]
// (amp already holds amp * T(omega t + phase)).  Values are stored in the
// library's physical layout: x-pairs, with the last axis in bricks of 2^lb rows.
constexpr int kMaxFlowTerms = 16;
template <typename T>
struct AnalyticArgs {
  int nterms;
  int comp[kMaxFlowTerms];
  double amp[kMaxFlowTerms];
  const double* tx;   // [nterms][nx]
  const double* ty;   // [nterms][ny]
  const double* tz;   // [nterms][nz]
  T* out[3];
  int ndim, nx, ny, nz, lb;
  long long npoints;
};

template <typename T>
__global__ void analytic_flow_kernel(AnalyticArgs<T> a) {
  const long long B = 1ll << a.lb, mask = B - 1;
  const long long mid = a.ndim == 3 ? a.ny : 1;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < a.npoints;
       i += (long long)gridDim.x * blockDim.x) {
    const int ix = (int)(i % a.nx);
    const long long r = i / a.nx;          // row: iy + ny iz
    const int iy = a.ndim >= 2 ? (int)(r % a.ny) : 0;
    const int iz = a.ndim == 3 ? (int)(r / a.ny) : 0;
    double v0 = 0.0, v1 = 0.0, v2 = 0.0;
    for (int k = 0; k < a.nterms; ++k) {
      double x = a.amp[k] * a.tx[(long long)k * a.nx + ix];
      if (a.ndim >= 2) x *= a.ty[(long long)k * a.ny + iy];
      if (a.ndim == 3) x *= a.tz[(long long)k * a.nz + iz];
      const int c = a.comp[k];
      if (c == 0) v0 += x;
      else if (c == 1) v1 += x;
      else v2 += x;
    }
    // element ix of row r: pair ix/2 (bricked rows), component ix & 1
    long long e;
    const long long rowlen = a.nx / 2;
    if (a.lb == 0) {
      e = r * rowlen + ix / 2;
    } else {
      const long long l = r / mid, m = r % mid;
      const long long b = ((l >> a.lb) * mid + m) * B + (l & mask);
      e = ((b >> a.lb) * rowlen + ix / 2) * B + (b & mask);
    }
    const long long o = 2 * e + (ix & 1);
    a.out[0][o] = (T)v0;
    if (a.ndim >= 2) a.out[1][o] = (T)v1;
    if (a.ndim == 3) a.out[2][o] = (T)v2;
  }
}

// dealias!: zero the spectral elements whose kr index lies in [rlo, rhi) or whose
// y (z) index lies in [ylo, yhi) ([zlo, zhi)); plain (unbricked) half-spectrum
// layout (nx/2+1, ny, nz, batch), x fastest.
template <typename T>
__global__ void dealias_kernel(cpx<T>* __restrict__ f, long long n, int nxh, int ny, int nz,
                               int rlo, int rhi, int ylo, int yhi, int zlo, int zhi) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x) {
    const int kx = (int)(i % nxh);
    const long long r = i / nxh;
    const int iy = (int)(r % ny);
    const int iz = (int)((r / ny) % nz);
    if ((kx >= rlo && kx < rhi) || (iy >= ylo && iy < yhi) || (iz >= zlo && iz < zhi))
      f[i] = cpx<T>{(T)0, (T)0};
  }
}

// Mass and variance record from the partial sums of variance_partial_kernel:
// rec[0] = Re sol[0] (mass = sum of c; sol + sol_lo in the mixed form), rec[1] = variance.  One thread sums the
// partials in order with Neumaier compensation, as the host does for a single
// diagnostics call, so both give the same value.
template <typename T>
__global__ void diag_record_kernel(const cpx<T>* __restrict__ sol, const cpx<T>* __restrict__ sol_lo,
                                   const double* __restrict__ partial, int npartial, double npoints,
                                   double* __restrict__ rec) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  double sum = 0.0, comp = 0.0;
  for (int i = 0; i < npartial; ++i) {
    const double x = partial[i];
    const double t = sum + x;
    comp += (fabs(sum) >= fabs(x)) ? (sum - t) + x : (x - t) + sum;
    sum = t;
  }
  double s0re = (double)sol[0].re, s0im = (double)sol[0].im;
  if (sol_lo) {
    s0re += (double)sol_lo[0].re;
    s0im += (double)sol_lo[0].im;
  }
  const double dc = s0re * s0re + s0im * s0im;
  rec[0] = s0re;
  rec[1] = ((sum - dc) + comp) / (npoints * npoints);
}

}  // namespace ptf
