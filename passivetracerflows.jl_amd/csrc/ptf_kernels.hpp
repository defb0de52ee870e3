// Kernel templates of libptf_hip.
//
// Spectral field: (nx/2+1, ny, nz) complex, x fastest.  Physical field:
// (nx, ny, nz) real.  Ensemble members of one size are stored back to back
// (member stride mstride) and indexed by the last grid dimension.
//
// Column kernels transform along y or z: a workgroup owns TX consecutive kx
// columns of one (y or z) line set, thread (c, j) = (t % TX, t / TX), so a
// wavefront reads TX contiguous complex values per line position.
// Row kernels transform along x with the half-length complex FFT of
// M = nx/2 points plus the real<->half-complex pre/post twiddle; a workgroup
// owns NL rows, thread (c, j) = (t / P, t % P) so lanes walk along a row.
#pragma once
#include "ptf_device.hpp"

namespace ptf {

enum {
  COL_PLAIN = 0,   // out0 = FFT_DIR(in0)
  COL_D1 = 1,      // out0 = IFFT(in0), out1 = IFFT(i k in0)                   (first inverse axis)
  COL_D2 = 2,      // out0 = IFFT(in0), out1 = IFFT(i k in0), out2 = IFFT(in1)  (3D y inverse)
  COL_STAGE = 3,   // FFT_fwd(in0) -> stage epilogue                            (last forward axis)
  COL_FLOWSYN = 4  // psi(t) = psi0 exp(-i w t): out0 = IFFT_y(-i l psi), out1 = IFFT_y(i kr psi)
};

template <typename T>
struct ColArgs {
  const cpx<T>* in0;
  const cpx<T>* in1;
  cpx<T>* out0;
  cpx<T>* out1;
  cpx<T>* out2;
  const cpx<T>* tw;     // table of length N * twstride
  int twstride;
  int ncols;            // nx/2 + 1
  long long stride;     // element stride along the transformed axis
  long long ostride;    // stride of blockIdx.y (the other non-x axis)
  long long mstride;    // member stride (blockIdx.z)
  double kscale;        // 2 pi / L along the transformed axis
  int axis;             // 1 = y, 2 = z  (for wavenumber indices in the epilogue)
  StageArgs<T> st;
  const cpx<double>* psih0;   // COL_FLOWSYN
  const double* omega;
  double t;
  double kscale_x;
};

template <typename T, int N, int R, int TX, int MODE, int DIR>
__global__ __launch_bounds__(TX*(N / R)) void col_kernel(ColArgs<T> a) {
  constexpr int P = N / R;
  using Lay = LayInterleaved<TX>;
  __shared__ cpx<T> lds[(N > R) ? N * TX : 1];
  const int t = threadIdx.x;
  const int c = t % TX;
  const int j = t / TX;
  const int kx = blockIdx.x * TX + c;
  const bool valid = kx < a.ncols;
  const long long o = blockIdx.y;
  const int member = blockIdx.z;
  const long long base = (long long)member * a.mstride + o * a.ostride + kx;
  const long long s = a.stride;
  const cpx<T> zero = {(T)0, (T)0};

  cpx<T> v[R];
  if constexpr (MODE == COL_FLOWSYN) {
    cpx<T> w[R];
#pragma unroll
    for (int q = 0; q < R; ++q) {
      const int iy = j + q * P;
      if (valid) {
        const cpx<double> p0 = a.psih0[base + iy * s];
        const double ph = -a.omega[base + iy * s] * a.t;
        double sn, cs;
        sincos(ph, &sn, &cs);
        const cpx<double> p = {p0.re * cs - p0.im * sn, p0.re * sn + p0.im * cs};
        const double l = fftfreq_index(iy, N) * a.kscale;
        const double kr = kx * a.kscale_x;
        // -i l p = (l p.im, -l p.re);  i kr p = (-kr p.im, kr p.re)
        v[q] = {(T)(l * p.im), (T)(-(l * p.re))};
        w[q] = {(T)(-(kr * p.im)), (T)(kr * p.re)};
      } else {
        v[q] = zero;
        w[q] = zero;
      }
    }
    line_fft<T, N, R, +1, Lay>(v, lds, c, j, a.tw, a.twstride);
    if (valid) {
#pragma unroll
      for (int q = 0; q < R; ++q) a.out0[base + (j + q * P) * s] = v[q];
    }
    line_fft<T, N, R, +1, Lay>(w, lds, c, j, a.tw, a.twstride);
    if (valid) {
#pragma unroll
      for (int q = 0; q < R; ++q) a.out1[base + (j + q * P) * s] = w[q];
    }
    return;
  } else {
#pragma unroll
    for (int q = 0; q < R; ++q) v[q] = valid ? a.in0[base + (j + q * P) * s] : zero;

    if constexpr (MODE == COL_PLAIN) {
      line_fft<T, N, R, DIR, Lay>(v, lds, c, j, a.tw, a.twstride);
      if (valid) {
#pragma unroll
        for (int q = 0; q < R; ++q) a.out0[base + (j + q * P) * s] = v[q];
      }
    } else if constexpr (MODE == COL_D1 || MODE == COL_D2) {
      cpx<T> w[R];
#pragma unroll
      for (int q = 0; q < R; ++q) {
        const T k = (T)(fftfreq_index(j + q * P, N) * a.kscale);
        w[q] = times_ik(v[q], k);
      }
      line_fft<T, N, R, +1, Lay>(v, lds, c, j, a.tw, a.twstride);
      if (valid) {
#pragma unroll
        for (int q = 0; q < R; ++q) a.out0[base + (j + q * P) * s] = v[q];
      }
      line_fft<T, N, R, +1, Lay>(w, lds, c, j, a.tw, a.twstride);
      if (valid) {
#pragma unroll
        for (int q = 0; q < R; ++q) a.out1[base + (j + q * P) * s] = w[q];
      }
      if constexpr (MODE == COL_D2) {
#pragma unroll
        for (int q = 0; q < R; ++q) v[q] = valid ? a.in1[base + (j + q * P) * s] : zero;
        line_fft<T, N, R, +1, Lay>(v, lds, c, j, a.tw, a.twstride);
        if (valid) {
#pragma unroll
          for (int q = 0; q < R; ++q) a.out2[base + (j + q * P) * s] = v[q];
        }
      }
    } else if constexpr (MODE == COL_STAGE) {
      line_fft<T, N, R, -1, Lay>(v, lds, c, j, a.tw, a.twstride);
      if (valid) {
#pragma unroll
        for (int q = 0; q < R; ++q) {
          const int ia = j + q * P;
          const long long idx = base + ia * s;
          if (a.st.kind == EPI_STORE) {
            a.out0[idx] = v[q];
          } else if (a.axis == 2) {
            stage_epilogue(a.st, idx, member, kx, (int)o, ia, v[q]);
          } else {
            stage_epilogue(a.st, idx, member, kx, ia, (int)o, v[q]);
          }
        }
      }
    }
  }
}

// --------------------------------------------------------------------------
// row (x) kernels
// --------------------------------------------------------------------------
enum {
  ROW_C2R = 0,    // physical row = C2R(spectral row) * scale
  ROW_R2C = 1,    // spectral row = R2C(physical row)
  ROW_ADV = 2     // spectral row = R2C( -sum_d flow_d * C2R(D_d in_d) ), D_0 = i kr, D_d = 1
};

template <typename T>
struct RowArgs {
  const cpx<T>* in[3];     // spectral inputs (ROW_C2R / ROW_ADV)
  const T* flow[3];        // physical flows (ROW_ADV)
  const T* rin;            // physical input (ROW_R2C)
  T* rout;                 // physical output (ROW_C2R)
  cpx<T>* out;             // spectral output (ROW_R2C / ROW_ADV with EPI_STORE)
  const cpx<T>* tw;        // exp(-2 pi i m / nx), m < nx  (length nx * twstride)
  int twstride;
  long long nrows;         // rows per member (ny * nz)
  long long smstride;      // spectral member stride
  long long pmstride;      // physical member stride
  double scale;            // C2R scale (1 / (nx ny nz))
  double kscale_x;
  int ny;                  // for the epilogue's (iky, ikz) of a row
  StageArgs<T> st;         // ROW_ADV epilogue (1D problems)
};

// C2R pre-processing of one element pair: Z[k] from X[k], X[M-k] (k in [0, M)),
// Z = s [(X_k + conj X_{M-k}) + i conj(W^k) (X_k - conj X_{M-k})], W = exp(-2 pi i/nx);
// only the real parts of X_0 and X_M are used.
template <typename T>
__device__ __forceinline__ cpx<T> c2r_pre(cpx<T> Xk, cpx<T> Xm, cpx<T> w, T s, bool k0) {
  if (k0) {
    Xk.im = (T)0;
    Xm.im = (T)0;
  }
  const cpx<T> Xmc = cconj(Xm);
  const cpx<T> e = Xk + Xmc;
  const cpx<T> d = Xk - Xmc;
  const cpx<T> wc = cconj(w);
  const cpx<T> od = cmul(wc, d);
  // e + i od
  return {s * (e.re - od.im), s * (e.im + od.re)};
}

// R2C post-processing: X[k] = 1/2 [(Z_k + conj Z_{M-k}) - i W^k (Z_k - conj Z_{M-k})]
template <typename T>
__device__ __forceinline__ cpx<T> r2c_post(cpx<T> Zk, cpx<T> Zm, cpx<T> w) {
  const cpx<T> Zmc = cconj(Zm);
  const cpx<T> e = Zk + Zmc;
  const cpx<T> d = Zk - Zmc;
  const cpx<T> od = cmul(w, d);
  // e - i od = (e.re + od.im, e.im - od.re)
  return {(T)0.5 * (e.re + od.im), (T)0.5 * (e.im - od.re)};
}

template <typename T, int NX, int R, int NL, int MODE, int NIN>
__global__ __launch_bounds__(NL*((NX / 2) / R)) void row_kernel(RowArgs<T> a) {
  constexpr int M = NX / 2;
  constexpr int NXH = M + 1;
  constexpr int P = M / R;
  constexpr int PAD = 1;
  using Lay = LayRows<M, PAD>;
  __shared__ cpx<T> lds[NL * (M + PAD)];
  const int t = threadIdx.x;
  const int j = t % P;
  const int c = t / P;
  const long long row = (long long)blockIdx.x * NL + c;
  const bool valid = row < a.nrows;
  const int member = blockIdx.y;
  const long long sbase = (long long)member * a.smstride + row * NXH;
  const long long pbase = (long long)member * a.pmstride + row * NX;
  const cpx<T> zero = {(T)0, (T)0};

  if constexpr (MODE == ROW_C2R) {
    cpx<T> z[R];
    const T s = (T)a.scale;
#pragma unroll
    for (int q = 0; q < R; ++q) {
      const int k = j + q * P;
      if (valid) {
        const cpx<T> Xk = a.in[0][sbase + k];
        const cpx<T> Xm = a.in[0][sbase + (M - k)];
        z[q] = c2r_pre(Xk, Xm, a.tw[k * a.twstride], s, k == 0);
      } else {
        z[q] = zero;
      }
    }
    line_fft<T, M, R, +1, Lay>(z, lds, c, j, a.tw, 2 * a.twstride);
    if (valid) {
#pragma unroll
      for (int q = 0; q < R; ++q) {
        const int m = j + q * P;
        using T2 = cpx<T>;
        *reinterpret_cast<T2*>(a.rout + pbase + 2 * m) = z[q];
      }
    }
    return;
  } else {
    cpx<T> f[R];
    if constexpr (MODE == ROW_R2C) {
#pragma unroll
      for (int q = 0; q < R; ++q) {
        const int m = j + q * P;
        f[q] = valid ? *reinterpret_cast<const cpx<T>*>(a.rin + pbase + 2 * m) : zero;
      }
    } else {
      // ROW_ADV: f = ((-u cx) - v cy) - w cz, cx = C2R(i kr Y0), cy = C2R(Y1), cz = C2R(Y2)
      const T s = (T)a.scale;
#pragma unroll
      for (int d = 0; d < NIN; ++d) {
        cpx<T> z[R];
#pragma unroll
        for (int q = 0; q < R; ++q) {
          const int k = j + q * P;
          if (valid) {
            cpx<T> Xk = a.in[d][sbase + k];
            cpx<T> Xm = a.in[d][sbase + (M - k)];
            if (d == 0) {
              Xk = times_ik(Xk, (T)(k * a.kscale_x));
              Xm = times_ik(Xm, (T)((M - k) * a.kscale_x));
            }
            z[q] = c2r_pre(Xk, Xm, a.tw[k * a.twstride], s, k == 0);
          } else {
            z[q] = zero;
          }
        }
        line_fft<T, M, R, +1, Lay>(z, lds, c, j, a.tw, 2 * a.twstride);
#pragma unroll
        for (int q = 0; q < R; ++q) {
          const int m = j + q * P;
          const cpx<T> uu = valid ? *reinterpret_cast<const cpx<T>*>(a.flow[d] + pbase + 2 * m) : zero;
          if (d == 0) {
            f[q] = {-uu.re * z[q].re, -uu.im * z[q].im};
          } else {
            f[q] = {f[q].re - uu.re * z[q].re, f[q].im - uu.im * z[q].im};
          }
        }
      }
    }
    // forward half-length FFT of f[2m] + i f[2m+1]
    line_fft<T, M, R, -1, Lay>(f, lds, c, j, a.tw, 2 * a.twstride);
    // exchange Z[M-k] through LDS
    __syncthreads();
#pragma unroll
    for (int q = 0; q < R; ++q) lds[Lay::idx(c, j + q * P)] = f[q];
    __syncthreads();
    if (valid) {
#pragma unroll
      for (int q = 0; q < R; ++q) {
        const int k = j + q * P;
        cpx<T> X;
        if (k == 0) {
          X = {f[q].re + f[q].im, (T)0};
        } else {
          const cpx<T> Zm = lds[Lay::idx(c, M - k)];
          X = r2c_post(f[q], Zm, a.tw[k * a.twstride]);
        }
        if constexpr (MODE == ROW_ADV) {
          if (a.st.kind != EPI_STORE) {
            const int iky = (int)(row % a.ny), ikz = (int)(row / a.ny);
            stage_epilogue(a.st, sbase + k, member, k, iky, ikz, X);
          } else {
            a.out[sbase + k] = X;
          }
        } else {
          a.out[sbase + k] = X;
        }
        if (k == 0) {
          const cpx<T> XM = {f[q].re - f[q].im, (T)0};
          if constexpr (MODE == ROW_ADV) {
            if (a.st.kind != EPI_STORE) {
              const int iky = (int)(row % a.ny), ikz = (int)(row / a.ny);
              stage_epilogue(a.st, sbase + M, member, M, iky, ikz, XM);
            } else {
              a.out[sbase + M] = XM;
            }
          } else {
            a.out[sbase + M] = XM;
          }
        }
      }
    }
  }
}

}  // namespace ptf
