// Device building blocks of libptf_hip (gfx950): complex arithmetic, radix-2/4/8/16
// register butterflies, and a Stockham autosort line FFT whose exchanges go
// through LDS.
//
// Register contract of line_fft: a line of length N is processed by P = N/R
// threads; thread j holds elements j + q*P (q = 0..R-1) on entry AND on exit
// ("natural" layout).  Because input and output positions of a thread
// coincide, a kernel can read a line, transform it, and write it back in place
// without a barrier, and an inverse FFT can feed a forward FFT with no data
// exchange (the fused advection kernels rely on this).
//
// Stockham formulation (Van Loan, "Computational Frameworks for the FFT",
// 1992, sect. 1.7): pass with radix Rp and span Ns (product of earlier radices):
//   butterfly b reads in[b + r*N/Rp], multiplies element r by
//   exp(DIR*2*pi*i*r*(b mod Ns)/(Ns*Rp)), does an Rp-point DFT and writes
//   out[(b / Ns)*Ns*Rp + (b mod Ns) + r*Ns].
// A thread holds E = R elements; every radix of the line's plan divides E, and a
// pass of radix Rp gives each thread the E/Rp butterflies b = j + s*P
// (s < E/Rp), whose inputs j + (s + r*E/Rp)*P are exactly the thread's own
// natural-layout registers.  So the first pass (Ns = 1, no twiddles) works on
// registers, and the last pass (Ns*Rp = N) writes its outputs back to the
// registers it read.  Powers of two use N = R0 * R^K (R0 <= R first); other
// lengths 2^a 3^b 5^c use radices 2, 3, 4, 5, 8, 16 that divide E.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace ptf {

template <typename T>
struct alignas(2 * sizeof(T)) cpx {
  T re, im;
};

template <typename T>
__device__ __forceinline__ cpx<T> operator+(cpx<T> a, cpx<T> b) { return {a.re + b.re, a.im + b.im}; }
template <typename T>
__device__ __forceinline__ cpx<T> operator-(cpx<T> a, cpx<T> b) { return {a.re - b.re, a.im - b.im}; }
template <typename T>
__device__ __forceinline__ cpx<T> cmul(cpx<T> a, cpx<T> b) {
  return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
template <typename T>
__device__ __forceinline__ cpx<T> cconj(cpx<T> a) { return {a.re, -a.im}; }
template <typename T>
__device__ __forceinline__ cpx<T> cscale(T s, cpx<T> a) { return {s * a.re, s * a.im}; }
// a * (i*k) for real k: (re + i im)(i k) = (-k im) + i (k re)
template <typename T>
__device__ __forceinline__ cpx<T> times_ik(cpx<T> a, T k) { return {-(k * a.im), k * a.re}; }

// multiply by DIR*i  (DIR = -1: forward kernel exp(-2 pi i/N); +1: inverse)
template <int DIR, typename T>
__device__ __forceinline__ cpx<T> mul_i(cpx<T> a) {
  if constexpr (DIR < 0) return {a.im, -a.re};
  else return {-a.im, a.re};
}

// exp(DIR * 2 pi i m / 16) * a, m a compile-time constant
template <int M, int DIR, typename T>
__device__ __forceinline__ cpx<T> rot16(cpx<T> a) {
  constexpr int m = M & 15;
  if constexpr (m == 0) {
    return a;
  } else if constexpr (m == 4) {
    return mul_i<DIR>(a);
  } else if constexpr (m == 8) {
    return {-a.re, -a.im};
  } else if constexpr (m == 12) {
    return mul_i<-DIR>(a);
  } else {
    // cos and sin of 2 pi m / 16
    constexpr double C1 = 0.92387953251128675613;  // cos(pi/8)
    constexpr double S1 = 0.38268343236508977173;  // sin(pi/8)
    constexpr double H = 0.70710678118654752440;   // cos(pi/4)
    constexpr double CT[16] = {1, C1, H, S1, 0, -S1, -H, -C1, -1, -C1, -H, -S1, 0, S1, H, C1};
    constexpr double ST[16] = {0, S1, H, C1, 1, C1, H, S1, 0, -S1, -H, -C1, -1, -C1, -H, -S1};
    const T c = (T)CT[m];
    const T s = (T)(DIR * ST[m]);
    return {a.re * c - a.im * s, a.re * s + a.im * c};
  }
}

template <int DIR, typename T>
__device__ __forceinline__ void dft2(cpx<T>& a0, cpx<T>& a1) {
  cpx<T> t = a0;
  a0 = t + a1;
  a1 = t - a1;
}

template <int DIR, typename T>
__device__ __forceinline__ void dft4(cpx<T>& a0, cpx<T>& a1, cpx<T>& a2, cpx<T>& a3) {
  cpx<T> t0 = a0 + a2, t1 = a0 - a2, t2 = a1 + a3, t3 = mul_i<DIR>(a1 - a3);
  a0 = t0 + t2;
  a2 = t0 - t2;
  a1 = t1 + t3;
  a3 = t1 - t3;
}

// 8-point DFT, natural order in and out: n = 2 n1 + n2, k = k1 + 4 k2
template <int DIR, typename T>
__device__ __forceinline__ void dft8(cpx<T>* a) {
  dft4<DIR>(a[0], a[2], a[4], a[6]);
  dft4<DIR>(a[1], a[3], a[5], a[7]);
  a[3] = rot16<2, DIR>(a[3]);
  a[5] = rot16<4, DIR>(a[5]);
  a[7] = rot16<6, DIR>(a[7]);
  cpx<T> o[8];
#pragma unroll
  for (int k1 = 0; k1 < 4; ++k1) {
    o[k1] = a[2 * k1] + a[2 * k1 + 1];
    o[k1 + 4] = a[2 * k1] - a[2 * k1 + 1];
  }
#pragma unroll
  for (int i = 0; i < 8; ++i) a[i] = o[i];
}

// 16-point DFT, natural order in and out: n = 4 n1 + n2, k = k1 + 4 k2
template <int DIR, typename T>
__device__ __forceinline__ void dft16(cpx<T>* a) {
  dft4<DIR>(a[0], a[4], a[8], a[12]);
  dft4<DIR>(a[1], a[5], a[9], a[13]);
  dft4<DIR>(a[2], a[6], a[10], a[14]);
  dft4<DIR>(a[3], a[7], a[11], a[15]);
  // a[4 k1 + n2] *= w16^(n2 k1)
  a[5] = rot16<1, DIR>(a[5]);
  a[6] = rot16<2, DIR>(a[6]);
  a[7] = rot16<3, DIR>(a[7]);
  a[9] = rot16<2, DIR>(a[9]);
  a[10] = rot16<4, DIR>(a[10]);
  a[11] = rot16<6, DIR>(a[11]);
  a[13] = rot16<3, DIR>(a[13]);
  a[14] = rot16<6, DIR>(a[14]);
  a[15] = rot16<9, DIR>(a[15]);
  dft4<DIR>(a[0], a[1], a[2], a[3]);
  dft4<DIR>(a[4], a[5], a[6], a[7]);
  dft4<DIR>(a[8], a[9], a[10], a[11]);
  dft4<DIR>(a[12], a[13], a[14], a[15]);
  // X[k1 + 4 k2] sits in a[4 k1 + k2]: transpose the 4x4 register block
  cpx<T> o[16];
#pragma unroll
  for (int k1 = 0; k1 < 4; ++k1)
#pragma unroll
    for (int k2 = 0; k2 < 4; ++k2) o[k1 + 4 * k2] = a[4 * k1 + k2];
#pragma unroll
  for (int i = 0; i < 16; ++i) a[i] = o[i];
}

// Prime-length DFTs (3, 5) by the symmetric pairs sp_n = a_n + a_{Q-n},
// dm_n = a_n - a_{Q-n}:
//   X_k     = a_0 + sum_n cos(2 pi n k/Q) sp_n + DIR i sum_n sin(2 pi n k/Q) dm_n,
//   X_{Q-k} = the same with the sine sum negated  (n, k = 1 .. (Q-1)/2).
template <typename T>
__device__ __forceinline__ cpx<T> axpy(T c, cpx<T> x, cpx<T> y) { return {y.re + c * x.re, y.im + c * x.im}; }

template <int DIR, typename T>
__device__ __forceinline__ void dft3(cpx<T>* a) {
  constexpr double S1 = 0.8660254037844386467637;   // sin(2 pi/3)
  const cpx<T> sp = a[1] + a[2], dm = a[1] - a[2];
  const cpx<T> x0 = a[0] + sp;
  const cpx<T> re = axpy((T)-0.5, sp, a[0]);
  const cpx<T> t = mul_i<DIR>(cscale((T)S1, dm));
  a[0] = x0;
  a[1] = re + t;
  a[2] = re - t;
}

template <int DIR, typename T>
__device__ __forceinline__ void dft5(cpx<T>* a) {
  constexpr double C1 = 0.3090169943749474241023, S1 = 0.9510565162951535721164;    // 2 pi/5
  constexpr double C2 = -0.8090169943749474241023, S2 = 0.5877852522924731291687;   // 4 pi/5
  const cpx<T> sp1 = a[1] + a[4], dm1 = a[1] - a[4];
  const cpx<T> sp2 = a[2] + a[3], dm2 = a[2] - a[3];
  const cpx<T> x0 = a[0] + sp1 + sp2;
  const cpx<T> re1 = axpy((T)C2, sp2, axpy((T)C1, sp1, a[0]));
  const cpx<T> im1 = axpy((T)S2, dm2, cscale((T)S1, dm1));
  const cpx<T> re2 = axpy((T)C1, sp2, axpy((T)C2, sp1, a[0]));
  const cpx<T> im2 = axpy((T)-S1, dm2, cscale((T)S2, dm1));
  const cpx<T> t1 = mul_i<DIR>(im1), t2 = mul_i<DIR>(im2);
  a[0] = x0;
  a[1] = re1 + t1;
  a[4] = re1 - t1;
  a[2] = re2 + t2;
  a[3] = re2 - t2;
}

template <int R, int DIR, typename T>
__device__ __forceinline__ void dft(cpx<T>* a) {
  if constexpr (R == 2) dft2<DIR>(a[0], a[1]);
  else if constexpr (R == 4) dft4<DIR>(a[0], a[1], a[2], a[3]);
  else if constexpr (R == 8) dft8<DIR>(a);
  else if constexpr (R == 16) dft16<DIR>(a);
  else if constexpr (R == 3) dft3<DIR>(a);
  else if constexpr (R == 5) dft5<DIR>(a);
  else static_assert(R == 2, "unsupported radix");
}

// Pass radices of a line of N points held E per thread (each dividing E, product N),
// smallest first; n = 0 when N cannot be split that way.  Greedy from the largest radix
// that divides both E and what is left: for powers of two with E = R this is
// R0 * R^K with the remainder R0 first.
struct RadixPlan {
  int n;
  int r[24];
};
constexpr RadixPlan radix_plan(int N, int E) {
  RadixPlan p{0, {}};
  if (N < 2 || E < 2 || N % E != 0) return p;
  const int pref[6] = {16, 8, 5, 4, 3, 2};
  int tmp[24] = {};
  int k = 0, rest = N;
  while (rest > 1) {
    int pick = 0;
    for (int i = 0; i < 6; ++i) {
      if (E % pref[i] == 0 && rest % pref[i] == 0) {
        pick = pref[i];
        break;
      }
    }
    if (pick == 0 || k == 24) return RadixPlan{0, {}};
    tmp[k++] = pick;
    rest /= pick;
  }
  for (int i = 0; i < k; ++i) p.r[i] = tmp[k - 1 - i];
  p.n = k;
  return p;
}
constexpr int radix_span(const RadixPlan& p, int pass) {   // product of the radices before `pass`
  int s = 1;
  for (int i = 0; i < pass; ++i) s *= p.r[i];
  return s;
}
constexpr bool fft_ok(int N, int E) { return radix_plan(N, E).n > 0; }
// workgroup barriers one line_fft<N, E> executes: one before the first pass's LDS writes,
// two per middle pass, one before the last pass's LDS reads (none for a single pass)
constexpr int fft_barriers(int N, int E) {
  const int n = radix_plan(N, E).n;
  return n >= 2 ? 2 * n - 2 : 0;
}

// Elements per thread of the one register layout compiled for a line length N that is
// not a power of two: N = 2^a 3^b 5^c with E = 2^min(a,2) times the product of the
// distinct odd primes (2^min(a,1) times it when both 3 and 5 divide N), at most 30;
// 0 when N has another prime factor or is a power of two.
constexpr int line_e(int N) {
  if (N < 2) return 0;
  int a = 0, n = N;
  while (n % 2 == 0) {
    n /= 2;
    ++a;
  }
  int odd = 1;
  const int primes[2] = {3, 5};
  for (int i = 0; i < 2; ++i) {
    if (n % primes[i] == 0) {
      odd *= primes[i];
      while (n % primes[i] == 0) n /= primes[i];
    }
  }
  if (n != 1 || odd == 1) return 0;
  const int e = (odd < 15 ? (1 << (a < 2 ? a : 2)) : (1 << (a < 1 ? a : 1))) * odd;
  return e <= 30 && fft_ok(N, e) ? e : 0;
}

// LDS layouts: line c, element i
// idx_split(c, i0, off) = idx(c, i0 + off) for a multiple of 8 `off` known at compile
// time (a layout may fold it into an immediate offset)
template <int NL>
struct LayInterleaved {  // lines interleaved: consecutive lanes = consecutive lines
  __device__ __forceinline__ int idx(int c, int i) const { return i * NL + c; }
  __device__ __forceinline__ int idx_split(int c, int i0, int off) const { return idx(c, i0 + off); }
};
struct LayStrided {      // runtime strides: line-major (rows contiguous) or element-major
  int sc, si;
  __device__ __forceinline__ int idx(int c, int i) const { return c * sc + i * si; }
  __device__ __forceinline__ int idx_split(int c, int i0, int off) const { return idx(c, i0 + off); }
};

// Orders a block's LDS accesses around an exchange: a workgroup barrier, or, when every
// line of the block lies within one wavefront (wave = true), only that wavefront's: a
// wavefront's LDS operations complete in issue order, so the fences (which keep the
// compiler from moving LDS accesses across) and the wavefront barrier suffice.
__device__ __forceinline__ void lds_sync(bool wave) {
  if (wave) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  } else {
    __syncthreads();
  }
}

// One pass of a line FFT (see the header): the E/RP butterflies of this thread on its
// natural-layout registers a (already read from LDS unless this is the first pass),
// outputs stored to LDS at their Stockham positions unless this is the last pass.
template <typename T, int N, int E, int DIR, int PASS, class Lay, bool TWH = false>
__device__ __forceinline__ void fft_pass(cpx<T> (&a)[E], cpx<T>* lds, const Lay& lay, int c, int j,
                                         const cpx<T>* __restrict__ tw, int twstride, bool wave) {
  constexpr RadixPlan pl = radix_plan(N, E);
  constexpr int P = N / E;
  constexpr int RP = pl.r[PASS];
  constexpr int B = E / RP;
  constexpr int Ns = radix_span(pl, PASS);
  constexpr bool kLast = PASS == pl.n - 1;
  static_assert(E % RP == 0, "a radix must divide the elements per thread");
  if constexpr (PASS > 0) {
    lds_sync(wave);
#pragma unroll
    for (int q = 0; q < E; ++q) {
      if constexpr (P % 8 == 0) a[q] = lds[lay.idx_split(c, j, q * P)];
      else a[q] = lds[lay.idx(c, j + q * P)];
    }
  }
#pragma unroll
  for (int s = 0; s < B; ++s) {
    cpx<T> t[RP];
#pragma unroll
    for (int r = 0; r < RP; ++r) t[r] = a[s + r * B];
    if constexpr (PASS > 0) {
      const int m = (j + s * P) % Ns;
      const int step = (N / (Ns * RP)) * twstride;
#pragma unroll
      for (int r = 1; r < RP; ++r) {
        cpx<T> w;
        if constexpr (TWH) {
          // half table: tw[k] for k < LT/2 (LT = N twstride), the rest is -tw[k - LT/2]
          const int k = r * m * step;
          const int half = (N / 2) * twstride;
          const bool hi = k >= half;
          w = tw[hi ? k - half : k];
          if (hi) w = {-w.re, -w.im};
        } else {
          w = tw[r * m * step];
        }
        if constexpr (DIR > 0) w.im = -w.im;
        t[r] = cmul(t[r], w);
      }
    }
    dft<RP, DIR>(t);
#pragma unroll
    for (int r = 0; r < RP; ++r) a[s + r * B] = t[r];
  }
  if constexpr (!kLast) {
    lds_sync(wave);
#pragma unroll
    for (int s = 0; s < B; ++s) {
      const int b = j + s * P;
      const int base = (b / Ns) * Ns * RP + b % Ns;
#pragma unroll
      for (int r = 0; r < RP; ++r) {
        if constexpr (Ns % 8 == 0) lds[lay.idx_split(c, base, r * Ns)] = a[s + r * B];
        else lds[lay.idx(c, base + r * Ns)] = a[s + r * B];
      }
    }
  }
}

template <typename T, int N, int E, int DIR, int PASS, class Lay, bool TWH = false>
__device__ __forceinline__ void fft_passes(cpx<T> (&a)[E], cpx<T>* lds, const Lay& lay, int c, int j,
                                           const cpx<T>* __restrict__ tw, int twstride, bool wave) {
  constexpr RadixPlan pl = radix_plan(N, E);
  if constexpr (PASS < pl.n) {
    fft_pass<T, N, E, DIR, PASS, Lay, TWH>(a, lds, lay, c, j, tw, twstride, wave);
    fft_passes<T, N, E, DIR, PASS + 1, Lay, TWH>(a, lds, lay, c, j, tw, twstride, wave);
  }
}

// Line FFT of length N with R elements per thread on P = N/R threads per line (see
// header).  tw[m] = exp(-2 pi i m / LT) with LT = N * twstride.  With more than one
// pass it starts with a barrier before its first LDS write and ends after its last LDS
// read (callers that write LDS afterwards must barrier first); every thread of the block
// must call it.  wave = true (every line of the block within one wavefront, lines never
// sharing LDS entries) orders the exchanges per wavefront instead (lds_sync); then a
// twiddle table in LDS must be complete before the call.  TWH: tw holds only the first
// LT/2 entries, the others follow from tw[k + LT/2] = -tw[k].
template <typename T, int N, int R, int DIR, class Lay, bool TWH = false>
__device__ __forceinline__ void line_fft(cpx<T> (&a)[R], cpx<T>* lds, const Lay& lay, int c, int j,
                                         const cpx<T>* __restrict__ tw, int twstride,
                                         bool wave = false) {
  constexpr RadixPlan pl = radix_plan(N, R);
  static_assert(pl.n > 0, "N must split into radices that divide R");
  static_assert(!TWH || N % 2 == 0, "a half twiddle table needs an even length");
  if constexpr (pl.n == 1) {
    dft<R, DIR>(a);
  } else {
    fft_passes<T, N, R, DIR, 0, Lay, TWH>(a, lds, lay, c, j, tw, twstride, wave);
  }
}

// --------------------------------------------------------------------------
// linear operator and stage updates
// --------------------------------------------------------------------------
// signed wavenumber index of fftfreq (Nyquist negative)
__device__ __forceinline__ int fftfreq_index(int i, int n) { return i < n / 2 ? i : i - n; }

__device__ __forceinline__ double ipow_d(double x, int n) {
  double r = 1.0;
  double b = x;
  while (n > 0) {
    if (n & 1) r *= b;
    b *= b;
    n >>= 1;
  }
  return r;
}

struct OpParams {   // per-member linear-operator coefficients
  double kappa, eta, iota, kappah;
};

struct GridInfo {
  int ndim, nx, ny, nz;     // physical sizes (ny = nz = 1 below ndim)
  double ksx, ksy, ksz;     // wavenumber scales 2 pi / L
};

// L = -kappa kr^2 - eta l^2 - iota m^2 - kappah Krsq^nkappah   (1D: (kr^2)^n)
__device__ __forceinline__ double linear_op(const GridInfo& g, const OpParams& p, int nkappah,
                                            int ikx, int iky, int ikz) {
  const double kr = ikx * g.ksx;
  const double kr2 = kr * kr;
  if (g.ndim == 1) return -p.kappa * kr2 - p.kappah * ipow_d(kr2, nkappah);
  const double l = fftfreq_index(iky, g.ny) * g.ksy;
  const double l2 = l * l;
  if (g.ndim == 2) return -p.kappa * kr2 - p.eta * l2 - p.kappah * ipow_d(kr2 + l2, nkappah);
  const double m = fftfreq_index(ikz, g.nz) * g.ksz;
  const double m2 = m * m;
  return -p.kappa * kr2 - p.eta * l2 - p.iota * m2 - p.kappah * ipow_d(kr2 + l2 + m2, nkappah);
}

enum { EPI_STORE = 0, EPI_RK4 = 1, EPI_ETDRK4 = 2 };

// Non-temporal access to fields that are touched once per stage and not again until
// the next stage, long after the caches have turned over (the stage state, the
// ETDRK4 coefficients, the steady flows): they should not displace the spectral
// fields that the next kernel reads.
template <typename T>
__device__ __forceinline__ cpx<T> ld_nt(const cpx<T>* p) {
  return {__builtin_nontemporal_load(&p->re), __builtin_nontemporal_load(&p->im)};
}
__device__ __forceinline__ double ld_nt(const double* p) { return __builtin_nontemporal_load(p); }
// A load that is non-temporal only where NT.  A block that covers part of each 128-byte
// line (the sibling block, on the same XCD, reads the rest) loads through the cache, so
// that the sibling finds the line in L2 instead of fetching it from memory again.
template <bool NT, typename T>
__device__ __forceinline__ cpx<T> ld_pol(const cpx<T>* p) {
  if constexpr (NT) return ld_nt(p);
  else return *p;
}
template <bool NT>
__device__ __forceinline__ double ld_pol(const double* p) {
  if constexpr (NT) return ld_nt(p);
  else return *p;
}
template <typename T>
__device__ __forceinline__ void st_nt(cpx<T>* p, cpx<T> v) {
  __builtin_nontemporal_store(v.re, &p->re);
  __builtin_nontemporal_store(v.im, &p->im);
}
// A store that is non-temporal only where NT (as ld_pol: a block that writes part of each
// 128-byte line stores through the cache, where the sibling block's half joins it).
template <bool NT, typename T>
__device__ __forceinline__ void st_pol(cpx<T>* p, cpx<T> v) {
  if constexpr (NT) st_nt(p, v);
  else *p = v;
}

template <typename T>
struct StageArgs {
  int kind;     // EPI_*
  int stage;    // 0..3
  int store_x;  // RK4 stage 0: also store the next stage input in stg (when the next
                // stage's first transform reads it from memory)
  cpx<T>* sol;  // state
  cpx<T>* sol_lo;  // mixed form: low part of the compensated state (sol + sol_lo), else NULL
  cpx<T>* stg;  // RK4: stage input sol1;  ETDRK4: sol1
  cpx<T>* stg2; // ETDRK4: sol2
  cpx<T>* acc;  // RK4: running RHS sum; ETDRK4: partial update
  cpx<T>* n1;   // ETDRK4: N1
  cpx<T>* n2;   // ETDRK4: N2
  const double* E;      // exp(L dt)
  const double* E2;     // exp(L dt / 2)
  const double* zeta;
  const double* alpha;
  const double* beta;
  const double* gam;
  double dt;
  const OpParams* op;   // per member
  int nkappah;
  GridInfo g;
  long long mstride;    // spectral elements per member
  // filtered steppers: sol *= filter(K) after the last stage's update
  int filtered;
  int f_order;
  double f_innerK;
  double f_decay;
};

// [FF] makefilter(grid; order, innerK, outerK, tol) at one spectral element:
// K = sqrt((kr dx/pi)^2 + (l dy/pi)^2 + (m dz/pi)^2), where k dx/pi = 2 i/n for the
// signed wavenumber index i; filter = 1 for K < innerK, else exp(-decay (K - innerK)^order).
__device__ __forceinline__ double filter_value(const GridInfo& g, int order, double innerK,
                                               double decay, int ikx, int iky, int ikz) {
  const double a = 2.0 * ikx / g.nx;
  double K2 = a * a;
  if (g.ndim >= 2) {
    const double b = 2.0 * fftfreq_index(iky, g.ny) / g.ny;
    K2 += b * b;
  }
  if (g.ndim == 3) {
    const double c = 2.0 * fftfreq_index(ikz, g.nz) / g.nz;
    K2 += c * c;
  }
  const double K = sqrt(K2);
  if (K < innerK) return 1.0;
  return exp(-decay * ipow_d(K - innerK, order));
}

// (hi, lo) + y -> (hi', lo') in working precision: TwoSum of hi and y, the error
// and lo folded together, then a FastTwoSum renormalisation (Knuth; Dekker).
template <typename T>
__device__ __forceinline__ void comp_add(T hi, T lo, T y, T& ho, T& lo_out) {
  const T s = hi + y;
  const T bb = s - hi;
  const T err = (hi - (s - bb)) + (y - bb);
  const T e = err + lo;
  ho = s + e;
  lo_out = e - (ho - s);
}

// Stage updates at spectral element ubase + off (ubase uniform over the block,
// off a 32-bit lane offset; member-relative wavenumber indices ikx, iky, ikz)
// given the nonlinear term N of the stage input.
// RK4 ([FF] RK4substeps!/RK4update!): RHS_s = N_s + L x_s;
//   x_2 = sol + dt/2 RHS_1, x_3 = sol + dt/2 RHS_2, x_4 = sol + dt RHS_3,
//   sol += dt (RHS_1/6 + RHS_2/3 + RHS_3/3 + RHS_4/6)   (RHS sum in this order)
// Storage between stages (acc, stg, sol):
//   stage 1: acc = RHS_1 (x_2 is recomputed from sol and acc in stage 2, with the
//            same expression, so it is not stored unless store_x is set);
//   stage 2: acc = RHS_1/6 + RHS_2/3, stg = x_3;
//   stage 3: stg = x_4; FP64 form: sol = sol + dt (RHS_1/6 + RHS_2/3 + RHS_3/3),
//            mixed form: acc = RHS_1/6 + RHS_2/3 + RHS_3/3 (sol stays compensated);
//   stage 4: FP64 form: sol = sol + dt RHS_4/6; mixed form: (sol, sol_lo) +=
//            dt (acc + RHS_4/6) with a compensated sum.
// The FP64 form thus adds the last term to sol in two roundings instead of one.
// ETDRK4 ([FF] ETDRK4substeps!/ETDRK4update!):
//   sol1 = E2 sol + zeta N1 ; sol2 = E2 sol + zeta N2 ; sol2 = E2 sol1 + zeta (2 N3 - N1);
//   sol = E sol + alpha N1 + 2 beta (N2 + N3) + Gamma N4
// Every update returns the value it stores as the next stage's input (for the
// last stage: the new sol), so a fused kernel can go on transforming it.

// RK4 epilogue split into an operand load (which callers may issue early) and
// the update.
template <typename T>
struct Rk4Ops {
  cpx<T> x, s, ac, lo;
};

// The RK4 stage operands and update.  ST >= 0 fixes the stage at compile time (callers
// that dispatch on the uniform stage once, so that the loads of several elements form
// one straight-line block); ST = -1 reads it from the arguments.  store = false (lanes
// of a tail column that only computes) suppresses the state stores.  NTS: the state
// loads are non-temporal (the caller's blocks cover whole 128-byte lines of the state).
template <typename T, int ST = -1, bool NTS = true>
__device__ __forceinline__ Rk4Ops<T> rk4_load(const StageArgs<T>& a, long long ubase, uint32_t off) {
  const int stage = ST >= 0 ? ST : a.stage;
  // every field is defined on every path so that the value stays in registers
  const cpx<T> z = {(T)0, (T)0};
  Rk4Ops<T> o = {z, z, z, z};
  if (stage == 0) {
    o.x = ld_pol<NTS>(a.sol + ubase + off);
  } else if (stage == 1) {
    o.s = ld_pol<NTS>(a.sol + ubase + off);
    o.ac = ld_pol<NTS>(a.acc + ubase + off);
  } else {
    o.x = ld_pol<NTS>(a.stg + ubase + off);
    o.s = ld_pol<NTS>(a.sol + ubase + off);
    // the compensated low part exists only in the mixed (FP32) form
    constexpr bool kMixed = sizeof(T) == 4;
    if (stage == 2 || (kMixed && a.sol_lo)) o.ac = ld_pol<NTS>(a.acc + ubase + off);
    if (kMixed && stage == 3 && a.sol_lo) o.lo = ld_pol<NTS>(a.sol_lo + ubase + off);
  }
  return o;
}

// rk4_finish with the element's linear operator L already evaluated (a caller that
// updates the same elements in every stage computes it once)
template <typename T, int ST = -1, bool NTS = true>
__device__ __forceinline__ cpx<T> rk4_finish_L(const StageArgs<T>& a, long long ubase, uint32_t off,
                                               T L, int ikx, int iky, int ikz, cpx<T> N,
                                               const Rk4Ops<T>& o, bool store = true) {
  const int stage = ST >= 0 ? ST : a.stage;
  const T dt = (T)a.dt;
  const T h2 = (T)(a.dt / 2);
  // stage 2 input x_2 = sol + dt/2 RHS_1, recomputed exactly as stage 1 formed it
  const cpx<T> x = stage == 1 ? cpx<T>{o.s.re + h2 * o.ac.re, o.s.im + h2 * o.ac.im} : o.x;
  const cpx<T> rhs = {N.re + L * x.re, N.im + L * x.im};
  if (stage == 0) {
    const cpx<T> nx = {x.re + h2 * rhs.re, x.im + h2 * rhs.im};
    if (store) {
      if (a.store_x) st_pol<NTS>(a.stg + ubase + off, nx);
      st_pol<NTS>(a.acc + ubase + off, rhs);
    }
    return nx;
  } else if (stage == 1) {
    const cpx<T> nx = {o.s.re + h2 * rhs.re, o.s.im + h2 * rhs.im};
    if (store) {
      st_pol<NTS>(a.stg + ubase + off, nx);
      st_pol<NTS>(a.acc + ubase + off, {o.ac.re / (T)6 + rhs.re / (T)3, o.ac.im / (T)6 + rhs.im / (T)3});
    }
    return nx;
  } else if (stage == 2) {
    const cpx<T> nx = {o.s.re + dt * rhs.re, o.s.im + dt * rhs.im};
    const cpx<T> tot = {o.ac.re + rhs.re / (T)3, o.ac.im + rhs.im / (T)3};
    if (store) {
      st_pol<NTS>(a.stg + ubase + off, nx);
      if (sizeof(T) == 4 && a.sol_lo) st_pol<NTS>(a.acc + ubase + off, tot);
      else st_pol<NTS>(a.sol + ubase + off, {o.s.re + dt * tot.re, o.s.im + dt * tot.im});
    }
    return nx;
  } else {
    if (sizeof(T) == 8 || !a.sol_lo) {
      // o.s = sol + dt (RHS_1/6 + RHS_2/3 + RHS_3/3), stored by stage 3
      cpx<T> ns = {o.s.re + dt * (rhs.re / (T)6), o.s.im + dt * (rhs.im / (T)6)};
      if (a.filtered) {
        const T f = (T)filter_value(a.g, a.f_order, a.f_innerK, a.f_decay, ikx, iky, ikz);
        ns = {f * ns.re, f * ns.im};
      }
      if (store) st_pol<NTS>(a.sol + ubase + off, ns);
      return ns;
    }
    // compensated accumulation: (hi, lo) += dt tot with TwoSum + renormalisation
    const cpx<T> tot = {o.ac.re + rhs.re / (T)6, o.ac.im + rhs.im / (T)6};
    T hr, lr, hi_, li_;
    comp_add(o.s.re, o.lo.re, dt * tot.re, hr, lr);
    comp_add(o.s.im, o.lo.im, dt * tot.im, hi_, li_);
    if (a.filtered) {
      // (hi + lo) f in FP64, split again into a working-precision pair
      const double f = filter_value(a.g, a.f_order, a.f_innerK, a.f_decay, ikx, iky, ikz);
      const double xr = ((double)hr + (double)lr) * f, xi = ((double)hi_ + (double)li_) * f;
      hr = (T)xr;
      hi_ = (T)xi;
      lr = (T)(xr - (double)hr);
      li_ = (T)(xi - (double)hi_);
    }
    if (store) {
      st_pol<NTS>(a.sol + ubase + off, {hr, hi_});
      st_pol<NTS>(a.sol_lo + ubase + off, {lr, li_});
    }
    return {hr, hi_};
  }
}

// op = the member's operator coefficients (a.op[member]), read once by the caller: a
// read through a.op between the state stores would be repeated for every element.
template <typename T, int ST = -1, bool NTS = true>
__device__ __forceinline__ cpx<T> rk4_finish(const StageArgs<T>& a, long long ubase, uint32_t off,
                                             const OpParams& op, int ikx, int iky, int ikz, cpx<T> N,
                                             const Rk4Ops<T>& o, bool store = true) {
  return rk4_finish_L<T, ST, NTS>(a, ubase, off, (T)linear_op(a.g, op, a.nkappah, ikx, iky, ikz), ikx,
                             iky, ikz, N, o, store);
}

// The ETDRK4 stage operands and update (coefficients are FP64 arrays; the update runs in
// FP64 and rounds once).  ST and store as for rk4_load / rk4_finish.
template <typename T>
struct EtdOps {
  double sr, si;          // sol (+ its low part) for stages 0-2; acc (+ low part) for stage 3
  cpx<T> s1, n1, n2;      // stage 2
  double e2, z, e, al, be, g;
};

// NTS, NTC: the loads of the state and of the (FP64 real) coefficients are non-temporal
// (as for rk4_load: where the caller's blocks cover whole 128-byte lines of them)
template <typename T, int ST = -1, bool NTS = true, bool NTC = true>
__device__ __forceinline__ EtdOps<T> etd_load(const StageArgs<T>& a, long long ubase, uint32_t off) {
  const int stage = ST >= 0 ? ST : a.stage;
  const cpx<T> zc = {(T)0, (T)0};
  EtdOps<T> o = {0.0, 0.0, zc, zc, zc, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  // the compensated low parts exist only in the mixed (FP32) form
  constexpr bool kMixed = sizeof(T) == 4;
  if (stage <= 2) {
    const cpx<T> s0 = ld_pol<NTS>(a.sol + ubase + off);
    o.sr = (double)s0.re;
    o.si = (double)s0.im;
    if (kMixed && a.sol_lo) {
      const cpx<T> lo = ld_pol<NTS>(a.sol_lo + ubase + off);
      o.sr += (double)lo.re;
      o.si += (double)lo.im;
    }
    o.e2 = ld_pol<NTC>(a.E2 + ubase + off);
    o.z = ld_pol<NTC>(a.zeta + ubase + off);
    if (stage == 2) {
      o.s1 = ld_pol<NTS>(a.stg + ubase + off);
      o.n1 = ld_pol<NTS>(a.n1 + ubase + off);
      o.n2 = ld_pol<NTS>(a.n2 + ubase + off);
      o.e = ld_pol<NTC>(a.E + ubase + off);
      o.al = ld_pol<NTC>(a.alpha + ubase + off);
      o.be = ld_pol<NTC>(a.beta + ubase + off);
    }
  } else {
    const cpx<T> ac = ld_pol<NTS>(a.acc + ubase + off);
    o.sr = (double)ac.re;
    o.si = (double)ac.im;
    if (kMixed && a.sol_lo) {
      const cpx<T> al = ld_pol<NTS>(a.n2 + ubase + off);   // the low part of acc
      o.sr += (double)al.re;
      o.si += (double)al.im;
    }
    o.g = ld_pol<NTC>(a.gam + ubase + off);
  }
  return o;
}

template <typename T, int ST = -1, bool NTS = true>
__device__ __forceinline__ cpx<T> etd_finish(const StageArgs<T>& a, long long ubase, uint32_t off,
                                             int ikx, int iky, int ikz, cpx<T> N,
                                             const EtdOps<T>& o, bool store = true) {
  const int stage = ST >= 0 ? ST : a.stage;
  constexpr bool kMixed = sizeof(T) == 4;
  const double nre = (double)N.re, nim = (double)N.im;
  if (stage == 0 || stage == 1) {
    const cpx<T> v = {(T)(o.e2 * o.sr + o.z * nre), (T)(o.e2 * o.si + o.z * nim)};
    if (store) {
      if (stage == 0) {
        st_pol<NTS>(a.stg + ubase + off, v);
        st_pol<NTS>(a.n1 + ubase + off, N);
      } else {
        st_pol<NTS>(a.stg2 + ubase + off, v);
        st_pol<NTS>(a.n2 + ubase + off, N);
      }
    }
    return v;
  } else if (stage == 2) {
    const double be2 = 2.0 * o.be;
    const cpx<T> v = {(T)(o.e2 * (double)o.s1.re + o.z * (2.0 * nre - (double)o.n1.re)),
                      (T)(o.e2 * (double)o.s1.im + o.z * (2.0 * nim - (double)o.n1.im))};
    const double ar = o.e * o.sr + o.al * (double)o.n1.re + be2 * ((double)o.n2.re + nre);
    const double ai = o.e * o.si + o.al * (double)o.n1.im + be2 * ((double)o.n2.im + nim);
    const cpx<T> hi = {(T)ar, (T)ai};
    if (store) {
      st_pol<NTS>(a.stg2 + ubase + off, v);
      st_pol<NTS>(a.acc + ubase + off, hi);
      if (kMixed && a.sol_lo)   // the low part of acc
        st_pol<NTS>(a.n2 + ubase + off, {(T)(ar - (double)hi.re), (T)(ai - (double)hi.im)});
    }
    return v;
  } else {
    double sr = o.sr + o.g * nre, si = o.si + o.g * nim;
    if (a.filtered) {
      const double f = filter_value(a.g, a.f_order, a.f_innerK, a.f_decay, ikx, iky, ikz);
      sr *= f;
      si *= f;
    }
    const cpx<T> hi = {(T)sr, (T)si};
    if (store) {
      st_pol<NTS>(a.sol + ubase + off, hi);
      if (kMixed && a.sol_lo)
        st_pol<NTS>(a.sol_lo + ubase + off, {(T)(sr - (double)hi.re), (T)(si - (double)hi.im)});
    }
    return hi;
  }
}

template <typename T>
__device__ __forceinline__ cpx<T> etdrk4_update(const StageArgs<T>& a, long long ubase,
                                                uint32_t off, int ikx, int iky, int ikz, cpx<T> N) {
  const EtdOps<T> o = etd_load(a, ubase, off);
  return etd_finish(a, ubase, off, ikx, iky, ikz, N, o);
}

// Stage update at spectral element ubase + off with the stage's nonlinear term N;
// returns the next stage's input at that element.
template <typename T>
__device__ __forceinline__ cpx<T> stage_epilogue(const StageArgs<T>& a, long long ubase,
                                                 uint32_t off, int member, int ikx, int iky,
                                                 int ikz, cpx<T> N) {
  if (a.kind == EPI_RK4) {
    const Rk4Ops<T> o = rk4_load(a, ubase, off);
    return rk4_finish(a, ubase, off, a.op[member], ikx, iky, ikz, N, o);
  }
  return etdrk4_update(a, ubase, off, ikx, iky, ikz, N);
}

}  // namespace ptf
