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
#include <type_traits>

#include "ptf_device.hpp"

namespace ptf {

enum {
  COL_PLAIN = 0,   // out0 = FFT_DIR(in0)
  COL_D1 = 1,      // out0 = IFFT(in0), out1 = IFFT(i k in0)                   (first inverse axis)
  COL_D2 = 2,      // out0 = IFFT(in0), out1 = IFFT(i k in0), out2 = IFFT(in1)  (3D y inverse)
  COL_STAGE = 3,   // FFT_fwd(in0) -> stage epilogue                            (last forward axis)
  COL_FLOWSYN = 4, // psi(t) = psi0 exp(-i w t): out0 = IFFT_y(-i l psi), out1 = IFFT_y(i kr psi)
  COL_STAGE_D1 = 5 // COL_STAGE, then COL_D1 on the next stage's input (stage boundary fusion)
};

// Epilogue families of the stage kernels (COL_STAGE, COL_STAGE_D1), each compiled on its
// own: a kernel then holds registers for one family's operands only (the RK4 operands
// prefetched ahead of the transform and the ETDRK4 chunks would otherwise share one
// allocation).  The host launches a family's kernel only for its epilogue kinds.
// EPF_ANY compiles every epilogue behind a runtime test of the kind (FP64 RK4 uses it:
// there the RK4-only build spilled more under the 3-waves/SIMD budget).
enum {
  EPF_NONE = 0,    // not a stage kernel
  EPF_RK4 = 1,     // EPI_RK4 (RK4 and filtered RK4)
  EPF_ETD = 2,     // EPI_ETDRK4 (ETDRK4 and filtered ETDRK4) and EPI_STORE (calcN! alone)
  EPF_ANY = 3      // any of them
};
template <int MODE, int EPF>
constexpr bool epf_ok() {
  return (MODE == COL_STAGE || MODE == COL_STAGE_D1) == (EPF != EPF_NONE);
}

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
  // Brick layout (the last grid axis stored in bricks of 2^lb planes, innermost):
  //   lb  > 0: the transformed line runs along the bricked axis; line element i
  //            sits at column*2^lb + (i mod 2^lb) + (i >> lb)*stride
  //   lbc > 0: the columns of this pass are (kx, i_last mod 2^lbc) pairs,
  //            column = kx*2^lbc + (i_last mod 2^lbc), and blockIdx.y is i_last >> lbc
  int lb;
  int lbc;
  // line sets along blockIdx.y and members (blockIdx.z): the tile space of the persistent
  // pipelined kernel, which is launched on a 1D grid
  int nother;
  int nmembers;
  // nonzero: a bricked line whose P threads divide a wavefront is held by lanes of one
  // wavefront (see col_kernel), and its exchanges are ordered per wavefront
  int wave_lines;
};

// Occupancy range for a block of `threads` threads using `lds` bytes of LDS:
// the upper end is the waves per SIMD that the LDS budget (160 KiB per CU)
// admits (at most 8); the lower end (at most 2) is what the register file must
// always allow.  Within the range the compiler raises occupancy only as far as
// it can without spilling.
constexpr int lds_waves(int threads, int lds) {
  int blocks = lds > 0 ? 163840 / lds : 16;
  if (blocks < 1) blocks = 1;
  if (blocks > 16) blocks = 16;
  int w = blocks * ((threads + 63) / 64) / 4;
  return w < 1 ? 1 : (w > 8 ? 8 : w);
}
constexpr int min_waves(int threads, int lds, int occ = 0) {
  const int w = lds_waves(threads, lds);
  const int want = occ > 0 ? occ : 2;   // occ > 0: a tighter register budget, spills allowed
  return w < want ? w : want;
}

// Workgroups are dispatched round-robin over the 8 XCDs of the device, each
// with its own L2.  Renumber the grid so that XCD k runs a contiguous range of
// logical blocks (x fastest, then y, then z): neighbouring tiles, which share
// rows of the same planes, then meet in one L2.
__device__ __forceinline__ void xcd_remap(int& bx, int& by, int& bz) {
  constexpr unsigned kXcd = 8;
  const unsigned gx = gridDim.x, gy = gridDim.y;
  const unsigned n = gx * gy * gridDim.z;
  const unsigned L = blockIdx.x + gx * (blockIdx.y + gy * blockIdx.z);
  const unsigned xcd = L % kXcd, local = L / kXcd;
  const unsigned per = n / kXcd, rem = n % kXcd;
  const unsigned logical = xcd * per + (xcd < rem ? xcd : rem) + local;
  bx = (int)(logical % gx);
  by = (int)((logical / gx) % gy);
  bz = (int)(logical / (gx * gy));
}

// Lines of up to 2048 points keep their twiddle table in LDS: the transform's twiddle
// reads then come back with its LDS reads instead of from L1/L2 after every exchange.
template <int N, int R>
constexpr bool col_tw_lds() { return N > R && N <= 2048; }

// Column exchange layouts.  Passes along a middle axis (lanes that share a line element i
// are consecutive) interleave the lines: element i of line c at i*TX + c, so those lanes
// hit consecutive entries.  The stage kernels run along the last axis, which is bricked
// from 64 planes on: there the lanes of one line are consecutive, and the lines are
// stored line-major with one skew entry every 8 elements, c*S + i + i/8 with
// S = N + N/8 + 8.  A Stockham pass writes a lane group's elements at stride Rp and reads
// them at stride 1; the skew spreads both over distinct banks (8 lanes per 128 bytes;
// S = 8 mod 16 puts the next line's lanes in the other half of a 256-byte group), where
// the interleaved layout maps a group of 8 lanes to 2 to 4 bank positions.
template <int N>
constexpr int col_line_stride() { return N + N / 8 + 8; }
template <int N, int TX, bool SKEW>
struct LayCol {
  __device__ __forceinline__ int idx(int c, int i) const {
    if constexpr (SKEW) return c * col_line_stride<N>() + i + (i >> 3);
    else return i * TX + c;
  }
  // off a multiple of 8: i0 + off skews by (i0 >> 3) + off / 8
  __device__ __forceinline__ int idx_split(int c, int i0, int off) const {
    if constexpr (SKEW) return c * col_line_stride<N>() + i0 + (i0 >> 3) + off + (off >> 3);
    else return idx(c, i0 + off);
  }
};
template <int MODE>
constexpr bool col_skew() { return MODE == COL_STAGE_D1; }

// Elements per operand chunk of the ETDRK4 epilogue (the operands of a chunk are loaded
// together: one memory round trip per chunk).  The third stage reads up to 9 operands per
// element (26 VGPRs in FP64): 4 elements at a time, 2 under the 3-waves/SIMD budget.  The
// others read 2 or 3 fields (6-8 VGPRs per element, after the transform has released its
// registers): 8 elements at a time, 4 under the 3-waves/SIMD budget.
template <int R, int ST, int OCC>
constexpr int etd_chunk() {
  const int c = ST == 2 ? (OCC > 0 ? 2 : 4) : (OCC > 0 ? 4 : 8);
  return R < c ? R : c;
}

template <typename T, int N, int R, int TX, int MODE>
constexpr int col_lds_line_elems() {
  return (N > R) ? (col_skew<MODE>() ? TX * col_line_stride<N>() : N * TX) : 1;
}
// FP64 lines of 2048 points keep half of the table in LDS (w[k + N/2] = -w[k]): 16 instead
// of 32 KB, so that three blocks of a fused stage kernel fit a CU instead of two
template <typename T, int N, int R>
constexpr bool col_tw_half() { return col_tw_lds<N, R>() && N * (int)sizeof(cpx<T>) > 16384; }
template <typename T, int N, int R>
constexpr int col_tw_elems() { return col_tw_lds<N, R>() ? (col_tw_half<T, N, R>() ? N / 2 : N) : 0; }
template <typename T, int N, int R, int TX, int MODE>
constexpr int col_lds_elems() {
  return col_lds_line_elems<T, N, R, TX, MODE>() + col_tw_elems<T, N, R>();
}

template <typename T, int N, int R, int TX, int MODE, int DIR, int OCC = 0, int EPF = EPF_NONE>
__global__ __attribute__((amdgpu_flat_work_group_size(1, TX * (N / R)),
                          amdgpu_waves_per_eu(
                              min_waves(TX * (N / R), (int)(col_lds_elems<T, N, R, TX, MODE>() * sizeof(cpx<T>)), OCC),
                              lds_waves(TX * (N / R), (int)(col_lds_elems<T, N, R, TX, MODE>() * sizeof(cpx<T>))))))
void col_kernel(ColArgs<T> a) {
  static_assert(epf_ok<MODE, EPF>(), "stage kernels take an epilogue family, others none");
  constexpr int P = N / R;
  // Non-temporal loads (and state stores) only where a block covers whole 128-byte lines
  // of a field: along the bricked axis a column's run is the 8 planes of a brick, so a
  // block reads TX * 8 consecutive elements; the columns of the 3D y pass are contiguous
  // (TX elements).
  constexpr int kRun = MODE == COL_D2 ? TX : TX * 8;
  constexpr bool kNtS = kRun * (int)sizeof(cpx<T>) >= 128;   // spectral fields, state
  constexpr bool kNtC = kRun * (int)sizeof(double) >= 128;   // FP64 real coefficients
  using Lay = LayCol<N, TX, col_skew<MODE>()>;
  const Lay lay{};
  __shared__ cpx<T> lds[col_lds_line_elems<T, N, R, TX, MODE>()];
  constexpr bool kTwh = col_tw_half<T, N, R>();
  __shared__ cpx<T> twl[col_tw_lds<N, R>() ? col_tw_elems<T, N, R>() : 1];
  const int t = threadIdx.x;
  const int lb = a.lb;
  // Lines within one wavefront: along a bricked axis, when the P threads of a line divide
  // a wavefront and a wavefront's 64 / P lines still cover whole 128-byte runs per load.
  constexpr bool kWaveOk = P <= 64 && 64 % P == 0 && (sizeof(T) == 8 || P <= 32);
  const bool wave = kWaveOk && lb > 0 && a.wave_lines != 0;
  if constexpr (col_tw_lds<N, R>()) {
    // twl[m] = exp(-2 pi i m / N); the first barrier of every transform orders these
    // writes before its first twiddle read (with per-wavefront exchanges, this barrier)
    for (int i = t; i < col_tw_elems<T, N, R>(); i += TX * P) twl[i] = a.tw[i * a.twstride];
    if (wave) __syncthreads();
  }
  // lane t = rb + 2^lb (c + TX jh): the brick's planes vary fastest, then the
  // columns, so a wavefront reads 2^lb * TX contiguous values; the line
  // element of register q is i = j + q P with j = rb + 2^lb jh (P % 2^lb == 0).
  // Per-wavefront lines: lane t = rb + 2^lb (jh + (P / 2^lb) c), so the P lanes of a line
  // are consecutive and a wavefront reads runs of 2^lb * 64 / P values.
  const int rb = t & ((1 << lb) - 1);
  const int rest = t >> lb;
  const int c = wave ? rest / (P >> lb) : rest % TX;
  const int jh = wave ? rest % (P >> lb) : rest / TX;
  const int j = rb + (jh << lb);
  // the line transform in direction D with this block's twiddle source
  auto fft = [&](auto dir_c, cpx<T>(&x)[R]) {
    constexpr int D = decltype(dir_c)::value;
    if constexpr (col_tw_lds<N, R>()) line_fft<T, N, R, D, Lay, kTwh>(x, lds, lay, c, j, twl, 1, wave);
    else line_fft<T, N, R, D, Lay>(x, lds, lay, c, j, a.tw, a.twstride, wave);
  };
  int bx, by, member;
  xcd_remap(bx, by, member);
  const int kx = bx * TX + c;
  const bool valid = kx < a.ncols;
  const int kxc = valid ? kx : a.ncols - 1;   // invalid lanes load a valid column, store nothing
  // block-uniform base (SGPRs) + 32-bit lane offsets
  const long long ubase = (long long)member * a.mstride + (long long)by * a.ostride;
  const uint32_t s = (uint32_t)a.stride;
  auto off = [&](int q) -> uint32_t {
    return ((uint32_t)kxc << lb) + (uint32_t)rb + (uint32_t)(jh + q * (P >> lb)) * s;
  };
  // wavenumber indices of this lane's column (3D y pass over folded columns)
  const int ikx_col = kx >> a.lbc;
  const int iklast_col = (by << a.lbc) + (kx & ((1 << a.lbc) - 1));

  cpx<T> v[R];
  if constexpr (MODE == COL_FLOWSYN) {
    const cpx<double>* ps = a.psih0 + ubase;
    const double* om = a.omega + ubase;
    // out1 = IFFT_y(i kr psi(t)) first, then out0 = IFFT_y(-i l psi(t)); psi(t) is formed
    // once (one read of psih0 and omega), both inputs at once
    cpx<T> w[R];
#pragma unroll
    for (int q = 0; q < R; ++q) {
      const int iy = j + q * P;
      const cpx<double> p0 = ps[off(q)];
      const double ph = -om[off(q)] * a.t;
      double sn, cs;
      sincos(ph, &sn, &cs);
      const cpx<double> p = {p0.re * cs - p0.im * sn, p0.re * sn + p0.im * cs};
      const double kr = ikx_col * a.kscale_x;
      v[q] = {(T)(-(kr * p.im)), (T)(kr * p.re)};          // i kr p
      const double l = fftfreq_index(iy, N) * a.kscale;
      w[q] = {(T)(l * p.im), (T)(-(l * p.re))};             // -i l p
    }
#pragma unroll
    for (int pass = 0; pass < 2; ++pass) {
      if (pass == 1) {
#pragma unroll
        for (int q = 0; q < R; ++q) v[q] = w[q];
      }
      fft(std::integral_constant<int, 1>{}, v);
      cpx<T>* o = (pass == 0 ? a.out1 : a.out0) + ubase;
      if (valid) {
#pragma unroll
        for (int q = 0; q < R; ++q) o[off(q)] = v[q];
      }
    }
    return;
  } else {
    const cpx<T>* in0 = a.in0 + ubase;
#pragma unroll
    for (int q = 0; q < R; ++q) {
      // the 3D y-inverse pass reads its inputs once: non-temporal loads there
      if constexpr (MODE == COL_D2) v[q] = ld_pol<kNtS>(in0 + off(q));
      else v[q] = in0[off(q)];
    }

    if constexpr (MODE == COL_PLAIN) {
      fft(std::integral_constant<int, DIR>{}, v);
      cpx<T>* o = a.out0 + ubase;
      if (valid) {
#pragma unroll
        for (int q = 0; q < R; ++q) o[off(q)] = v[q];
      }
    } else if constexpr (MODE == COL_D1 || MODE == COL_D2) {
      // The derivative is transformed and stored first, because out0 may alias in0
      // (in place).  The plain copy stays in registers from the first load, so the
      // input is read from memory once.  When registers allow (not FP64 at radix 16),
      // in1 is loaded together with in0, otherwise after the second transform.
      constexpr bool kKeep = sizeof(T) == 4 || R <= 8;
      cpx<T> u[R];
      cpx<T> p2[(MODE == COL_D2 && kKeep) ? R : 1];
#pragma unroll
      for (int q = 0; q < R; ++q) u[q] = v[q];
      if constexpr (kKeep && MODE == COL_D2) {
        const cpx<T>* in1 = a.in1 + ubase;
#pragma unroll
        for (int q = 0; q < R; ++q) p2[q] = ld_pol<kNtS>(in1 + off(q));
      }
#pragma unroll
      for (int q = 0; q < R; ++q) {
        const T k = (T)(fftfreq_index(j + q * P, N) * a.kscale);
        v[q] = times_ik(v[q], k);
      }
      fft(std::integral_constant<int, 1>{}, v);
      {
        cpx<T>* o = a.out1 + ubase;
        if (valid) {
#pragma unroll
          for (int q = 0; q < R; ++q) o[off(q)] = v[q];
        }
      }
#pragma unroll
      for (int q = 0; q < R; ++q) v[q] = u[q];
      fft(std::integral_constant<int, 1>{}, v);
      {
        cpx<T>* o = a.out0 + ubase;
        if (valid) {
#pragma unroll
          for (int q = 0; q < R; ++q) o[off(q)] = v[q];
        }
      }
      if constexpr (MODE == COL_D2) {
        const cpx<T>* in1 = a.in1 + ubase;
#pragma unroll
        for (int q = 0; q < R; ++q) {
          if constexpr (kKeep) v[q] = p2[q];
          else v[q] = ld_pol<kNtS>(in1 + off(q));
        }
        fft(std::integral_constant<int, 1>{}, v);
        cpx<T>* o = a.out2 + ubase;
        if (valid) {
#pragma unroll
          for (int q = 0; q < R; ++q) o[off(q)] = v[q];
        }
      }
    } else if constexpr (MODE == COL_STAGE || MODE == COL_STAGE_D1) {
      // FP32 and radix 4: the RK4 operands of the epilogue are loaded before the
      // transform so that their latency overlaps it (FP64 radix 8/16 would spill)
      constexpr bool kPre = sizeof(T) == 4 || R <= 4;
      Rk4Ops<T> ops[kPre ? R : 1];
      const bool rk4 = EPF == EPF_RK4 || (EPF == EPF_ANY && a.st.kind == EPI_RK4);
      // The stage is uniform: dispatch on it once, so that the operand loads of a chunk
      // of elements form one straight-line block and go out together (one memory round
      // trip per chunk, not one per element).  FP64 radix >= 8 (RK4) loads in two chunks;
      // ETDRK4 (up to 9 operands per element) in chunks of 4 or 2 elements.
      auto stage_dispatch = [&](auto body) {
        switch (a.st.stage) {
          case 0: body(std::integral_constant<int, 0>{}); break;
          case 1: body(std::integral_constant<int, 1>{}); break;
          case 2: body(std::integral_constant<int, 2>{}); break;
          default: body(std::integral_constant<int, 3>{}); break;
        }
      };
      if constexpr (kPre) {
        if (rk4) {
          stage_dispatch([&](auto stc) {
            constexpr int ST = decltype(stc)::value;
#pragma unroll
            for (int q = 0; q < R; ++q) ops[q] = rk4_load<T, ST, kNtS>(a.st, ubase, off(q));
          });
        }
      }
      fft(std::integral_constant<int, -1>{}, v);
      // invalid (tail) lanes compute on a clamped column and store nothing
      if (rk4) {
        const OpParams opm = a.st.op[member];
        stage_dispatch([&](auto stc) {
          constexpr int ST = decltype(stc)::value;
          constexpr int kChunk = (!kPre && sizeof(T) == 8 && R >= 8) ? R / 2 : R;
#pragma unroll
          for (int q0 = 0; q0 < R; q0 += kChunk) {
            Rk4Ops<T> o[kChunk];
            if constexpr (!kPre) {
#pragma unroll
              for (int i = 0; i < kChunk; ++i) o[i] = rk4_load<T, ST, kNtS>(a.st, ubase, off(q0 + i));
            }
#pragma unroll
            for (int i = 0; i < kChunk; ++i) {
              const int q = q0 + i;
              const int ia = j + q * P;
              const int iky = a.axis == 2 ? by : ia;
              const int ikz = a.axis == 2 ? ia : iklast_col;
              Rk4Ops<T> oq;
              if constexpr (kPre) oq = ops[q];
              else oq = o[i];
              v[q] = rk4_finish<T, ST, kNtS>(a.st, ubase, off(q), opm, ikx_col, iky, ikz, v[q], oq, valid);
            }
          }
        });
      } else if (EPF != EPF_RK4 && a.st.kind == EPI_ETDRK4) {
        stage_dispatch([&](auto stc) {
          constexpr int ST = decltype(stc)::value;
          constexpr int kChunk = etd_chunk<R, ST, OCC>();
#pragma unroll
          for (int q0 = 0; q0 < R; q0 += kChunk) {
            EtdOps<T> o[kChunk];
#pragma unroll
            for (int i = 0; i < kChunk; ++i) o[i] = etd_load<T, ST, kNtS, kNtC>(a.st, ubase, off(q0 + i));
#pragma unroll
            for (int i = 0; i < kChunk; ++i) {
              const int q = q0 + i;
              const int ia = j + q * P;
              const int iky = a.axis == 2 ? by : ia;
              const int ikz = a.axis == 2 ? ia : iklast_col;
              v[q] = etd_finish<T, ST, kNtS>(a.st, ubase, off(q), ikx_col, iky, ikz, v[q], o[i], valid);
            }
          }
        });
      } else if constexpr (MODE == COL_STAGE) {
        // plain store of the transformed line (calcN! without a stage update)
        if (valid) {
#pragma unroll
          for (int q = 0; q < R; ++q) (a.out0 + ubase)[off(q)] = v[q];
        }
      }
      if constexpr (MODE == COL_STAGE_D1) {
        // v holds the next stage's input along the line: transform it as COL_D1
        constexpr bool kKeep = sizeof(T) == 4 || R <= 8;
        cpx<T> u[kKeep ? R : 1];
        if constexpr (kKeep) {
#pragma unroll
          for (int q = 0; q < R; ++q) u[q] = v[q];
        }
        cpx<T> w[kKeep ? 1 : R];
        if constexpr (!kKeep) {
#pragma unroll
          for (int q = 0; q < R; ++q) w[q] = v[q];
        }
#pragma unroll
        for (int q = 0; q < R; ++q) {
          const T k = (T)(fftfreq_index(j + q * P, N) * a.kscale);
          v[q] = times_ik(v[q], k);
        }
        fft(std::integral_constant<int, 1>{}, v);
        {
          cpx<T>* op = a.out1 + ubase;
          if (valid) {
#pragma unroll
            for (int q = 0; q < R; ++q) op[off(q)] = v[q];
          }
        }
#pragma unroll
        for (int q = 0; q < R; ++q) {
          if constexpr (kKeep) v[q] = u[q];
          else v[q] = w[q];
        }
        fft(std::integral_constant<int, 1>{}, v);
        cpx<T>* op = a.out0 + ubase;
        if (valid) {
#pragma unroll
          for (int q = 0; q < R; ++q) op[off(q)] = v[q];
        }
      }
    }
  }
}

// --------------------------------------------------------------------------
// software-pipelined stage-boundary kernel (COL_STAGE_D1 work, persistent blocks)
// --------------------------------------------------------------------------
// The same work as col_kernel<..., COL_STAGE_D1, ...>: forward FFT of in0 along the line,
// the RK4/ETDRK4 stage update, then IFFT and IFFT(i k .) of the next stage input.  A
// block stays resident and walks over tiles (TX columns of one line set each); while it
// transforms tile i it already has the loads of tile i+1 in flight: the line and, for
// RK4, the stage operands, held in registers.  So a block's memory latency overlaps its
// own compute instead of needing other blocks to fill the gap, which the register
// budget of FP64 lines caps at a few waves per SIMD.
//
// Tiles are numbered x fastest (column tile), then blockIdx.y's dimension (`other`), then
// members; XCD k takes a contiguous range of tiles (blocks are dispatched round-robin over
// the 8 XCDs) and its blocks interleave over that range, so tiles that share cache lines
// run at the same time on one L2.
template <int N, int R, int TX>
constexpr int pipe_threads() { return TX * (N / R); }

template <typename T, int N, int R, int TX, int OCC = 0, int EPF = EPF_RK4>
__global__ __attribute__((amdgpu_flat_work_group_size(1, TX * (N / R)),
                          amdgpu_waves_per_eu(1, OCC > 0 ? OCC : 2)))
void col_pipe_kernel(ColArgs<T> a) {
  static_assert(EPF != EPF_NONE, "a stage kernel takes an epilogue family");
  constexpr int P = N / R;
  // non-temporal operand loads and state stores where a tile covers whole 128-byte lines
  // (as in col_kernel)
  constexpr bool kNtS = TX * 8 * (int)sizeof(cpx<T>) >= 128;
  constexpr bool kNtC = TX * 8 * (int)sizeof(double) >= 128;
  const int tiles_x = (a.ncols + TX - 1) / TX;
  const int tiles_y = a.nother;
  const int tiles_total = tiles_x * tiles_y * a.nmembers;
  using Lay = LayInterleaved<TX>;
  const Lay lay{};
  __shared__ cpx<T> lds[(N > R) ? N * TX : 1];
  __shared__ cpx<T> twl[col_tw_lds<N, R>() ? N : 1];
  const int t = threadIdx.x;
  if constexpr (col_tw_lds<N, R>()) {
    for (int i = t; i < N; i += TX * P) twl[i] = a.tw[i * a.twstride];
  }
  const int lb = a.lb;
  const int rb = t & ((1 << lb) - 1);
  const int rest = t >> lb;
  const int c = rest % TX;
  const int jh = rest / TX;
  const int j = rb + (jh << lb);
  auto fft = [&](auto dir_c, cpx<T>(&x)[R]) {
    constexpr int D = decltype(dir_c)::value;
    if constexpr (col_tw_lds<N, R>()) line_fft<T, N, R, D, Lay>(x, lds, lay, c, j, twl, 1);
    else line_fft<T, N, R, D, Lay>(x, lds, lay, c, j, a.tw, a.twstride);
  };
  const uint32_t s = (uint32_t)a.stride;

  // this block's tiles: XCD range [lo, hi), visited with stride = blocks on the XCD (with
  // fewer than 8 blocks, each block is its own range)
  const unsigned nb = gridDim.x;
  const unsigned kXcd = nb < 8u ? nb : 8u;
  const unsigned xcd = blockIdx.x % kXcd, local = blockIdx.x / kXcd;
  const unsigned per_xcd_blocks = nb / kXcd + (xcd < nb % kXcd ? 1u : 0u);
  const unsigned per = (unsigned)tiles_total / kXcd, remt = (unsigned)tiles_total % kXcd;
  const unsigned lo = xcd * per + (xcd < remt ? xcd : remt);
  const unsigned hi = lo + per + (xcd < remt ? 1u : 0u);

  struct Tile {
    int bx, by, member, kx, kxc;
    bool valid;
    long long ubase;
  };
  auto tile_at = [&](unsigned L) {
    Tile u;
    u.bx = (int)(L % (unsigned)tiles_x);
    u.by = (int)((L / (unsigned)tiles_x) % (unsigned)tiles_y);
    u.member = (int)(L / ((unsigned)tiles_x * (unsigned)tiles_y));
    u.kx = u.bx * TX + c;
    u.valid = u.kx < a.ncols;
    u.kxc = u.valid ? u.kx : a.ncols - 1;
    u.ubase = (long long)u.member * a.mstride + (long long)u.by * a.ostride;
    return u;
  };
  auto off = [&](const Tile& u, int q) -> uint32_t {
    return ((uint32_t)u.kxc << lb) + (uint32_t)rb + (uint32_t)(jh + q * (P >> lb)) * s;
  };
  auto load_line = [&](const Tile& u, cpx<T>(&v)[R]) {
    const cpx<T>* in0 = a.in0 + u.ubase;
#pragma unroll
    for (int q = 0; q < R; ++q) v[q] = in0[off(u, q)];
  };

  const bool rk4 = EPF == EPF_RK4 || (EPF == EPF_ANY && a.st.kind == EPI_RK4);
  auto stage_dispatch = [&](auto body) {
    switch (a.st.stage) {
      case 0: body(std::integral_constant<int, 0>{}); break;
      case 1: body(std::integral_constant<int, 1>{}); break;
      case 2: body(std::integral_constant<int, 2>{}); break;
      default: body(std::integral_constant<int, 3>{}); break;
    }
  };

  // the transform of the next stage input v (already updated) and its stores
  auto next_stage_inverse = [&](const Tile& u, cpx<T>(&v)[R]) {
    cpx<T> keep[R];
#pragma unroll
    for (int q = 0; q < R; ++q) keep[q] = v[q];
#pragma unroll
    for (int q = 0; q < R; ++q) {
      const T k = (T)(fftfreq_index(j + q * P, N) * a.kscale);
      v[q] = times_ik(v[q], k);
    }
    fft(std::integral_constant<int, 1>{}, v);
    if (u.valid) {
      cpx<T>* o1 = a.out1 + u.ubase;
#pragma unroll
      for (int q = 0; q < R; ++q) o1[off(u, q)] = v[q];
    }
#pragma unroll
    for (int q = 0; q < R; ++q) v[q] = keep[q];
    fft(std::integral_constant<int, 1>{}, v);
    if (u.valid) {
      cpx<T>* o0 = a.out0 + u.ubase;
#pragma unroll
      for (int q = 0; q < R; ++q) o0[off(u, q)] = v[q];
    }
  };

  unsigned L = lo + local;
  if (L >= hi) return;
  if (rk4) {
    stage_dispatch([&](auto stc) {
      constexpr int ST = decltype(stc)::value;
      Tile cur = tile_at(L);
      cpx<T> v[R];
      Rk4Ops<T> ops[R];
      load_line(cur, v);
#pragma unroll
      for (int q = 0; q < R; ++q) ops[q] = rk4_load<T, ST, kNtS>(a.st, cur.ubase, off(cur, q));
      for (;;) {
        const unsigned Ln = L + per_xcd_blocks;
        const bool more = Ln < hi;
        Tile nxt = cur;
        cpx<T> vn[R];
        Rk4Ops<T> opsn[R];
        if (more) {
          // the next tile's loads go out before this tile's work and stores
          nxt = tile_at(Ln);
          load_line(nxt, vn);
#pragma unroll
          for (int q = 0; q < R; ++q) opsn[q] = rk4_load<T, ST, kNtS>(a.st, nxt.ubase, off(nxt, q));
        }
        fft(std::integral_constant<int, -1>{}, v);
        const OpParams opm = a.st.op[cur.member];
        const int ikx = cur.kx >> a.lbc;
        const int iklast = (cur.by << a.lbc) + (cur.kx & ((1 << a.lbc) - 1));
#pragma unroll
        for (int q = 0; q < R; ++q) {
          const int ia = j + q * P;
          const int iky = a.axis == 2 ? cur.by : ia;
          const int ikz = a.axis == 2 ? ia : iklast;
          v[q] = rk4_finish<T, ST, kNtS>(a.st, cur.ubase, off(cur, q), opm, ikx, iky, ikz, v[q], ops[q],
                                   cur.valid);
        }
        next_stage_inverse(cur, v);
        if (!more) break;
        L = Ln;
        cur = nxt;
#pragma unroll
        for (int q = 0; q < R; ++q) {
          v[q] = vn[q];
          ops[q] = opsn[q];
        }
      }
    });
  } else {
    // ETDRK4: up to 9 operands per element, loaded after the transform; the next line is
    // still prefetched
    stage_dispatch([&](auto stc) {
      constexpr int ST = decltype(stc)::value;
      Tile cur = tile_at(L);
      cpx<T> v[R];
      load_line(cur, v);
      for (;;) {
        const unsigned Ln = L + per_xcd_blocks;
        const bool more = Ln < hi;
        Tile nxt = cur;
        cpx<T> vn[R];
        if (more) {
          nxt = tile_at(Ln);
          load_line(nxt, vn);
        }
        fft(std::integral_constant<int, -1>{}, v);
        const int ikx = cur.kx >> a.lbc;
        const int iklast = (cur.by << a.lbc) + (cur.kx & ((1 << a.lbc) - 1));
        constexpr int kChunk = etd_chunk<R, ST, 0>();
#pragma unroll
        for (int q0 = 0; q0 < R; q0 += kChunk) {
          EtdOps<T> o[kChunk];
#pragma unroll
          for (int i = 0; i < kChunk; ++i) o[i] = etd_load<T, ST, kNtS, kNtC>(a.st, cur.ubase, off(cur, q0 + i));
#pragma unroll
          for (int i = 0; i < kChunk; ++i) {
            const int q = q0 + i;
            const int ia = j + q * P;
            const int iky = a.axis == 2 ? cur.by : ia;
            const int ikz = a.axis == 2 ? ia : iklast;
            v[q] = etd_finish<T, ST, kNtS>(a.st, cur.ubase, off(cur, q), ikx, iky, ikz, v[q], o[i], cur.valid);
          }
        }
        next_stage_inverse(cur, v);
        if (!more) break;
        L = Ln;
        cur = nxt;
#pragma unroll
        for (int q = 0; q < R; ++q) v[q] = vn[q];
      }
    });
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
  // Brick layout of the last axis (2^lb planes innermost): rows are numbered
  // (last_hi, middle, last_lo) with last_lo fastest; spectral element k of row r
  // sits at (r >> lb) * nxh * 2^lb + k * 2^lb + (r mod 2^lb); physical values are
  // stored as x-pairs, pair m of row r at ((r >> lb) * nx/2 + m) * 2^lb + (r mod 2^lb).
  int lb;
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

// Partners X[M - k] of the row entries x[q] = X[j + q P] through LDS (X[M] is
// passed in `last` by lane j = 0).  Starts and ends with a barrier: LDS is
// free afterwards.
template <typename T, int M, int R, class Lay>
__device__ __forceinline__ void stage_row_pairs(const cpx<T> (&x)[R], cpx<T> last, cpx<T> (&xm)[R],
                                                cpx<T>* lds, const Lay& lay, int c, int j) {
  constexpr int P = M / R;
  __syncthreads();
#pragma unroll
  for (int q = 0; q < R; ++q) lds[lay.idx(c, j + q * P)] = x[q];
  if (j == 0) lds[lay.idx(c, M)] = last;
  __syncthreads();
#pragma unroll
  for (int q = 0; q < R; ++q) xm[q] = lds[lay.idx(c, M - (j + q * P))];
}

// Row element k is row[k << lb] (brick layout; lb = 0 for contiguous rows).  The x
// pass reads each spectral row once per stage: non-temporal loads (NT), where the
// block holds whole bricks of rows (see row_nt).
template <typename T, int M, int R, bool NT = true>
__device__ __forceinline__ void load_row(const cpx<T>* __restrict__ row, cpx<T> (&x)[R],
                                         cpx<T>& last, int j, int lb) {
  constexpr int P = M / R;
#pragma unroll
  for (int q = 0; q < R; ++q) x[q] = ld_pol<NT>(row + ((uint32_t)(j + q * P) << lb));
  if (j == 0) last = ld_pol<NT>(row + ((uint32_t)M << lb));
}

template <typename T, int M, int R, class Lay, bool NT = true>
__device__ __forceinline__ void load_row_pairs(const cpx<T>* __restrict__ row, cpx<T> (&x)[R],
                                               cpx<T> (&xm)[R], cpx<T>* lds, const Lay& lay, int c,
                                               int j, int lb) {
  cpx<T> last = {(T)0, (T)0};
  load_row<T, M, R, NT>(row, x, last, j, lb);
  stage_row_pairs<T, M, R, Lay>(x, last, xm, lds, lay, c, j);
}

constexpr int ilog2(int n) { return n <= 1 ? 0 : 1 + ilog2(n / 2); }

// A block of NL rows reads the spectral rows and flows of whole 8-plane bricks (every
// 128-byte line it touches entirely) when NL >= 8: non-temporal loads.  A block with part
// of a brick shares each line with its sibling blocks (renumbered onto one XCD), which
// find it in L2 when it was loaded through the cache.  (Contiguous rows, lb = 0, are
// whole lines either way; there the choice only affects what stays in the cache.)
template <int NL>
constexpr bool row_nt() { return NL >= 8; }

// Row r of a bricked array starts at ((r >> lb) * len << lb) + (r mod 2^lb), len = the
// row length in elements (nx/2 + 1 spectral values, or nx/2 physical x-pairs).  A block
// of NL rows starting at row0 holds whole bricks (NL >= 2^lb) or NL consecutive rows of
// one brick (NL < 2^lb).
__device__ __forceinline__ long long brick_row_base(long long row0, int lb, int len) {
  return (((row0 >> lb) * len) << lb) + (row0 & ((1LL << lb) - 1));
}

// Rows of up to 2048 points keep the twiddle table exp(-2 pi i k / nx), k < nx, in LDS
// (the half-length transform reads every other entry, the real pre/post-processing
// the first nx/2).
template <int NX>
constexpr bool row_tw_lds() { return NX <= 2048; }
// FP64 rows of 2048 points keep half of the table (w[k + nx/2] = -w[k]; the pre- and
// post-processing read only k < nx/2): 16 instead of 32 KB of LDS per block
template <typename T, int NX>
constexpr bool row_tw_half() { return row_tw_lds<NX>() && NX * (int)sizeof(cpx<T>) > 16384; }
template <typename T, int NX>
constexpr int row_tw_elems() { return row_tw_lds<NX>() ? (row_tw_half<T, NX>() ? NX / 2 : NX) : 0; }
template <typename T, int NX>
constexpr int row_tw_bytes() { return row_tw_elems<T, NX>() * (int)sizeof(cpx<T>); }

template <typename T, int NX, int R, int NL, int MODE, int NIN, int OCC = 0>
__global__ __attribute__((amdgpu_flat_work_group_size(1, NL * ((NX / 2) / R)),
                          amdgpu_waves_per_eu(
                              min_waves(NL * ((NX / 2) / R), (int)(NL * (NX / 2 + 1) * sizeof(cpx<T>)) + row_tw_bytes<T, NX>(), OCC),
                              lds_waves(NL * ((NX / 2) / R), (int)(NL * (NX / 2 + 1) * sizeof(cpx<T>)) + row_tw_bytes<T, NX>()))))
void row_kernel(RowArgs<T> a) {
  constexpr int M = NX / 2;
  constexpr int NXH = M + 1;
  constexpr int P = M / R;
  // LDS rows hold M + 1 entries (X[M] while staging a row).  Contiguous rows
  // (lanes walk a row) are stored row-major; bricked rows (lanes walk the rows
  // of a brick first) element-major, so a wavefront's accesses stay contiguous.
  constexpr int PAD = 1;
  using Lay = LayStrided;
  const Lay lay = a.lb ? Lay{1, NL} : Lay{M + PAD, 1};
  __shared__ cpx<T> lds[NL * (M + PAD)];
  constexpr bool kTwh = row_tw_half<T, NX>();
  __shared__ cpx<T> twl[row_tw_lds<NX>() ? row_tw_elems<T, NX>() : 1];
  const int t = threadIdx.x;
  if constexpr (row_tw_lds<NX>()) {
    // read only after the barriers of the first row exchange or transform
    for (int i = t; i < row_tw_elems<T, NX>(); i += (int)blockDim.x) twl[i] = a.tw[i * a.twstride];
  }
  // lane t = rb + 2^lt (j + P g), 2^lt = min(2^lb, NL): the block's rows of a brick vary
  // fastest.  With part of a brick per block (NL < 2^lb) the sibling blocks of a brick,
  // which share its cache lines, are renumbered onto one XCD.
  const int lb = a.lb;
  const int lt = lb < ilog2(NL) ? lb : ilog2(NL);
  int bx = blockIdx.x, member = blockIdx.y;
  if (lt < lb) {
    int bz;
    xcd_remap(bx, member, bz);
  }
  const int rb = t & ((1 << lt) - 1);
  const int rest = t >> lt;
  const int j = rest % P;
  const int g = rest / P;
  const int c = (g << lt) + rb;                                  // row within the block
  const long long row0 = (long long)bx * NL;
  const long long row = row0 + c;
  const bool valid = row < a.nrows;
  const int gmax = (int)((a.nrows - row0) >> lt) - 1;
  const int gc = valid ? g : (g < gmax ? g : gmax);              // clamped brick row
  // block-uniform bases + 32-bit lane offsets
  const long long sub = (long long)member * a.smstride + brick_row_base(row0, lb, NXH);
  const long long pub = (long long)member * a.pmstride + 2 * brick_row_base(row0, lb, M);
  const uint32_t so = ((uint32_t)gc * NXH << lb) + rb;           // spectral element 0 (loads)
  const uint32_t so_st = ((uint32_t)g * NXH << lb) + rb;         // (stores)
  const uint32_t po = ((uint32_t)gc * M << lb) + rb;             // physical pair 0 (loads)
  const uint32_t po_st = ((uint32_t)g * M << lb) + rb;           // (stores)
  const cpx<T> zero = {(T)0, (T)0};

  if constexpr (MODE == ROW_C2R) {
    cpx<T> z[R], zm[R];
    const T s = (T)a.scale;
    load_row_pairs<T, M, R, Lay, row_nt<NL>()>(a.in[0] + sub + so, z, zm, lds, lay, c, j, lb);
#pragma unroll
    for (int q = 0; q < R; ++q) {
      const int k = j + q * P;
      z[q] = c2r_pre(z[q], zm[q], (row_tw_lds<NX>() ? twl[k] : a.tw[k * a.twstride]), s, k == 0);
    }
    if constexpr (row_tw_lds<NX>()) line_fft<T, M, R, +1, Lay, kTwh>(z, lds, lay, c, j, twl, 2);
    else line_fft<T, M, R, +1, Lay>(z, lds, lay, c, j, a.tw, 2 * a.twstride);
    if (valid) {
      cpx<T>* out = reinterpret_cast<cpx<T>*>(a.rout + pub);
#pragma unroll
      for (int q = 0; q < R; ++q) out[po_st + ((uint32_t)(j + q * P) << lb)] = z[q];
    }
    return;
  } else {
    cpx<T> f[R];
    if constexpr (MODE == ROW_R2C) {
      const cpx<T>* rin = reinterpret_cast<const cpx<T>*>(a.rin + pub);
#pragma unroll
      for (int q = 0; q < R; ++q) f[q] = rin[po + ((uint32_t)(j + q * P) << lb)];
    } else {
      // ROW_ADV: f = ((-u cx) - v cy) - w cz, cx = C2R(i kr Y0), cy = C2R(Y1), cz = C2R(Y2)
      const T s = (T)a.scale;
      // select, not a.in[d]: a runtime index into the by-value kernel arguments
      // would copy them to scratch
      auto in_at = [&](int d) { return d == 0 ? a.in[0] : (d == 1 ? a.in[1] : a.in[2]); };
      auto flow_at = [&](int d) { return d == 0 ? a.flow[0] : (d == 1 ? a.flow[1] : a.flow[2]); };
      // FP32 and radix 4: the next input row and the flow values are loaded before
      // the current row is transformed (FP64 radix 8/16 would spill)
      constexpr bool kPre = sizeof(T) == 4 || R <= 4;
      cpx<T> nxt[kPre ? R : 1];
      cpx<T> nlast = zero;
      if constexpr (kPre) load_row<T, M, R, row_nt<NL>()>(a.in[0] + sub + so, nxt, nlast, j, lb);
#pragma unroll 1
      for (int d = 0; d < NIN; ++d) {
        cpx<T> z[R], zm[R];
        if constexpr (kPre) {
#pragma unroll
          for (int q = 0; q < R; ++q) z[q] = nxt[q];
          stage_row_pairs<T, M, R, Lay>(z, nlast, zm, lds, lay, c, j);
          if (d + 1 < NIN) load_row<T, M, R, row_nt<NL>()>(in_at(d + 1) + sub + so, nxt, nlast, j, lb);
        } else {
          load_row_pairs<T, M, R, Lay, row_nt<NL>()>(in_at(d) + sub + so, z, zm, lds, lay, c, j, lb);
        }
        const cpx<T>* fl = reinterpret_cast<const cpx<T>*>(flow_at(d) + pub);
        cpx<T> fv[kPre ? R : 1];
        if constexpr (kPre) {
#pragma unroll
          for (int q = 0; q < R; ++q) fv[q] = ld_pol<row_nt<NL>()>(fl + po + ((uint32_t)(j + q * P) << lb));
        }
#pragma unroll
        for (int q = 0; q < R; ++q) {
          const int k = j + q * P;
          cpx<T> Xk = z[q];
          cpx<T> Xm = zm[q];
          if (d == 0) {
            Xk = times_ik(Xk, (T)(k * a.kscale_x));
            Xm = times_ik(Xm, (T)((M - k) * a.kscale_x));
          }
          z[q] = c2r_pre(Xk, Xm, (row_tw_lds<NX>() ? twl[k] : a.tw[k * a.twstride]), s, k == 0);
        }
        if constexpr (row_tw_lds<NX>()) line_fft<T, M, R, +1, Lay, kTwh>(z, lds, lay, c, j, twl, 2);
        else line_fft<T, M, R, +1, Lay>(z, lds, lay, c, j, a.tw, 2 * a.twstride);
#pragma unroll
        for (int q = 0; q < R; ++q) {
          cpx<T> uu;
          if constexpr (kPre) uu = fv[q];
          else uu = ld_pol<row_nt<NL>()>(fl + po + ((uint32_t)(j + q * P) << lb));
          if (d == 0) {
            f[q] = {-uu.re * z[q].re, -uu.im * z[q].im};
          } else {
            f[q] = {f[q].re - uu.re * z[q].re, f[q].im - uu.im * z[q].im};
          }
        }
      }
    }
    // forward half-length FFT of f[2m] + i f[2m+1]
    if constexpr (row_tw_lds<NX>()) line_fft<T, M, R, -1, Lay, kTwh>(f, lds, lay, c, j, twl, 2);
    else line_fft<T, M, R, -1, Lay>(f, lds, lay, c, j, a.tw, 2 * a.twstride);
    // exchange Z[M-k] through LDS
    __syncthreads();
#pragma unroll
    for (int q = 0; q < R; ++q) lds[lay.idx(c, j + q * P)] = f[q];
    __syncthreads();
    // the post-processed outputs of all elements before any store: a twiddle load placed
    // after a store could not be issued ahead of it (the arrays may alias), which would
    // cost one memory round trip per element
    cpx<T> Xs[R];
#pragma unroll
    for (int q = 0; q < R; ++q) {
      const int k = j + q * P;
      const cpx<T> w = (row_tw_lds<NX>() ? twl[k] : a.tw[k * a.twstride]);
      const cpx<T> Zm = lds[lay.idx(c, M - k)];   // row entry M exists (k = 0 discards it)
      Xs[q] = k == 0 ? cpx<T>{f[q].re + f[q].im, (T)0} : r2c_post(f[q], Zm, w);
    }
    if (valid) {
      const int iky = (int)(row % a.ny), ikz = (int)(row / a.ny);
#pragma unroll
      for (int q = 0; q < R; ++q) {
        const int k = j + q * P;
        const cpx<T> X = Xs[q];
        // only 1D problems finish a stage in the x kernel (x is their last forward axis)
        constexpr bool kEpi = MODE == ROW_ADV && NIN == 1;
        bool epi = false;
        if constexpr (kEpi) epi = a.st.kind != EPI_STORE;
        if constexpr (kEpi) {
          if (epi) (void)stage_epilogue(a.st, sub, so_st + ((uint32_t)k << lb), member, k, iky, ikz, X);
        }
        if (!epi) (a.out + sub)[so_st + ((uint32_t)k << lb)] = X;
        if (k == 0) {
          const cpx<T> XM = {f[q].re - f[q].im, (T)0};
          if constexpr (kEpi) {
            if (epi) (void)stage_epilogue(a.st, sub, so_st + ((uint32_t)M << lb), member, M, iky, ikz, XM);
          }
          if (!epi) (a.out + sub)[so_st + ((uint32_t)M << lb)] = XM;
        }
      }
    }
    (void)zero;
  }
}

// ROW_ADV (spectral output, no stage epilogue) with the NIN input rows of a block
// transformed side by side.  The block is NIN groups of NL * M/R threads (whole
// wavefronts); group d loads its spectral rows and flow values at once and
// forms the product flow_d * C2R(D_d in_d) for the block's NL rows.  Groups
// 1.. hand their products over through LDS; group 0 forms ((-u cx) - v cy) - w cz
// and performs the R2C, while the other groups only take part in its workgroup
// barriers (the same number, counted from the R2C's radix plan), so every wavefront
// of the block reaches every barrier and none ends early.  Compared with row_kernel
// this trades idle lanes in the R2C for fewer registers per lane and all loads of a
// block in flight together.
template <typename T, int NX, int R, int NL, int NIN>
constexpr int adv_par_threads() { return NIN * NL * ((NX / 2) / R); }
template <typename T, int NX, int R, int NL, int NIN>
constexpr int adv_par_lds() { return NIN * NL * (NX / 2 + 1) * (int)sizeof(cpx<T>) + row_tw_bytes<T, NX>(); }

template <typename T, int NX, int R, int NL, int NIN, int OCC = 0>
__global__ __attribute__((amdgpu_flat_work_group_size(1, adv_par_threads<T, NX, R, NL, NIN>()),
                          amdgpu_waves_per_eu(
                              min_waves(adv_par_threads<T, NX, R, NL, NIN>(), adv_par_lds<T, NX, R, NL, NIN>(), OCC),
                              lds_waves(adv_par_threads<T, NX, R, NL, NIN>(), adv_par_lds<T, NX, R, NL, NIN>()))))
void row_adv_par_kernel(RowArgs<T> a) {
  constexpr int M = NX / 2;
  constexpr int NXH = M + 1;
  constexpr int P = M / R;
  constexpr int G = NL * P;   // threads per group
  static_assert(G % 64 == 0, "a group must be whole wavefronts");
  static_assert(NIN >= 2 && NIN <= 3, "two or three inputs");
  using Lay = LayStrided;
  const Lay lay = a.lb ? Lay{1, NL} : Lay{M + 1, 1};
  __shared__ cpx<T> lds_all[NIN * NL * (M + 1)];
  constexpr bool kTwh = row_tw_half<T, NX>();
  __shared__ cpx<T> twl[row_tw_lds<NX>() ? row_tw_elems<T, NX>() : 1];
  if constexpr (row_tw_lds<NX>()) {
    // every group fills (before any group ends); read after the first row exchange
    for (int i = (int)threadIdx.x; i < row_tw_elems<T, NX>(); i += (int)blockDim.x) twl[i] = a.tw[i * a.twstride];
  }
  const int grp = (int)threadIdx.x / G;   // uniform over a wavefront
  const int t = (int)threadIdx.x - grp * G;
  cpx<T>* lds = lds_all + grp * (NL * (M + 1));
  // lane and block mapping as in row_kernel
  const int lb = a.lb;
  const int lt = lb < ilog2(NL) ? lb : ilog2(NL);
  int bx = blockIdx.x, member = blockIdx.y;
  if (lt < lb) {
    int bz;
    xcd_remap(bx, member, bz);
  }
  const int rb = t & ((1 << lt) - 1);
  const int rest = t >> lt;
  const int j = rest % P;
  const int g = rest / P;
  const int c = (g << lt) + rb;
  const long long row0 = (long long)bx * NL;
  const long long row = row0 + c;
  const bool valid = row < a.nrows;
  const int gmax = (int)((a.nrows - row0) >> lt) - 1;
  const int gc = valid ? g : (g < gmax ? g : gmax);
  const long long sub = (long long)member * a.smstride + brick_row_base(row0, lb, NXH);
  const long long pub = (long long)member * a.pmstride + 2 * brick_row_base(row0, lb, M);
  const uint32_t so = ((uint32_t)gc * NXH << lb) + rb;
  const uint32_t so_st = ((uint32_t)g * NXH << lb) + rb;
  const uint32_t po = ((uint32_t)gc * M << lb) + rb;
  const T s = (T)a.scale;
  // select, not a.in[grp]: a runtime index into the by-value arguments would go to scratch
  const cpx<T>* in = grp == 0 ? a.in[0] : (grp == 1 ? a.in[1] : a.in[2]);
  const T* flp = grp == 0 ? a.flow[0] : (grp == 1 ? a.flow[1] : a.flow[2]);
  cpx<T> z[R], last = {(T)0, (T)0};
  load_row<T, M, R, row_nt<NL>()>(in + sub + so, z, last, j, lb);
  const cpx<T>* fl = reinterpret_cast<const cpx<T>*>(flp + pub);
  cpx<T> fv[R];
#pragma unroll
  for (int q = 0; q < R; ++q) fv[q] = ld_pol<row_nt<NL>()>(fl + po + ((uint32_t)(j + q * P) << lb));
  cpx<T> zm[R];
  stage_row_pairs<T, M, R, Lay>(z, last, zm, lds, lay, c, j);
#pragma unroll
  for (int q = 0; q < R; ++q) {
    const int k = j + q * P;
    cpx<T> Xk = z[q];
    cpx<T> Xm = zm[q];
    if (grp == 0) {
      Xk = times_ik(Xk, (T)(k * a.kscale_x));
      Xm = times_ik(Xm, (T)((M - k) * a.kscale_x));
    }
    z[q] = c2r_pre(Xk, Xm, (row_tw_lds<NX>() ? twl[k] : a.tw[k * a.twstride]), s, k == 0);
  }
  if constexpr (row_tw_lds<NX>()) line_fft<T, M, R, +1, Lay, kTwh>(z, lds, lay, c, j, twl, 2);
  else line_fft<T, M, R, +1, Lay>(z, lds, lay, c, j, a.tw, 2 * a.twstride);
  cpx<T> f[R];
#pragma unroll
  for (int q = 0; q < R; ++q) f[q] = {fv[q].re * z[q].re, fv[q].im * z[q].im};
  __syncthreads();  // every group's last line_fft reads are done
  if (grp != 0) {
#pragma unroll
    for (int q = 0; q < R; ++q) lds[lay.idx(c, j + q * P)] = f[q];
  }
  __syncthreads();
  if (grp != 0) {
    // the barriers of group 0's forward transform and of its Z[M-k] exchange below
    constexpr int kBarriers = fft_barriers(M, R) + 2;
    for (int i = 0; i < kBarriers; ++i) __syncthreads();
    return;
  }
#pragma unroll
  for (int q = 0; q < R; ++q) {
    const int i = lay.idx(c, j + q * P);
    cpx<T> v = {-f[q].re, -f[q].im};
    const cpx<T> p1 = lds_all[NL * (M + 1) + i];
    v = {v.re - p1.re, v.im - p1.im};
    if constexpr (NIN == 3) {
      const cpx<T> p2 = lds_all[2 * NL * (M + 1) + i];
      v = {v.re - p2.re, v.im - p2.im};
    }
    f[q] = v;
  }
  // forward half-length FFT of f[2m] + i f[2m+1] (the other groups match its barriers)
  if constexpr (row_tw_lds<NX>()) line_fft<T, M, R, -1, Lay, kTwh>(f, lds, lay, c, j, twl, 2);
  else line_fft<T, M, R, -1, Lay>(f, lds, lay, c, j, a.tw, 2 * a.twstride);
  __syncthreads();
#pragma unroll
  for (int q = 0; q < R; ++q) lds[lay.idx(c, j + q * P)] = f[q];
  __syncthreads();
  // all outputs before any store (twiddle loads ahead of the stores, as in row_kernel)
  cpx<T> Xs[R];
#pragma unroll
  for (int q = 0; q < R; ++q) {
    const int k = j + q * P;
    const cpx<T> w = (row_tw_lds<NX>() ? twl[k] : a.tw[k * a.twstride]);
    const cpx<T> Zm = lds[lay.idx(c, M - k)];   // row entry M exists (k = 0 discards it)
    Xs[q] = k == 0 ? cpx<T>{f[q].re + f[q].im, (T)0} : r2c_post(f[q], Zm, w);
  }
  if (valid) {
    if (j == 0) (a.out + sub)[so_st + ((uint32_t)M << lb)] = cpx<T>{f[0].re - f[0].im, (T)0};
#pragma unroll
    for (int q = 0; q < R; ++q) (a.out + sub)[so_st + ((uint32_t)(j + q * P) << lb)] = Xs[q];
  }
}

}  // namespace ptf
