// Whole-member step kernel for small square 2D grids (ensembles of many small members).
//
// One workgroup advances one member by a whole RK4 or ETDRK4 step: the y-inverse,
// x-advection and y-forward passes of all four stages run on two LDS-resident
// half-spectrum planes, so only the stage state (sol, acc, stage inputs, ETDRK4
// coefficients: the same operands, read and written by the same rk4_/etd_ functions as
// the column stage kernels) and the flow cross HBM.  The transform intermediates that the
// pass kernels exchange through memory (two spectral fields out of every inverse pass,
// the x pass's product) stay in LDS.
//
// LDS: planes A and B of N rows of NXH = N/2 + 1 values, row-major (kx fastest) with a
// row stride SR >= NXH chosen so that the lanes of a column exchange (stride SR) and of a
// row exchange (stride 1 and the Stockham strides) spread over the banks.  Column kx is
// read as A[ky * SR + kx], row y as A[y * SR + k].  Column lines (ky) and row lines
// (half-length x transforms of M = N/2 points) are held by consecutive lanes of one
// wavefront, and a line's own entries of A or B serve as its exchange space, so the
// transforms order their LDS accesses per wavefront; workgroup barriers separate the
// column and row phases.
//
// Per stage (calcN! S:756-769 for a steady 2D flow, then the stage update):
//   columns: A = IFFT_y(x), B = IFFT_y(i l x)       (stage input x; stage 1: x = sol)
//   rows:    A = R2C_x(-(u C2R_x(i kr A) + v C2R_x(B)))
//   columns: N = FFT_y(A) -> stage update -> next stage input x (kept in registers)
#pragma once
#include "ptf_kernels.hpp"

namespace ptf {

template <typename T>
struct MemberArgs {
  StageArgs<T> st;         // state, operators, stepper (as the column stage kernels)
  const T* flow[2];        // physical u, v of the group, members back to back
  const cpx<T>* tw;        // exp(-2 pi i m / N), m < N
  long long pmstride;      // physical values per member
  double kscale_x, kscale_y;
  double scale;            // C2R scale 1 / (N N)
  int lb;                  // y stored in bricks of 2^lb rows (spectral and physical)
};

// threads: enough for every column (NXH lines of N/RC lanes) and every row (N lines of
// (N/2)/RR lanes) at once, in whole wavefronts
template <int N, int RC, int RR>
constexpr int member_threads() {
  constexpr int nct = (N / 2 + 1) * (N / RC);
  constexpr int nrt = N * ((N / 2) / RR);
  constexpr int m = nct > nrt ? nct : nrt;
  return (m + 63) / 64 * 64;
}
// LDS row stride of the planes (from a 64-bank model of the column and row exchanges of
// the compiled configurations; N/2 + 1 otherwise)
template <typename T, int N>
constexpr int member_row_stride() {
  return sizeof(T) == 4 ? (N == 64 ? 39 : (N == 128 ? 67 : N / 2 + 1)) : (N == 64 ? 35 : N / 2 + 1);
}
template <typename T, int N>
constexpr int member_lds_bytes() {
  return (int)((2 * N * member_row_stride<T, N>() + N) * sizeof(cpx<T>));
}
// Register budget: two workgroups per CU where the LDS and the waves admit it (so that one
// hides the other's latency and a group of more members than CUs runs in one round),
// otherwise one.
template <typename T, int N, int RC, int RR>
constexpr int member_min_waves() {
  constexpr int waves = member_threads<N, RC, RR>() / 64;
  constexpr int blocks = (2 * member_lds_bytes<T, N>() <= 163840 && 2 * waves <= 12) ? 2 : 1;
  return (blocks * waves + 3) / 4;
}

// the LDS bound on waves per SIMD, rounded up where a workgroup's waves do not divide
// evenly over the 4 SIMDs
template <typename T, int N, int RC, int RR>
constexpr int member_max_waves() {
  constexpr int w = lds_waves(member_threads<N, RC, RR>(), member_lds_bytes<T, N>());
  constexpr int m = member_min_waves<T, N, RC, RR>();
  return w > m ? w : m;
}

template <typename T, int N, int RC, int RR>
__global__ __attribute__((amdgpu_flat_work_group_size(1, member_threads<N, RC, RR>()),
                          amdgpu_waves_per_eu(member_min_waves<T, N, RC, RR>(),
                                              member_max_waves<T, N, RC, RR>())))
void member_step_kernel(MemberArgs<T> a) {
  constexpr int M = N / 2;
  constexpr int NXH = M + 1;
  constexpr int SR = member_row_stride<T, N>();
  constexpr int PC = N / RC;      // lanes per column line
  constexpr int PR = M / RR;      // lanes per row line
  constexpr int NCT = NXH * PC;
  constexpr int NRT = N * PR;
  static_assert(64 % PC == 0 && 64 % PR == 0, "a line within one wavefront");
  static_assert(fft_ok(N, RC) && fft_ok(M, RR), "radix plans");
  __shared__ cpx<T> A[N * SR];
  __shared__ cpx<T> B[N * SR];
  __shared__ cpx<T> twl[N];
  const int t = threadIdx.x;
  const int member = blockIdx.x;
  for (int i = t; i < N; i += (int)blockDim.x) twl[i] = a.tw[i];
  __syncthreads();

  const long long ubase = (long long)member * a.st.mstride;
  const long long pbase = (long long)member * a.pmstride;
  const int lb = a.lb;
  const int mask = (1 << lb) - 1;
  const LayStrided layc{1, SR};   // column kx, element ky
  const LayStrided layr{SR, 1};   // row y, element k
  const bool colt = t < NCT;       // column lanes: kx = t / PC, elements ky = jc + q PC
  const bool rowt = t < NRT;       // row lanes: y = t / PR, elements k = jr + q PR
  const OpParams opm = a.st.op[member];
  // Each phase derives its lane indices from an opaque copy of the thread index, so that
  // the compiler recomputes the (cheap) addresses of a phase inside it instead of keeping
  // those of all phases live across the stage loop.
  auto lane = [&]() {
    int tt = t;
    __asm__ volatile("" : "+v"(tt));
    return tt;
  };
  // spectral element (kx, ky) in the group's layout (y in bricks of 2^lb rows)
  auto soff = [&](int kx, int ky) -> uint32_t {
    return ((uint32_t)((ky >> lb) * NXH + kx) << lb) + (uint32_t)(ky & mask);
  };

  // A[:, kx] = IFFT_y(x), B[:, kx] = IFFT_y(i l x); each transform exchanges through its
  // own column of the plane it then writes
  auto col_inverse = [&](int kx, int jc, cpx<T>(&x)[RC]) {
    cpx<T> w[RC];
#pragma unroll
    for (int q = 0; q < RC; ++q)
      w[q] = times_ik(x[q], (T)(fftfreq_index(jc + q * PC, N) * a.kscale_y));
    line_fft<T, N, RC, +1, LayStrided>(w, B, layc, kx, jc, twl, 1, true);
    lds_sync(true);
#pragma unroll
    for (int q = 0; q < RC; ++q) B[(jc + q * PC) * SR + kx] = w[q];
    line_fft<T, N, RC, +1, LayStrided>(x, A, layc, kx, jc, twl, 1, true);
    lds_sync(true);
#pragma unroll
    for (int q = 0; q < RC; ++q) A[(jc + q * PC) * SR + kx] = x[q];
  };

  // the linear operator of this lane's column elements, the same in every stage (kept in
  // registers for up to 8 elements per lane; longer columns evaluate it per stage)
  constexpr bool kLq = RC <= 8;
  T Lq[kLq ? RC : 1];
  if (kLq && colt) {
    const int tt = lane();
    const int kx = tt / PC, jc = tt % PC;
#pragma unroll
    for (int q = 0; q < (kLq ? RC : 0); ++q)
      Lq[q] = (T)linear_op(a.st.g, opm, a.st.nkappah, kx, jc + q * PC, 0);
  }
  // Column phase of stage ST: N = FFT_y(A[:, kx]), the stage update of this lane's
  // elements, and (ST < 3) the next stage's inverse pair.  Where registers allow (at most
  // 128 bytes of RK4 operands per lane), the RK4 operands are loaded before the transform,
  // so that their latency overlaps it; otherwise in chunks after it (one memory round trip
  // per chunk; FP64 and ETDRK4, up to 9 operands per element, in chunks of 2).
  auto col_stage = [&](auto stc) {
    constexpr int ST = decltype(stc)::value;
    constexpr bool kPre = RC * 4 * (int)sizeof(cpx<T>) <= 128;
    constexpr int kChunk = RC < 2 ? RC : (sizeof(T) == 8 ? 2 : 4);
    constexpr int kChunkE = RC < 2 ? RC : 2;
    const int tt = lane();
    const int kx = tt / PC, jc = tt % PC;
    const bool rk4 = a.st.kind == EPI_RK4;
    cpx<T> x[RC];
#pragma unroll
    for (int q = 0; q < RC; ++q) x[q] = A[(jc + q * PC) * SR + kx];
    Rk4Ops<T> pre[kPre ? RC : 1];
    if constexpr (kPre) {
      if (rk4) {
#pragma unroll
        for (int q = 0; q < RC; ++q) pre[q] = rk4_load<T, ST>(a.st, ubase, soff(kx, jc + q * PC));
      }
    }
    line_fft<T, N, RC, -1, LayStrided>(x, A, layc, kx, jc, twl, 1, true);
    if (rk4) {
#pragma unroll
      for (int q0 = 0; q0 < RC; q0 += kChunk) {
        Rk4Ops<T> o[kChunk];
        if constexpr (!kPre) {
#pragma unroll
          for (int i = 0; i < kChunk; ++i) o[i] = rk4_load<T, ST>(a.st, ubase, soff(kx, jc + (q0 + i) * PC));
        }
#pragma unroll
        for (int i = 0; i < kChunk; ++i) {
          const int ky = jc + (q0 + i) * PC;
          Rk4Ops<T> oq;
          if constexpr (kPre) oq = pre[q0 + i];
          else oq = o[i];
          const T L = kLq ? Lq[kLq ? q0 + i : 0] : (T)linear_op(a.st.g, opm, a.st.nkappah, kx, ky, 0);
          x[q0 + i] = rk4_finish_L<T, ST>(a.st, ubase, soff(kx, ky), L, kx, ky, 0, x[q0 + i], oq);
        }
      }
    } else {
#pragma unroll
      for (int q0 = 0; q0 < RC; q0 += kChunkE) {
        EtdOps<T> o[kChunkE];
#pragma unroll
        for (int i = 0; i < kChunkE; ++i) o[i] = etd_load<T, ST>(a.st, ubase, soff(kx, jc + (q0 + i) * PC));
#pragma unroll
        for (int i = 0; i < kChunkE; ++i) {
          const int ky = jc + (q0 + i) * PC;
          x[q0 + i] = etd_finish<T, ST>(a.st, ubase, soff(kx, ky), kx, ky, 0, x[q0 + i], o[i]);
        }
      }
    }
    if constexpr (ST < 3) col_inverse(kx, jc, x);
  };

  if (colt) {
    const int tt = lane();
    const int kx = tt / PC, jc = tt % PC;
    cpx<T> x[RC];
#pragma unroll
    for (int q = 0; q < RC; ++q) x[q] = a.st.sol[ubase + soff(kx, jc + q * PC)];
    col_inverse(kx, jc, x);
  }
  __syncthreads();
#pragma unroll 1
  for (int stage = 0; stage < 4; ++stage) {
    // rows: A[y, :] = R2C(-(u c_x + v c_y)), c_x = C2R(i kr A[y, :]), c_y = C2R(B[y, :])
    if (rowt) {
      const int tt = lane();
      const int y = tt / PR, jr = tt % PR;
      auto poff = [&](int m) -> uint32_t {   // physical x-pair m of row y
        return ((uint32_t)((y >> lb) * M + m) << lb) + (uint32_t)(y & mask);
      };
      const T s = (T)a.scale;
      const cpx<T>* u = reinterpret_cast<const cpx<T>*>(a.flow[0] + pbase);
      const cpx<T>* v = reinterpret_cast<const cpx<T>*>(a.flow[1] + pbase);
      // The row's flow values are loaded before the transforms that need them where
      // registers allow (u and v: at most 64 bytes per lane each), else after them.
      constexpr bool kPreFlow = RR * (int)sizeof(cpx<T>) <= 64;
      cpx<T> uu[RR], vv[RR];
      if constexpr (kPreFlow) {
#pragma unroll
        for (int q = 0; q < RR; ++q) uu[q] = u[poff(jr + q * PR)];
#pragma unroll
        for (int q = 0; q < RR; ++q) vv[q] = v[poff(jr + q * PR)];
      }
      cpx<T> z[RR], f[RR];
      // c_x from the row of A (whose entries are the transform's exchange space afterwards)
#pragma unroll
      for (int q = 0; q < RR; ++q) {
        const int k = jr + q * PR;
        const cpx<T> Xk = times_ik(A[y * SR + k], (T)(k * a.kscale_x));
        const cpx<T> Xm = times_ik(A[y * SR + (M - k)], (T)((M - k) * a.kscale_x));
        z[q] = c2r_pre(Xk, Xm, twl[k], s, k == 0);
      }
      line_fft<T, M, RR, +1, LayStrided>(z, A, layr, y, jr, twl, 2, true);
      if constexpr (!kPreFlow) {
#pragma unroll
        for (int q = 0; q < RR; ++q) uu[q] = u[poff(jr + q * PR)];
      }
#pragma unroll
      for (int q = 0; q < RR; ++q) f[q] = {-uu[q].re * z[q].re, -uu[q].im * z[q].im};
      if constexpr (!kPreFlow) {
#pragma unroll
        for (int q = 0; q < RR; ++q) vv[q] = v[poff(jr + q * PR)];
      }
      // c_y from the row of B
#pragma unroll
      for (int q = 0; q < RR; ++q) {
        const int k = jr + q * PR;
        z[q] = c2r_pre(B[y * SR + k], B[y * SR + (M - k)], twl[k], s, k == 0);
      }
      line_fft<T, M, RR, +1, LayStrided>(z, B, layr, y, jr, twl, 2, true);
#pragma unroll
      for (int q = 0; q < RR; ++q) f[q] = {f[q].re - vv[q].re * z[q].re, f[q].im - vv[q].im * z[q].im};
      // forward half-length FFT of f[2m] + i f[2m+1], then X[k] from Z[k] and Z[M - k]
      line_fft<T, M, RR, -1, LayStrided>(f, A, layr, y, jr, twl, 2, true);
      lds_sync(true);
#pragma unroll
      for (int q = 0; q < RR; ++q) A[y * SR + jr + q * PR] = f[q];
      lds_sync(true);
      cpx<T> X[RR];
#pragma unroll
      for (int q = 0; q < RR; ++q) {
        const int k = jr + q * PR;
        const cpx<T> Zm = A[y * SR + (M - k)];   // entry M exists (k = 0 discards it)
        X[q] = k == 0 ? cpx<T>{f[q].re + f[q].im, (T)0} : r2c_post(f[q], Zm, twl[k]);
      }
      lds_sync(true);
#pragma unroll
      for (int q = 0; q < RR; ++q) {
        const int k = jr + q * PR;
        A[y * SR + k] = X[q];
        if (k == 0) A[y * SR + M] = cpx<T>{f[q].re - f[q].im, (T)0};
      }
    }
    __syncthreads();
    // columns: FFT_y, stage update, and the next stage's inverse pair
    if (colt) {
      switch (stage) {
        case 0: col_stage(std::integral_constant<int, 0>{}); break;
        case 1: col_stage(std::integral_constant<int, 1>{}); break;
        case 2: col_stage(std::integral_constant<int, 2>{}); break;
        default: col_stage(std::integral_constant<int, 3>{}); break;
      }
    }
    __syncthreads();
  }
}

// The compiled (N, RC, RR) of each precision; null when N has none.  threads = the
// launch's workgroup size.
const void* member_kernel(bool dbl, int n, int* threads);

}  // namespace ptf
