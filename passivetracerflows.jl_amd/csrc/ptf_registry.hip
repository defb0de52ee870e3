// Kernel-variant registry.  This file is compiled once per (precision, kernel kind,
// part), with -DPTF_T=<double|float> -DPTF_TNAME=<f64|f32> -DPTF_KIND=<k>
// -DPTF_PART=<p> -DPTF_NPARTS=<n> (the line lengths of a kind are dealt round-robin
// over its n parts so that the heavy kinds compile in parallel), and exports
//   const void* ptf_reg_<TNAME>_<KIND>_<PART>(int N, int R, int W)
// returning the kernel for line length N (column kinds: axis length; row kinds: nx),
// radix R and width W (columns per block TX, or rows per block NL; plus
// 100 for the input-parallel advection form and 1000 * occ for the variants
// built with a 3-waves/SIMD budget), or NULL
// when that variant does not exist.  Only variants that satisfy the launch
// and LDS limits are instantiated.
#include "ptf_kernels.hpp"

using namespace ptf;

#ifndef PTF_T
#error "PTF_T must be defined"
#endif
#ifndef PTF_PART
#define PTF_PART 0
#endif
#ifndef PTF_NPARTS
#define PTF_NPARTS 1
#endif

namespace {

// kinds
// (the stage kinds 4 and 6 are the ETDRK4 / plain-store family, 7 and 8 the same kernels
// compiled for the RK4 family)
constexpr int K_COL_FWD = 0, K_COL_INV = 1, K_COL_D1 = 2, K_COL_D2 = 3, K_COL_STAGE = 4,
              K_COL_FLOWSYN = 5, K_COL_STAGE_D1 = 6, K_COL_STAGE_RK4 = 7, K_COL_STAGE_D1_RK4 = 8;
constexpr int K_ROW_C2R = 10, K_ROW_R2C = 11, K_ROW_ADV1 = 12, K_ROW_ADV2 = 13, K_ROW_ADV3 = 14;

constexpr int col_mode(int k) {
  return k == K_COL_FWD || k == K_COL_INV ? COL_PLAIN
         : k == K_COL_D1                  ? COL_D1
         : k == K_COL_D2                  ? COL_D2
         : k == K_COL_STAGE || k == K_COL_STAGE_RK4       ? COL_STAGE
         : k == K_COL_STAGE_D1 || k == K_COL_STAGE_D1_RK4 ? COL_STAGE_D1
                                          : COL_FLOWSYN;
}
constexpr int col_dir(int k) {
  return (k == K_COL_FWD || col_mode(k) == COL_STAGE || col_mode(k) == COL_STAGE_D1) ? -1 : +1;
}
// the epilogue family a kind is compiled for (FP64 RK4: every epilogue, which spilled
// less than the RK4-only build under the 3-waves/SIMD budget)
template <typename T>
constexpr int col_epf(int k) {
  return (k == K_COL_STAGE_RK4 || k == K_COL_STAGE_D1_RK4) ? (sizeof(T) == 8 ? EPF_ANY : EPF_RK4)
         : (k == K_COL_STAGE || k == K_COL_STAGE_D1)       ? EPF_ETD
                                                           : EPF_NONE;
}

constexpr int row_mode(int k) { return k == K_ROW_C2R ? ROW_C2R : k == K_ROW_R2C ? ROW_R2C : ROW_ADV; }
constexpr int row_nin(int k) { return k == K_ROW_ADV2 ? 2 : k == K_ROW_ADV3 ? 3 : 1; }

// kinds that are also built with a 3-waves/SIMD register budget (OCC = 3)
constexpr bool has_occ(int k) { return k == 3 || k == 6 || k == 8 || k == 13 || k == 14; }

// Lengths that are not powers of two are compiled with one register layout (R =
// line_e(N) elements per thread), the default register budget, and one width: the
// largest power of two (at most wmax) that keeps a block within 128 threads.
constexpr bool is_pow2_len(int n) { return n > 0 && (n & (n - 1)) == 0; }
constexpr int reduced_width(int P, int wmax) {
  int w = 1;
  while (w * 2 <= wmax && w * 2 * P <= 128) w *= 2;
  return w;
}
constexpr bool reduced_ok(int N, int R, int width, int wmax, int P, int OCC) {
  return R == line_e(N) && OCC == 0 && width == reduced_width(P, wmax);
}

template <typename T, int N, int R, int TX, int KIND, int OCC>
const void* colp() {
  constexpr int P = N / R;
  constexpr bool ok = R <= N && (R >= 4 || !is_pow2_len(N)) && N % R == 0 && fft_ok(N, R) &&
                      TX * P >= 16 && TX * P <= 1024 &&
                      (is_pow2_len(N) || reduced_ok(N, R, TX, 64, P, OCC)) &&
                      (long long)col_lds_elems<T, N, R, TX, col_mode(KIND)>() * (long long)sizeof(cpx<T>) <= 163840;
  if constexpr (ok) {
    return reinterpret_cast<const void*>(
        &col_kernel<T, N, R, TX, col_mode(KIND), col_dir(KIND), OCC, col_epf<T>(KIND)>);
  } else {
    return nullptr;
  }
}

template <typename T, int NX, int R, int NL, int KIND, int OCC>
const void* rowp() {
  constexpr int M = NX / 2;
  constexpr int P = M / R;
  constexpr bool ok = R <= M && (R >= 4 || !is_pow2_len(M)) && M % R == 0 && fft_ok(M, R) &&
                      NL * P >= 16 && NL * P <= 1024 &&
                      (is_pow2_len(M) || reduced_ok(M, R, NL, 16, P, OCC)) &&
                      (long long)NL * (M + 1) * (long long)sizeof(cpx<T>) + row_tw_bytes<T, NX>() <= 163840;
  if constexpr (ok) {
    return reinterpret_cast<const void*>(&row_kernel<T, NX, R, NL, row_mode(KIND), row_nin(KIND), OCC>);
  } else {
    return nullptr;
  }
}

// input-parallel advection (row_adv_par_kernel), kinds 13 and 14
template <typename T, int NX, int R, int NL, int KIND, int OCC>
const void* rowpar() {
  if constexpr (KIND == K_ROW_ADV2 || KIND == K_ROW_ADV3) {
    constexpr int NIN = row_nin(KIND);
    constexpr int M = NX / 2;
    constexpr int P = M / R;
    constexpr bool ok = is_pow2_len(M) && R <= M && R >= 4 && (NL * P) % 64 == 0 &&
                        NIN * NL * P <= 1024 &&
                        adv_par_lds<T, NX, R, NL, NIN>() <= 163840;
    if constexpr (ok) {
      return reinterpret_cast<const void*>(&row_adv_par_kernel<T, NX, R, NL, NIN, OCC>);
    } else {
      return nullptr;
    }
  } else {
    return nullptr;
  }
}

// pipelined persistent stage-boundary kernel (col_pipe_kernel), kinds 6 and 8 with par = 1:
// power-of-two lines of 64..1024 points, radix 4 or 8, up to 4 columns per block, the
// default register budget (it holds two tiles' data)
template <typename T, int N, int R, int TX, int OCC, int EPF>
const void* pipep() {
  constexpr int P = N / R;
  constexpr bool ok = OCC == 0 && is_pow2_len(N) && N >= 64 && N <= 1024 && (R == 4 || R == 8) && TX <= 4 &&
                      TX * P >= 32 && TX * P <= 512 &&
                      (long long)col_lds_elems<T, N, R, TX, COL_PLAIN>() * (long long)sizeof(cpx<T>) <= 65536;
  if constexpr (ok) {
    return reinterpret_cast<const void*>(&col_pipe_kernel<T, N, R, TX, OCC, EPF>);
  } else {
    return nullptr;
  }
}

template <typename T, int N, int R, int KIND, int OCC>
const void* by_w_par(int W) {
  if constexpr (KIND == K_COL_STAGE_D1 || KIND == K_COL_STAGE_D1_RK4) {
    switch (W) {
      case 1: return pipep<T, N, R, 1, OCC, col_epf<T>(KIND)>();
      case 2: return pipep<T, N, R, 2, OCC, col_epf<T>(KIND)>();
      case 4: return pipep<T, N, R, 4, OCC, col_epf<T>(KIND)>();
      default: return nullptr;
    }
  } else if constexpr (KIND >= 10) {
    switch (W) {
      case 1: return rowpar<T, N, R, 1, KIND, OCC>();
      case 2: return rowpar<T, N, R, 2, KIND, OCC>();
      case 4: return rowpar<T, N, R, 4, KIND, OCC>();
      case 8: return rowpar<T, N, R, 8, KIND, OCC>();
      case 16: return rowpar<T, N, R, 16, KIND, OCC>();
      default: return nullptr;
    }
  } else {
    return nullptr;
  }
}

template <typename T, int N, int R, int KIND, int OCC>
const void* by_w_occ(int W) {
  if constexpr (KIND < 10) {
    switch (W) {
      case 1:   // one line per block: long lines only (N >= 512), where LDS caps the blocks per CU
        if constexpr (N >= 512) return colp<T, N, R, 1, KIND, OCC>();
        return nullptr;
      case 2: return colp<T, N, R, 2, KIND, OCC>();
      case 4: return colp<T, N, R, 4, KIND, OCC>();
      case 8: return colp<T, N, R, 8, KIND, OCC>();
      case 16: return colp<T, N, R, 16, KIND, OCC>();
      case 32: return colp<T, N, R, 32, KIND, OCC>();
      case 64: return colp<T, N, R, 64, KIND, OCC>();
      default: return nullptr;
    }
  } else {
    switch (W) {
      case 1: return rowp<T, N, R, 1, KIND, OCC>();
      case 2: return rowp<T, N, R, 2, KIND, OCC>();
      case 4: return rowp<T, N, R, 4, KIND, OCC>();
      case 8: return rowp<T, N, R, 8, KIND, OCC>();
      case 16: return rowp<T, N, R, 16, KIND, OCC>();
      default: return nullptr;
    }
  }
}

// W encodes the width, the form and the occupancy request:
// W = width + 100 * par + 1000 * occ  (par = 1: input-parallel advection)
template <typename T, int N, int R, int KIND>
const void* by_w(int W) {
  const int occ = W / 1000, par = (W / 100) % 10, w = W % 100;
  if (par) {
    if (occ == 0) return by_w_par<T, N, R, KIND, 0>(w);
    if constexpr (has_occ(KIND)) {
      if (occ == 3) return by_w_par<T, N, R, KIND, 3>(w);
    }
    return nullptr;
  }
  if (occ == 0) return by_w_occ<T, N, R, KIND, 0>(w);
  if constexpr (has_occ(KIND)) {
    if (occ == 3) return by_w_occ<T, N, R, KIND, 3>(w);
  }
  return nullptr;
}

template <typename T, int N, int KIND>
const void* by_r(int R, int W) {
  constexpr int L = KIND < 10 ? N : N / 2;   // the transformed length (rows: nx/2)
  if constexpr (is_pow2_len(L)) {
    switch (R) {
      case 4: return by_w<T, N, 4, KIND>(W);
      case 8: return by_w<T, N, 8, KIND>(W);
      case 16: return by_w<T, N, 16, KIND>(W);
      default: return nullptr;
    }
  } else if constexpr (line_e(L) > 0) {
    if (R == line_e(L)) return by_w<T, N, line_e(L), KIND>(W);
    return nullptr;
  } else {
    return nullptr;
  }
}

// line length N, the I-th of its kind's list: compiled into part I mod PTF_NPARTS
template <typename T, int N, int KIND, int I>
const void* in_part(int R, int W) {
  if constexpr (I % PTF_NPARTS == PTF_PART) return by_r<T, N, KIND>(R, W);
  else return nullptr;
}

template <typename T, int KIND>
const void* by_n(int N, int R, int W) {
  if constexpr (KIND < 10) {
    switch (N) {
      case 8: return in_part<T, 8, KIND, 0>(R, W);
      case 16: return in_part<T, 16, KIND, 1>(R, W);
      case 32: return in_part<T, 32, KIND, 2>(R, W);
      case 64: return in_part<T, 64, KIND, 3>(R, W);
      case 128: return in_part<T, 128, KIND, 4>(R, W);
      case 256: return in_part<T, 256, KIND, 5>(R, W);
      case 512: return in_part<T, 512, KIND, 6>(R, W);
      case 1024: return in_part<T, 1024, KIND, 7>(R, W);
      case 2048: return in_part<T, 2048, KIND, 8>(R, W);
      // lengths 2^a m, m in {3, 5, 9, 25, 27, 125} (one register layout each)
      case 12: return in_part<T, 12, KIND, 9>(R, W);
      case 18: return in_part<T, 18, KIND, 10>(R, W);
      case 20: return in_part<T, 20, KIND, 11>(R, W);
      case 24: return in_part<T, 24, KIND, 12>(R, W);
      case 36: return in_part<T, 36, KIND, 13>(R, W);
      case 40: return in_part<T, 40, KIND, 14>(R, W);
      case 48: return in_part<T, 48, KIND, 15>(R, W);
      case 50: return in_part<T, 50, KIND, 16>(R, W);
      case 54: return in_part<T, 54, KIND, 17>(R, W);
      case 72: return in_part<T, 72, KIND, 18>(R, W);
      case 80: return in_part<T, 80, KIND, 19>(R, W);
      case 96: return in_part<T, 96, KIND, 20>(R, W);
      case 100: return in_part<T, 100, KIND, 21>(R, W);
      case 108: return in_part<T, 108, KIND, 22>(R, W);
      case 144: return in_part<T, 144, KIND, 23>(R, W);
      case 160: return in_part<T, 160, KIND, 24>(R, W);
      case 192: return in_part<T, 192, KIND, 25>(R, W);
      case 200: return in_part<T, 200, KIND, 26>(R, W);
      case 216: return in_part<T, 216, KIND, 27>(R, W);
      case 250: return in_part<T, 250, KIND, 28>(R, W);
      case 288: return in_part<T, 288, KIND, 29>(R, W);
      case 320: return in_part<T, 320, KIND, 30>(R, W);
      case 384: return in_part<T, 384, KIND, 31>(R, W);
      case 400: return in_part<T, 400, KIND, 32>(R, W);
      case 432: return in_part<T, 432, KIND, 33>(R, W);
      case 500: return in_part<T, 500, KIND, 34>(R, W);
      case 576: return in_part<T, 576, KIND, 35>(R, W);
      case 640: return in_part<T, 640, KIND, 36>(R, W);
      case 768: return in_part<T, 768, KIND, 37>(R, W);
      case 800: return in_part<T, 800, KIND, 38>(R, W);
      case 864: return in_part<T, 864, KIND, 39>(R, W);
      case 1000: return in_part<T, 1000, KIND, 40>(R, W);
      case 1152: return in_part<T, 1152, KIND, 41>(R, W);
      case 1280: return in_part<T, 1280, KIND, 42>(R, W);
      case 1536: return in_part<T, 1536, KIND, 43>(R, W);
      case 1600: return in_part<T, 1600, KIND, 44>(R, W);
      case 1728: return in_part<T, 1728, KIND, 45>(R, W);
      case 2000: return in_part<T, 2000, KIND, 46>(R, W);
      default: return nullptr;
    }
  } else {
    switch (N) {  // nx
      case 16: return in_part<T, 16, KIND, 0>(R, W);
      case 32: return in_part<T, 32, KIND, 1>(R, W);
      case 64: return in_part<T, 64, KIND, 2>(R, W);
      case 128: return in_part<T, 128, KIND, 3>(R, W);
      case 256: return in_part<T, 256, KIND, 4>(R, W);
      case 512: return in_part<T, 512, KIND, 5>(R, W);
      case 1024: return in_part<T, 1024, KIND, 6>(R, W);
      case 2048: return in_part<T, 2048, KIND, 7>(R, W);
      case 4096: return in_part<T, 4096, KIND, 8>(R, W);
      // nx = 2 M with M = 2^a m, m in {3, 5, 9, 25, 27, 125} (one register layout each)
      case 24: return in_part<T, 24, KIND, 9>(R, W);
      case 36: return in_part<T, 36, KIND, 10>(R, W);
      case 40: return in_part<T, 40, KIND, 11>(R, W);
      case 48: return in_part<T, 48, KIND, 12>(R, W);
      case 50: return in_part<T, 50, KIND, 13>(R, W);
      case 54: return in_part<T, 54, KIND, 14>(R, W);
      case 72: return in_part<T, 72, KIND, 15>(R, W);
      case 80: return in_part<T, 80, KIND, 16>(R, W);
      case 96: return in_part<T, 96, KIND, 17>(R, W);
      case 100: return in_part<T, 100, KIND, 18>(R, W);
      case 108: return in_part<T, 108, KIND, 19>(R, W);
      case 144: return in_part<T, 144, KIND, 20>(R, W);
      case 160: return in_part<T, 160, KIND, 21>(R, W);
      case 192: return in_part<T, 192, KIND, 22>(R, W);
      case 200: return in_part<T, 200, KIND, 23>(R, W);
      case 216: return in_part<T, 216, KIND, 24>(R, W);
      case 250: return in_part<T, 250, KIND, 25>(R, W);
      case 288: return in_part<T, 288, KIND, 26>(R, W);
      case 320: return in_part<T, 320, KIND, 27>(R, W);
      case 384: return in_part<T, 384, KIND, 28>(R, W);
      case 400: return in_part<T, 400, KIND, 29>(R, W);
      case 432: return in_part<T, 432, KIND, 30>(R, W);
      case 500: return in_part<T, 500, KIND, 31>(R, W);
      case 576: return in_part<T, 576, KIND, 32>(R, W);
      case 640: return in_part<T, 640, KIND, 33>(R, W);
      case 768: return in_part<T, 768, KIND, 34>(R, W);
      case 800: return in_part<T, 800, KIND, 35>(R, W);
      case 864: return in_part<T, 864, KIND, 36>(R, W);
      case 1000: return in_part<T, 1000, KIND, 37>(R, W);
      case 1152: return in_part<T, 1152, KIND, 38>(R, W);
      case 1280: return in_part<T, 1280, KIND, 39>(R, W);
      case 1536: return in_part<T, 1536, KIND, 40>(R, W);
      case 1600: return in_part<T, 1600, KIND, 41>(R, W);
      case 1728: return in_part<T, 1728, KIND, 42>(R, W);
      case 2000: return in_part<T, 2000, KIND, 43>(R, W);
      case 2304: return in_part<T, 2304, KIND, 44>(R, W);
      case 2560: return in_part<T, 2560, KIND, 45>(R, W);
      case 3072: return in_part<T, 3072, KIND, 46>(R, W);
      case 3200: return in_part<T, 3200, KIND, 47>(R, W);
      case 3456: return in_part<T, 3456, KIND, 48>(R, W);
      case 4000: return in_part<T, 4000, KIND, 49>(R, W);
      default: return nullptr;
    }
  }
}

}  // namespace

#define PTF_JOIN4_(a, b, c, d) a##_##b##_##c##_##d
#define PTF_JOIN4(a, b, c, d) PTF_JOIN4_(a, b, c, d)

extern "C" const void* PTF_JOIN4(ptf_reg, PTF_TNAME, PTF_KIND, PTF_PART)(int N, int R, int W) {
  return by_n<PTF_T, PTF_KIND>(N, R, W);
}
