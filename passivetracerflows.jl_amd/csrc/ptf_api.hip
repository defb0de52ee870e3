// Host runtime of libptf_hip: the C ABI declared in include/ptf.h.
//
// A "group" is B independent problems that share one grid (a single problem
// is a group with B = 1; an ensemble is one group per distinct grid size).
// Every group owns its state fields, scratch fields, flows, FFT twiddle tables
// and a per-stage list of kernel launches ("slots") whose arguments are fixed
// at creation, so one step is a plain walk over the slot lists.
#include "ptf.h"
#include "ptf_member.hpp"
#include "ptf_setup.hpp"

#include <dlfcn.h>
#include <hip/hip_runtime.h>
#include <hip/hip_version.h>

#include <algorithm>
#include <cmath>
#include <cstddef>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

using namespace ptf;

// ---------------------------------------------------------------------------
// errors
// ---------------------------------------------------------------------------
static thread_local std::string g_err;

static int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

#define HIPCHK(x)                                                                          \
  do {                                                                                     \
    hipError_t e_ = (x);                                                                   \
    if (e_ != hipSuccess) return fail(PTF_ERR_DEVICE, std::string(#x) + ": " + hipGetErrorString(e_)); \
  } while (0)

#define CHK(x)                  \
  do {                          \
    int r_ = (x);               \
    if (r_ != PTF_OK) return r_; \
  } while (0)

// ---------------------------------------------------------------------------
// kernel registry (ptf_registry.hip, one object per precision and kind)
// ---------------------------------------------------------------------------
enum {
  K_ANALYTIC = 20,  // analytic_flow_kernel (not in the variant registry)
  K_MEMBER = 21,    // member_step_kernel (ptf_member.hip, not in the variant registry)
  K_COL_FWD = 0, K_COL_INV = 1, K_COL_D1 = 2, K_COL_D2 = 3, K_COL_STAGE = 4, K_COL_FLOWSYN = 5,
  K_COL_STAGE_D1 = 6,
  // the stage kernels are compiled per epilogue family: kinds 4 and 6 for ETDRK4 and the
  // plain store, 7 and 8 (the same two kernels) for RK4; stage_col_kind picks one
  K_COL_STAGE_RK4 = 7, K_COL_STAGE_D1_RK4 = 8,
  K_ROW_C2R = 10, K_ROW_R2C = 11, K_ROW_ADV1 = 12, K_ROW_ADV2 = 13, K_ROW_ADV3 = 14
};

// Each (precision, kind) is compiled in up to kMaxParts parts (line lengths dealt
// round-robin); a part that the build does not produce is a weak symbol left NULL.
static constexpr int kMaxParts = 6;
#define PTF_DECL1(t, k, p) extern "C" __attribute__((weak)) const void* ptf_reg_##t##_##k##_##p(int, int, int);
#define PTF_DECL(t, k) \
  PTF_DECL1(t, k, 0) PTF_DECL1(t, k, 1) PTF_DECL1(t, k, 2) PTF_DECL1(t, k, 3) PTF_DECL1(t, k, 4) \
  PTF_DECL1(t, k, 5)
#define PTF_DECL_ALL(t) \
  PTF_DECL(t, 0) PTF_DECL(t, 1) PTF_DECL(t, 2) PTF_DECL(t, 3) PTF_DECL(t, 4) PTF_DECL(t, 5) \
  PTF_DECL(t, 6) PTF_DECL(t, 7) PTF_DECL(t, 8) \
  PTF_DECL(t, 10) PTF_DECL(t, 11) PTF_DECL(t, 12) PTF_DECL(t, 13) PTF_DECL(t, 14)
PTF_DECL_ALL(f64)
PTF_DECL_ALL(f32)

typedef const void* (*RegFn)(int, int, int);

// the registry function of one part (NULL when that part was not built)
static RegFn reg_part(bool dbl, int kind, int part) {
#define PTF_P(t, k) \
  case k: { \
    const RegFn f[kMaxParts] = {ptf_reg_##t##_##k##_0, ptf_reg_##t##_##k##_1, ptf_reg_##t##_##k##_2, \
                                ptf_reg_##t##_##k##_3, ptf_reg_##t##_##k##_4, ptf_reg_##t##_##k##_5}; \
    return f[part]; \
  }
#define PTF_SW(t) \
  switch (kind) { \
    PTF_P(t, 0) PTF_P(t, 1) PTF_P(t, 2) PTF_P(t, 3) PTF_P(t, 4) PTF_P(t, 5) PTF_P(t, 6) \
    PTF_P(t, 7) PTF_P(t, 8) PTF_P(t, 10) PTF_P(t, 11) PTF_P(t, 12) PTF_P(t, 13) PTF_P(t, 14) \
    default: return nullptr; \
  }
  if (part < 0 || part >= kMaxParts) return nullptr;
  if (dbl) {
    PTF_SW(f64)
  } else {
    PTF_SW(f32)
  }
#undef PTF_SW
#undef PTF_P
}

// the kernel of variant (N, R, W) of a kind, from whichever part holds line length N
static const void* reg_lookup(bool dbl, int kind, int N, int R, int W) {
  for (int part = 0; part < kMaxParts; ++part) {
    const RegFn f = reg_part(dbl, kind, part);
    if (!f) continue;
    if (const void* p = f(N, R, W)) return p;
  }
  return nullptr;
}

static bool kind_known(int kind) {
  return (kind >= 0 && kind <= 8) || (kind >= 10 && kind <= 14);
}

static const char* kind_name(int kind) {
  switch (kind) {
    case K_COL_FWD: return "col_fft_fwd";
    case K_COL_INV: return "col_fft_inv";
    case K_COL_D1: return "col_inv_deriv1";
    case K_COL_D2: return "col_inv_deriv2";
    case K_COL_STAGE: return "col_fwd_stage";
    case K_COL_FLOWSYN: return "col_flow_synth";
    case K_COL_STAGE_D1: return "col_fwd_stage_inv";
    case K_COL_STAGE_RK4: return "col_fwd_stage_rk4";
    case K_COL_STAGE_D1_RK4: return "col_fwd_stage_inv_rk4";
    case K_ROW_C2R: return "row_c2r";
    case K_ROW_R2C: return "row_r2c";
    case K_ROW_ADV1: return "row_advect1";
    case K_ROW_ADV2: return "row_advect2";
    case K_ROW_ADV3: return "row_advect3";
    case K_ANALYTIC: return "analytic_flow";
    case K_MEMBER: return "member_step";
  }
  return "?";
}

// elements per thread, tried in this order: the power-of-two layouts, then the one layout
// each other length 2^a 3^b 5^c is compiled with (line_e)
static const int kRadices[] = {16, 8, 4, 12, 20, 30, 6, 10, 15, 3, 5};
static const int kColWidths[] = {16, 8, 32, 64, 4, 2, 1};
static const int kNumColWidths = 7;
static const int kRowWidths[] = {8, 16, 4, 2, 1, 1};
static const int kNumRowWidths = 5;

// compute units of the current device (persistent kernels size their grids by it)
static int device_cus() {
  static int cus = 0;
  if (cus == 0) {
    int dev = 0, n = 0;
    if (hipGetDevice(&dev) == hipSuccess &&
        hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess && n > 0)
      cus = n;
    else
      cus = 256;
  }
  return cus;
}

// ---------------------------------------------------------------------------
// launch slots
// ---------------------------------------------------------------------------
struct Slot {
  std::string name;
  bool dbl = true;
  int kind = 0;
  int N = 0;            // axis length (columns) or nx (rows)
  int R = 0, W = 0;     // radix, columns/rows per block
  int occ = 0;          // 0: default register budget; 3: 3 waves/SIMD
  int par = 0;          // 1: input-parallel advection (row_adv_par_kernel)
  int wg = 0;           // 1: a bricked column line spread over the workgroup (workgroup
                        // barriers) where it would fit one wavefront (see wave_ok)
  const void* fn = nullptr;
  long long lines = 0;  // ncols (columns) or nrows (rows)
  int other = 1;        // grid.y for columns
  int members = 1;
  std::vector<unsigned char> args;  // ColArgs<T> or RowArgs<T>
  bool writes_state = false;        // launch modifies sol (timing must restore it)
  int brick = 0;                    // log2 brick: columns along a bricked line / bricked rows
  int fixed_threads = 0;            // > 0: a kernel outside the registry with this launch shape
  unsigned fixed_blocks = 0;

  bool is_col() const { return kind < 10; }
  // Per-wavefront lines apply to a column kernel along a bricked axis whose P = N/R threads
  // per line divide a wavefront (FP32: at most 32, so that a wavefront still reads whole
  // 128-byte runs); the same test as kWaveOk in col_kernel.  The persistent pipelined
  // kernel keeps its own lane mapping.
  bool wave_ok(int r) const {
    if (!is_col() || brick <= 0 || fixed_threads > 0 || r <= 0 || ((kind == 6 || kind == 8) && par)) return false;
    const int P = N / r;
    return P <= 64 && 64 % P == 0 && (dbl || P <= 32);
  }
  int wave_lines() const { return wave_ok(R) && !wg ? 1 : 0; }
  bool has_occ() const { return kind == 3 || kind == 6 || kind == 8 || kind == 13 || kind == 14; }
  // par: the input-parallel advection form (kinds 13, 14) or the pipelined persistent
  // stage-boundary kernel (kinds 6, 8)
  bool has_par() const { return kind == 6 || kind == 8 || kind == 13 || kind == 14; }
  int threads() const {
    if (fixed_threads > 0) return fixed_threads;
    if (is_col()) return W * (N / R);
    return (par ? (kind == 14 ? 3 : 2) : 1) * W * ((N / 2) / R);
  }
  dim3 grid() const {
    if (fixed_threads > 0) return dim3(fixed_blocks, 1, 1);
    if (is_col() && par) {
      // persistent blocks: two waves per SIMD on every CU, at most one block per tile
      const long long tiles = ((lines + W - 1) / W) * (long long)other * members;
      const long long wpb = (threads() + 63) / 64;
      const long long want = (long long)device_cus() * 8 / wpb;
      return dim3((unsigned)std::max(1LL, std::min(tiles, want)), 1, 1);
    }
    if (is_col()) return dim3((unsigned)((lines + W - 1) / W), (unsigned)other, (unsigned)members);
    return dim3((unsigned)((lines + W - 1) / W), (unsigned)members, 1);
  }
};

static bool choose_variant(Slot& s, int R, int W, int occ = 0, int par = 0, int wg = 0) {
  if (!kind_known(s.kind)) return false;
  if (occ && !s.has_occ()) return false;
  if (par && !s.has_par()) return false;
  if (wg) {
    Slot t = s;
    t.par = par;
    if (!t.wave_ok(R)) return false;   // the workgroup form only where the wavefront one applies
  }
  if (s.brick > 0) {
    // a bricked line needs whole bricks per thread group; bricked rows go in whole bricks
    // or in power-of-two parts of one brick per block (every row width divides or is a
    // multiple of the brick)
    if (s.is_col() && (R > s.N || (s.N / R) % (1 << s.brick) != 0)) return false;
    if (!s.is_col() && W % (1 << s.brick) != 0 && (1 << s.brick) % W != 0) return false;
  }
  const void* p = reg_lookup(s.dbl, s.kind, s.N, R, W + 100 * par + 1000 * occ);
  if (!p) return false;
  s.R = R;
  s.W = W;
  s.occ = occ;
  s.par = par;
  s.wg = wg;
  s.fn = p;
  return true;
}

// default variant: first valid (R, W) in preference order
static int pick_default(Slot& s) {
  const int* widths = s.is_col() ? kColWidths : kRowWidths;
  const int nw = s.is_col() ? kNumColWidths : kNumRowWidths;
  for (int R : kRadices) {
    for (int wi = 0; wi < nw; ++wi) {
      int W = widths[wi];
      if (!s.is_col() && s.lines < W && W > 1) continue;  // no idle rows
      if (choose_variant(s, R, W)) return PTF_OK;
    }
  }
  // fall back to any width
  for (int R : kRadices)
    for (int wi = 0; wi < nw; ++wi)
      if (choose_variant(s, R, widths[wi])) return PTF_OK;
  return fail(PTF_ERR_UNSUPPORTED, std::string("no kernel variant for ") + kind_name(s.kind) +
                                        " with line length " + std::to_string(s.N));
}

static_assert(sizeof(ColArgs<float>) == sizeof(ColArgs<double>) &&
                  offsetof(ColArgs<float>, wave_lines) == offsetof(ColArgs<double>, wave_lines),
              "column arguments share one layout");

static int launch(const Slot& s, hipStream_t stream) {
  if (s.is_col() && s.fixed_threads == 0 && s.args.size() == sizeof(ColArgs<double>)) {
    // the lane mapping of the selected variant (the kernel arguments are copied at launch)
    alignas(16) unsigned char buf[sizeof(ColArgs<double>)];
    std::memcpy(buf, s.args.data(), sizeof(buf));
    const int wl = s.wave_lines();
    std::memcpy(buf + offsetof(ColArgs<double>, wave_lines), &wl, sizeof(wl));
    void* argv[1] = {(void*)buf};
    HIPCHK(hipLaunchKernel(s.fn, s.grid(), dim3(s.threads()), argv, 0, stream));
    return PTF_OK;
  }
  void* argv[1] = {(void*)s.args.data()};
  HIPCHK(hipLaunchKernel(s.fn, s.grid(), dim3(s.threads()), argv, 0, stream));
  return PTF_OK;
}

template <typename A>
static void set_args(Slot& s, const A& a) {
  s.args.resize(sizeof(A));
  std::memcpy(s.args.data(), &a, sizeof(A));
}
template <typename A>
static A& args_of(Slot& s) {
  return *reinterpret_cast<A*>(s.args.data());
}

// ---------------------------------------------------------------------------
// twiddles: tw[m] = exp(-2 pi i m / L), m < L, from an octant-reduced long
// double sin/cos rounded once to the working precision
// ---------------------------------------------------------------------------
static void cs_exact(long n, long L, long double* c, long double* s) {
  n %= L;
  if (n < 0) n += L;
  if (2 * n > L) {  // (pi, 2 pi): conjugate
    cs_exact(L - n, L, c, s);
    *s = -*s;
    return;
  }
  if (4 * n > L) {  // (pi/2, pi]
    cs_exact(L / 2 - n, L, c, s);
    *c = -*c;
    return;
  }
  if (8 * n > L) {  // (pi/4, pi/2]
    long double cc, ss;
    cs_exact(L / 4 - n, L, &cc, &ss);
    *c = ss;
    *s = cc;
    return;
  }
  const long double pi = 3.141592653589793238462643383279502884L;
  const long double th = 2.0L * pi * (long double)n / (long double)L;
  *c = cosl(th);
  *s = sinl(th);
}

template <typename T>
static std::vector<cpx<T>> twiddle_table(int L) {
  std::vector<cpx<T>> tw(L);
  for (int m = 0; m < L; ++m) {
    long double c, s;
    if (L % 8 == 0) {
      cs_exact(m, L, &c, &s);
    } else {
      const long double pi = 3.141592653589793238462643383279502884L;
      c = cosl(2.0L * pi * m / L);
      s = sinl(2.0L * pi * m / L);
    }
    tw[m].re = (T)(double)c;
    tw[m].im = (T)(double)(-s);
  }
  return tw;
}

// ---------------------------------------------------------------------------
// device memory helpers
// ---------------------------------------------------------------------------
struct DevBuf {
  void* p = nullptr;
  size_t n = 0;
  ~DevBuf() {
    if (p) (void)hipFree(p);
  }
};

static int dalloc(DevBuf& b, size_t nbytes) {
  if (b.p) {
    (void)hipFree(b.p);
    b.p = nullptr;
  }
  if (nbytes == 0) nbytes = 16;
  hipError_t e = hipMalloc(&b.p, nbytes);
  if (e != hipSuccess) {
    b.p = nullptr;
    (void)hipGetLastError();
    return fail(PTF_ERR_NOMEM, "hipMalloc of " + std::to_string(nbytes) + " bytes failed: " +
                                   hipGetErrorString(e));
  }
  b.n = nbytes;
  return PTF_OK;
}


// ---------------------------------------------------------------------------
// group of B problems on one grid
// ---------------------------------------------------------------------------
struct Group {
  int ndim = 1, nx = 1, ny = 1, nz = 1, nxh = 1;
  int B = 1;
  int lb = 0;  // the last axis is stored in bricks of 2^lb planes (2D: y, 3D: z)
  bool dbl = true;
  size_t esz = 8;
  int stepper = PTF_RK4;          // PTF_RK4 or PTF_ETDRK4 (the unfiltered form)
  bool filtered = false;          // FilteredRK4 / FilteredETDRK4
  int f_order = 4;
  double f_innerK = 0.65, f_decay = 0.0;
  int flow_kind = PTF_FLOW_STEADY;
  long long nS = 0, nR = 0;  // elements per member
  double dt = 0.01;
  int nkappah = 0;
  double Ls[3] = {0, 0, 0}, ks[3] = {0, 0, 0};
  std::vector<OpParams> ops;

  std::vector<double> L_host;  // L in the public layout, all members
  DevBuf d_ops, sol, sol_lo, stg, stg2, acc, n1, n2, w[3], flow[3], d_L, etd[6], tw[3], psih0, omega;
  DevBuf stage;   // FP64 staging of one member in the public layout (host transfers)
  DevBuf phys, spec, red, backup;  // set_c/updatevars/calcN staging, reductions, timing backup
  DevBuf fs[2];                    // spectral-flow synthesis scratch
  std::vector<Slot> prologue;      // first inverse pass of a step when w0/w1 are not primed
  std::vector<Slot> stages[4];
  std::vector<Slot> flowsyn;
  std::vector<ptf_flow_term> aterms;   // PTF_FLOW_ANALYTIC
  bool plan_only = false;              // a transform plan: no state, scratch for B members
  DevBuf atab;                         // x, y, z factor tables of the terms
  // w0/w1 hold the first inverse pass of the current state (the fused last
  // kernel of each step leaves them so); anything else that writes sol or the
  // transform scratch clears this
  bool primed = false;
  // plan slots (B = 1 view, pointers set per call)
  Slot p_r2c, p_c2r, p_yf, p_yi, p_zf, p_zi;
  hipStream_t stream = nullptr;
};

template <typename T>
static GridInfo grid_info(const Group& g) {
  GridInfo gi;
  gi.ndim = g.ndim;
  gi.nx = g.nx;
  gi.ny = g.ny;
  gi.nz = g.nz;
  gi.ksx = g.ks[0];
  gi.ksy = g.ks[1];
  gi.ksz = g.ks[2];
  return gi;
}

// wavenumber scale as the grid's rfftfreq(n, 2 pi/dx) computes it: (2 pi/dx)/n
static double kscale(int n, double L) {
  const double dx = L / n;
  const double fs = 2.0 * M_PI / dx;
  return fs / n;
}

// whether any kernel variant of `kind` was compiled for line length N (the registry
// answers on the host, without a device)
static bool len_compiled(bool dbl, int kind, int N) {
  for (int R : kRadices)
    for (int W : {1, 2, 4, 8, 16, 32, 64})
      if (reg_lookup(dbl, kind, N, R, W)) return true;
  return false;
}

static int validate_desc(const ptf_desc* d) {
  if (!d) return fail(PTF_ERR_INVALID, "desc is NULL");
  if (d->ndim < 1 || d->ndim > 3) return fail(PTF_ERR_INVALID, "ndim must be 1, 2 or 3");
  if (d->precision != PTF_FP64 && d->precision != PTF_MIXED)
    return fail(PTF_ERR_INVALID, "precision must be PTF_FP64 or PTF_MIXED");
  if (d->stepper < PTF_RK4 || d->stepper > PTF_FILTERED_ETDRK4)
    return fail(PTF_ERR_INVALID, "stepper must be PTF_RK4, PTF_ETDRK4 or a filtered form");
  if ((d->stepper == PTF_FILTERED_RK4 || d->stepper == PTF_FILTERED_ETDRK4) && d->filter_order) {
    if (d->filter_order < 1) return fail(PTF_ERR_INVALID, "filter_order must be >= 1");
    if (!(d->filter_outerK > d->filter_innerK))
      return fail(PTF_ERR_INVALID, "filter_outerK must exceed filter_innerK");
    if (!(d->filter_tol > 0 && d->filter_tol < 1))
      return fail(PTF_ERR_INVALID, "filter_tol must be in (0, 1)");
  }
  if (d->flow_kind < PTF_FLOW_STEADY || d->flow_kind > PTF_FLOW_ANALYTIC)
    return fail(PTF_ERR_INVALID, "unknown flow_kind");
  if (d->flow_kind == PTF_FLOW_SPECTRAL && d->ndim != 2)
    return fail(PTF_ERR_UNSUPPORTED, "the built-in spectral flow is 2D only");
  if (d->nkappah < 0) return fail(PTF_ERR_INVALID, "nkappah must be >= 0");
  if (!(d->dt > 0) || !std::isfinite(d->dt)) return fail(PTF_ERR_INVALID, "dt must be positive");
  for (int a = 0; a < d->ndim; ++a) {
    const long long n = d->n[a];
    if (n <= 0 || (n % 2)) return fail(PTF_ERR_INVALID, "grid sizes must be positive and even");
    if (!(d->L[a] > 0)) return fail(PTF_ERR_INVALID, "domain lengths must be positive");
    const bool dbl = d->precision == PTF_FP64;
    if (a == 0 && (n < 16 || n > 4096))
      return fail(PTF_ERR_UNSUPPORTED, "nx must be in [16, 4096], got " + std::to_string(n));
    if (a > 0 && (n < 8 || n > 2048))
      return fail(PTF_ERR_UNSUPPORTED, "ny and nz must be in [8, 2048], got " + std::to_string(n));
    if (!len_compiled(dbl, a == 0 ? K_ROW_C2R : K_COL_FWD, (int)n))
      return fail(PTF_ERR_UNSUPPORTED,
                  "grid size " + std::to_string(n) + " has no compiled transform (supported: "
                  "powers of two, and 2^a m for m in {3, 5, 9, 25, 27, 125}; "
                  "for nx the half length nx/2 must be one of these)");
  }
  return PTF_OK;
}

// ----- slot builders ---------------------------------------------------------
template <typename T>
static ColArgs<T> col_base(const Group& g, int axis) {
  ColArgs<T> a;
  std::memset(&a, 0, sizeof(a));
  a.tw = (const cpx<T>*)g.tw[axis].p;
  a.twstride = 1;
  a.mstride = g.nS;
  a.kscale = g.ks[axis];
  a.axis = axis;
  const long long B = 1LL << g.lb;
  if (g.ndim == 3 && axis == 1) {
    // y pass of a 3D field: columns are (kx, kz_lo) pairs, blockIdx.y = kz_hi
    a.ncols = (int)(g.nxh * B);
    a.stride = g.nxh * B;
    a.ostride = (long long)g.nxh * g.ny * B;
    a.lbc = g.lb;
  } else if (g.ndim == 3) {
    // z pass: the line runs along the bricked axis, blockIdx.y = ky
    a.ncols = g.nxh;
    a.stride = (long long)g.nxh * g.ny * B;
    a.ostride = g.nxh * B;
    a.lb = g.lb;
  } else {
    // y pass of a 2D field: the line runs along the bricked axis
    a.ncols = g.nxh;
    a.stride = g.nxh * B;
    a.ostride = 0;
    a.lb = g.lb;
  }
  a.kscale_x = g.ks[0];
  // blockIdx.y extent of this pass (see make_slot) and the member count
  a.nother = g.ndim == 3 ? (axis == 1 ? (g.nz >> g.lb) : g.ny) : 1;
  a.nmembers = g.B;
  a.wave_lines = 0;   // set per launch from the slot's variant
  return a;
}

template <typename T>
static RowArgs<T> row_base(const Group& g) {
  RowArgs<T> a;
  std::memset(&a, 0, sizeof(a));
  a.tw = (const cpx<T>*)g.tw[0].p;
  a.twstride = 1;
  a.nrows = (long long)g.ny * g.nz;
  a.smstride = g.nS;
  a.pmstride = g.nR;
  a.scale = 1.0 / ((double)g.nx * g.ny * g.nz);
  a.kscale_x = g.ks[0];
  a.ny = g.ny;
  a.lb = g.lb;
  return a;
}

static Slot make_slot(const Group& g, int kind, int axis, int members) {
  Slot s;
  s.dbl = g.dbl;
  s.kind = kind;
  s.members = members;
  if (kind < 10) {
    if (g.ndim == 3 && axis == 1) {
      s.N = g.ny;
      s.lines = (long long)g.nxh << g.lb;
      s.other = g.nz >> g.lb;
    } else if (g.ndim == 3) {
      s.N = g.nz;
      s.lines = g.nxh;
      s.other = g.ny;
      s.brick = g.lb;
    } else {
      s.N = g.ny;
      s.lines = g.nxh;
      s.other = 1;
      s.brick = g.lb;
    }
  } else {
    s.N = g.nx;
    s.lines = (long long)g.ny * g.nz;
    s.brick = g.lb;
  }
  return s;
}

template <typename T>
static StageArgs<T> stage_args(const Group& g, int stage) {
  StageArgs<T> st;
  std::memset(&st, 0, sizeof(st));
  st.kind = g.stepper == PTF_RK4 ? EPI_RK4 : EPI_ETDRK4;
  st.stage = stage;
  st.sol = (cpx<T>*)g.sol.p;
  st.sol_lo = (cpx<T>*)g.sol_lo.p;  // NULL in the FP64 form
  st.stg = (cpx<T>*)g.stg.p;
  st.stg2 = (cpx<T>*)g.stg2.p;
  st.acc = (cpx<T>*)g.acc.p;
  st.n1 = (cpx<T>*)g.n1.p;
  st.n2 = (cpx<T>*)g.n2.p;
  st.E = (const double*)g.etd[0].p;
  st.E2 = (const double*)g.etd[1].p;
  st.zeta = (const double*)g.etd[2].p;
  st.alpha = (const double*)g.etd[3].p;
  st.beta = (const double*)g.etd[4].p;
  st.gam = (const double*)g.etd[5].p;
  st.dt = g.dt;
  st.op = (const OpParams*)g.d_ops.p;
  st.nkappah = g.nkappah;
  st.g = grid_info<T>(g);
  st.mstride = g.nS;
  st.filtered = g.filtered && stage == 3 ? 1 : 0;
  st.f_order = g.f_order;
  st.f_innerK = g.f_innerK;
  st.f_decay = g.f_decay;
  return st;
}

// whether a stage epilogue writes sol (RK4 stages 3 and 4 in the FP64 form; the last
// stage otherwise)
template <typename T>
static bool writes_sol(const StageArgs<T>& st) {
  if (st.kind == EPI_STORE) return false;
  return st.stage == 3 || (st.kind == EPI_RK4 && st.stage == 2);
}

// The kernel kind of a stage's last column pass: with or without the fused first inverse
// pass of the next stage, compiled for the epilogue family of st.kind.
static int stage_col_kind(bool fused, int epi) {
  if (epi == EPI_RK4) return fused ? K_COL_STAGE_D1_RK4 : K_COL_STAGE_RK4;
  return fused ? K_COL_STAGE_D1 : K_COL_STAGE;
}

// Launch list of one calcN! evaluation of `in`, ending in `st` (a stage
// epilogue or a plain store to `out`).
// With fused = true the first inverse pass is left out (the previous stage's
// last kernel produced w0/w1) and the last kernel also performs the first
// inverse pass of the next stage's input (K_COL_STAGE_D1 or K_COL_STAGE_D1_RK4).
template <typename T>
static int build_calcN(Group& g, const cpx<T>* in, const StageArgs<T>& st, cpx<T>* out,
                       std::vector<Slot>& list, const std::string& prefix, bool fused = false) {
  cpx<T>* w0 = (cpx<T>*)g.w[0].p;
  cpx<T>* w1 = (cpx<T>*)g.w[1].p;
  cpx<T>* w2 = (cpx<T>*)g.w[2].p;
  const T* u = (const T*)g.flow[0].p;
  const T* v = (const T*)g.flow[1].p;
  const T* w = (const T*)g.flow[2].p;
  if (g.ndim == 1) {
    Slot s = make_slot(g, K_ROW_ADV1, 0, g.B);
    RowArgs<T> a = row_base<T>(g);
    a.in[0] = in;
    a.flow[0] = u;
    a.out = out;
    a.st = st;
    set_args(s, a);
    s.name = prefix + "x_advect_stage";
    s.writes_state = writes_sol(st);
    CHK(pick_default(s));
    list.push_back(s);
    return PTF_OK;
  }
  if (g.ndim == 2) {
    if (!fused) {
      Slot s1 = make_slot(g, K_COL_D1, 1, g.B);
      ColArgs<T> a1 = col_base<T>(g, 1);
      a1.in0 = in;
      a1.out0 = w0;
      a1.out1 = w1;
      set_args(s1, a1);
      s1.name = prefix + "y_inverse_deriv";
      CHK(pick_default(s1));
      list.push_back(s1);
    }

    Slot s2 = make_slot(g, K_ROW_ADV2, 0, g.B);
    RowArgs<T> a2 = row_base<T>(g);
    a2.in[0] = w0;
    a2.in[1] = w1;
    a2.flow[0] = u;
    a2.flow[1] = v;
    a2.out = w0;
    a2.st.kind = EPI_STORE;
    set_args(s2, a2);
    s2.name = prefix + "x_advect";
    CHK(pick_default(s2));
    list.push_back(s2);

    Slot s3 = make_slot(g, stage_col_kind(fused, st.kind), 1, g.B);
    ColArgs<T> a3 = col_base<T>(g, 1);
    a3.in0 = w0;
    a3.out0 = fused ? w0 : out;
    a3.out1 = fused ? w1 : nullptr;
    a3.st = st;
    set_args(s3, a3);
    s3.name = prefix + (fused ? "y_forward_stage_inverse" : "y_forward_stage");
    s3.writes_state = writes_sol(st);
    CHK(pick_default(s3));
    list.push_back(s3);
    return PTF_OK;
  }
  // 3D
  if (!fused) {
    Slot s1 = make_slot(g, K_COL_D1, 2, g.B);
    ColArgs<T> a1 = col_base<T>(g, 2);
    a1.in0 = in;
    a1.out0 = w0;
    a1.out1 = w1;
    set_args(s1, a1);
    s1.name = prefix + "z_inverse_deriv";
    CHK(pick_default(s1));
    list.push_back(s1);
  }

  Slot s2 = make_slot(g, K_COL_D2, 1, g.B);
  ColArgs<T> a2 = col_base<T>(g, 1);
  a2.in0 = w0;
  a2.in1 = w1;
  a2.out0 = w0;
  a2.out1 = w2;
  a2.out2 = w1;
  set_args(s2, a2);
  s2.name = prefix + "y_inverse_deriv2";
  CHK(pick_default(s2));
  list.push_back(s2);

  Slot s3 = make_slot(g, K_ROW_ADV3, 0, g.B);
  RowArgs<T> a3 = row_base<T>(g);
  a3.in[0] = w0;   // IFFT_yz(c)        -> multiplied by i kr in the kernel
  a3.in[1] = w2;   // IFFT_y(i l IFFT_z(c))
  a3.in[2] = w1;   // IFFT_y(IFFT_z(i m c))
  a3.flow[0] = u;
  a3.flow[1] = v;
  a3.flow[2] = w;
  a3.out = w0;
  a3.st.kind = EPI_STORE;
  set_args(s3, a3);
  s3.name = prefix + "x_advect";
  CHK(pick_default(s3));
  list.push_back(s3);

  Slot s4 = make_slot(g, K_COL_FWD, 1, g.B);
  ColArgs<T> a4 = col_base<T>(g, 1);
  a4.in0 = w0;
  a4.out0 = w0;
  set_args(s4, a4);
  s4.name = prefix + "y_forward";
  CHK(pick_default(s4));
  list.push_back(s4);

  Slot s5 = make_slot(g, stage_col_kind(fused, st.kind), 2, g.B);
  ColArgs<T> a5 = col_base<T>(g, 2);
  a5.in0 = w0;
  a5.out0 = fused ? w0 : out;
  a5.out1 = fused ? w1 : nullptr;
  a5.st = st;
  set_args(s5, a5);
  s5.name = prefix + (fused ? "z_forward_stage_inverse" : "z_forward_stage");
  s5.writes_state = writes_sol(st);
  CHK(pick_default(s5));
  list.push_back(s5);
  return PTF_OK;
}

// Ensemble groups of small square 2D members take the whole step in one launch, a
// workgroup per member (member_step_kernel) where one is compiled for the size, up to
// members of kMemberMax^2.  PTF_MEMBER_KERNEL=<n> moves that limit (0: the pass kernels
// for every group), to measure the two.
static constexpr int kMemberMax = 128;
template <typename T>
static bool build_member_slot(Group& g) {
  if (g.ndim != 2 || g.nx != g.ny || g.B < 2 || g.plan_only) return false;
  const char* e = std::getenv("PTF_MEMBER_KERNEL");
  const int limit = e ? std::atoi(e) : kMemberMax;
  if (g.nx > limit) return false;
  int threads = 0;
  const void* fn = member_kernel(g.dbl, g.nx, &threads);
  if (!fn) return false;
  Slot s;
  s.name = "step_member";
  s.dbl = g.dbl;
  s.kind = K_MEMBER;
  s.N = g.nx;
  s.lines = g.ny;
  s.members = g.B;
  s.fn = fn;
  s.fixed_threads = threads;
  s.fixed_blocks = (unsigned)g.B;
  s.writes_state = true;
  MemberArgs<T> a;
  std::memset(&a, 0, sizeof(a));
  a.st = stage_args<T>(g, 0);   // the kernel runs all four stages
  a.st.store_x = 0;             // a stage input stays on chip for the next stage
  a.st.filtered = g.filtered ? 1 : 0;   // applied by the last stage's update only
  a.flow[0] = (const T*)g.flow[0].p;
  a.flow[1] = (const T*)g.flow[1].p;
  a.tw = (const cpx<T>*)g.tw[0].p;   // exp(-2 pi i m / n): the x and y tables (nx = ny)
  a.pmstride = g.nR;
  a.kscale_x = g.ks[0];
  a.kscale_y = g.ks[1];
  a.scale = 1.0 / ((double)g.nx * g.ny);
  a.lb = g.lb;
  set_args(s, a);
  for (int k = 0; k < 4; ++k) g.stages[k].clear();
  g.stages[0].push_back(s);
  g.prologue.clear();
  return true;
}

template <typename T>
static int build_group_slots(Group& g) {
  // Stage-boundary fusion (2D/3D): the last kernel of a stage also performs the first inverse
  // pass of the next stage's input.  PTF_STAGE_FUSION=0 builds the unfused schedule (each
  // stage starts with its own inverse pass, which reads the stored stage input), to measure
  // the two against each other.
  const char* fz = std::getenv("PTF_STAGE_FUSION");
  const bool fuse = g.ndim >= 2 && !(fz && fz[0] == '0');
  const bool member = build_member_slot<T>(g);
  for (int s = 0; s < 4 && !member; ++s) {
    g.stages[s].clear();
    const cpx<T>* in;
    if (g.stepper == PTF_RK4) {
      in = (const cpx<T>*)(s == 0 ? g.sol.p : g.stg.p);
    } else {
      in = (const cpx<T>*)(s == 0 ? g.sol.p : s == 1 ? g.stg.p : g.stg2.p);
    }
    StageArgs<T> st = stage_args<T>(g, s);
    // the fused forms hand the next stage input on in registers
    st.store_x = fuse ? 0 : 1;
    CHK(build_calcN<T>(g, in, st, nullptr, g.stages[s], "stage" + std::to_string(s + 1) + "_",
                       fuse));
  }
  if (!member) g.prologue.clear();
  if (fuse && !member) {
    const int axis = g.ndim == 3 ? 2 : 1;
    Slot s1 = make_slot(g, K_COL_D1, axis, g.B);
    ColArgs<T> a1 = col_base<T>(g, axis);
    a1.in0 = (const cpx<T>*)g.sol.p;
    a1.out0 = (cpx<T>*)g.w[0].p;
    a1.out1 = (cpx<T>*)g.w[1].p;
    set_args(s1, a1);
    s1.name = axis == 2 ? "prologue_z_inverse_deriv" : "prologue_y_inverse_deriv";
    CHK(pick_default(s1));
    g.prologue.push_back(s1);
  }
  g.primed = false;
  g.flowsyn.clear();
  if (g.flow_kind == PTF_FLOW_SPECTRAL) {
    Slot s1 = make_slot(g, K_COL_FLOWSYN, 1, 1);
    ColArgs<T> a1 = col_base<T>(g, 1);
    a1.psih0 = (const cpx<double>*)g.psih0.p;
    a1.omega = (const double*)g.omega.p;
    a1.out0 = (cpx<T>*)g.fs[0].p;
    a1.out1 = (cpx<T>*)g.fs[1].p;
    set_args(s1, a1);
    s1.name = "flow_y_synthesis";
    CHK(pick_default(s1));
    g.flowsyn.push_back(s1);
    for (int c = 0; c < 2; ++c) {
      Slot s2 = make_slot(g, K_ROW_C2R, 0, 1);
      RowArgs<T> a2 = row_base<T>(g);
      a2.in[0] = (const cpx<T>*)g.fs[c].p;
      a2.rout = (T*)g.flow[c].p;
      set_args(s2, a2);
      s2.name = c == 0 ? "flow_x_c2r_u" : "flow_x_c2r_v";
      CHK(pick_default(s2));
      g.flowsyn.push_back(s2);
    }
  }
  // plan slots: all members of a batched plan; one member of a problem or ensemble
  // group (pointers patched per call)
  const int pm = g.plan_only ? g.B : 1;
  g.p_r2c = make_slot(g, K_ROW_R2C, 0, pm);
  set_args(g.p_r2c, row_base<T>(g));
  g.p_r2c.name = "plan_x_r2c";
  CHK(pick_default(g.p_r2c));
  g.p_c2r = make_slot(g, K_ROW_C2R, 0, pm);
  set_args(g.p_c2r, row_base<T>(g));
  g.p_c2r.name = "plan_x_c2r";
  CHK(pick_default(g.p_c2r));
  if (g.ndim >= 2) {
    g.p_yf = make_slot(g, K_COL_FWD, 1, pm);
    set_args(g.p_yf, col_base<T>(g, 1));
    g.p_yf.name = "plan_y_fwd";
    CHK(pick_default(g.p_yf));
    g.p_yi = make_slot(g, K_COL_INV, 1, pm);
    set_args(g.p_yi, col_base<T>(g, 1));
    g.p_yi.name = "plan_y_inv";
    CHK(pick_default(g.p_yi));
  }
  if (g.ndim == 3) {
    g.p_zf = make_slot(g, K_COL_FWD, 2, pm);
    set_args(g.p_zf, col_base<T>(g, 2));
    g.p_zf.name = "plan_z_fwd";
    CHK(pick_default(g.p_zf));
    g.p_zi = make_slot(g, K_COL_INV, 2, pm);
    set_args(g.p_zi, col_base<T>(g, 2));
    g.p_zi.name = "plan_z_inv";
    CHK(pick_default(g.p_zi));
  }
  return PTF_OK;
}

template <typename T>
static int upload_twiddles(Group& g) {
  const int lens[3] = {g.nx, g.ny, g.nz};
  for (int a = 0; a < g.ndim; ++a) {
    std::vector<cpx<T>> tw = twiddle_table<T>(lens[a]);
    CHK(dalloc(g.tw[a], tw.size() * sizeof(cpx<T>)));
    HIPCHK(hipMemcpy(g.tw[a].p, tw.data(), tw.size() * sizeof(cpx<T>), hipMemcpyHostToDevice));
  }
  return PTF_OK;
}

// ---------------------------------------------------------------------------
// public layout <-> brick layout (host side)
// A member's rows are (l, m) in the public order r = l*mid + m (l = last-axis
// index, m = middle index; mid = ny in 3D, 1 in 2D).  In the brick layout row
// (l, m) becomes b = ((l >> lb) mid + m) 2^lb + (l mod 2^lb) and element k of
// row b sits at ((b >> lb) rowlen + k) 2^lb + (b mod 2^lb).
// ---------------------------------------------------------------------------
static size_t brick_index(const Group& g, size_t r, size_t k, size_t rowlen) {
  if (!g.lb) return r * rowlen + k;
  const size_t mid = g.ndim == 3 ? (size_t)g.ny : 1;
  const size_t l = r / mid, m = r % mid;
  const size_t B = (size_t)1 << g.lb, mask = B - 1;
  const size_t b = ((l >> g.lb) * mid + m) * B + (l & mask);
  return ((b >> g.lb) * rowlen + k) * B + (b & mask);
}

static size_t member_rows(const Group& g) { return (size_t)g.ny * g.nz; }

// w = doubles per element (2: complex values or physical x-pairs; 1: reals)
static void to_brick(const Group& g, const double* src, double* dst, size_t rowlen, int w) {
  const size_t rows = member_rows(g);
  for (size_t r = 0; r < rows; ++r)
    for (size_t k = 0; k < rowlen; ++k) {
      const size_t d = brick_index(g, r, k, rowlen);
      for (int e = 0; e < w; ++e) dst[d * w + e] = src[(r * rowlen + k) * w + e];
    }
}

static int group_init(Group& g, const ptf_desc* descs, int count, bool allow_brick = true,
                      bool plan_only = false);

static int group_init_once(Group& g, const ptf_desc* descs, int count, int lb, bool plan_only);

static int runtime_check();

static int group_init(Group& g, const ptf_desc* descs, int count, bool allow_brick,
                      bool plan_only) {
  CHK(runtime_check());
  const ptf_desc& d = descs[0];
  int lb = 0;
  if (allow_brick && d.ndim >= 2) {
    const long long nlast = d.ndim == 3 ? d.n[2] : d.n[1];
    if (nlast >= 64 && nlast % 8 == 0) lb = 3;   // whole bricks of 8 planes
  }
  int rc = group_init_once(g, descs, count, lb, plan_only);
  if (rc == PTF_ERR_UNSUPPORTED && lb > 0) rc = group_init_once(g, descs, count, 0, plan_only);
  return rc;
}

static int group_init_once(Group& g, const ptf_desc* descs, int count, int lb, bool plan_only) {
  g.lb = lb;
  g.plan_only = plan_only;
  const ptf_desc& d = descs[0];
  g.ndim = d.ndim;
  g.nx = (int)d.n[0];
  g.ny = d.ndim >= 2 ? (int)d.n[1] : 1;
  g.nz = d.ndim >= 3 ? (int)d.n[2] : 1;
  g.nxh = g.nx / 2 + 1;
  g.B = count;
  g.dbl = d.precision == PTF_FP64;
  g.esz = g.dbl ? 8 : 4;
  g.filtered = d.stepper == PTF_FILTERED_RK4 || d.stepper == PTF_FILTERED_ETDRK4;
  g.stepper = (d.stepper == PTF_ETDRK4 || d.stepper == PTF_FILTERED_ETDRK4) ? PTF_ETDRK4 : PTF_RK4;
  if (g.filtered) {
    const bool dflt = d.filter_order == 0;
    const int order = dflt ? 4 : d.filter_order;
    const double innerK = dflt ? 0.65 : d.filter_innerK;
    const double outerK = dflt ? 1.0 : d.filter_outerK;
    const double tol = dflt ? 1e-15 : d.filter_tol;
    double span = 1.0;
    for (int i = 0; i < order; ++i) span *= (outerK - innerK);
    g.f_order = order;
    g.f_innerK = innerK;
    g.f_decay = -std::log(tol) / span;
  }
  g.flow_kind = d.flow_kind;
  g.nS = (long long)g.nxh * g.ny * g.nz;
  g.nR = (long long)g.nx * g.ny * g.nz;
  g.dt = d.dt;
  g.nkappah = d.nkappah;
  for (int a = 0; a < 3; ++a) {
    g.Ls[a] = a < d.ndim ? d.L[a] : 1.0;
    g.ks[a] = a < d.ndim ? kscale((int)d.n[a], d.L[a]) : 0.0;
  }
  g.ops.resize(count);
  for (int i = 0; i < count; ++i) {
    g.ops[i].kappa = descs[i].kappa;
    g.ops[i].eta = descs[i].eta;
    g.ops[i].iota = descs[i].iota;
    g.ops[i].kappah = descs[i].kappah;
  }
  const size_t cs = 2 * g.esz;  // complex element bytes
  const size_t S = (size_t)g.nS * g.B * cs;
  const size_t Rb = (size_t)g.nR * g.B * g.esz;
  CHK(dalloc(g.d_ops, sizeof(OpParams) * count));
  HIPCHK(hipMemcpy(g.d_ops.p, g.ops.data(), sizeof(OpParams) * count, hipMemcpyHostToDevice));
  if (plan_only) {
    // a transform plan keeps only the inverse transform's scratch (all members)
    CHK(dalloc(g.spec, S));
    if (g.dbl) {
      CHK(upload_twiddles<double>(g));
      CHK(build_group_slots<double>(g));
    } else {
      CHK(upload_twiddles<float>(g));
      CHK(build_group_slots<float>(g));
    }
    return PTF_OK;
  }
  CHK(dalloc(g.sol, S));
  if (!g.dbl) {
    CHK(dalloc(g.sol_lo, S));
    HIPCHK(hipMemset(g.sol_lo.p, 0, S));
  }
  CHK(dalloc(g.stg, S));
  CHK(dalloc(g.acc, S));
  if (g.stepper == PTF_ETDRK4) {
    CHK(dalloc(g.stg2, S));
    CHK(dalloc(g.n1, S));
    CHK(dalloc(g.n2, S));
  }
  for (int i = 0; i < (g.ndim == 3 ? 3 : 2); ++i) CHK(dalloc(g.w[i], S));
  for (int i = 0; i < g.ndim; ++i) CHK(dalloc(g.flow[i], Rb));
  HIPCHK(hipMemset(g.sol.p, 0, S));
  for (int i = 0; i < g.ndim; ++i) HIPCHK(hipMemset(g.flow[i].p, 0, Rb));
  CHK(dalloc(g.phys, (size_t)g.nR * g.esz));
  CHK(dalloc(g.spec, (size_t)g.nS * cs));
  CHK(dalloc(g.red, 2 * sizeof(double) * 1024));
  if (g.flow_kind == PTF_FLOW_SPECTRAL) {
    CHK(dalloc(g.fs[0], (size_t)g.nS * cs));
    CHK(dalloc(g.fs[1], (size_t)g.nS * cs));
    CHK(dalloc(g.psih0, (size_t)g.nS * 2 * sizeof(double)));
    CHK(dalloc(g.omega, (size_t)g.nS * sizeof(double)));
    HIPCHK(hipMemset(g.psih0.p, 0, (size_t)g.nS * 2 * sizeof(double)));
    HIPCHK(hipMemset(g.omega.p, 0, (size_t)g.nS * sizeof(double)));
  }
  // linear operator (FP64) for get_L and the ETDRK4 coefficients
  {
    std::vector<double> L((size_t)g.nS * g.B);
    for (int b = 0; b < g.B; ++b) {
      const OpParams& p = g.ops[b];
      for (int kz = 0; kz < g.nz; ++kz)
        for (int ky = 0; ky < g.ny; ++ky)
          for (int kx = 0; kx < g.nxh; ++kx) {
            const double kr = kx * g.ks[0];
            const double kr2 = kr * kr;
            double v;
            if (g.ndim == 1) {
              v = -p.kappa * kr2 - p.kappah * std::pow(kr2, g.nkappah);
            } else {
              const int iy = ky < g.ny / 2 ? ky : ky - g.ny;
              const double l = iy * g.ks[1];
              const double l2 = l * l;
              if (g.ndim == 2) {
                v = -p.kappa * kr2 - p.eta * l2 - p.kappah * std::pow(kr2 + l2, g.nkappah);
              } else {
                const int iz = kz < g.nz / 2 ? kz : kz - g.nz;
                const double m = iz * g.ks[2];
                const double m2 = m * m;
                v = -p.kappa * kr2 - p.eta * l2 - p.iota * m2 -
                    p.kappah * std::pow(kr2 + l2 + m2, g.nkappah);
              }
            }
            L[(size_t)b * g.nS + kx + (size_t)g.nxh * (ky + (size_t)g.ny * kz)] = v;
          }
    }
    g.L_host = L;
    if (g.lb) {
      std::vector<double> Lb(L.size());
      for (int b = 0; b < g.B; ++b)
        to_brick(g, L.data() + (size_t)b * g.nS, Lb.data() + (size_t)b * g.nS, (size_t)g.nxh, 1);
      L.swap(Lb);
    }
    CHK(dalloc(g.d_L, L.size() * sizeof(double)));
    HIPCHK(hipMemcpy(g.d_L.p, L.data(), L.size() * sizeof(double), hipMemcpyHostToDevice));
  }
  if (g.stepper == PTF_ETDRK4) {
    const long long n = g.nS * g.B;
    for (int i = 0; i < 6; ++i) CHK(dalloc(g.etd[i], (size_t)n * sizeof(double)));
    const int th = 256;
    hipLaunchKernelGGL(etd_coeff_kernel, dim3((unsigned)((n + th - 1) / th)), dim3(th), 0, 0,
                       (const double*)g.d_L.p, n, g.dt, 32, 1.0, (double*)g.etd[0].p,
                       (double*)g.etd[1].p, (double*)g.etd[2].p, (double*)g.etd[3].p,
                       (double*)g.etd[4].p, (double*)g.etd[5].p);
    HIPCHK(hipGetLastError());
    HIPCHK(hipDeviceSynchronize());
  }
  if (g.dbl) {
    CHK(upload_twiddles<double>(g));
    CHK(build_group_slots<double>(g));
  } else {
    CHK(upload_twiddles<float>(g));
    CHK(build_group_slots<float>(g));
  }
  return PTF_OK;
}

// ----- plan-style transforms on one member -------------------------------------
template <typename T>
static int group_rfft(Group& g, const void* in_real, void* out_complex, hipStream_t st) {
  {
    RowArgs<T>& a = args_of<RowArgs<T>>(g.p_r2c);
    a.rin = (const T*)in_real;
    a.out = (cpx<T>*)out_complex;
    CHK(launch(g.p_r2c, st));
  }
  if (g.ndim >= 2) {
    ColArgs<T>& a = args_of<ColArgs<T>>(g.p_yf);
    a.in0 = (const cpx<T>*)out_complex;
    a.out0 = (cpx<T>*)out_complex;
    CHK(launch(g.p_yf, st));
  }
  if (g.ndim == 3) {
    ColArgs<T>& a = args_of<ColArgs<T>>(g.p_zf);
    a.in0 = (const cpx<T>*)out_complex;
    a.out0 = (cpx<T>*)out_complex;
    CHK(launch(g.p_zf, st));
  }
  return PTF_OK;
}

// The inverse transforms along z and y write the intermediate spectrum into `spec` (the
// plan's own scratch, or the static scratch of ptf_irfft_noalloc); the input is preserved.
template <typename T>
static int group_irfft(Group& g, const void* in_complex, void* out_real, hipStream_t st,
                       void* spec = nullptr) {
  cpx<T>* scr = (cpx<T>*)(spec ? spec : g.spec.p);
  const cpx<T>* src = (const cpx<T>*)in_complex;
  if (g.ndim == 3) {
    ColArgs<T>& a = args_of<ColArgs<T>>(g.p_zi);
    a.in0 = src;
    a.out0 = scr;
    CHK(launch(g.p_zi, st));
    src = scr;
  }
  if (g.ndim >= 2) {
    ColArgs<T>& a = args_of<ColArgs<T>>(g.p_yi);
    a.in0 = src;
    a.out0 = scr;
    CHK(launch(g.p_yi, st));
    src = scr;
  }
  RowArgs<T>& a = args_of<RowArgs<T>>(g.p_c2r);
  a.in[0] = src;
  a.rout = (T*)out_real;
  CHK(launch(g.p_c2r, st));
  return PTF_OK;
}

static int g_rfft(Group& g, const void* in, void* out, hipStream_t st) {
  return g.dbl ? group_rfft<double>(g, in, out, st) : group_rfft<float>(g, in, out, st);
}
static int g_irfft(Group& g, const void* in, void* out, hipStream_t st, void* spec = nullptr) {
  return g.dbl ? group_irfft<double>(g, in, out, st, spec) : group_irfft<float>(g, in, out, st, spec);
}

// ---------------------------------------------------------------------------
// static scratch of the allocation-free inverse transform (ptf_irfft_noalloc)
// ---------------------------------------------------------------------------
// Callers that cannot allocate device memory (code under HIP graph capture, serving
// loops, harnesses that forbid hipMalloc) get the intermediate spectrum of a 2D/3D inverse
// transform in this fixed array instead of the plan's scratch buffer.  Every device the
// library runs on holds its own copy until the process ends.  In host code the name
// `scratch` is not a device address: the address is taken once per device with
// hipGetSymbolAddress (with that device current) and only that address is used.
static constexpr size_t kStaticScratchBytes = size_t(64) << 20;   // 64 MiB
static __device__ __align__(256) unsigned char scratch[kStaticScratchBytes];
static constexpr int kMaxDevices = 64;
static void* g_scratch_addr[kMaxDevices];
static std::mutex g_scratch_mu;

// device address of `scratch` on device `dev` (fetched on first use with `dev` current)
static int static_scratch(int dev, void** out) {
  if (dev < 0 || dev >= kMaxDevices)
    return fail(PTF_ERR_INVALID, "device " + std::to_string(dev) + " out of range for the static scratch");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_addr[dev]) {
    int cur = 0;
    HIPCHK(hipGetDevice(&cur));
    if (cur != dev) HIPCHK(hipSetDevice(dev));
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (cur != dev) HIPCHK(hipSetDevice(cur));
    if (e != hipSuccess || !p)
      return fail(PTF_ERR_DEVICE, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch_addr[dev] = p;
  }
  *out = g_scratch_addr[dev];
  return PTF_OK;
}

// time factor T(omega t + phase) of an analytic flow term
static double term_time_factor(const ptf_flow_term& k, double t) {
  const double a = k.omega * t + k.tphase;
  return k.time_fn == PTF_FN_SIN ? std::sin(a) : (k.time_fn == PTF_FN_COS ? std::cos(a) : 1.0);
}

template <typename T>
static void patch_analytic(Group& g, Slot& s, double t) {
  AnalyticArgs<T>& a = args_of<AnalyticArgs<T>>(s);
  for (size_t k = 0; k < g.aterms.size(); ++k)
    a.amp[k] = g.aterms[k].amp * term_time_factor(g.aterms[k], t);
}

// set the clock of the per-step flow synthesis launches to t
static void patch_flowsyn(Group& g, double t) {
  if (g.flowsyn.empty()) return;
  Slot& s = g.flowsyn[0];
  if (s.kind == K_ANALYTIC) {
    if (g.dbl) patch_analytic<double>(g, s, t);
    else patch_analytic<float>(g, s, t);
  } else if (g.dbl) {
    args_of<ColArgs<double>>(s).t = t;
  } else {
    args_of<ColArgs<float>>(s).t = t;
  }
}

// ---- transfers in the public layout, converted on the GPU --------------------------
// One member's field crosses PCIe in the reference layout (FP64); the brick layout
// and the working precision are applied on the device.  rowlen = entries per x-row
// (nx/2 x-pairs of a physical field, nx/2+1 spectral values), w = doubles per entry.
static int ensure_stage(Group& g) {
  const size_t need = (size_t)std::max<long long>(g.nR, 2 * g.nS) * sizeof(double);
  if (!g.stage.p || g.stage.n < need) CHK(dalloc(g.stage, need));
  return PTF_OK;
}

static unsigned layout_blocks(long long n) {
  return (unsigned)std::min<long long>((n + 255) / 256, 16384);
}

// host public FP64 -> device working layout (dst, and dst_lo for a hi + lo split);
// dbl selects the device element type (FP64 fields of the mixed form pass true)
static int put_public(Group& g, bool dbl, void* dst, void* dst_lo, const double* src,
                      long long rowlen, int w, hipStream_t st) {
  CHK(ensure_stage(g));
  const long long rows = (long long)g.ny * g.nz;
  const long long n = rows * rowlen * w;
  HIPCHK(hipMemcpyAsync(g.stage.p, src, (size_t)n * sizeof(double), hipMemcpyHostToDevice, st));
  if (dbl)
    hipLaunchKernelGGL(layout_in_kernel<double>, dim3(layout_blocks(n)), dim3(256), 0, st,
                       (const double*)g.stage.p, (double*)dst, (double*)dst_lo, rows, (int)rowlen, w,
                       g.lb, g.ndim, g.ny);
  else
    hipLaunchKernelGGL(layout_in_kernel<float>, dim3(layout_blocks(n)), dim3(256), 0, st,
                       (const double*)g.stage.p, (float*)dst, (float*)dst_lo, rows, (int)rowlen, w,
                       g.lb, g.ndim, g.ny);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(st));
  return PTF_OK;
}

// device working layout (src, plus src_lo) -> host public FP64
static int get_public(Group& g, bool dbl, double* dst, const void* src, const void* src_lo,
                      long long rowlen, int w, hipStream_t st) {
  CHK(ensure_stage(g));
  const long long rows = (long long)g.ny * g.nz;
  const long long n = rows * rowlen * w;
  if (dbl)
    hipLaunchKernelGGL(layout_out_kernel<double>, dim3(layout_blocks(n)), dim3(256), 0, st,
                       (const double*)src, (const double*)src_lo, (double*)g.stage.p, rows,
                       (int)rowlen, w, g.lb, g.ndim, g.ny);
  else
    hipLaunchKernelGGL(layout_out_kernel<float>, dim3(layout_blocks(n)), dim3(256), 0, st,
                       (const float*)src, (const float*)src_lo, (double*)g.stage.p, rows,
                       (int)rowlen, w, g.lb, g.ndim, g.ny);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(dst, g.stage.p, (size_t)n * sizeof(double), hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  return PTF_OK;
}

static int group_step_launches(Group& g, hipStream_t st, double t) {
  if (!g.flowsyn.empty()) {
    patch_flowsyn(g, t);
    for (auto& fs : g.flowsyn) CHK(launch(fs, st));
  }
  if (!g.primed)
    for (auto& sl : g.prologue) CHK(launch(sl, st));
  for (int s = 0; s < 4; ++s)
    for (auto& sl : g.stages[s]) CHK(launch(sl, st));
  g.primed = true;
  return PTF_OK;
}

// launch the prologue if needed so that a captured step can assume primed scratch
static int group_prime(Group& g, hipStream_t st) {
  if (!g.primed)
    for (auto& sl : g.prologue) CHK(launch(sl, st));
  g.primed = true;
  return PTF_OK;
}

// all slots of one step, in order (flow synthesis first)
static std::vector<Slot*> step_slots(Group& g) {
  std::vector<Slot*> v;
  for (auto& s : g.flowsyn) v.push_back(&s);
  for (int s = 0; s < 4; ++s)
    for (auto& sl : g.stages[s]) v.push_back(&sl);
  return v;
}

static int time_slot(Group& g, Slot& s, int reps, hipStream_t st, double* avg_ms) {
  if (reps < 1) reps = 1;
  g.primed = false;  // trial launches overwrite the transform scratch
  const size_t S = (size_t)g.nS * g.B * 2 * g.esz;
  if (s.writes_state) {
    if (!g.backup.p) CHK(dalloc(g.backup, 2 * S));
    HIPCHK(hipMemcpyAsync(g.backup.p, g.sol.p, S, hipMemcpyDeviceToDevice, st));
    if (g.sol_lo.p)
      HIPCHK(hipMemcpyAsync((char*)g.backup.p + S, g.sol_lo.p, S, hipMemcpyDeviceToDevice, st));
  }
  CHK(launch(s, st));  // warm-up
  hipEvent_t e0, e1;
  HIPCHK(hipEventCreate(&e0));
  HIPCHK(hipEventCreate(&e1));
  HIPCHK(hipEventRecord(e0, st));
  for (int r = 0; r < reps; ++r) CHK(launch(s, st));
  HIPCHK(hipEventRecord(e1, st));
  HIPCHK(hipEventSynchronize(e1));
  float ms = 0;
  HIPCHK(hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (s.writes_state) {
    HIPCHK(hipMemcpyAsync(g.sol.p, g.backup.p, S, hipMemcpyDeviceToDevice, st));
    if (g.sol_lo.p)
      HIPCHK(hipMemcpyAsync(g.sol_lo.p, (char*)g.backup.p + S, S, hipMemcpyDeviceToDevice, st));
    HIPCHK(hipStreamSynchronize(st));
  }
  *avg_ms = (double)ms / reps;
  return PTF_OK;
}

// autotune: for every distinct (kind, N, lines) pick the fastest valid (R, W, form).  A
// candidate is scored by the sum of its times over the distinct launches of that key
// (the four stages of a stage kernel read different operands), each timed in isolation.
// Launches that share a key get one variant.  The stage kernels of the four stages read
// different operands, so each stage's kernel is tuned on its own (its slot name carries
// the stage); other kernels of a step repeat per stage with the same arguments.
static std::string tune_key(const Slot& s) {
  const bool stage_kind = s.kind == K_COL_STAGE || s.kind == K_COL_STAGE_D1 || s.kind == K_COL_STAGE_RK4 ||
                          s.kind == K_COL_STAGE_D1_RK4 || s.kind == K_ROW_ADV1;
  return std::to_string(s.kind) + ":" + std::to_string(s.N) + ":" + std::to_string(s.lines) +
         ":" + std::to_string(s.members) + (stage_kind ? ":" + s.name : std::string());
}

// In-step durations: `nsteps` whole steps of the group with an event after every launch
// (after one unmeasured step); out[key] = the summed duration per step of the launches of
// each tune key.  The state advances (callers restore it).
static int instep_times(Group& g, hipStream_t st, int nsteps,
                        std::vector<std::pair<std::string, double>>* out);

// one candidate of a tune key: (R, W, occ, par, wg) with its isolated-launch score
struct TuneCand {
  int R, W, occ, par, wg;
  double ms;
};

static int autotune_group(Group& g, hipStream_t st) {
  std::vector<Slot*> all = step_slots(g);
  std::vector<std::string> done;
  std::vector<std::pair<std::string, std::vector<TuneCand>>> shortlist;
  for (Slot* s : all) {
    const std::string key = tune_key(*s);
    if (std::find(done.begin(), done.end(), key) != done.end()) continue;
    done.push_back(key);
    std::vector<Slot*> same;   // launches of this key with distinct arguments
    for (Slot* o : all) {
      if (tune_key(*o) != key) continue;
      bool dup = false;
      for (Slot* u : same) dup = dup || u->args == o->args;
      if (!dup) same.push_back(o);
    }
    const int* widths = s->is_col() ? kColWidths : kRowWidths;
    const int nw = s->is_col() ? kNumColWidths : kNumRowWidths;
    TuneCand bestc = {s->R, s->W, s->occ, s->par, s->wg, 1e30};
    const bool verbose = std::getenv("PTF_AUTOTUNE_VERBOSE") != nullptr;
    std::vector<TuneCand> cands;
    for (int wg : {0, 1}) {
      for (int par : {0, 1}) {
        for (int occ : {0, 3}) {
          for (int R : kRadices) {
            for (int wi = 0; wi < nw; ++wi) {
              double tot = 0;
              bool ok = true;
              for (Slot* o : same) {
                Slot trial = *o;
                if (!choose_variant(trial, R, widths[wi], occ, par, wg)) {
                  ok = false;
                  break;
                }
                double ms = 0;
                CHK(time_slot(g, trial, 5, st, &ms));
                tot += ms;
              }
              if (!ok) continue;
              if (verbose)
                std::fprintf(stderr, "autotune %s N=%d R=%d W=%d occ=%d par=%d wg=%d: %.4f ms (%d launches)\n",
                             kind_name(s->kind), s->N, R, widths[wi], occ, par, wg, tot, (int)same.size());
              const TuneCand c = {R, widths[wi], occ, par, wg, tot};
              cands.push_back(c);
              if (tot < bestc.ms) bestc = c;
            }
          }
        }
      }
    }
    for (Slot* o : all)
      if (tune_key(*o) == key) choose_variant(*o, bestc.R, bestc.W, bestc.occ, bestc.par, bestc.wg);
    std::sort(cands.begin(), cands.end(), [](const TuneCand& a, const TuneCand& b) { return a.ms < b.ms; });
    if (cands.size() > 4) cands.resize(4);
    shortlist.push_back({key, cands});
  }
  // Second pass, inside running steps: a launch timed alone re-reads data that its previous
  // repetition left in the caches, which favours some variants; here each key's shortlist
  // (the four fastest alone) is timed in whole steps, where every kernel reads what the
  // previous one wrote, and the fastest there is kept.  The state is restored afterwards.
  if (std::getenv("PTF_AUTOTUNE_ISOLATED") == nullptr && !g.plan_only && g.sol.p) {
    const size_t S = (size_t)g.nS * g.B * 2 * g.esz;
    if (!g.backup.p) CHK(dalloc(g.backup, 2 * S));
    HIPCHK(hipMemcpyAsync(g.backup.p, g.sol.p, S, hipMemcpyDeviceToDevice, st));
    if (g.sol_lo.p)
      HIPCHK(hipMemcpyAsync((char*)g.backup.p + S, g.sol_lo.p, S, hipMemcpyDeviceToDevice, st));
    const bool verbose = std::getenv("PTF_AUTOTUNE_VERBOSE") != nullptr;
    for (auto& ent : shortlist) {
      const std::string& key = ent.first;
      if (ent.second.size() < 2) continue;
      double best = 1e30;
      TuneCand pick = ent.second[0];
      for (const TuneCand& c : ent.second) {
        for (Slot* o : all)
          if (tune_key(*o) == key) choose_variant(*o, c.R, c.W, c.occ, c.par, c.wg);
        std::vector<std::pair<std::string, double>> t;
        CHK(instep_times(g, st, 3, &t));
        double ms = 0;
        for (auto& kv : t)
          if (kv.first == key) ms = kv.second;
        if (verbose)
          std::fprintf(stderr, "autotune in-step %s R=%d W=%d occ=%d par=%d wg=%d: %.4f ms/step (alone %.4f)\n",
                       key.c_str(), c.R, c.W, c.occ, c.par, c.wg, ms, c.ms);
        if (ms < best) {
          best = ms;
          pick = c;
        }
      }
      for (Slot* o : all)
        if (tune_key(*o) == key) choose_variant(*o, pick.R, pick.W, pick.occ, pick.par, pick.wg);
    }
    HIPCHK(hipMemcpyAsync(g.sol.p, g.backup.p, S, hipMemcpyDeviceToDevice, st));
    if (g.sol_lo.p)
      HIPCHK(hipMemcpyAsync(g.sol_lo.p, (char*)g.backup.p + S, S, hipMemcpyDeviceToDevice, st));
    HIPCHK(hipStreamSynchronize(st));
    g.primed = false;
  }
  return PTF_OK;
}

static int instep_times(Group& g, hipStream_t st, int nsteps,
                        std::vector<std::pair<std::string, double>>* out) {
  const std::vector<Slot*> v = step_slots(g);
  const size_t ns = v.size();
  g.primed = false;
  CHK(group_step_launches(g, st, 0.0));   // one unmeasured step (also primes)
  std::vector<hipEvent_t> ev((size_t)nsteps * ns + 1, nullptr);
  int rc = PTF_OK;
  for (auto& e : ev) {
    if (hipEventCreate(&e) != hipSuccess) {
      e = nullptr;
      rc = fail(PTF_ERR_DEVICE, "hipEventCreate failed");
      break;
    }
  }
  size_t k = 0;
  if (rc == PTF_OK && hipEventRecord(ev[k++], st) != hipSuccess) rc = fail(PTF_ERR_DEVICE, "hipEventRecord");
  for (int i = 0; i < nsteps && rc == PTF_OK; ++i) {
    patch_flowsyn(g, 0.0);
    for (Slot* s : v) {
      if ((rc = launch(*s, st)) != PTF_OK) break;
      if (hipEventRecord(ev[k++], st) != hipSuccess) {
        rc = fail(PTF_ERR_DEVICE, "hipEventRecord");
        break;
      }
    }
  }
  if (rc == PTF_OK && hipEventSynchronize(ev[k - 1]) != hipSuccess)
    rc = fail(PTF_ERR_DEVICE, "hipEventSynchronize");
  if (rc == PTF_OK) {
    out->clear();
    for (size_t e = 1; e < k && rc == PTF_OK; ++e) {
      float ms = 0;
      if (hipEventElapsedTime(&ms, ev[e - 1], ev[e]) != hipSuccess) {
        rc = fail(PTF_ERR_DEVICE, "hipEventElapsedTime");
        break;
      }
      const std::string key = tune_key(*v[(e - 1) % ns]);
      bool found = false;
      for (auto& kv : *out)
        if (kv.first == key) {
          kv.second += ms / nsteps;
          found = true;
        }
      if (!found) out->push_back({key, ms / nsteps});
    }
  }
  for (auto& e : ev)
    if (e) (void)hipEventDestroy(e);
  g.primed = false;
  return rc;
}

// ---------------------------------------------------------------------------
// public handles
// ---------------------------------------------------------------------------
struct ptf_problem {
  ptf_desc d;
  Group g;
  double t = 0.0;
  long long step = 0;
  hipStream_t stream = nullptr;
  bool own_stream = false;
  bool flow_set = false;
  // replay of kGraphSteps captured steps (steady flows; rebuilt after variant changes)
  bool use_graph = true;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  void drop_graph() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    exec = nullptr;
    graph = nullptr;
  }
  ~ptf_problem() {
    drop_graph();
    if (own_stream && stream) (void)hipStreamDestroy(stream);
  }
};
static constexpr int kGraphSteps = 8;

struct ptf_plan {
  Group g;
  hipStream_t stream = nullptr;
  bool own_stream = false;
  int device = 0;   // the device current at creation (holds the plan's buffers and stream)
  ~ptf_plan() {
    if (own_stream && stream) (void)hipStreamDestroy(stream);
  }
};

struct ptf_ensemble {
  std::vector<std::unique_ptr<Group>> groups;
  std::vector<std::pair<int, int>> where;  // member -> (group, index in group)
  // the members' clocks: every member steps with the ensemble; t advances by the member's
  // dt after each step (t += dt, as FourierFlows' clock), kept per group (one dt per group)
  std::vector<double> gt;
  int64_t step = 0;
  hipStream_t stream = nullptr;
  bool own_stream = false;
  std::vector<hipStream_t> gstreams;
  std::vector<hipEvent_t> events;
  // captured step graphs: [0] one step, [1] kEnsGraphSteps steps with each group's
  // launches in one chain
  hipGraph_t graph[2] = {nullptr, nullptr};
  hipGraphExec_t exec[2] = {nullptr, nullptr};
  ~ptf_ensemble() {
    for (int i = 0; i < 2; ++i) {
      if (exec[i]) (void)hipGraphExecDestroy(exec[i]);
      if (graph[i]) (void)hipGraphDestroy(graph[i]);
    }
    for (auto e : events) (void)hipEventDestroy(e);
    for (auto s : gstreams) (void)hipStreamDestroy(s);
    if (own_stream && stream) (void)hipStreamDestroy(stream);
  }
};

// sol of one member as FP64 (hi + lo in the mixed form)
static int get_member_sol(Group& g, int member, double* out_pub, hipStream_t st) {
  const size_t off = (size_t)member * g.nS * 2 * g.esz;
  const void* lo = g.sol_lo.p ? (const char*)g.sol_lo.p + off : nullptr;
  return get_public(g, g.dbl, out_pub, (const char*)g.sol.p + off, lo, g.nxh, 2, st);
}

// set sol of one member from FP64 (split into hi + lo in the mixed form)
static int set_member_sol(Group& g, int member, const double* in_pub, hipStream_t st) {
  g.primed = false;
  const size_t off = (size_t)member * g.nS * 2 * g.esz;
  void* lo = g.sol_lo.p ? (char*)g.sol_lo.p + off : nullptr;
  return put_public(g, g.dbl, (char*)g.sol.p + off, lo, in_pub, g.nxh, 2, st);
}

// after a set_c the low part of the member's compensated state is zero
static int clear_member_lo(Group& g, int member, hipStream_t st) {
  if (!g.sol_lo.p) return PTF_OK;
  const size_t nb = (size_t)g.nS * 2 * g.esz;
  HIPCHK(hipMemsetAsync((char*)g.sol_lo.p + (size_t)member * nb, 0, nb, st));
  return PTF_OK;
}

// The library is compiled against one HIP runtime (HIP_VERSION_MAJOR.HIP_VERSION_MINOR).
// A process that has already loaded another libamdhip64.so.7 (a framework that bundles
// its own ROCm release) binds this library to that runtime as well, because the dynamic
// loader reuses a loaded soname; code objects and runtime then come from different
// releases, which has ended in host crashes.  Every entry point that first touches the
// device checks the runtime it is bound to and refuses a different release.
static const std::string& runtime_mismatch() {
  static const std::string msg = [] {
    int v = 0;
    const hipError_t e = hipRuntimeGetVersion(&v);
    Dl_info info;
    std::memset(&info, 0, sizeof(info));
    const char* path = "(unknown)";
    if (dladdr(reinterpret_cast<void*>(&hipRuntimeGetVersion), &info) && info.dli_fname)
      path = info.dli_fname;
    const int major = v / 10000000, minor = (v / 100000) % 100;
    if (e == hipSuccess && major == HIP_VERSION_MAJOR && minor == HIP_VERSION_MINOR)
      return std::string();
    char buf[512];
    std::snprintf(buf, sizeof(buf),
                  "HIP runtime mismatch: libptf_hip was built against HIP %d.%d but is bound to "
                  "%s (HIP %d.%d, version %d). Another HIP runtime was loaded into this process "
                  "first; load libptf_hip before any framework that bundles its own ROCm, or use "
                  "a process without one",
                  HIP_VERSION_MAJOR, HIP_VERSION_MINOR, path, major, minor, v);
    return std::string(buf);
  }();
  return msg;
}

static int runtime_check() {
  const std::string& m = runtime_mismatch();
  return m.empty() ? PTF_OK : fail(PTF_ERR_DEVICE, m);
}

extern "C" {

const char* ptf_last_error(void) { return g_err.c_str(); }

int ptf_version(int32_t* major, int32_t* minor) {
  if (major) *major = 0;
  if (minor) *minor = 1;
  return PTF_OK;
}

int ptf_device_init(int32_t ordinal) {
  CHK(runtime_check());
  int n = 0;
  HIPCHK(hipGetDeviceCount(&n));
  if (ordinal < 0 || ordinal >= n)
    return fail(PTF_ERR_INVALID, "device ordinal " + std::to_string(ordinal) + " out of range (" +
                                     std::to_string(n) + " devices)");
  HIPCHK(hipSetDevice(ordinal));
  void* scr = nullptr;
  CHK(static_scratch(ordinal, &scr));   // the allocation-free entry point's scratch address
  return PTF_OK;
}

int ptf_device_synchronize(void) {
  HIPCHK(hipDeviceSynchronize());
  return PTF_OK;
}

int ptf_buffer_alloc(uint64_t nbytes, void** dptr) {
  if (!dptr) return fail(PTF_ERR_INVALID, "dptr is NULL");
  CHK(runtime_check());
  hipError_t e = hipMalloc(dptr, nbytes ? nbytes : 16);
  if (e != hipSuccess) {
    *dptr = nullptr;
    (void)hipGetLastError();
    return fail(PTF_ERR_NOMEM, std::string("hipMalloc failed: ") + hipGetErrorString(e));
  }
  return PTF_OK;
}

int ptf_buffer_free(void* dptr) {
  if (dptr) HIPCHK(hipFree(dptr));
  return PTF_OK;
}

int ptf_buffer_upload(void* dptr, const void* host, uint64_t nbytes) {
  if (!dptr || !host) return fail(PTF_ERR_INVALID, "NULL pointer");
  HIPCHK(hipMemcpy(dptr, host, nbytes, hipMemcpyHostToDevice));
  return PTF_OK;
}

int ptf_buffer_download(void* host, const void* dptr, uint64_t nbytes) {
  if (!dptr || !host) return fail(PTF_ERR_INVALID, "NULL pointer");
  HIPCHK(hipMemcpy(host, dptr, nbytes, hipMemcpyDeviceToHost));
  return PTF_OK;
}

// ----- plans -----------------------------------------------------------------
int ptf_plan_create(int32_t ndim, const int64_t n[3], int32_t precision, ptf_plan** out) {
  return ptf_plan_create_batched(ndim, n, 1, precision, out);
}

int ptf_plan_create_batched(int32_t ndim, const int64_t n[3], int64_t batch, int32_t precision,
                            ptf_plan** out) {
  if (!out || !n) return fail(PTF_ERR_INVALID, "NULL argument");
  *out = nullptr;
  if (batch < 1 || batch > 65535) return fail(PTF_ERR_INVALID, "batch must be in [1, 65535]");
  ptf_desc d;
  std::memset(&d, 0, sizeof(d));
  d.ndim = ndim;
  d.precision = precision;
  for (int a = 0; a < 3; ++a) {
    d.n[a] = a < ndim ? n[a] : 1;
    d.L[a] = 2 * M_PI;
  }
  d.dt = 1.0;
  CHK(validate_desc(&d));
  std::unique_ptr<ptf_plan> p(new ptf_plan());
  HIPCHK(hipGetDevice(&p->device));
  std::vector<ptf_desc> ds((size_t)batch, d);
  CHK(group_init(p->g, ds.data(), (int)batch, /*allow_brick=*/false, /*plan_only=*/true));
  HIPCHK(hipStreamCreate(&p->stream));
  p->own_stream = true;
  *out = p.release();
  return PTF_OK;
}

int ptf_plan_destroy(ptf_plan* plan) {
  delete plan;
  return PTF_OK;
}

int ptf_plan_set_stream(ptf_plan* plan, void* s) {
  if (!plan) return fail(PTF_ERR_INVALID, "plan is NULL");
  if (plan->own_stream && plan->stream) (void)hipStreamDestroy(plan->stream);
  plan->own_stream = false;
  plan->stream = (hipStream_t)s;
  return PTF_OK;
}

int ptf_rfft(ptf_plan* plan, const void* in_real, void* out_complex) {
  if (!plan || !in_real || !out_complex) return fail(PTF_ERR_INVALID, "NULL argument");
  CHK(g_rfft(plan->g, in_real, out_complex, plan->stream));
  return PTF_OK;
}

int ptf_irfft(ptf_plan* plan, const void* in_complex, void* out_real) {
  if (!plan || !in_complex || !out_real) return fail(PTF_ERR_INVALID, "NULL argument");
  CHK(g_irfft(plan->g, in_complex, out_real, plan->stream));
  return PTF_OK;
}

int ptf_irfft_noalloc(ptf_plan* plan, const void* in_complex, void* out_real) {
  if (!plan || !in_complex || !out_real) return fail(PTF_ERR_INVALID, "NULL argument");
  const Group& g = plan->g;
  // 2D/3D: one intermediate spectrum of every member; 1D needs none
  const size_t need = g.ndim >= 2 ? (size_t)g.nS * (size_t)g.B * 2 * g.esz : 0;
  if (need > kStaticScratchBytes)
    return fail(PTF_ERR_INVALID, "ptf_irfft_noalloc: the plan needs " + std::to_string(need) +
                                     " bytes of scratch; the static scratch holds " +
                                     std::to_string(kStaticScratchBytes) + " (use ptf_irfft)");
  void* scr = nullptr;
  if (need) CHK(static_scratch(plan->device, &scr));
  CHK(g_irfft(plan->g, in_complex, out_real, plan->stream, scr));
  return PTF_OK;
}

// [FF] getaliasedwavenumbers(nk, nkr, f): 1-based iL = floor((1-f)/2 nk) + 1,
// iR = ceil((1+f)/2 nk); full axis aliased indices [iL-1, iR), half axis [iL-1, nkr)
static void aliased_range(int nk, double f, int* lo, int* hi) {
  const int iL = (int)std::floor((1.0 - f) / 2.0 * nk) + 1;
  const int iR = (int)std::ceil((1.0 + f) / 2.0 * nk);
  *lo = iL - 1;
  *hi = iR;
}

int ptf_dealias(ptf_plan* plan, void* spec, double aliased_fraction) {
  if (!plan || !spec) return fail(PTF_ERR_INVALID, "NULL argument");
  if (!(aliased_fraction < 1.0)) return fail(PTF_ERR_INVALID, "aliased_fraction must be < 1");
  if (!(aliased_fraction > 0.0)) return PTF_OK;   // [FF] dealias! does nothing then
  Group& g = plan->g;
  int rlo, rhi, ylo = 0, yhi = 0, zlo = 0, zhi = 0;
  aliased_range(g.nx, aliased_fraction, &rlo, &rhi);
  rhi = g.nxh;
  if (g.ndim >= 2) aliased_range(g.ny, aliased_fraction, &ylo, &yhi);
  if (g.ndim == 3) aliased_range(g.nz, aliased_fraction, &zlo, &zhi);
  const long long n = g.nS * g.B;
  const int th = 256;
  const unsigned blocks = (unsigned)std::min<long long>((n + th - 1) / th, 65536);
  if (g.dbl)
    hipLaunchKernelGGL(dealias_kernel<double>, dim3(blocks), dim3(th), 0, plan->stream,
                       (cpx<double>*)spec, n, g.nxh, g.ny, g.nz, rlo, rhi, ylo, yhi, zlo, zhi);
  else
    hipLaunchKernelGGL(dealias_kernel<float>, dim3(blocks), dim3(th), 0, plan->stream,
                       (cpx<float>*)spec, n, g.nxh, g.ny, g.nz, rlo, rhi, ylo, yhi, zlo, zhi);
  HIPCHK(hipGetLastError());
  return PTF_OK;
}

// ----- problems ----------------------------------------------------------------
int ptf_problem_create(const ptf_desc* desc, ptf_problem** out) {
  if (!out) return fail(PTF_ERR_INVALID, "out is NULL");
  *out = nullptr;
  CHK(validate_desc(desc));
  std::unique_ptr<ptf_problem> p(new ptf_problem());
  p->d = *desc;
  CHK(group_init(p->g, desc, 1));
  HIPCHK(hipStreamCreate(&p->stream));
  p->own_stream = true;
  *out = p.release();
  return PTF_OK;
}

int ptf_problem_destroy(ptf_problem* prob) {
  if (prob && prob->stream) (void)hipStreamSynchronize(prob->stream);
  delete prob;
  return PTF_OK;
}

int ptf_problem_set_stream(ptf_problem* prob, void* s) {
  if (!prob) return fail(PTF_ERR_INVALID, "problem is NULL");
  if (prob->own_stream && prob->stream) {
    (void)hipStreamSynchronize(prob->stream);
    (void)hipStreamDestroy(prob->stream);
  }
  prob->own_stream = false;
  prob->stream = (hipStream_t)s;
  return PTF_OK;
}

int ptf_problem_get_stream(ptf_problem* prob, void** s) {
  if (!prob || !s) return fail(PTF_ERR_INVALID, "NULL argument");
  *s = (void*)prob->stream;
  return PTF_OK;
}

int ptf_set_flow(ptf_problem* prob, const double* u, const double* v, const double* w) {
  if (!prob) return fail(PTF_ERR_INVALID, "problem is NULL");
  if (prob->d.flow_kind == PTF_FLOW_SPECTRAL)
    return fail(PTF_ERR_INVALID, "problem uses the built-in spectral flow; use ptf_set_spectral_flow");
  const double* f[3] = {u, v, w};
  Group& g = prob->g;
  for (int i = 0; i < g.ndim; ++i) {
    if (!f[i]) return fail(PTF_ERR_INVALID, "flow component " + std::to_string(i) + " is NULL");
    CHK(put_public(g, g.dbl, g.flow[i].p, nullptr, f[i], g.nx / 2, 2, prob->stream));
  }
  HIPCHK(hipStreamSynchronize(prob->stream));
  prob->flow_set = true;
  return PTF_OK;
}

int ptf_set_spectral_flow(ptf_problem* prob, const double* psih0, const double* omega) {
  if (!prob || !psih0 || !omega) return fail(PTF_ERR_INVALID, "NULL argument");
  if (prob->d.flow_kind != PTF_FLOW_SPECTRAL)
    return fail(PTF_ERR_INVALID, "problem was not created with PTF_FLOW_SPECTRAL");
  Group& g = prob->g;
  CHK(put_public(g, true, g.psih0.p, nullptr, psih0, g.nxh, 2, prob->stream));
  CHK(put_public(g, true, g.omega.p, nullptr, omega, g.nxh, 1, prob->stream));
  prob->flow_set = true;
  return PTF_OK;
}

}  // extern "C"

template <typename T>
static int set_analytic_impl(ptf_problem* prob, const ptf_flow_term* terms, int nterms) {
  Group& g = prob->g;
  const int ns[3] = {g.nx, g.ny, g.nz};
  // factor tables at the grid points x_j = -L/2 + j L/n (rounded once from long double)
  size_t off[3];
  size_t pos = 0;
  for (int d = 0; d < 3; ++d) {
    off[d] = pos;
    pos += (size_t)nterms * (d < g.ndim ? ns[d] : 1);
  }
  std::vector<double> tab(pos, 1.0);
  for (int d = 0; d < g.ndim; ++d) {
    const long double L = g.Ls[d];
    for (int k = 0; k < nterms; ++k) {
      const ptf_flow_term& tm = terms[k];
      const int fn = d == 0 ? tm.fx : (d == 1 ? tm.fy : tm.fz);
      const double kk = d == 0 ? tm.kx : (d == 1 ? tm.ky : tm.kz);
      const double ph = d == 0 ? tm.px : (d == 1 ? tm.py : tm.pz);
      for (int j = 0; j < ns[d]; ++j) {
        const double x = (double)(-L / 2 + (long double)j * (L / ns[d]));
        const double a = kk * x + ph;
        tab[off[d] + (size_t)k * ns[d] + j] =
            fn == PTF_FN_SIN ? std::sin(a) : (fn == PTF_FN_COS ? std::cos(a) : 1.0);
      }
    }
  }
  CHK(dalloc(g.atab, tab.size() * sizeof(double)));
  HIPCHK(hipMemcpy(g.atab.p, tab.data(), tab.size() * sizeof(double), hipMemcpyHostToDevice));
  g.aterms.assign(terms, terms + nterms);
  AnalyticArgs<T> a;
  std::memset(&a, 0, sizeof(a));
  a.nterms = nterms;
  for (int k = 0; k < nterms; ++k) {
    a.comp[k] = terms[k].component;
    a.amp[k] = terms[k].amp;
  }
  a.tx = (const double*)g.atab.p + off[0];
  a.ty = (const double*)g.atab.p + off[1];
  a.tz = (const double*)g.atab.p + off[2];
  for (int c = 0; c < g.ndim; ++c) a.out[c] = (T*)g.flow[c].p;
  a.ndim = g.ndim;
  a.nx = g.nx;
  a.ny = g.ny;
  a.nz = g.nz;
  a.lb = g.lb;
  a.npoints = g.nR;
  Slot s;
  s.name = "flow_analytic_synthesis";
  s.dbl = g.dbl;
  s.kind = K_ANALYTIC;
  s.fn = reinterpret_cast<const void*>(&analytic_flow_kernel<T>);
  s.fixed_threads = 256;
  s.fixed_blocks = (unsigned)std::min<long long>((g.nR + 255) / 256, 16384);
  set_args(s, a);
  bool tdep = false;
  for (int k = 0; k < nterms; ++k) tdep = tdep || terms[k].time_fn != PTF_FN_ONE;
  g.flowsyn.clear();
  if (tdep) {
    g.flowsyn.push_back(s);  // synthesised at clock.t before every step
  } else {
    CHK(launch(s, prob->stream));
    HIPCHK(hipStreamSynchronize(prob->stream));
  }
  return PTF_OK;
}

extern "C" {

int ptf_set_analytic_flow(ptf_problem* prob, const ptf_flow_term* terms, int32_t nterms) {
  if (!prob || !terms) return fail(PTF_ERR_INVALID, "NULL argument");
  if (prob->d.flow_kind != PTF_FLOW_ANALYTIC)
    return fail(PTF_ERR_INVALID, "problem was not created with PTF_FLOW_ANALYTIC");
  if (nterms < 1 || nterms > PTF_MAX_FLOW_TERMS)
    return fail(PTF_ERR_INVALID, "nterms must be in [1, PTF_MAX_FLOW_TERMS]");
  for (int k = 0; k < nterms; ++k) {
    const ptf_flow_term& t = terms[k];
    if (t.component < 0 || t.component >= prob->g.ndim)
      return fail(PTF_ERR_INVALID, "term " + std::to_string(k) + ": component out of range");
    for (int f : {t.fx, t.fy, t.fz, t.time_fn})
      if (f < PTF_FN_ONE || f > PTF_FN_COS)
        return fail(PTF_ERR_INVALID, "term " + std::to_string(k) + ": unknown function");
    for (double v : {t.kx, t.ky, t.kz, t.px, t.py, t.pz, t.amp, t.omega, t.tphase})
      if (!std::isfinite(v))
        return fail(PTF_ERR_INVALID, "term " + std::to_string(k) + ": not finite");
  }
  prob->g.primed = false;
  CHK(prob->g.dbl ? set_analytic_impl<double>(prob, terms, nterms)
                  : set_analytic_impl<float>(prob, terms, nterms));
  prob->flow_set = true;
  return PTF_OK;
}

int ptf_set_c(ptf_problem* prob, const double* c) {
  if (!prob || !c) return fail(PTF_ERR_INVALID, "NULL argument");
  Group& g = prob->g;
  CHK(put_public(g, g.dbl, g.phys.p, nullptr, c, g.nx / 2, 2, prob->stream));
  CHK(g_rfft(g, g.phys.p, g.sol.p, prob->stream));
  CHK(clear_member_lo(g, 0, prob->stream));
  g.primed = false;
  HIPCHK(hipStreamSynchronize(prob->stream));
  return PTF_OK;
}

// Capture kGraphSteps steps of a primed problem into a graph (kernels only).
static int build_step_graph(ptf_problem* prob) {
  prob->drop_graph();
  Group& g = prob->g;
  HIPCHK(hipStreamBeginCapture(prob->stream, hipStreamCaptureModeThreadLocal));
  int rc = PTF_OK;
  for (int i = 0; i < kGraphSteps && rc == PTF_OK; ++i) {
    for (int s = 0; s < 4 && rc == PTF_OK; ++s)
      for (auto& sl : g.stages[s])
        if ((rc = launch(sl, prob->stream)) != PTF_OK) break;
  }
  hipGraph_t graph = nullptr;
  const hipError_t ce = hipStreamEndCapture(prob->stream, &graph);
  if (rc != PTF_OK || ce != hipSuccess) {
    if (graph) (void)hipGraphDestroy(graph);
    return rc != PTF_OK ? rc : fail(PTF_ERR_DEVICE, std::string("graph capture: ") + hipGetErrorString(ce));
  }
  prob->graph = graph;
  if (hipGraphInstantiate(&prob->exec, prob->graph, nullptr, nullptr, 0) != hipSuccess) {
    prob->drop_graph();
    return fail(PTF_ERR_DEVICE, "hipGraphInstantiate failed");
  }
  return PTF_OK;
}

int ptf_step(ptf_problem* prob, int64_t nsteps) {
  if (!prob) return fail(PTF_ERR_INVALID, "problem is NULL");
  if (nsteps < 0) return fail(PTF_ERR_INVALID, "nsteps must be >= 0");
  Group& g = prob->g;
  int64_t done = 0;
  // steady flows: whole blocks of kGraphSteps steps replay a captured graph (the
  // steps' kernels and arguments do not depend on the clock)
  if (prob->use_graph && g.flowsyn.empty() && nsteps >= kGraphSteps) {
    CHK(group_prime(g, prob->stream));
    if (!prob->exec) CHK(build_step_graph(prob));
    for (; done + kGraphSteps <= nsteps; done += kGraphSteps) {
      HIPCHK(hipGraphLaunch(prob->exec, prob->stream));
      for (int i = 0; i < kGraphSteps; ++i) {
        prob->t += g.dt;
        prob->step += 1;
      }
    }
  }
  for (int64_t i = done; i < nsteps; ++i) {
    CHK(group_step_launches(g, prob->stream, prob->t));
    prob->t += g.dt;
    prob->step += 1;
  }
  HIPCHK(hipGetLastError());
  return PTF_OK;
}

int ptf_problem_set_graph(ptf_problem* prob, int32_t enable) {
  if (!prob) return fail(PTF_ERR_INVALID, "problem is NULL");
  prob->use_graph = enable != 0;
  prob->drop_graph();
  return PTF_OK;
}

int ptf_step_profiled(ptf_problem* prob, int64_t nsteps, double* slot_ms) {
  if (!prob || !slot_ms) return fail(PTF_ERR_INVALID, "NULL argument");
  if (nsteps < 1) return fail(PTF_ERR_INVALID, "nsteps must be >= 1");
  Group& g = prob->g;
  hipStream_t st = prob->stream;
  const std::vector<Slot*> v = step_slots(g);
  const size_t ns = v.size();
  CHK(group_prime(g, st));
  // one event before the first launch and one after every launch of every step
  std::vector<hipEvent_t> ev((size_t)nsteps * ns + 1, nullptr);
  int rc = PTF_OK;
  for (auto& e : ev) {
    if (hipEventCreate(&e) != hipSuccess) {
      e = nullptr;
      rc = fail(PTF_ERR_DEVICE, "hipEventCreate failed");
      break;
    }
  }
  size_t k = 0;
  if (rc == PTF_OK && hipEventRecord(ev[k++], st) != hipSuccess) rc = fail(PTF_ERR_DEVICE, "hipEventRecord");
  for (int64_t i = 0; i < nsteps && rc == PTF_OK; ++i) {
    patch_flowsyn(g, prob->t);
    for (Slot* s : v) {
      if ((rc = launch(*s, st)) != PTF_OK) break;
      if (hipEventRecord(ev[k++], st) != hipSuccess) {
        rc = fail(PTF_ERR_DEVICE, "hipEventRecord");
        break;
      }
    }
    if (rc != PTF_OK) break;
    prob->t += g.dt;
    prob->step += 1;
  }
  if (rc == PTF_OK && hipEventSynchronize(ev[k - 1]) != hipSuccess)
    rc = fail(PTF_ERR_DEVICE, "hipEventSynchronize");
  if (rc == PTF_OK) {
    for (size_t j = 0; j < ns; ++j) slot_ms[j] = 0.0;
    for (size_t e = 1; e < k && rc == PTF_OK; ++e) {
      float ms = 0;
      if (hipEventElapsedTime(&ms, ev[e - 1], ev[e]) != hipSuccess)
        rc = fail(PTF_ERR_DEVICE, "hipEventElapsedTime");
      slot_ms[(e - 1) % ns] += ms;
    }
    for (size_t j = 0; j < ns; ++j) slot_ms[j] /= (double)nsteps;
  }
  for (auto& e : ev)
    if (e) (void)hipEventDestroy(e);
  return rc;
}

int ptf_get_clock(ptf_problem* prob, double* t, int64_t* step) {
  if (!prob) return fail(PTF_ERR_INVALID, "problem is NULL");
  if (t) *t = prob->t;
  if (step) *step = prob->step;
  return PTF_OK;
}

int ptf_set_clock(ptf_problem* prob, double t, int64_t step) {
  if (!prob) return fail(PTF_ERR_INVALID, "problem is NULL");
  prob->t = t;
  prob->step = step;
  return PTF_OK;
}

int ptf_get_sol(ptf_problem* prob, double* sol) {
  if (!prob || !sol) return fail(PTF_ERR_INVALID, "NULL argument");
  return get_member_sol(prob->g, 0, sol, prob->stream);
}

int ptf_set_sol(ptf_problem* prob, const double* sol) {
  if (!prob || !sol) return fail(PTF_ERR_INVALID, "NULL argument");
  return set_member_sol(prob->g, 0, sol, prob->stream);
}

int ptf_updatevars(ptf_problem* prob, double* c) {
  if (!prob || !c) return fail(PTF_ERR_INVALID, "NULL argument");
  Group& g = prob->g;
  CHK(g_irfft(g, g.sol.p, g.phys.p, prob->stream));
  return get_public(g, g.dbl, c, g.phys.p, nullptr, g.nx / 2, 2, prob->stream);
}

}  // extern "C"

template <typename T>
static int calcN_impl(ptf_problem* prob, const double* sol, double* N) {
  Group& g = prob->g;
  hipStream_t st = prob->stream;
  if (!g.flowsyn.empty()) {
    patch_flowsyn(g, prob->t);
    for (auto& fs : g.flowsyn) CHK(launch(fs, st));
  }
  g.primed = false;  // the evaluation below uses the transform scratch
  CHK(put_public(g, g.dbl, g.stg.p, nullptr, sol, g.nxh, 2, st));
  StageArgs<T> none;
  std::memset(&none, 0, sizeof(none));
  none.kind = EPI_STORE;
  std::vector<Slot> list;
  CHK(build_calcN<T>(g, (const cpx<T>*)g.stg.p, none, (cpx<T>*)g.acc.p, list, "calcN_"));
  for (auto& s : list) CHK(launch(s, st));
  return get_public(g, g.dbl, N, g.acc.p, nullptr, g.nxh, 2, st);
}

extern "C" {

int ptf_calcN(ptf_problem* prob, const double* sol, double* N) {
  if (!prob || !sol || !N) return fail(PTF_ERR_INVALID, "NULL argument");
  return prob->g.dbl ? calcN_impl<double>(prob, sol, N) : calcN_impl<float>(prob, sol, N);
}

int ptf_get_L(ptf_problem* prob, double* L) {
  if (!prob || !L) return fail(PTF_ERR_INVALID, "NULL argument");
  std::memcpy(L, prob->g.L_host.data(), (size_t)prob->g.nS * sizeof(double));
  return PTF_OK;
}

static int diagnostics_group(Group& g, int member, hipStream_t st, double* mass, double* var) {
  const size_t cs = 2 * g.esz;
  const char* base = (const char*)g.sol.p + (size_t)member * g.nS * cs;
  // the mixed form's state is the compensated pair sol + sol_lo: both enter the FP64 sums
  const char* lo = g.sol_lo.p ? (const char*)g.sol_lo.p + (size_t)member * g.nS * cs : nullptr;
  double s0re = 0, s0im = 0;
  if (g.dbl) {
    double v[2];
    HIPCHK(hipMemcpyAsync(v, base, 16, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    s0re = v[0];
    s0im = v[1];
  } else {
    float v[2], w[2] = {0.f, 0.f};
    HIPCHK(hipMemcpyAsync(v, base, 8, hipMemcpyDeviceToHost, st));
    if (lo) HIPCHK(hipMemcpyAsync(w, lo, 8, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    s0re = (double)v[0] + (double)w[0];
    s0im = (double)v[1] + (double)w[1];
  }
  if (mass) *mass = s0re;
  if (var) {
    const int blocks = 1024;
    if (g.dbl)
      hipLaunchKernelGGL(variance_partial_kernel<double>, dim3(blocks), dim3(256), 0, st,
                         (const cpx<double>*)base, (const cpx<double>*)nullptr, g.nS, g.nxh, g.lb,
                         (double*)g.red.p);
    else
      hipLaunchKernelGGL(variance_partial_kernel<float>, dim3(blocks), dim3(256), 0, st,
                         (const cpx<float>*)base, (const cpx<float>*)lo, g.nS, g.nxh, g.lb,
                         (double*)g.red.p);
    HIPCHK(hipGetLastError());
    std::vector<double> part(2 * blocks);
    HIPCHK(hipMemcpyAsync(part.data(), g.red.p, part.size() * sizeof(double), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    double sum = 0, comp = 0;
    for (int i = 0; i < 2 * blocks; ++i) {
      const double x = part[i];
      const double t = sum + x;
      comp += (std::fabs(sum) >= std::fabs(x)) ? (sum - t) + x : (x - t) + sum;
      sum = t;
    }
    const double dc = s0re * s0re + s0im * s0im;
    const double tot = (sum - dc) + comp;
    const double N = (double)g.nR;
    *var = tot / (N * N);
  }
  return PTF_OK;
}

int ptf_diagnostics(ptf_problem* prob, double* mass, double* variance) {
  if (!prob) return fail(PTF_ERR_INVALID, "problem is NULL");
  return diagnostics_group(prob->g, 0, prob->stream, mass, variance);
}

int ptf_step_record(ptf_problem* prob, int64_t nsteps, int64_t every, int64_t capacity,
                    int64_t* nrec_out, double* t_out, int64_t* step_out, double* mass_out,
                    double* var_out, double* c_out) {
  if (!prob || !nrec_out) return fail(PTF_ERR_INVALID, "NULL argument");
  if (nsteps < 0 || every < 1) return fail(PTF_ERR_INVALID, "need nsteps >= 0 and every >= 1");
  Group& g = prob->g;
  hipStream_t st = prob->stream;
  // records after the steps whose new clock.step is a multiple of `every`
  auto floordiv = [](int64_t a, int64_t b) { return a >= 0 ? a / b : -((-a + b - 1) / b); };
  const int64_t s0 = prob->step;
  const int64_t nrec = floordiv(s0 + nsteps, every) - floordiv(s0, every);
  *nrec_out = nrec;
  if (nrec > capacity)
    return fail(PTF_ERR_INVALID, "capacity " + std::to_string(capacity) + " < " +
                                     std::to_string(nrec) + " records");
  const bool want_diag = mass_out || var_out;
  DevBuf rec, snaps;   // freed on return (hipFree waits for the device)
  const size_t rsz = (size_t)g.nR * g.esz;   // one physical snapshot
  int rc = PTF_OK;
  if (want_diag && nrec > 0) rc = dalloc(rec, (size_t)nrec * 2 * sizeof(double));
  if (rc == PTF_OK && c_out && nrec > 0) rc = dalloc(snaps, (size_t)nrec * rsz);
  if (rc != PTF_OK) return rc;
  const int blocks = 1024;
  int64_t r = 0;
  for (int64_t i = 0; i < nsteps && rc == PTF_OK; ++i) {
    rc = group_step_launches(g, st, prob->t);
    if (rc != PTF_OK) break;
    prob->t += g.dt;
    prob->step += 1;
    if (prob->step % every != 0) continue;
    if (t_out) t_out[r] = prob->t;
    if (step_out) step_out[r] = prob->step;
    if (want_diag) {
      double* dst = (double*)rec.p + 2 * r;
      if (g.dbl) {
        hipLaunchKernelGGL(variance_partial_kernel<double>, dim3(blocks), dim3(256), 0, st,
                           (const cpx<double>*)g.sol.p, (const cpx<double>*)nullptr, g.nS, g.nxh,
                           g.lb, (double*)g.red.p);
        hipLaunchKernelGGL(diag_record_kernel<double>, dim3(1), dim3(64), 0, st,
                           (const cpx<double>*)g.sol.p, (const cpx<double>*)nullptr,
                           (const double*)g.red.p, 2 * blocks, (double)g.nR, dst);
      } else {
        hipLaunchKernelGGL(variance_partial_kernel<float>, dim3(blocks), dim3(256), 0, st,
                           (const cpx<float>*)g.sol.p, (const cpx<float>*)g.sol_lo.p, g.nS, g.nxh,
                           g.lb, (double*)g.red.p);
        hipLaunchKernelGGL(diag_record_kernel<float>, dim3(1), dim3(64), 0, st,
                           (const cpx<float>*)g.sol.p, (const cpx<float>*)g.sol_lo.p,
                           (const double*)g.red.p, 2 * blocks, (double)g.nR, dst);
      }
      if (hipGetLastError() != hipSuccess) rc = fail(PTF_ERR_DEVICE, "diagnostics launch failed");
    }
    if (rc == PTF_OK && c_out) rc = g_irfft(g, g.sol.p, (char*)snaps.p + (size_t)r * rsz, st);
    ++r;
  }
  if (rc == PTF_OK && want_diag && nrec > 0) {
    std::vector<double> h((size_t)nrec * 2);
    if (hipMemcpyAsync(h.data(), rec.p, h.size() * sizeof(double), hipMemcpyDeviceToHost, st) !=
            hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess) {
      rc = fail(PTF_ERR_DEVICE, "download of the diagnostics failed");
    } else {
      for (int64_t k = 0; k < nrec; ++k) {
        if (mass_out) mass_out[k] = h[2 * k];
        if (var_out) var_out[k] = h[2 * k + 1];
      }
    }
  }
  if (rc == PTF_OK && c_out) {
    for (int64_t k = 0; k < nrec && rc == PTF_OK; ++k) {
      double* dst = c_out + (size_t)k * g.nR;
      const void* src = (const char*)snaps.p + (size_t)k * rsz;
      rc = get_public(g, g.dbl, dst, src, nullptr, g.nx / 2, 2, st);
    }
  }
  if (rc == PTF_OK && hipStreamSynchronize(st) != hipSuccess)
    rc = fail(PTF_ERR_DEVICE, "hipStreamSynchronize failed");
  return rc;
}

int ptf_problem_sol_ptr(ptf_problem* prob, void** dptr) {
  if (!prob || !dptr) return fail(PTF_ERR_INVALID, "NULL argument");
  *dptr = prob->g.sol.p;
  return PTF_OK;
}

int ptf_autotune(ptf_problem* prob) {
  if (!prob) return fail(PTF_ERR_INVALID, "problem is NULL");
  prob->drop_graph();
  return autotune_group(prob->g, prob->stream);
}

int ptf_problem_num_slots(ptf_problem* prob, int32_t* n) {
  if (!prob || !n) return fail(PTF_ERR_INVALID, "NULL argument");
  *n = (int32_t)step_slots(prob->g).size();
  return PTF_OK;
}

static int slot_name_into(Slot& s, char* buf, int32_t buflen) {
  if (!buf || buflen <= 0) return fail(PTF_ERR_INVALID, "bad buffer");
  char tmp[256];
  std::snprintf(tmp, sizeof(tmp), "%s [%s N=%d R=%d W=%d%s%s%s %s]", s.name.c_str(),
                kind_name(s.kind), s.N, s.R, s.W, s.occ ? " occ=3" : "", s.par ? " par" : "",
                s.wg ? " wg" : "", s.dbl ? "f64" : "f32");
  std::strncpy(buf, tmp, (size_t)buflen - 1);
  buf[buflen - 1] = 0;
  return PTF_OK;
}

int ptf_problem_slot_name(ptf_problem* prob, int32_t slot, char* buf, int32_t buflen) {
  if (!prob) return fail(PTF_ERR_INVALID, "problem is NULL");
  auto v = step_slots(prob->g);
  if (slot < 0 || slot >= (int)v.size()) return fail(PTF_ERR_INVALID, "slot out of range");
  return slot_name_into(*v[slot], buf, buflen);
}

int ptf_problem_time_slot(ptf_problem* prob, int32_t slot, int32_t reps, double* avg_ms) {
  if (!prob || !avg_ms) return fail(PTF_ERR_INVALID, "NULL argument");
  auto v = step_slots(prob->g);
  if (slot < 0 || slot >= (int)v.size()) return fail(PTF_ERR_INVALID, "slot out of range");
  return time_slot(prob->g, *v[slot], reps, prob->stream, avg_ms);
}

int ptf_problem_set_variant(ptf_problem* prob, int32_t slot, int32_t R, int32_t W, int32_t flags) {
  if (!prob) return fail(PTF_ERR_INVALID, "problem is NULL");
  auto v = step_slots(prob->g);
  if (slot < 0 || slot >= (int)v.size()) return fail(PTF_ERR_INVALID, "slot out of range");
  if (flags & ~(PTF_VARIANT_OCC3 | PTF_VARIANT_PAR | PTF_VARIANT_WG_LINES))
    return fail(PTF_ERR_INVALID, "unknown flags");
  Slot trial = *v[slot];
  if (!choose_variant(trial, R, W, (flags & PTF_VARIANT_OCC3) ? 3 : 0,
                      (flags & PTF_VARIANT_PAR) ? 1 : 0, (flags & PTF_VARIANT_WG_LINES) ? 1 : 0))
    return fail(PTF_ERR_UNSUPPORTED, "no compiled variant R=" + std::to_string(R) +
                                         " W=" + std::to_string(W) + " flags=" +
                                         std::to_string(flags) + " for slot " + trial.name);
  *v[slot] = trial;
  prob->drop_graph();
  return PTF_OK;
}

int ptf_problem_get_variant(ptf_problem* prob, int32_t slot, int32_t* R, int32_t* W, int32_t* flags) {
  if (!prob || !R || !W || !flags) return fail(PTF_ERR_INVALID, "NULL argument");
  auto v = step_slots(prob->g);
  if (slot < 0 || slot >= (int)v.size()) return fail(PTF_ERR_INVALID, "slot out of range");
  *R = v[slot]->R;
  *W = v[slot]->W;
  *flags = (v[slot]->occ ? PTF_VARIANT_OCC3 : 0) | (v[slot]->par ? PTF_VARIANT_PAR : 0) |
           (v[slot]->wg ? PTF_VARIANT_WG_LINES : 0);
  return PTF_OK;
}

// ----- ensembles ---------------------------------------------------------------
int ptf_ensemble_create(const ptf_desc* descs, int32_t count, ptf_ensemble** out) {
  if (!out || !descs || count <= 0) return fail(PTF_ERR_INVALID, "bad arguments");
  *out = nullptr;
  for (int i = 0; i < count; ++i) {
    CHK(validate_desc(&descs[i]));
    if (descs[i].flow_kind == PTF_FLOW_SPECTRAL || descs[i].flow_kind == PTF_FLOW_ANALYTIC)
      return fail(PTF_ERR_UNSUPPORTED, "ensemble members take steady or per-step flows");
  }
  std::unique_ptr<ptf_ensemble> e(new ptf_ensemble());
  e->where.resize(count);
  // group members with identical grid, precision, stepper (and filter), dt and nkappah
  std::vector<std::vector<int>> groups;
  for (int i = 0; i < count; ++i) {
    const ptf_desc& a = descs[i];
    int gi = -1;
    for (size_t k = 0; k < groups.size(); ++k) {
      const ptf_desc& b = descs[groups[k][0]];
      bool same = a.ndim == b.ndim && a.precision == b.precision && a.stepper == b.stepper &&
                  a.dt == b.dt && a.nkappah == b.nkappah && a.filter_order == b.filter_order &&
                  a.filter_innerK == b.filter_innerK && a.filter_outerK == b.filter_outerK &&
                  a.filter_tol == b.filter_tol;
      for (int x = 0; x < a.ndim && same; ++x) same = a.n[x] == b.n[x] && a.L[x] == b.L[x];
      if (same) {
        gi = (int)k;
        break;
      }
    }
    if (gi < 0) {
      groups.push_back({});
      gi = (int)groups.size() - 1;
    }
    e->where[i] = {gi, (int)groups[gi].size()};
    groups[gi].push_back(i);
  }
  for (auto& members : groups) {
    std::vector<ptf_desc> ds;
    for (int i : members) ds.push_back(descs[i]);
    std::unique_ptr<Group> g(new Group());
    CHK(group_init(*g, ds.data(), (int)ds.size()));
    e->groups.push_back(std::move(g));
    e->gt.push_back(0.0);
  }
  HIPCHK(hipStreamCreate(&e->stream));
  e->own_stream = true;
  for (size_t k = 0; k < e->groups.size(); ++k) {
    hipStream_t s;
    HIPCHK(hipStreamCreate(&s));
    e->gstreams.push_back(s);
    hipEvent_t ev;
    HIPCHK(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
    e->events.push_back(ev);
  }
  hipEvent_t ev;
  HIPCHK(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
  e->events.push_back(ev);
  *out = e.release();
  return PTF_OK;
}

int ptf_ensemble_destroy(ptf_ensemble* ens) {
  if (ens && ens->stream) (void)hipStreamSynchronize(ens->stream);
  delete ens;
  return PTF_OK;
}

static void drop_graph(ptf_ensemble* e) {
  for (int i = 0; i < 2; ++i) {
    if (e->exec[i]) (void)hipGraphExecDestroy(e->exec[i]);
    if (e->graph[i]) (void)hipGraphDestroy(e->graph[i]);
    e->exec[i] = nullptr;
    e->graph[i] = nullptr;
  }
}

int ptf_ensemble_set_stream(ptf_ensemble* ens, void* s) {
  if (!ens) return fail(PTF_ERR_INVALID, "ensemble is NULL");
  if (ens->own_stream && ens->stream) {
    (void)hipStreamSynchronize(ens->stream);
    (void)hipStreamDestroy(ens->stream);
  }
  ens->own_stream = false;
  ens->stream = (hipStream_t)s;
  drop_graph(ens);
  return PTF_OK;
}

static int member_of(ptf_ensemble* e, int member, Group** g, int* idx) {
  if (!e) return fail(PTF_ERR_INVALID, "ensemble is NULL");
  if (member < 0 || member >= (int)e->where.size()) return fail(PTF_ERR_INVALID, "member out of range");
  *g = e->groups[e->where[member].first].get();
  *idx = e->where[member].second;
  return PTF_OK;
}

int ptf_ensemble_set_flow(ptf_ensemble* ens, int32_t member, const double* u, const double* v) {
  Group* g;
  int i;
  CHK(member_of(ens, member, &g, &i));
  const double* f[3] = {u, v, nullptr};
  for (int c = 0; c < g->ndim && c < 2; ++c) {
    if (!f[c]) return fail(PTF_ERR_INVALID, "NULL flow component");
    char* dst = (char*)g->flow[c].p + (size_t)i * g->nR * g->esz;
    CHK(put_public(*g, g->dbl, dst, nullptr, f[c], g->nx / 2, 2, ens->stream));
  }
  HIPCHK(hipStreamSynchronize(ens->stream));
  return PTF_OK;
}

int ptf_ensemble_set_c(ptf_ensemble* ens, int32_t member, const double* c) {
  Group* g;
  int i;
  CHK(member_of(ens, member, &g, &i));
  if (!c) return fail(PTF_ERR_INVALID, "c is NULL");
  CHK(put_public(*g, g->dbl, g->phys.p, nullptr, c, g->nx / 2, 2, ens->stream));
  char* dst = (char*)g->sol.p + (size_t)i * g->nS * 2 * g->esz;
  CHK(g_rfft(*g, g->phys.p, dst, ens->stream));
  CHK(clear_member_lo(*g, i, ens->stream));
  g->primed = false;
  HIPCHK(hipStreamSynchronize(ens->stream));
  return PTF_OK;
}

int ptf_ensemble_get_c(ptf_ensemble* ens, int32_t member, double* c) {
  Group* g;
  int i;
  CHK(member_of(ens, member, &g, &i));
  if (!c) return fail(PTF_ERR_INVALID, "c is NULL");
  char* src = (char*)g->sol.p + (size_t)i * g->nS * 2 * g->esz;
  CHK(g_irfft(*g, src, g->phys.p, ens->stream));
  return get_public(*g, g->dbl, c, g->phys.p, nullptr, g->nx / 2, 2, ens->stream);
}

int ptf_ensemble_get_sol(ptf_ensemble* ens, int32_t member, double* sol) {
  Group* g;
  int i;
  CHK(member_of(ens, member, &g, &i));
  if (!sol) return fail(PTF_ERR_INVALID, "sol is NULL");
  return get_member_sol(*g, i, sol, ens->stream);
}

// nsteps steps of every group: each group's launches go to its own stream, forked from
// and joined to the ensemble's stream.  The groups share no buffers, so with nsteps > 1 a
// group with less work per step runs ahead instead of waiting for the slowest group at
// every step.
static int ensemble_record_steps(ptf_ensemble* e, int nsteps) {
  hipStream_t s0 = e->stream;
  const int ng = (int)e->groups.size();
  hipEvent_t fork = e->events[ng];
  HIPCHK(hipEventRecord(fork, s0));
  for (int k = 0; k < ng; ++k) {
    HIPCHK(hipStreamWaitEvent(e->gstreams[k], fork, 0));
    for (int i = 0; i < nsteps; ++i) CHK(group_step_launches(*e->groups[k], e->gstreams[k], 0.0));
    HIPCHK(hipEventRecord(e->events[k], e->gstreams[k]));
  }
  for (int k = 0; k < ng; ++k) HIPCHK(hipStreamWaitEvent(s0, e->events[k], 0));
  return PTF_OK;
}

// Steps per replay of the multi-step ensemble graph (the groups join once per block).
static constexpr int kEnsGraphSteps = 16;

// capture graph `which` (0: one step, 1: kEnsGraphSteps steps) if it is not there yet
static int ensemble_graph(ptf_ensemble* ens, int which) {
  if (ens->exec[which]) return PTF_OK;
  HIPCHK(hipStreamBeginCapture(ens->stream, hipStreamCaptureModeThreadLocal));
  int rc = ensemble_record_steps(ens, which ? kEnsGraphSteps : 1);
  hipGraph_t graph = nullptr;
  hipError_t ce = hipStreamEndCapture(ens->stream, &graph);
  if (rc != PTF_OK) {
    if (graph) (void)hipGraphDestroy(graph);
    return rc;
  }
  HIPCHK(ce);
  ens->graph[which] = graph;
  if (hipGraphInstantiate(&ens->exec[which], graph, nullptr, nullptr, 0) != hipSuccess) {
    ens->exec[which] = nullptr;
    return fail(PTF_ERR_DEVICE, "hipGraphInstantiate failed");
  }
  return PTF_OK;
}

static void ensemble_advance_clock(ptf_ensemble* e, int64_t nsteps) {
  for (size_t k = 0; k < e->groups.size(); ++k)
    for (int64_t i = 0; i < nsteps; ++i) e->gt[k] += e->groups[k]->dt;
  e->step += nsteps;
}

static int ensemble_step_impl(ptf_ensemble* ens, int64_t nsteps, int32_t use_graph);

int ptf_ensemble_step(ptf_ensemble* ens, int64_t nsteps, int32_t use_graph) {
  if (!ens) return fail(PTF_ERR_INVALID, "ensemble is NULL");
  if (nsteps < 0) return fail(PTF_ERR_INVALID, "nsteps must be >= 0");
  CHK(ensemble_step_impl(ens, nsteps, use_graph));
  ensemble_advance_clock(ens, nsteps);
  return PTF_OK;
}

int ptf_ensemble_get_clock(ptf_ensemble* ens, int32_t member, double* t, int64_t* step) {
  Group* g;
  int i;
  CHK(member_of(ens, member, &g, &i));
  if (!t || !step) return fail(PTF_ERR_INVALID, "NULL argument");
  *t = ens->gt[ens->where[member].first];
  *step = ens->step;
  return PTF_OK;
}

int ptf_ensemble_set_clock(ptf_ensemble* ens, double t, int64_t step) {
  if (!ens) return fail(PTF_ERR_INVALID, "ensemble is NULL");
  for (double& x : ens->gt) x = t;
  ens->step = step;
  return PTF_OK;
}

int ptf_ensemble_diagnostics(ptf_ensemble* ens, int32_t member, double* mass, double* variance) {
  Group* g;
  int i;
  CHK(member_of(ens, member, &g, &i));
  if (!g->red.p) CHK(dalloc(g->red, 2 * 1024 * sizeof(double)));
  return diagnostics_group(*g, i, ens->stream, mass, variance);
}

static int ensemble_step_impl(ptf_ensemble* ens, int64_t nsteps, int32_t use_graph) {
  for (auto& g : ens->groups) CHK(group_prime(*g, ens->stream));
  if (!use_graph) {
    for (int64_t i = 0; i < nsteps; ++i) CHK(ensemble_record_steps(ens, 1));
    HIPCHK(hipGetLastError());
    return PTF_OK;
  }
  // blocks of kEnsGraphSteps steps, then single steps
  const int64_t blocks = nsteps / kEnsGraphSteps, rest = nsteps % kEnsGraphSteps;
  if (blocks > 0) CHK(ensemble_graph(ens, 1));
  if (rest > 0) CHK(ensemble_graph(ens, 0));
  for (int64_t i = 0; i < blocks; ++i) HIPCHK(hipGraphLaunch(ens->exec[1], ens->stream));
  for (int64_t i = 0; i < rest; ++i) HIPCHK(hipGraphLaunch(ens->exec[0], ens->stream));
  return PTF_OK;
}

int ptf_ensemble_autotune(ptf_ensemble* ens) {
  if (!ens) return fail(PTF_ERR_INVALID, "ensemble is NULL");
  for (auto& g : ens->groups) CHK(autotune_group(*g, ens->stream));
  drop_graph(ens);
  return PTF_OK;
}

static std::vector<std::pair<Group*, Slot*>> ens_slots(ptf_ensemble* e) {
  std::vector<std::pair<Group*, Slot*>> v;
  for (auto& g : e->groups)
    for (Slot* s : step_slots(*g)) v.push_back({g.get(), s});
  return v;
}

int ptf_ensemble_num_slots(ptf_ensemble* ens, int32_t* n) {
  if (!ens || !n) return fail(PTF_ERR_INVALID, "NULL argument");
  *n = (int32_t)ens_slots(ens).size();
  return PTF_OK;
}

int ptf_ensemble_slot_name(ptf_ensemble* ens, int32_t slot, char* buf, int32_t buflen) {
  if (!ens) return fail(PTF_ERR_INVALID, "ensemble is NULL");
  auto v = ens_slots(ens);
  if (slot < 0 || slot >= (int)v.size()) return fail(PTF_ERR_INVALID, "slot out of range");
  return slot_name_into(*v[slot].second, buf, buflen);
}

int ptf_ensemble_time_slot(ptf_ensemble* ens, int32_t slot, int32_t reps, double* avg_ms) {
  if (!ens || !avg_ms) return fail(PTF_ERR_INVALID, "NULL argument");
  auto v = ens_slots(ens);
  if (slot < 0 || slot >= (int)v.size()) return fail(PTF_ERR_INVALID, "slot out of range");
  return time_slot(*v[slot].first, *v[slot].second, reps, ens->stream, avg_ms);
}

int ptf_ensemble_set_variant(ptf_ensemble* ens, int32_t slot, int32_t R, int32_t W, int32_t flags) {
  if (!ens) return fail(PTF_ERR_INVALID, "ensemble is NULL");
  auto v = ens_slots(ens);
  if (slot < 0 || slot >= (int)v.size()) return fail(PTF_ERR_INVALID, "slot out of range");
  if (flags & ~(PTF_VARIANT_OCC3 | PTF_VARIANT_PAR | PTF_VARIANT_WG_LINES))
    return fail(PTF_ERR_INVALID, "unknown flags");
  Slot trial = *v[slot].second;
  if (!choose_variant(trial, R, W, (flags & PTF_VARIANT_OCC3) ? 3 : 0,
                      (flags & PTF_VARIANT_PAR) ? 1 : 0, (flags & PTF_VARIANT_WG_LINES) ? 1 : 0))
    return fail(PTF_ERR_UNSUPPORTED, "no compiled variant R=" + std::to_string(R) +
                                         " W=" + std::to_string(W) + " flags=" +
                                         std::to_string(flags) + " for slot " + trial.name);
  *v[slot].second = trial;
  drop_graph(ens);
  return PTF_OK;
}

int ptf_ensemble_get_variant(ptf_ensemble* ens, int32_t slot, int32_t* R, int32_t* W,
                             int32_t* flags) {
  if (!ens || !R || !W || !flags) return fail(PTF_ERR_INVALID, "NULL argument");
  auto v = ens_slots(ens);
  if (slot < 0 || slot >= (int)v.size()) return fail(PTF_ERR_INVALID, "slot out of range");
  const Slot& s = *v[slot].second;
  *R = s.R;
  *W = s.W;
  *flags = (s.occ ? PTF_VARIANT_OCC3 : 0) | (s.par ? PTF_VARIANT_PAR : 0) |
           (s.wg ? PTF_VARIANT_WG_LINES : 0);
  return PTF_OK;
}

}  // extern "C"
