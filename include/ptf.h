/*
 * libptf_hip -- C ABI of the MI355X TracerAdvectionDiffusion time step.
 *
 * Every entry point returns an int status (PTF_OK == 0); ptf_last_error()
 * returns the message of the last failure on the calling thread.  No C++
 * exception crosses this boundary.  Handles are not thread-safe; each
 * problem, plan and ensemble owns one HIP stream (or uses the one given with
 * *_set_stream) and every call is ordered on it.
 *
 * Layouts are the reference's (Julia column-major, x fastest):
 *   physical  R-field : nx*ny*nz reals
 *   spectral  S-field : (nx/2+1)*ny*nz complex, interleaved re/im
 * Host-side arrays crossing the ABI are always FP64 (double / interleaved
 * double pairs), whatever the precision the problem computes in.
 *
 * Reference interfaces replaced (PassiveTracerFlows.jl v0.9.1,
 * src/traceradvectiondiffusion.jl abbreviated S; FourierFlows.jl 0.10.x [FF],
 * not vendored in the reference, cited through its call sites):
 */
#ifndef PTF_H
#define PTF_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  PTF_OK = 0,
  PTF_ERR_INVALID = 1,     /* bad argument: null handle, unsupported size, ... */
  PTF_ERR_DEVICE = 2,      /* HIP runtime failure                              */
  PTF_ERR_NOMEM = 3,       /* device allocation failed                         */
  PTF_ERR_UNSUPPORTED = 4  /* valid request outside what this build implements */
};

enum { PTF_FP64 = 0, PTF_MIXED = 1 };            /* arithmetic of the bulk fields   */
enum { PTF_RK4 = 0, PTF_ETDRK4 = 1,                 /* stepper="RK4" / "ETDRK4"        */
       PTF_FILTERED_RK4 = 2, PTF_FILTERED_ETDRK4 = 3 }; /* "FilteredRK4" / "FilteredETDRK4" */
enum {
  PTF_FLOW_STEADY = 0,    /* u,v,w sampled once on the grid (S:426-452)              */
  PTF_FLOW_PER_STEP = 1,  /* u,v,w re-uploaded by the caller at every step boundary, */
                          /* i.e. evaluated at clock.t (S:701, S:718, S:737)         */
  PTF_FLOW_SPECTRAL = 2,  /* built-in: psi_hat(t) = psi0_hat exp(-i omega t),        */
                          /* u = -d_y psi, v = d_x psi, synthesised on the GPU at     */
                          /* clock.t once per step (2D only)                         */
  PTF_FLOW_ANALYTIC = 3   /* built-in: sums of separable terms (ptf_flow_term),     */
                          /* evaluated on the GPU at clock.t once per step (at set   */
                          /* time only when no term depends on t)                    */
};

/* One separable term of an analytic flow, added to velocity component
 * `component` (0 = u, 1 = v, 2 = w):
 *   amp * T(omega t + tphase) * X(kx x + px) * Y(ky y + py) * Z(kz z + pz),
 * with X, Y, Z selected by fx, fy, fz and T by time_fn among
 * PTF_FN_ONE (1), PTF_FN_SIN, PTF_FN_COS.  The factors are multiplied in the
 * order ((amp T) X) Y) Z and the terms of a component summed in array order.
 * Examples: the cellular flow psi = (U/k) sin kx sin ky is
 * u = -U sin(kx) cos(ky), v = U cos(kx) sin(ky); the ABC flow is
 * u = A sin z + C cos y, v = B sin x + A cos z, w = C sin y + B cos x. */
enum { PTF_FN_ONE = 0, PTF_FN_SIN = 1, PTF_FN_COS = 2 };
typedef struct ptf_flow_term {
  int32_t component;
  int32_t fx, fy, fz;
  double kx, ky, kz;
  double px, py, pz;
  double amp;
  int32_t time_fn;
  int32_t reserved;
  double omega, tphase;
} ptf_flow_term;
#define PTF_MAX_FLOW_TERMS 16

typedef struct ptf_desc {
  int32_t ndim;          /* 1, 2 or 3                                                */
  int32_t precision;     /* PTF_FP64 or PTF_MIXED                                    */
  int64_t n[3];          /* nx, ny, nz (unused trailing entries ignored)             */
  double L[3];           /* Lx, Ly, Lz                                               */
  double kappa;          /* x-diffusivity kappa                                      */
  double eta;            /* y-diffusivity eta (reference default eta = kappa)        */
  double iota;           /* z-diffusivity iota (reference default iota = kappa)      */
  double kappah;         /* hyperdiffusivity kappa_h                                 */
  int32_t nkappah;       /* hyperdiffusivity order n_kappa_h (Krsq^n)                */
  int32_t stepper;       /* PTF_RK4, PTF_ETDRK4, PTF_FILTERED_RK4, PTF_FILTERED_ETDRK4 */
  double dt;             /* time step                                                */
  int32_t flow_kind;     /* PTF_FLOW_*                                               */
  /* Filtered steppers ([FF] makefilter(grid; order, innerK, outerK, tol)):
   * filter(K) = exp(-decay (K - innerK)^order) for K >= innerK, else 1, with
   * K = |(kr dx, l dy, m dz)| / pi and decay = -log(tol) / (outerK - innerK)^order;
   * sol is multiplied by it after every step's update.  filter_order = 0 selects
   * the defaults order 4, innerK 0.65, outerK 1, tol 1e-15. */
  int32_t filter_order;
  double filter_innerK;
  double filter_outerK;
  double filter_tol;
} ptf_desc;

typedef struct ptf_problem ptf_problem;
typedef struct ptf_plan ptf_plan;
typedef struct ptf_ensemble ptf_ensemble;

/* ---- library / device ---------------------------------------------------- */
const char* ptf_last_error(void);
int ptf_version(int32_t* major, int32_t* minor);
/* Select the device for the calling thread ([FF] Device / GPU() at S:505, S:646). */
int ptf_device_init(int32_t ordinal);
int ptf_device_synchronize(void);

/* ---- device buffers (the arrays of the ROCm device; [FF] zeros(dev, T, dims),
 *      @devzeros at S:505, S:646 and device_array(dev)(c) at S:847) ---------- */
int ptf_buffer_alloc(uint64_t nbytes, void** dptr);
int ptf_buffer_free(void* dptr);
int ptf_buffer_upload(void* dptr, const void* host, uint64_t nbytes);
int ptf_buffer_download(void* host, const void* dptr, uint64_t nbytes);

/* ---- FFT plans: grid.rfftplan with mul!/ldiv! (S:20, S:698-783, S:818, S:847) */
/* ndim/n as the grid; precision PTF_FP64 (double) or PTF_MIXED (float).        */
int ptf_plan_create(int32_t ndim, const int64_t n[3], int32_t precision, ptf_plan** out);
/* A plan over `batch` consecutive fields of the same grid, transformed
 * independently in one launch per pass: the layered (nx, ny, nlayers)
 * transforms over dims 1:2 of the MQG-coupled tracer (Vars at S:678-679,
 * fwdtransform! at S:865-866) are ndim = 2, batch = nlayers.  Member b of the
 * real input/output starts at b*nx*ny*nz reals, of the spectral one at
 * b*(nx/2+1)*ny*nz complex values; ptf_rfft/ptf_irfft/ptf_dealias act on all
 * members. */
int ptf_plan_create_batched(int32_t ndim, const int64_t n[3], int64_t batch, int32_t precision,
                            ptf_plan** out);
int ptf_plan_destroy(ptf_plan* plan);
int ptf_plan_set_stream(ptf_plan* plan, void* hip_stream);
/* mul!(out, rfftplan, in): unnormalised R2C; device pointers; input preserved. */
int ptf_rfft(ptf_plan* plan, const void* in_real, void* out_complex);
/* ldiv!(out, rfftplan, in): C2R scaled by 1/(nx*ny*nz); input preserved (a
 * superset of FFTW's destroy-input contract); only the real parts of the
 * kx = 0 and kx = nx/2 entries are used after the y/z inverse passes.        */
int ptf_irfft(ptf_plan* plan, const void* in_complex, void* out_real);
/* ptf_irfft without any device allocation by the library (no counterpart in the
 * reference: for callers that cannot allocate, such as code under HIP graph capture,
 * serving loops and harnesses that forbid hipMalloc).  The intermediate spectrum of a
 * 2D/3D inverse goes into a static 64 MiB device array of the library (one per device)
 * instead of the plan's scratch buffer: PTF_ERR_INVALID when the plan's batch of half
 * spectra needs more (e.g. 3D 256^3 FP64).  Results are identical to ptf_irfft.  Every
 * call on a device shares that device's array and it keeps the last call's contents:
 * calls that overlap on one device (other streams, other threads) race. */
int ptf_irfft_noalloc(ptf_plan* plan, const void* in_complex, void* out_real);
/* dealias!(fh, grid) ([FF], used by the MQG-coupled path S:225-250): zero the
 * aliased wavenumbers of a device half spectrum of the plan's grid, kr indices
 * from floor((1-f)/2 nx) (0-based) to the end, l and m indices in
 * [floor((1-f)/2 n), ceil((1+f)/2 n)), f = aliased_fraction (1/3: the 2/3 rule);
 * f <= 0 leaves fh unchanged.  Asynchronous on the plan's stream. */
int ptf_dealias(ptf_plan* plan, void* spec, double aliased_fraction);

/* ---- problems: TracerAdvectionDiffusion.Problem (S:143-216) + the
 *      [FF] FourierFlows.Problem/Clock/RK4TimeStepper/ETDRK4TimeStepper ---- */
int ptf_problem_create(const ptf_desc* desc, ptf_problem** out);
int ptf_problem_destroy(ptf_problem* prob);
int ptf_problem_set_stream(ptf_problem* prob, void* hip_stream);
int ptf_problem_get_stream(ptf_problem* prob, void** hip_stream);
/* Steady fields (S:426-452) or the per-step refresh of a time-varying flow.
 * Host FP64 R-fields; v, w may be NULL beyond ndim. */
int ptf_set_flow(ptf_problem* prob, const double* u, const double* v, const double* w);
/* PTF_FLOW_SPECTRAL: psi0 (interleaved complex S-field) and omega (real S-field). */
int ptf_set_spectral_flow(ptf_problem* prob, const double* psih0, const double* omega);
/* PTF_FLOW_ANALYTIC: the flow as nterms (<= PTF_MAX_FLOW_TERMS) separable terms;
 * the x, y, z factors are tabulated on the host in FP64 once (the grid points of
 * the problem), the time factors at clock.t before every step.  This replaces
 * the per-step host evaluation and upload of a time-varying user flow (the
 * u(x, y, t) closures of S:701, S:718, S:737) for flows of this form. */
int ptf_set_analytic_flow(ptf_problem* prob, const ptf_flow_term* terms, int32_t nterms);
/* set_c!(prob, c) (S:844-852, S:872): sol = R2C(c); host FP64 R-field. */
int ptf_set_c(ptf_problem* prob, const double* c);
/* stepforward!(prob, nsteps) [FF]: clock.t += dt, clock.step += 1 after each step. */
int ptf_step(ptf_problem* prob, int64_t nsteps);
/* With a steady flow, ptf_step replays whole blocks of 8 steps from a captured
 * hipGraph (kernels only; the captured arguments do not depend on the clock);
 * enable = 0 launches every kernel directly.  Default: enabled. */
int ptf_problem_set_graph(ptf_problem* prob, int32_t enable);
int ptf_get_clock(ptf_problem* prob, double* t, int64_t* step);
int ptf_set_clock(ptf_problem* prob, double t, int64_t step);
/* prob.sol read/write (host interleaved FP64 S-field). */
int ptf_get_sol(ptf_problem* prob, double* sol);
int ptf_set_sol(ptf_problem* prob, const double* sol);
/* updatevars!(prob) (S:815-821, S:837): c = C2R(sol) into a host FP64 R-field. */
int ptf_updatevars(ptf_problem* prob, double* c);
/* calcN!(N, sol, t, clock, vars, params, grid) (S:695-786) on host arrays: the
 * operator plug-in evaluated once with the current flow (FP64 host S-fields). */
int ptf_calcN(ptf_problem* prob, const double* sol, double* N);
/* The diagonal linear operator L of Equation(params, grid) (S:502-554), host
 * real S-shaped array. */
int ptf_get_L(ptf_problem* prob, double* L);
/* FP64 diagnostics: mass = sum(c) = Re sol[0]; variance = mean((c - mean c)^2)
 * by a compensated Parseval sum over the half spectrum. */
int ptf_diagnostics(ptf_problem* prob, double* mass, double* variance);
/* stepforward!(prob, diags, nsteps) with the diagnostics kept on the GPU
 * ([FF] Diagnostic/increment!, used with Output/saveoutput as in
 * examples/onedim_gaussiandiffusion.jl:66-95): after every step whose new
 * clock.step is a multiple of `every`, record t, step, mass and variance (as
 * ptf_diagnostics) and, when c_out != NULL, the physical field c (as
 * ptf_updatevars) into device memory; everything is copied to the host once,
 * after the steps.  *nrec_out = number of records; capacity = entries the
 * output arrays hold (nrec > capacity is an error, raised before any step).
 * Any output pointer may be NULL; c_out holds nrec R-fields (FP64). */
int ptf_step_record(ptf_problem* prob, int64_t nsteps, int64_t every, int64_t capacity,
                    int64_t* nrec_out, double* t_out, int64_t* step_out, double* mass_out,
                    double* var_out, double* c_out);
/* Device pointer of sol: the S-field in the problem's precision and working
 * layout (when the last axis has >= 64 points it is stored in bricks of 8
 * planes, element (kx, y, z) at ((z>>3) ny + y)(nx/2+1) 8 + 8 kx + (z&7); use
 * ptf_get_sol for the reference layout).  For in-HBM consumers of the state. */
int ptf_problem_sol_ptr(ptf_problem* prob, void** dptr);
/* Pick the fastest compiled kernel variant for every launch slot. */
int ptf_autotune(ptf_problem* prob);
/* Kernel-slot introspection for measurement: number of launches per step,
 * name of slot i, and the average duration of slot i over reps launches
 * (HIP events on the problem's stream).  Timing launches write only scratch
 * fields: the state is unchanged. */
int ptf_problem_num_slots(ptf_problem* prob, int32_t* nslots);
int ptf_problem_slot_name(ptf_problem* prob, int32_t slot, char* buf, int32_t buflen);
int ptf_problem_time_slot(ptf_problem* prob, int32_t slot, int32_t reps, double* avg_ms);
/* Kernel variant of slot i: radix R, width W (columns or rows per block) and
 * flags (PTF_VARIANT_*).  set_variant selects any compiled variant (as the
 * autotuner does); PTF_ERR_UNSUPPORTED when that variant was not compiled.
 * OCC3: a 3-waves/SIMD register budget; PAR: the input-parallel advection form
 * (row slots) or the persistent pipelined stage kernel (fused column slots);
 * WG_LINES: a column line along the bricked axis spread over the workgroup with
 * workgroup barriers, where by default it is held by one wavefront (accepted only
 * for the slots and radices where the per-wavefront form applies). */
enum { PTF_VARIANT_OCC3 = 1, PTF_VARIANT_PAR = 2, PTF_VARIANT_WG_LINES = 4 };
int ptf_problem_set_variant(ptf_problem* prob, int32_t slot, int32_t R, int32_t W, int32_t flags);
int ptf_problem_get_variant(ptf_problem* prob, int32_t slot, int32_t* R, int32_t* W, int32_t* flags);
/* ptf_step with a HIP event recorded around every launch of every step:
 * slot_ms[i] (num_slots entries) receives the average duration of slot i over
 * the nsteps steps, measured inside the running step sequence.  The state
 * advances exactly as with ptf_step. */
int ptf_step_profiled(ptf_problem* prob, int64_t nsteps, double* slot_ms);

/* ---- ensembles: many independent 2D problems, one graph-captured launch
 *      sequence per step (no reference counterpart: a batch of
 *      TracerAdvectionDiffusion.Problem(dev, TwoDAdvectingFlow(...)), S:165-188,
 *      each member with its own grid size, diffusivities and flow; RK4, ETDRK4 or
 *      their filtered forms; steady flows, or per-step flows the caller refreshes
 *      with ptf_ensemble_set_flow at the member's clock.t) ------------------------ */
int ptf_ensemble_create(const ptf_desc* descs, int32_t count, ptf_ensemble** out);
int ptf_ensemble_destroy(ptf_ensemble* ens);
int ptf_ensemble_set_stream(ptf_ensemble* ens, void* hip_stream);
int ptf_ensemble_set_flow(ptf_ensemble* ens, int32_t member, const double* u, const double* v);
int ptf_ensemble_set_c(ptf_ensemble* ens, int32_t member, const double* c);
int ptf_ensemble_get_c(ptf_ensemble* ens, int32_t member, double* c);
int ptf_ensemble_get_sol(ptf_ensemble* ens, int32_t member, double* sol);
/* Advance every member nsteps steps; use_graph != 0 replays captured hipGraphs (blocks
 * of 16 steps in which each size group runs ahead on its own, then single steps). */
int ptf_ensemble_step(ptf_ensemble* ens, int64_t nsteps, int32_t use_graph);
/* A member's clock (every member steps with the ensemble; t advances by that member's dt
 * after each step, as [FF] stepforward! advances prob.clock), and a shared reset. */
int ptf_ensemble_get_clock(ptf_ensemble* ens, int32_t member, double* t, int64_t* step);
int ptf_ensemble_set_clock(ptf_ensemble* ens, double t, int64_t step);
/* FP64 mass and variance of one member (as ptf_diagnostics for a problem). */
int ptf_ensemble_diagnostics(ptf_ensemble* ens, int32_t member, double* mass, double* variance);
int ptf_ensemble_autotune(ptf_ensemble* ens);
int ptf_ensemble_num_slots(ptf_ensemble* ens, int32_t* nslots);
int ptf_ensemble_slot_name(ptf_ensemble* ens, int32_t slot, char* buf, int32_t buflen);
int ptf_ensemble_time_slot(ptf_ensemble* ens, int32_t slot, int32_t reps, double* avg_ms);
/* The ensemble's kernel variant of launch slot `slot` (as ptf_problem_set_variant /
 * ptf_problem_get_variant): lets a caller store the autotuner's choice and apply it in
 * another process without timing anything. */
int ptf_ensemble_set_variant(ptf_ensemble* ens, int32_t slot, int32_t R, int32_t W, int32_t flags);
int ptf_ensemble_get_variant(ptf_ensemble* ens, int32_t slot, int32_t* R, int32_t* W, int32_t* flags);

#ifdef __cplusplus
}
#endif
#endif /* PTF_H */
