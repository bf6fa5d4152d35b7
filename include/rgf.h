/*
 * rgf.h -- C ABI of librgf, an MI355X (gfx950) implementation of the
 * recursive Green's function (RGF) S-matrix solve of chiaweihsu/RGF.
 *
 * The reference has no FFI; its boundary is the MATLAB functions
 *   [smatrix, channels] = cal_smatrix_RGF(syst, out, in)     src/cal_smatrix_RGF.m:1
 *   [smatrix, channels] = cal_smatrix_RGF_3D(syst, out, in)  src/cal_smatrix_RGF_3D.m:1
 *   channels = setup_channels(ny, BC, k0dx, eps_L, eps_R)     src/setup_channels.m:1
 *   channels = setup_channels_3D(nx, ny, BC, k0dx, eps_L, eps_R)  src/setup_channels_3D.m:1
 *   A = build_laplacian_1d(n, BC)                             src/build_laplacian_1d.m:1
 * and the entry points below are what a MEX gateway
 * (rgf_amd/mex/cal_smatrix_RGF_mex.c) or a ctypes binding binds to reproduce them.
 *
 * Conventions
 *  - Plain C types only; inputs are borrowed for the call and never mutated;
 *    outputs are caller-allocated.  No exceptions cross this boundary: every
 *    function returns an rgf_status, and rgf_last_error() gives the message
 *    (the reference's error texts are reproduced where it has one).
 *  - Complex outputs are interleaved (re, im) doubles, matrices column-major,
 *    as MATLAB's mxGetComplexDoubles layout.
 *  - Channel indices are 1-based, as in the reference.
 *  - A handle owns one HIP stream and its device workspace; a handle is not
 *    safe for concurrent use from several threads.
 */
#ifndef RGF_H
#define RGF_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define RGF_ABI_VERSION 5  /* 2: rgf_options.refine_s_tol, rgf_info.refine_s_err_estimate;
                              3: rgf_info.max_multiplier;
                              4: rgf_options.autotune and factor_slots, rgf_info.block_size,
                                 autotuned, factor_slices_resident and factor_slices_rebuilt;
                              5: rgf_options.pivoting, rgf_info.pivoted */
#define RGF_ALL (-1) /* selection count meaning "all propagating channels 1..N_prop" */

typedef enum {
    RGF_OK = 0,
    RGF_ERR_INVALID_ARGUMENT = 1,   /* malformed syst / sizes / options          */
    RGF_ERR_BOUNDARY_CONDITION = 2, /* BC not eligible, complex Bloch phase      */
    RGF_ERR_CHANNEL_INDEX = 3,      /* in/out index outside 1..N_prop            */
    RGF_ERR_OUT_OF_MEMORY = 4,      /* device or host allocation failed          */
    RGF_ERR_DEVICE = 5,             /* HIP runtime error / no device             */
    RGF_ERR_NUMERICAL = 6,          /* zero or non-finite pivot in a slice inverse */
    RGF_ERR_NOT_CONVERGED = 7,      /* mixed precision refinement did not reach tolerance */
    RGF_ERR_BUFFER_TOO_SMALL = 8
} rgf_status;

/* Transverse boundary conditions (setup_channels.m:10-20). 'periodic' is
 * RGF_BC_BLOCH with phase 0. */
typedef enum {
    RGF_BC_BLOCH = 0,
    RGF_BC_DIRICHLET = 1,          /* 'Dirichlet' / 'PEC'                      */
    RGF_BC_NEUMANN = 2,            /* 'Neumann' / 'PMC'                        */
    RGF_BC_DIRICHLET_NEUMANN = 3,  /* 'DirichletNeumann' / 'PECPMC'            */
    RGF_BC_NEUMANN_DIRICHLET = 4   /* 'NeumannDirichlet' / 'PMCPEC'            */
} rgf_bc;

typedef enum {
    RGF_PRECISION_FP64 = 0,        /* complex FP64 throughout                   */
    RGF_PRECISION_FP32_REFINED = 1 /* complex FP32 slice inverses on MFMA, FP64
                                      residual iterative refinement              */
} rgf_precision;

/* The reference's `syst` struct (cal_smatrix_RGF.m:12-34; _3D.m:12-31). */
typedef struct {
    int32_t ndim;          /* 2 or 3                                                   */
    int64_t dims[3];       /* 2D: {ny, nx, 1}, eps is ny x nx, slices are columns;
                              3D: {nx, ny, nz}, eps is nx x ny x nz, slices along z  */
    const double* eps_re;  /* prod(dims) values, column-major                          */
    const double* eps_im;  /* may be NULL: real permittivity                           */
    double eps_L[2];       /* epsilon_L (re, im)                                       */
    double eps_R[2];       /* epsilon_R (re, im)                                       */
    double k0dx[2];        /* k0dx (re, im)                                            */
    int32_t bc[2];         /* rgf_bc; 2D uses bc[0] (y); 3D bc[0] = x, bc[1] = y        */
    double bloch_ka[2];    /* Bloch phase k*a for each transverse direction            */
    int32_t eps_on_device; /* 1: eps_re/eps_im are device pointers (FP64)              */
} rgf_problem;

/* The reference's `out`/`in` arguments (cal_smatrix_RGF.m:35-50). A count of
 * RGF_ALL selects every propagating channel of that side (cell {'L'} / {'R'});
 * a count of 0 selects none. Indices are 1-based; order and duplicates are kept. */
typedef struct {
    const int32_t* ind_in_L;  int32_t n_in_L;
    const int32_t* ind_in_R;  int32_t n_in_R;
    const int32_t* ind_out_L; int32_t n_out_L;
    const int32_t* ind_out_R; int32_t n_out_R;
} rgf_selection;

typedef struct {
    int32_t precision;        /* rgf_precision                                         */
    int32_t max_refine_iter;  /* FP32_REFINED: correction sweeps (default 6)           */
    double refine_tol;        /* FP32_REFINED: stop at max|r|/max|b| below (1e-13)     */
    int32_t block_size;       /* slice-inverse panel width, 0 = tuned default           */
    int32_t complex_3m;       /* FP32 GEMMs with 3 real products per complex product:
                                 bit 0 = in the slice inverses, bit 1 = in the
                                 refinement sweeps (default 3: both; the 3D N=4096
                                 workload reaches 1e-13 in 3 sweeps either way)     */
    int32_t profile;          /* 1: time the slice-inverse trailing-update kernel with
                                 HIP events on its stream, one launch in four (fills
                                 the kernel_* fields)                                  */
    int32_t single_stream;    /* 1: no look-ahead second stream in the slice inverse
                                 (same kernels, one stream; for counter collection)   */
    int32_t batch_concurrency;/* rgf_solve_batched, problems outside the lockstep
                                 engine: solves in flight at once, each on its own
                                 streams and workspace; 0 = 8                        */
    int32_t batch_engine;     /* rgf_solve_batched: 0 = mixed-precision problems up to
                                 1024 transverse sites run in the lockstep engine, its
                                 block steps replayed as hipGraphs; 1 = the same with
                                 plain launches; 2 = every problem on worker contexts */
    int32_t batch_graph_rounds;/* lockstep engine: block-step rounds per captured
                                 graph (0 = 16)                                      */
    int32_t autotune;         /* block_size 0: 1 (default) = the slice-inverse block width
                                 and stream use are timed once per process for each
                                 (device, N, precision, complex_3m bit 0,
                                 single_stream) on a synthetic slice matrix (FP32 64
                                 or 128, FP64 32 or 64; look-ahead streams or one
                                 stream) and the fastest kept; 0 = fixed (FP32 128,
                                 FP64 64, look-ahead unless single_stream)           */
    double refine_s_tol;      /* FP32_REFINED, > 0: also stop once the estimated max
                                 error of the S entries is below this (see
                                 rgf_info.refine_s_err_estimate); 0 = off (default)  */
    int32_t factor_slots;     /* FP32_REFINED: at most this many N x N FP32 slices of
                                 device memory for the slice inverses (0 = as many as
                                 free memory allows).  With fewer than L + 2, the
                                 first R inverses stay resident and the others are
                                 rebuilt in segments by each refinement sweep from
                                 stored checkpoints (memory O((R + L/C + C) N^2)
                                 instead of O(L N^2), more inverses computed)     */
    int32_t pivoting;         /* slice inverses: 0 (default) = blocked Gauss-Jordan with
                                 pivoting inside each diagonal block; a slice it
                                 refuses (zero pivot, or multipliers above 1e6 in
                                 FP64; in FP32+refine a refinement that stalls after
                                 multipliers above 1e3) has the solve run again with
                                 every slice inverted with partial pivoting over all
                                 rows (LAPACK getrf + getri, as MATLAB's inv; much
                                 slower); 1 = that from the start; 2 = no fallback
                                 (the refusal is returned)                         */
} rgf_options;

typedef struct {
    int64_t n_transverse;     /* N                                                     */
    int64_t n_slices;         /* L                                                     */
    int32_t n_prop_L, n_prop_R;
    int32_t n_out, n_in;
    int32_t swept_right_to_left; /* 1 when the slice loop ran from the right lead       */
    int32_t refine_iters;     /* correction sweeps performed                           */
    double refine_residual;   /* final max|b - H psi| / max|b|                          */
    double seconds_total, seconds_setup, seconds_slices, seconds_refine;
    double flops_slices;      /* algorithmic flops executed in the slice loop          */
    /* with opt.profile: the dominant kernel, over the launches timed (one in
       four): summed event-timed duration, number of timed launches, their
       executed flops.  rgf_solve: the rank-b trailing update of the slice
       inverse; the lockstep batched engine: the refinement sweep product
       z_n = g_n t over the whole batch (the same figures in every info of
       the batch) */
    double kernel_seconds;
    double kernel_flops;
    int64_t kernel_launches;
    /* with opt.refine_s_tol: estimated max error of the S entries at the
       stop, 4 q/(1-q) max|S_k - S_(k-1)| with q the ratio of the last two
       residuals (linear convergence, a factor 2 to spare); 0 otherwise */
    double refine_s_err_estimate;
    /* largest multiplier |(A[:, P] A_PP^-1)_ij| of the slice inverses (the
       growth the rows outside a diagonal block see; pivoting is confined to
       the block).  The FP64 paths fail with RGF_ERR_NUMERICAL above 1e6. */
    double max_multiplier;
    int32_t block_size;       /* slice-inverse block width used                         */
    int32_t autotuned;        /* 1: this call timed the block-width candidates         */
    /* FP32_REFINED: slice inverses kept in device memory (L + 2 when all fit;
       fewer when the rest were rebuilt by the sweeps, see factor_slots) and
       how many the sweeps rebuilt */
    int64_t factor_slices_resident;
    int64_t factor_slices_rebuilt;
    int32_t pivoted;          /* 1: the slice inverses used partial pivoting over all
                                 rows (rgf_options.pivoting)                          */
} rgf_info;

/* Lead channels (setup_channels.m:30-82; setup_channels_3D.m:30-64).
 * Every array has N = transverse size entries (caller-allocated); complex
 * arrays are interleaved with 2N doubles.  Sides share data when
 * epsilon_R == epsilon_L, as in the reference (setup_channels.m:166-167). */
typedef struct {
    double* kt1;              /* 2D: kydx; 3D: kxdx (column-major nx x ny)             */
    double* kt2;              /* 3D: kydx (column-major nx x ny); ignored in 2D          */
    double* kl_L;             /* 2D: L.kxdx; 3D: L.kzdx (complex)                        */
    double* kl_R;
    int32_t* ind_prop_L;      /* first n_prop_L entries used, 1-based                    */
    int32_t* ind_prop_R;
    double* sqrt_mu_L;        /* complex, first n_prop_L used                             */
    double* sqrt_mu_R;
    int32_t* ind_prop_conj_L; /* filled when has_conj (2D only)                          */
    int32_t* ind_prop_conj_R;
    int32_t n_prop_L, n_prop_R;
    int32_t has_conj;         /* 0 for 3D and for Bloch with ka != 0 (setup_channels.m:80-82) */
} rgf_channels;

typedef struct rgf_context* rgf_handle;

int32_t rgf_abi_version(void);
const char* rgf_status_string(rgf_status s);
void rgf_default_options(rgf_options* opt);

/* Handle lifetime.  device < 0 picks the current device. */
rgf_status rgf_create(rgf_handle* h, int32_t device);
rgf_status rgf_destroy(rgf_handle h);
const char* rgf_last_error(rgf_handle h);   /* NULL handle: last error of a handle-less call */
/* Use an external HIP stream (e.g. torch's current stream); NULL restores the own stream. */
rgf_status rgf_set_stream(rgf_handle h, void* hip_stream);
/* Release cached device workspace. */
rgf_status rgf_trim(rgf_handle h);

/* Host-only (no GPU): channel counts and the S shape for a problem and
 * selection; replaces the sizing part of cal_smatrix_RGF.m:97-184. */
rgf_status rgf_query(const rgf_problem* p, const rgf_selection* sel,
                     int32_t* n_prop_L, int32_t* n_prop_R, int32_t* n_out, int32_t* n_in);

/* Host-only: lead channels; replaces setup_channels.m / setup_channels_3D.m. */
rgf_status rgf_setup_channels(const rgf_problem* p, rgf_channels* out);

/* Host-only: transverse mode profiles at given wavenumbers, what
 * channels.fun_phi returns (setup_channels.m:121-156 for 2D,
 * setup_channels_3D.m:101-123 for 3D).  Reads p's ndim, transverse dims, bc
 * and bloch_ka (and checks p as rgf_setup_channels does).  phi_out: complex,
 * interleaved; 2D: dims[0] x count column-major, column c the profile at
 * k1[c]; 3D: (dims[0] * dims[1]) x count, column c the dims[0] x dims[1]
 * profile (x fast) at (k1[c], k2[c]). */
rgf_status rgf_transverse_profiles(const rgf_problem* p, int64_t count, const double* k1, const double* k2,
                                   double* phi_out);

/* Host-only: 1D Laplacian of build_laplacian_1d.m as a dense n x n complex
 * matrix (interleaved, column-major, 2 n^2 doubles). */
rgf_status rgf_build_laplacian_1d(int64_t n, int32_t bc, double ka, double* A_out);

/* The S-matrix solve: replaces cal_smatrix_RGF.m:1-291 (ndim 2) and
 * cal_smatrix_RGF_3D.m:1-306 (ndim 3).  S_out: n_out x n_in complex,
 * interleaved, column-major (sizes from rgf_query).  opt may be NULL. */
rgf_status rgf_solve(rgf_handle h, const rgf_problem* p, const rgf_selection* sel,
                     const rgf_options* opt, double* S_out, rgf_info* info);

/* rgf_solve plus the field profiles: the total field in every slice for each
 * selected input (no counterpart in the reference, whose recursion never
 * forms them).  fields_out: N x (L + 2) x n_in complex, interleaved,
 * column-major, caller-allocated; slice 0 is the left lead's slice next to
 * the region, slices 1..L are the reference's slices (epsilon's slice
 * order), slice L + 1 is the right lead's; inputs are ordered as S's
 * columns; transverse sites as in epsilon.  fields[:, n, a] = G(n, source_a)
 * B_a, so S = -2i C fields[:, boundary, :] - delta.  Requires precision
 * RGF_PRECISION_FP32_REFINED: the fields are the refined FP64 solution of the
 * whole slab, within refine_tol of the FP64 residual. */
rgf_status rgf_solve_fields(rgf_handle h, const rgf_problem* p, const rgf_selection* sel,
                            const rgf_options* opt, double* S_out, double* fields_out,
                            rgf_info* info);

/* Many independent solves (frequencies, Bloch phases, widths) in one call.
 * Mixed-precision problems with up to 1024 transverse sites (batch_engine 0
 * or 1) run together in the lockstep engine: every launch advances all of
 * them by one 64-wide block step of their slice inverses (or one slice of
 * their refinement sweeps), so small slices fill the GPU.  The others run
 * up to opt->batch_concurrency at once, each on its own worker context
 * (streams and device workspace, kept by the handle for later calls).
 * Each S equals rgf_solve's to its precision's tolerance.  On failure the
 * status of the first failing problem is returned and rgf_last_error names
 * it. */
rgf_status rgf_solve_batched(rgf_handle h, int32_t count, const rgf_problem* p,
                             const rgf_selection* sel, const rgf_options* opt,
                             double* const* S_out, rgf_info* info);

/* Transmission T = sum_ij |S_ij|^2 over the selected block (inputs on one
 * side, outputs on the other) and its gradient with respect to eps: the
 * adjoint of the slice recursion.  dT_deps_re[k] = dT/d Re(eps_k),
 * dT_deps_im[k] = dT/d Im(eps_k) (either may be NULL), column-major like eps. */
rgf_status rgf_transmission_gradient(rgf_handle h, const rgf_problem* p,
                                     const rgf_selection* sel, const rgf_options* opt,
                                     int32_t checkpoint_every, double* T_out,
                                     double* dT_deps_re, double* dT_deps_im, rgf_info* info);

/* Kernel-level entry points (used by the tests and the benchmark).
 * rgf_dense_inverse: in-place inverse of an n x n complex matrix
 * (interleaved, column-major) with the slice-inverse kernel of `precision`;
 * RGF_PRECISION_FP32_REFINED | RGF_DENSE_3M selects the FP32 kernels with
 * three real products per complex product (complex_3m bit 0).  A matrix the
 * blocked kernels refuse is inverted again with partial pivoting over all
 * rows (rgf_options.pivoting 0); block_size < 0 takes that path directly.
 * rgf_dense_gemm: C = A B, (m x k)(k x n), interleaved column-major. */
#define RGF_DENSE_3M 0x100
rgf_status rgf_dense_inverse(rgf_handle h, int32_t precision, int64_t n, double* A, int32_t block_size);

/* No counterpart in the reference.  rgf_dense_inverse for callers that cannot
 * allocate (code under HIP graph capture, tight serving loops): the same
 * slice-inverse kernels on a matrix already in device memory, with their
 * workspace in a static device array of RGF_NOALLOC_SCRATCH_BYTES per GPU
 * (held until the process ends); the call allocates nothing.
 *   A_re, A_im: the n x n matrix, planar (real and imaginary planes), row-major
 *     with row stride lda >= n elements, inverted in place; elements are
 *     double for RGF_PRECISION_FP64, float for RGF_PRECISION_FP32_REFINED
 *     (| RGF_DENSE_3M as in rgf_dense_inverse).
 *   status: device memory for two int32 words, or NULL.  Given, the call is
 *     asynchronous on `stream` and leaves bit 1 of [0] set on a zero or
 *     non-finite pivot, bit 2 if a hand-over between the work groups of one
 *     launch timed out (an internal error: the result is not valid), and in
 *     [1] the largest multiplier squared (float bits, see
 *     rgf_info.max_multiplier); NULL: the call waits for the inverse and
 *     returns RGF_ERR_NUMERICAL on a zero pivot (or, in FP64, a multiplier
 *     above 1e6), RGF_ERR_DEVICE on a hand-over time-out, as
 *     rgf_dense_inverse does.
 *   stream: a HIP stream (hipStream_t), or NULL for the handle's stream.
 * n beyond the workspace (about 8160 at the default block sizes) is refused
 * with RGF_ERR_INVALID_ARGUMENT.  Every call on one GPU shares that GPU's
 * array: calls that overlap on a GPU (from any handle) race. */
#define RGF_NOALLOC_SCRATCH_BYTES (32u << 20)
rgf_status rgf_dense_inverse_noalloc(rgf_handle h, int32_t precision, int64_t n, void* A_re, void* A_im, int64_t lda,
                                     int32_t block_size, int32_t* status, void* stream);
rgf_status rgf_dense_gemm(rgf_handle h, int32_t precision, int64_t m, int64_t n, int64_t k,
                          const double* A, const double* B, double* C);

#ifdef __cplusplus
}
#endif
#endif /* RGF_H */
