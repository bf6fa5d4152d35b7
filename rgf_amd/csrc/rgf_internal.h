// Internal declarations shared by the host driver and the gfx950 kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <complex>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/rgf.h"

namespace rgf {

using cd = std::complex<double>;

struct Error : std::runtime_error {
    rgf_status status;
    Error(rgf_status s, const std::string& m) : std::runtime_error(m), status(s) {}
};

#define RGF_HIP(expr)                                                                      \
    do {                                                                                   \
        hipError_t _e = (expr);                                                            \
        if (_e != hipSuccess)                                                              \
            throw ::rgf::Error(_e == hipErrorOutOfMemory ? RGF_ERR_OUT_OF_MEMORY : RGF_ERR_DEVICE, \
                               std::string(#expr) + ": " + hipGetErrorString(_e));          \
    } while (0)

// Planar complex matrix view on the device: element (i, j) of a row-major
// matrix is re[i*ld + j] + 1i*im[i*ld + j].
template <typename T>
struct CMat {
    T* re = nullptr;
    T* im = nullptr;
    int64_t ld = 0;
    CMat() = default;
    CMat(T* r, T* i, int64_t l) : re(r), im(i), ld(l) {}
    CMat offset(int64_t i, int64_t j) const { return CMat(re + i * ld + j, im + i * ld + j, ld); }
};

// Sparse slice operator A0 (CSR) on the device, FP64.
struct DevCSR {
    int32_t* rowptr = nullptr;
    int32_t* col = nullptr;
    double* vre = nullptr;
    double* vim = nullptr;
    int32_t n = 0;
    int32_t nnz = 0;
};

// Event timing of one kernel class on the stream it is launched on.
// Times one launch in every `stride` with a pair of HIP events on its
// stream (events add latency to the stream); flops and launches count the
// timed launches only, so flops / seconds estimates the kernel's rate.
// The events come from `pool` when one is given (kept by the caller across
// timers: creating and destroying a few thousand events per solve costs
// milliseconds of host time), else from the timer's own list.
struct KernelTimer {
    std::vector<hipEvent_t> own;
    std::vector<hipEvent_t>& ev;
    explicit KernelTimer(std::vector<hipEvent_t>* pool = nullptr) : ev(pool ? *pool : own) {}
    KernelTimer(const KernelTimer&) = delete;
    KernelTimer& operator=(const KernelTimer&) = delete;
    size_t used = 0;
    double flops = 0;
    int64_t launches = 0;  // timed launches
    int64_t seen = 0;      // all launches
    int stride = 4;
    bool active = false;
    void begin(hipStream_t s);
    void end(hipStream_t s, double kernel_flops);
    double seconds();  // synchronizes on the recorded events
    ~KernelTimer();
};

// Second and third streams and the events of the look-ahead of the slice
// inverse.
struct Overlap {
    hipStream_t s2 = nullptr, s3 = nullptr;
    hipEvent_t ev[8] = {};
};

// ---- kernel launchers (rgf_kernels.hip) -------------------------------------
// C = alpha * A(MxK) * B(KxN) + beta * C, beta in {0,1}.  Tiles overlapping
// rows or columns [skip_lo, skip_hi) are left untouched (block Gauss-Jordan).
// With a workspace (ws, ws_elems elements of T), thin products with a long
// K are split along K over several work groups and reduced deterministically.
template <typename T>
void cgemm(hipStream_t s, int64_t M, int64_t N, int64_t K, CMat<T> A, CMat<T> B, CMat<T> C,
           cd alpha, int beta, int64_t skip_lo = 0, int64_t skip_hi = 0, int use3m = 0, T* ws = nullptr,
           size_t ws_elems = 0);

// In-place inverse of the N x N matrix A by blocked Gauss-Jordan elimination
// with partial pivoting inside each diagonal block.  ws must hold
// inverse_workspace_elems<T>(N, b) elements of T.  dev_status points to two
// words: dev_status[0] != 0 after the call signals a zero or non-finite
// pivot; dev_status[1] receives (atomic max) the largest |multiplier|^2 of
// the column panels A[:, P] A_PP^-1 as non-negative float bits.  Both are
// accumulated, not cleared.
template <typename T>
size_t inverse_workspace_elems(int64_t N, int b);
template <typename T>
void invert_inplace(hipStream_t s, int64_t N, CMat<T> A, int b, T* ws, int* dev_status, int use3m = 0,
                    KernelTimer* timer = nullptr, const Overlap* ov = nullptr);

// M <- A0 - diag(k2*eps) - M - Gextra  (Gextra may be null).  eps is a
// planar FP64 vector (re, im; im may be null) or, when eps_re is null, the
// constant eps_const.  With src, src takes the place of M on the right (M is
// only written: one pass instead of a copy and an update).
template <typename T>
void assemble_slice(hipStream_t s, int64_t N, CMat<T> M, const double* eps_re, const double* eps_im,
                    cd eps_const, cd k2, const CMat<T>* Gextra, const DevCSR& A0, const CMat<T>* src = nullptr);

// Dense helpers.
template <typename T>
void fill_zero(hipStream_t s, int64_t rows, int64_t cols, CMat<T> A);
// a well-conditioned synthetic slice matrix (for timing the slice inverse)
template <typename T>
void tune_fill(hipStream_t s, int64_t N, CMat<T> A, uint32_t seed);
template <typename T>
void set_identity(hipStream_t s, int64_t N, CMat<T> A);
template <typename Td, typename Ts>
void convert(hipStream_t s, int64_t rows, int64_t cols, CMat<Ts> src, CMat<Td> dst, double scale = 1.0);
template <typename T>
void copy_mat(hipStream_t s, int64_t rows, int64_t cols, CMat<T> src, CMat<T> dst);
// dst(FP64) += scale * src(T)
template <typename T>
void accumulate(hipStream_t s, int64_t rows, int64_t cols, CMat<T> src, CMat<double> dst, double scale);
// dst = a + b (same type)
template <typename T>
void add_mat(hipStream_t s, int64_t rows, int64_t cols, CMat<T> a, CMat<T> b, CMat<T> dst);

// Mode-matrix fill: P1[i][a] = phi(i,a) * g[a], P2[a][i] = conj(phi(i,a)),
// phi((i1,i2),(a1,a2)) = f1[i1][a1] f2[i2][a2]; f1 is n1 x n1, f2 n2 x n2
// (row-major interleaved FP64), g length N interleaved (may be null -> 1).
void fill_modes(hipStream_t s, int64_t n1, int64_t n2, const double* f1, const double* f2,
                const double* g, CMat<double> P1, CMat<double> P2);
// G0 = phi diag(g) phi^H for phi = f1 (x) f2 (n1 x n1 and n2 x n2 tables,
// g length n1 n2, interleaved FP64) in two contractions; M is scratch of
// 2 n1^2 n2 doubles.
void lead_green_separable(hipStream_t s, int64_t n1, int64_t n2, const double* f1, const double* f2, const double* g,
                          double* M, CMat<double> G0);
// Selected mode columns q = 0..n-1 of phi, scaled: cols holds (a1, a2, s_re,
// s_im) per column (doubles).  rows = 0: D(i, q) = phi(i, a_q) s_q (N x n);
// rows = 1: D(q, i) = s_q conj(phi(i, a_q)) (n x N).
void fill_mode_cols(hipStream_t s, int64_t n1, int64_t n2, const double* f1, const double* f2, int64_t n,
                    const double* cols, int rows, CMat<double> D);

// One forward-sweep product of the mixed refinement: D = G t (FP32; G N x N,
// t and D N x m); with `next`, also t = (float)(sc * next) + D, the right-hand
// side of the following slice (fused into the split-K reduction).
void sweep_forward(hipStream_t s, int64_t N, int64_t m, CMat<float> G, CMat<float> t, CMat<float> D,
                   const CMat<double>* next, double sc, int use3m, float* ws, size_t ws_elems);

// Field profiles from the refined solution psi ((L+2) slices of N x m planar
// FP64): out (device, interleaved complex, column-major N x (L+2) x nin) takes
// column col[a] for input a, slices in physical order (reversed if mirrored).
void pack_fields(hipStream_t s, int64_t N, int64_t L, int64_t m, int64_t nin, const double* psi_re,
                 const double* psi_im, const int32_t* col, int mirrored, double* out);

// Residual of the block-tridiagonal system for slices 0..L+1 (FP64):
// r_n = b_n - (A0 - k2 diag(eps_n)) psi_n + psi_{n-1} + psi_{n+1}, with m
// columns; b_0 holds B_L in columns [0, mL), b_{L+1} holds B_R in columns
// [mL, mL+mR), all else is zero.  The lead self-energy terms at slices 0 and
// L+1 are added by the caller.
void residual_stencil(hipStream_t s, int64_t N, int64_t L, int64_t m, const double* psi_re,
                      const double* psi_im, double* r_re, double* r_im, const double* eps_re,
                      const double* eps_im, cd epsL, cd epsR, cd k2, const DevCSR& A0,
                      const double* b0_re, const double* b0_im, int64_t mL, const double* b1_re,
                      const double* b1_im, int64_t mR);

// Residual of the row (adjoint) system y H = c over slices 0..L+1, y_n and
// r_n m x N (FP64, slice stride N m): r_n = c_n - y_n D_n + y_{n-1} + y_{n+1},
// c non-zero at slice L+1 only (m x N); AT = A0 transposed.  The lead terms
// y_0 G0L and y_{L+1} G0R are added by the caller.
void residual_rows(hipStream_t s, int64_t N, int64_t L, int64_t m, const double* y_re, const double* y_im,
                   double* r_re, double* r_im, const double* eps_re, const double* eps_im, cd epsL, cd epsR, cd k2,
                   const DevCSR& AT, const double* c_re, const double* c_im);

// Adjoint contraction for one slice: with u_i = k2 * sum_j psi[i][j] phi[j][i]
// (psi N x m, phi m x N), dT/dRe(eps_i) = 4 Im(u_i) and dT/dIm(eps_i) = 4 Re(u_i).
void grad_contract(hipStream_t s, int64_t N, int64_t m, CMat<double> psi, CMat<double> phi, cd k2, double* gre,
                   double* gim);

// Address of the no-allocation workspace (RGF_NOALLOC_SCRATCH_BYTES) on a
// device, looked up once per device and kept.
void* noalloc_scratch(int device);
// p[0..n) = 0 with a kernel (n <= 64; capturable, unlike a memset node)
void zero_words(hipStream_t s, int* p, int n);
// planar (ld) <-> interleaved complex (ld = N) N x N row-major copies
template <typename T>
void to_interleaved(hipStream_t s, int64_t N, CMat<T> A, T* out);
template <typename T>
void from_interleaved(hipStream_t s, int64_t N, const T* in, CMat<T> A);
// *status |= 1 when info[0] or info[1] is non-zero
void fold_info(hipStream_t s, const int* info, int* status);
// slice-inverse block size per precision (FP64 64, FP32 128)
template <typename T>
constexpr int diag_block_max() {
    return sizeof(T) == 8 ? 64 : 128;
}

// max |x| over a planar FP64 array (result to host, synchronizes the stream).
double max_abs(hipStream_t s, int64_t count, const double* re, const double* im, double* dev_scratch);

// ---- batched engine (many independent mixed-precision solves in lockstep) ----
// One problem of the batch, device-resident.  The transverse size is padded
// to NP (a multiple of 64) with the identity: the padded slice matrices are
// diag(M, I), their inverses diag(g, I), and the padded rows of every
// right-hand side, correction and solution stay zero.  Complex arrays are
// planar: the re plane of the whole array, then the im plane.
struct BatchProb {
    int32_t N, NP, nb, L;           // transverse size, padded size, 64-wide blocks, slices
    int32_t m, mL, mR, npo, nqo;    // right-hand sides (mL left + mR right inputs, padded to 4), outputs L / R
    int32_t pad0;
    const int32_t* rowptr;          // slice stencil A0 (CSR, N rows, FP64)
    const int32_t* col;
    const double* vre;
    const double* vim;
    const double* eps_re;           // N x L, slice n at n N; eps_im may be null
    const double* eps_im;
    double k2r, k2i, eLr, eLi, eRr, eRi;
    float* F;                       // factor store: L + 2 slices of 2 x NP x NP
    float* ws;                      // D (2 x 64 x 64), R (2 x 64 x NP), C (2 x NP x 64)
    int32_t* status;                // [0] pivot flags, [1] max multiplier^2 (float bits)
    double* G0L;                    // 2 x NP x NP, zero outside N x N
    double* G0R;
    double* B;                      // 2 x NP x m: B_L in columns [0, mL), B_R in [mL, mL + mR)
    double* CL;                     // 2 x npo x NP
    double* CR;                     // 2 x nqo x NP
    double* psi;                    // 2 x (L + 2) x NP x m
    float* z;                       // 2 x (L + 2) x NP x m: forward z_n, then the correction
    float* t;                       // 2 x NP x m
    double* W;                      // 2 x 2 x NP x m: G0L psi_0 (re, im), then G0R psi_{L+1}
    double* T;                      // 2 x (npo + nqo) x m: C_L psi_0 and C_R psi_{L+1}
    unsigned long long* rmax;       // max |r| of the last residual (double bits)
};

// Work maps: entry k of a map is (problem << 20) | local index.
// Dyson round r (read as *rbase + k): the diagonal-block inverses (one work
// group per table entry) on s beside the write-back / assembly row pass
// (k1map: groups of kBatchRows rows, n1) on s2 (forked and joined with the
// two events), then the panel products (k2map, n2 tiles) and the trailing
// updates (k3map, n3 tiles) on s.
void batch_round(hipStream_t s, hipStream_t s2, hipEvent_t fork, hipEvent_t join, const BatchProb* P, int np,
                 const int* rbase, int k, const int32_t* k1map, int n1, const int32_t* k2map, int n2,
                 const int32_t* k3map, int n3, int use3m);
constexpr int kBatchRows = 4;  // rows per work group of the row-wise passes (a wavefront per row)
// Refinement.  map entries index the sweep-product tiles (NP/64 x m/64 per
// problem), the row groups of the residual pass or the kBatchChunk-element
// chunks of an elementwise pass.
// residual pass: r = b - H psi in FP32 into z, max |r| per problem into
// rmax; first = psi is zero (only the lead slices, groups of 64 rows of the
// 2 N; otherwise of the S N slice rows)
void batch_resmax(hipStream_t s, const BatchProb* P, const int32_t* map, int nchunks, int first);
constexpr int kBatchResRows = 64;  // slice rows per work group of the residual pass
// lead terms W = G0 psi at the two lead slices (FP64 products)
void batch_lead_terms(hipStream_t s, const BatchProb* P, const int32_t* map, int ntiles);
// forward step n: t = r_n + z_{n-1} (chunks of N m); then z_n = g_n t
void batch_fwd_rhs(hipStream_t s, const BatchProb* P, const int32_t* map, int nchunks, int n);
void batch_fwd_gemm(hipStream_t s, const BatchProb* P, const int32_t* map, int ntiles, int n, int use3m, int first);
// backward step k: z_n += g_n z_{n+1} with n = L - k (each problem's own end)
void batch_bwd_gemm(hipStream_t s, const BatchProb* P, const int32_t* map, int ntiles, int k, int use3m);
// psi += z over all slices (first: psi = z)
void batch_accumulate(hipStream_t s, const BatchProb* P, const int32_t* map, int nchunks, int first);
// T = [C_L psi_0; C_R psi_{L+1}]
void batch_project(hipStream_t s, const BatchProb* P, const int32_t* map, int ntiles);
// F[0] = (float) G0L with the identity on the padding
void batch_first_slice(hipStream_t s, const BatchProb* P, const int32_t* map, int nchunks);
void batch_advance(hipStream_t s, int* rbase, int k);
constexpr int kBatchChunk = 16384;  // elements per work group of the elementwise passes

}  // namespace rgf
