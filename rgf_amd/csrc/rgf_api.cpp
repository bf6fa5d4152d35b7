// C ABI of librgf and the slice-recursion drivers.
//
// Two drivers share the kernels of rgf_kernels.hip:
//
//  * carry (FP64): the reference's slice loop (cal_smatrix_RGF.m:215-280)
//    with the propagators carried projected onto the requested channels:
//      g_n = (A_n - g_{n-1})^-1,  X = g_n X,  Z += Y X,  Y = Y g_n
//    where X = G_RL B, Y = C G_LR, Z = C G_LL B.  B and C are linear, so
//    this equals the reference's G_RL/G_LR/G_LL products followed by its
//    Fisher-Lee projection, at O(N^2 p) instead of O(N^3) per slice.  The
//    final lead attach (I - G0_R g)\G0_R equals (A_lead - G0_R - g)^-1
//    because G0_R = (A_lead - G0_R)^-1, so it is one more Dyson step.  The
//    loop runs from the side with fewer requested channels, which reproduces
//    the reference's choice of loop direction for all five schemes.
//
//  * refined (FP32 + FP64 refinement): the same Dyson recursion in complex
//    FP32 on MFMA with every slice inverse g_n stored; the S-matrix columns
//    are the block-tridiagonal solve H psi = b (slices 0..L+1, leads folded
//    into the end blocks), corrected by FP64 residual iterative refinement
//    psi += H~^-1 (b - H psi), where H~^-1 is the forward/backward sweep
//    through the stored FP32 inverses and the residual uses the FP64 slice
//    operators and FP64 lead Green's functions.
#include <hip/hip_runtime.h>
#include <rocsolver/rocsolver.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <functional>
#include <initializer_list>
#include <map>
#include <memory>
#include <atomic>
#include <mutex>
#include <thread>
#include <string>

#include "rgf_context.h"

using namespace rgf;


static std::string g_last_error;
static std::mutex g_err_mu;

namespace rgf {

double now() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}


// bytes the named buffers already hold (dbuf reuses or grows them in place)
size_t held_bytes(rgf_context* c, std::initializer_list<const char*> names) {
    size_t t = 0;
    for (const char* n : names) {
        auto it = c->bufs.find(n);
        if (it != c->bufs.end()) t += it->second.bytes;
    }
    return t;
}

void release(rgf_context* c, const std::string& name) {
    auto it = c->bufs.find(name);
    if (it == c->bufs.end()) return;
    if (it->second.p) {
        RGF_HIP(hipStreamSynchronize(c->stream));
        RGF_HIP(hipFree(it->second.p));
    }
    c->bufs.erase(it);
}


// host complex row-major -> planar device
void upload(rgf_context* c, const std::vector<cd>& h, int64_t rows, int64_t cols, CMat<double> d) {
    if (rows <= 0 || cols <= 0) return;
    std::vector<double> re((size_t)rows * cols), im((size_t)rows * cols);
    for (int64_t i = 0; i < rows; ++i)
        for (int64_t j = 0; j < cols; ++j) {
            re[i * cols + j] = h[i * cols + j].real();
            im[i * cols + j] = h[i * cols + j].imag();
        }
    if (d.ld == cols) {
        RGF_HIP(hipMemcpyAsync(d.re, re.data(), re.size() * sizeof(double), hipMemcpyHostToDevice, c->stream));
        RGF_HIP(hipMemcpyAsync(d.im, im.data(), im.size() * sizeof(double), hipMemcpyHostToDevice, c->stream));
    } else {
        RGF_HIP(hipMemcpy2DAsync(d.re, d.ld * sizeof(double), re.data(), cols * sizeof(double), cols * sizeof(double),
                                 rows, hipMemcpyHostToDevice, c->stream));
        RGF_HIP(hipMemcpy2DAsync(d.im, d.ld * sizeof(double), im.data(), cols * sizeof(double), cols * sizeof(double),
                                 rows, hipMemcpyHostToDevice, c->stream));
    }
    RGF_HIP(hipStreamSynchronize(c->stream));
}

std::vector<cd> download(rgf_context* c, CMat<double> d, int64_t rows, int64_t cols) {
    if (rows <= 0 || cols <= 0) return std::vector<cd>();
    std::vector<double> re((size_t)rows * cols), im((size_t)rows * cols);
    if (d.ld == cols) {
        RGF_HIP(hipMemcpyAsync(re.data(), d.re, re.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
        RGF_HIP(hipMemcpyAsync(im.data(), d.im, im.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    } else {
        RGF_HIP(hipMemcpy2DAsync(re.data(), cols * sizeof(double), d.re, d.ld * sizeof(double), cols * sizeof(double),
                                 rows, hipMemcpyDeviceToHost, c->stream));
        RGF_HIP(hipMemcpy2DAsync(im.data(), cols * sizeof(double), d.im, d.ld * sizeof(double), cols * sizeof(double),
                                 rows, hipMemcpyDeviceToHost, c->stream));
    }
    RGF_HIP(hipStreamSynchronize(c->stream));
    std::vector<cd> h((size_t)rows * cols);
    for (size_t q = 0; q < h.size(); ++q) h[q] = cd(re[q], im[q]);
    return h;
}

// Largest multiplier |(A[:, P] A_PP^-1)_ij| the slice inverses may apply to
// the rows outside a diagonal block before the result is reported as
// inaccurate: pivoting is confined to the block, so an ill-conditioned block
// (in a well-conditioned matrix) shows up as large multipliers, which
// amplify the rounding of every later update by about that factor.
double multiplier_limit(bool fp64) { return fp64 ? 1e6 : 1e3; }

// A slice inverse the blocked Gauss-Jordan kernels refuse: a zero or
// non-finite pivot, or multipliers above multiplier_limit (their pivoting is
// confined to the diagonal block).  The solve drivers then invert every
// slice again with partial pivoting over all rows (with_pivoting_fallback).
struct SliceRefused : Error {
    using Error::Error;
};

// The flags word of the slice-inverse status: bit 1 a zero or non-finite
// pivot, bit 2 a hand-over between the work groups of a panel launch that
// timed out (the block-two-ahead step then read unfinished panel tiles).
void status_error(int flags) {
    if (flags & 2) throw Error(RGF_ERR_DEVICE, "a work-group hand-over in a slice inverse timed out");
    if (flags) throw SliceRefused(RGF_ERR_NUMERICAL, "zero or non-finite pivot in a slice inverse");
}

// Status of the slice inverses since the last reset_status: a zero or
// non-finite pivot is an error; so is a multiplier above multiplier_limit
// when `growth` (the FP64 path and the dense inverse; the mixed path's FP64
// refinement measures the accuracy it reaches and does not need it).
double check_status(rgf_context* c, bool growth = true, bool fp64 = true) {
    int st[2] = {0, 0};
    RGF_HIP(hipMemcpyAsync(st, c->d_status, sizeof st, hipMemcpyDeviceToHost, c->stream));
    RGF_HIP(hipStreamSynchronize(c->stream));
    status_error(st[0]);
    float m2;
    memcpy(&m2, &st[1], sizeof m2);
    const double m = std::sqrt((double)m2);
    if (growth && !(m <= multiplier_limit(fp64))) {
        char buf[256];
        snprintf(buf, sizeof buf,
                 "ill-conditioned diagonal block in a slice inverse: multiplier %.3g above %.0e (pivoting is "
                 "confined to the diagonal block, so the inverse would be inaccurate)",
                 m, multiplier_limit(fp64));
        throw SliceRefused(RGF_ERR_NUMERICAL, buf);
    }
    return m;
}

void reset_status(rgf_context* c) { RGF_HIP(hipMemsetAsync(c->d_status, 0, 2 * sizeof(int), c->stream)); }

void rocblas_ok(rocblas_status st, const char* what) {
    if (st != rocblas_status_success) {
        char buf[160];
        snprintf(buf, sizeof buf, "%s failed (rocblas status %d)", what, (int)st);
        throw Error(st == rocblas_status_memory_error ? RGF_ERR_OUT_OF_MEMORY : RGF_ERR_DEVICE, buf);
    }
}

// The slice inverse with partial pivoting over all rows, as MATLAB's inv
// (cal_smatrix_RGF.m:265; LAPACK getrf + getri), through the vendor's dense
// solver library: the fallback for slices the blocked Gauss-Jordan kernels
// refuse, and every slice with rgf_options.pivoting = 1.  The row-major
// matrix is the column-major transpose and inv(A^T) = inv(A)^T, so its
// interleaved copy goes to the solver as it is.  A zero pivot sets bit 1 of
// the status word.
template <typename T>
void invert_pivoted(rgf_context* c, hipStream_t s, int64_t N, CMat<T> A) {
    if (N > 46340) throw Error(RGF_ERR_INVALID_ARGUMENT, "matrix too large for the pivoted slice inverse");
    if (!c->rb) {
        rocblas_handle h = nullptr;
        rocblas_ok(rocblas_create_handle(&h), "rocblas_create_handle");
        c->rb = h;
    }
    rocblas_handle rb = (rocblas_handle)c->rb;
    rocblas_ok(rocblas_set_stream(rb, s), "rocblas_set_stream");
    T* M = dbuf<T>(c, "piv_m", (size_t)2 * N * N);
    rocblas_int* ipiv = dbuf<rocblas_int>(c, "piv_ipiv", (size_t)N);
    rocblas_int* info = dbuf<rocblas_int>(c, "piv_info", 2);
    const rocblas_int n = (rocblas_int)N;
    to_interleaved<T>(s, N, A, M);
    if constexpr (sizeof(T) == 8) {
        rocblas_ok(rocsolver_zgetrf(rb, n, n, (rocblas_double_complex*)M, n, ipiv, info), "rocsolver_zgetrf");
        rocblas_ok(rocsolver_zgetri(rb, n, (rocblas_double_complex*)M, n, ipiv, info + 1), "rocsolver_zgetri");
    } else {
        rocblas_ok(rocsolver_cgetrf(rb, n, n, (rocblas_float_complex*)M, n, ipiv, info), "rocsolver_cgetrf");
        rocblas_ok(rocsolver_cgetri(rb, n, (rocblas_float_complex*)M, n, ipiv, info + 1), "rocsolver_cgetri");
    }
    from_interleaved<T>(s, N, M, A);
    fold_info(s, info, c->d_status);
}

// one slice inverse of a solve: the blocked kernels, or with w.pivot_all the
// pivoted inverse
template <typename T>
void invert_slice(rgf_context* c, const Work& w, hipStream_t s, int64_t N, CMat<T> A, int b, T* ws, int use3m,
                  KernelTimer* timer, const Overlap* ov) {
    if (w.pivot_all)
        invert_pivoted<T>(c, s, N, A);
    else
        invert_inplace<T>(s, N, A, b, ws, c->d_status, use3m, timer, ov);
}

// Runs a solve driver f on w.  rgf_options.pivoting 0 (default): when the
// blocked slice inverses refuse a slice, or in the mixed precision when the
// refinement stalls after multipliers above the FP32 limit, f runs again with
// every slice inverted with partial pivoting over all rows; 1: pivoted from
// the start; 2: no fallback (the refusal is the caller's error).
void with_pivoting_fallback(const rgf_options& o, Work& w, const std::function<void()>& f) {
    w.pivot_all = o.pivoting == 1;
    try {
        f();
        return;
    } catch (const Error& e) {
        const bool refused = dynamic_cast<const SliceRefused*>(&e) != nullptr;
        const bool stalled = e.status == RGF_ERR_NOT_CONVERGED && o.precision == RGF_PRECISION_FP32_REFINED &&
                             !(w.max_multiplier <= multiplier_limit(false));
        if (o.pivoting != 0 || !(refused || stalled)) throw;
    }
    w.pivot_all = 1;
    f();
}


void upload_eps(rgf_context* c, const rgf_problem* p, Work& w) {
    const int64_t N = w.N, L = w.L;
    w.eps_re = dbuf<double>(c, "eps_re", (size_t)N * std::max<int64_t>(L, 1));
    w.eps_im = p->eps_im ? dbuf<double>(c, "eps_im", (size_t)N * std::max<int64_t>(L, 1)) : nullptr;
    const hipMemcpyKind kind = p->eps_on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
    if (L == 0) return;
    if (!w.mirrored) {
        RGF_HIP(hipMemcpyAsync(w.eps_re, p->eps_re, N * L * sizeof(double), kind, c->stream));
        if (w.eps_im) RGF_HIP(hipMemcpyAsync(w.eps_im, p->eps_im, N * L * sizeof(double), kind, c->stream));
    } else {
        for (int64_t n = 0; n < L; ++n) {
            RGF_HIP(hipMemcpyAsync(w.eps_re + n * N, p->eps_re + (L - 1 - n) * N, N * sizeof(double), kind, c->stream));
            if (w.eps_im)
                RGF_HIP(hipMemcpyAsync(w.eps_im + n * N, p->eps_im + (L - 1 - n) * N, N * sizeof(double), kind,
                                       c->stream));
        }
    }
}

void upload_stencil(rgf_context* c, const Channels& ch, Work& w) {
    HostCSR h = slice_stencil(ch);
    DevCSR& d = w.A0;
    d.n = h.n;
    d.nnz = (int32_t)h.col.size();
    d.rowptr = dbuf<int32_t>(c, "A0_rowptr", h.rowptr.size());
    d.col = dbuf<int32_t>(c, "A0_col", h.col.size());
    d.vre = dbuf<double>(c, "A0_vre", h.vre.size());
    d.vim = dbuf<double>(c, "A0_vim", h.vim.size());
    RGF_HIP(hipMemcpyAsync(d.rowptr, h.rowptr.data(), h.rowptr.size() * 4, hipMemcpyHostToDevice, c->stream));
    RGF_HIP(hipMemcpyAsync(d.col, h.col.data(), h.col.size() * 4, hipMemcpyHostToDevice, c->stream));
    RGF_HIP(hipMemcpyAsync(d.vre, h.vre.data(), h.vre.size() * 8, hipMemcpyHostToDevice, c->stream));
    RGF_HIP(hipMemcpyAsync(d.vim, h.vim.data(), h.vim.size() * 8, hipMemcpyHostToDevice, c->stream));
    RGF_HIP(hipStreamSynchronize(c->stream));
}


ModeTables upload_modes(rgf_context* c, const Channels& ch) {
    double* df1 = dbuf<double>(c, "f1", 2 * ch.f1.size());
    double* df2 = dbuf<double>(c, "f2", 2 * ch.f2.size());
    // std::complex<double> is two contiguous doubles (re, im)
    RGF_HIP(hipMemcpyAsync(df1, ch.f1.data(), ch.f1.size() * 16, hipMemcpyHostToDevice, c->stream));
    RGF_HIP(hipMemcpyAsync(df2, ch.f2.data(), ch.f2.size() * 16, hipMemcpyHostToDevice, c->stream));
    RGF_HIP(hipStreamSynchronize(c->stream));
    return {df1, df2};
}

// G0 = phi diag(exp(i kl)) phi^H, FP64 (cal_smatrix_RGF.m:198-203)
void lead_green(rgf_context* c, const Channels& ch, const ModeTables& f, const Side& side, CMat<double> G0) {
    const int64_t N = ch.N;
    std::vector<double> g(2 * N);
    for (int64_t a = 0; a < N; ++a) {
        const cd e = std::exp(cd(0, 1) * side.kl[a]);
        g[2 * a] = e.real();
        g[2 * a + 1] = e.imag();
    }
    double* dg = dbuf<double>(c, "gdiag", g.size());
    RGF_HIP(hipMemcpyAsync(dg, g.data(), g.size() * 8, hipMemcpyHostToDevice, c->stream));
    if (ch.n2 > 1) {
        // 3D lead: phi = f1 (x) f2, two contractions (N^2 (n1 + n2) work, no N x N temporaries)
        double* M = dbuf<double>(c, "g0_stage", (size_t)2 * ch.n1 * ch.n1 * ch.n2);
        lead_green_separable(c->stream, ch.n1, ch.n2, f.f1, f.f2, dg, M, G0);
    } else {
        CMat<double> P1 = dmat<double>(c, "tmpA", N, N), P2 = dmat<double>(c, "tmpB", N, N);
        fill_modes(c->stream, ch.n1, ch.n2, f.f1, f.f2, dg, P1, P2);
        cgemm<double>(c->stream, N, N, N, P1, P2, G0, cd(1, 0), 0);
    }
    RGF_HIP(hipStreamSynchronize(c->stream));
}

// B = phi(:, prop(in)) diag(sqrt_mu(in)) (N x n, rows = 0); C = diag(sqrt_mu(out))
// phi(:, prop(out))^H (n x N, rows = 1) (cal_smatrix_RGF.m:189-195; sqrt_mu is
// not conjugated in C).  Filled on the device from the mode tables.
void mode_cols(rgf_context* c, const Channels& ch, const ModeTables& f, const Side& s,
               const std::vector<int32_t>& idx, int rows, const std::string& name, CMat<double> D) {
    const int64_t n = idx.size();
    if (n == 0) return;
    std::vector<double> cols(4 * n);
    for (int64_t q = 0; q < n; ++q) {
        const int64_t a = s.ind_prop[idx[q]];
        cols[4 * q] = (double)(a % ch.n1);
        cols[4 * q + 1] = (double)(a / ch.n1);
        cols[4 * q + 2] = s.sqrt_mu[idx[q]].real();
        cols[4 * q + 3] = s.sqrt_mu[idx[q]].imag();
    }
    double* dcols = dbuf<double>(c, name, cols.size());
    RGF_HIP(hipMemcpyAsync(dcols, cols.data(), cols.size() * 8, hipMemcpyHostToDevice, c->stream));
    fill_mode_cols(c->stream, ch.n1, ch.n2, f.f1, f.f2, n, dcols, rows, D);
    RGF_HIP(hipStreamSynchronize(c->stream));
}

void prepare(rgf_context* c, const rgf_problem* p, const Channels& ch, const Selection& sel, Work& w,
             int force_mirror = -1) {
    w.ch = &ch;
    w.N = ch.N;
    w.L = ch.L_slices;
    w.k2 = ch.k2;
    w.mirrored = force_mirror >= 0 ? force_mirror == 1
                                   : (sel.inR.size() + sel.outR.size()) < (sel.inL.size() + sel.outL.size());
    if (!w.mirrored) {
        w.sL = &ch.L;
        w.sR = &ch.R;
        w.epsL = ch.epsL;
        w.epsR = ch.epsR;
        w.inL = sel.inL;
        w.inR = sel.inR;
        w.outL = sel.outL;
        w.outR = sel.outR;
    } else {
        w.sL = &ch.R;
        w.sR = &ch.L;
        w.epsL = ch.epsR;
        w.epsR = ch.epsL;
        w.inL = sel.inR;
        w.inR = sel.inL;
        w.outL = sel.outR;
        w.outR = sel.outL;
    }
    const int64_t N = w.N;
    upload_eps(c, p, w);
    upload_stencil(c, ch, w);
    const ModeTables f = upload_modes(c, ch);
    w.G0L = dmat<double>(c, "G0L", N, N);
    lead_green(c, ch, f, *w.sL, w.G0L);
    if (ch.same_sides) {
        w.G0R = w.G0L;
    } else {
        w.G0R = dmat<double>(c, "G0R", N, N);
        lead_green(c, ch, f, *w.sR, w.G0R);
    }
    release(c, "tmpA");
    release(c, "tmpB");
    w.BL = dmat<double>(c, "BL", N, w.inL.size());
    w.BR = dmat<double>(c, "BR", N, w.inR.size());
    w.CL = dmat<double>(c, "CL", w.outL.size(), N);
    w.CR = dmat<double>(c, "CR", w.outR.size(), N);
    mode_cols(c, ch, f, *w.sL, w.inL, 0, "BL_cols", w.BL);
    mode_cols(c, ch, f, *w.sR, w.inR, 0, "BR_cols", w.BR);
    mode_cols(c, ch, f, *w.sL, w.outL, 1, "CL_cols", w.CL);
    mode_cols(c, ch, f, *w.sR, w.outR, 1, "CR_cols", w.CR);
}

// split-K workspace for the thin GEMMs (elements of T)
template <typename T>
std::pair<T*, size_t> splitk_ws(rgf_context* c) {
    const size_t elems = (size_t)(64u << 20) / sizeof(T);  // 64 MiB
    return {dbuf<T>(c, sizeof(T) == 8 ? "splitk64" : "splitk32", elems), elems};
}

int default_block(int precision) { return precision == RGF_PRECISION_FP64 ? 64 : 128; }

// ---------------------------------------------------------------- slice-inverse autotuner
// The slice inverse's block width and stream use, timed once per process for
// each (device, N, precision, three-product bit, single_stream option) on a
// synthetic slice matrix of that size (tune_fill): widths FP32 {64, 128},
// FP64 {32, 64}, each with the look-ahead streams and on one stream (only
// the latter when the caller asked for one stream).  Each candidate runs
// three times, interleaved with the others (HIP events on the handle's
// stream); the first run warms up and the faster of the other two counts.
// A candidate whose inverse reports a zero pivot is dropped.
struct SliceCfg {
    int b = 0;
    bool lookahead = true;
};
struct TuneKey {
    int dev;
    int64_t n;
    int prec, use3m, single;
    bool operator<(const TuneKey& o) const {
        return std::tie(dev, n, prec, use3m, single) < std::tie(o.dev, o.n, o.prec, o.use3m, o.single);
    }
};
std::mutex g_tune_mu;
std::map<TuneKey, SliceCfg> g_tuned;

template <typename T>
SliceCfg tune_slice(rgf_context* c, int64_t N, int use3m, const std::vector<SliceCfg>& cand) {
    hipStream_t s = c->stream;
    size_t wse = 0;
    for (const SliceCfg& k : cand) wse = std::max(wse, inverse_workspace_elems<T>(N, k.b));
    CMat<T> A = dmat<T>(c, "tune_A", N, N);
    T* ws = dbuf<T>(c, "tune_ws", wse);
    hipEvent_t e0, e1;
    RGF_HIP(hipEventCreate(&e0));
    RGF_HIP(hipEventCreate(&e1));
    std::vector<double> best(cand.size(), INFINITY);
    for (int rep = 0; rep < 3; ++rep)
        for (size_t k = 0; k < cand.size(); ++k) {
            tune_fill<T>(s, N, A, 17u + (uint32_t)rep);
            reset_status(c);
            RGF_HIP(hipEventRecord(e0, s));
            invert_inplace<T>(s, N, A, cand[k].b, ws, c->d_status, use3m, nullptr,
                              cand[k].lookahead ? &c->ov : nullptr);
            RGF_HIP(hipEventRecord(e1, s));
            int st[2] = {0, 0};
            RGF_HIP(hipMemcpyAsync(st, c->d_status, sizeof st, hipMemcpyDeviceToHost, s));
            RGF_HIP(hipStreamSynchronize(s));
            float ms = 0;
            RGF_HIP(hipEventElapsedTime(&ms, e0, e1));
            if (st[0]) best[k] = -1;  // unusable
            else if (rep > 0 && best[k] >= 0) best[k] = std::min(best[k], (double)ms);
        }
    RGF_HIP(hipEventDestroy(e0));
    RGF_HIP(hipEventDestroy(e1));
    release(c, "tune_A");
    release(c, "tune_ws");
    reset_status(c);
    int pick = 0;
    for (size_t k = 1; k < cand.size(); ++k)
        if (best[k] >= 0 && (best[pick] < 0 || best[k] < best[pick])) pick = (int)k;
    if (getenv("RGF_TUNE_VERBOSE")) {
        fprintf(stderr, "rgf autotune: N=%lld %s%s:", (long long)N, sizeof(T) == 8 ? "fp64" : "fp32",
                use3m ? " 3m" : "");
        for (size_t k = 0; k < cand.size(); ++k)
            fprintf(stderr, " b=%d%s %.3f ms", cand[k].b, cand[k].lookahead ? "" : "/1s", best[k]);
        fprintf(stderr, " -> b=%d%s\n", cand[pick].b, cand[pick].lookahead ? "" : "/1s");
    }
    return best[pick] < 0 ? SliceCfg{} : cand[pick];
}

// The block width and stream use of the slice inverses of a solve:
// opt.block_size when set, else the tuned width (opt.autotune) or the fixed
// default; one stream when opt.single_stream is set or the tuner chose it.
SliceCfg slice_config(rgf_context* c, int64_t N, int precision, int use3m, const rgf_options& o, Work& w) {
    SliceCfg cfg{o.block_size > 0 ? o.block_size : default_block(precision), !o.single_stream};
    if (o.block_size <= 0 && o.autotune && N > 64) {
        const TuneKey key{c->device, N, precision, use3m, o.single_stream ? 1 : 0};
        std::lock_guard<std::mutex> lk(g_tune_mu);
        auto it = g_tuned.find(key);
        if (it == g_tuned.end()) {
            std::vector<SliceCfg> cand;
            for (int b : precision == RGF_PRECISION_FP64 ? std::vector<int>{32, 64} : std::vector<int>{64, 128}) {
                if (!o.single_stream) cand.push_back({b, true});
                cand.push_back({b, false});
            }
            SliceCfg t = precision == RGF_PRECISION_FP64 ? tune_slice<double>(c, N, 0, cand)
                                                         : tune_slice<float>(c, N, use3m, cand);
            it = g_tuned.emplace(key, t.b > 0 ? t : cfg).first;
            w.tuned = 1;
        }
        cfg = it->second;
    }
    w.block_size = cfg.b;
    return cfg;
}

// ---------------------------------------------------------------- FP64 carry driver
void run_carry(rgf_context* c, Work& w, const rgf_options& o) {
    const int64_t N = w.N, L = w.L;
    const int64_t pi = w.inL.size(), po = w.outL.size();
    const SliceCfg cfg = slice_config(c, N, RGF_PRECISION_FP64, 0, o, w);
    const int b = cfg.b;
    hipStream_t s = c->stream;
    CMat<double> G = dmat<double>(c, "G", N, N);
    double* ws = dbuf<double>(c, "inv_ws", inverse_workspace_elems<double>(N, b));
    // X rows padded to an even length (16-byte rows for the FP64 GEMMs); the
    // padding column is never read into a stored result
    const int64_t pip = (pi + 1) / 2 * 2;
    CMat<double> X = dmat<double>(c, "X", N, pip), X2 = dmat<double>(c, "X2", N, pip);
    CMat<double> Y = dmat<double>(c, "Y", po, N), Y2 = dmat<double>(c, "Y2", po, N);
    CMat<double> Z = dmat<double>(c, "Z", po, pi);
    auto wk = splitk_ws<double>(c);
    copy_mat<double>(s, N, N, w.G0L, G);
    // thin products with K = N: split along K over the workspace
    if (pi) cgemm<double>(s, N, pi, N, w.G0L, w.BL, X, cd(1, 0), 0, 0, 0, 0, wk.first, wk.second);
    if (po) cgemm<double>(s, po, N, N, w.CL, w.G0L, Y, cd(1, 0), 0, 0, 0, 0, wk.first, wk.second);
    if (po && pi) cgemm<double>(s, po, pi, N, w.CL, X, Z, cd(1, 0), 0, 0, 0, 0, wk.first, wk.second);
    reset_status(c);
    RGF_HIP(hipStreamSynchronize(s));
    const double t0 = now();
    for (int64_t n = 1; n <= L + 1; ++n) {
        if (n <= L)
            assemble_slice<double>(s, N, G, w.eps_re + (n - 1) * N, w.eps_im ? w.eps_im + (n - 1) * N : nullptr,
                                   cd(0, 0), w.k2, nullptr, w.A0);
        else
            assemble_slice<double>(s, N, G, nullptr, nullptr, w.epsR, w.k2, &w.G0R, w.A0);
        invert_slice<double>(c, w, s, N, G, b, ws, 0, w.timer, cfg.lookahead ? &c->ov : nullptr);
        if (pi) {
            cgemm<double>(s, N, pi, N, G, X, X2, cd(1, 0), 0, 0, 0, 0, wk.first, wk.second);
            std::swap(X, X2);
        }
        if (po && pi) cgemm<double>(s, po, pi, N, Y, X, Z, cd(1, 0), 1, 0, 0, 0, wk.first, wk.second);
        if (po) {
            cgemm<double>(s, po, N, N, Y, G, Y2, cd(1, 0), 0, 0, 0, 0, wk.first, wk.second);
            std::swap(Y, Y2);
        }
    }
    RGF_HIP(hipStreamSynchronize(s));
    w.t_slices = now() - t0;
    const double Nd = (double)N;
    w.flops = (double)(L + 1) * (8 * Nd * Nd * Nd + 8 * Nd * Nd * (pi + po) + 8.0 * po * Nd * pi);
    w.max_multiplier = check_status(c);
    const int64_t qi = w.inR.size(), qo = w.outR.size();
    if (po && pi) w.TLL = download(c, Z, po, pi);
    if (qo && pi) {
        CMat<double> T = dmat<double>(c, "T", qo, pi);
        cgemm<double>(s, qo, pi, N, w.CR, X, T, cd(1, 0), 0, 0, 0, 0, wk.first, wk.second);
        w.TRL = download(c, T, qo, pi);
    }
    if (po && qi) {
        CMat<double> T = dmat<double>(c, "T", po, qi);
        cgemm<double>(s, po, qi, N, Y, w.BR, T, cd(1, 0), 0, 0, 0, 0, wk.first, wk.second);
        w.TLR = download(c, T, po, qi);
    }
    if (qo && qi) {
        CMat<double> GB = dmat<double>(c, "GB", N, qi), T = dmat<double>(c, "T", qo, qi);
        cgemm<double>(s, N, qi, N, G, w.BR, GB, cd(1, 0), 0, 0, 0, 0, wk.first, wk.second);
        cgemm<double>(s, qo, qi, N, w.CR, GB, T, cd(1, 0), 0, 0, 0, 0, wk.first, wk.second);
        w.TRR = download(c, T, qo, qi);
    }
}

// ---------------------------------------------------------------- mixed-path factor store
// The FP32 slice inverses g_0 .. g_{S-1} of the mixed path.  When all S fit
// in device memory they stay resident.  Otherwise g_0 .. g_{R-1} stay
// resident and the slices from R on are grouped in segments of C slices that
// the refinement sweeps rebuild when they reach them: segment k (slices
// R + kC .. R + (k+1)C - 1) is rebuilt by the Dyson steps of the first pass
// from the slice before it, g_{R-1} (resident) for k = 0 and a stored
// checkpoint for k > 0, with the same kernels and inputs, so the rebuilt
// inverses are the same numbers.  Device memory: R + (segments - 1) + C
// slices instead of S.  A forward sweep rebuilds every segment once and a
// backward sweep every segment but the one it starts in.
struct FactorStore {
    rgf_context* c = nullptr;
    const Work* w = nullptr;
    int64_t N = 0, S = 0, R = 0, C = 1, nseg = 0;
    int b = 0, use3m = 0;
    const Overlap* ov = nullptr;
    float* res = nullptr;  // R slots: g_0 .. g_{R-1}
    float* ck = nullptr;   // nseg - 1 slots: ck[k - 1] = g_{R + kC - 1}
    float* seg = nullptr;  // C slots: the segment `cur`
    float* ws = nullptr;   // slice-inverse workspace
    CMat<float> G0R32;     // the right lead's G0 (assembly of slice S - 1)
    int64_t cur = -1;
    int64_t rebuilt = 0;   // slices rebuilt by the sweeps

    bool hybrid() const { return R < S; }
    CMat<float> slot(float* base, int64_t k) const {
        float* p = base + (size_t)k * N * N * 2;
        return CMat<float>(p, p + (size_t)N * N, N);
    }
    // g_n = (D_n - g_{n-1})^-1 into out (the first pass's step)
    void step(int64_t n, CMat<float> prev, CMat<float> out, KernelTimer* timer) const {
        hipStream_t s = c->stream;
        const int64_t L = S - 2;
        if (n <= L) {
            assemble_slice<float>(s, N, out, w->eps_re + (n - 1) * N, w->eps_im ? w->eps_im + (n - 1) * N : nullptr,
                                  cd(0, 0), w->k2, nullptr, w->A0, &prev);
        } else {
            copy_mat<float>(s, N, N, prev, out);
            assemble_slice<float>(s, N, out, nullptr, nullptr, w->epsR, w->k2, &G0R32, w->A0);
        }
        invert_slice<float>(c, *w, s, N, out, b, ws, use3m, timer, ov);
    }
    // segment k into the segment slots; in the first pass, also store the
    // checkpoint of segment k + 1
    void build(int64_t k, bool first, KernelTimer* timer) {
        const int64_t n0 = R + k * C, n1 = std::min(n0 + C, S);
        for (int64_t n = n0; n < n1; ++n)
            step(n, n > n0 ? slot(seg, n - 1 - n0) : (k == 0 ? slot(res, R - 1) : slot(ck, k - 1)), slot(seg, n - n0),
                 timer);
        if (first && k + 1 < nseg) copy_mat<float>(c->stream, N, N, slot(seg, n1 - 1 - n0), slot(ck, k));
        if (!first) rebuilt += n1 - n0;
        cur = k;
    }
    // g_n, rebuilding its segment first when needed
    CMat<float> at(int64_t n) {
        if (n < R) return slot(res, n);
        const int64_t k = (n - R) / C;
        if (k != cur) build(k, false, nullptr);
        return slot(seg, n - R - k * C);
    }
};

// The resident slices R and segment length C for at most K slices of
// device memory (R + (segments - 1) + C <= K): the most resident slices,
// then the longest segments.  Returns false when even R = 1 does not fit.
bool plan_factor_store(int64_t S, int64_t K, int64_t& R, int64_t& C) {
    if (K >= S) {
        R = S;
        C = 1;
        return true;
    }
    R = 0;
    C = 1;
    for (int64_t c = 1; c + 1 <= K; ++c) {
        // the largest r with r + ceil((S - r) / c) - 1 + c <= K (the left side
        // does not decrease with r)
        int64_t lo = 1, hi = S - 1, best = 0;
        while (lo <= hi) {
            const int64_t r = (lo + hi) / 2;
            if (r + (S - r + c - 1) / c - 1 + c <= K) {
                best = r;
                lo = r + 1;
            } else {
                hi = r - 1;
            }
        }
        if (best > R || (best == R && best > 0)) {
            R = best;
            C = c;
        }
    }
    return R >= 1;
}

// ---------------------------------------------------------------- FP32 + refinement driver
void run_refined(rgf_context* c, Work& w, const rgf_options& o, FactorStore* keep = nullptr, double extra_bytes = 0) {
    const int64_t N = w.N, L = w.L, S = L + 2;
    // right-hand sides: mL left columns, mR right columns, zero columns up to
    // a multiple of 4 (16-byte rows for the FP32 sweep GEMMs; the padding
    // solves H psi = 0 and stays zero)
    const int64_t mL = w.inL.size(), mR = w.inR.size(), m = (mL + mR + 3) / 4 * 4;
    // three real products per complex product: bit 0 in the slice inverses,
    // bit 1 in the refinement sweeps
    const int use3m = o.complex_3m & 1, sweep3m = (o.complex_3m >> 1) & 1;
    const SliceCfg cfg = slice_config(c, N, RGF_PRECISION_FP32_REFINED, use3m, o, w);
    const int b = cfg.b;
    hipStream_t s = c->stream;
    // factor store: up to S slices of N x N complex FP32, beside the FP64
    // right-hand sides, residuals and FP32 corrections (and extra_bytes the
    // caller needs after the solve)
    size_t freeb = 0, totalb = 0;
    RGF_HIP(hipMemGetInfo(&freeb, &totalb));
    const double slice_bytes = (double)N * N * 8;
    const double fixed = (double)S * N * m * (16 + 16 + 8) + slice_bytes +
                         (double)inverse_workspace_elems<float>(N, b) * sizeof(float) + extra_bytes;
    // buffers kept from an earlier solve on this handle are reused (or freed
    // and grown), so what they hold counts as available
    const double avail =
        (double)freeb + (double)held_bytes(c, {"factors", "factor_ck", "factor_seg", "psi", "res", "delta", "G0R32",
                                               "inv_ws32", "adj_y", "adj_r", "adj_d"});
    int64_t K = (int64_t)std::max(0.0, std::floor((0.92 * avail - fixed) / slice_bytes));
    if (o.factor_slots > 0) K = std::min<int64_t>(K, o.factor_slots);
    FactorStore local;
    FactorStore& fs = keep ? *keep : local;
    fs = FactorStore();
    if (!plan_factor_store(S, K, fs.R, fs.C))
        throw Error(RGF_ERR_OUT_OF_MEMORY, "mixed-precision factor store does not fit in device memory");
    fs.c = c;
    fs.w = &w;
    fs.N = N;
    fs.S = S;
    fs.nseg = fs.hybrid() ? (S - fs.R + fs.C - 1) / fs.C : 0;
    fs.b = b;
    fs.use3m = use3m;
    fs.ov = cfg.lookahead ? &c->ov : nullptr;
    fs.res = dbuf<float>(c, "factors", (size_t)fs.R * N * N * 2);
    if (fs.hybrid()) {
        fs.ck = fs.nseg > 1 ? dbuf<float>(c, "factor_ck", (size_t)(fs.nseg - 1) * N * N * 2) : nullptr;
        fs.seg = dbuf<float>(c, "factor_seg", (size_t)fs.C * N * N * 2);
    } else {
        release(c, "factor_ck");
        release(c, "factor_seg");
    }
    w.factors_resident = fs.R;
    fs.ws = dbuf<float>(c, "inv_ws32", inverse_workspace_elems<float>(N, b));
    fs.G0R32 = dmat<float>(c, "G0R32", N, N);
    convert<float, double>(s, N, N, w.G0L, fs.slot(fs.res, 0), 1.0);
    convert<float, double>(s, N, N, w.G0R, fs.G0R32, 1.0);
    reset_status(c);
    RGF_HIP(hipStreamSynchronize(s));
    const double t0 = now();
    for (int64_t n = 1; n < fs.R; ++n) fs.step(n, fs.slot(fs.res, n - 1), fs.slot(fs.res, n), w.timer);
    for (int64_t k = 0; k < fs.nseg; ++k) fs.build(k, true, w.timer);
    RGF_HIP(hipStreamSynchronize(s));
    w.t_slices = now() - t0;
    w.max_multiplier = check_status(c, false, false);
    const double Nd = (double)N;
    w.flops = (double)(L + 1) * 8 * Nd * Nd * Nd * (use3m ? 0.75 : 1.0);
    if (!fs.hybrid()) release(c, "G0R32");
    auto Fn = [&](int64_t n) { return fs.at(n); };

    // refinement: psi (FP64), r (FP64), d (FP32), slices of N x m
    const double t1 = now();
    const size_t SE = (size_t)N * m;
    double* psi = dbuf<double>(c, "psi", 2 * S * SE);
    double* r = dbuf<double>(c, "res", 2 * S * SE);
    float* d = dbuf<float>(c, "delta", 2 * S * SE);
    CMat<float> t = dmat<float>(c, "tmp32", N, m);
    auto P = [&](double* base, int64_t n) { return CMat<double>(base + n * SE, base + S * SE + n * SE, m); };
    auto D = [&](int64_t n) { return CMat<float>(d + n * SE, d + S * SE + n * SE, m); };
    RGF_HIP(hipMemsetAsync(psi, 0, 2 * S * SE * sizeof(double), s));
    const double b0 = std::max(max_abs(s, N * mL, w.BL.re, w.BL.im, c->d_scratch),
                               max_abs(s, N * mR, w.BR.re, w.BR.im, c->d_scratch));
    const double bmax = b0 == 0 ? 1.0 : b0;
    auto wk = splitk_ws<float>(c);
    auto wk64 = splitk_ws<double>(c);
    const int maxit = o.max_refine_iter > 0 ? o.max_refine_iter : 6;
    const double tol = o.refine_tol > 0 ? o.refine_tol : 1e-13;
    double rel = 1.0;
    int it = 0;
    int64_t skipped_cols = 0;  // sweep product columns left out in the first correction
    // optional stop on the S error: the boundary projections C psi of the
    // last two iterates and the ratio q of their residuals give the error of
    // the latest as q/(1-q) times their difference (linear convergence)
    const int64_t npo = (int64_t)w.outL.size(), nqo = (int64_t)w.outR.size();
    const bool s_stop = o.refine_s_tol > 0 && npo + nqo > 0;
    CMat<double> Tcur, Tprev;
    if (s_stop) {
        Tcur = dmat<double>(c, "Tcur", npo + nqo, m);
        Tprev = dmat<double>(c, "Tprev", npo + nqo, m);
    }
    auto project = [&](CMat<double> T) {
        if (npo) cgemm<double>(s, npo, m, N, w.CL, P(psi, 0), T, cd(1, 0), 0, 0, 0, 0, wk64.first, wk64.second);
        if (nqo)
            cgemm<double>(s, nqo, m, N, w.CR, P(psi, L + 1), T.offset(npo, 0), cd(1, 0), 0, 0, 0, 0, wk64.first,
                          wk64.second);
    };
    double rel_prev = 1.0;
    w.refine_s_err = 0;
    bool s_done = false;
    for (;; ++it) {
        double rmax;
        if (it == 0) {
            // psi = 0: r = b (B_L in slice 0, B_R in slice L+1), max|r| = bmax
            RGF_HIP(hipMemsetAsync(r, 0, 2 * S * SE * sizeof(double), s));
            const CMat<double> r0 = P(r, 0), r1 = P(r, L + 1);
            if (mL) {
                RGF_HIP(hipMemcpy2DAsync(r0.re, m * sizeof(double), w.BL.re, w.BL.ld * sizeof(double),
                                         mL * sizeof(double), N, hipMemcpyDeviceToDevice, s));
                RGF_HIP(hipMemcpy2DAsync(r0.im, m * sizeof(double), w.BL.im, w.BL.ld * sizeof(double),
                                         mL * sizeof(double), N, hipMemcpyDeviceToDevice, s));
            }
            if (mR) {
                RGF_HIP(hipMemcpy2DAsync(r1.re + mL, m * sizeof(double), w.BR.re, w.BR.ld * sizeof(double),
                                         mR * sizeof(double), N, hipMemcpyDeviceToDevice, s));
                RGF_HIP(hipMemcpy2DAsync(r1.im + mL, m * sizeof(double), w.BR.im, w.BR.ld * sizeof(double),
                                         mR * sizeof(double), N, hipMemcpyDeviceToDevice, s));
            }
            rmax = b0;
        } else {
            // r = b - H psi
            residual_stencil(s, N, L, m, psi, psi + S * SE, r, r + S * SE, w.eps_re, w.eps_im, w.epsL, w.epsR, w.k2,
                             w.A0, w.BL.re, w.BL.im, mL, w.BR.re, w.BR.im, mR);
            cgemm<double>(s, N, m, N, w.G0L, P(psi, 0), P(r, 0), cd(1, 0), 1, 0, 0, 0, wk64.first, wk64.second);
            cgemm<double>(s, N, m, N, w.G0R, P(psi, L + 1), P(r, L + 1), cd(1, 0), 1, 0, 0, 0, wk64.first,
                          wk64.second);
            rmax = max_abs(s, S * SE, r, r + S * SE, c->d_scratch);
        }
        rel = rmax / bmax;
        if (!std::isfinite(rel)) throw Error(RGF_ERR_NUMERICAL, "non-finite residual in refinement");
        if (s_stop && it >= 1) {
            project(Tcur);
            if (it >= 2) {
                // Tprev <- Tprev - Tcur; S = -2i T - delta, so |dS| = 2 |dT|
                accumulate<double>(s, npo + nqo, m, Tcur, Tprev, -1.0);
                const double dS = 2.0 * max_abs(s, (npo + nqo) * m, Tprev.re, Tprev.im, c->d_scratch);
                const double q = rel / rel_prev;
                if (q < 0.5) {
                    w.refine_s_err = 2.0 * dS * q / (1.0 - q);
                    s_done = w.refine_s_err <= o.refine_s_tol;
                }
            }
            copy_mat<double>(s, npo + nqo, m, Tcur, Tprev);
        }
        rel_prev = rel;
        if (rel <= tol || s_done || it >= maxit) break;
        const double sc = 1.0 / rmax;
        // forward sweep: d_n = g_n t_n, t_0 = r_0, t_{n+1} = r_{n+1} + d_n
        // (scaled by sc; each product's reduction also forms the next t).  In
        // the first correction r = b, so t_n and d_n for n < S-2 have only the
        // left-input columns: those products run on the first mf columns (the
        // others of d stay zero; t keeps its zero columns until the product at
        // S-2, which runs in full and adds B_R to the last t)
        const int64_t mf = it == 0 ? std::min<int64_t>(m, (mL + 3) / 4 * 4) : m;
        if (mf < m) RGF_HIP(hipMemsetAsync(d, 0, 2 * S * SE * sizeof(float), s));
        convert<float, double>(s, N, m, P(r, 0), t, sc);
        for (int64_t n = 0; n < S; ++n) {
            const CMat<double> rn = P(r, n + 1);
            sweep_forward(s, N, n + 2 < S ? mf : m, Fn(n), t, D(n), n + 1 < S ? &rn : nullptr, sc, sweep3m,
                          wk.first, wk.second);
            if (n + 2 < S) skipped_cols += m - mf;
        }
        // backward sweep: d_n += g_n d_{n+1}
        for (int64_t n = S - 2; n >= 0; --n)
            cgemm<float>(s, N, m, N, Fn(n), D(n + 1), D(n), cd(1, 0), 1, 0, 0, sweep3m, wk.first, wk.second);
        accumulate<float>(s, S * N, m, CMat<float>(d, d + S * SE, m), CMat<double>(psi, psi + S * SE, m), rmax);
    }
    w.refine_iters = it;
    w.refine_residual = rel;
    w.flops += ((double)it * (2.0 * S - 1) * m - (double)skipped_cols) * 8 * Nd * Nd * (sweep3m ? 0.75 : 1.0);
    w.factors_rebuilt = fs.rebuilt;
    w.flops += (double)fs.rebuilt * 8 * Nd * Nd * Nd * (use3m ? 0.75 : 1.0);
    // S blocks from the boundary slices: C_L psi_0, C_R psi_{L+1}
    const int64_t po = w.outL.size(), qo = w.outR.size();
    if (po) {
        CMat<double> T = dmat<double>(c, "T", po, m);
        cgemm<double>(s, po, m, N, w.CL, P(psi, 0), T, cd(1, 0), 0, 0, 0, 0, wk64.first, wk64.second);
        auto h = download(c, T, po, m);
        w.TLL.assign((size_t)po * mL, 0);
        w.TLR.assign((size_t)po * mR, 0);
        for (int64_t i = 0; i < po; ++i) {
            for (int64_t j = 0; j < mL; ++j) w.TLL[i * mL + j] = h[i * m + j];
            for (int64_t j = 0; j < mR; ++j) w.TLR[i * mR + j] = h[i * m + mL + j];
        }
    }
    if (qo) {
        CMat<double> T = dmat<double>(c, "T", qo, m);
        cgemm<double>(s, qo, m, N, w.CR, P(psi, L + 1), T, cd(1, 0), 0, 0, 0, 0, wk64.first, wk64.second);
        auto h = download(c, T, qo, m);
        w.TRL.assign((size_t)qo * mL, 0);
        w.TRR.assign((size_t)qo * mR, 0);
        for (int64_t i = 0; i < qo; ++i) {
            for (int64_t j = 0; j < mL; ++j) w.TRL[i * mL + j] = h[i * m + j];
            for (int64_t j = 0; j < mR; ++j) w.TRR[i * mR + j] = h[i * m + mL + j];
        }
    }
    if (w.fields) {
        // field profiles in the caller's order: inputs L then R, physical
        // slices 0..L+1 (the sweep may run mirrored)
        const int64_t nin = mL + mR;
        const int64_t nIL = w.mirrored ? mR : mL;  // physical left inputs
        std::vector<int32_t> col((size_t)nin);
        for (int64_t a = 0; a < nin; ++a) {
            const bool physR = a >= nIL;
            const int64_t k = physR ? a - nIL : a;
            const bool sweepL = w.mirrored ? physR : !physR;
            col[(size_t)a] = (int32_t)(sweepL ? k : mL + k);
        }
        int32_t* dcol = dbuf<int32_t>(c, "field_col", (size_t)nin);
        RGF_HIP(hipMemcpyAsync(dcol, col.data(), nin * sizeof(int32_t), hipMemcpyHostToDevice, s));
        const size_t cnt = (size_t)2 * N * (L + 2) * nin;
        double* pk = dbuf<double>(c, "field_pack", cnt);
        pack_fields(s, N, L, m, nin, psi, psi + S * SE, dcol, w.mirrored ? 1 : 0, pk);
        RGF_HIP(hipMemcpyAsync(w.fields, pk, cnt * sizeof(double), hipMemcpyDeviceToHost, s));
        RGF_HIP(hipStreamSynchronize(s));
        release(c, "field_pack");
    }
    w.t_refine = now() - t1;
    if (s_stop) {
        release(c, "Tcur");
        release(c, "Tprev");
    }
    if (rel > tol && !s_done) {
        char buf[160];
        snprintf(buf, sizeof buf, "mixed-precision refinement stopped at relative residual %.3e (tolerance %.1e)", rel,
                 tol);
        throw Error(RGF_ERR_NOT_CONVERGED, buf);
    }
}

void fill_info(rgf_info* info, const Channels& ch, const Selection& sel) {
    if (!info) return;
    memset(info, 0, sizeof *info);
    info->n_transverse = ch.N;
    info->n_slices = ch.L_slices;
    info->n_prop_L = ch.L.n_prop;
    info->n_prop_R = ch.R.n_prop;
    info->n_out = (int32_t)(sel.outL.size() + sel.outR.size());
    info->n_in = (int32_t)(sel.inL.size() + sel.inR.size());
}

// ---------------------------------------------------------------- adjoint gradient (FP64)
// T = sum_ij |t_ij|^2 with t = -2i C_R G_{L+1,0} B_L (sweep-side names).
// Forward: g_n = (D_n - g_{n-1})^-1 and z_n = g_n z_{n-1} (z_0 = G0_L B_L);
// the pairs (g_n, z_n) are kept at every `every`-th slice.  Backward, with
// each segment recomputed from its checkpoint:
//   psi_n = z_n + g_n psi_{n+1}    (psi_{L+1} = z_{L+1}: forward fields)
//   phi_n = phi_{n+1} g_n           (phi_{L+1} = t^H C_R g_{L+1}: adjoint rows)
// and dT/d eps_{m,n} from g = -2i k^2 sum_j psi_n[m,j] phi_n[j,m].
void run_gradient(rgf_context* c, Work& w, const rgf_options& o, int every, double& T, std::vector<double>& gre,
                  std::vector<double>& gim) {
    const int64_t N = w.N, L = w.L, m = w.inL.size(), no = w.outR.size();
    const SliceCfg cfg = slice_config(c, N, RGF_PRECISION_FP64, 0, o, w);
    const int b = cfg.b;
    hipStream_t s = c->stream;
    const int64_t C = (every <= 0 || every > L + 1) ? L + 2 : every;
    const int64_t nck = L / C + 1;  // checkpoints at n = 0, C, 2C, ... <= L
    const size_t NN = (size_t)N * N, NM = (size_t)N * m;
    double* ckg = dbuf<double>(c, "ck_g", 2 * NN * nck);
    double* ckz = dbuf<double>(c, "ck_z", 2 * NM * nck);
    double* sg = dbuf<double>(c, "seg_g", 2 * NN * std::max<int64_t>(C, 1));
    double* sz = dbuf<double>(c, "seg_z", 2 * NM * std::max<int64_t>(C, 1));
    auto mat = [&](double* base, size_t per, int64_t k, int64_t cols) {
        return CMat<double>(base + 2 * per * k, base + 2 * per * k + per, cols);
    };
    auto slot_g = [&](int64_t n) { return (n % C == 0) ? mat(ckg, NN, n / C, N) : mat(sg, NN, n % C, N); };
    auto slot_z = [&](int64_t n) { return (n % C == 0) ? mat(ckz, NM, n / C, m) : mat(sz, NM, n % C, m); };
    CMat<double> G = dmat<double>(c, "G", N, N);
    CMat<double> za = dmat<double>(c, "za", N, m), zb = dmat<double>(c, "zb", N, m);
    double* ws = dbuf<double>(c, "inv_ws", inverse_workspace_elems<double>(N, b));
    auto wk = splitk_ws<double>(c);
    auto dyson = [&](int64_t n, CMat<double> prev, CMat<double> out) {
        if (n <= L) {
            assemble_slice<double>(s, N, out, w.eps_re + (n - 1) * N, w.eps_im ? w.eps_im + (n - 1) * N : nullptr,
                                   cd(0, 0), w.k2, nullptr, w.A0, &prev);
        } else {
            copy_mat<double>(s, N, N, prev, out);
            assemble_slice<double>(s, N, out, nullptr, nullptr, w.epsR, w.k2, &w.G0R, w.A0);
        }
        invert_slice<double>(c, w, s, N, out, b, ws, 0, w.timer, cfg.lookahead ? &c->ov : nullptr);
    };
    reset_status(c);
    RGF_HIP(hipStreamSynchronize(s));
    const double t0 = now();
    // forward sweep
    copy_mat<double>(s, N, N, w.G0L, slot_g(0));
    cgemm<double>(s, N, m, N, w.G0L, w.BL, slot_z(0), cd(1, 0), 0, 0, 0, 0, wk.first, wk.second);
    copy_mat<double>(s, N, N, w.G0L, G);
    copy_mat<double>(s, N, m, slot_z(0), za);
    for (int64_t n = 1; n <= L + 1; ++n) {
        dyson(n, G, G);
        cgemm<double>(s, N, m, N, G, za, zb, cd(1, 0), 0, 0, 0, 0, wk.first, wk.second);
        std::swap(za, zb);
        if (n <= L && n % C == 0) {
            copy_mat<double>(s, N, N, G, slot_g(n));
            copy_mat<double>(s, N, m, za, slot_z(n));
        }
    }
    // t = -2i C_R z_{L+1}
    CMat<double> Tm = dmat<double>(c, "T", no, m);
    cgemm<double>(s, no, m, N, w.CR, za, Tm, cd(0, -2), 0, 0, 0, 0, wk.first, wk.second);
    std::vector<cd> t = download(c, Tm, no, m);
    T = 0;
    for (auto& v : t) T += std::norm(v);
    // phi_{L+1} = t^H C_R g_{L+1} (m x N)
    std::vector<cd> tH((size_t)m * no);
    for (int64_t i = 0; i < no; ++i)
        for (int64_t j = 0; j < m; ++j) tH[j * no + i] = std::conj(t[i * m + j]);
    CMat<double> TH = dmat<double>(c, "TH", m, no), W = dmat<double>(c, "W", m, N);
    upload(c, tH, m, no, TH);
    CMat<double> pha = dmat<double>(c, "pha", m, N), phb = dmat<double>(c, "phb", m, N);
    cgemm<double>(s, m, N, no, TH, w.CR, W, cd(1, 0), 0);
    cgemm<double>(s, m, N, N, W, G, pha, cd(1, 0), 0, 0, 0, 0, wk.first, wk.second);
    CMat<double> psa = dmat<double>(c, "psa", N, m), psb = dmat<double>(c, "psb", N, m);
    copy_mat<double>(s, N, m, za, psa);
    double* dg = dbuf<double>(c, "grad", 2 * (size_t)N * std::max<int64_t>(L, 1));
    // backward sweep
    int64_t cur = -1;
    for (int64_t n = L; n >= 1; --n) {
        const int64_t k = n / C;
        if (k != cur) {
            for (int64_t q = k * C + 1; q <= std::min(k * C + C - 1, L); ++q) {
                dyson(q, slot_g(q - 1), slot_g(q));
                cgemm<double>(s, N, m, N, slot_g(q), slot_z(q - 1), slot_z(q), cd(1, 0), 0, 0, 0, 0, wk.first,
                              wk.second);
            }
            cur = k;
        }
        copy_mat<double>(s, N, m, slot_z(n), psb);
        cgemm<double>(s, N, m, N, slot_g(n), psa, psb, cd(1, 0), 1, 0, 0, 0, wk.first, wk.second);
        std::swap(psa, psb);
        cgemm<double>(s, m, N, N, pha, slot_g(n), phb, cd(1, 0), 0, 0, 0, 0, wk.first, wk.second);
        std::swap(pha, phb);
        grad_contract(s, N, m, psa, pha, w.k2, dg + (n - 1) * N, dg + (size_t)N * L + (n - 1) * N);
    }
    RGF_HIP(hipStreamSynchronize(s));
    w.t_slices = now() - t0;
    w.max_multiplier = check_status(c);
    const double Nd = (double)N;
    const int64_t recomputed = L - (L / C);
    w.flops = (double)(L + 1 + recomputed) * (8 * Nd * Nd * Nd + 8 * Nd * Nd * m) + (double)L * 16 * Nd * Nd * m;
    gre.assign((size_t)N * L, 0);
    gim.assign((size_t)N * L, 0);
    if (L > 0) {
        RGF_HIP(hipMemcpyAsync(gre.data(), dg, sizeof(double) * N * L, hipMemcpyDeviceToHost, s));
        RGF_HIP(hipMemcpyAsync(gim.data(), dg + (size_t)N * L, sizeof(double) * N * L, hipMemcpyDeviceToHost, s));
        RGF_HIP(hipStreamSynchronize(s));
    }
}

// ---------------------------------------------------------------- adjoint gradient (FP32 + refinement)
// A0 transposed (CSR): the row (adjoint) residual needs A0's columns.
HostCSR transpose_csr(const HostCSR& a) {
    HostCSR t;
    t.n = a.n;
    t.rowptr.assign((size_t)a.n + 1, 0);
    for (int32_t q : a.col) ++t.rowptr[(size_t)q + 1];
    for (int32_t i = 0; i < a.n; ++i) t.rowptr[(size_t)i + 1] += t.rowptr[(size_t)i];
    t.col.resize(a.col.size());
    t.vre.resize(a.col.size());
    t.vim.resize(a.col.size());
    std::vector<int32_t> at(t.rowptr.begin(), t.rowptr.end() - 1);
    for (int32_t i = 0; i < a.n; ++i)
        for (int32_t q = a.rowptr[(size_t)i]; q < a.rowptr[(size_t)i + 1]; ++q) {
            const int32_t k = a.col[(size_t)q], dst = at[(size_t)k]++;
            t.col[(size_t)dst] = i;
            t.vre[(size_t)dst] = a.vre[(size_t)q];
            t.vim[(size_t)dst] = a.vim[(size_t)q];
        }
    return t;
}

// T = sum |t_ij|^2 and dT/d eps with the mixed path (sweep-side names as in
// run_gradient).  The forward fields psi = H^-1 b (b = B_L at slice 0) are
// run_refined's refined solution (named buffer "psi", S x N x m).  The
// adjoint rows y = c H^-1, c = t^H C_R at slice L+1 (so that y_n =
// t^H C_R G_{L+1,n}), are refined the same way in row form through the same
// stored FP32 inverses: forward z_n = (r_n + z_{n-1}) g_n, backward
// d_n = z_n + d_{n+1} g_n, with the FP64 residual r = c - y H (the stencil
// transposed, the lead terms y_0 G0L and y_{L+1} G0R).  The gradient is the
// contraction of run_gradient: u_i = k^2 sum_j psi_n[i,j] y_n[j,i].
void run_gradient_refined(rgf_context* c, Work& w, const rgf_options& o, double& T, std::vector<double>& gre,
                          std::vector<double>& gim) {
    const int64_t N = w.N, L = w.L, S = L + 2;
    const int64_t mL = (int64_t)w.inL.size(), m = (mL + 3) / 4 * 4, no = (int64_t)w.outR.size();
    // the factor store is sized with room for the adjoint rows, residuals
    // and corrections allocated below
    FactorStore fs;
    run_refined(c, w, o, &fs, (double)S * N * m * (16 + 16 + 8));
    const int64_t rebuilt0 = fs.rebuilt;
    hipStream_t s = c->stream;
    const double t0 = now();
    T = 0;
    std::vector<cd> tH((size_t)m * no, cd(0, 0));  // rows >= mL stay zero
    for (int64_t i = 0; i < no; ++i)
        for (int64_t j = 0; j < mL; ++j) {
            const cd t = cd(0, -2) * w.TRL[(size_t)(i * mL + j)];
            T += std::norm(t);
            tH[(size_t)(j * no + i)] = std::conj(t);
        }
    CMat<double> TH = dmat<double>(c, "TH", m, no), Cadj = dmat<double>(c, "adj_c", m, N);
    upload(c, tH, m, no, TH);
    cgemm<double>(s, m, N, no, TH, w.CR, Cadj, cd(1, 0), 0);
    DevCSR AT;
    {
        const HostCSR h = transpose_csr(slice_stencil(*w.ch));
        AT.n = h.n;
        AT.nnz = (int32_t)h.col.size();
        AT.rowptr = dbuf<int32_t>(c, "AT_rowptr", h.rowptr.size());
        AT.col = dbuf<int32_t>(c, "AT_col", h.col.size());
        AT.vre = dbuf<double>(c, "AT_vre", h.vre.size());
        AT.vim = dbuf<double>(c, "AT_vim", h.vim.size());
        RGF_HIP(hipMemcpyAsync(AT.rowptr, h.rowptr.data(), h.rowptr.size() * 4, hipMemcpyHostToDevice, s));
        RGF_HIP(hipMemcpyAsync(AT.col, h.col.data(), h.col.size() * 4, hipMemcpyHostToDevice, s));
        RGF_HIP(hipMemcpyAsync(AT.vre, h.vre.data(), h.vre.size() * 8, hipMemcpyHostToDevice, s));
        RGF_HIP(hipMemcpyAsync(AT.vim, h.vim.data(), h.vim.size() * 8, hipMemcpyHostToDevice, s));
        RGF_HIP(hipStreamSynchronize(s));
    }
    const size_t SE = (size_t)N * m;
    auto Fn = [&](int64_t n) { return fs.at(n); };
    double* y = dbuf<double>(c, "adj_y", 2 * S * SE);
    double* r = dbuf<double>(c, "adj_r", 2 * S * SE);
    float* d = dbuf<float>(c, "adj_d", 2 * S * SE);
    CMat<float> tt = dmat<float>(c, "adj_t", m, N);
    auto Y = [&](double* base, int64_t n) { return CMat<double>(base + n * SE, base + S * SE + n * SE, N); };
    auto D = [&](int64_t n) { return CMat<float>(d + n * SE, d + S * SE + n * SE, N); };
    auto wk = splitk_ws<float>(c);
    auto wk64 = splitk_ws<double>(c);
    const int sweep3m = (o.complex_3m >> 1) & 1;
    const double cmax0 = max_abs(s, (int64_t)SE, Cadj.re, Cadj.im, c->d_scratch);
    const double cmax = cmax0 == 0 ? 1.0 : cmax0;
    const int maxit = o.max_refine_iter > 0 ? o.max_refine_iter : 6;
    const double tol = o.refine_tol > 0 ? o.refine_tol : 1e-13;
    double rel = 1.0;
    int it = 0;
    for (;; ++it) {
        double rmax;
        if (it == 0) {
            rmax = cmax0;  // y = 0: r = c, at slice L+1 only
        } else {
            residual_rows(s, N, L, m, y, y + S * SE, r, r + S * SE, w.eps_re, w.eps_im, w.epsL, w.epsR, w.k2, AT,
                          Cadj.re, Cadj.im);
            cgemm<double>(s, m, N, N, Y(y, 0), w.G0L, Y(r, 0), cd(1, 0), 1, 0, 0, 0, wk64.first, wk64.second);
            cgemm<double>(s, m, N, N, Y(y, L + 1), w.G0R, Y(r, L + 1), cd(1, 0), 1, 0, 0, 0, wk64.first,
                          wk64.second);
            rmax = max_abs(s, S * (int64_t)SE, r, r + S * SE, c->d_scratch);
        }
        rel = rmax / cmax;
        if (!std::isfinite(rel)) throw Error(RGF_ERR_NUMERICAL, "non-finite residual in the adjoint refinement");
        if (rel <= tol || it >= maxit || rmax == 0) break;
        const double sc = 1.0 / rmax;
        if (it == 0) {
            // r vanishes before slice L+1: z_n = 0 there, z_{L+1} = (c / rmax) g_{L+1}
            RGF_HIP(hipMemsetAsync(d, 0, 2 * S * SE * sizeof(float), s));
            convert<float, double>(s, m, N, Cadj, tt, sc);
            cgemm<float>(s, m, N, N, tt, Fn(S - 1), D(S - 1), cd(1, 0), 0, 0, 0, sweep3m, wk.first, wk.second);
        } else {
            convert<float, double>(s, m, N, Y(r, 0), tt, sc);
            for (int64_t n = 0; n < S; ++n) {
                cgemm<float>(s, m, N, N, tt, Fn(n), D(n), cd(1, 0), 0, 0, 0, sweep3m, wk.first, wk.second);
                if (n + 1 < S) {
                    convert<float, double>(s, m, N, Y(r, n + 1), tt, sc);
                    add_mat<float>(s, m, N, tt, D(n), tt);
                }
            }
        }
        for (int64_t n = S - 2; n >= 0; --n)
            cgemm<float>(s, m, N, N, D(n + 1), Fn(n), D(n), cd(1, 0), 1, 0, 0, sweep3m, wk.first, wk.second);
        if (it == 0) RGF_HIP(hipMemsetAsync(y, 0, 2 * S * SE * sizeof(double), s));
        accumulate<float>(s, S * m, N, CMat<float>(d, d + S * SE, N), CMat<double>(y, y + S * SE, N), rmax);
    }
    if (rel > tol && it >= maxit) {
        char buf[160];
        snprintf(buf, sizeof buf, "adjoint refinement stopped at relative residual %.3e (tolerance %.1e)", rel, tol);
        throw Error(RGF_ERR_NOT_CONVERGED, buf);
    }
    // gradient: psi_n (N x m, ld m) and y_n (m x N, ld N), slices 1..L
    double* psi = (double*)c->bufs["psi"].p;
    double* dg = dbuf<double>(c, "grad", 2 * (size_t)N * std::max<int64_t>(L, 1));
    for (int64_t n = 1; n <= L; ++n)
        grad_contract(s, N, m, CMat<double>(psi + n * SE, psi + S * SE + n * SE, m), Y(y, n), w.k2,
                      dg + (n - 1) * N, dg + (size_t)N * L + (n - 1) * N);
    gre.assign((size_t)N * L, 0);
    gim.assign((size_t)N * L, 0);
    if (L > 0) {
        RGF_HIP(hipMemcpyAsync(gre.data(), dg, sizeof(double) * N * L, hipMemcpyDeviceToHost, s));
        RGF_HIP(hipMemcpyAsync(gim.data(), dg + (size_t)N * L, sizeof(double) * N * L, hipMemcpyDeviceToHost, s));
    }
    RGF_HIP(hipStreamSynchronize(s));
    w.t_refine += now() - t0;
    w.refine_iters += it;  // forward and adjoint correction sweeps
    w.factors_rebuilt = fs.rebuilt;
    w.flops += (double)(fs.rebuilt - rebuilt0) * 8 * (double)N * N * N * ((o.complex_3m & 1) ? 0.75 : 1.0);
    if (fs.hybrid()) release(c, "G0R32");
}

void gradient_one(rgf_context* c, const rgf_problem* p, const rgf_selection* s, const rgf_options* opt, int every,
                  double* T_out, double* gre_out, double* gim_out, rgf_info* info) {
    const double t0 = now();
    rgf_options o;
    rgf_default_options(&o);
    if (opt) o = *opt;
    if (o.precision != RGF_PRECISION_FP64 && o.precision != RGF_PRECISION_FP32_REFINED)
        throw Error(RGF_ERR_INVALID_ARGUMENT, "unknown precision");
    Channels ch = setup_channels(p);
    Selection sel = make_selection(s, ch);
    fill_info(info, ch, sel);
    const bool l2r = !sel.inL.empty() && sel.inR.empty() && !sel.outR.empty() && sel.outL.empty();
    const bool r2l = sel.inL.empty() && !sel.inR.empty() && sel.outR.empty() && !sel.outL.empty();
    if (!l2r && !r2l)
        throw Error(RGF_ERR_INVALID_ARGUMENT,
                    "the transmission gradient needs inputs on one side and outputs on the other side");
    RGF_HIP(hipSetDevice(c->device));
    Work w;
    std::unique_ptr<KernelTimer> timer;
    if (o.profile) {
        timer.reset(new KernelTimer(&c->timer_events));
        w.timer = timer.get();
    }
    prepare(c, p, ch, sel, w, r2l ? 1 : 0);
    w.t_setup = now() - t0;
    double T = 0;
    std::vector<double> gre, gim;
    with_pivoting_fallback(o, w, [&] {
        if (o.precision == RGF_PRECISION_FP64)
            run_gradient(c, w, o, every, T, gre, gim);
        else
            run_gradient_refined(c, w, o, T, gre, gim);
    });
    if (T_out) *T_out = T;
    const int64_t N = w.N, L = w.L;
    for (int64_t n = 0; n < L; ++n) {
        const int64_t un = w.mirrored ? L - 1 - n : n;  // user slice index
        for (int64_t i = 0; i < N; ++i) {
            if (gre_out) gre_out[un * N + i] = gre[n * N + i];
            if (gim_out) gim_out[un * N + i] = gim[n * N + i];
        }
    }
    if (info) {
        info->swept_right_to_left = w.mirrored;
        info->seconds_setup = w.t_setup;
        info->seconds_slices = w.t_slices;
        info->seconds_refine = w.t_refine;
        info->refine_iters = w.refine_iters;
        info->refine_residual = w.refine_residual;
        info->flops_slices = w.flops;
        info->max_multiplier = w.max_multiplier;
        info->pivoted = w.pivot_all;
        info->block_size = w.block_size;
        info->autotuned = w.tuned;
        info->factor_slices_resident = w.factors_resident;
        info->factor_slices_rebuilt = w.factors_rebuilt;
        if (timer) {
            info->kernel_seconds = timer->seconds();
            info->kernel_flops = timer->flops;
            info->kernel_launches = timer->launches;
        }
        info->seconds_total = now() - t0;
    }
}

// Fisher-Lee relation S = -2i C G B minus the Kronecker delta of matching
// indices (cal_smatrix_RGF.m:279, 286-289), written column-major interleaved.
void write_smatrix(const Work& w, const Selection& sel, double* S_out) {
    const int64_t nOL = sel.outL.size(), nOR = sel.outR.size(), nIL = sel.inL.size(), nIR = sel.inR.size();
    const int64_t nout = nOL + nOR, nin = nIL + nIR;
    // block of (row side, col side) in user (original) sides
    auto block = [&](bool rowR, bool colR, int64_t i, int64_t j) -> cd {
        bool sr = rowR, sc = colR;
        if (w.mirrored) {
            sr = !sr;
            sc = !sc;
        }
        const std::vector<cd>* T;
        int64_t ncols;
        if (!sr && !sc) {
            T = &w.TLL;
            ncols = w.inL.size();
        } else if (!sr && sc) {
            T = &w.TLR;
            ncols = w.inR.size();
        } else if (sr && !sc) {
            T = &w.TRL;
            ncols = w.inL.size();
        } else {
            T = &w.TRR;
            ncols = w.inR.size();
        }
        if (T->empty()) return cd(0, 0);
        return (*T)[i * ncols + j];
    };
    for (int64_t j = 0; j < nin; ++j) {
        const bool colR = j >= nIL;
        const int64_t jj = colR ? j - nIL : j;
        for (int64_t i = 0; i < nout; ++i) {
            const bool rowR = i >= nOL;
            const int64_t ii = rowR ? i - nOL : i;
            const cd v = cd(0, -2) * block(rowR, colR, ii, jj);
            S_out[2 * (i + j * nout)] = v.real();
            S_out[2 * (i + j * nout) + 1] = v.imag();
        }
    }
    // delta via intersect: first occurrence of each common index value
    auto sub_delta = [&](const std::vector<int32_t>& out, const std::vector<int32_t>& in, int64_t ro, int64_t co) {
        std::map<int32_t, int64_t> first_in;
        for (int64_t q = (int64_t)in.size() - 1; q >= 0; --q) first_in[in[q]] = q;
        std::map<int32_t, int64_t> first_out;
        for (int64_t q = (int64_t)out.size() - 1; q >= 0; --q) first_out[out[q]] = q;
        for (auto& e : first_out) {
            auto it = first_in.find(e.first);
            if (it == first_in.end()) continue;
            S_out[2 * ((ro + e.second) + (co + it->second) * nout)] -= 1.0;
        }
    };
    sub_delta(sel.outL, sel.inL, 0, 0);
    sub_delta(sel.outR, sel.inR, nOL, nIL);
}

void solve_one(rgf_context* c, const rgf_problem* p, const rgf_selection* s, const rgf_options* opt, double* S_out,
               rgf_info* info, double* fields = nullptr) {
    const double t0 = now();
    rgf_options o;
    rgf_default_options(&o);
    if (opt) o = *opt;
    if (o.precision != RGF_PRECISION_FP64 && o.precision != RGF_PRECISION_FP32_REFINED)
        throw Error(RGF_ERR_INVALID_ARGUMENT, "unknown precision");
    Channels ch = setup_channels(p);
    Selection sel = make_selection(s, ch);
    fill_info(info, ch, sel);
    const int64_t nout = sel.outL.size() + sel.outR.size(), nin = sel.inL.size() + sel.inR.size();
    if (fields && o.precision != RGF_PRECISION_FP32_REFINED)
        throw Error(RGF_ERR_INVALID_ARGUMENT,
                    "field profiles come from the refined solution: use precision RGF_PRECISION_FP32_REFINED");
    if (nout == 0 || nin == 0) return;  // cal_smatrix_RGF.m:181-184
    if (!S_out) throw Error(RGF_ERR_INVALID_ARGUMENT, "S_out is NULL");
    RGF_HIP(hipSetDevice(c->device));
    Work w;
    std::unique_ptr<KernelTimer> timer;
    if (o.profile) {
        timer.reset(new KernelTimer(&c->timer_events));
        w.timer = timer.get();
    }
    prepare(c, p, ch, sel, w);
    w.fields = fields;
    w.t_setup = now() - t0;
    with_pivoting_fallback(o, w, [&] {
        if (o.precision == RGF_PRECISION_FP64)
            run_carry(c, w, o);
        else
            run_refined(c, w, o);
    });
    write_smatrix(w, sel, S_out);
    if (info) {
        info->swept_right_to_left = w.mirrored;
        info->refine_iters = w.refine_iters;
        info->refine_residual = w.refine_residual;
        info->refine_s_err_estimate = w.refine_s_err;
        info->max_multiplier = w.max_multiplier;
        info->pivoted = w.pivot_all;
        info->block_size = w.block_size;
        info->autotuned = w.tuned;
        info->factor_slices_resident = w.factors_resident;
        info->factor_slices_rebuilt = w.factors_rebuilt;
        info->seconds_setup = w.t_setup;
        info->seconds_slices = w.t_slices;
        info->seconds_refine = w.t_refine;
        info->flops_slices = w.flops;
        if (timer) {
            info->kernel_seconds = timer->seconds();
            info->kernel_flops = timer->flops;
            info->kernel_launches = timer->launches;
        }
        info->seconds_total = now() - t0;
    }
}

rgf_status guard(rgf_context* c, const std::function<void()>& f) {
    try {
        f();
        if (c) c->err.clear();
        return RGF_OK;
    } catch (const Error& e) {
        if (c) c->err = e.what();
        std::lock_guard<std::mutex> lk(g_err_mu);
        g_last_error = e.what();
        return e.status;
    } catch (const std::bad_alloc&) {
        if (c) c->err = "host out of memory";
        std::lock_guard<std::mutex> lk(g_err_mu);
        g_last_error = "host out of memory";
        return RGF_ERR_OUT_OF_MEMORY;
    } catch (const std::exception& e) {
        if (c) c->err = e.what();
        std::lock_guard<std::mutex> lk(g_err_mu);
        g_last_error = e.what();
        return RGF_ERR_INVALID_ARGUMENT;
    }
}

// interleaved column-major (host) <-> planar row-major (device)
void upload_colmajor(rgf_context* c, const double* A, int64_t rows, int64_t cols, CMat<double> d) {
    std::vector<cd> h((size_t)rows * cols);
    for (int64_t i = 0; i < rows; ++i)
        for (int64_t j = 0; j < cols; ++j) h[i * cols + j] = cd(A[2 * (i + j * rows)], A[2 * (i + j * rows) + 1]);
    upload(c, h, rows, cols, d);
}

void download_colmajor(rgf_context* c, CMat<double> d, int64_t rows, int64_t cols, double* A) {
    auto h = download(c, d, rows, cols);
    for (int64_t i = 0; i < rows; ++i)
        for (int64_t j = 0; j < cols; ++j) {
            A[2 * (i + j * rows)] = h[i * cols + j].real();
            A[2 * (i + j * rows) + 1] = h[i * cols + j].imag();
        }
}

}  // namespace rgf

// =================================================================== C ABI
extern "C" {

int32_t rgf_abi_version(void) { return RGF_ABI_VERSION; }

const char* rgf_status_string(rgf_status s) {
    switch (s) {
        case RGF_OK: return "ok";
        case RGF_ERR_INVALID_ARGUMENT: return "invalid argument";
        case RGF_ERR_BOUNDARY_CONDITION: return "invalid boundary condition";
        case RGF_ERR_CHANNEL_INDEX: return "channel index out of range";
        case RGF_ERR_OUT_OF_MEMORY: return "out of memory";
        case RGF_ERR_DEVICE: return "device error";
        case RGF_ERR_NUMERICAL: return "numerical failure";
        case RGF_ERR_NOT_CONVERGED: return "refinement not converged";
        case RGF_ERR_BUFFER_TOO_SMALL: return "buffer too small";
    }
    return "unknown status";
}

void rgf_default_options(rgf_options* o) {
    if (!o) return;
    memset(o, 0, sizeof *o);
    o->precision = RGF_PRECISION_FP64;
    o->max_refine_iter = 6;
    o->refine_tol = 1e-13;
    o->block_size = 0;
    o->complex_3m = 3;  // three real products in the slice inverses and the refinement sweeps
    o->autotune = 1;
}

rgf_status rgf_create(rgf_handle* h, int32_t device) {
    return guard(nullptr, [&] {
        if (!h) throw Error(RGF_ERR_INVALID_ARGUMENT, "handle pointer is NULL");
        *h = nullptr;
        int count = 0;
        if (hipGetDeviceCount(&count) != hipSuccess || count == 0) throw Error(RGF_ERR_DEVICE, "no HIP device");
        std::unique_ptr<rgf_context> c(new rgf_context);
        if (device < 0) RGF_HIP(hipGetDevice(&device));
        if (device >= count) throw Error(RGF_ERR_DEVICE, "device index out of range");
        c->device = device;
        RGF_HIP(hipSetDevice(device));
        RGF_HIP(hipStreamCreateWithFlags(&c->own, hipStreamNonBlocking));
        c->stream = c->own;
        RGF_HIP(hipStreamCreateWithFlags(&c->ov.s2, hipStreamNonBlocking));
        RGF_HIP(hipStreamCreateWithFlags(&c->ov.s3, hipStreamNonBlocking));
        // the events order kernels of one device (a device-scope release is
        // enough); they may complete with a kernel launch (hipExtLaunchKernel)
        for (auto& e : c->ov.ev) RGF_HIP(hipEventCreateWithFlags(&e, hipEventReleaseToDevice));
        RGF_HIP(hipMalloc(&c->d_status, 256));
        RGF_HIP(hipMalloc(&c->d_scratch, 256));
        c->scratch = noalloc_scratch(device);  // once per device, with it current
        *h = c.release();
    });
}

rgf_status rgf_destroy(rgf_handle h) {
    if (!h) return RGF_OK;
    for (auto w : h->workers) (void)rgf_destroy(w);
    h->workers.clear();
    rgf_status st = guard(nullptr, [&] {
        RGF_HIP(hipSetDevice(h->device));
        RGF_HIP(hipStreamSynchronize(h->stream));
        for (auto& e : h->bufs)
            if (e.second.p) RGF_HIP(hipFree(e.second.p));
        h->bufs.clear();
        if (h->d_status) RGF_HIP(hipFree(h->d_status));
        if (h->d_scratch) RGF_HIP(hipFree(h->d_scratch));
        if (h->rb) rocblas_ok(rocblas_destroy_handle((rocblas_handle)h->rb), "rocblas_destroy_handle");
        if (h->own) RGF_HIP(hipStreamDestroy(h->own));
        for (auto& e : h->ov.ev)
            if (e) RGF_HIP(hipEventDestroy(e));
        for (auto e : h->timer_events) RGF_HIP(hipEventDestroy(e));
        h->timer_events.clear();
        if (h->ov.s2) RGF_HIP(hipStreamDestroy(h->ov.s2));
        if (h->ov.s3) RGF_HIP(hipStreamDestroy(h->ov.s3));
    });
    delete h;
    return st;
}

const char* rgf_last_error(rgf_handle h) {
    if (h) return h->err.c_str();
    std::lock_guard<std::mutex> lk(g_err_mu);
    return g_last_error.c_str();
}

rgf_status rgf_set_stream(rgf_handle h, void* stream) {
    return guard(h, [&] {
        if (!h) throw Error(RGF_ERR_INVALID_ARGUMENT, "NULL handle");
        h->stream = stream ? (hipStream_t)stream : h->own;
    });
}

rgf_status rgf_trim(rgf_handle h) {
    return guard(h, [&] {
        if (!h) throw Error(RGF_ERR_INVALID_ARGUMENT, "NULL handle");
        RGF_HIP(hipSetDevice(h->device));
        RGF_HIP(hipStreamSynchronize(h->stream));
        for (auto& e : h->bufs)
            if (e.second.p) RGF_HIP(hipFree(e.second.p));
        h->bufs.clear();
    });
}

rgf_status rgf_query(const rgf_problem* p, const rgf_selection* s, int32_t* nL, int32_t* nR, int32_t* nout,
                     int32_t* nin) {
    return guard(nullptr, [&] {
        Channels ch = setup_channels(p, false);
        if (nL) *nL = ch.L.n_prop;
        if (nR) *nR = ch.R.n_prop;
        if (s) {
            Selection sel = make_selection(s, ch);
            if (nout) *nout = (int32_t)(sel.outL.size() + sel.outR.size());
            if (nin) *nin = (int32_t)(sel.inL.size() + sel.inR.size());
        }
    });
}

rgf_status rgf_setup_channels(const rgf_problem* p, rgf_channels* out) {
    return guard(nullptr, [&] {
        if (!out) throw Error(RGF_ERR_INVALID_ARGUMENT, "channels output is NULL");
        Channels ch = setup_channels(p, false);
        const int64_t N = ch.N;
        for (int64_t a = 0; a < N; ++a) {
            if (out->kt1) out->kt1[a] = ch.kt1[a];
            if (out->kt2 && ch.ndim == 3) out->kt2[a] = ch.kt2[a];
            if (out->kl_L) {
                out->kl_L[2 * a] = ch.L.kl[a].real();
                out->kl_L[2 * a + 1] = ch.L.kl[a].imag();
            }
            if (out->kl_R) {
                out->kl_R[2 * a] = ch.R.kl[a].real();
                out->kl_R[2 * a + 1] = ch.R.kl[a].imag();
            }
        }
        const Side* sides[2] = {&ch.L, &ch.R};
        int32_t* ip[2] = {out->ind_prop_L, out->ind_prop_R};
        double* mu[2] = {out->sqrt_mu_L, out->sqrt_mu_R};
        int32_t* cj[2] = {out->ind_prop_conj_L, out->ind_prop_conj_R};
        for (int k = 0; k < 2; ++k)
            for (int32_t q = 0; q < sides[k]->n_prop; ++q) {
                if (ip[k]) ip[k][q] = sides[k]->ind_prop[q] + 1;
                if (mu[k]) {
                    mu[k][2 * q] = sides[k]->sqrt_mu[q].real();
                    mu[k][2 * q + 1] = sides[k]->sqrt_mu[q].imag();
                }
                if (cj[k] && sides[k]->has_conj) cj[k][q] = sides[k]->conj[q] + 1;
            }
        out->n_prop_L = ch.L.n_prop;
        out->n_prop_R = ch.R.n_prop;
        out->has_conj = ch.L.has_conj ? 1 : 0;
    });
}

rgf_status rgf_transverse_profiles(const rgf_problem* p, int64_t count, const double* k1, const double* k2,
                                   double* phi_out) {
    return guard(nullptr, [&] {
        validate_problem(p);
        if (count < 0) throw Error(RGF_ERR_INVALID_ARGUMENT, "count must be non-negative");
        if (count > 0 && (!k1 || !phi_out || (p->ndim == 3 && !k2)))
            throw Error(RGF_ERR_INVALID_ARGUMENT, "wavenumber or output array is NULL");
        const int64_t n1 = p->dims[0], n2 = p->ndim == 3 ? p->dims[1] : 1;
        const int b1 = p->bc[0], b2 = p->ndim == 3 ? p->bc[1] : RGF_BC_DIRICHLET;
        const double s = mode_norm(n1, b1) * (p->ndim == 3 ? mode_norm(n2, b2) : 1.0);
        std::vector<cd> f2((size_t)n2);
        for (int64_t c = 0; c < count; ++c) {
            for (int64_t y = 0; y < n2; ++y) f2[y] = p->ndim == 3 ? mode_shape(b2, y, k2[c]) * s : cd(s, 0);
            double* col = phi_out + 2 * c * n1 * n2;
            for (int64_t y = 0; y < n2; ++y)
                for (int64_t x = 0; x < n1; ++x) {
                    const cd v = mode_shape(b1, x, k1[c]) * f2[y];  // x fast (setup_channels_3D.m:118,123)
                    col[2 * (x + n1 * y)] = v.real();
                    col[2 * (x + n1 * y) + 1] = v.imag();
                }
        }
    });
}

rgf_status rgf_build_laplacian_1d(int64_t n, int32_t bc, double ka, double* A) {
    return guard(nullptr, [&] {
        if (n < 0) throw Error(RGF_ERR_INVALID_ARGUMENT, "The number of sites n must be a non-negative integer scalar");
        if (bc < RGF_BC_BLOCH || bc > RGF_BC_NEUMANN_DIRICHLET)
            throw Error(RGF_ERR_BOUNDARY_CONDITION, "Boundary condition BC is not an eligible option");
        if (!A) throw Error(RGF_ERR_INVALID_ARGUMENT, "output is NULL");
        memset(A, 0, sizeof(double) * 2 * n * n);
        for (auto& e : laplacian_1d(n, bc, ka)) {
            A[2 * (e.first.first + e.first.second * n)] = e.second.real();
            A[2 * (e.first.first + e.first.second * n) + 1] = e.second.imag();
        }
    });
}

rgf_status rgf_solve(rgf_handle h, const rgf_problem* p, const rgf_selection* s, const rgf_options* opt,
                     double* S_out, rgf_info* info) {
    return guard(h, [&] {
        if (!h) throw Error(RGF_ERR_INVALID_ARGUMENT, "NULL handle");
        solve_one(h, p, s, opt, S_out, info);
    });
}

rgf_status rgf_solve_fields(rgf_handle h, const rgf_problem* p, const rgf_selection* s, const rgf_options* opt,
                            double* S_out, double* fields_out, rgf_info* info) {
    return guard(h, [&] {
        if (!h) throw Error(RGF_ERR_INVALID_ARGUMENT, "NULL handle");
        if (!fields_out) throw Error(RGF_ERR_INVALID_ARGUMENT, "fields_out is NULL");
        solve_one(h, p, s, opt, S_out, info, fields_out);
    });
}

rgf_status rgf_solve_batched(rgf_handle h, int32_t count, const rgf_problem* p, const rgf_selection* s,
                             const rgf_options* opt, double* const* S_out, rgf_info* info) {
    return guard(h, [&] {
        if (!h) throw Error(RGF_ERR_INVALID_ARGUMENT, "NULL handle");
        if (count < 0 || (count > 0 && (!p || !s || !S_out))) throw Error(RGF_ERR_INVALID_ARGUMENT, "bad batch");
        if (count == 0) return;
        rgf_options o;
        rgf_default_options(&o);
        if (opt) o = *opt;
        std::vector<rgf_status> sts((size_t)count, RGF_OK);
        std::vector<std::string> msgs((size_t)count);
        // the lockstep engine takes the small mixed-precision problems
        std::vector<int32_t> eng, rest;
        for (int32_t q = 0; q < count; ++q) (batch_engine_takes(&p[q], o) ? eng : rest).push_back(q);
        if (!eng.empty()) solve_batched_engine(h, eng, p, s, o, S_out, info, sts, msgs);
        // problems whose slice inverses the engine's kernels refused (pivoting
        // inside each diagonal block only) run again on the worker contexts
        // with the inverses pivoted over all rows
        rgf_options opiv = o;
        opiv.pivoting = 1;
        std::vector<char> pivot_first((size_t)count, 0);
        if (o.pivoting == 0)
            for (int32_t q : eng)
                if (sts[(size_t)q] == RGF_ERR_NUMERICAL || sts[(size_t)q] == RGF_ERR_NOT_CONVERGED) {
                    pivot_first[(size_t)q] = 1;
                    sts[(size_t)q] = RGF_OK;
                    msgs[(size_t)q].clear();
                    rest.push_back(q);
                }
        const int32_t nrest = (int32_t)rest.size();
        const int want = o.batch_concurrency > 0 ? o.batch_concurrency : 8;
        const int nw = std::max(1, std::min<int>(want, nrest));
        while ((int)h->workers.size() < nw) {
            rgf_handle w = nullptr;
            const rgf_status cs = rgf_create(&w, h->device);
            if (cs != RGF_OK) throw Error(cs, std::string("batched solve: worker context: ") + rgf_last_error(nullptr));
            h->workers.push_back(w);
        }
        // largest problems first (N^2 L): balances the workers and lets each
        // worker's workspace reach its final size on its first problem
        std::vector<int32_t> order = rest;
        std::vector<double> cost((size_t)count);
        for (int32_t q : rest) {
            const int nd = p[q].ndim == 3 ? 3 : 2;  // invalid ndim is reported by the solve itself
            double n = 1;
            for (int d = 0; d < nd - 1; ++d) n *= (double)std::max<int64_t>(p[q].dims[d], 1);
            cost[(size_t)q] = n * n * (double)std::max<int64_t>(p[q].dims[nd - 1], 1);
        }
        std::stable_sort(order.begin(), order.end(), [&](int32_t a, int32_t b) { return cost[a] > cost[b]; });
        // block widths tuned here, one size at a time, before the workers
        // share the GPU (timings taken beside other solves would mislead)
        if (nw > 1)
            for (int32_t q : order) {
                if ((p[q].ndim != 2 && p[q].ndim != 3) ||
                    (o.precision != RGF_PRECISION_FP64 && o.precision != RGF_PRECISION_FP32_REFINED))
                    continue;
                const int64_t N = p[q].dims[0] * (p[q].ndim == 3 ? p[q].dims[1] : 1);
                if (N <= 0) continue;
                RGF_HIP(hipSetDevice(h->device));
                Work tw;
                slice_config(h, N, o.precision, o.precision == RGF_PRECISION_FP64 ? 0 : (o.complex_3m & 1), o, tw);
            }
        std::atomic<int32_t> next{0};
        auto run = [&](int wi) {
            rgf_context* c = h->workers[(size_t)wi];
            if (hipSetDevice(c->device) != hipSuccess) {
                for (int32_t k = next++; k < nrest; k = next++) sts[(size_t)order[(size_t)k]] = RGF_ERR_DEVICE;
                return;
            }
            for (int32_t k = next++; k < nrest; k = next++) {
                const int32_t q = order[(size_t)k];
                const rgf_options* oq = pivot_first[(size_t)q] ? &opiv : opt;
                sts[(size_t)q] = guard(c, [&] { solve_one(c, &p[q], &s[q], oq, S_out[q], info ? &info[q] : nullptr); });
                if (sts[(size_t)q] != RGF_OK) msgs[(size_t)q] = c->err;
            }
        };
        if (nrest == 0) {
        } else if (nw == 1) {
            run(0);
        } else {
            std::vector<std::thread> th;
            th.reserve((size_t)nw);
            for (int wi = 0; wi < nw; ++wi) th.emplace_back(run, wi);
            for (auto& t : th) t.join();
        }
        for (int32_t q = 0; q < count; ++q)
            if (sts[(size_t)q] != RGF_OK)
                throw Error(sts[(size_t)q], "batched solve, problem " + std::to_string(q) + ": " + msgs[(size_t)q]);
    });
}

rgf_status rgf_transmission_gradient(rgf_handle h, const rgf_problem* p, const rgf_selection* s,
                                     const rgf_options* opt, int32_t every, double* T_out, double* gre,
                                     double* gim, rgf_info* info) {
    return guard(h, [&] {
        if (!h) throw Error(RGF_ERR_INVALID_ARGUMENT, "NULL handle");
        gradient_one(h, p, s, opt, every, T_out, gre, gim, info);
    });
}

rgf_status rgf_dense_inverse(rgf_handle h, int32_t precision, int64_t n, double* A, int32_t block_size) {
    return guard(h, [&] {
        if (!h || !A || n <= 0) throw Error(RGF_ERR_INVALID_ARGUMENT, "bad dense inverse arguments");
        const int use3m = (precision & RGF_DENSE_3M) ? 1 : 0;
        precision &= ~RGF_DENSE_3M;
        if (use3m && precision == RGF_PRECISION_FP64)
            throw Error(RGF_ERR_INVALID_ARGUMENT, "RGF_DENSE_3M applies to the FP32 kernels");
        RGF_HIP(hipSetDevice(h->device));
        CMat<double> M = dmat<double>(h, "dense_A", n, n);
        // the blocked kernels; a refused matrix is inverted again with
        // partial pivoting over all rows, as the solves do
        Work w;
        w.pivot_all = block_size < 0;
        for (;;) {
            upload_colmajor(h, A, n, n, M);
            reset_status(h);
            if (precision == RGF_PRECISION_FP64) {
                const int b = block_size > 0 ? block_size : default_block(precision);
                double* ws = w.pivot_all ? nullptr : dbuf<double>(h, "inv_ws", inverse_workspace_elems<double>(n, b));
                invert_slice<double>(h, w, h->stream, n, M, b, ws, 0, nullptr, &h->ov);
            } else {
                const int b = block_size > 0 ? block_size : default_block(precision);
                CMat<float> M32 = dmat<float>(h, "dense_A32", n, n);
                convert<float, double>(h->stream, n, n, M, M32, 1.0);
                float* ws = w.pivot_all ? nullptr : dbuf<float>(h, "inv_ws32", inverse_workspace_elems<float>(n, b));
                invert_slice<float>(h, w, h->stream, n, M32, b, ws, use3m, nullptr, &h->ov);
                convert<double, float>(h->stream, n, n, M32, M, 1.0);
            }
            try {
                check_status(h, true, precision == RGF_PRECISION_FP64);
            } catch (const SliceRefused&) {
                if (w.pivot_all) throw;
                w.pivot_all = 1;
                continue;
            }
            break;
        }
        download_colmajor(h, M, n, n, A);
    });
}

rgf_status rgf_dense_inverse_noalloc(rgf_handle h, int32_t precision, int64_t n, void* A_re, void* A_im, int64_t lda,
                                     int32_t block_size, int32_t* status, void* stream) {
    return guard(h, [&] {
        if (!h || !A_re || !A_im || n <= 0 || lda < n)
            throw Error(RGF_ERR_INVALID_ARGUMENT, "bad no-allocation dense inverse arguments");
        const int use3m = (precision & RGF_DENSE_3M) ? 1 : 0;
        precision &= ~RGF_DENSE_3M;
        if (precision != RGF_PRECISION_FP64 && precision != RGF_PRECISION_FP32_REFINED)
            throw Error(RGF_ERR_INVALID_ARGUMENT, "unknown precision");
        if (use3m && precision == RGF_PRECISION_FP64)
            throw Error(RGF_ERR_INVALID_ARGUMENT, "RGF_DENSE_3M applies to the FP32 kernels");
        const bool fp64 = precision == RGF_PRECISION_FP64;
        const int b = block_size > 0 ? block_size : default_block(precision);
        if (b > (fp64 ? diag_block_max<double>() : diag_block_max<float>()) || b % 16 != 0)
            throw Error(RGF_ERR_INVALID_ARGUMENT, "invalid inverse block size");
        const size_t need = 256 + (fp64 ? sizeof(double) * inverse_workspace_elems<double>(n, b)
                                        : sizeof(float) * inverse_workspace_elems<float>(n, b));
        if (need > RGF_NOALLOC_SCRATCH_BYTES) {
            char buf[200];
            snprintf(buf, sizeof buf,
                     "n = %lld needs %zu bytes of workspace; the no-allocation entry point has %u "
                     "(use rgf_dense_inverse)",
                     (long long)n, need, (unsigned)RGF_NOALLOC_SCRATCH_BYTES);
            throw Error(RGF_ERR_INVALID_ARGUMENT, buf);
        }
        if (!h->scratch) throw Error(RGF_ERR_DEVICE, "no-allocation workspace unavailable on this device");
        RGF_HIP(hipSetDevice(h->device));
        hipStream_t s = stream ? (hipStream_t)stream : h->stream;
        int* st = status ? status : (int*)h->scratch;
        zero_words(s, st, 2);
        void* ws = (char*)h->scratch + 256;
        if (fp64)
            invert_inplace<double>(s, n, CMat<double>((double*)A_re, (double*)A_im, lda), b, (double*)ws, st, 0,
                                   nullptr, &h->ov);
        else
            invert_inplace<float>(s, n, CMat<float>((float*)A_re, (float*)A_im, lda), b, (float*)ws, st, use3m,
                                  nullptr, &h->ov);
        if (!status) {
            int w[2] = {0, 0};
            RGF_HIP(hipMemcpyAsync(w, st, sizeof w, hipMemcpyDeviceToHost, s));
            RGF_HIP(hipStreamSynchronize(s));
            status_error(w[0]);
            float m2;
            memcpy(&m2, &w[1], sizeof m2);
            if (fp64 && !(std::sqrt((double)m2) <= multiplier_limit(true)))
                throw Error(RGF_ERR_NUMERICAL,
                            "ill-conditioned diagonal block in a slice inverse (pivoting is confined to the "
                            "diagonal block, so the inverse would be inaccurate)");
        }
    });
}

rgf_status rgf_dense_gemm(rgf_handle h, int32_t precision, int64_t m, int64_t n, int64_t k, const double* A,
                          const double* B, double* C) {
    return guard(h, [&] {
        if (!h || !A || !B || !C || m <= 0 || n <= 0 || k <= 0) throw Error(RGF_ERR_INVALID_ARGUMENT, "bad gemm");
        RGF_HIP(hipSetDevice(h->device));
        CMat<double> dA = dmat<double>(h, "gA", m, k), dB = dmat<double>(h, "gB", k, n), dC = dmat<double>(h, "gC", m, n);
        upload_colmajor(h, A, m, k, dA);
        upload_colmajor(h, B, k, n, dB);
        // with the solver's split-K workspace: thin products with a long inner
        // dimension take the split and its reduction, as in the sweeps
        if (precision == RGF_PRECISION_FP64) {
            const auto wk = splitk_ws<double>(h);
            cgemm<double>(h->stream, m, n, k, dA, dB, dC, cd(1, 0), 0, 0, 0, 0, wk.first, wk.second);
        } else {
            CMat<float> a = dmat<float>(h, "gA32", m, k), b = dmat<float>(h, "gB32", k, n), c = dmat<float>(h, "gC32", m, n);
            convert<float, double>(h->stream, m, k, dA, a, 1.0);
            convert<float, double>(h->stream, k, n, dB, b, 1.0);
            const auto wk = splitk_ws<float>(h);
            cgemm<float>(h->stream, m, n, k, a, b, c, cd(1, 0), 0, 0, 0, 0, wk.first, wk.second);
            convert<double, float>(h->stream, m, n, c, dC, 1.0);
        }
        download_colmajor(h, dC, m, n, C);
    });
}

}  // extern "C"
