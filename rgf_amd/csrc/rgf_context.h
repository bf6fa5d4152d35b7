// Internal state of a librgf handle and the host helpers that the solve
// drivers (rgf_api.cpp) and the batched engine (rgf_batch.cpp) share.
#pragma once

#include <functional>
#include <initializer_list>
#include <map>
#include <string>
#include <vector>

#include "rgf_host.h"

struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
};

struct rgf_context {
    int device = 0;
    hipStream_t own = nullptr;
    hipStream_t stream = nullptr;
    rgf::Overlap ov;  // look-ahead stream and events of the slice inverse
    std::string err;
    std::map<std::string, DevBuf> bufs;
    int* d_status = nullptr;
    double* d_scratch = nullptr;
    std::vector<rgf_context*> workers;  // worker contexts of rgf_solve_batched
    std::vector<hipEvent_t> timer_events;  // kept across solves (options.profile)
    void* scratch = nullptr;               // the device's no-allocation workspace
    void* rb = nullptr;                    // dense-solver library handle (pivoted slice inverses)
};

namespace rgf {

double now();

// A named device buffer of the handle, grown (freed and allocated again) when
// a call needs more than it holds.
template <typename T>
T* dbuf(rgf_context* c, const std::string& name, size_t count) {
    DevBuf& b = c->bufs[name];
    const size_t need = std::max<size_t>(count, 1) * sizeof(T);
    if (b.bytes < need) {
        if (b.p) {
            RGF_HIP(hipStreamSynchronize(c->stream));
            RGF_HIP(hipFree(b.p));
            b.p = nullptr;
            b.bytes = 0;
        }
        RGF_HIP(hipMalloc(&b.p, need));
        b.bytes = need;
    }
    return (T*)b.p;
}

// planar device matrix rows x cols, row-major, named buffer
template <typename T>
CMat<T> dmat(rgf_context* c, const std::string& name, int64_t rows, int64_t cols) {
    T* p = dbuf<T>(c, name, (size_t)2 * std::max<int64_t>(rows, 1) * std::max<int64_t>(cols, 1));
    return CMat<T>(p, p + (size_t)std::max<int64_t>(rows, 1) * std::max<int64_t>(cols, 1), std::max<int64_t>(cols, 1));
}

// bytes the named buffers already hold (dbuf reuses or grows them in place)
size_t held_bytes(rgf_context* c, std::initializer_list<const char*> names);
void release(rgf_context* c, const std::string& name);

// ---------------------------------------------------------------- problem on the device
// Everything in sweep order: "L" is the lead the slice loop starts from.
struct Work {
    const Channels* ch = nullptr;
    bool mirrored = false;
    int64_t N = 0, L = 0;
    cd k2, epsL, epsR;
    const Side* sL = nullptr;
    const Side* sR = nullptr;
    std::vector<int32_t> inL, inR, outL, outR;  // 0-based, sweep-side
    // device (FP64)
    double* eps_re = nullptr;
    double* eps_im = nullptr;  // N x L planar, slice n at n*N
    DevCSR A0;
    CMat<double> G0L, G0R, BL, BR, CL, CR;
    // host results (FP64, row-major), projected Green's functions (no -2i)
    std::vector<cd> TLL, TLR, TRL, TRR;
    double flops = 0;
    double t_setup = 0, t_slices = 0, t_refine = 0;
    KernelTimer* timer = nullptr;
    int refine_iters = 0;
    double refine_residual = 0;
    double refine_s_err = 0;  // estimated S error at the stop (refine_s_tol)
    double max_multiplier = 0;  // largest slice-inverse multiplier
    int block_size = 0;         // slice-inverse block width used
    int tuned = 0;              // 1: this solve ran the block-width tuner
    int64_t factors_resident = 0;  // mixed path: slice inverses kept resident
    int64_t factors_rebuilt = 0;   // mixed path: slice inverses rebuilt by the sweeps
    double* fields = nullptr;  // host: field profiles wanted (rgf_solve_fields)
    int pivot_all = 0;         // 1: every slice inverted with partial pivoting over all rows
};

// Transverse mode tables f1 (n1 x n1) and f2 (n2 x n2), interleaved FP64,
// on the device: phi((i1,i2),(a1,a2)) = f1[i1][a1] f2[i2][a2].
struct ModeTables {
    const double* f1;
    const double* f2;
};

ModeTables upload_modes(rgf_context* c, const Channels& ch);
// G0 = phi diag(exp(i kl)) phi^H (FP64) into G0 (its ld may exceed N)
void lead_green(rgf_context* c, const Channels& ch, const ModeTables& f, const Side& side, CMat<double> G0);
// selected mode columns (rows = 0: B, N x n) or conjugated rows (rows = 1: C, n x N)
void mode_cols(rgf_context* c, const Channels& ch, const ModeTables& f, const Side& s,
               const std::vector<int32_t>& idx, int rows, const std::string& name, CMat<double> D);
// the reference's S from the projected blocks of w (-2i, then the delta)
void write_smatrix(const Work& w, const Selection& sel, double* S_out);
void fill_info(rgf_info* info, const Channels& ch, const Selection& sel);
rgf_status guard(rgf_context* c, const std::function<void()>& f);

// ---- batched engine (rgf_batch.cpp)
// Whether a problem runs in the lockstep engine: the mixed precision with a
// transverse size up to kBatchMaxN.
constexpr int64_t kBatchMaxN = 1024;
bool batch_engine_takes(const rgf_problem* p, const rgf_options& o);
// Solves problems idx[0..count) of the batch arrays; per-problem status and
// message go to sts / msgs (indexed like the batch arrays).
void solve_batched_engine(rgf_context* h, const std::vector<int32_t>& idx, const rgf_problem* p,
                          const rgf_selection* s, const rgf_options& o, double* const* S_out, rgf_info* info,
                          std::vector<rgf_status>& sts, std::vector<std::string>& msgs);

}  // namespace rgf
