// Batched engine: many independent mixed-precision solves (frequencies,
// Bloch phases, widths) advanced in lockstep, so that every launch carries
// the work of the whole batch.
//
// Dyson phase.  Every problem's slice inverses are blocked Gauss-Jordan
// eliminations with 64-wide blocks (the transverse size padded with the
// identity to NP, a multiple of 64).  Round r advances every problem by one
// block step (problem p: slice 1 + r / nb_p, block r % nb_p) with three
// launches over the ragged batch: diagonal-block inverses beside the
// write-back of the previous step and the assembly of the next slice; the
// panel products; the trailing updates.  The round index is read from
// device memory, so a run of rounds is captured once as a hipGraph and
// replayed (the graph's last node advances the index).
//
// Refinement.  The requested S columns are the solution of the whole
// block-tridiagonal system, refined in FP64 exactly as in the single-solve
// mixed path (run_refined in rgf_api.cpp): residual maximum, forward and
// backward sweeps through the stored FP32 inverses, accumulation, each a
// launch over all problems still iterating.  The residual is recomputed
// slice by slice inside the forward sweep instead of stored, which holds the
// batch's memory to the factor store, psi (FP64) and the FP32 correction.
//
// The engine reproduces the reference's S (cal_smatrix_RGF.m:279-289,
// cal_smatrix_RGF_3D.m:294-304) for each problem, as rgf_solve does.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <memory>
#include <thread>

#include "rgf_context.h"

namespace rgf {

namespace {

int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }
size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

// (problem << 20) | local index
int32_t map_entry(size_t prob, int64_t loc) {
    if (prob >= 2048 || loc >= (1 << 20) || loc < 0) throw Error(RGF_ERR_INVALID_ARGUMENT, "batched engine: work map overflow");
    return (int32_t)((prob << 20) | (size_t)loc);
}

// One problem of the batch on the host: channels, selection, stencil, sizes
// and its region of the batch's device arena.
struct HostProb {
    int32_t user = 0;  // index in the caller's arrays
    Channels ch;
    Selection sel;
    HostCSR csr;
    int64_t N = 0, NP = 0, nb = 0, L = 0, m = 0, mL = 0, mR = 0, npo = 0, nqo = 0;
    // byte offsets in the problem's region: [zeroed | eps | csr | F | psi | z]
    size_t o_ws = 0, o_G0L = 0, o_G0R = 0, o_B = 0, o_CL = 0, o_CR = 0, o_t = 0, o_W = 0, o_T = 0, zero_bytes = 0;
    size_t o_eps = 0, o_csr = 0, o_F = 0, o_psi = 0, o_z = 0, bytes = 0;
    size_t base = 0;  // offset of the region in the arena
    bool cplx_eps = false;  // eps has an imaginary part
    bool ready = false, failed = false, converged = false;
    int iters = 0;
    double bmax = 0, rel = 1;
    // refine_s_tol: the S blocks of the last sweep and the error estimate
    std::vector<double> Tprev;
    double s_err = 0;
    bool s_done = false;
    int64_t T() const { return (L + 1) * nb; }
    int64_t S() const { return L + 2; }
};

// Stencil arrays in the problem's region: rowptr, col, vre, vim at 256-byte
// aligned offsets; returns the bytes they take.
size_t csr_layout(const HostCSR& h, size_t off[4]) {
    const size_t b[4] = {4 * h.rowptr.size(), 4 * h.col.size(), 8 * h.vre.size(), 8 * h.vim.size()};
    size_t o = 0;
    for (int k = 0; k < 4; ++k) {
        off[k] = o;
        o = align256(o + b[k]);
    }
    return o;
}

void plan_sizes(HostProb& hp) {
    const Channels& ch = hp.ch;
    hp.N = ch.N;
    hp.L = ch.L_slices;
    hp.NP = ceil_div(hp.N, 64) * 64;
    hp.nb = hp.NP / 64;
    hp.mL = (int64_t)hp.sel.inL.size();
    hp.mR = (int64_t)hp.sel.inR.size();
    hp.m = ceil_div(hp.mL + hp.mR, 4) * 4;
    hp.npo = (int64_t)hp.sel.outL.size();
    hp.nqo = (int64_t)hp.sel.outR.size();
    hp.csr = slice_stencil(ch);
    const size_t NP = hp.NP, m = hp.m, S = hp.S(), NN = NP * NP;
    size_t o = 0;
    auto take = [&](size_t bytes) {
        const size_t at = o;
        o = align256(o + bytes);
        return at;
    };
    hp.o_ws = take(sizeof(float) * 2 * (64 * 64 + 2 * 64 * NP));
    hp.o_G0L = take(sizeof(double) * 2 * NN);
    hp.o_G0R = hp.ch.same_sides ? hp.o_G0L : take(sizeof(double) * 2 * NN);
    hp.o_B = take(sizeof(double) * 2 * NP * m);
    hp.o_CL = take(sizeof(double) * 2 * hp.npo * NP);
    hp.o_CR = take(sizeof(double) * 2 * hp.nqo * NP);
    hp.o_t = take(sizeof(float) * 2 * NP * m);
    hp.o_W = take(sizeof(double) * 4 * NP * m);
    hp.o_T = take(sizeof(double) * 2 * (hp.npo + hp.nqo) * m);
    hp.zero_bytes = o;
    hp.o_eps = take(sizeof(double) * 2 * hp.N * hp.L);
    size_t co[4];
    hp.o_csr = take(csr_layout(hp.csr, co));
    hp.o_F = take(sizeof(float) * 2 * S * NN);
    hp.o_psi = take(sizeof(double) * 2 * S * NP * m);
    hp.o_z = take(sizeof(float) * 2 * S * NP * m);
    hp.bytes = o;
}

// Device set-up of one problem on a worker context: its arena region, eps,
// stencil, lead Green's functions and the lead projections B and C.
void setup_one(rgf_context* c, HostProb& hp, char* arena, const rgf_problem* p) {
    hipStream_t s = c->stream;
    char* base = arena + hp.base;
    const int64_t N = hp.N, NP = hp.NP, L = hp.L, m = hp.m;
    RGF_HIP(hipMemsetAsync(base, 0, hp.zero_bytes, s));
    // the first correction's forward products leave the right inputs' columns
    // of z untouched before the last slice: they must read as zero
    RGF_HIP(hipMemsetAsync(base + hp.o_z, 0, sizeof(float) * 2 * hp.S() * NP * m, s));
    double* eps = (double*)(base + hp.o_eps);
    if (L > 0) {
        const hipMemcpyKind kind = p->eps_on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
        RGF_HIP(hipMemcpyAsync(eps, p->eps_re, sizeof(double) * N * L, kind, s));
        if (p->eps_im) RGF_HIP(hipMemcpyAsync(eps + N * L, p->eps_im, sizeof(double) * N * L, kind, s));
    }
    char* cs = base + hp.o_csr;
    const HostCSR& h = hp.csr;
    size_t co[4];
    csr_layout(h, co);
    RGF_HIP(hipMemcpyAsync(cs + co[0], h.rowptr.data(), 4 * h.rowptr.size(), hipMemcpyHostToDevice, s));
    RGF_HIP(hipMemcpyAsync(cs + co[1], h.col.data(), 4 * h.col.size(), hipMemcpyHostToDevice, s));
    RGF_HIP(hipMemcpyAsync(cs + co[2], h.vre.data(), 8 * h.vre.size(), hipMemcpyHostToDevice, s));
    RGF_HIP(hipMemcpyAsync(cs + co[3], h.vim.data(), 8 * h.vim.size(), hipMemcpyHostToDevice, s));
    build_modes(hp.ch);
    const ModeTables f = upload_modes(c, hp.ch);
    const size_t NN = (size_t)NP * NP;
    double* G0L = (double*)(base + hp.o_G0L);
    lead_green(c, hp.ch, f, hp.ch.L, CMat<double>(G0L, G0L + NN, NP));
    if (!hp.ch.same_sides) {
        double* G0R = (double*)(base + hp.o_G0R);
        lead_green(c, hp.ch, f, hp.ch.R, CMat<double>(G0R, G0R + NN, NP));
    }
    double* B = (double*)(base + hp.o_B);
    const size_t SE = (size_t)NP * m;
    mode_cols(c, hp.ch, f, hp.ch.L, hp.sel.inL, 0, "BL_cols", CMat<double>(B, B + SE, m));
    mode_cols(c, hp.ch, f, hp.ch.R, hp.sel.inR, 0, "BR_cols", CMat<double>(B + hp.mL, B + SE + hp.mL, m));
    double* CL = (double*)(base + hp.o_CL);
    double* CR = (double*)(base + hp.o_CR);
    mode_cols(c, hp.ch, f, hp.ch.L, hp.sel.outL, 1, "CL_cols", CMat<double>(CL, CL + hp.npo * NP, NP));
    mode_cols(c, hp.ch, f, hp.ch.R, hp.sel.outR, 1, "CR_cols", CMat<double>(CR, CR + hp.nqo * NP, NP));
    RGF_HIP(hipStreamSynchronize(s));
    hp.ready = true;
}

BatchProb device_entry(const HostProb& hp, char* arena, int32_t* status, unsigned long long* rmax) {
    char* base = arena + hp.base;
    BatchProb b;
    memset(&b, 0, sizeof b);
    b.N = (int32_t)hp.N;
    b.NP = (int32_t)hp.NP;
    b.nb = (int32_t)hp.nb;
    b.L = (int32_t)hp.L;
    b.m = (int32_t)hp.m;
    b.mL = (int32_t)hp.mL;
    b.mR = (int32_t)hp.mR;
    b.npo = (int32_t)hp.npo;
    b.nqo = (int32_t)hp.nqo;
    size_t co[4];
    csr_layout(hp.csr, co);
    char* cs = base + hp.o_csr;
    b.rowptr = (const int32_t*)(cs + co[0]);
    b.col = (const int32_t*)(cs + co[1]);
    b.vre = (const double*)(cs + co[2]);
    b.vim = (const double*)(cs + co[3]);
    double* eps = (double*)(base + hp.o_eps);
    b.eps_re = eps;
    b.eps_im = hp.cplx_eps ? eps + hp.N * hp.L : nullptr;
    b.k2r = hp.ch.k2.real();
    b.k2i = hp.ch.k2.imag();
    b.eLr = hp.ch.epsL.real();
    b.eLi = hp.ch.epsL.imag();
    b.eRr = hp.ch.epsR.real();
    b.eRi = hp.ch.epsR.imag();
    b.F = (float*)(base + hp.o_F);
    b.ws = (float*)(base + hp.o_ws);
    b.status = status;
    b.G0L = (double*)(base + hp.o_G0L);
    b.G0R = (double*)(base + hp.o_G0R);
    b.B = (double*)(base + hp.o_B);
    b.CL = (double*)(base + hp.o_CL);
    b.CR = (double*)(base + hp.o_CR);
    b.psi = (double*)(base + hp.o_psi);
    b.z = (float*)(base + hp.o_z);
    b.t = (float*)(base + hp.o_t);
    b.W = (double*)(base + hp.o_W);
    b.T = (double*)(base + hp.o_T);
    b.rmax = rmax;
    return b;
}

// Work maps of a set of problems (positions in the device table).
struct Maps {
    std::vector<int32_t> v;
    int append(const std::vector<size_t>& probs, const std::function<int64_t(size_t)>& count) {
        const size_t at = v.size();
        for (size_t k : probs)
            for (int64_t c = 0, n = count(k); c < n; ++c) v.push_back(map_entry(k, c));
        return (int)(v.size() - at);
    }
};

// The engine on one chunk of problems that fits in device memory together.
void run_chunk(rgf_context* h, std::vector<HostProb*>& ps, const rgf_problem* p, const rgf_options& o,
               double* const* S_out, rgf_info* info, std::vector<rgf_status>& sts, std::vector<std::string>& msgs) {
    const double t0 = now();
    const size_t np = ps.size();
    hipStream_t s = h->stream;
    size_t total = 0;
    for (HostProb* hp : ps) {
        hp->base = total;
        total += hp->bytes;
    }
    char* arena = dbuf<char>(h, "batch_arena", total);
    int32_t* status = dbuf<int32_t>(h, "batch_status", 2 * np);
    unsigned long long* rmax = dbuf<unsigned long long>(h, "batch_rmax", np);
    RGF_HIP(hipMemsetAsync(status, 0, 8 * np, s));
    RGF_HIP(hipStreamSynchronize(s));

    // device set-up on the worker contexts (their own streams and scratch)
    {
        const int nw = (int)std::min<size_t>(8, np);
        while ((int)h->workers.size() < nw) {
            rgf_handle w = nullptr;
            const rgf_status cs = rgf_create(&w, h->device);
            if (cs != RGF_OK) throw Error(cs, std::string("batched engine: worker context: ") + rgf_last_error(nullptr));
            h->workers.push_back(w);
        }
        std::atomic<size_t> next{0};
        auto run = [&](int wi) {
            rgf_context* c = h->workers[(size_t)wi];
            if (hipSetDevice(c->device) != hipSuccess) return;
            for (size_t k = next++; k < np; k = next++) {
                HostProb& hp = *ps[k];
                const rgf_status st = guard(c, [&] { setup_one(c, hp, arena, &p[hp.user]); });
                if (st != RGF_OK) {
                    hp.failed = true;
                    sts[(size_t)hp.user] = st;
                    msgs[(size_t)hp.user] = c->err;
                }
            }
        };
        std::vector<std::thread> th;
        for (int wi = 0; wi < nw; ++wi) th.emplace_back(run, wi);
        for (auto& t : th) t.join();
        for (size_t k = 0; k < np; ++k)
            if (!ps[k]->ready && !ps[k]->failed) {  // no worker could take it
                ps[k]->failed = true;
                sts[(size_t)ps[k]->user] = RGF_ERR_DEVICE;
                msgs[(size_t)ps[k]->user] = "batched engine: no worker context could set the problem up";
            }
    }
    // the table (problems in chunk order; a failed set-up runs no work)
    std::vector<BatchProb> tab(np);
    std::vector<size_t> live;
    for (size_t k = 0; k < np; ++k) {
        tab[k] = device_entry(*ps[k], arena, status + 2 * k, rmax + k);
        if (!ps[k]->failed) {
            live.push_back(k);
        } else {
            tab[k].L = -1;  // no block steps: every launch skips it
            tab[k].nb = 1;
        }
    }
    if (live.empty()) return;
    BatchProb* dtab = dbuf<BatchProb>(h, "batch_table", np);
    RGF_HIP(hipMemcpyAsync(dtab, tab.data(), sizeof(BatchProb) * np, hipMemcpyHostToDevice, s));

    // Dyson rounds
    Maps mp;
    const int64_t CH = kBatchChunk;
    // problems without work keep their diagonal groups (they exit at once):
    // the k1 grid indexes the table directly
    auto nn = [&](size_t k) { return ps[k]->NP * ps[k]->NP; };
    const int n1 = mp.append(live, [&](size_t k) { return ps[k]->NP / kBatchRows; });
    const int n2 = mp.append(live, [&](size_t k) { return ps[k]->nb > 1 ? 2 * ps[k]->nb : (int64_t)0; });
    const int n3 = mp.append(live, [&](size_t k) { return (ps[k]->nb - 1) * (ps[k]->nb - 1); });
    const int nfs = mp.append(live, [&](size_t k) { return ceil_div(nn(k), CH); });
    int32_t* dmaps = dbuf<int32_t>(h, "batch_maps", mp.v.size() + 64);
    RGF_HIP(hipMemcpyAsync(dmaps, mp.v.data(), 4 * mp.v.size(), hipMemcpyHostToDevice, s));
    const int32_t* k1map = dmaps;
    const int32_t* k2map = dmaps + n1;
    const int32_t* k3map = k2map + n2;
    const int32_t* fsmap = k3map + n3;
    int* rbase = dbuf<int>(h, "batch_round", 64);
    RGF_HIP(hipMemsetAsync(rbase, 0, sizeof(int), s));
    const int use3m = o.complex_3m & 1, sweep3m = (o.complex_3m >> 1) & 1;
    batch_first_slice(s, dtab, fsmap, nfs);
    int64_t rounds = 0;
    for (size_t k : live) rounds = std::max(rounds, ps[k]->T() + 1);
    const int graph_mode = o.batch_engine;  // 0: hipGraph replays, 1: plain launches
    const int K = o.batch_graph_rounds > 0 ? o.batch_graph_rounds : 16;
    // fork / join of the row pass on the second stream
    struct Ev {
        hipEvent_t e = nullptr;
        ~Ev() {
            if (e) (void)hipEventDestroy(e);
        }
    } fork, join;
    RGF_HIP(hipEventCreateWithFlags(&fork.e, hipEventDisableTiming));
    RGF_HIP(hipEventCreateWithFlags(&join.e, hipEventDisableTiming));
    hipStream_t s2 = h->ov.s2;
    RGF_HIP(hipStreamSynchronize(s));
    const double t_setup = now() - t0;
    const double t1 = now();
    if (graph_mode == 1) {
        for (int64_t r = 0; r < rounds; ++r)
            batch_round(s, s2, fork.e, join.e, dtab, (int)np, rbase, (int)r, k1map, n1, k2map, n2, k3map, n3,
                        use3m);
    } else {
        hipGraph_t graph = nullptr;
        hipGraphExec_t exec = nullptr;
        RGF_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        for (int k = 0; k < K; ++k)
            batch_round(s, s2, fork.e, join.e, dtab, (int)np, rbase, k, k1map, n1, k2map, n2, k3map, n3, use3m);
        batch_advance(s, rbase, K);
        RGF_HIP(hipStreamEndCapture(s, &graph));
        const hipError_t ie = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
        if (ie != hipSuccess) {
            (void)hipGraphDestroy(graph);
            RGF_HIP(ie);
        }
        hipError_t le = hipSuccess;
        for (int64_t r = 0; r < rounds && le == hipSuccess; r += K) le = hipGraphLaunch(exec, s);
        const hipError_t se = hipStreamSynchronize(s);
        (void)hipGraphExecDestroy(exec);
        (void)hipGraphDestroy(graph);
        RGF_HIP(le);
        RGF_HIP(se);
    }
    RGF_HIP(hipStreamSynchronize(s));
    const double t_slices = now() - t1;
    {
        std::vector<int32_t> st(2 * np);
        RGF_HIP(hipMemcpy(st.data(), status, 8 * np, hipMemcpyDeviceToHost));
        for (size_t k : live) {
            float m2;
            memcpy(&m2, &st[2 * k + 1], sizeof m2);
            if (info) info[ps[k]->user].max_multiplier = std::sqrt((double)m2);
            if (st[2 * k]) {
                ps[k]->failed = true;
                sts[(size_t)ps[k]->user] = RGF_ERR_NUMERICAL;
                msgs[(size_t)ps[k]->user] = "zero or non-finite pivot in a slice inverse";
            }
        }
    }

    // refinement
    const double t2 = now();
    // with opt.profile: one forward sweep product in four timed with HIP
    // events on the stream (the engine's dominant kernel)
    std::unique_ptr<KernelTimer> timer;
    if (o.profile) timer.reset(new KernelTimer(&h->timer_events));
    const double fl3 = sweep3m ? 6.0 : 8.0;
    const int maxit = o.max_refine_iter > 0 ? o.max_refine_iter : 6;
    const double tol = o.refine_tol > 0 ? o.refine_tol : 1e-13;
    const bool s_stop = o.refine_s_tol > 0;
    std::vector<unsigned long long> rm(np);
    for (int it = 0;; ++it) {
        std::vector<size_t> act;
        for (size_t k : live)
            if (!ps[k]->failed && !ps[k]->converged) act.push_back(k);
        if (act.empty()) break;
        const int first = it == 0 ? 1 : 0;
        Maps im;
        const int nlead = first ? 0 : im.append(act, [&](size_t k) {
            return 2 * (ps[k]->NP / 64) * ceil_div(ps[k]->m, 64);
        });
        const int nres = im.append(act, [&](size_t k) {
            return ceil_div((first ? 2 : ps[k]->S()) * ps[k]->N, kBatchResRows);
        });
        int32_t* imaps = dbuf<int32_t>(h, "batch_imaps", im.v.size() + 64);
        RGF_HIP(hipMemcpyAsync(imaps, im.v.data(), 4 * im.v.size(), hipMemcpyHostToDevice, s));
        batch_lead_terms(s, dtab, imaps, nlead);
        RGF_HIP(hipMemsetAsync(rmax, 0, 8 * np, s));
        batch_resmax(s, dtab, imaps + nlead, nres, first);
        RGF_HIP(hipMemcpyAsync(rm.data(), rmax, 8 * np, hipMemcpyDeviceToHost, s));
        RGF_HIP(hipStreamSynchronize(s));
        if (s_stop && it >= 1) {
            // the S-error estimate of run_refined: with linear convergence the
            // latest S is within q / (1 - q) of the change over the last sweep
            // (q the ratio of the last two residuals), doubled; S = -2i T
            Maps pm;
            const int npj = pm.append(act, [&](size_t k) {
                const HostProb& hp = *ps[k];
                return (ceil_div(hp.npo, 64) + ceil_div(hp.nqo, 64)) * ceil_div(hp.m, 64);
            });
            int32_t* pmaps = dbuf<int32_t>(h, "batch_pmaps", pm.v.size() + 64);
            RGF_HIP(hipMemcpyAsync(pmaps, pm.v.data(), 4 * pm.v.size(), hipMemcpyHostToDevice, s));
            batch_project(s, dtab, pmaps, npj);
            RGF_HIP(hipStreamSynchronize(s));
            for (size_t k : act) {
                HostProb& hp = *ps[k];
                std::vector<double> Tc((size_t)2 * (hp.npo + hp.nqo) * hp.m);
                if (!Tc.empty())
                    RGF_HIP(hipMemcpy(Tc.data(), arena + hp.base + hp.o_T, sizeof(double) * Tc.size(),
                                      hipMemcpyDeviceToHost));
                double r;
                memcpy(&r, &rm[k], 8);
                const double rel = r / hp.bmax;
                if (it >= 2 && Tc.size() == hp.Tprev.size()) {
                    double dT = 0;
                    const size_t h2 = Tc.size() / 2;
                    for (size_t e = 0; e < h2; ++e)
                        dT = std::max(dT, std::hypot(Tc[e] - hp.Tprev[e], Tc[h2 + e] - hp.Tprev[h2 + e]));
                    const double q = rel / hp.rel;  // hp.rel: the previous sweep's residual
                    if (q < 0.5) {
                        hp.s_err = 2.0 * (2.0 * dT) * q / (1.0 - q);
                        hp.s_done = hp.s_err <= o.refine_s_tol;
                    }
                }
                hp.Tprev = std::move(Tc);
            }
        }
        std::vector<size_t> go;
        for (size_t k : act) {
            HostProb& hp = *ps[k];
            double r;
            memcpy(&r, &rm[k], 8);
            if (first) hp.bmax = r == 0 ? 1.0 : r;
            hp.rel = r / hp.bmax;
            hp.iters = it;
            if (!std::isfinite(hp.rel)) {
                hp.failed = true;
                sts[(size_t)hp.user] = RGF_ERR_NUMERICAL;
                msgs[(size_t)hp.user] = "non-finite residual in refinement";
            } else if (hp.rel <= tol || r == 0 || hp.s_done) {
                hp.converged = true;
            } else if (it >= maxit) {
                hp.failed = true;
                char buf[160];
                snprintf(buf, sizeof buf, "mixed-precision refinement stopped at relative residual %.3e (tolerance %.1e)",
                         hp.rel, tol);
                sts[(size_t)hp.user] = RGF_ERR_NOT_CONVERGED;
                msgs[(size_t)hp.user] = buf;
            } else {
                go.push_back(k);
            }
        }
        if (go.empty()) break;
        Maps sm;
        const int nrhs = sm.append(go, [&](size_t k) { return ceil_div(ps[k]->N * ps[k]->m, CH); });
        const int ngem = sm.append(go, [&](size_t k) { return (ps[k]->NP / 64) * ceil_div(ps[k]->m, 64); });
        const int nacc = sm.append(go, [&](size_t k) { return ceil_div(ps[k]->S() * ps[k]->NP * ps[k]->m, CH); });
        int32_t* smaps = dbuf<int32_t>(h, "batch_smaps", sm.v.size() + 64);
        RGF_HIP(hipMemcpyAsync(smaps, sm.v.data(), 4 * sm.v.size(), hipMemcpyHostToDevice, s));
        int64_t smax = 0;
        for (size_t k : go) smax = std::max(smax, ps[k]->S());
        for (int64_t n = 0; n < smax; ++n) {
            batch_fwd_rhs(s, dtab, smaps, nrhs, (int)n);
            if (timer) timer->begin(s);
            batch_fwd_gemm(s, dtab, smaps + nrhs, ngem, (int)n, sweep3m, first);
            if (timer && timer->active) {
                double fl = 0;  // executed flops of the launch: NP x m' x NP per problem
                for (size_t k : go) {
                    const HostProb& hp = *ps[k];
                    if (n >= hp.S()) continue;
                    const int64_t mc = (first && n < hp.S() - 1) ? std::min<int64_t>(hp.m, ceil_div(hp.mL, 4) * 4)
                                                                  : hp.m;
                    fl += fl3 * (double)hp.NP * (double)hp.NP * (double)mc;
                }
                timer->end(s, fl);
            }
        }
        for (int64_t k = 0; k + 1 < smax; ++k) batch_bwd_gemm(s, dtab, smaps + nrhs, ngem, (int)k, sweep3m);
        batch_accumulate(s, dtab, smaps + nrhs + ngem, nacc, first);
        RGF_HIP(hipGetLastError());
    }
    // S blocks from the boundary slices
    std::vector<size_t> done;
    for (size_t k : live)
        if (!ps[k]->failed) done.push_back(k);
    if (!done.empty()) {
        Maps pm;
        const int npj = pm.append(done, [&](size_t k) {
            const HostProb& hp = *ps[k];
            return (ceil_div(hp.npo, 64) + ceil_div(hp.nqo, 64)) * ceil_div(hp.m, 64);
        });
        int32_t* pmaps = dbuf<int32_t>(h, "batch_pmaps", pm.v.size() + 64);
        RGF_HIP(hipMemcpyAsync(pmaps, pm.v.data(), 4 * pm.v.size(), hipMemcpyHostToDevice, s));
        batch_project(s, dtab, pmaps, npj);
        RGF_HIP(hipStreamSynchronize(s));
    }
    const double t_refine = now() - t2;
    double k_sec = 0, k_fl = 0;
    int64_t k_n = 0;
    if (timer) {
        k_sec = timer->seconds();
        k_fl = timer->flops;
        k_n = timer->launches;
    }
    for (size_t k : done) {
        HostProb& hp = *ps[k];
        const int64_t rows = hp.npo + hp.nqo, m = hp.m;
        std::vector<double> Th((size_t)2 * rows * m);
        if (!Th.empty())
            RGF_HIP(hipMemcpy(Th.data(), (char*)arena + hp.base + hp.o_T, sizeof(double) * Th.size(),
                              hipMemcpyDeviceToHost));
        Work w;
        w.ch = &hp.ch;
        w.mirrored = false;
        w.inL = hp.sel.inL;
        w.inR = hp.sel.inR;
        w.outL = hp.sel.outL;
        w.outR = hp.sel.outR;
        auto block = [&](int64_t r0, int64_t nr, int64_t c0, int64_t nc, std::vector<cd>& out) {
            out.assign((size_t)(nr * nc), cd(0, 0));
            for (int64_t i = 0; i < nr; ++i)
                for (int64_t j = 0; j < nc; ++j) {
                    const size_t e = (size_t)((r0 + i) * m + c0 + j);
                    out[(size_t)(i * nc + j)] = cd(Th[e], Th[(size_t)(rows * m) + e]);
                }
        };
        if (hp.npo) {
            block(0, hp.npo, 0, hp.mL, w.TLL);
            block(0, hp.npo, hp.mL, hp.mR, w.TLR);
        }
        if (hp.nqo) {
            block(hp.npo, hp.nqo, 0, hp.mL, w.TRL);
            block(hp.npo, hp.nqo, hp.mL, hp.mR, w.TRR);
        }
        write_smatrix(w, hp.sel, S_out[hp.user]);
        if (info) {
            rgf_info& in = info[hp.user];
            in.refine_iters = hp.iters;
            in.refine_residual = hp.rel;
            in.refine_s_err_estimate = hp.s_err;
            in.block_size = 64;  // the engine's block step
            in.seconds_setup = t_setup;
            in.seconds_slices = t_slices;
            in.seconds_refine = t_refine;
            in.seconds_total = now() - t0;
            const double Nd = (double)hp.NP;
            in.flops_slices = (double)(hp.L + 1) * 8 * Nd * Nd * Nd * (use3m ? 0.75 : 1.0);
            // the timed kernel is shared by the whole chunk
            in.kernel_seconds = k_sec;
            in.kernel_flops = k_fl;
            in.kernel_launches = k_n;
        }
    }
}

}  // namespace

bool batch_engine_takes(const rgf_problem* p, const rgf_options& o) {
    if (o.precision != RGF_PRECISION_FP32_REFINED || o.batch_engine == 2 || o.pivoting == 1) return false;
    if (!p || (p->ndim != 2 && p->ndim != 3)) return false;
    int64_t N = 1;
    for (int d = 0; d < p->ndim - 1; ++d) N *= std::max<int64_t>(p->dims[d], 0);
    return N >= 1 && N <= kBatchMaxN;
}

void solve_batched_engine(rgf_context* h, const std::vector<int32_t>& idx, const rgf_problem* p, const rgf_selection* s,
                          const rgf_options& o, double* const* S_out, rgf_info* info, std::vector<rgf_status>& sts,
                          std::vector<std::string>& msgs) {
    RGF_HIP(hipSetDevice(h->device));
    std::vector<HostProb> hps;
    hps.reserve(idx.size());
    for (int32_t q : idx) {
        HostProb hp;
        hp.user = q;
        hp.cplx_eps = p[q].eps_im != nullptr;
        const rgf_status st = guard(h, [&] {
            hp.ch = setup_channels(&p[q], false);  // mode tables: on the worker threads
            hp.sel = make_selection(&s[q], hp.ch);
            if (info) fill_info(&info[q], hp.ch, hp.sel);
            if (hp.sel.outL.size() + hp.sel.outR.size() == 0 || hp.sel.inL.size() + hp.sel.inR.size() == 0)
                return;  // cal_smatrix_RGF.m:181-184: nothing to compute
            if (!S_out[q]) throw Error(RGF_ERR_INVALID_ARGUMENT, "S_out is NULL");
            plan_sizes(hp);
            hps.push_back(std::move(hp));
        });
        if (st != RGF_OK) {
            sts[(size_t)q] = st;
            msgs[(size_t)q] = h->err;
        }
    }
    // longest first: the problems still stepping in a round are a prefix
    std::vector<HostProb*> order;
    for (auto& hp : hps) order.push_back(&hp);
    std::stable_sort(order.begin(), order.end(), [](const HostProb* a, const HostProb* b) { return a->T() > b->T(); });
    // chunks that fit in device memory with what the handle already holds
    size_t freeb = 0, totalb = 0;
    RGF_HIP(hipMemGetInfo(&freeb, &totalb));
    const double avail = 0.9 * ((double)freeb + (double)held_bytes(h, {"batch_arena"}));
    size_t at = 0;
    while (at < order.size()) {
        std::vector<HostProb*> chunk;
        double need = 0;
        while (at < order.size() && chunk.size() < 2047) {
            const double b = (double)order[at]->bytes;
            if (!chunk.empty() && need + b > avail) break;
            need += b;
            chunk.push_back(order[at++]);
        }
        if (need > avail) {
            for (HostProb* hp : chunk) {
                sts[(size_t)hp->user] = RGF_ERR_OUT_OF_MEMORY;
                msgs[(size_t)hp->user] = "batched solve: the problem's factor store does not fit in device memory";
            }
            continue;
        }
        run_chunk(h, chunk, p, o, S_out, info, sts, msgs);
    }
}

}  // namespace rgf
