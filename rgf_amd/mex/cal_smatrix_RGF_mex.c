/*
 * MATLAB/Octave MEX gateway over librgf (include/rgf.h).
 *
 *   [smatrix, info] = cal_smatrix_RGF_mex(syst, out, in [, opts])
 *   [T, dT_deps]     = cal_smatrix_RGF_mex('gradient', syst, out, in [, opts])
 *   channels         = cal_smatrix_RGF_mex('channels', syst)
 *   phi              = cal_smatrix_RGF_mex('phi', syst, kydx)          (2D)
 *   phi              = cal_smatrix_RGF_mex('phi', syst, kxdx, kydx)    (3D)
 *
 * It reproduces the calling convention of cal_smatrix_RGF.m:1 and
 * cal_smatrix_RGF_3D.m:1 of chiaweihsu/RGF: `syst` is the same struct
 * (epsilon, epsilon_L, epsilon_R, k0dx and yBC (2D) or xyBC (3D)); `out`
 * and `in` are cells {'L','R'} or structs with ind_out_L/ind_out_R and
 * ind_in_L/ind_in_R.  A 3D epsilon (or a syst with xyBC) selects the 3D
 * solver.  `opts` may hold precision ('fp64' | 'fp32_refined'), refine_tol, refine_s_tol,
 * max_refine_iter and checkpoint_every.  'channels' returns the struct of
 * setup_channels.m:30-82 / setup_channels_3D.m:30-64 (kydx, or kxdx and
 * kydx; L and R with kxdx or kzdx, ind_prop, N_prop, sqrt_mu and, in 2D
 * where defined, ind_prop_conj) from rgf_setup_channels, with an empty
 * fun_phi field; 'phi' evaluates the transverse profiles fun_phi returns
 * (rgf_transverse_profiles) for the transverse size and boundary condition
 * of syst (its epsilon may have no slices).  The wrappers cal_smatrix_RGF.m /
 * cal_smatrix_RGF_3D.m next to this file assemble the reference's `channels`
 * output from the two.  'channels' and 'phi' use no GPU.
 *
 * The gateway only marshals: every check and every number comes from
 * librgf.  Device resources live in one handle created on first use and
 * released by mexAtExit.  librgf frees its per-call state before it returns
 * a status, so raising an error through mexErrMsgIdAndTxt (which does not
 * return) leaves the handle valid; host buffers are mxMalloc'ed and are
 * reclaimed by MATLAB on error.
 *
 * Build (MATLAB R2018a+ interleaved complex API, or Octave's mkoctfile):
 *   mex -R2018a -I../../include cal_smatrix_RGF_mex.c -L.. -lrgf
 *   mkoctfile --mex -I../../include cal_smatrix_RGF_mex.c -L.. -lrgf
 */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "mex.h"
#include "rgf.h"

static rgf_handle g_handle = NULL;

static void release_handle(void) {
    if (g_handle) {
        rgf_destroy(g_handle);
        g_handle = NULL;
    }
}

static void fail(const char* id, const char* msg) { mexErrMsgIdAndTxt(id, "%s", msg); }

static void check(rgf_status st, rgf_handle h) {
    if (st != RGF_OK) {
        const char* m = rgf_last_error(h);
        char buf[512];
        snprintf(buf, sizeof buf, "%s", m && *m ? m : rgf_status_string(st));
        fail("RGF:error", buf);
    }
}

static double scalar_part(const mxArray* a, int imag_part, const char* name) {
    if (!a || !mxIsNumeric(a) || mxGetNumberOfElements(a) != 1) {
        char buf[128];
        snprintf(buf, sizeof buf, "%s must be a numeric scalar", name);
        fail("RGF:input", buf);
    }
    if (mxIsComplex(a)) {
        mxComplexDouble* z = mxGetComplexDoubles(a);
        return imag_part ? z[0].imag : z[0].real;
    }
    return imag_part ? 0.0 : mxGetScalar(a);
}

static int streq_ci(const char* a, const char* b) {
    for (; *a && *b; ++a, ++b) {
        char x = (*a >= 'A' && *a <= 'Z') ? *a + 32 : *a;
        char y = (*b >= 'A' && *b <= 'Z') ? *b + 32 : *b;
        if (x != y) return 0;
    }
    return *a == *b;
}

/* setup_channels.m:102-159: names are case-insensitive; numbers are Bloch phases */
static void parse_bc(const mxArray* bc, int ndim, rgf_problem* p) {
    if (mxIsChar(bc)) {
        char s[64];
        mxGetString(bc, s, sizeof s);
        int code = -1;
        double ka = 0.0;
        if (streq_ci(s, "periodic")) code = RGF_BC_BLOCH;
        else if (streq_ci(s, "dirichlet") || streq_ci(s, "pec")) code = RGF_BC_DIRICHLET;
        else if (ndim == 2 && (streq_ci(s, "neumann") || streq_ci(s, "pmc"))) code = RGF_BC_NEUMANN;
        else if (ndim == 2 && (streq_ci(s, "dirichletneumann") || streq_ci(s, "pecpmc"))) code = RGF_BC_DIRICHLET_NEUMANN;
        else if (ndim == 2 && (streq_ci(s, "neumanndirichlet") || streq_ci(s, "pmcpec"))) code = RGF_BC_NEUMANN_DIRICHLET;
        else if (streq_ci(s, "bloch"))
            fail("RGF:input", "To use Bloch periodic boundary condition, set BC to k*a where k is the Bloch wave number and a is the periodicity");
        if (code < 0) {
            char buf[128];
            snprintf(buf, sizeof buf, "Boundary condition BC = %s is not an eligible option", s);
            fail("RGF:input", buf);
        }
        p->bc[0] = p->bc[1] = code;
        p->bloch_ka[0] = p->bloch_ka[1] = ka;
        return;
    }
    if (!mxIsNumeric(bc) || mxIsComplex(bc)) fail("RGF:input", "Bloch phase k*a has to be real");
    size_t n = mxGetNumberOfElements(bc);
    if ((ndim == 2 && n != 1) || (ndim == 3 && n != 2))
        fail("RGF:input", ndim == 2 ? "yBC must be a character vector or a numeric scalar"
                                    : "xyBC must be a character vector or a numeric vector with 2 elements");
    double* v = mxGetDoubles(bc);
    p->bc[0] = p->bc[1] = RGF_BC_BLOCH;
    p->bloch_ka[0] = v[0];
    p->bloch_ka[1] = ndim == 3 ? v[1] : 0.0;
}

typedef struct {
    double* re;
    double* im;
} eps_buf;

static int read_syst(const mxArray* syst, rgf_problem* p, eps_buf* buf) {
    if (!mxIsStruct(syst) || mxGetNumberOfElements(syst) != 1) fail("RGF:input", "syst must be a structure");
    const char* req[] = {"epsilon", "epsilon_L", "epsilon_R", "k0dx"};
    for (int q = 0; q < 4; ++q)
        if (!mxGetField(syst, 0, req[q])) {
            char b[96];
            snprintf(b, sizeof b, "syst.%s must be provided", req[q]);
            fail("RGF:input", b);
        }
    const mxArray* eps = mxGetField(syst, 0, "epsilon");
    const mxArray* xy = mxGetField(syst, 0, "xyBC");
    const mxArray* y = mxGetField(syst, 0, "yBC");
    const mwSize nd = mxGetNumberOfDimensions(eps);
    const mwSize* d = mxGetDimensions(eps);
    memset(p, 0, sizeof *p);
    int ndim = (xy && !y) || nd == 3 ? 3 : 2;
    if (ndim == 3 && !xy) fail("RGF:input", "syst.xyBC must be provided");
    if (ndim == 2 && !y) fail("RGF:input", "syst.yBC must be provided");
    if (!mxIsDouble(eps)) fail("RGF:input", "syst.epsilon must be a double array");
    p->ndim = ndim;
    p->dims[0] = (int64_t)d[0];
    p->dims[1] = (int64_t)d[1];
    p->dims[2] = ndim == 3 ? (nd == 3 ? (int64_t)d[2] : 1) : 1;
    size_t n = mxGetNumberOfElements(eps);
    buf->re = NULL;
    buf->im = NULL;
    if (mxIsComplex(eps)) {
        /* interleaved complex -> planar, as the C ABI expects */
        mxComplexDouble* z = mxGetComplexDoubles(eps);
        buf->re = (double*)mxMalloc(n * sizeof(double));
        buf->im = (double*)mxMalloc(n * sizeof(double));
        for (size_t k = 0; k < n; ++k) {
            buf->re[k] = z[k].real;
            buf->im[k] = z[k].imag;
        }
        p->eps_re = buf->re;
        p->eps_im = buf->im;
    } else {
        p->eps_re = mxGetDoubles(eps);
        p->eps_im = NULL;
    }
    p->eps_L[0] = scalar_part(mxGetField(syst, 0, "epsilon_L"), 0, "syst.epsilon_L");
    p->eps_L[1] = scalar_part(mxGetField(syst, 0, "epsilon_L"), 1, "syst.epsilon_L");
    p->eps_R[0] = scalar_part(mxGetField(syst, 0, "epsilon_R"), 0, "syst.epsilon_R");
    p->eps_R[1] = scalar_part(mxGetField(syst, 0, "epsilon_R"), 1, "syst.epsilon_R");
    p->k0dx[0] = scalar_part(mxGetField(syst, 0, "k0dx"), 0, "syst.k0dx");
    p->k0dx[1] = scalar_part(mxGetField(syst, 0, "k0dx"), 1, "syst.k0dx");
    parse_bc(ndim == 3 ? xy : y, ndim, p);
    return ndim;
}

typedef struct {
    int32_t* a[4];
} sel_buf;

/* cal_smatrix_RGF.m:103-175: cell {'left','L','right','R'} or struct of index vectors */
static void read_side(const mxArray* spec, const char* kind, int32_t** ind_L, int32_t* n_L, int32_t** ind_R,
                      int32_t* n_R, int32_t** keepL, int32_t** keepR) {
    *ind_L = *ind_R = NULL;
    *n_L = *n_R = 0;
    if (mxIsCell(spec)) {
        for (size_t k = 0; k < mxGetNumberOfElements(spec); ++k) {
            const mxArray* el = mxGetCell(spec, k);
            char s[16] = {0};
            if (!el || !mxIsChar(el)) {
                mexWarnMsgIdAndTxt("RGF:cell", "element %s{%d} has wrong format and is ignored; should be one of {'left', 'right', 'L', 'R'}", kind, (int)k + 1);
                continue;
            }
            mxGetString(el, s, sizeof s);
            if (!strcmp(s, "left") || !strcmp(s, "L")) *n_L = RGF_ALL;
            else if (!strcmp(s, "right") || !strcmp(s, "R")) *n_R = RGF_ALL;
            else mexWarnMsgIdAndTxt("RGF:cell", "element %s{%d} = %s ignored; should be one of {'left', 'right', 'L', 'R'}", kind, (int)k + 1, s);
        }
        return;
    }
    if (!mxIsStruct(spec)) {
        char b[96];
        snprintf(b, sizeof b, "input argument '%s' must be a structure or a cell array", kind);
        fail("RGF:input", b);
    }
    const char* names[2] = {NULL, NULL};
    char fL[16], fR[16];
    snprintf(fL, sizeof fL, "ind_%s_L", kind);
    snprintf(fR, sizeof fR, "ind_%s_R", kind);
    names[0] = fL;
    names[1] = fR;
    for (int side = 0; side < 2; ++side) {
        const mxArray* v = mxGetField(spec, 0, names[side]);
        if (!v) continue;
        size_t n = mxGetNumberOfElements(v);
        int32_t* a = (int32_t*)mxMalloc((n ? n : 1) * sizeof(int32_t));
        double* src = mxIsDouble(v) ? mxGetDoubles(v) : NULL;
        for (size_t q = 0; q < n; ++q) {
            double x = src ? src[q] : 0.0;
            if (!src || !(x >= -2147483648.0 && x <= 2147483647.0) || x != (double)(int32_t)x) {
                /* non-integers are rejected like out-of-range indices (cal_smatrix_RGF.m:110) */
                a[q] = 0;
            } else {
                a[q] = (int32_t)x;
            }
        }
        if (side == 0) {
            *ind_L = a;
            *n_L = (int32_t)n;
            *keepL = a;
        } else {
            *ind_R = a;
            *n_R = (int32_t)n;
            *keepR = a;
        }
    }
}

static void get_opts(const mxArray* s, rgf_options* o, int32_t* every) {
    rgf_default_options(o);
    *every = 32;
    if (!s || !mxIsStruct(s)) return;
    const mxArray* f;
    if ((f = mxGetField(s, 0, "precision")) && mxIsChar(f)) {
        char b[32];
        mxGetString(f, b, sizeof b);
        o->precision = streq_ci(b, "fp64") ? RGF_PRECISION_FP64 : RGF_PRECISION_FP32_REFINED;
    }
    if ((f = mxGetField(s, 0, "refine_tol"))) o->refine_tol = mxGetScalar(f);
    if ((f = mxGetField(s, 0, "refine_s_tol"))) o->refine_s_tol = mxGetScalar(f);
    if ((f = mxGetField(s, 0, "max_refine_iter"))) o->max_refine_iter = (int32_t)mxGetScalar(f);
    if ((f = mxGetField(s, 0, "checkpoint_every"))) *every = (int32_t)mxGetScalar(f);
}

/* n complex values (interleaved) as an r x c array; real when every
 * imaginary part is zero, which is how MATLAB stores such results */
static mxArray* complex_array(mwSize r, mwSize c, const double* z) {
    const size_t n = (size_t)r * c;
    int cx = 0;
    for (size_t k = 0; k < n; ++k) cx |= z[2 * k + 1] != 0.0;
    mxArray* a = mxCreateDoubleMatrix(r, c, cx ? mxCOMPLEX : mxREAL);
    if (cx) {
        mxComplexDouble* d = mxGetComplexDoubles(a);
        for (size_t k = 0; k < n; ++k) {
            d[k].real = z[2 * k];
            d[k].imag = z[2 * k + 1];
        }
    } else {
        double* d = mxGetDoubles(a);
        for (size_t k = 0; k < n; ++k) d[k] = z[2 * k];
    }
    return a;
}

static mxArray* real_array(mwSize r, mwSize c, const double* v) {
    mxArray* a = mxCreateDoubleMatrix(r, c, mxREAL);
    if (r > 0 && c > 0) memcpy(mxGetDoubles(a), v, sizeof(double) * r * c);
    return a;
}

static mxArray* index_row(int32_t n, const int32_t* ind) {
    mxArray* a = mxCreateDoubleMatrix(1, n, mxREAL);
    double* d = mxGetDoubles(a);
    for (int32_t k = 0; k < n; ++k) d[k] = ind[k];
    return a;
}

/* setup_longitudinal's struct (setup_channels.m:179-236,
 * setup_channels_3D.m:146-188); field order as the reference builds it */
static mxArray* side_struct(const rgf_problem* p, const double* kl, const int32_t* ip, int32_t np,
                            const double* mu, const int32_t* conj, int has_conj) {
    const int three = p->ndim == 3;
    const char* f2[] = {"kxdx", "ind_prop", "N_prop", "sqrt_mu", "ind_prop_conj"};
    const char* f3[] = {"kzdx", "ind_prop", "N_prop", "sqrt_mu"};
    mxArray* s = mxCreateStructMatrix(1, 1, three ? 4 : (has_conj ? 5 : 4), three ? f3 : f2);
    const mwSize r = three ? (mwSize)p->dims[0] : 1, c = three ? (mwSize)p->dims[1] : (mwSize)p->dims[0];
    mxSetField(s, 0, three ? "kzdx" : "kxdx", complex_array(r, c, kl));
    mxSetField(s, 0, "ind_prop", index_row(np, ip));
    mxSetField(s, 0, "N_prop", mxCreateDoubleScalar(np));
    mxSetField(s, 0, "sqrt_mu", complex_array(1, np, mu));
    if (!three && has_conj) mxSetField(s, 0, "ind_prop_conj", index_row(np, conj));
    return s;
}

/* cal_smatrix_RGF_mex('channels', syst) */
static mxArray* channels_struct(const rgf_problem* p) {
    const int64_t N = p->dims[0] * (p->ndim == 3 ? p->dims[1] : 1);
    double* kt1 = (double*)mxMalloc(sizeof(double) * N);
    double* kt2 = (double*)mxMalloc(sizeof(double) * N);
    double* kl[2];
    double* mu[2];
    int32_t* ip[2];
    int32_t* cj[2];
    for (int k = 0; k < 2; ++k) {
        kl[k] = (double*)mxMalloc(sizeof(double) * 2 * N);
        mu[k] = (double*)mxMalloc(sizeof(double) * 2 * N);
        ip[k] = (int32_t*)mxMalloc(sizeof(int32_t) * N);
        cj[k] = (int32_t*)mxMalloc(sizeof(int32_t) * N);
    }
    rgf_channels ch = {kt1, kt2, kl[0], kl[1], ip[0], ip[1], mu[0], mu[1], cj[0], cj[1], 0, 0, 0};
    check(rgf_setup_channels(p, &ch), NULL);
    mxArray* s;
    if (p->ndim == 3) {
        const char* f[] = {"kxdx", "kydx", "fun_phi", "L", "R"};
        s = mxCreateStructMatrix(1, 1, 5, f);
        mxSetField(s, 0, "kxdx", real_array((mwSize)p->dims[0], (mwSize)p->dims[1], kt1));
        mxSetField(s, 0, "kydx", real_array((mwSize)p->dims[0], (mwSize)p->dims[1], kt2));
    } else {
        const char* f[] = {"kydx", "fun_phi", "L", "R"};
        s = mxCreateStructMatrix(1, 1, 4, f);
        mxSetField(s, 0, "kydx", real_array(1, (mwSize)N, kt1));
    }
    mxSetField(s, 0, "fun_phi", mxCreateDoubleMatrix(0, 0, mxREAL));
    mxSetField(s, 0, "L", side_struct(p, kl[0], ip[0], ch.n_prop_L, mu[0], cj[0], ch.has_conj));
    mxSetField(s, 0, "R", side_struct(p, kl[1], ip[1], ch.n_prop_R, mu[1], cj[1], ch.has_conj));
    for (int k = 0; k < 2; ++k) {
        mxFree(kl[k]);
        mxFree(mu[k]);
        mxFree(ip[k]);
        mxFree(cj[k]);
    }
    mxFree(kt1);
    mxFree(kt2);
    return s;
}

static const double* real_vector(const mxArray* a, const char* name, size_t* n) {
    if (!a || !mxIsDouble(a) || mxIsComplex(a)) {
        char b[96];
        snprintf(b, sizeof b, "%s must be a real numeric array", name);
        fail("RGF:input", b);
    }
    *n = mxGetNumberOfElements(a);
    return mxGetDoubles(a);
}

/* cal_smatrix_RGF_mex('phi', syst, kydx) / ('phi', syst, kxdx, kydx):
 * 2D: ny x numel(kydx); 3D: nx x ny (x fast) for scalar wavenumbers, nx x ny x K
 * for K pairs */
static mxArray* profiles(const rgf_problem* p, int nrhs, const mxArray* prhs[]) {
    const int three = p->ndim == 3;
    if (nrhs != (three ? 4 : 3))
        fail("RGF:input", three ? "usage: cal_smatrix_RGF_mex('phi', syst, kxdx, kydx)"
                                : "usage: cal_smatrix_RGF_mex('phi', syst, kydx)");
    size_t n1, n2 = 0;
    const double* k1 = real_vector(prhs[2], three ? "kxdx" : "kydx", &n1);
    const double* k2 = three ? real_vector(prhs[3], "kydx", &n2) : NULL;
    if (three && n1 != n2) fail("RGF:input", "kxdx and kydx must have the same number of elements");
    const int64_t N = p->dims[0] * (three ? p->dims[1] : 1);
    double* z = (double*)mxMalloc(sizeof(double) * 2 * (size_t)N * (n1 ? n1 : 1));
    check(rgf_transverse_profiles(p, (int64_t)n1, k1, k2, z), NULL);
    mxArray* a;
    if (three) {
        const mwSize d[3] = {(mwSize)p->dims[0], (mwSize)p->dims[1], (mwSize)n1};
        int cx = 0;
        for (size_t k = 0; k < (size_t)N * n1; ++k) cx |= z[2 * k + 1] != 0.0;
        a = mxCreateNumericArray(n1 == 1 ? 2 : 3, d, mxDOUBLE_CLASS, cx ? mxCOMPLEX : mxREAL);
        if (cx) {
            mxComplexDouble* o = mxGetComplexDoubles(a);
            for (size_t k = 0; k < (size_t)N * n1; ++k) {
                o[k].real = z[2 * k];
                o[k].imag = z[2 * k + 1];
            }
        } else {
            double* o = mxGetDoubles(a);
            for (size_t k = 0; k < (size_t)N * n1; ++k) o[k] = z[2 * k];
        }
    } else {
        a = complex_array((mwSize)N, (mwSize)n1, z);
    }
    mxFree(z);
    return a;
}

void mexFunction(int nlhs, mxArray* plhs[], int nrhs, const mxArray* prhs[]) {
    int gradient = nrhs >= 1 && mxIsChar(prhs[0]);
    int a0 = gradient ? 1 : 0;
    if (gradient) {
        char mode[16] = {0};
        mxGetString(prhs[0], mode, sizeof mode);
        const int chan = !strcmp(mode, "channels"), phi = !strcmp(mode, "phi");
        if (chan || phi) {
            if (nrhs < 2) fail("RGF:input", chan ? "usage: cal_smatrix_RGF_mex('channels', syst)"
                                                 : "usage: cal_smatrix_RGF_mex('phi', syst, kydx)");
            rgf_problem p;
            eps_buf eb;
            read_syst(prhs[1], &p, &eb);
            plhs[0] = chan ? channels_struct(&p) : profiles(&p, nrhs, prhs);
            if (eb.re) mxFree(eb.re);
            if (eb.im) mxFree(eb.im);
            return;
        }
        if (strcmp(mode, "gradient")) fail("RGF:input", "the mode argument must be 'gradient', 'channels' or 'phi'");
    }
    if (nrhs < a0 + 3) fail("RGF:input", "usage: cal_smatrix_RGF_mex(syst, out, in [, opts])");
    rgf_problem p;
    eps_buf eb;
    read_syst(prhs[a0], &p, &eb);
    rgf_selection sel;
    int32_t *k1 = NULL, *k2 = NULL, *k3 = NULL, *k4 = NULL;
    read_side(prhs[a0 + 1], "out", (int32_t**)&sel.ind_out_L, &sel.n_out_L, (int32_t**)&sel.ind_out_R,
              &sel.n_out_R, &k1, &k2);
    read_side(prhs[a0 + 2], "in", (int32_t**)&sel.ind_in_L, &sel.n_in_L, (int32_t**)&sel.ind_in_R, &sel.n_in_R,
              &k3, &k4);
    rgf_options opt;
    int32_t every;
    get_opts(nrhs > a0 + 3 ? prhs[a0 + 3] : NULL, &opt, &every);
    rgf_info info;
    memset(&info, 0, sizeof info);
    int32_t nL, nR, nout, nin;
    /* the problem and the selection are checked before any device is touched */
    check(rgf_query(&p, &sel, &nL, &nR, &nout, &nin), NULL);
    info.n_prop_L = nL;
    info.n_prop_R = nR;
    if (!gradient && (nout == 0 || nin == 0)) {
        plhs[0] = mxCreateDoubleMatrix(nout, nin, mxREAL); /* cal_smatrix_RGF.m:181-184 */
    } else {
        if (!g_handle) {
            rgf_status st = rgf_create(&g_handle, -1);
            if (st != RGF_OK) fail("RGF:device", rgf_last_error(NULL));
            mexAtExit(release_handle);
        }
        if (!gradient) {
            plhs[0] = mxCreateDoubleMatrix(nout, nin, mxCOMPLEX);
            check(rgf_solve(g_handle, &p, &sel, &opt, (double*)mxGetComplexDoubles(plhs[0]), &info), g_handle);
        } else {
            size_t n = (size_t)(p.dims[0] * p.dims[1] * p.dims[2]);
            double T = 0;
            mxArray* g = mxCreateNumericArray(p.ndim == 3 ? 3 : 2, (const mwSize[]){(mwSize)p.dims[0], (mwSize)p.dims[1], (mwSize)p.dims[2]}, mxDOUBLE_CLASS, p.eps_im ? mxCOMPLEX : mxREAL);
            double* gre = (double*)mxMalloc(n * sizeof(double));
            double* gim = (double*)mxMalloc(n * sizeof(double));
            check(rgf_transmission_gradient(g_handle, &p, &sel, &opt, every, &T, gre, gim, &info), g_handle);
            if (p.eps_im) {
                mxComplexDouble* z = mxGetComplexDoubles(g);
                for (size_t k = 0; k < n; ++k) {
                    z[k].real = gre[k];
                    z[k].imag = gim[k];
                }
            } else {
                memcpy(mxGetDoubles(g), gre, n * sizeof(double));
            }
            mxFree(gre);
            mxFree(gim);
            plhs[0] = mxCreateDoubleScalar(T);
            if (nlhs > 1) plhs[1] = g;
            else mxDestroyArray(g);
        }
    }
    if (!gradient && nlhs > 1) {
        const char* f[] = {"n_prop_L", "n_prop_R", "refine_iters", "refine_residual", "seconds_total"};
        plhs[1] = mxCreateStructMatrix(1, 1, 5, f);
        mxSetField(plhs[1], 0, "n_prop_L", mxCreateDoubleScalar(info.n_prop_L));
        mxSetField(plhs[1], 0, "n_prop_R", mxCreateDoubleScalar(info.n_prop_R));
        mxSetField(plhs[1], 0, "refine_iters", mxCreateDoubleScalar(info.refine_iters));
        mxSetField(plhs[1], 0, "refine_residual", mxCreateDoubleScalar(info.refine_residual));
        mxSetField(plhs[1], 0, "seconds_total", mxCreateDoubleScalar(info.seconds_total));
    }
    if (k1) mxFree(k1);
    if (k2) mxFree(k2);
    if (k3) mxFree(k3);
    if (k4) mxFree(k4);
    if (eb.re) mxFree(eb.re);
    if (eb.im) mxFree(eb.im);
}
