/*
 * Host implementation of the MEX API subset in mex.h, plus constructors and a
 * protected call for driving the gateway from Python (ctypes).  Arrays are
 * column-major; complex data is interleaved (mxComplexDouble), as in MATLAB
 * R2018a and later.  Only 1x1 structs are supported (all the gateway uses).
 */
#include <setjmp.h>
#include <stdarg.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "mex.h"

#define MAXF 32

struct mxArray_tag {
    mxClassID cls;
    mwSize ndim;
    mwSize dims[8];
    int cplx;
    double* re;               /* real double data (cplx == 0) */
    mxComplexDouble* z;       /* interleaved complex data (cplx == 1) */
    char* str;                /* char arrays */
    int nf;                   /* struct fields */
    char* fname[MAXF];
    mxArray* fval[MAXF];
    mxArray** cell;           /* cell elements */
};

static jmp_buf g_jmp;
static int g_in_call = 0;
static char g_err_id[128], g_err[1024], g_warn[4096];
static void (*g_atexit)(void) = NULL;

static mxArray* alloc_array(mxClassID cls, mwSize ndim, const mwSize* dims) {
    mxArray* a = (mxArray*)calloc(1, sizeof *a);
    a->cls = cls;
    a->ndim = ndim < 2 ? 2 : ndim;
    a->dims[0] = a->dims[1] = 1;
    for (mwSize k = 0; k < ndim && k < 8; ++k) a->dims[k] = dims[k];
    /* trailing singleton dimensions are dropped, as MATLAB does */
    while (a->ndim > 2 && a->dims[a->ndim - 1] == 1) a->ndim--;
    return a;
}

void* mxMalloc(mwSize n) { return malloc(n ? n : 1); }
void mxFree(void* p) { free(p); }

mxArray* mxCreateNumericArray(mwSize ndim, const mwSize* dims, mxClassID cls, mxComplexity c) {
    (void)cls;
    mxArray* a = alloc_array(mxDOUBLE_CLASS, ndim, dims);
    a->cplx = c == mxCOMPLEX;
    size_t n = mxGetNumberOfElements(a);
    if (a->cplx)
        a->z = (mxComplexDouble*)calloc(n ? n : 1, sizeof(mxComplexDouble));
    else
        a->re = (double*)calloc(n ? n : 1, sizeof(double));
    return a;
}

mxArray* mxCreateDoubleMatrix(mwSize m, mwSize n, mxComplexity c) {
    mwSize d[2] = {m, n};
    return mxCreateNumericArray(2, d, mxDOUBLE_CLASS, c);
}

mxArray* mxCreateDoubleScalar(double v) {
    mxArray* a = mxCreateDoubleMatrix(1, 1, mxREAL);
    a->re[0] = v;
    return a;
}

mxArray* mxCreateStructMatrix(mwSize m, mwSize n, int nfields, const char** names) {
    mwSize d[2] = {m, n};
    mxArray* a = alloc_array(mxSTRUCT_CLASS, 2, d);
    for (int k = 0; k < nfields && k < MAXF; ++k) {
        a->fname[k] = strdup(names[k]);
        a->fval[k] = NULL;
    }
    a->nf = nfields;
    return a;
}

void mxDestroyArray(mxArray* a) {
    if (!a) return;
    free(a->re);
    free(a->z);
    free(a->str);
    for (int k = 0; k < a->nf; ++k) {
        free(a->fname[k]);
        mxDestroyArray(a->fval[k]);
    }
    if (a->cell) {
        size_t n = mxGetNumberOfElements(a);
        for (size_t k = 0; k < n; ++k) mxDestroyArray(a->cell[k]);
        free(a->cell);
    }
    free(a);
}

bool mxIsChar(const mxArray* a) { return a && a->cls == mxCHAR_CLASS; }
bool mxIsStruct(const mxArray* a) { return a && a->cls == mxSTRUCT_CLASS; }
bool mxIsCell(const mxArray* a) { return a && a->cls == mxCELL_CLASS; }
bool mxIsNumeric(const mxArray* a) { return a && a->cls == mxDOUBLE_CLASS; }
bool mxIsDouble(const mxArray* a) { return a && a->cls == mxDOUBLE_CLASS; }
bool mxIsComplex(const mxArray* a) { return a && a->cls == mxDOUBLE_CLASS && a->cplx; }

size_t mxGetNumberOfElements(const mxArray* a) {
    size_t n = 1;
    for (mwSize k = 0; k < a->ndim; ++k) n *= a->dims[k];
    return n;
}
mwSize mxGetNumberOfDimensions(const mxArray* a) { return a->ndim; }
const mwSize* mxGetDimensions(const mxArray* a) { return a->dims; }

double mxGetScalar(const mxArray* a) {
    if (!a || a->cls != mxDOUBLE_CLASS || mxGetNumberOfElements(a) == 0) return 0.0;
    return a->cplx ? a->z[0].real : a->re[0];
}
double* mxGetDoubles(const mxArray* a) { return a && a->cls == mxDOUBLE_CLASS && !a->cplx ? a->re : NULL; }
mxComplexDouble* mxGetComplexDoubles(const mxArray* a) {
    return a && a->cls == mxDOUBLE_CLASS && a->cplx ? a->z : NULL;
}

int mxGetString(const mxArray* a, char* buf, mwSize len) {
    if (!a || a->cls != mxCHAR_CLASS || len == 0) return 1;
    size_t n = strlen(a->str);
    size_t c = n < len - 1 ? n : len - 1;
    memcpy(buf, a->str, c);
    buf[c] = 0;
    return n > c;
}

mxArray* mxGetField(const mxArray* s, mwIndex i, const char* name) {
    if (!mxIsStruct(s) || i != 0) return NULL;
    for (int k = 0; k < s->nf; ++k)
        if (!strcmp(s->fname[k], name)) return s->fval[k];
    return NULL;
}

void mxSetField(mxArray* s, mwIndex i, const char* name, mxArray* v) {
    if (!mxIsStruct(s) || i != 0) return;
    for (int k = 0; k < s->nf; ++k)
        if (!strcmp(s->fname[k], name)) {
            s->fval[k] = v;
            return;
        }
}

mxArray* mxGetCell(const mxArray* c, mwIndex i) { return mxIsCell(c) ? c->cell[i] : NULL; }

void mexErrMsgIdAndTxt(const char* id, const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    snprintf(g_err_id, sizeof g_err_id, "%s", id);
    vsnprintf(g_err, sizeof g_err, fmt, ap);
    va_end(ap);
    if (g_in_call) longjmp(g_jmp, 1);
    fprintf(stderr, "mexErrMsgIdAndTxt outside a call: %s: %s\n", g_err_id, g_err);
    abort();
}

void mexWarnMsgIdAndTxt(const char* id, const char* fmt, ...) {
    char m[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(m, sizeof m, fmt, ap);
    va_end(ap);
    size_t used = strlen(g_warn);
    snprintf(g_warn + used, sizeof g_warn - used, "%s: %s\n", id, m);
}

int mexAtExit(void (*fn)(void)) {
    g_atexit = fn;
    return 0;
}

/* ---- constructors and access for the Python side ---- */

mxArray* h_double(int ndim, const int64_t* dims, int cplx, const double* re, const double* im) {
    mwSize d[8];
    for (int k = 0; k < ndim && k < 8; ++k) d[k] = (mwSize)dims[k];
    mxArray* a = mxCreateNumericArray((mwSize)ndim, d, mxDOUBLE_CLASS, cplx ? mxCOMPLEX : mxREAL);
    size_t n = mxGetNumberOfElements(a);
    for (size_t k = 0; k < n; ++k) {
        if (cplx) {
            a->z[k].real = re[k];
            a->z[k].imag = im ? im[k] : 0.0;
        } else {
            a->re[k] = re[k];
        }
    }
    return a;
}

mxArray* h_char(const char* s) {
    mwSize d[2] = {1, strlen(s)};
    mxArray* a = alloc_array(mxCHAR_CLASS, 2, d);
    a->str = strdup(s);
    return a;
}

mxArray* h_struct(void) { return mxCreateStructMatrix(1, 1, 0, NULL); }

void h_set(mxArray* s, const char* name, mxArray* v) {
    for (int k = 0; k < s->nf; ++k)
        if (!strcmp(s->fname[k], name)) {
            s->fval[k] = v;
            return;
        }
    if (s->nf < MAXF) {
        s->fname[s->nf] = strdup(name);
        s->fval[s->nf] = v;
        s->nf++;
    }
}

mxArray* h_cell(int n) {
    mwSize d[2] = {1, (mwSize)n};
    mxArray* a = alloc_array(mxCELL_CLASS, 2, d);
    a->cell = (mxArray**)calloc(n ? n : 1, sizeof(mxArray*));
    return a;
}

void h_cell_set(mxArray* c, int k, mxArray* v) { c->cell[k] = v; }

/* 0: mexFunction returned; 1: it raised an error (h_error_id / h_error) */
int h_call(int nlhs, mxArray** plhs, int nrhs, mxArray** prhs) {
    g_err_id[0] = g_err[0] = g_warn[0] = 0;
    for (int k = 0; k < nlhs; ++k) plhs[k] = NULL;
    g_in_call = 1;
    if (setjmp(g_jmp)) {
        g_in_call = 0;
        return 1;
    }
    mexFunction(nlhs, plhs, nrhs, (const mxArray**)prhs);
    g_in_call = 0;
    return 0;
}

const char* h_error_id(void) { return g_err_id; }
const char* h_error(void) { return g_err; }
const char* h_warnings(void) { return g_warn; }
void h_exit(void) {
    if (g_atexit) g_atexit();
    g_atexit = NULL;
}

int h_class(const mxArray* a) { return a ? (int)a->cls : -1; }
int h_ndim(const mxArray* a) { return (int)a->ndim; }
int64_t h_dim(const mxArray* a, int k) { return k < (int)a->ndim ? (int64_t)a->dims[k] : 1; }
int h_is_complex(const mxArray* a) { return a->cplx; }
mxArray* h_field(const mxArray* s, const char* name) { return mxGetField(s, 0, name); }

void h_read(const mxArray* a, double* re, double* im) {
    size_t n = mxGetNumberOfElements(a);
    for (size_t k = 0; k < n; ++k) {
        re[k] = a->cplx ? a->z[k].real : a->re[k];
        im[k] = a->cplx ? a->z[k].imag : 0.0;
    }
}

int h_nfields(const mxArray* s) { return mxIsStruct(s) ? s->nf : 0; }
const char* h_fieldname(const mxArray* s, int k) { return k < s->nf ? s->fname[k] : NULL; }
