/*
 * The part of MATLAB's MEX API (R2018a interleaved-complex flavour) that the
 * gateway rgf_amd/mex/cal_smatrix_RGF_mex.c uses, implemented on the host by
 * mex_harness.c so that the gateway can be compiled and exercised without
 * MATLAB.  Errors raised through mexErrMsgIdAndTxt return to the harness's
 * h_call (as MATLAB unwinds to its prompt); warnings are collected.
 */
#ifndef RGF_TEST_MEX_H
#define RGF_TEST_MEX_H
#include <stdbool.h>
#include <stddef.h>

typedef size_t mwSize;
typedef size_t mwIndex;
typedef enum { mxUNKNOWN_CLASS = 0, mxCELL_CLASS, mxSTRUCT_CLASS, mxCHAR_CLASS, mxDOUBLE_CLASS } mxClassID;
typedef enum { mxREAL = 0, mxCOMPLEX } mxComplexity;
typedef struct {
    double real, imag;
} mxComplexDouble;
typedef struct mxArray_tag mxArray;

void* mxMalloc(mwSize n);
void mxFree(void* p);
mxArray* mxCreateDoubleScalar(double v);
mxArray* mxCreateDoubleMatrix(mwSize m, mwSize n, mxComplexity c);
mxArray* mxCreateNumericArray(mwSize ndim, const mwSize* dims, mxClassID cls, mxComplexity c);
mxArray* mxCreateStructMatrix(mwSize m, mwSize n, int nfields, const char** names);
void mxDestroyArray(mxArray* a);
bool mxIsChar(const mxArray* a);
bool mxIsStruct(const mxArray* a);
bool mxIsCell(const mxArray* a);
bool mxIsNumeric(const mxArray* a);
bool mxIsDouble(const mxArray* a);
bool mxIsComplex(const mxArray* a);
size_t mxGetNumberOfElements(const mxArray* a);
mwSize mxGetNumberOfDimensions(const mxArray* a);
const mwSize* mxGetDimensions(const mxArray* a);
double mxGetScalar(const mxArray* a);
double* mxGetDoubles(const mxArray* a);
mxComplexDouble* mxGetComplexDoubles(const mxArray* a);
int mxGetString(const mxArray* a, char* buf, mwSize len);
mxArray* mxGetField(const mxArray* s, mwIndex i, const char* name);
void mxSetField(mxArray* s, mwIndex i, const char* name, mxArray* v);
mxArray* mxGetCell(const mxArray* c, mwIndex i);
void mexErrMsgIdAndTxt(const char* id, const char* fmt, ...);
void mexWarnMsgIdAndTxt(const char* id, const char* fmt, ...);
int mexAtExit(void (*fn)(void));
void mexFunction(int nlhs, mxArray* plhs[], int nrhs, const mxArray* prhs[]);

#endif
