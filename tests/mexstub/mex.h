/* Minimal stand-in for MATLAB's mex.h, enough to compile and drive the repo's
 * own totem_amd/mex/waveCuda.cpp gateway in a CPU test (no MATLAB here). */
#pragma once
#include <cstddef>
#include <cstdlib>
#include <string>
typedef size_t mwSize;
enum mxClassID { mxDOUBLE_CLASS = 6 };
enum mxComplexity { mxREAL = 0 };
struct mxArray {
  double* data;
  size_t m, n;
};
inline double* mxGetPr(const mxArray* a) { return a->data; }
inline double mxGetScalar(const mxArray* a) { return a->data[0]; }
inline mxArray* mxCreateNumericMatrix(mwSize m, mwSize n, mxClassID, mxComplexity) {
  mxArray* a = new mxArray;
  a->m = m;
  a->n = n;
  a->data = static_cast<double*>(std::calloc(m * n + 1, sizeof(double)));
  return a;
}
extern std::string g_mex_error;
inline void mexErrMsgTxt(const char* s) { g_mex_error = s; }
