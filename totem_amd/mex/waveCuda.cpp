// MATLAB mex gateway for the MI355X engine: the drop-in for the reference's
// waveCuda.cpp (waveCuda.cpp:4-101).  Same 41 inputs and 4 outputs, same
// prhs -> myProjectfun mapping (absorption flags at prhs[30..31], cimlen at
// prhs[40], 1-based beginRow/beginCol converted to 0-based), linked against one
// of libwaveCudalobato / libwaveCudaPengAndCorr / ...P.  Unlike the reference,
// an engine failure raises a MATLAB error instead of returning zeros.
#include "mex.h"
#include "totem_abi.h"

static double* pr(const mxArray* a) { return (double*)mxGetPr(a); }

void mexFunction(int nlhs, mxArray* plhs[], int nrhs, const mxArray* prhs[]) {
  if (nrhs != 41) {
    mexErrMsgTxt("Invalid number of input arguments ");
    return;
  }
  if (nlhs != 4) {
    mexErrMsgTxt("Invalid number of outputs ");
    return;
  }
  const int absorp_flag = (int)mxGetScalar(prhs[30]);
  const int absorp_i_flag = (int)mxGetScalar(prhs[31]);
  const int Height = (int)mxGetScalar(prhs[15]);
  const int Width = (int)mxGetScalar(prhs[16]);
  const int Depth = (int)mxGetScalar(prhs[17]);
  const int P_Height = (int)mxGetScalar(prhs[18]);
  const int P_Width = (int)mxGetScalar(prhs[19]);
  const int Slices = (int)mxGetScalar(prhs[20]);
  const int Step = (int)mxGetScalar(prhs[21]);
  const int beginRow = (int)mxGetScalar(prhs[22]);
  const int beginCol = (int)mxGetScalar(prhs[23]);
  const int width_red = (int)mxGetScalar(prhs[24]);
  const double sigma = mxGetScalar(prhs[25]);
  const int w_num = (int)mxGetScalar(prhs[26]);
  const int aper_num = (int)mxGetScalar(prhs[27]);
  const double parameter = mxGetScalar(prhs[28]);
  const int ceng_len = (int)mxGetScalar(prhs[33]);
  const int cimlen = (int)mxGetScalar(prhs[40]);

  plhs[0] = mxCreateNumericMatrix(Depth, aper_num, mxDOUBLE_CLASS, mxREAL);
  plhs[1] = mxCreateNumericMatrix((mwSize)Depth * aper_num, ceng_len, mxDOUBLE_CLASS, mxREAL);
  plhs[2] = mxCreateNumericMatrix((mwSize)P_Height * P_Width, Slices, mxDOUBLE_CLASS, mxREAL);
  plhs[3] = mxCreateNumericMatrix((mwSize)P_Height * P_Width, Slices, mxDOUBLE_CLASS, mxREAL);

  myProjectfun(pr(prhs[0]), pr(prhs[1]), pr(prhs[2]), pr(prhs[3]), absorp_flag == 1 ? pr(prhs[4]) : NULL,
               pr(prhs[5]), absorp_i_flag == 1 ? pr(prhs[6]) : NULL, pr(prhs[7]), pr(prhs[8]), pr(prhs[9]),
               pr(prhs[10]), pr(prhs[11]), pr(prhs[12]), pr(prhs[13]), pr(prhs[14]), Height, Width, Depth,
               P_Height, P_Width, Slices, Step, beginRow - 1, beginCol - 1, width_red, sigma, w_num, aper_num,
               parameter, pr(prhs[29]), (int*)mxGetPr(prhs[32]), ceng_len, pr(prhs[34]), pr(prhs[35]),
               pr(prhs[36]), pr(prhs[37]), pr(prhs[38]), cimlen != 0 ? pr(prhs[39]) : NULL, cimlen, pr(plhs[0]),
               pr(plhs[1]), pr(plhs[2]), pr(plhs[3]));
  const char* err = totem_last_error();
  if (err && err[0]) mexErrMsgTxt(err);
}
