// Drives mexFunction from totem_amd/mex/waveCuda.cpp with synthetic inputs and a
// recording myProjectfun, then prints what the engine would have received.
#include <cstdio>
#include <vector>

#include "mex.h"
#include "totem_abi.h"

std::string g_mex_error;
static std::string g_engine_error;
extern "C" const char* totem_last_error(void) { return g_engine_error.c_str(); }

struct Seen {
  int H, W, D, PH, PW, S, step, br, bc, wr, wn, an, cl, cim;
  double sigma, parameter;
  bool absorp_null, absorp_i_null, corr_null;
  double first_real_w, first_aper2, first_mid;
} seen;

extern "C" void myProjectfun(TOTEM_MYPROJECTFUN_PARAMS) {
  seen = Seen{Height, Width, Depth, P_Height, P_Width, Slices, step, beginRow, beginCol, width_red, w_num,
              aper_num, ceng_len, cim_len, sigma, parameter, absorp_n == nullptr, absorp_n_i == nullptr,
              corr_info_matrix == nullptr, my_real_w[0], aper2[0], (double)mid_ceng[0]};
  return_result[0] = 42.0;
  if (Slices > 0) potentialx[0] = 7.0;
  if (beginRow == 99) g_engine_error = "engine failure";
}

void mexFunction(int nlhs, mxArray* plhs[], int nrhs, const mxArray* prhs[]);

int main(int argc, char** argv) {
  const int fail = argc > 1;
  std::vector<mxArray*> in;
  std::vector<double> scal = {0};
  for (int i = 0; i < 41; ++i) in.push_back(mxCreateNumericMatrix(4, 4, mxDOUBLE_CLASS, mxREAL));
  auto setv = [&](int i, double v) { in[i]->data[0] = v; };
  setv(0, 0.25);                 // my_real_w
  setv(15, 64); setv(16, 64); setv(17, 12); setv(18, 256); setv(19, 256); setv(20, 9); setv(21, 3);
  setv(22, fail ? 100 : 11); setv(23, 21); setv(24, 4); setv(25, 0.7288); setv(26, 3); setv(27, 2);
  setv(28, 0.05); setv(30, 0); setv(31, 1); setv(33, 1); setv(34, 0.5); setv(40, 0);
  // mid_ceng arrives as uint32 data reinterpreted as int (cuda_stem.m:67)
  int* mid = reinterpret_cast<int*>(in[32]->data);
  mid[0] = 5;
  mxArray* out[4];
  mexFunction(4, out, 41, const_cast<const mxArray**>(in.data()));
  std::printf("H=%d W=%d D=%d PH=%d PW=%d S=%d step=%d br=%d bc=%d wr=%d wn=%d an=%d cl=%d cim=%d\n", seen.H, seen.W,
              seen.D, seen.PH, seen.PW, seen.S, seen.step, seen.br, seen.bc, seen.wr, seen.wn, seen.an, seen.cl,
              seen.cim);
  std::printf("sigma=%g parameter=%g absorp_null=%d absorp_i_null=%d corr_null=%d w0=%g aper2=%g mid=%g\n",
              seen.sigma, seen.parameter, seen.absorp_null, seen.absorp_i_null, seen.corr_null, seen.first_real_w,
              seen.first_aper2, seen.first_mid);
  std::printf("out0=%zux%zu:%g out1=%zux%zu out2=%zux%zu:%g\n", out[0]->m, out[0]->n, out[0]->data[0], out[1]->m,
              out[1]->n, out[2]->m, out[2]->n, out[2]->data[0]);
  std::printf("error=%s\n", g_mex_error.c_str());
  // wrong argument count
  g_mex_error.clear();
  mexFunction(4, out, 40, const_cast<const mxArray**>(in.data()));
  std::printf("nrhs40=%s\n", g_mex_error.c_str());
  return 0;
}
