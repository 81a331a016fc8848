// One of the four drop-in libraries: exports myProjectfun (waveCuda.h:10-21) with
// C linkage and forwards to the engine entry point named by TOTEM_VARIANT
// (totem_project_lobato, totem_project_peng, totem_potential_lobato or
// totem_potential_peng).  Like the reference it returns nothing; a failure is
// printed and left in totem_last_error() for the mex wrapper to raise.
#include <cstdio>

#include "../../include/totem_abi.h"

#ifndef TOTEM_VARIANT
#error "define TOTEM_VARIANT to the engine entry point"
#endif

extern "C" void myProjectfun(TOTEM_MYPROJECTFUN_PARAMS) {
  const int rc = TOTEM_VARIANT(my_real_w, my_imag_w, aper, atom_slice, absorp_n, atom_slice_i, absorp_n_i,
                               proj_coff_mat, proj_coff_mat_i, kx, ky, s_2, my_real_const, my_img_const, gfsf,
                               Height, Width, Depth, P_Height, P_Width, Slices, step, beginRow, beginCol,
                               width_red, sigma, w_num, aper_num, parameter, AperTrue, mid_ceng, ceng_len, aper2,
                               series_n_corr, series_n_i_corr, ele_n_corr, ele_n_i_corr, corr_info_matrix,
                               cim_len, return_result, mid_ceng_mat, potentialx, potentialy);
  if (rc != TOTEM_OK) std::fprintf(stderr, "myProjectfun: %s\n", totem_last_error());
}
