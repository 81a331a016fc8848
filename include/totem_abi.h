/*
 * totem_abi.h -- C ABI of the MI355X multislice STEM engine.
 *
 * Plain C types only: host pointers are double/int arrays laid out exactly as
 * the reference's MATLAB driver hands them to the mex layer (row-major after
 * the transposes in cuda_stem.m:22-79); device pointers are void*.
 *
 * Status convention for the totem_* functions: 0 on success, a negative code
 * on failure, with a message available from totem_last_error().
 */
#ifndef TOTEM_ABI_H
#define TOTEM_ABI_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define TOTEM_OK 0
#define TOTEM_ERR_ARG (-1)      /* invalid argument / shape / window */
#define TOTEM_ERR_DEVICE (-2)   /* HIP runtime error */
#define TOTEM_ERR_MEMORY (-3)   /* device allocation failed or workspace too small */
#define TOTEM_ERR_UNSUPPORTED (-4) /* grid size without a compiled FFT plan, non-separable k grids */

/* ------------------------------------------------------------------ drop-in
 * Replaces the reference's
 *   extern void myProjectfun(...)           waveCuda.h:10-21
 * implemented four times:
 *   cuda_source_lobato.cu:1018               (STEM, Lobato factors)
 *   cuda_source_peng_and_corr.cu:1217        (STEM, Peng + correction)
 *   cuda_source_lobato_potential.cu:747      (transmission only, Lobato)
 *   cuda_source_peng_and_corr_potential.cu:952 (transmission only, Peng + correction)
 * Each variant library (libwaveCudalobato.so, libwaveCudaPengAndCorr.so,
 * libwaveCudalobatoP.so, libwaveCudaPengAndCorrP.so) exports this one symbol
 * with C linkage, as the four mex products of mex_cuda.m:19-38 do.  The
 * function returns nothing (like the reference); failures are reported through
 * totem_last_error() and leave the outputs zeroed.
 */
void myProjectfun(double *my_real_w, double *my_imag_w, double *aper,
                  double *atom_slice, double *absorp_n, double *atom_slice_i,
                  double *absorp_n_i, double *proj_coff_mat, double *proj_coff_mat_i,
                  double *kx, double *ky, double *s_2, double *my_real_const,
                  double *my_img_const, double *gfsf, int Height, int Width, int Depth,
                  int P_Height, int P_Width, int Slices, int step, int beginRow, int beginCol,
                  int width_red, double sigma, int w_num, int aper_num, double parameter,
                  double *AperTrue, int *mid_ceng, int ceng_len, double *aper2,
                  double *series_n_corr, double *series_n_i_corr,
                  double *ele_n_corr, double *ele_n_i_corr,
                  double *corr_info_matrix, int cim_len,
                  double *return_result, double *mid_ceng_mat, double *potentialx, double *potentialy);

/* The four variants with a status return (same 43 parameters as above).
 * STEM variants write return_result [aper_num][Depth] and mid_ceng_mat
 * [ceng_len][aper_num][Depth]; potential variants write the transmission
 * function t = exp(i sigma/1000 V) as potentialx/potentialy [Slices][P_Height][P_Width]. */
#define TOTEM_MYPROJECTFUN_PARAMS \
  double *my_real_w, double *my_imag_w, double *aper, double *atom_slice, double *absorp_n, \
  double *atom_slice_i, double *absorp_n_i, double *proj_coff_mat, double *proj_coff_mat_i, \
  double *kx, double *ky, double *s_2, double *my_real_const, double *my_img_const, double *gfsf, \
  int Height, int Width, int Depth, int P_Height, int P_Width, int Slices, int step, int beginRow, \
  int beginCol, int width_red, double sigma, int w_num, int aper_num, double parameter, \
  double *AperTrue, int *mid_ceng, int ceng_len, double *aper2, double *series_n_corr, \
  double *series_n_i_corr, double *ele_n_corr, double *ele_n_i_corr, double *corr_info_matrix, \
  int cim_len, double *return_result, double *mid_ceng_mat, double *potentialx, double *potentialy

int totem_project_lobato(TOTEM_MYPROJECTFUN_PARAMS);   /* cuda_source_lobato.cu:1018 */
int totem_project_peng(TOTEM_MYPROJECTFUN_PARAMS);     /* cuda_source_peng_and_corr.cu:1217 */
int totem_potential_lobato(TOTEM_MYPROJECTFUN_PARAMS); /* cuda_source_lobato_potential.cu:747 */
int totem_potential_peng(TOTEM_MYPROJECTFUN_PARAMS);   /* cuda_source_peng_and_corr_potential.cu:952 */

/* Message of the last failure on the calling thread ("" if none). */
const char *totem_last_error(void);
/* ABI version of this library. */
int totem_abi_version(void);
/* 1 if an n-point FFT plan is compiled in. */
int totem_fft_size_supported(int n);

/* ------------------------------------------------------------ potential build
 * The potential part of myProjectfun (cal_p/initKernel_P, cuda_source_lobato.cu:585-956;
 * p_corr, cuda_source_peng_and_corr.cu:652-783) with the transmission stack
 * left on the device.  Host arrays in the myProjectfun layouts. */
typedef struct totem_potential_desc {
  int paraflag;                   /* 'l' Lobato (14/26-column rows), 'p' Peng, 'n' Peng + correction (13/25) */
  int np;                         /* P_Height == P_Width */
  int slices;
  const double *atom_slice;       /* [S] rows of proj_coff_mat per slice */
  const double *proj_coff_mat;    /* [sum][14 | 13] */
  const double *absorp_n;         /* [sum][10] or NULL */
  const double *atom_slice_i;     /* [S] or NULL */
  const double *proj_coff_mat_i;  /* [sum_i][26 | 25] or NULL */
  const double *absorp_n_i;       /* [sum_i][10] or NULL */
  const double *series_n_corr;    /* [S] or NULL */
  const double *ele_n_corr;       /* [sum][5] = x y B occ*w class(1-based) */
  const double *series_n_i_corr;  /* [S] or NULL */
  const double *ele_n_i_corr;     /* [sum][8] = x y B occ w1 class1 w2 class2 */
  const double *corr_info_matrix; /* [cim_len][np][np] or NULL */
  int cim_len;
  const double *kx, *ky, *s_2;    /* [np][np] centred grids (kx varies along columns) */
  const double *aper_true;        /* [np][np] centred, multiplies V(k) */
  double sigma;                   /* 1/(kV Angstrom); the engine uses sigma/1000 */
  double parameter;               /* PARAMETER */
} totem_potential_desc;

/* Builds t = exp(i sigma/1000 V) into t_dev (complex64 [S][np][np], device).
 * If v_re/v_im are non-NULL the real-space potential V (FP32 values) is also
 * copied to them (host, [S][np][np]). */
int totem_build_transmission(const totem_potential_desc *desc, void *t_dev, double *v_re, double *v_im,
                             void *stream);

/* Host-side plan of the same build, no device work: the form-factor classes, the
 * terms over all slices, split = 1 when Peng rows were split into one class per
 * Gaussian term (done only when that lowers the accumulation's cost), and
 * paired = 1 when kx/ky pair up (k[np - i] = -k[i]) so that the accumulation forms
 * a quarter of the reciprocal-space tiles and mirrors the rest.  Diagnostic; no
 * counterpart in the reference.  Any output pointer may be NULL. */
int totem_potential_plan_stats(const totem_potential_desc *desc, int *classes, long long *terms, int *split,
                               int *paired);

/* --------------------------------------------------------- device STEM job
 * The propagation part of myProjectfun with every array already in HBM. */
typedef struct totem_stem_job {
  int n;                 /* probe window side (Height == Width) */
  int np;                /* potential side */
  int slices;
  const void *t;         /* complex64 [S][np][np] */
  const void *probes;    /* complex64 [M][n][n] */
  int n_probes;          /* M (focal-spread members) */
  const void *prop;      /* complex64 [n][n], zero frequency at [0][0] */
  const double *aper;    /* [A][n][n] centred masks */
  const double *aper2;   /* [A][n][n] ifftshifted masks, required when n_mid > 0 */
  int n_aper;            /* A */
  const int *mid;        /* host [L], 1-based slice numbers */
  int n_mid;             /* L */
  int depth;             /* D scan positions */
  int width_red;         /* scan columns */
  int step;              /* scan step in pixels */
  int begin_row, begin_col; /* 0-based window origin of position 0 */
  int wrap;              /* 1: periodic window (mod np); 0: window must lie inside the potential */
  double *det;           /* device [(L+1)][A][M*D]: mid layers 0..L-1, exit face at L;
                            column m*D + d (no focal-spread weighting applied) */
  void *exit_wave;       /* device complex64 [M*D][n][n] or NULL: the exit-face wave in reciprocal
                            space, Psi(k) = FFT2(psi) (zero frequency at [0][0], unnormalised),
                            i.e. what the reference's final forward FFT produces before its
                            quadrant swap (cuda_source_lobato.cu:1405-1415) */
} totem_stem_job;

#define TOTEM_RUN_GRAPH 1     /* capture the slice loop of a batch in a hipGraph and replay it */
#define TOTEM_RUN_TIMING 2    /* time every rows/columns pass with HIP events (no graph; synchronises) */

/* Probes in flight per launch that keep their n x n waves inside the GPU's
 * memory-side cache (the batch myProjectfun uses; no counterpart in the
 * reference, which fills free memory, cuda_source_lobato.cu:1249-1260). */
int totem_stem_preferred_batch(int n);
/* Measures every plan of n x n grids with probes per launch around the cache-fit
 * rule on the current device (a synthetic job through `slices` <= 64 slices, its
 * propagator zero beyond the 2/3 band limit as ToTEM's is, so that the runs move the
 * band a real run moves) and
 * makes the fastest plan and batch the device's for n from then on; returns the
 * batch and its probe.slice/s (may be NULL); totem_stem_plan reports the plan.
 * Allocates and synchronises; no counterpart in the reference. */
int totem_stem_autotune(int n, int slices, int *best_batch, double *probe_slices_per_s);
/* Pass schedules ("plans") compiled for n-point grids: each fixes the radix plan of
 * the rows and columns passes, rows and columns per workgroup and the LDS exchange
 * form.  totem_stem_autotune measures every plan with the batch sizes and keeps the
 * fastest; totem_stem_set_plan chooses one explicitly.  A run (or a served grid
 * side) uses the plan current when it starts.  No counterpart in the reference,
 * whose FFT schedules are cuFFT's (cuda_source_lobato.cu:1324-1327). */
int totem_stem_plan_count(int n);  /* 0 if n has no plan */
int totem_stem_plan(int n, int *plan, char *desc, size_t cap);  /* plan in use on the current device */
int totem_stem_set_plan(int n, int plan);
/* Columns of the columns pass's workgroups under plan `plan` of n-point grids (0 if
 * there is no such plan). */
int totem_stem_plan_columns(int n, int plan);
/* Band of a run: totem_stem_run flags the columns of job->prop that hold a nonzero
 * entry and processes only the shortest wrapped run of `group`-column groups
 * [lo, lo + len) (mod n) covering them (group = totem_stem_plan_columns of the plan
 * in use; len = 0 if prop is zero, n if no group is empty), extended at its end to a
 * multiple of 8 groups (at most n) for the columns pass's XCD order.  The wave in every other
 * column is zero after each propagator multiply, so neither pass reads or writes
 * it; the detectors and the exit wave see exact zeros there, as with the full
 * transform.  This is that rule on the host, from per-column flags (nonzero:
 * the column holds a nonzero entry).  No counterpart in the reference, whose cuFFT
 * transforms transform every column (cuda_source_lobato.cu:1324-1327, 1388-1391). */
int totem_stem_band(const int *col_flags, int n, int group, int *lo, int *len);
/* Device bytes of workspace totem_stem_run needs for `batch` probes in flight. */
size_t totem_stem_workspace_bytes(const totem_stem_job *job, int batch);
/* Runs all M*D probes of the job, `batch` at a time (batch = 1 is the plain
 * one-probe-per-launch form).  workspace may be NULL (the library allocates).
 * Asynchronous on `stream` (NULL: the library's stream) unless it allocates. */
int totem_stem_run(const totem_stem_job *job, int batch, int flags, void *workspace, size_t ws_bytes,
                   void *stream);

/* totem_stem_run without any device allocation (no hipMalloc/hipFree, no allocating
 * library call).  No counterpart in the reference: for callers that cannot allocate
 * device memory (serving loops and harnesses that forbid hipMalloc).  The workspace
 * (waves, detector partials, work items) is the library's static device array of
 * TOTEM_STATIC_SCRATCH_BYTES per GPU, held from the library's first use on that GPU
 * until the process ends; a job whose totem_stem_workspace_bytes(job, batch) exceeds
 * it is refused with TOTEM_ERR_MEMORY.  Every call on one GPU shares that array and
 * it keeps the last call's contents: calls that overlap on one GPU (other host
 * threads) race.  Plain launches only (no graph, no timing).  Returns when the run
 * has finished on `stream` (NULL: the device's null stream). */
#define TOTEM_STATIC_SCRATCH_BYTES (128u << 20)
int totem_stem_run_static(const totem_stem_job *job, int batch, void *stream);

/* Accumulated durations (ms) and launch counts of the rows and columns passes
 * over TOTEM_RUN_TIMING runs on the calling thread; reset != 0 clears them after reading. */
void totem_kernel_times(int reset, double *rows_ms, long *rows_n, double *cols_ms, long *cols_n);

/* ----------------------------------------------- specimen on the device
 * Frozen-phonon displacement and slicing on the GPU (the host steps before
 * myProjectfun in the reference: randshiftpos.m:1-44,
 * CalAllEquilent20210223_lobato_peng.m:23-57,213-226 and
 * CalAllEquilent20210305_peng_corr.m:25-70).  `pot`
 * holds every atom's coefficient rows as one slice (slices = 1; grids, sigma,
 * parameter, correction tables as for totem_build_transmission); term_atom maps
 * each of its rows, in the order proj_coff_mat, proj_coff_mat_i, ele_n_corr,
 * ele_n_i_corr, to its atom. */
typedef struct totem_specimen_desc {
  totem_potential_desc pot;
  int n_atoms;
  const double *xyz;      /* [n_atoms][3] positions (Angstrom) */
  const double *rms;      /* [n_atoms] displacement rms per axis (Angstrom), or NULL (static specimen) */
  const int *term_atom;   /* [n_terms] */
  int n_terms;
  int slices;             /* S */
  double slice_origin;    /* st[0]; slice i is st[i] <= z < st[i] + slice_thickness, st[i] = st[0] + i dz */
  double slice_thickness;
  const int *mid;         /* [n_mid] 1-based slices whose upper face st[m-1] + dz atoms must not cross */
  int n_mid;
  double multiceng;       /* Peng: atoms within multiceng * dz of a face also feed the neighbouring
                             slices, erf-weighted per Gaussian term; corrections within 0.015 A of a
                             face are shared half and half when multiceng >= 1.  Lobato: ignored (0). */
} totem_specimen_desc;
typedef struct totem_specimen totem_specimen;
int totem_specimen_create(const totem_specimen_desc *desc, totem_specimen **out);
/* Displaces the atoms for (seed, config) -- Gaussian per axis, redrawn (up to 64 times, then
 * left in place) while an atom leaves [st[0], st[S-1] + dz) or crosses a mid-slice face --
 * slices them and builds t = exp(i sigma/1000 V) into t_dev (complex64 [S][np][np]). */
int totem_specimen_transmission(totem_specimen *spec, unsigned long long seed, int config, void *t_dev,
                                void *stream);
/* The displaced positions of (seed, config), host [n_atoms][3]. */
int totem_specimen_positions(totem_specimen *spec, unsigned long long seed, int config, double *xyz);
int totem_specimen_destroy(totem_specimen *spec);

/* ------------------------------------------------------------- served engine
 * Continuous batching of many device STEM jobs (no counterpart in the reference,
 * which runs one myProjectfun call at a time): jobs of any supported grid side are
 * submitted at any time, every step advances each in-flight probe by one slice,
 * probes of all jobs with the same grid side share launches, and a job's probes
 * are admitted into wave slots (first come, first served) as slots free up.  A
 * job's arrays must stay valid, and be ready on the device, from submit until
 * poll reports it; its det (and exit_wave) are then final.  Results equal those
 * of the same job run alone.  One host thread per server. */
typedef struct totem_server totem_server;
/* arena_bytes of HBM for in-flight waves (a page holds one 2048^2 slot or several
 * smaller ones); jobs may use up to max_aper detectors and max_mid mid slices. */
int totem_server_create(size_t arena_bytes, int max_aper, int max_mid, totem_server **out);
int totem_server_submit(totem_server *srv, const totem_stem_job *job, long long *id);
/* Admits waiting probes and advances every in-flight probe by one slice
 * (asynchronous); in_flight receives the probes in flight after admission. */
int totem_server_step(totem_server *srv, int *in_flight);
/* Ids of up to cap finished jobs (each reported once). */
int totem_server_poll(totem_server *srv, long long *ids, int cap, int *n_done);
int totem_server_sync(totem_server *srv);
int totem_server_destroy(totem_server *srv);

/* ------------------------------------------------------- batch HRTEM imaging
 * Replaces the computation of the reference's Python tool (cuda_stem64.py:164-704,
 * class stem: createNewPP2, WholeTCC2D_newTEM_forpar, the product with the exit
 * wave, ifftshift, batched inverse FFT, |.|^2 and the gfsf weighting).  The random
 * draws come from the caller (the reference draws them with curand, :667). */
typedef struct totem_hrtem_desc {
  int n;                 /* grid side N */
  int batches;           /* B images */
  int mtotal;            /* focal members (only member 0 reaches the output, as in the reference) */
  double lambda_nm;      /* para_part1['lambda'] */
  double alafa;          /* para_part1['alafa'], convergence semi-angle of the source (rad) */
  double delta_yita;     /* para_part1['delta_yita'] */
  double gfsf;           /* para_part1['gfsf'] (weight of the single accumulated member) */
  const double *u, *v;   /* [N][N] reciprocal grids U, V (1/nm) */
  const double *wave;    /* [N][N] complex128 as (re, im) pairs: centred reciprocal-space exit wave */
  const double *pp2;     /* [49] nominal parameters (para_part2) */
  const double *mrr;     /* [21] draw ranges (myrandrange) */
  const double *ai;      /* [B][21] uniform draws in [0, 1) */
} totem_hrtem_desc;
/* images: host [B][N][N] doubles; paras (may be NULL): host [B][49], the drawn
 * parameter rows (newPp2).  Synchronous on `stream` (NULL: the library's). */
int totem_hrtem_images(const totem_hrtem_desc *desc, double *paras, double *images, void *stream);

/* HRTEM image of an exit wave (ToTEM_submit.m:1327-1335):
 * image = sum_m gfsf[m] |IFFT2(ifftshift(wave * tcc[m] * ap))|^2 with wave the centred
 * reciprocal-space exit wave (complex128 [n][n] as (re, im) pairs), tcc the centred
 * transfer functions (complex128 [m][n][n]), ap the objective aperture [n][n]; all host
 * arrays, image host [n][n].  Synchronous. */
int totem_tcc_image(int n, int m, const double *wave, const double *tcc, const double *ap, const double *gfsf,
                    double *image, void *stream);
/* out = sum_m weights[m] |fields[m]|^2 over device complex64 fields [m][n][n], quadrant-swapped
 * (fftshift) when shift != 0; out host [n][n] (the CBED pattern of ToTEM_submit.m:1459-1472 from
 * a job's exit waves).  Synchronous. */
int totem_weighted_pattern(int n, int m, const void *fields, const double *weights, int shift, double *out,
                           void *stream);

/* Batched in-place 2-D FFT of `batch` n x n complex64 matrices on the device
 * (sign -1 forward, +1 inverse, unnormalised).  Test/diagnostic entry point. */
int totem_fft2d(void *data, int n, int batch, int sign, void *stream);

#ifdef __cplusplus
}
#endif

#endif /* TOTEM_ABI_H */
