// Host-side launch interface of the kernels in kernels.hip.
#pragma once
#include <hip/hip_runtime.h>
#include "kernels.h"

#include <cstdint>
#include <string>
#include <vector>

namespace totem {

// Device pointer (current device) of the n-point plan's twiddle table, created on
// first use from twiddle_host() and kept for the process.
const float2* tw_table(int n);
// Host copy of the n-point plan's stage twiddle tables (see fft_block.h), FP64
// values rounded to unit-modulus FP32 pairs.
void twiddle_host(int n, std::vector<float2>& out);

bool size_supported(int n);   // FFT lengths with a compiled radix plan
// Pass variants ("plans") of n-point grids: schedules of the rows and columns passes
// (radix plans, rows / columns per workgroup, exchange form).  Both passes of a run
// must use the same variant (it fixes the in-flight wave layout).
int pass_variant_count(int n);
std::string pass_variant_desc(int n, int variant);
int detector_groups(int n, int variant);  // column groups per probe in k_pass_cols
int detector_groups_max(int n);           // over the variants (sizes detector partials)

// One slice step of every listed item (slice = step - item.start_step; items outside
// [0, S) of their job are skipped), with pass variant `variant` of n.
// Columns per columns-pass workgroup of plan v of grid side n.
int pass_cols_width(int n, int v);
// prop [n][n] row-major -> the column-group order JobDev::prop_cols describes.
void prop_column_groups(hipStream_t st, const float2* prop, float2* out, int n, int C);
// flags[c] = 1 if column c of prop [n][n] (row-major) holds a nonzero entry, else 0.
void column_flags(hipStream_t st, const float2* prop, int* flags, int n);
// The shortest wrapped run of C-column groups covering every flagged column (len = 0
// without a flag, the whole grid without an empty group).
void column_band(const int* flags, int n, int C, int& lo, int& len);
// The band a run processes: column_band widened at its end to a multiple of 8
// groups (at most the grid), in columns and in groups, into the job entry.
void set_job_band(JobDev& jd, const int* flags, int n, int C);
// The job's band of prop (column flags on `st`, which is synchronised) into jd;
// flags: n ints of device scratch.
void job_band(hipStream_t st, JobDev& jd, const float2* prop, int* flags, int n, int C);
void pass_rows(int n, int variant, hipStream_t st, const ItemDev* items, int n_items, const JobDev* jobs, int step);
// The columns pass runs `groups` column groups per item, from the first group of each
// job's band on (all n / pass_cols_width(n, variant) of them for the default band).
void pass_cols(int n, int variant, hipStream_t st, const ItemDev* items, int n_items, const JobDev* jobs, int step,
               int groups);
// Detector sums of the items items[list[e]] (list may be null: items[e]), e < n, into their
// job's det; lmax1 and amax bound (L + 1) and A over the jobs involved.
void reduce_det(hipStream_t st, const ItemDev* items, const int* list, int n, const JobDev* jobs, int groups,
                int lmax1, int amax);
// In-place batched 2-D FFT of `batch` n x n complex64 matrices; the column pass
// writes to dst (may equal data).  epi: 0 plain, 1 transmission exp(i sigma V), 2 scale 1/n^2.
void fft2d(int n, hipStream_t st, float2* data, float2* dst, int batch, int sign, int epi, float sigma);
// p[0..n) = value (small fill kernels; graph-capture safe, unlike memset nodes)
void fill(hipStream_t st, float2* p, long n, float re, float im);
void fill(hipStream_t st, double* p, long n, double v);
// True when the grid's frequencies pair up (ky1[np - r] = -ky1[r], kx1[np - c] = -kx1[c]
// for 1 <= r, c <= np/2 - 16) and np is a multiple of 32: the accumulation then
// forms a quarter of the tiles and writes each one's mirrored pixels too.
bool hermitian_pairs(const double* kx1, const double* ky1, int np);
// Entries per phase table the accumulation reads (np/2 + 16 with paired frequencies).
int phase_columns(int np, bool sym);
void phase_tables(hipStream_t st, const double* tx, const double* ty, const double* wt, int n_terms,
                  const double* kx1, const double* ky1, int np, bool sym, float2* ex, float2* ey);
void accumulate(hipStream_t st, const float2* ex, const float2* ey, const PotSeg* segs, int n_segs,
                const PotClass* classes, const double* s2, const double* aper_true, const double* corr,
                double parameter, int np, bool sym, float2* vshift);

// Specimen slicing on the device (specimen.hip).
void displace(hipStream_t st, int n, const double* xyz, const double* rms, uint64_t seed, int config, double lo,
              double hi, const double* faces, int n_faces, double* out);
// Clears counts[S * Ct], then writes keys / vals / wslot[n * K] (K slots per
// term) and counts the slots per (slice, class) key.
struct TermKeyArgs {
  int n = 0, K = 2, S = 0, Ct = 0, split = 0;
  const int *term_atom = nullptr, *term_cls = nullptr, *term_mode = nullptr, *sub_base = nullptr;
  const double *term_w = nullptr, *xyz = nullptr;
  const PotClass* classes = nullptr;
  double origin = 0, dz = 0, m = 0, discon = 0;
  unsigned *keys = nullptr, *vals = nullptr;
  double* wslot = nullptr;
  int* counts = nullptr;
};
void term_keys(hipStream_t st, const TermKeyArgs& a);
// Stable LSD radix sort of n (key, value) pairs on key bits [0, bits), 8 bits per
// pass, ping-ponging between (keys, vals) and (keys2, vals2); returns the values
// buffer that holds the sorted order.  temp: sort_pairs_temp_bytes(n) bytes.
size_t sort_pairs_temp_bytes(int n);
unsigned* sort_pairs(hipStream_t st, void* temp, unsigned* keys, unsigned* keys2, unsigned* vals, unsigned* vals2,
                     int n, int bits);
void term_gather(hipStream_t st, int n, int K, const unsigned* vals, const int* term_atom, const double* wslot,
                 const double* xyz, double* tx, double* ty, double* wt);

}  // namespace totem
