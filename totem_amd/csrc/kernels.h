// Device-side data structures shared by the kernels and the host engine.
#pragma once
#include <hip/hip_runtime.h>
#include <cstddef>

namespace totem {

constexpr int kMaxAper = 8;  // detectors per job handled in one pass

// One simulation (one potential, one optics set-up) as the kernels see it.
struct JobDev {
  const float2* t;            // transmission stack [S][np][np], complex64
  const float2* prop;         // Fresnel propagator [N][N], zero frequency at [0][0]
  const double* aper;         // centred detector masks [A][N][N] (final detector)
  const double* aper2;        // ifftshifted masks [A][N][N] (mid-slice detector) or null
  const int* mid_of_slice;    // [S]: mid-layer index written at this slice, -1 if none (or null)
  const int* layer_hit;       // [L]: 1 if some slice writes mid-layer l
  double* det;                // detector output [(L+1)][A][n_out]; layer L is the exit face
  int np;                     // potential side
  int S;                      // slices
  int A;                      // detectors
  int L;                      // mid layers
  float2* exit;               // exit-face wave Psi(k) [n_out][N][N] (zero frequency at [0][0]) or null
  int n_out;                  // output columns (probe positions x focal members)
  int wrap;                   // periodic window into the potential
  int prop_cols;              // C > 0: prop stored column group by column group, element (r, g*C + c)
                              // at (g*N + r)*C + c (the columns pass's read order); 0: row-major
  // Columns c of the reciprocal grid with ((c - band_lo) mod N) < band_len: every
  // column outside holds zeros only in prop, so the wave there is zero after every
  // propagator multiply and neither pass reads or writes it (slice 0's rows pass
  // still reads the whole probe; the columns pass is launched for the band's groups
  // only).  band_lo and band_len are multiples of the columns pass's group width.
  int band_lo = 0;
  int band_len = 1 << 30;
  // the same band in column groups of the columns pass: groups (g - band_g0) mod G <
  // band_ng; the columns pass runs and the detector reduction sums only these
  int band_g0 = 0;
  int band_ng = 1 << 30;
};

// One probe position of one job (a "work item") and the device memory it owns
// while in flight: a wave buffer and its detector partials.
struct ItemDev {
  const float2* probe;        // initial wave [N][N], complex64
  float2* wave;               // wave buffer [N][N]
  double* part;               // detector partials [lmax1][amax][G] (G = column groups of the N grid)
  int job;                    // index into the job table
  int r0, c0;                 // window origin (row, column) in the potential
  int start_step;             // launch step at which the item's slice 0 runs
  int out;                    // output column in job.det / job.exit
};

// Form-factor class of the potential builder.
enum ClassKind : int {
  kPeng = 0,        // sum a exp(-b s^2)
  kLobato = 1,      // sum a (2 + b g^2)/(1 + b g^2)^2 * exp(-B s^2), g^2 = 4 s^2
  kPengIon = 2,     // w1 * Peng(ion) + w2 * Peng(atom)
  kLobatoIon = 3,   // (w1 * Lobato(ion) + w2 * Lobato(atom)) * exp(-B s^2)
  kCorr = 4,        // corr_table[i1] * exp(-B s^2)
  kCorrIon = 5,     // (w1 * corr_table[i1] + w2 * corr_table[i2]) * exp(-B s^2)
};

struct PotClass {
  int kind;
  int has_abs;      // adds i * sum C exp(-D s^2)
  int i1, i2;       // correction-table indices (0-based)
  double w1, w2, B;
  double p1[10];    // a1 b1 ... a5 b5
  double p2[10];    // second coefficient set (ion forms)
  double ab[10];    // absorptive C1 D1 ... C5 D5
};

// A run of terms of one class inside one slice.
struct PotSeg {
  int cls;          // class index
  int begin, end;   // term range relative to the slice's first term
};

}  // namespace totem
