// FP64 C++ restatement of the reference's multicore CPU STEM path -- ORACLE, TEST
// INFRASTRUCTURE ONLY (the CPU baseline of bench.py and a second check of the
// numpy oracle; never part of the product path).
//
// Follows the same reference code as oracle/potential.py and oracle/multislice.py:
//   * potential: GetPotential4AllSlice_multicore_lotabo_peng_corr.m:15-65 (correction
//     terms), GetPotential4AllSlice_multicore_lotabo_peng.m:25-181 (Peng / Lobato
//     atoms, absorptive parts, ion rows), :187-190 (V = ifft2(ifftshift(Vk .* APERTURE))
//     * Np^2, t = exp(i V sigma / 1000)); slices in parallel like the reference's
//     parfor over slices (:28);
//   * propagation: cuda_STEM_CPU_Mcore_lobato_peng.m:50-97 (window multiply, fft2,
//     propagator, mid-slice detectors, ifft2, final fft2 and detectors, gfsf
//     weighting); probe positions in parallel like its parfor over scan columns (:55).
//
// Arithmetic differences from the MATLAB code that do not change results beyond FP64
// rounding: the phase exp(-2 pi i (gx x + gy y)) is the product of per-atom row and
// column factors (gx varies along columns, gy along rows, as the meshgrid gives); atoms
// with identical coefficient rows share one form-factor evaluation per pixel; the FFT
// is this file's own mixed-radix (2, 3, 4, 5) Stockham transform instead of FFTW.
//
// C ABI (ctypes): oracle_fft2, oracle_potential, oracle_stem.  Parity with the numpy
// oracle is checked in tests/test_oracle_cpp.py.
#include <algorithm>
#include <cmath>
#include <complex>
#include <cstring>
#include <map>
#include <vector>

#include <omp.h>

namespace {

using cd = std::complex<double>;

// ---------------------------------------------------------------- FFT
// Stockham autosort, decimation in frequency: a stage of radix R on sub-transforms of
// length n with s interleaved sequences reads x[q + s (p + r m)] (m = n / R), forms the
// R-point DFT b_k, and writes y[q + s (R p + k)] = b_k w^(p k), w = exp(sign 2 pi i / n);
// the next stage runs on length m with stride s R.  With s0 > 1 the same code transforms
// s0 interleaved sequences (the columns of a row-major matrix).
struct Stage {
  int R, n, m;
  std::vector<cd> tw;  // [p][k], k = 1..R-1
};

struct Plan {
  int n = 0;
  std::vector<Stage> st;
  cd w3[3], w5[5];
};

static bool factor(int n, std::vector<int>& rad) {
  rad.clear();
  while (n % 4 == 0) { rad.push_back(4); n /= 4; }
  while (n % 2 == 0) { rad.push_back(2); n /= 2; }
  while (n % 3 == 0) { rad.push_back(3); n /= 3; }
  while (n % 5 == 0) { rad.push_back(5); n /= 5; }
  return n == 1;
}

static Plan make_plan(int n, int sign) {
  Plan P;
  P.n = n;
  std::vector<int> rad;
  factor(n, rad);
  int len = n;
  for (int R : rad) {
    Stage s;
    s.R = R;
    s.n = len;
    s.m = len / R;
    s.tw.resize((size_t)s.m * (R - 1));
    for (int p = 0; p < s.m; ++p)
      for (int k = 1; k < R; ++k) {
        const double a = sign * 2.0 * M_PI * (double)((long)p * k % len) / (double)len;
        s.tw[(size_t)p * (R - 1) + k - 1] = cd(std::cos(a), std::sin(a));
      }
    P.st.push_back(std::move(s));
    len /= R;
  }
  for (int k = 0; k < 3; ++k) P.w3[k] = cd(std::cos(sign * 2.0 * M_PI * k / 3), std::sin(sign * 2.0 * M_PI * k / 3));
  for (int k = 0; k < 5; ++k) P.w5[k] = cd(std::cos(sign * 2.0 * M_PI * k / 5), std::sin(sign * 2.0 * M_PI * k / 5));
  return P;
}

// One transform set: s0 interleaved sequences of length P.n in x (work buffer y of the
// same size); the result ends in x.
static void fft_run(const Plan& P, int sign, cd* x, cd* y, int s0) {
  cd* a = x;
  cd* b = y;
  int s = s0;
  const cd I(0.0, (double)sign);
  for (const Stage& sg : P.st) {
    const int R = sg.R, m = sg.m;
    for (int p = 0; p < m; ++p) {
      const cd* w = &sg.tw[(size_t)p * (R - 1)];
      for (int q = 0; q < s; ++q) {
        const cd* in = a + q + (size_t)s * p;
        cd* out = b + q + (size_t)s * R * p;
        const size_t sm = (size_t)s * m;
        if (R == 4) {
          const cd a0 = in[0], a1 = in[sm], a2 = in[2 * sm], a3 = in[3 * sm];
          const cd t0 = a0 + a2, t1 = a0 - a2, t2 = a1 + a3, t3 = I * (a1 - a3);
          out[0] = t0 + t2;
          out[s] = (t1 + t3) * w[0];
          out[2 * s] = (t0 - t2) * w[1];
          out[3 * s] = (t1 - t3) * w[2];
        } else if (R == 2) {
          const cd a0 = in[0], a1 = in[sm];
          out[0] = a0 + a1;
          out[s] = (a0 - a1) * w[0];
        } else if (R == 3) {
          const cd a0 = in[0], a1 = in[sm], a2 = in[2 * sm];
          out[0] = a0 + a1 + a2;
          out[s] = (a0 + a1 * P.w3[1] + a2 * P.w3[2]) * w[0];
          out[2 * s] = (a0 + a1 * P.w3[2] + a2 * P.w3[1]) * w[1];
        } else {  // 5
          cd v[5];
          for (int r = 0; r < 5; ++r) v[r] = in[r * sm];
          for (int k = 0; k < 5; ++k) {
            cd acc = 0.0;
            for (int r = 0; r < 5; ++r) acc += v[r] * P.w5[(r * k) % 5];
            out[(size_t)k * s] = k == 0 ? acc : acc * w[k - 1];
          }
        }
      }
    }
    std::swap(a, b);
    s *= R;
  }
  if (a != x) std::memcpy(x, a, sizeof(cd) * (size_t)P.n * s0);
}

// Plans per (n, sign), built once per thread.
struct Plans {
  std::map<std::pair<int, int>, Plan> m;
  const Plan& get(int n, int sign) {
    auto k = std::make_pair(n, sign);
    auto it = m.find(k);
    if (it == m.end()) it = m.emplace(k, make_plan(n, sign)).first;
    return it->second;
  }
};

// Unnormalised 2-D transform of a row-major n x n matrix (sign -1 forward, +1 inverse).
// Rows, then columns, are transformed kBlock at a time, interleaved in a block buffer
// that stays in cache through all stages, so that every stage's inner loop runs over
// the block (vectorised) and the matrix is read and written once per direction.
constexpr int kBlock = 16;
static void fft2(Plans& pl, int n, int sign, cd* x, std::vector<cd>& work) {
  const Plan& P = pl.get(n, sign);
  work.resize(2 * (size_t)n * kBlock);
  cd* blk = work.data();
  cd* tmp = work.data() + (size_t)n * kBlock;
  for (int r0 = 0; r0 < n; r0 += kBlock) {
    const int b = std::min(kBlock, n - r0);
    for (int i = 0; i < b; ++i)
      for (int j = 0; j < n; ++j) blk[(size_t)j * b + i] = x[(size_t)(r0 + i) * n + j];
    fft_run(P, sign, blk, tmp, b);
    for (int i = 0; i < b; ++i)
      for (int j = 0; j < n; ++j) x[(size_t)(r0 + i) * n + j] = blk[(size_t)j * b + i];
  }
  for (int c0 = 0; c0 < n; c0 += kBlock) {
    const int b = std::min(kBlock, n - c0);
    for (int j = 0; j < n; ++j)
      for (int i = 0; i < b; ++i) blk[(size_t)j * b + i] = x[(size_t)j * n + c0 + i];
    fft_run(P, sign, blk, tmp, b);
    for (int j = 0; j < n; ++j)
      for (int i = 0; i < b; ++i) x[(size_t)j * n + c0 + i] = blk[(size_t)j * b + i];
  }
}

// ---------------------------------------------------------------- potential

double peng_sum(const double* c, double s2) {
  double f = 0.0;
  for (int q = 0; q < 5; ++q) f += c[2 * q] * std::exp(-s2 * c[2 * q + 1]);
  return f;
}
double lobato_sum(const double* c, double g2) {
  double f = 0.0;
  for (int q = 0; q < 5; ++q) {
    const double d = 1.0 + c[2 * q + 1] * g2;
    f += c[2 * q] * (2.0 + c[2 * q + 1] * g2) / (d * d);
  }
  return f;
}

}  // namespace

extern "C" {

// Inputs of the potential build, in the layouts of oracle/potential.py (row-major
// [rows][cols] doubles; gx/gy/s2/aperture [np][np] centred).
typedef struct oracle_potential_desc {
  int paraflag;  // 'l', 'p' or 'n'
  int np, slices;
  const double *series_n, *ele_n, *absorp_n;          // [S], [rows][13|14], [rows][10] or NULL
  const double *series_n_i, *ele_n_i, *absorp_n_i;    // [S] or NULL, [rows][25|26], [rows][10] or NULL
  const double *series_n_corr, *ele_n_corr;           // [S], [rows][5]
  const double *series_n_i_corr, *ele_n_i_corr;       // [S], [rows][8]
  const double *corr_matrix;                          // [cim][np][np]
  int cim;
  const double *gx, *gy, *s2, *aperture;
  double sigma, parameter;
} oracle_potential_desc;

int oracle_fft2(double* data, int n, int sign) {
  std::vector<int> rad;
  if (n <= 0 || !factor(n, rad) || (sign != 1 && sign != -1)) return -1;
  Plans pl;
  std::vector<cd> work;
  fft2(pl, n, sign, reinterpret_cast<cd*>(data), work);
  return 0;
}

// Builds slices [s0, s1): t (complex [S][np][np] as (re, im) pairs) and, if v is not
// NULL, the real-space potential V (complex, same layout).  `threads` OpenMP threads
// over slices.  Returns 0, or -1 for a grid without a (2, 3, 5)-smooth size.
int oracle_potential(const oracle_potential_desc* d, int s0, int s1, double* t, double* v, int threads) {
  const int n = d->np;
  std::vector<int> rad;
  if (n <= 0 || !factor(n, rad) || s0 < 0 || s1 > d->slices || s0 > s1) return -1;
  const size_t n2 = (size_t)n * n;
  const bool lob = d->paraflag == 'l';
  const int ncol = lob ? 14 : 13, ncol_i = lob ? 26 : 25;
  auto offsets = [&](const double* series) {
    std::vector<long> off(d->slices + 1, 0);
    for (int s = 0; s < d->slices; ++s) off[s + 1] = off[s] + (series ? (long)series[s] : 0L);
    return off;
  };
  const std::vector<long> o1 = offsets(d->series_n), o2 = offsets(d->series_n_i);
  const bool corr = d->paraflag == 'n' && d->cim > 0 && d->corr_matrix;
  const std::vector<long> oc1 = offsets(corr ? d->series_n_corr : nullptr);
  const std::vector<long> oc2 = offsets(corr ? d->series_n_i_corr : nullptr);
  std::vector<double> kx(n), ky(n);
  for (int c = 0; c < n; ++c) kx[c] = d->gx[c];
  for (int r = 0; r < n; ++r) ky[r] = d->gy[(size_t)r * n];

#pragma omp parallel num_threads(threads)
  {
    Plans pl;
    std::vector<cd> Vk(n2), acc(n2), F(n2), work;
    std::vector<double> ar(n2), ai(n2), exr, exi, eyr, eyi;
#pragma omp for schedule(dynamic, 1)
    for (int s = s0; s < s1; ++s) {
      std::fill(Vk.begin(), Vk.end(), cd(0.0));
      // one "term": weight w at (x, y) with form factor F (per pixel) shared by a group
      // of consecutive terms with identical coefficient rows
      struct Term {
        double x, y, w;
        const double* key;  // coefficient row (compared for grouping)
        int klen, kind;     // kind: 0 atom, 1 ion, 2 corr, 3 corr ion
        long row;
      };
      std::vector<Term> terms;
      if (corr) {
        for (long a = oc1[s]; a < oc1[s + 1]; ++a) {
          const double* r = d->ele_n_corr + a * 5;
          terms.push_back({r[0], r[1], r[3], r + 2, 3, 2, a});  // key: B, occ*w (unused), class
        }
        for (long a = oc2[s]; a < oc2[s + 1]; ++a) {
          const double* r = d->ele_n_i_corr + a * 8;
          terms.push_back({r[0], r[1], r[3], r + 2, 6, 3, a});
        }
      }
      for (long a = o1[s]; a < o1[s + 1]; ++a) {
        const double* r = d->ele_n + a * ncol;
        terms.push_back({r[0], r[1], r[2], r + 3, ncol - 3, 0, a});
      }
      for (long a = o2[s]; a < o2[s + 1]; ++a) {
        const double* r = d->ele_n_i + a * ncol_i;
        terms.push_back({r[0], r[1], r[2], r + 3, ncol_i - 3, 1, a});
      }
      // the values that fix a term's form factor, compared lexicographically: terms
      // with equal values share one evaluation per pixel
      auto ff_values = [&](const Term& p, std::vector<double>& v) {
        v.assign(1, (double)p.kind);
        if (p.kind == 2) {  // B and class; the weight is the term's own
          v.push_back(p.key[0]);
          v.push_back(p.key[2]);
        } else if (p.kind == 3) {  // B, w1, class1, w2, class2 (occ is the weight)
          v.push_back(p.key[0]);
          v.insert(v.end(), p.key + 2, p.key + 6);
        } else {
          v.insert(v.end(), p.key, p.key + p.klen);
          const double* ab = p.kind == 0 ? d->absorp_n : d->absorp_n_i;
          if (ab) v.insert(v.end(), ab + p.row * 10, ab + p.row * 10 + 10);
        }
      };
      {
        std::vector<std::pair<std::vector<double>, size_t>> keyed(terms.size());
        for (size_t i = 0; i < terms.size(); ++i) {
          ff_values(terms[i], keyed[i].first);
          keyed[i].second = i;
        }
        std::stable_sort(keyed.begin(), keyed.end(),
                         [](const auto& a, const auto& b) { return a.first < b.first; });
        std::vector<Term> sorted;
        sorted.reserve(terms.size());
        for (auto& k : keyed) sorted.push_back(terms[k.second]);
        terms.swap(sorted);
      }
      std::vector<double> kv0, kv1;
      size_t i0 = 0;
      while (i0 < terms.size()) {
        size_t i1 = i0 + 1;
        ff_values(terms[i0], kv0);
        while (i1 < terms.size()) {
          ff_values(terms[i1], kv1);
          if (kv1 != kv0) break;
          ++i1;
        }
        // form factor of the group per pixel
        const Term& k = terms[i0];
        for (size_t p = 0; p < n2; ++p) {
          const double s2v = d->s2[p], g2 = 4.0 * s2v;
          double re = 0.0, im = 0.0;
          if (k.kind == 2) {
            re = d->corr_matrix[((size_t)k.key[2] - 1) * n2 + p] * std::exp(-s2v * k.key[0]);
          } else if (k.kind == 3) {
            re = (k.key[2] * d->corr_matrix[((size_t)k.key[3] - 1) * n2 + p] +
                  k.key[4] * d->corr_matrix[((size_t)k.key[5] - 1) * n2 + p]) *
                 std::exp(-s2v * k.key[0]);
          } else if (k.kind == 0) {
            re = lob ? lobato_sum(k.key, g2) * std::exp(-s2v * k.key[10]) : peng_sum(k.key, s2v);
            if (d->absorp_n) im = peng_sum(d->absorp_n + k.row * 10, s2v);
          } else {
            re = lob ? (k.key[0] * lobato_sum(k.key + 1, g2) + k.key[11] * lobato_sum(k.key + 12, g2)) *
                           std::exp(-s2v * k.key[22])
                     : k.key[0] * peng_sum(k.key + 1, s2v) + k.key[11] * peng_sum(k.key + 12, s2v);
            if (d->absorp_n_i) im = peng_sum(d->absorp_n_i + k.row * 10, s2v);
          }
          F[p] = cd(re, im);
        }
        // structure factor of the group: sum_j w_j ey_j[r] ex_j[c], the phase tables of
        // every term first, then tiles of kTileRows rows that stay in cache over all terms
        const size_t g = i1 - i0;
        exr.resize(g * n);
        exi.resize(g * n);
        eyr.resize(g * n);
        eyi.resize(g * n);
        for (size_t i = 0; i < g; ++i) {
          const Term& t = terms[i0 + i];
          for (int c = 0; c < n; ++c) {
            const double a = -2.0 * M_PI * kx[c] * t.x;
            exr[i * n + c] = std::cos(a);
            exi[i * n + c] = std::sin(a);
          }
          for (int r = 0; r < n; ++r) {
            const double a = -2.0 * M_PI * ky[r] * t.y;
            eyr[i * n + r] = t.w * std::cos(a);
            eyi[i * n + r] = t.w * std::sin(a);
          }
        }
        constexpr int kTileRows = 8;
        for (int rt = 0; rt < n; rt += kTileRows) {
          const int re = std::min(n, rt + kTileRows);
          std::fill(ar.begin() + (size_t)rt * n, ar.begin() + (size_t)re * n, 0.0);
          std::fill(ai.begin() + (size_t)rt * n, ai.begin() + (size_t)re * n, 0.0);
          for (size_t i = 0; i < g; ++i) {
            const double* xr = &exr[i * n];
            const double* xi = &exi[i * n];
            for (int r = rt; r < re; ++r) {
              const double er = eyr[i * n + r], ei = eyi[i * n + r];
              double* pr = &ar[(size_t)r * n];
              double* pi = &ai[(size_t)r * n];
              for (int c = 0; c < n; ++c) {
                pr[c] += er * xr[c] - ei * xi[c];
                pi[c] += er * xi[c] + ei * xr[c];
              }
            }
          }
        }
        for (size_t p = 0; p < n2; ++p) Vk[p] += d->parameter * F[p] * cd(ar[p], ai[p]);
        i0 = i1;
      }
      // V = ifft2(ifftshift(Vk .* APERTURE)) * Np^2: the unnormalised inverse transform
      const int h = n / 2;
      for (int r = 0; r < n; ++r)
        for (int c = 0; c < n; ++c) {
          const size_t src = (size_t)r * n + c;
          acc[(size_t)((r + h) % n) * n + (c + h) % n] = Vk[src] * d->aperture[src];
        }
      fft2(pl, n, +1, acc.data(), work);
      double* ts = t + 2 * (size_t)s * n2;
      for (size_t p = 0; p < n2; ++p) {
        const cd tv = std::exp(cd(0.0, 1.0) * acc[p] * (d->sigma / 1000.0));
        ts[2 * p] = tv.real();
        ts[2 * p + 1] = tv.imag();
        if (v) {
          v[2 * ((size_t)s * n2 + p)] = acc[p].real();
          v[2 * ((size_t)s * n2 + p) + 1] = acc[p].imag();
        }
      }
    }
  }
  return 0;
}

// Multislice STEM of D probe positions (cuda_STEM_CPU_Mcore_lobato_peng.m:50-97).
// t: complex [S][np][np]; probe: complex [M][n][n]; psf: complex [n][n] (ifftshifted);
// aper_shifted: [A][n][n]; r0/c0: 0-based window origins per position; periodic != 0
// takes windows modulo np (else they must lie inside the potential); mid: 1-based
// slices [L].  out: [D][A] final detectors, mid_out: [D][L][A] (may be NULL if L = 0).
// `threads` OpenMP threads over positions.
int oracle_stem(const double* t, int np, int S, const double* probe, int M, const double* gfsf, const double* psf,
                const double* aper_shifted, int A, int n, const int* r0, const int* c0, int D, int periodic,
                const int* mid, int L, double* out, double* mid_out, int threads) {
  std::vector<int> rad;
  if (n <= 0 || !factor(n, rad) || np < n || S <= 0) return -1;
  for (int d = 0; d < D; ++d)
    if (!periodic && (r0[d] < 0 || c0[d] < 0 || r0[d] + n > np || c0[d] + n > np)) return -2;
  const size_t n2 = (size_t)n * n, np2 = (size_t)np * np;
  const cd* T = reinterpret_cast<const cd*>(t);
  const cd* PR = reinterpret_cast<const cd*>(probe);
  const cd* H = reinterpret_cast<const cd*>(psf);
  std::vector<int> mid_of(S + 1, -1);
  for (int l = 0; l < L; ++l)  // a slice listed twice feeds its first entry, as list.index does
    if (mid[l] >= 1 && mid[l] <= S && mid_of[mid[l]] < 0) mid_of[mid[l]] = l;
  const double inv = 1.0 / (double)n2;
#pragma omp parallel num_threads(threads)
  {
    Plans pl;
    std::vector<cd> w(n2), work;
    std::vector<int> ri(n), ci(n);
#pragma omp for schedule(dynamic, 1)
    for (int d = 0; d < D; ++d) {
      for (int i = 0; i < n; ++i) {
        ri[i] = periodic ? (r0[d] + i) % np : r0[d] + i;
        ci[i] = periodic ? (c0[d] + i) % np : c0[d] + i;
      }
      std::vector<double> res(A, 0.0), mres((size_t)std::max(L, 1) * A, 0.0);
      for (int m = 0; m < M; ++m) {
        std::copy(PR + (size_t)m * n2, PR + (size_t)(m + 1) * n2, w.begin());
        for (int k = 1; k <= S; ++k) {
          const cd* tk = T + (size_t)(k - 1) * np2;
          for (int r = 0; r < n; ++r) {
            const cd* tr = tk + (size_t)ri[r] * np;
            cd* wr = w.data() + (size_t)r * n;
            for (int c = 0; c < n; ++c) wr[c] *= tr[ci[c]];
          }
          fft2(pl, n, -1, w.data(), work);
          for (size_t p = 0; p < n2; ++p) w[p] *= H[p];
          if (mid_of[k] >= 0)
            for (int a = 0; a < A; ++a) {
              double s = 0.0;
              const double* ap = aper_shifted + (size_t)a * n2;
              for (size_t p = 0; p < n2; ++p) s += std::norm(w[p] * ap[p]);
              mres[(size_t)mid_of[k] * A + a] += gfsf[m] * s;
            }
          fft2(pl, n, +1, w.data(), work);
          for (size_t p = 0; p < n2; ++p) w[p] *= inv;
        }
        fft2(pl, n, -1, w.data(), work);
        for (int a = 0; a < A; ++a) {
          double s = 0.0;
          const double* ap = aper_shifted + (size_t)a * n2;
          for (size_t p = 0; p < n2; ++p) s += std::norm(w[p] * ap[p]);
          res[a] += gfsf[m] * s;
        }
      }
      for (int a = 0; a < A; ++a) out[(size_t)d * A + a] = res[a];
      if (mid_out)
        for (int l = 0; l < L; ++l)
          for (int a = 0; a < A; ++a) mid_out[((size_t)d * L + l) * A + a] = mres[(size_t)l * A + a];
    }
  }
  return 0;
}

int oracle_max_threads(void) { return omp_get_max_threads(); }

}  // extern "C"
