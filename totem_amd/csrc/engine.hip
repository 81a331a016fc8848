// Host orchestration and C ABI of the multislice STEM engine.
//
// myProjectfun semantics (reference cuda_source_lobato.cu:1018-1562 and
// cuda_source_peng_and_corr.cu:1217-1786):
//   1. transmission stack t[s] = exp(i sigma/1000 V_s) from the per-slice atom rows
//   2. for every focal member m and scan position d: multislice through all
//      slices, annular detectors at the exit face and at the listed mid slices
//   3. return_result[a][d] = sum_m gfsf[m] I_m[a][d], mid_ceng_mat[l][a][d] likewise
//
// Design differences from the reference (results agree within FP32 rounding):
//   * all probes of a batch advance one slice per pair of launches (rows pass,
//     columns pass) instead of ~6 synchronised launches + 2 FFT plan builds;
//   * the exit-face detector reads Psi*H inside the last columns pass
//     (FFT(IFFT(Psi*H)) == Psi*H), so the last inverse and forward FFTs are skipped;
//   * detector sums are per-workgroup FP64 partials reduced in a fixed order
//     (deterministic), not a log2(N^2)-launch tree;
//   * the potential is accumulated per pixel tile over atoms grouped by
//     form-factor class (no cross-atom read-modify-write race).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/totem_abi.h"
#include "host_common.h"
#include "kernels.h"
#include "launch.h"
#include "unit_round.h"

namespace totem {

std::string& last_error() {
  static thread_local std::string msg;
  return msg;
}

// Grow-only device buffer: keeps its memory across calls, so steady-state calls
// neither allocate nor free (hipFree would wait for all work on the device).
static void reserve(DevBuf& b, size_t bytes) {
  if (b.bytes < bytes) b.alloc(std::max<size_t>(bytes, 256));
}

template <class T> static void upload(DevBuf& b, const std::vector<T>& v, hipStream_t st) {
  reserve(b, std::max<size_t>(v.size(), 1) * sizeof(T));
  if (!v.empty()) HIP_OK(hipMemcpyAsync(b.p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, st));
}

// Process-wide resident context: one stream per device.
struct Context {
  std::mutex mu;
  std::map<int, hipStream_t> streams;
};
static Context& ctx() {
  static Context* c = new Context();  // never destroyed: streams live until process exit
  return *c;
}

hipStream_t lib_stream() {
  int dev = 0;
  HIP_OK(hipGetDevice(&dev));
  Context& c = ctx();
  std::lock_guard<std::mutex> lk(c.mu);
  auto it = c.streams.find(dev);
  if (it != c.streams.end()) return it->second;
  hipStream_t s;
  HIP_OK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
  c.streams[dev] = s;
  return s;
}

// The reference picks the GPU with the most free memory and makes it current
// (cuda_source_lobato.cu:1063-1081), which creates a context on every visible
// GPU.  Here the caller's current device is used unless TOTEM_DEVICE names one
// (an index) or asks for the reference's choice ("max_free").
static void select_device() {
  int count = 0;
  HIP_OK(hipGetDeviceCount(&count));
  require(count > 0, "no HIP device", TOTEM_ERR_DEVICE);
  const char* env = std::getenv("TOTEM_DEVICE");
  if (!env || !*env) return;
  if (std::strcmp(env, "max_free") != 0) {
    char* end = nullptr;
    const long d = std::strtol(env, &end, 10);
    require(end && *end == 0 && d >= 0 && d < count,
            std::string("TOTEM_DEVICE must be a device index below ") + std::to_string(count) + " or max_free");
    HIP_OK(hipSetDevice((int)d));
    return;
  }
  int best = 0;
  size_t best_free = 0;
  for (int i = 0; i < count; ++i) {
    HIP_OK(hipSetDevice(i));
    size_t fr = 0, tot = 0;
    HIP_OK(hipMemGetInfo(&fr, &tot));
    if (fr > best_free) {
      best_free = fr;
      best = i;
    }
  }
  HIP_OK(hipSetDevice(best));
}

// -------------------------------------------------------------- potential build

// A form-factor class: kind, absorption flag, correction-table indices, weights,
// B and the three coefficient sets, as 37 doubles compared lexicographically.
using ClassKeyV = std::array<double, 37>;
struct ClassKey {
  ClassKeyV v;
  bool operator<(const ClassKey& o) const { return v < o.v; }
};

struct Term {
  double x, y, w;
  int cls;
};

static PotClass make_class(const ClassKeyV& key) {
  PotClass k;
  std::memset(&k, 0, sizeof(k));
  k.kind = (int)key[0];
  k.has_abs = (int)key[1];
  k.i1 = (int)key[2];
  k.i2 = (int)key[3];
  k.w1 = key[4];
  k.w2 = key[5];
  k.B = key[6];
  for (int q = 0; q < 10; ++q) {
    k.p1[q] = key[7 + q];
    k.p2[q] = key[17 + q];
    k.ab[q] = key[27 + q];
  }
  return k;
}

static ClassKeyV blank_key(int kind) {
  ClassKeyV k{};
  k[0] = kind;
  return k;
}

static void counts_to_offsets(const double* counts, int S, std::vector<long>& off, const char* what) {
  off.assign(S + 1, 0);
  for (int s = 0; s < S; ++s) {
    const double c = counts ? counts[s] : 0.0;
    require(c >= 0 && c == std::floor(c), std::string(what) + ": per-slice counts must be non-negative integers");
    off[s + 1] = off[s] + (long)c;
  }
}

struct PotentialPlan {
  int np = 0, S = 0;
  std::vector<double> kx1, ky1;
  std::vector<std::vector<Term>> slices;
  std::vector<PotClass> classes;
  bool split = false;  // Peng rows split into per-Gaussian classes
};

// Relative cost of a plan's accumulation: per slice and class present, the
// class's terms padded to the 16-term groups of the matrix-core sum, plus one
// FP64 form-factor evaluation per pixel counted as kFormFactorTerms terms.
constexpr long kFormFactorTerms = 32;
static long plan_cost(const std::vector<std::vector<Term>>& slices, size_t n_classes) {
  std::vector<int> cnt(n_classes, 0);
  long cost = 0;
  for (const auto& sl : slices) {
    for (const Term& t : sl) ++cnt[t.cls];
    for (const Term& t : sl) {
      int& c = cnt[t.cls];
      if (c > 0) {
        cost += 16L * ((c + 15) / 16) + kFormFactorTerms;
        c = 0;
      }
    }
  }
  return cost;
}

static PotentialPlan plan_potential(const totem_potential_desc& d, bool allow_split = true) {
  PotentialPlan P;
  const int np = d.np, S = d.slices;
  require(d.paraflag == 'l' || d.paraflag == 'p' || d.paraflag == 'n', "paraflag must be 'l', 'p' or 'n'");
  require(np > 0 && np % 2 == 0, "P_Height/P_Width must be positive and even");
  require(S > 0, "Slices must be positive");
  require(d.kx && d.ky && d.s_2 && d.aper_true, "kx, ky, s_2 and AperTrue are required");
  require(size_supported(np), "potential grid " + std::to_string(np) + " has no compiled FFT plan",
          TOTEM_ERR_UNSUPPORTED);
  P.np = np;
  P.S = S;
  // The reference grids come from meshgrid: kx depends on the column, ky on the row.
  P.kx1.assign(d.kx, d.kx + np);
  P.ky1.resize(np);
  for (int r = 0; r < np; ++r) P.ky1[r] = d.ky[(size_t)r * np];
  for (int r = 0; r < np; ++r)
    for (int c = 0; c < np; ++c) {
      const size_t i = (size_t)r * np + c;
      require(d.kx[i] == P.kx1[c] && d.ky[i] == P.ky1[r],
              "kx/ky are not meshgrid-separable (kx must vary along columns, ky along rows)",
              TOTEM_ERR_UNSUPPORTED);
    }
  const bool lob = d.paraflag == 'l';
  const int ncol = lob ? 14 : 13, ncol_i = lob ? 26 : 25;
  std::map<ClassKey, int> cmap;
  // rows of one element arrive in runs, so the previous row's class is checked first
  ClassKeyV last{};
  int last_id = -1;
  auto class_of = [&](const ClassKeyV& key) {
    if (last_id >= 0 && key == last) return last_id;
    ClassKey k{key};
    auto it = cmap.find(k);
    int id;
    if (it != cmap.end()) {
      id = it->second;
    } else {
      id = (int)P.classes.size();
      P.classes.push_back(make_class(k.v));
      cmap.emplace(k, id);
    }
    last = key;
    last_id = id;
    return id;
  };
  P.slices.assign(S, {});
  std::vector<long> off;
  // atoms with a single (atomic or ionic) form factor
  counts_to_offsets(d.atom_slice, S, off, "atom_slice");
  if (off[S] > 0) require(d.proj_coff_mat != nullptr, "proj_coff_mat missing");
  for (int s = 0; s < S; ++s)
    for (long a = off[s]; a < off[s + 1]; ++a) {
      const double* r = d.proj_coff_mat + a * ncol;
      auto key = blank_key(lob ? kLobato : kPeng);
      for (int q = 0; q < 10; ++q) key[7 + q] = r[3 + q];
      if (lob) key[6] = r[13];
      if (d.absorp_n) {
        key[1] = 1;
        for (int q = 0; q < 10; ++q) key[27 + q] = d.absorp_n[a * 10 + q];
      }
      P.slices[s].push_back({r[0], r[1], r[2], class_of(key)});
    }
  // atoms with mixed atomic + ionic form factors
  counts_to_offsets(d.atom_slice_i, S, off, "atom_slice_i");
  if (off[S] > 0) require(d.proj_coff_mat_i != nullptr, "proj_coff_mat_i missing");
  for (int s = 0; s < S; ++s)
    for (long a = off[s]; a < off[s + 1]; ++a) {
      const double* r = d.proj_coff_mat_i + a * ncol_i;
      auto key = blank_key(lob ? kLobatoIon : kPengIon);
      key[4] = r[3];
      for (int q = 0; q < 10; ++q) key[7 + q] = r[4 + q];
      key[5] = r[14];
      for (int q = 0; q < 10; ++q) key[17 + q] = r[15 + q];
      if (lob) key[6] = r[25];
      if (d.absorp_n_i) {
        key[1] = 1;
        for (int q = 0; q < 10; ++q) key[27 + q] = d.absorp_n_i[a * 10 + q];
      }
      P.slices[s].push_back({r[0], r[1], r[2], class_of(key)});
    }
  // correction terms (Peng "extended"), p_corr: cuda_source_peng_and_corr.cu:652-783
  if (d.cim_len != 0) {
    require(d.cim_len > 0 && d.corr_info_matrix != nullptr, "corr_info_matrix missing");
    auto check_cls = [&](double c) {
      require(c >= 1 && c <= d.cim_len && c == std::floor(c), "correction class index out of range");
      return (int)c - 1;
    };
    counts_to_offsets(d.series_n_corr, S, off, "series_n_corr");
    if (off[S] > 0) require(d.ele_n_corr != nullptr, "ele_n_corr missing");
    for (int s = 0; s < S; ++s)
      for (long a = off[s]; a < off[s + 1]; ++a) {
        const double* r = d.ele_n_corr + a * 5;
        auto key = blank_key(kCorr);
        key[2] = check_cls(r[4]);
        key[6] = r[2];
        P.slices[s].push_back({r[0], r[1], r[3], class_of(key)});
      }
    counts_to_offsets(d.series_n_i_corr, S, off, "series_n_i_corr");
    if (off[S] > 0) require(d.ele_n_i_corr != nullptr, "ele_n_i_corr missing");
    for (int s = 0; s < S; ++s)
      for (long a = off[s]; a < off[s + 1]; ++a) {
        const double* r = d.ele_n_i_corr + a * 8;
        auto key = blank_key(kCorrIon);
        key[6] = r[2];
        key[4] = r[4];
        key[2] = check_cls(r[5]);
        key[5] = r[6];
        key[3] = check_cls(r[7]);
        P.slices[s].push_back({r[0], r[1], r[3], class_of(key)});
      }
  }
  // Rows with their own coefficient sets (erf-weighted slicing near faces scales
  // Peng's a_q per row) leave about one class per row, and every class costs a
  // form-factor evaluation per pixel.  Peng rows can then be split into their
  // Gaussian terms: a_q exp(-b_q s^2) becomes a class keyed by b_q alone
  // (likewise C_q exp(-D_q s^2) of the absorptive part) with a_q moved into the
  // term's weight.  That helps only when the b_q repeat across rows; when they do
  // not (a per-atom Debye-Waller B folded into every b_q) it multiplies the
  // classes, so the split is kept only if it lowers plan_cost by a fifth.
  if (allow_split && P.classes.size() > 16) {
    std::vector<ClassKeyV> key_of(P.classes.size());
    for (const auto& kv : cmap) key_of[kv.second] = kv.first.v;
    std::map<ClassKey, int> smap;
    std::vector<PotClass> sclasses;
    auto sclass = [&](const ClassKeyV& key) {
      ClassKey k{key};
      auto it = smap.find(k);
      if (it != smap.end()) return it->second;
      const int id = (int)sclasses.size();
      sclasses.push_back(make_class(key));
      smap.emplace(k, id);
      return id;
    };
    std::vector<std::vector<Term>> sslices(P.slices.size());
    for (size_t si = 0; si < P.slices.size(); ++si) {
      const auto& sl = P.slices[si];
      std::vector<Term>& out = sslices[si];
      out.reserve(sl.size() * 5);
      for (const Term& t : sl) {
        const PotClass& k = P.classes[t.cls];
        if (k.kind != kPeng && k.kind != kPengIon) {
          out.push_back({t.x, t.y, t.w, sclass(key_of[t.cls])});
          continue;
        }
        auto gaussians = [&](const double* p, double f) {
          for (int q = 0; q < 5; ++q) {
            if (p[2 * q] == 0.0) continue;
            ClassKeyV key = blank_key(kPeng);
            key[7] = 1.0;
            key[8] = p[2 * q + 1];
            out.push_back({t.x, t.y, t.w * f * p[2 * q], sclass(key)});
          }
        };
        gaussians(k.p1, k.kind == kPengIon ? k.w1 : 1.0);
        if (k.kind == kPengIon) gaussians(k.p2, k.w2);
        if (k.has_abs)
          for (int q = 0; q < 5; ++q) {
            if (k.ab[2 * q] == 0.0) continue;
            ClassKeyV key = blank_key(kPeng);
            key[1] = 1.0;
            key[27] = 1.0;
            key[28] = k.ab[2 * q + 1];
            out.push_back({t.x, t.y, t.w * k.ab[2 * q], sclass(key)});
          }
      }
    }
    if (5 * plan_cost(sslices, sclasses.size()) < 4 * plan_cost(P.slices, P.classes.size())) {
      P.slices.swap(sslices);
      P.classes.swap(sclasses);
      cmap.swap(smap);
      P.split = true;
    }
  }
  // Class ids in key order: any subset of the same classes keeps the same
  // relative order, so the per-slice segment order (and with it the summation
  // order) does not depend on which other atoms the specimen holds.
  std::vector<int> remap(P.classes.size());
  int next = 0;
  for (const auto& kv : cmap) remap[kv.second] = next++;
  std::vector<PotClass> sorted(P.classes.size());
  for (size_t c = 0; c < P.classes.size(); ++c) sorted[remap[c]] = P.classes[c];
  P.classes.swap(sorted);
  for (auto& sl : P.slices)
    for (auto& t : sl) t.cls = remap[t.cls];
  return P;
}

// Device-side inputs of one potential build: terms flattened slice by slice and
// grouped by form-factor class within a slice (segments), class table, grids.
struct PotDev {
  int np = 0, S = 0;
  const double *kx1 = nullptr, *ky1 = nullptr, *s2 = nullptr, *aper = nullptr, *corr = nullptr;
  const PotClass* classes = nullptr;
  const double *tx = nullptr, *ty = nullptr, *wt = nullptr;
  const PotSeg* segs = nullptr;
  std::vector<long> term_off, seg_off;  // host, [S + 1]
  size_t max_terms = 0;
  double sigma = 0.0, parameter = 0.0;
  bool sym = false;  // paired frequencies: a quarter of the tiles (hermitian_pairs)
};

// Per slice: phase tables, reciprocal-space accumulation, unnormalised inverse FFT
// with exp(i sigma/1000 V) fused into its column pass (or V copied out when asked).
void potential_run(const PotDev& p, float2* t_dev, double* v_re, double* v_im, hipStream_t st, DevBuf& dex,
                   DevBuf& dey) {
  const int np = p.np, S = p.S;
  const size_t np2 = (size_t)np * np;
  reserve(dex, std::max<size_t>(p.max_terms, 1) * np * sizeof(float2));
  reserve(dey, std::max<size_t>(p.max_terms, 1) * np * sizeof(float2));
  tw_table(np);
  const float sig = (float)(p.sigma / 1000.0);
  const bool want_v = v_re && v_im;
  std::vector<float2> hv;
  if (want_v) hv.resize(np2);
  for (int s = 0; s < S; ++s) {
    const int nt = (int)(p.term_off[s + 1] - p.term_off[s]);
    float2* ts = t_dev + (size_t)s * np2;
    phase_tables(st, p.tx + p.term_off[s], p.ty + p.term_off[s], p.wt + p.term_off[s], nt, p.kx1, p.ky1, np, p.sym,
                 dex.as<float2>(), dey.as<float2>());
    accumulate(st, dex.as<float2>(), dey.as<float2>(), p.segs + p.seg_off[s],
               (int)(p.seg_off[s + 1] - p.seg_off[s]), p.classes, p.s2, p.aper, p.corr, p.parameter, np, p.sym, ts);
    if (want_v) {
      fft2d(np, st, ts, ts, 1, +1, 0, 0.f);
      HIP_OK(hipMemcpyAsync(hv.data(), ts, np2 * sizeof(float2), hipMemcpyDeviceToHost, st));
      HIP_OK(hipStreamSynchronize(st));
      for (size_t i = 0; i < np2; ++i) {
        v_re[(size_t)s * np2 + i] = hv[i].x;
        v_im[(size_t)s * np2 + i] = hv[i].y;
        const float amp = std::exp(-sig * hv[i].y);
        hv[i] = make_float2(amp * std::cos(sig * hv[i].x), amp * std::sin(sig * hv[i].x));
      }
      HIP_OK(hipMemcpyAsync(ts, hv.data(), np2 * sizeof(float2), hipMemcpyHostToDevice, st));
    } else {
      fft2d(np, st, ts, ts, 1, +1, 1, sig);
    }
    HIP_OK(hipGetLastError());
  }
}

// Flattens a plan's terms slice by slice, stably grouped by class within a slice.
static void flatten_plan(PotentialPlan& P, std::vector<double>& tx, std::vector<double>& ty, std::vector<double>& wt,
                         std::vector<PotSeg>& segs, std::vector<long>& term_off, std::vector<long>& seg_off,
                         size_t& max_terms) {
  const int S = P.S;
  term_off.assign(S + 1, 0);
  seg_off.assign(S + 1, 0);
  max_terms = 0;
  for (int s = 0; s < S; ++s) {
    auto& v = P.slices[s];
    std::stable_sort(v.begin(), v.end(), [](const Term& a, const Term& b) { return a.cls < b.cls; });
    for (size_t i = 0; i < v.size(); ++i) {
      tx.push_back(v[i].x);
      ty.push_back(v[i].y);
      wt.push_back(v[i].w);
      if (i == 0 || v[i].cls != v[i - 1].cls) segs.push_back({v[i].cls, (int)i, (int)i + 1});
      else segs.back().end = (int)i + 1;
    }
    term_off[s + 1] = (long)tx.size();
    seg_off[s + 1] = (long)segs.size();
    max_terms = std::max(max_terms, v.size());
  }
}

// Phase-table scratch (terms of the largest slice x np x 16 bytes) above this
// size is released after the build instead of kept for the next one.
constexpr size_t kKeepBuilderScratch = 256ull << 20;

static void build_transmission_impl(const totem_potential_desc& d, float2* t_dev, double* v_re, double* v_im,
                                    hipStream_t st) {
  PotentialPlan P = plan_potential(d);
  const size_t np2 = (size_t)P.np * P.np;
  std::vector<double> tx, ty, wt;
  std::vector<PotSeg> segs;
  PotDev pd;
  flatten_plan(P, tx, ty, wt, segs, pd.term_off, pd.seg_off, pd.max_terms);
  // per-device builder scratch, one build at a time per device
  struct Scratch {
    std::mutex mu;
    DevBuf dtx, dty, dwt, dsegs, dcls, dkx, dky, ds2, dap, dcorr, dex, dey;
  };
  static std::mutex scratch_mu;
  static std::map<int, std::unique_ptr<Scratch>> scratch;
  int dev = 0;
  HIP_OK(hipGetDevice(&dev));
  Scratch* sc;
  {
    std::lock_guard<std::mutex> lk(scratch_mu);
    auto& p = scratch[dev];
    if (!p) p.reset(new Scratch());
    sc = p.get();
  }
  std::lock_guard<std::mutex> build_lock(sc->mu);
  upload(sc->dtx, tx, st);
  upload(sc->dty, ty, st);
  upload(sc->dwt, wt, st);
  upload(sc->dsegs, segs, st);
  upload(sc->dcls, P.classes, st);
  upload(sc->dkx, P.kx1, st);
  upload(sc->dky, P.ky1, st);
  reserve(sc->ds2, np2 * sizeof(double));
  HIP_OK(hipMemcpyAsync(sc->ds2.p, d.s_2, np2 * sizeof(double), hipMemcpyHostToDevice, st));
  reserve(sc->dap, np2 * sizeof(double));
  HIP_OK(hipMemcpyAsync(sc->dap.p, d.aper_true, np2 * sizeof(double), hipMemcpyHostToDevice, st));
  if (d.cim_len > 0) {
    reserve(sc->dcorr, np2 * d.cim_len * sizeof(double));
    HIP_OK(hipMemcpyAsync(sc->dcorr.p, d.corr_info_matrix, np2 * d.cim_len * sizeof(double), hipMemcpyHostToDevice,
                          st));
  }
  pd.np = P.np;
  pd.S = P.S;
  pd.kx1 = sc->dkx.as<double>();
  pd.ky1 = sc->dky.as<double>();
  pd.s2 = sc->ds2.as<double>();
  pd.aper = sc->dap.as<double>();
  pd.corr = sc->dcorr.as<double>();
  pd.classes = sc->dcls.as<PotClass>();
  pd.tx = sc->dtx.as<double>();
  pd.ty = sc->dty.as<double>();
  pd.wt = sc->dwt.as<double>();
  pd.segs = sc->dsegs.as<PotSeg>();
  pd.sigma = d.sigma;
  pd.parameter = d.parameter;
  pd.sym = hermitian_pairs(P.kx1.data(), P.ky1.data(), P.np);
  potential_run(pd, t_dev, v_re, v_im, st, sc->dex, sc->dey);
  HIP_OK(hipStreamSynchronize(st));
  // the phase tables grow with the terms of the largest slice; a large set is not
  // kept for the rest of the process
  if (sc->dex.bytes + sc->dey.bytes > kKeepBuilderScratch) {
    sc->dex.release();
    sc->dey.release();
  }
}

// ------------------------------------------------------- specimen on the device

// Correction terms this close (Angstrom) to a slice face are shared with the
// neighbouring slice (CalAllEquilent20210305_peng_corr.m:41-58, discon).
constexpr double kCorrDiscon = 0.015;

}  // namespace totem

// Resident specimen: static atoms and per-term data on the device; per phonon
// configuration the atoms are displaced, sliced and the transmission built
// without leaving the GPU except for the per-(slice, class) counts.
struct totem_specimen {
  std::mutex mu;
  int dev = 0;
  int np = 0, S = 0, Ct = 0, n_atoms = 0, n_terms = 0, n_faces = 0, bits = 1, K = 2, split = 0;
  double origin = 0, dz = 0, lo = 0, hi = 0, sigma = 0, parameter = 0, m = 0;
  bool has_rms = false, sym = false;
  size_t sort_bytes = 0;
  totem::DevBuf xyz, rms, pos, faces, term_atom, term_cls, term_w, term_mode, sub_base, keys, keys2, vals, vals2,
      wslot, counts, sort_tmp;
  totem::DevBuf tx, ty, wt, segs, classes, kx, ky, s2, aper, corr, ex, ey;
  std::vector<int> h_counts;
  std::vector<totem::PotSeg> h_segs;
};

namespace totem {

static totem_specimen* specimen_create_impl(const totem_specimen_desc& d) {
  require(d.pot.slices == 1, "specimen: pot must hold every atom's rows as one slice (slices = 1)");
  require(d.slices > 0, "specimen: slices must be positive");
  require(d.slice_thickness > 0 && std::isfinite(d.slice_thickness) && std::isfinite(d.slice_origin),
          "specimen: slice_thickness must be positive and finite");
  require(d.n_atoms >= 0 && (d.n_atoms == 0 || d.xyz), "specimen: xyz missing");
  require(d.n_terms >= 0 && (d.n_terms == 0 || d.term_atom), "specimen: term_atom missing");
  require(d.n_mid >= 0 && (d.n_mid == 0 || d.mid), "specimen: mid missing");
  require(d.multiceng >= 0 && std::isfinite(d.multiceng), "specimen: multiceng must be finite and >= 0");
  PotentialPlan P = plan_potential(d.pot, false);  // one term per row: term_atom indexes them
  const auto& terms = P.slices[0];
  require((long)terms.size() == (long)d.n_terms,
          "specimen: n_terms = " + std::to_string(d.n_terms) + " but pot holds " + std::to_string(terms.size()) +
              " rows");
  const int np = P.np, S = d.slices, C = (int)P.classes.size();
  // Lobato slicing is single-slice whatever multiceng says (ToTEM_submit.m:874-877).
  const double m = d.pot.paraflag == 'l' ? 0.0 : d.multiceng;
  // One-term classes for erf-weighted rows: class c's Gaussian q alone at sub_base[c] + q.
  std::vector<int> sub_base(std::max(C, 1), -1);
  std::vector<PotClass> classes = P.classes;
  int nq_max = 0;
  if (m > 0)
    for (int c = 0; c < C; ++c) {
      const PotClass k = P.classes[c];
      if (k.kind != kPeng && k.kind != kPengIon) continue;
      require(!k.has_abs, "specimen: absorptive coefficients with multiceng > 0 are not supported",
              TOTEM_ERR_UNSUPPORTED);
      const int nq = k.kind == kPengIon ? 10 : 5;
      nq_max = std::max(nq_max, nq);
      sub_base[c] = (int)classes.size();
      for (int q = 0; q < nq; ++q) {
        PotClass one = k;
        for (int j = 0; j < 5; ++j) {
          if (j != q) one.p1[2 * j] = 0.0;
          if (j + 5 != q) one.p2[2 * j] = 0.0;
        }
        classes.push_back(one);
      }
    }
  const int Ct = std::max<int>(1, (int)classes.size());
  require((long)S * Ct < (1L << 30), "specimen: slices x form-factor classes exceeds 2^30");
  std::vector<int> t_atom(d.term_atom, d.term_atom + d.n_terms), t_cls(d.n_terms), t_mode(d.n_terms);
  std::vector<double> t_w(d.n_terms);
  for (int i = 0; i < d.n_terms; ++i) {
    const int a = t_atom[i];
    require(a >= 0 && a < d.n_atoms, "specimen: term_atom out of range");
    require(terms[i].x == d.xyz[3 * a] && terms[i].y == d.xyz[3 * a + 1],
            "specimen: row " + std::to_string(i) + " does not sit at its atom's x, y");
    t_cls[i] = terms[i].cls;
    t_w[i] = terms[i].w;
    const int kind = P.classes[terms[i].cls].kind;
    t_mode[i] = kind == kPeng ? 1 : kind == kPengIon ? 2 : (kind == kCorr || kind == kCorrIon) ? 3 : 0;
  }
  // key slots per term: two slices at a rounding face, each split in two for
  // corrections, or every slice of the widened range times the Gaussian terms
  const int K = std::max(4, m > 0 ? ((int)std::floor(2.0 * m + 1.0) + 2) * nq_max : 0);
  require((long)d.n_terms * K < (1L << 31), "specimen: too many terms for the slot index");
  std::vector<double> faces;
  for (int m = 0; m < d.n_mid; ++m) {
    require(d.mid[m] >= 1 && d.mid[m] <= S, "specimen: mid slice out of range");
    faces.push_back(d.slice_origin + d.slice_thickness * (double)(d.mid[m] - 1) + d.slice_thickness);
  }
  std::vector<double> xyz(d.xyz, d.xyz + 3 * (size_t)d.n_atoms), rms;
  if (d.rms) {
    rms.assign(d.rms, d.rms + d.n_atoms);
    for (double r : rms) require(r >= 0 && std::isfinite(r), "specimen: rms must be finite and >= 0");
  }
  std::unique_ptr<totem_specimen> sp(new totem_specimen());
  HIP_OK(hipGetDevice(&sp->dev));
  sp->np = np;
  sp->S = S;
  sp->Ct = Ct;
  sp->K = K;
  sp->m = m;
  sp->split = (m >= 1.0 && d.slice_thickness > 2.0 * kCorrDiscon) ? 1 : 0;
  sp->n_atoms = d.n_atoms;
  sp->n_terms = d.n_terms;
  sp->n_faces = (int)faces.size();
  sp->origin = d.slice_origin;
  sp->dz = d.slice_thickness;
  sp->lo = d.slice_origin;                                                         // st[0]
  sp->hi = d.slice_origin + d.slice_thickness * (double)(S - 1) + d.slice_thickness;  // st[S-1] + dz
  sp->sigma = d.pot.sigma;
  sp->parameter = d.pot.parameter;
  sp->has_rms = d.rms != nullptr;
  sp->sym = hermitian_pairs(P.kx1.data(), P.ky1.data(), np);
  while ((1L << sp->bits) <= (long)S * Ct) ++sp->bits;
  hipStream_t st = lib_stream();
  const size_t np2 = (size_t)np * np, n2 = (size_t)K * std::max(d.n_terms, 1);
  upload(sp->xyz, xyz, st);
  upload(sp->rms, rms, st);
  upload(sp->faces, faces, st);
  upload(sp->term_atom, t_atom, st);
  upload(sp->term_cls, t_cls, st);
  upload(sp->term_w, t_w, st);
  upload(sp->term_mode, t_mode, st);
  upload(sp->sub_base, sub_base, st);
  upload(sp->classes, classes, st);
  upload(sp->kx, P.kx1, st);
  upload(sp->ky, P.ky1, st);
  reserve(sp->pos, std::max<size_t>(3 * (size_t)d.n_atoms, 1) * sizeof(double));
  reserve(sp->s2, np2 * sizeof(double));
  HIP_OK(hipMemcpyAsync(sp->s2.p, d.pot.s_2, np2 * sizeof(double), hipMemcpyHostToDevice, st));
  reserve(sp->aper, np2 * sizeof(double));
  HIP_OK(hipMemcpyAsync(sp->aper.p, d.pot.aper_true, np2 * sizeof(double), hipMemcpyHostToDevice, st));
  if (d.pot.cim_len > 0) {
    reserve(sp->corr, np2 * d.pot.cim_len * sizeof(double));
    HIP_OK(hipMemcpyAsync(sp->corr.p, d.pot.corr_info_matrix, np2 * d.pot.cim_len * sizeof(double),
                          hipMemcpyHostToDevice, st));
  }
  for (DevBuf* b : {&sp->keys, &sp->keys2, &sp->vals, &sp->vals2}) reserve(*b, n2 * sizeof(unsigned));
  reserve(sp->counts, (size_t)S * Ct * sizeof(int));
  for (DevBuf* b : {&sp->wslot, &sp->tx, &sp->ty, &sp->wt}) reserve(*b, n2 * sizeof(double));
  reserve(sp->segs, std::max<size_t>((size_t)S * Ct, 1) * sizeof(PotSeg));
  sp->sort_bytes = sort_pairs_temp_bytes((int)n2);
  reserve(sp->sort_tmp, sp->sort_bytes);
  sp->h_counts.resize((size_t)S * Ct);
  HIP_OK(hipStreamSynchronize(st));
  return sp.release();
}

static void specimen_positions_impl(totem_specimen& sp, uint64_t seed, int config, hipStream_t st) {
  displace(st, sp.n_atoms, sp.xyz.as<double>(), sp.has_rms ? sp.rms.as<double>() : nullptr, seed, config, sp.lo,
           sp.hi, sp.faces.as<double>(), sp.n_faces, sp.pos.as<double>());
  HIP_OK(hipGetLastError());
}

static void specimen_transmission_impl(totem_specimen& sp, uint64_t seed, int config, float2* t_dev,
                                       hipStream_t st) {
  const int S = sp.S, C = sp.Ct, n2 = sp.K * sp.n_terms;
  specimen_positions_impl(sp, seed, config, st);
  TermKeyArgs ka;
  ka.n = sp.n_terms;
  ka.K = sp.K;
  ka.S = S;
  ka.Ct = C;
  ka.split = sp.split;
  ka.term_atom = sp.term_atom.as<int>();
  ka.term_cls = sp.term_cls.as<int>();
  ka.term_mode = sp.term_mode.as<int>();
  ka.sub_base = sp.sub_base.as<int>();
  ka.term_w = sp.term_w.as<double>();
  ka.xyz = sp.pos.as<double>();
  ka.classes = sp.classes.as<PotClass>();
  ka.origin = sp.origin;
  ka.dz = sp.dz;
  ka.m = sp.m;
  ka.discon = kCorrDiscon;
  ka.keys = sp.keys.as<unsigned>();
  ka.vals = sp.vals.as<unsigned>();
  ka.wslot = sp.wslot.as<double>();
  ka.counts = sp.counts.as<int>();
  term_keys(st, ka);
  HIP_OK(hipGetLastError());
  unsigned* sorted_vals = sp.vals.as<unsigned>();
  if (n2 > 0)
    sorted_vals = sort_pairs(st, sp.sort_tmp.p, sp.keys.as<unsigned>(), sp.keys2.as<unsigned>(),
                             sp.vals.as<unsigned>(), sp.vals2.as<unsigned>(), n2, sp.bits);
  HIP_OK(hipMemcpyAsync(sp.h_counts.data(), sp.counts.p, sp.h_counts.size() * sizeof(int), hipMemcpyDeviceToHost,
                        st));
  HIP_OK(hipStreamSynchronize(st));
  // Per slice: terms [term_off[s], term_off[s+1]) of the sorted order, one
  // segment per class present, classes ascending -- the host build's layout.
  PotDev pd;
  pd.term_off.assign(S + 1, 0);
  pd.seg_off.assign(S + 1, 0);
  sp.h_segs.clear();
  long total = 0;
  for (int s = 0; s < S; ++s) {
    int in_slice = 0;
    for (int c = 0; c < C; ++c) {
      const int k = sp.h_counts[(size_t)s * C + c];
      if (k > 0) sp.h_segs.push_back({c, in_slice, in_slice + k});
      in_slice += k;
    }
    total += in_slice;
    pd.term_off[s + 1] = total;
    pd.seg_off[s + 1] = (long)sp.h_segs.size();
    pd.max_terms = std::max<size_t>(pd.max_terms, (size_t)in_slice);
  }
  require(total <= n2, "specimen: internal count mismatch");
  if (!sp.h_segs.empty())
    HIP_OK(hipMemcpyAsync(sp.segs.p, sp.h_segs.data(), sp.h_segs.size() * sizeof(PotSeg), hipMemcpyHostToDevice,
                          st));
  term_gather(st, (int)total, sp.K, sorted_vals, sp.term_atom.as<int>(), sp.wslot.as<double>(),
              sp.pos.as<double>(), sp.tx.as<double>(), sp.ty.as<double>(), sp.wt.as<double>());
  HIP_OK(hipGetLastError());
  pd.np = sp.np;
  pd.S = S;
  pd.kx1 = sp.kx.as<double>();
  pd.ky1 = sp.ky.as<double>();
  pd.s2 = sp.s2.as<double>();
  pd.aper = sp.aper.as<double>();
  pd.corr = sp.corr.as<double>();
  pd.classes = sp.classes.as<PotClass>();
  pd.tx = sp.tx.as<double>();
  pd.ty = sp.ty.as<double>();
  pd.wt = sp.wt.as<double>();
  pd.segs = sp.segs.as<PotSeg>();
  pd.sigma = sp.sigma;
  pd.parameter = sp.parameter;
  pd.sym = sp.sym;
  potential_run(pd, t_dev, nullptr, nullptr, st, sp.ex, sp.ey);
  HIP_OK(hipStreamSynchronize(st));
}

// ---------------------------------------------------------------- STEM propagation

struct StemWorkspace {
  float2* psi;
  float2* prop;  // the job's propagator in the columns pass's column-group order
  double* partial;
  ItemDev* items;
  JobDev* job;
  int* mid_of_slice;
  int* layer_hit;
  int* col_flags;  // [n]: column of the propagator holds a nonzero entry
};

// Workspace of `batch` probes in flight out of `total`: waves and detector
// partials per slot, the work items of every probe of the job, and `batch` more
// items at a fixed address for graph replays.
static size_t stem_ws_bytes(int n, int S, int A, int L, int batch, long total) {
  const int G = detector_groups_max(n);
  size_t b = 0;
  b += align_up((size_t)batch * n * n * sizeof(float2));
  b += align_up((size_t)n * n * sizeof(float2));
  b += align_up((size_t)batch * (L + 1) * A * G * sizeof(double));
  b += align_up((size_t)(total + batch) * sizeof(ItemDev));
  b += align_up(sizeof(JobDev));
  b += align_up((size_t)std::max(S, 1) * sizeof(int));
  b += align_up((size_t)std::max(L, 1) * sizeof(int));
  b += align_up((size_t)n * sizeof(int));
  return b;
}

static StemWorkspace carve(void* base, int n, int S, int A, int L, int batch, long total) {
  const int G = detector_groups_max(n);
  char* p = static_cast<char*>(base);
  StemWorkspace w;
  w.psi = reinterpret_cast<float2*>(p);
  p += align_up((size_t)batch * n * n * sizeof(float2));
  w.prop = reinterpret_cast<float2*>(p);
  p += align_up((size_t)n * n * sizeof(float2));
  w.partial = reinterpret_cast<double*>(p);
  p += align_up((size_t)batch * (L + 1) * A * G * sizeof(double));
  w.items = reinterpret_cast<ItemDev*>(p);
  p += align_up((size_t)(total + batch) * sizeof(ItemDev));
  w.job = reinterpret_cast<JobDev*>(p);
  p += align_up(sizeof(JobDev));
  w.mid_of_slice = reinterpret_cast<int*>(p);
  p += align_up((size_t)std::max(S, 1) * sizeof(int));
  w.layer_hit = reinterpret_cast<int*>(p);
  p += align_up((size_t)std::max(L, 1) * sizeof(int));
  w.col_flags = reinterpret_cast<int*>(p);
  return w;
}

void validate_job(const totem_stem_job& j) {
  require(j.n > 0 && j.np > 0 && j.slices > 0, "grid sizes and slice count must be positive");
  require(j.n % 2 == 0, "probe window side must be even");
  require(size_supported(j.n), "probe grid " + std::to_string(j.n) + " has no compiled FFT plan",
          TOTEM_ERR_UNSUPPORTED);
  require(j.t && j.probes && j.prop && j.aper && j.det, "t, probes, prop, aper and det are required");
  require(j.n_probes > 0 && j.n_aper > 0 && j.depth > 0 && j.width_red > 0, "w_num, aper_num, Depth, width_red > 0");
  require(j.n_mid >= 0 && (j.n_mid == 0 || (j.mid && j.aper2)), "mid slices need mid_ceng and aper2");
  require(j.step >= 0, "step must be non-negative");
  const long rows = (long)(j.depth - 1) / j.width_red;
  const long cols = std::min<long>(j.depth, j.width_red) - 1;
  const long rmax = (long)j.begin_row + rows * j.step, cmax = (long)j.begin_col + cols * j.step;
  require(j.begin_row >= 0 && j.begin_col >= 0, "beginRow/beginCol must be >= 1 (1-based)");
  if (j.wrap) {
    require(j.n <= j.np, "periodic windows need Height <= P_Height");
  } else {
    require(rmax + j.n <= j.np && cmax + j.n <= j.np,
            "probe window leaves the potential: last window ends at row " + std::to_string(rmax + j.n) +
                ", column " + std::to_string(cmax + j.n) + " of " + std::to_string(j.np));
  }
}

void mid_walk(const totem_stem_job& j, std::vector<int>& mos, std::vector<int>& hit) {
  // the reference advances one pointer through the sorted 1-based list
  // (cuda_source_lobato.cu:1340-1385): unmatched or repeated entries are never written
  mos.assign(j.slices, -1);
  hit.assign(std::max(j.n_mid, 1), 0);
  int li = 0;
  for (int s = 0; s < j.slices; ++s)
    if (li < j.n_mid && s + 1 == j.mid[li]) {
      mos[s] = li;
      hit[li] = 1;
      ++li;
    }
}

void window_origin(const totem_stem_job& j, int d, int& r0, int& c0) {
  long r = (long)j.begin_row + (long)(d / j.width_red) * j.step;
  long c = (long)j.begin_col + (long)(d % j.width_red) * j.step;
  if (j.wrap) {
    r %= j.np;
    c %= j.np;
  }
  r0 = (int)r;
  c0 = (int)c;
}

// Probes in flight that keep a batch's waves (n^2 complex64 each) within the
// 256 MB memory-side cache of the MI355X, leaving room for the slice and the
// propagator: then every slice step's wave traffic stays in that cache and only
// the transmission slices stream from HBM.  Measured optimum 192-224 MiB of waves
// for 256^2 to 2048^2 grids (1.1-1.2x the throughput of HBM-resident batches).
constexpr size_t kCacheWaveBytes = 208ull << 20;
static int cache_fit_batch(int n) { return (int)std::max<size_t>(1, kCacheWaveBytes / ((size_t)n * n * sizeof(float2))); }

// Batches measured by totem_stem_autotune and pass variants (plans) chosen by it or
// set by totem_stem_set_plan, per (device, n); they replace the cache-fit rule and
// variant 0.
static std::mutex g_tune_mu;
static std::map<std::pair<int, int>, int> g_tuned, g_plan;

int preferred_batch(int n) {
  int dev = 0;
  if (hipGetDevice(&dev) == hipSuccess) {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    auto it = g_tuned.find({dev, n});
    if (it != g_tuned.end()) return it->second;
  }
  return cache_fit_batch(n);
}

int current_plan(int n) {
  int dev = 0;
  if (hipGetDevice(&dev) == hipSuccess) {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    auto it = g_plan.find({dev, n});
    if (it != g_plan.end()) return it->second;
  }
  return 0;
}

static void set_plan(int n, int plan) {
  require(size_supported(n), "grid " + std::to_string(n) + " has no compiled FFT plan", TOTEM_ERR_UNSUPPORTED);
  require(plan >= 0 && plan < pass_variant_count(n),
          "plan " + std::to_string(plan) + " out of range for n=" + std::to_string(n) + " (" +
              std::to_string(pass_variant_count(n)) + " compiled)");
  int dev = 0;
  HIP_OK(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lk(g_tune_mu);
  g_plan[{dev, n}] = plan;
}

// The shortest wrapped run of C-column groups [lo, lo + len) (mod n) that covers
// every flagged column: the complement of the longest wrapped run of groups without
// a flag (the first such run if several tie).  No flag: len = 0; no empty group:
// the whole grid.
void column_band(const int* flags, int n, int C, int& lo, int& len) {
  const int G = n / C;
  std::vector<char> on(G, 0);
  for (int c = 0; c < n; ++c)
    if (flags[c]) on[c / C] = 1;
  int n_on = 0;
  for (int g = 0; g < G; ++g) n_on += on[g];
  if (n_on == 0) {
    lo = 0;
    len = 0;
    return;
  }
  if (n_on == G) {
    lo = 0;
    len = n;
    return;
  }
  int best_start = 0, best = 0;
  for (int g = 0; g < G; ++g) {
    if (on[g] || on[(g + G - 1) % G] == 0) continue;  // runs start after a flagged group
    int r = 0;
    while (r < G && !on[(g + r) % G]) ++r;
    if (r > best) {
      best = r;
      best_start = g;
    }
  }
  lo = ((best_start + best) % G) * C;
  len = (G - best) * C;
}

void set_job_band(JobDev& jd, const int* flags, int n, int C) {
  int lo = 0, len = 0;
  column_band(flags, n, C, lo, len);
  // rounded up to a multiple of 8 groups so that the columns pass's XCD order applies
  // (the added groups hold zeros and are transformed as such)
  int ng = len / C;
  if (ng < n / C) {
    ng = std::min(n / C, (ng + 7) / 8 * 8);
    len = ng * C;
  }
  jd.band_lo = lo;
  jd.band_len = len;
  jd.band_g0 = lo / C;
  jd.band_ng = ng;
}

void job_band(hipStream_t st, JobDev& jd, const float2* prop, int* flags, int n, int C) {
  column_flags(st, prop, flags, n);
  HIP_OK(hipGetLastError());
  std::vector<int> h(n);
  HIP_OK(hipMemcpyAsync(h.data(), flags, n * sizeof(int), hipMemcpyDeviceToHost, st));
  HIP_OK(hipStreamSynchronize(st));
  set_job_band(jd, h.data(), n, C);
}

// Per-kernel durations of the last timed run (TOTEM_RUN_TIMING), from HIP events
// recorded on the launch stream around every rows/columns pass.
struct KernelTimes {
  double rows_ms = 0.0, cols_ms = 0.0;
  long rows_n = 0, cols_n = 0;
};
static thread_local KernelTimes g_times;

static void run_stem_impl(const totem_stem_job& j, int batch, int flags, void* ws, size_t ws_bytes, hipStream_t st) {
  validate_job(j);
  const int n = j.n, S = j.slices, A = j.n_aper, L = j.n_mid, M = j.n_probes, D = j.depth;
  const long total = (long)M * D;
  require(batch > 0, "batch must be positive");
  batch = (int)std::min<long>(batch, total);
  const size_t need = stem_ws_bytes(n, S, A, L, batch, total);
  DevBuf own;
  if (!ws) {
    own.alloc(need);
    ws = own.p;
  } else {
    require(ws_bytes >= need, "workspace too small: need " + std::to_string(need) + " bytes", TOTEM_ERR_MEMORY);
  }
  StemWorkspace w = carve(ws, n, S, A, L, batch, total);
  const int plan = current_plan(n);            // one pass variant for the whole run
  const int G = detector_groups(n, plan), Gslot = detector_groups_max(n);
  tw_table(n);
  std::vector<int> mos, hit;
  mid_walk(j, mos, hit);
  // the propagator regrouped once per run into the columns pass's read order, and
  // the columns it leaves nonzero
  const int pc = pass_cols_width(n, plan);
  prop_column_groups(st, static_cast<const float2*>(j.prop), w.prop, n, pc);
  JobDev jd;
  job_band(st, jd, static_cast<const float2*>(j.prop), w.col_flags, n, pc);
  const int band_groups = std::min(jd.band_ng, n / pc);  // columns-pass groups per probe
  jd.t = static_cast<const float2*>(j.t);
  jd.prop = w.prop;
  jd.prop_cols = pc;
  jd.aper = j.aper;
  jd.aper2 = j.aper2;
  jd.mid_of_slice = L > 0 ? w.mid_of_slice : nullptr;
  jd.layer_hit = w.layer_hit;
  jd.det = j.det;
  jd.exit = static_cast<float2*>(j.exit_wave);
  jd.np = j.np;
  jd.S = S;
  jd.A = A;
  jd.L = L;
  jd.n_out = (int)total;
  jd.wrap = j.wrap;
  HIP_OK(hipMemcpyAsync(w.job, &jd, sizeof(jd), hipMemcpyHostToDevice, st));
  HIP_OK(hipMemcpyAsync(w.mid_of_slice, mos.data(), S * sizeof(int), hipMemcpyHostToDevice, st));
  HIP_OK(hipMemcpyAsync(w.layer_hit, hit.data(), hit.size() * sizeof(int), hipMemcpyHostToDevice, st));
  // the exit wave's columns outside the band are never written by the passes (the
  // detector reduction reads the band groups' partials only)
  if (jd.band_len < n && j.exit_wave) fill(st, static_cast<float2*>(j.exit_wave), total * n * n, 0.f, 0.f);
  HIP_OK(hipStreamSynchronize(st));  // host staging buffers above are locals
  // every probe's work item, uploaded once: chunk c runs items [c*batch, ...) in
  // wave slots 0..batch-1, which consecutive chunks reuse in stream order
  const float2* probes = static_cast<const float2*>(j.probes);
  std::vector<ItemDev> items(total);
  for (long o = 0; o < total; ++o) {
    const int m = (int)(o / D), d = (int)(o % D);
    const long k = o % batch;
    ItemDev& it = items[o];
    it.probe = probes + (size_t)m * n * n;
    it.job = 0;
    window_origin(j, d, it.r0, it.c0);
    it.start_step = 0;
    it.wave = w.psi + (size_t)k * n * n;
    it.part = w.partial + (size_t)k * (L + 1) * A * Gslot;
    it.out = (int)o;
  }
  ItemDev* const stage = w.items + total;  // graph replays read the chunk's items here
  HIP_OK(hipMemcpyAsync(w.items, items.data(), total * sizeof(ItemDev), hipMemcpyHostToDevice, st));
  HIP_OK(hipStreamSynchronize(st));  // the host list is a local
  auto record = [&](const ItemDev* it, int cnt) {
    for (int s = 0; s < S; ++s) {
      pass_rows(n, plan, st, it, cnt, w.job, s);
      if (band_groups > 0) pass_cols(n, plan, st, it, cnt, w.job, s, band_groups);
    }
    reduce_det(st, it, nullptr, cnt, w.job, G, L + 1, A);
  };
  std::vector<hipEvent_t> ev;
  if (flags & TOTEM_RUN_TIMING) {
    ev.resize(2 * S + 1);
    for (auto& e : ev) HIP_OK(hipEventCreate(&e));
  }
  hipGraphExec_t gexec = nullptr;
  int graph_count = -1;
  for (long i0 = 0; i0 < total; i0 += batch) {
    const int cnt = (int)std::min<long>(batch, total - i0);
    const ItemDev* ci = w.items + i0;
    if (flags & TOTEM_RUN_TIMING) {
      HIP_OK(hipEventRecord(ev[0], st));
      for (int s = 0; s < S; ++s) {
        pass_rows(n, plan, st, ci, cnt, w.job, s);
        HIP_OK(hipEventRecord(ev[2 * s + 1], st));
        if (band_groups > 0) pass_cols(n, plan, st, ci, cnt, w.job, s, band_groups);
        HIP_OK(hipEventRecord(ev[2 * s + 2], st));
      }
      reduce_det(st, ci, nullptr, cnt, w.job, G, L + 1, A);
      HIP_OK(hipEventSynchronize(ev[2 * S]));
      for (int s = 0; s < S; ++s) {
        float a = 0.f, b = 0.f;
        HIP_OK(hipEventElapsedTime(&a, ev[2 * s], ev[2 * s + 1]));
        HIP_OK(hipEventElapsedTime(&b, ev[2 * s + 1], ev[2 * s + 2]));
        g_times.rows_ms += a;
        g_times.cols_ms += b;
        g_times.rows_n += 1;
        g_times.cols_n += 1;
      }
    } else if (flags & TOTEM_RUN_GRAPH) {
      HIP_OK(hipMemcpyAsync(stage, ci, cnt * sizeof(ItemDev), hipMemcpyDeviceToDevice, st));
      if (graph_count != cnt) {
        if (gexec) HIP_OK(hipGraphExecDestroy(gexec));
        hipGraph_t g;
        HIP_OK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        record(stage, cnt);
        HIP_OK(hipStreamEndCapture(st, &g));
        HIP_OK(hipGraphInstantiate(&gexec, g, nullptr, nullptr, 0));
        HIP_OK(hipGraphDestroy(g));
        graph_count = cnt;
      }
      HIP_OK(hipGraphLaunch(gexec, st));
    } else {
      record(ci, cnt);
    }
    HIP_OK(hipGetLastError());
  }
  for (auto& e : ev) HIP_OK(hipEventDestroy(e));
  if (gexec) {
    HIP_OK(hipStreamSynchronize(st));
    HIP_OK(hipGraphExecDestroy(gexec));
  }
  if (own.p) HIP_OK(hipStreamSynchronize(st));
}

// Workspace of totem_stem_run_static: one array per GPU, in the library's code
// object, held from the library's first use on that GPU until the process ends.
// Host code never uses the array's name as an address (on the host it is only a
// stand-in for the device symbol): static_scratch() looks up each device's own
// copy once with hipGetSymbolAddress and every launch and copy uses that pointer.
// Nothing in it must start at zero: every wave slot is written at slice 0 and
// every detector partial a reduction reads is written earlier in the same run.
static __device__ __align__(256) unsigned char scratch[TOTEM_STATIC_SCRATCH_BYTES];

static void* static_scratch() {
  int dev = 0;
  HIP_OK(hipGetDevice(&dev));
  static std::mutex mu;
  static std::map<int, void*> ptr;  // device -> address of its scratch
  std::lock_guard<std::mutex> lk(mu);
  auto it = ptr.find(dev);
  if (it != ptr.end()) return it->second;
  void* p = nullptr;
  HIP_OK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
  require(p != nullptr, "static scratch has no device address", TOTEM_ERR_DEVICE);
  ptr.emplace(dev, p);
  return p;
}

// totem_stem_run without allocation: waves, detector partials, work items and the
// job entry in the static scratch; synchronous, so that sequential calls from any
// streams never share the array while a run is in flight.
static void run_stem_static_impl(const totem_stem_job& j, int batch, hipStream_t st) {
  validate_job(j);
  require(batch > 0, "batch must be positive");
  const long total = (long)j.n_probes * j.depth;
  batch = (int)std::min<long>(batch, total);
  const size_t need = stem_ws_bytes(j.n, j.slices, j.n_aper, std::max(j.n_mid, 0), batch, total);
  require(need <= (size_t)TOTEM_STATIC_SCRATCH_BYTES,
          "job needs " + std::to_string(need) + " bytes of workspace for " + std::to_string(batch) +
              " probes in flight; the static scratch holds " + std::to_string((size_t)TOTEM_STATIC_SCRATCH_BYTES) +
              " (use fewer probes per launch or totem_stem_run)",
          TOTEM_ERR_MEMORY);
  void* ws = static_scratch();
  run_stem_impl(j, batch, 0, ws, (size_t)TOTEM_STATIC_SCRATCH_BYTES, st);
  HIP_OK(hipStreamSynchronize(st));
}

// Pass variant and probes per launch by measurement: a synthetic job (unit
// transmission and probe, one detector) of 2 x b probes through `slices` slices,
// for every compiled variant of n and b around the cache-fit rule; the fastest
// per probe.slice becomes the (device, n) plan and batch.
static int autotune_impl(int n, int slices, double* best_rate) {
  require(size_supported(n), "grid " + std::to_string(n) + " has no compiled FFT plan", TOTEM_ERR_UNSUPPORTED);
  require(slices > 0 && slices <= 64, "autotune slices must be in 1..64");
  hipStream_t st = lib_stream();
  const int base = cache_fit_batch(n);
  std::vector<int> cand;
  for (double f : {0.5, 0.75, 1.0, 1.25, 1.5, 2.0}) {
    const int b = std::max(1, (int)std::lround(base * f));
    if (cand.empty() || cand.back() != b) cand.push_back(b);
  }
  const size_t nn = (size_t)n * n;
  const int bmax = cand.back();
  DevBuf t(slices * nn * sizeof(float2)), probe(nn * sizeof(float2)), prop(nn * sizeof(float2));
  DevBuf aper(nn * sizeof(double)), det((size_t)2 * bmax * sizeof(double));
  fill(st, t.as<float2>(), (long)(slices * nn), 1.f, 0.f);
  fill(st, probe.as<float2>(), (long)nn, 1.f, 0.f);
  fill(st, aper.as<double>(), (long)nn, 0.0);
  {
    // a propagator with the support of ToTEM's (the 2/3 band limit, |k| <= n/3 in
    // FFT order), so that the tuned runs move and transform the band a real run does
    std::vector<float2> h(nn);
    for (int r = 0; r < n; ++r)
      for (int c = 0; c < n; ++c) {
        const long kr = r < n / 2 ? r : r - n, kc = c < n / 2 ? c : c - n;
        const bool in = 9 * (kr * kr + kc * kc) <= (long)n * n;
        h[(size_t)r * n + c] = make_float2(in ? 1.f : 0.f, 0.f);
      }
    HIP_OK(hipMemcpyAsync(prop.p, h.data(), nn * sizeof(float2), hipMemcpyHostToDevice, st));
    HIP_OK(hipStreamSynchronize(st));
  }
  totem_stem_job j;
  std::memset(&j, 0, sizeof(j));
  j.n = n;
  j.np = n;
  j.slices = slices;
  j.t = t.p;
  j.probes = probe.p;
  j.n_probes = 1;
  j.prop = prop.p;
  j.aper = aper.as<double>();
  j.n_aper = 1;
  j.width_red = 1;
  j.wrap = 1;
  j.det = det.as<double>();
  DevBuf ws;
  hipEvent_t e0, e1;
  HIP_OK(hipEventCreate(&e0));
  HIP_OK(hipEventCreate(&e1));
  const int prev_plan = current_plan(n);
  int best = base, best_plan = 0;
  double best_r = 0.0;
  try {
    for (int plan = 0; plan < pass_variant_count(n); ++plan) {
      set_plan(n, plan);
      for (int b : cand) {
        j.depth = 2 * b;
        const size_t need = stem_ws_bytes(n, slices, 1, 0, b, 2 * b);
        if (ws.bytes < need) ws.alloc(need);
        run_stem_impl(j, b, 0, ws.p, ws.bytes, st);  // warm-up
        HIP_OK(hipEventRecord(e0, st));
        run_stem_impl(j, b, 0, ws.p, ws.bytes, st);
        HIP_OK(hipEventRecord(e1, st));
        HIP_OK(hipEventSynchronize(e1));
        float ms = 0.f;
        HIP_OK(hipEventElapsedTime(&ms, e0, e1));
        const double rate = 2.0 * b * slices / (ms * 1e-3);
        if (rate > best_r) {
          best_r = rate;
          best = b;
          best_plan = plan;
        }
      }
    }
  } catch (...) {
    set_plan(n, prev_plan);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    throw;
  }
  HIP_OK(hipEventDestroy(e0));
  HIP_OK(hipEventDestroy(e1));
  HIP_OK(hipStreamSynchronize(st));
  set_plan(n, best_plan);
  int dev = 0;
  HIP_OK(hipGetDevice(&dev));
  {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    g_tuned[{dev, n}] = best;
  }
  if (best_rate) *best_rate = best_r;
  return best;
}

// --------------------------------------------------------------------- myProjectfun

struct ProjectArgs {
  double *my_real_w, *my_imag_w, *aper, *atom_slice, *absorp_n, *atom_slice_i, *absorp_n_i, *proj_coff_mat,
      *proj_coff_mat_i, *kx, *ky, *s_2, *my_real_const, *my_img_const, *gfsf;
  int Height, Width, Depth, P_Height, P_Width, Slices, step, beginRow, beginCol, width_red;
  double sigma;
  int w_num, aper_num;
  double parameter;
  double* AperTrue;
  int* mid_ceng;
  int ceng_len;
  double *aper2, *series_n_corr, *series_n_i_corr, *ele_n_corr, *ele_n_i_corr, *corr_info_matrix;
  int cim_len;
  double *return_result, *mid_ceng_mat, *potentialx, *potentialy;
};

static totem_potential_desc potential_desc(const ProjectArgs& a, int paraflag) {
  totem_potential_desc d;
  std::memset(&d, 0, sizeof(d));
  d.paraflag = paraflag;
  d.np = a.P_Height;
  d.slices = a.Slices;
  d.atom_slice = a.atom_slice;
  d.proj_coff_mat = a.proj_coff_mat;
  d.absorp_n = a.absorp_n;
  d.atom_slice_i = a.atom_slice_i;
  d.proj_coff_mat_i = a.proj_coff_mat_i;
  d.absorp_n_i = a.absorp_n_i;
  if (paraflag != 'l' && a.cim_len != 0) {
    d.series_n_corr = a.series_n_corr;
    d.ele_n_corr = a.ele_n_corr;
    d.series_n_i_corr = a.series_n_i_corr;
    d.ele_n_i_corr = a.ele_n_i_corr;
    d.corr_info_matrix = a.corr_info_matrix;
    d.cim_len = a.cim_len;
  }
  d.kx = a.kx;
  d.ky = a.ky;
  d.s_2 = a.s_2;
  d.aper_true = a.AperTrue;
  d.sigma = a.sigma;
  d.parameter = a.parameter;
  return d;
}

static void zero_outputs(const ProjectArgs& a, bool potential_only) {
  const size_t np2 = (size_t)std::max(a.P_Height, 0) * std::max(a.P_Width, 0);
  if (!potential_only) {
    if (a.return_result && a.aper_num > 0 && a.Depth > 0)
      std::memset(a.return_result, 0, sizeof(double) * a.aper_num * a.Depth);
    if (a.mid_ceng_mat && a.ceng_len > 0 && a.aper_num > 0 && a.Depth > 0)
      std::memset(a.mid_ceng_mat, 0, sizeof(double) * a.aper_num * a.Depth * a.ceng_len);
  } else {
    if (a.potentialx && a.Slices > 0) std::memset(a.potentialx, 0, sizeof(double) * np2 * a.Slices);
    if (a.potentialy && a.Slices > 0) std::memset(a.potentialy, 0, sizeof(double) * np2 * a.Slices);
  }
}

static void project_impl(const ProjectArgs& a, int paraflag, bool potential_only) {
  zero_outputs(a, potential_only);
  require(a.P_Height == a.P_Width, "P_Height must equal P_Width (square potential grid)");
  require(a.Slices > 0, "Slices must be positive");
  select_device();
  hipStream_t st = lib_stream();
  const int np = a.P_Height, S = a.Slices;
  const size_t np2 = (size_t)np * np;
  totem_potential_desc pd = potential_desc(a, paraflag);
  DevBuf t(np2 * S * sizeof(float2));
  if (potential_only) {
    require(a.potentialx && a.potentialy, "potentialx/potentialy outputs are required");
    build_transmission_impl(pd, t.as<float2>(), nullptr, nullptr, st);
    std::vector<float2> h(np2 * S);
    HIP_OK(hipMemcpyAsync(h.data(), t.p, h.size() * sizeof(float2), hipMemcpyDeviceToHost, st));
    HIP_OK(hipStreamSynchronize(st));
    for (size_t i = 0; i < h.size(); ++i) {
      a.potentialx[i] = h[i].x;
      a.potentialy[i] = h[i].y;
    }
    return;
  }
  require(a.Height == a.Width, "Height must equal Width (square probe window)");
  require(a.my_real_w && a.my_imag_w && a.aper && a.my_real_const && a.my_img_const && a.gfsf,
          "probe, aperture, propagator and gfsf arrays are required");
  require(a.return_result && (a.ceng_len == 0 || a.mid_ceng_mat), "outputs are required");
  require(a.aper_num > 0 && a.w_num > 0 && a.Depth > 0, "aper_num, w_num and Depth must be positive");
  const int n = a.Height, M = a.w_num, A = a.aper_num, D = a.Depth, L = std::max(a.ceng_len, 0);
  const size_t n2 = (size_t)n * n;
  // Validate the scan geometry before the (expensive) potential build.
  totem_stem_job j;
  std::memset(&j, 0, sizeof(j));
  j.n = n;
  j.np = np;
  j.slices = S;
  j.n_probes = M;
  j.n_aper = A;
  j.mid = a.mid_ceng;
  j.n_mid = L;
  j.depth = D;
  j.width_red = a.width_red;
  j.step = a.step;
  j.begin_row = a.beginRow;
  j.begin_col = a.beginCol;
  j.wrap = 0;
  {
    totem_stem_job probe_check = j;
    probe_check.t = probe_check.probes = probe_check.prop = (const void*)1;
    probe_check.aper = probe_check.aper2 = (const double*)1;
    probe_check.det = (double*)1;
    validate_job(probe_check);
  }
  build_transmission_impl(pd, t.as<float2>(), nullptr, nullptr, st);
  // probes and propagator: double -> complex64 (initKernel_W, cuda_source_lobato.cu:39-57; :1207-1220)
  std::vector<float2> hp(n2 * M), hh(n2);
  for (size_t i = 0; i < n2 * M; ++i) hp[i] = make_float2((float)a.my_real_w[i], (float)a.my_imag_w[i]);
  for (size_t i = 0; i < n2; ++i) {
    const double re = a.my_real_const[i], im = a.my_img_const[i];
    hh[i] = std::fabs(re * re + im * im - 1.0) < 1e-12 ? unit_float2(re, im) : make_float2((float)re, (float)im);
  }
  DevBuf dprobe, dprop, daper, daper2, ddet;
  upload(dprobe, hp, st);
  upload(dprop, hh, st);
  daper.alloc(n2 * A * sizeof(double));
  HIP_OK(hipMemcpyAsync(daper.p, a.aper, n2 * A * sizeof(double), hipMemcpyHostToDevice, st));
  if (L > 0) {
    require(a.aper2 != nullptr, "aper2 is required when ceng_len > 0");
    daper2.alloc(n2 * A * sizeof(double));
    HIP_OK(hipMemcpyAsync(daper2.p, a.aper2, n2 * A * sizeof(double), hipMemcpyHostToDevice, st));
  }
  const size_t det_n = (size_t)(L + 1) * A * M * D;
  ddet.alloc(det_n * sizeof(double));
  j.t = t.p;
  j.probes = dprobe.p;
  j.prop = dprop.p;
  j.aper = daper.as<double>();
  j.aper2 = daper2.as<double>();
  j.det = ddet.as<double>();
  // probes in flight: as many as keep their waves inside the memory-side cache
  // (preferred_batch), and never more than half the free HBM (the reference sizes
  // its batches by free memory alone, cuda_source_lobato.cu:1249-1260)
  size_t fr = 0, tot = 0;
  HIP_OK(hipMemGetInfo(&fr, &tot));
  const size_t per = stem_ws_bytes(n, S, A, L, 2, 2) - stem_ws_bytes(n, S, A, L, 1, 1);
  long batch = (long)(fr / 2 / std::max<size_t>(per, 1));
  batch = std::max<long>(1, std::min<long>(batch, std::min<long>((long)M * D, preferred_batch(n))));
  DevBuf ws(stem_ws_bytes(n, S, A, L, (int)batch, (long)M * D));
  // more than one batch: the slice loop is captured once and replayed per batch
  run_stem_impl(j, (int)batch, (long)M * D > batch ? TOTEM_RUN_GRAPH : 0, ws.p, ws.bytes, st);
  std::vector<double> det(det_n);
  HIP_OK(hipMemcpyAsync(det.data(), ddet.p, det_n * sizeof(double), hipMemcpyDeviceToHost, st));
  HIP_OK(hipStreamSynchronize(st));
  // focal-spread weighting (accumulate_result / accumulate_result2, cuda_source_lobato.cu:981-1015)
  const size_t MD = (size_t)M * D;
  for (int aa = 0; aa < A; ++aa)
    for (int d = 0; d < D; ++d) {
      double s = 0.0;
      for (int m = 0; m < M; ++m) s += det[((size_t)L * A + aa) * MD + (size_t)m * D + d] * a.gfsf[m];
      a.return_result[(size_t)aa * D + d] = s;
    }
  for (int l = 0; l < L; ++l)
    for (int aa = 0; aa < A; ++aa)
      for (int d = 0; d < D; ++d) {
        double s = 0.0;
        for (int m = 0; m < M; ++m) s += det[((size_t)l * A + aa) * MD + (size_t)m * D + d] * a.gfsf[m];
        a.mid_ceng_mat[((size_t)l * A + aa) * D + d] = s;
      }
}

}  // namespace totem

using namespace totem;

#define TOTEM_PACK_ARGS                                                                                         \
  ProjectArgs a{my_real_w, my_imag_w, aper, atom_slice, absorp_n, atom_slice_i, absorp_n_i, proj_coff_mat,       \
                proj_coff_mat_i, kx, ky, s_2, my_real_const, my_img_const, gfsf, Height, Width, Depth, P_Height, \
                P_Width, Slices, step, beginRow, beginCol, width_red, sigma, w_num, aper_num, parameter,         \
                AperTrue, mid_ceng, ceng_len, aper2, series_n_corr, series_n_i_corr, ele_n_corr, ele_n_i_corr,   \
                corr_info_matrix, cim_len, return_result, mid_ceng_mat, potentialx, potentialy}

extern "C" {

int totem_project_lobato(TOTEM_MYPROJECTFUN_PARAMS) {
  TOTEM_PACK_ARGS;
  return guarded([&] { project_impl(a, 'l', false); });
}
int totem_project_peng(TOTEM_MYPROJECTFUN_PARAMS) {
  TOTEM_PACK_ARGS;
  return guarded([&] { project_impl(a, 'n', false); });
}
int totem_potential_lobato(TOTEM_MYPROJECTFUN_PARAMS) {
  TOTEM_PACK_ARGS;
  return guarded([&] { project_impl(a, 'l', true); });
}
int totem_potential_peng(TOTEM_MYPROJECTFUN_PARAMS) {
  TOTEM_PACK_ARGS;
  return guarded([&] { project_impl(a, 'n', true); });
}

const char* totem_last_error(void) { return totem::last_error().c_str(); }
int totem_abi_version(void) { return 2; }
int totem_fft_size_supported(int n) { return size_supported(n) ? 1 : 0; }

int totem_build_transmission(const totem_potential_desc* desc, void* t_dev, double* v_re, double* v_im,
                             void* stream) {
  return guarded([&] {
    require(desc && t_dev, "desc and t_dev are required");
    hipStream_t st = stream ? static_cast<hipStream_t>(stream) : lib_stream();
    build_transmission_impl(*desc, static_cast<float2*>(t_dev), v_re, v_im, st);
  });
}

int totem_potential_plan_stats(const totem_potential_desc* desc, int* classes, long long* terms, int* split,
                               int* paired) {
  return guarded([&] {
    require(desc != nullptr, "desc is required");
    PotentialPlan P = plan_potential(*desc);
    long long n = 0;
    for (const auto& sl : P.slices) n += (long long)sl.size();
    if (classes) *classes = (int)P.classes.size();
    if (terms) *terms = n;
    if (split) *split = P.split ? 1 : 0;
    if (paired) *paired = totem::hermitian_pairs(P.kx1.data(), P.ky1.data(), P.np) ? 1 : 0;
  });
}

int totem_hrtem_images(const totem_hrtem_desc* desc, double* paras, double* images, void* stream) {
  return guarded([&] {
    require(desc != nullptr, "desc is required");
    hipStream_t st = stream ? static_cast<hipStream_t>(stream) : totem::lib_stream();
    totem::hrtem_images_impl(*desc, paras, images, st);
  });
}

int totem_tcc_image(int n, int m, const double* wave, const double* tcc, const double* ap, const double* gfsf,
                    double* image, void* stream) {
  return guarded([&] {
    hipStream_t st = stream ? static_cast<hipStream_t>(stream) : totem::lib_stream();
    totem::tcc_image_impl(n, m, wave, tcc, ap, gfsf, image, st);
  });
}

int totem_weighted_pattern(int n, int m, const void* fields, const double* weights, int shift, double* out,
                           void* stream) {
  return guarded([&] {
    hipStream_t st = stream ? static_cast<hipStream_t>(stream) : totem::lib_stream();
    totem::weighted_pattern_impl(n, m, fields, weights, shift, out, st);
  });
}

int totem_specimen_create(const totem_specimen_desc* desc, totem_specimen** out) {
  return guarded([&] {
    require(desc && out, "null argument");
    *out = nullptr;
    *out = specimen_create_impl(*desc);
  });
}

int totem_specimen_transmission(totem_specimen* spec, unsigned long long seed, int config, void* t_dev,
                                void* stream) {
  return guarded([&] {
    require(spec && t_dev, "null argument");
    std::lock_guard<std::mutex> lk(spec->mu);
    HIP_OK(hipSetDevice(spec->dev));
    hipStream_t st = stream ? static_cast<hipStream_t>(stream) : lib_stream();
    specimen_transmission_impl(*spec, seed, config, static_cast<float2*>(t_dev), st);
  });
}

int totem_specimen_positions(totem_specimen* spec, unsigned long long seed, int config, double* xyz) {
  return guarded([&] {
    require(spec && xyz, "null argument");
    std::lock_guard<std::mutex> lk(spec->mu);
    HIP_OK(hipSetDevice(spec->dev));
    hipStream_t st = lib_stream();
    specimen_positions_impl(*spec, seed, config, st);
    if (spec->n_atoms > 0)
      HIP_OK(hipMemcpyAsync(xyz, spec->pos.p, 3 * (size_t)spec->n_atoms * sizeof(double), hipMemcpyDeviceToHost, st));
    HIP_OK(hipStreamSynchronize(st));
  });
}

int totem_specimen_destroy(totem_specimen* spec) {
  return guarded([&] {
    if (!spec) return;
    HIP_OK(hipSetDevice(spec->dev));
    HIP_OK(hipDeviceSynchronize());
    delete spec;
  });
}

int totem_stem_preferred_batch(int n) { return n > 0 ? totem::preferred_batch(n) : 0; }

int totem_stem_autotune(int n, int slices, int* best_batch, double* probe_slices_per_s) {
  return guarded([&] {
    require(best_batch != nullptr, "best_batch is required");
    *best_batch = totem::autotune_impl(n, slices, probe_slices_per_s);
  });
}

int totem_stem_plan_count(int n) { return n > 0 ? totem::pass_variant_count(n) : 0; }

int totem_stem_plan(int n, int* plan, char* desc, size_t cap) {
  return guarded([&] {
    require(size_supported(n), "grid " + std::to_string(n) + " has no compiled FFT plan", TOTEM_ERR_UNSUPPORTED);
    const int p = totem::current_plan(n);
    if (plan) *plan = p;
    if (desc && cap > 0) {
      const std::string d = totem::pass_variant_desc(n, p);
      std::snprintf(desc, cap, "%s", d.c_str());
    }
  });
}

int totem_stem_set_plan(int n, int plan) {
  return guarded([&] { totem::set_plan(n, plan); });
}

int totem_stem_band(const int* col_flags, int n, int group, int* lo, int* len) {
  return guarded([&] {
    require(col_flags && lo && len, "col_flags, lo and len are required");
    require(n > 0 && group > 0 && n % group == 0, "n must be a positive multiple of group");
    totem::column_band(col_flags, n, group, *lo, *len);
  });
}

int totem_stem_plan_columns(int n, int plan) {
  if (!size_supported(n) || plan < 0 || plan >= pass_variant_count(n)) return 0;
  return pass_cols_width(n, plan);
}

size_t totem_stem_workspace_bytes(const totem_stem_job* job, int batch) {
  if (!job || batch <= 0 || !size_supported(job->n)) return 0;
  return stem_ws_bytes(job->n, job->slices, job->n_aper, std::max(job->n_mid, 0), batch,
                       (long)job->n_probes * job->depth);
}

int totem_stem_run(const totem_stem_job* job, int batch, int flags, void* workspace, size_t ws_bytes, void* stream) {
  return guarded([&] {
    require(job != nullptr, "job is required");
    hipStream_t st = stream ? static_cast<hipStream_t>(stream) : lib_stream();
    run_stem_impl(*job, batch, flags, workspace, ws_bytes, st);
  });
}

int totem_stem_run_static(const totem_stem_job* job, int batch, void* stream) {
  return guarded([&] {
    require(job != nullptr, "job is required");
    run_stem_static_impl(*job, batch, static_cast<hipStream_t>(stream));
  });
}

void totem_kernel_times(int reset, double* rows_ms, long* rows_n, double* cols_ms, long* cols_n) {
  if (rows_ms) *rows_ms = g_times.rows_ms;
  if (rows_n) *rows_n = g_times.rows_n;
  if (cols_ms) *cols_ms = g_times.cols_ms;
  if (cols_n) *cols_n = g_times.cols_n;
  if (reset) g_times = KernelTimes();
}

int totem_fft2d(void* data, int n, int batch, int sign, void* stream) {
  return guarded([&] {
    require(data && batch > 0 && (sign == 1 || sign == -1), "bad fft2d arguments");
    require(size_supported(n), "no FFT plan for n=" + std::to_string(n), TOTEM_ERR_UNSUPPORTED);
    hipStream_t st = stream ? static_cast<hipStream_t>(stream) : lib_stream();
    tw_table(n);
    fft2d(n, st, static_cast<float2*>(data), static_cast<float2*>(data), batch, sign, 0, 0.f);
    HIP_OK(hipGetLastError());
    if (!stream) HIP_OK(hipStreamSynchronize(st));
  });
}

}  // extern "C"
