// Hot-path kernels of the multislice STEM engine (gfx950).
//
//  * k_pass_rows  -- per slice, along rows:  [inverse row FFT, 1/N^2] -> x t(window) -> forward row FFT
//  * k_pass_cols  -- per slice, along columns: forward column FFT -> x propagator ->
//                    [annular detectors] -> inverse column FFT   (exit slice: detectors only)
//  * k_reduce_det -- per finished probe: fixed-order compensated FP64 sum of detector partials
//  * k_phase / k_accumulate_mfma16 -- projected potential in reciprocal space per slice
//                    (structure-factor sums on the matrix cores, 16 x 16 pixel tiles)
//  * k_fft_rows / k_fft_cols -- plain batched 2-D FFT passes (potential IFFT, tests)
//
// A full 2-D forward+inverse FFT per slice therefore costs two read+write
// passes over the probe wave in HBM: 32 f N^2 bytes per probe.slice, where f is the
// fraction of reciprocal-space columns in which the propagator has a nonzero entry
// (the wave is zero elsewhere after every propagator multiply, so neither pass
// reads or writes those columns; 2/3 with the 2/3 band limit).
#include "fft_block.h"
#include "host_common.h"
#include "kernels.h"
#include "launch.h"
#include "unit_round.h"

#include <algorithm>
#include <cmath>
#include <map>
#include <mutex>
#include <string>
#include <type_traits>

// Tuning knobs (compile-time).
#ifndef TOTEM_ROWS_MINWAVES
#define TOTEM_ROWS_MINWAVES 1  // minimum waves per SIMD the rows pass is compiled for
#endif
#ifndef TOTEM_COLS_MINWAVES
#define TOTEM_COLS_MINWAVES 1
#endif
// Columns per columns-pass workgroup.  The workgroup exchanges C whole columns
// through LDS ((N + N/16) * C float2 with the bank padding), which must stay within
// 64 KiB: 16 columns of 512 points would need 68 KiB, and 4 columns of 2048 points
// fit only with the exchange split into two rounds of half the buffer.
#ifndef TOTEM_COLS_C512
#define TOTEM_COLS_C512 8
#endif
#ifndef TOTEM_COLS_C1024
#define TOTEM_COLS_C1024 4
#endif
#ifndef TOTEM_COLS_C1536
#define TOTEM_COLS_C1536 4
#endif
#ifndef TOTEM_COLS_C2048
#define TOTEM_COLS_C2048 4
#endif
#ifndef TOTEM_COLS_SPLIT2048
#define TOTEM_COLS_SPLIT2048 1  // 2048-point columns exchange through half buffers
#endif
#ifndef TOTEM_ROWS_SPLIT2048
#define TOTEM_ROWS_SPLIT2048 0  // 2048-point rows: half-buffer exchanges, 4 rows per workgroup
#endif
#ifndef TOTEM_ROWS_PP
#define TOTEM_ROWS_PP 0        // rows pass: ping-pong LDS (1) or one buffer (0)
#endif
#ifndef TOTEM_COLS_PP
#define TOTEM_COLS_PP 0
#endif
#ifndef TOTEM_XCD_REMAP
#define TOTEM_XCD_REMAP 1      // XCD-aware workgroup -> (probe, row/column group) mapping
#endif

namespace totem {

// ------------------------------------------------------------------------- helpers

#ifndef TOTEM_ROWS_K
#define TOTEM_ROWS_K 1  // row groups per workgroup (loads of the next one overlap the FFTs of this one)
#endif
#ifndef TOTEM_ROWS_TPREFETCH
#define TOTEM_ROWS_TPREFETCH 1  // load the transmission window together with the wave
#endif
#ifndef TOTEM_ROWS_THREADS
#define TOTEM_ROWS_THREADS 256  // threads per rows-pass workgroup
#endif
#ifndef TOTEM_ROWS_THREADS_SMALL
#define TOTEM_ROWS_THREADS_SMALL 128  // ... at 768 and 1024 points
#endif
// Rows per rows-pass workgroup also sets the wave tile (TR = min(RW, 4) rows of
// 16 / TR columns).  At 1024 points two rows per workgroup (2 x 8 tiles) measured
// best over the two passes together (rows 110 -> 97 us, columns 90 -> 98 us per
// launch of 26 probes); at 768 points 128-thread workgroups (4 rows, tiles
// unchanged) take the rows pass from 100.5 to 96.5 us; 2048 points keep 2 x 8
// tiles from 128 lanes per row.
// Pass variants ("plans"): each grid side has one or more schedules for its two
// passes -- the radix plan of each pass, rows per rows workgroup (which also sets
// the wave tile shared by both passes), columns per columns workgroup and whether
// the column exchanges run through half buffers.  Variant 0 is the default; the
// others are chosen at run time per device (totem_stem_set_plan, or measured by
// totem_stem_autotune), for a whole run or served grid side at a time.
template <int N> struct PassSpecBase {
  using RowPlan = Plan<N>;
  using ColPlan = Plan<N>;
  static constexpr int rows_threads = (N == 1024 || N == 768) ? TOTEM_ROWS_THREADS_SMALL
                                      : (N == 2048 && TOTEM_ROWS_SPLIT2048) ? 2 * TOTEM_ROWS_THREADS
                                                                             : TOTEM_ROWS_THREADS;
  static constexpr bool rows_split = N == 2048 && TOTEM_ROWS_SPLIT2048;
  static constexpr int cols_c = N < 512 ? 16
                                : N == 512 ? TOTEM_COLS_C512
                                : N == 1024 ? TOTEM_COLS_C1024
                                : N == 1536 ? TOTEM_COLS_C1536
                                : N == 2048 ? TOTEM_COLS_C2048 : 8;
  static constexpr bool cols_split = N == 2048 && TOTEM_COLS_SPLIT2048;
  static constexpr bool rows_tprefetch = TOTEM_ROWS_TPREFETCH;  // transmission loaded with the wave
  static constexpr int rows_k = TOTEM_ROWS_K;                   // row groups per rows workgroup, in turn
};
template <int N, int V> struct PassSpec : PassSpecBase<N> {};
// The plain 2-D FFT kernels (potential build, fft2d) use the base schedule of each
// side whatever the pass variants specialise.
constexpr int kBaseSpec = 15;
// Variant 1 of most sizes: the rows pass loads its transmission window after the
// inverse transform, not with the wave (fewer registers live across the
// transform: 1024 points 118 -> 88 VGPRs, the load latency then hidden by more waves).
template <int N> struct PassSpecLateT : PassSpecBase<N> {
  static constexpr bool rows_tprefetch = false;
};
template <> struct PassSpec<256, 1> : PassSpecLateT<256> {};
template <> struct PassSpec<512, 1> : PassSpecLateT<512> {};
template <> struct PassSpec<1024, 1> : PassSpecLateT<1024> {};
template <> struct PassSpec<1536, 1> : PassSpecLateT<1536> {};
template <> struct PassSpec<2048, 1> : PassSpecLateT<2048> {};
// 768: rows transformed by the 12-element plan 4.4.4.4.3 on 64 lanes (one row per
// wave, 2 x 8 tiles); variant 0 also the columns, 4 per columns workgroup, variant 1
// the columns by the 24-element plan, variant 2 both passes by the 24-element plan.
template <> struct PassSpec<768, 1> : PassSpecBase<768> {
  using RowPlan = PlanB<768>;
  static constexpr int rows_threads = 128;
};
template <> struct PassSpec<768, 0> : PassSpec<768, 1> {
  using ColPlan = PlanB<768>;
  static constexpr int cols_c = 4;
};
template <> struct PassSpec<768, 2> : PassSpecBase<768> {};
// 1024, variant 2: rows transformed by 8.8.8.2 on 128 lanes (8 elements per lane,
// fewer registers per wave, one more exchange), two rows per workgroup (2 x 8 tiles).
template <> struct PassSpec<1024, 2> : PassSpecBase<1024> {
  using RowPlan = PlanB<1024>;
  static constexpr int rows_threads = 256;
};
// ... and with the transmission loaded after the inverse transform (59 VGPRs, 8 waves/SIMD).
template <> struct PassSpec<1024, 3> : PassSpec<1024, 2> {
  static constexpr bool rows_tprefetch = false;
};
// (Each workgroup taking two row groups in turn, the second group's wave loads
// issued before the first group's transforms -- rows_k = 2 on this plan, 95 VGPRs
// -- measured 90.2 against 88.9 us per 26-probe launch and is not compiled.)
// 512, variant 2: rows transformed by 8.8.8 on 64 lanes (8 elements per lane), four
// rows per workgroup (4 x 4 tiles, as variant 0).
template <> struct PassSpec<512, 2> : PassSpecBase<512> {
  using RowPlan = PlanB<512>;
  static constexpr int rows_threads = 256;
};
// 1536, variant 2: rows by 4.4.4.4.2.3 on 128 lanes (12 elements per lane), two rows
// per rows workgroup (2 x 8 tiles); columns by the 24-element plan, 4 per workgroup.
// Per 26-probe launch: rows 102.6, columns 121.6 us against variant 0's 110.1 and
// 113.7 (the 2 x 8 tiles cost the columns pass what the rows pass gains; four rows
// per workgroup keep variant 0's tiles but take the rows pass to 136.8 us).
template <> struct PassSpec<1536, 2> : PassSpecBase<1536> {
  using RowPlan = PlanB<1536>;
  static constexpr int rows_threads = 256;
};
template <int N> constexpr int pass_variants() {
  return N == 1024 ? 4 : (N == 512 || N == 768 || N == 1536) ? 3 : (N >= 256 && N != 320) ? 2 : 1;
}

// Calls f(std::integral_constant<int, V>) for the compiled variant V == v of grid side S.
template <int S, int V = 0, class F> static auto on_variant(int v, F&& f) {
  if constexpr (V + 1 < pass_variants<S>()) {
    if (v != V) return on_variant<S, V + 1>(v, static_cast<F&&>(f));
  }
  return f(std::integral_constant<int, V>{});
}

template <int N, int V = 0> struct RowCfg {
  using Spec = PassSpec<N, V>;
  using PL = typename Spec::RowPlan;
  static constexpr bool TPREFETCH = Spec::rows_tprefetch;
  static constexpr int K = Spec::rows_k;
  static constexpr int T = PL::T;
  static constexpr int RW = (Spec::rows_threads / T) > 0 ? Spec::rows_threads / T : 1;  // rows per workgroup
  static constexpr int THREADS = RW * T;
  static constexpr bool SPLIT = Spec::rows_split;
  static constexpr int LL = SPLIT ? row_lds_len<N / 2>() : row_lds_len<N>();
  static constexpr bool PP = TOTEM_ROWS_PP != 0 && !SPLIT;
  static constexpr int NBUF = PP ? 2 : 1;
};
template <int N, int V = 0> struct ColCfg {
  using Spec = PassSpec<N, V>;
  using PL = typename Spec::ColPlan;
  static constexpr bool SPLIT = Spec::cols_split;
  static constexpr bool PP = TOTEM_COLS_PP != 0 && !SPLIT;
  static constexpr int NBUF = PP ? 2 : 1;
  static constexpr int T = PL::T;
  static constexpr int C = Spec::cols_c;  // columns per workgroup
  static constexpr int THREADS = T * C;
  static constexpr int LC = SPLIT ? col_lds_len<N / 2, C>() : col_lds_len<N, C>();  // one exchange buffer
};
template <int N> using FftRowCfg = RowCfg<N, kBaseSpec>;
template <int N> using FftColCfg = ColCfg<N, kBaseSpec>;
// Every pass kernel's static LDS (transform buffers plus the detector's wave sums)
// stays within 64 KiB per workgroup.
#ifndef TOTEM_LDS_LIMIT
#define TOTEM_LDS_LIMIT (64 * 1024)
#endif
constexpr int kLdsLimit = TOTEM_LDS_LIMIT;

#ifndef TOTEM_TILED
#define TOTEM_TILED 1  // in-flight waves in the tiled layout below (0: row-major)
#endif

// Layout of an in-flight wave [N][N] (the probe, the transmission, the propagator
// and the outputs stay row-major).  Tiled: TR x TC tiles of 16 elements (one
// 128-byte line), tiles row-major, elements row-major inside a tile.  A rows-pass
// workgroup (TR rows at a time) then streams whole lines, and a columns-pass
// wave (C >= TC columns x 64/C consecutive rows) touches whole lines instead of
// C x 8-byte segments of 64/C rows.  Element (r, c) sits at
//   ((r / TR) * (N / TC) + c / TC) * 16 + (r % TR) * TC + c % TC.
template <int N, int V = 0> struct WaveLay {
  static constexpr int RW = RowCfg<N, V>::RW, RT = RowCfg<N, V>::T, CT = ColCfg<N, V>::T;
  static constexpr int TR = TOTEM_TILED ? (RW < 4 ? RW : 4) : 1;
  static constexpr int TC = TOTEM_TILED ? 16 / TR : 1;
  static_assert(!TOTEM_TILED || (TR * TC == 16 && N % TC == 0 && RT % TC == 0 && RT % TR == 0 && RW % TR == 0 &&
                                 CT % TR == 0),
                "tile shape");
  __device__ static int at(int r, int c) {
    if constexpr (!TOTEM_TILED) return r * N + c;
    return ((r / TR) * (N / TC) + c / TC) * 16 + (r % TR) * TC + c % TC;
  }
  // element stride between (r, c + T*v) and (r, c + T*(v+1)) for T a multiple of TC
  static constexpr int row_step(int T) { return TOTEM_TILED ? 16 * (T / TC) : T; }
};

#ifndef TOTEM_ROWS_WAVESYNC
#define TOTEM_ROWS_WAVESYNC 1  // rows that fit in one wave exchange with wave-local barriers
#endif
#ifndef TOTEM_ROWS_LANE_TILED
#define TOTEM_ROWS_LANE_TILED 2  // 0: never, 1: always, 2: the sizes that keep workgroup barriers
#endif
// Rows-pass lane mapping, two forms:
//  * row-major lanes: T consecutive lanes per row, so a row with T <= 64 stays
//    inside one wave and its LDS exchanges need only wave-local barriers; the TR
//    rows of a tile are read by TR waves of the workgroup and share each line in
//    the caches;
//  * lane-tiled: the 64 lanes of a wave cover TR rows x 64/TR columns, whole
//    128-byte tiles per load and store instruction; a row then spans TR waves
//    (workgroup barriers), and the LDS rows are padded by 4 float2 so that the
//    TR rows of a wave fall on different banks.
// Measured with cache-fit batches (us per launch of the rows pass, workgroup
// barriers -> wave-local row-major -> lane-tiled): 256: 94 -> 88 -> 95,
// 512: 103.5 -> 100 -> 104, 768: 177 -> 100 -> 133, 1024: 110.5 -> 113 -> 107,
// 1536: 185 -> 116 -> 137, 2048 (T = 128, no wave-local form): 111 -> - -> 97.
template <int N, int V = 0> constexpr bool rows_lane_tiled() {
  return TOTEM_ROWS_LANE_TILED == 1 ||
         (TOTEM_ROWS_LANE_TILED == 2 && !(TOTEM_ROWS_WAVESYNC && RowCfg<N, V>::T <= 64 && N != 1024));
}
// LDS row pad of the lane-tiled form: the TR rows of a wave start 16 / TR float2
// apart modulo 16, so one instruction's accesses of the TR rows fall on disjoint banks.
#ifndef TOTEM_ROWS_PAD
#define TOTEM_ROWS_PAD 0  // 0: 16 / TR
#endif
template <int N, int V = 0> constexpr int rows_lds_pad();
template <int N, int V = 0> constexpr bool rows_wave_local() {
  return TOTEM_ROWS_WAVESYNC && !rows_lane_tiled<N, V>() && RowCfg<N, V>::T <= 64 && 64 % RowCfg<N, V>::T == 0;
}
template <int N, int V> constexpr int rows_lds_pad() {
  return !rows_lane_tiled<N, V>() ? 0 : TOTEM_ROWS_PAD ? TOTEM_ROWS_PAD : 16 / WaveLay<N, V>::TR;
}
template <int N, int V> __device__ __forceinline__ void row_lane(int tid, int& lr, int& t) {
  constexpr int T = RowCfg<N, V>::T, TR = rows_lane_tiled<N, V>() ? WaveLay<N, V>::TR : 1;
  const int q = tid / (TR * T), rem = tid - q * (TR * T);
  lr = q * TR + rem % TR;
  t = rem / TR;
}

// Workgroups are dispatched round-robin over the 8 XCDs, each with its own L2.
// With the remap (unit_map below), XCD x works through the groups
// [x*G/8, (x+1)*G/8) of every probe in order: neighbouring column groups (which
// share 128-byte lines) run back to back on one L2, and each L2 only ever sees 1/8
// of the propagator / a band of the transmission slice.  Any G with G/8 divisible
// by the units per workgroup gives a bijection; otherwise the plain probe-major
// order is used.  Taking one group of every probe of the launch in turn instead
// (so that consecutive workgroups of an XCD share a group's transmission rows or
// propagator columns) measured slower at every size (2048 points, 6 probes per
// launch: rows 91.5 -> 105 us, columns 108 -> 158 us; 1024: columns 88 -> 120 us).

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// ------------------------------------------------------------------------ pass R

#ifndef TOTEM_COLS_HPREFETCH
#define TOTEM_COLS_HPREFETCH 1  // load the propagator before the forward transform
#endif
#ifndef TOTEM_COLS_K
#define TOTEM_COLS_K 1
#endif

// Unit u (< K) of workgroup `bid` -> (probe item, group g of G per probe).
template <int N, int K>
__device__ __forceinline__ void unit_map(int bid, int u, int G, int& item, int& g) {
#if TOTEM_XCD_REMAP
  if ((G & 7) == 0 && ((G >> 3) % K) == 0) {
    const int xcd = bid & 7, gp = G >> 3;
    const int q = (bid >> 3) * K + u;
    item = q / gp;
    g = xcd * gp + (q - item * gp);
    return;
  }
#endif
  const int L = bid * K + u;
  item = L / G;
  g = L - item * G;
}

// The wave source (the probe, row-major, at slice 0; the item's tiled wave after)
// and destination are buffer resources: lanes hold 32-bit byte offsets and the
// offsets between a lane's elements are uniform (scalar registers), so the wave
// accesses need no 64-bit address arithmetic per element.
struct RowUnit {
  bool active;
  int slice;
  __amdgpu_buffer_rsrc_t src, dst;
  int src_off, src_step;  // bytes: the lane's first element, stride between consecutive v
  int dst_off;
  __amdgpu_buffer_rsrc_t trs;  // the transmission slice
  int t_off;                   // bytes: the lane's row in it
  int c0, np;
  int lo, len;                 // the job's column band
};

// A lane offset beyond every buffer resource's range: buffer loads there return 0
// and stores are dropped, without a memory access.
constexpr int kNoAccess = (int)0x80000000u;
// off if column c0 + T*v lies in a job's band [lo, lo + len) (mod N) (base = c0 - lo),
// else kNoAccess
template <int N, int T> __device__ __forceinline__ int band_off(int base, int v, int len, int off) {
  int d = base + T * v;
  if constexpr ((N & (N - 1)) == 0) {
    d &= N - 1;
  } else {
    d = d < 0 ? d + N : d;
  }
  return d < len ? off : kNoAccess;
}

template <int N, int V, int K>
__device__ __forceinline__ RowUnit row_unit(const ItemDev* __restrict__ items, const JobDev* __restrict__ jobs,
                                            int step, int u, int lr, int t) {
  using Cfg = RowCfg<N, V>;
  using Lay = WaveLay<N, V>;
  constexpr int RW = Cfg::RW, G = N / RW;
  RowUnit r;
  int item, rg;
  unit_map<N, K>(blockIdx.x, u, G, item, rg);
  const ItemDev it = items[item];
  const JobDev& jb = jobs[it.job];
  r.slice = step - it.start_step;
  r.active = r.slice >= 0 && r.slice < jb.S;
  const int row = rg * RW + lr;
  constexpr unsigned kBytes = (unsigned)N * N * sizeof(float2);
  r.dst = buf_rsrc(it.wave, kBytes);
  r.dst_off = Lay::at(row, t) * (int)sizeof(float2);
  const bool probe = r.slice == 0;
  r.src = buf_rsrc(probe ? it.probe : it.wave, kBytes);
  r.src_off = probe ? (row * N + t) * (int)sizeof(float2) : r.dst_off;
  r.src_step = (probe ? Cfg::T : Lay::row_step(Cfg::T)) * (int)sizeof(float2);
  r.np = jb.np;
  int pr = it.r0 + row;
  if (pr >= r.np) pr -= r.np;  // only reachable with wrap (validated on the host)
  const size_t np2 = (size_t)r.np * r.np;
  r.trs = buf_rsrc(jb.t + (size_t)(r.active ? r.slice : 0) * np2, (unsigned)(np2 * sizeof(float2)));
  r.t_off = pr * r.np * (int)sizeof(float2);
  r.c0 = it.c0;
  r.lo = jb.band_lo;
  r.len = jb.band_len;
  return r;
}

template <int N, int V>
__global__ void __launch_bounds__((RowCfg<N, V>::THREADS), TOTEM_ROWS_MINWAVES)
k_pass_rows(const ItemDev* __restrict__ items, const JobDev* __restrict__ jobs, int step, const float2* __restrict__ tw) {
  using Cfg = RowCfg<N, V>;
  using PL = typename Cfg::PL;
  constexpr int T = Cfg::T, RW = Cfg::RW, E = N / T, K = Cfg::K;
  constexpr int LLP = Cfg::LL + rows_lds_pad<N, V>();
  static_assert(Cfg::NBUF * RW * LLP * sizeof(float2) <= kLdsLimit, "rows-pass LDS above 64 KiB");
  __shared__ float2 lds[Cfg::NBUF * RW * LLP];
  float2* const lds1 = lds + (Cfg::NBUF - 1) * RW * LLP;
  int lr, t;
  row_lane<N, V>(threadIdx.x, lr, t);
  const RowLayT<rows_wave_local<N, V>()> lay{lr * LLP};
  constexpr int DSTEP = WaveLay<N, V>::row_step(T) * (int)sizeof(float2);
  float2 xa[E], xb[E];
  RowUnit ua = row_unit<N, V, K>(items, jobs, step, 0, lr, t);
  // After slice 0 the wave is zero outside the job's band and only its band columns are read.
  if (ua.active) {
    if (ua.slice == 0) {
#pragma unroll
      for (int v = 0; v < E; ++v) xa[v] = bld(ua.src, ua.src_off, ua.src_step * v);
    } else {
#pragma unroll
      for (int v = 0; v < E; ++v)
        xa[v] = bld(ua.src, band_off<N, T>(t - ua.lo, v, ua.len, ua.src_off), ua.src_step * v);
    }
  }
#pragma unroll
  for (int u = 0; u < K; ++u) {
    RowUnit ub;
    if (u + 1 < K) {
      ub = row_unit<N, V, K>(items, jobs, step, u + 1, lr, t);
      if (ub.active) {
#pragma unroll
        for (int v = 0; v < E; ++v)
          xb[v] = bld(ub.src, ub.slice == 0 ? ub.src_off : band_off<N, T>(t - ub.lo, v, ub.len, ub.src_off),
                      ub.src_step * v);
      }
    }
    if (ua.active) {  // uniform over the workgroup
      float2 tv[Cfg::TPREFETCH ? E : 1];
      if constexpr (Cfg::TPREFETCH) {
#pragma unroll
        for (int v = 0; v < E; ++v) {
          int pc = ua.c0 + t + T * v;
          if (pc >= ua.np) pc -= ua.np;
          tv[v] = bld(ua.trs, ua.t_off + pc * (int)sizeof(float2), 0);
        }
      }
      if (ua.slice > 0) {
        block_fft<N, +1, Cfg::PP, 0, Cfg::SPLIT, PL>(xa, lds, lds1, lay, t, tw);
        const float s = 1.0f / (float)(N * N);
#pragma unroll
        for (int v = 0; v < E; ++v) xa[v] = cscale(xa[v], s);
      }
#pragma unroll
      for (int v = 0; v < E; ++v) {
        if constexpr (Cfg::TPREFETCH) {
          xa[v] = cmul(xa[v], tv[v]);
        } else {
          int pc = ua.c0 + t + T * v;
          if (pc >= ua.np) pc -= ua.np;
          xa[v] = cmul(xa[v], bld(ua.trs, ua.t_off + pc * (int)sizeof(float2), 0));
        }
      }
      block_fft<N, -1, Cfg::PP, next_start<N, Cfg::PP, PL>(0), Cfg::SPLIT, PL>(xa, lds, lds1, lay, t, tw);
#pragma unroll
      for (int v = 0; v < E; ++v)  // the columns pass reads band columns only
        bst(ua.dst, band_off<N, T>(t - ua.lo, v, ua.len, ua.dst_off), DSTEP * v, xa[v]);
      if (u + 1 < K) __syncthreads();  // LDS reuse by the next unit
    }
    if (u + 1 < K) {
      ua = ub;
#pragma unroll
      for (int v = 0; v < E; ++v) xa[v] = xb[v];
    }
  }
}

// ------------------------------------------------------------------------ pass C

template <int N, int E, int T, int C>
__device__ __forceinline__ void detector_partial(const float2* x, const double* __restrict__ masks, bool centred,
                                                 int A, int t, int col, double* red, double* out, int G, int cg) {
  // sum_k |Psi(k)|^2 * mask(k) over this workgroup's columns, FP64
  double acc[kMaxAper];
#pragma unroll
  for (int a = 0; a < kMaxAper; ++a) acc[a] = 0.0;
  const int mcol = centred ? ((col + N / 2) % N) : col;
#pragma unroll
  for (int v = 0; v < E; ++v) {
    const int row = t + T * v;
    const int mrow = centred ? ((row + N / 2) % N) : row;
    const double inten = (double)x[v].x * (double)x[v].x + (double)x[v].y * (double)x[v].y;
    const double* m = masks + (size_t)mrow * N + mcol;
#pragma unroll
    for (int a = 0; a < kMaxAper; ++a)
      if (a < A) acc[a] += inten * gldd(m + (size_t)a * N * N);
  }
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  constexpr int NW = (T * C + 63) / 64;
#pragma unroll
  for (int a = 0; a < kMaxAper; ++a) {
    if (a < A) {
      double s = wave_sum(acc[a]);
      if (lane == 0) red[wave * kMaxAper + a] = s;
    }
  }
  __syncthreads();
  if ((int)threadIdx.x < A) {
    double s = 0.0;
    for (int q = 0; q < NW; ++q) s += red[q * kMaxAper + threadIdx.x];
    out[(size_t)threadIdx.x * G + cg] = s;
  }
  __syncthreads();
}

struct ColUnit {
  bool active, last;
  int slice, mid, cg, out;
  __amdgpu_buffer_rsrc_t wr;
  const JobDev* jb;
  double* pbase;
};

template <int N, int V, int K>
__device__ __forceinline__ ColUnit col_unit(const ItemDev* __restrict__ items, const JobDev* __restrict__ jobs,
                                            int step, int u, int groups) {
  using Cfg = ColCfg<N, V>;
  constexpr int G = N / Cfg::C;
  ColUnit r;
  int item, gi;
  // `groups` groups per item from the first group of the job's band on (all G groups
  // with the default band)
  unit_map<N, K>(blockIdx.x, u, groups, item, gi);
  const ItemDev it = items[item];
  r.jb = &jobs[it.job];
  r.cg = gi + r.jb->band_g0;
  if (r.cg >= G) r.cg -= G;
  r.slice = step - it.start_step;
  // a launch runs the largest band of its jobs; groups beyond a job's own band are idle
  r.active = r.slice >= 0 && r.slice < r.jb->S && gi < r.jb->band_ng;
  r.wr = buf_rsrc(it.wave, (unsigned)N * N * sizeof(float2));
  r.last = r.slice == r.jb->S - 1;
  r.mid = (r.active && r.jb->mid_of_slice) ? r.jb->mid_of_slice[r.slice] : -1;
  r.pbase = it.part;
  r.out = it.out;
  return r;
}

// One kernel for every slice; the detector work is a run-time branch (uniform
// per workgroup).  A variant compiled without the detector code was measured
// slower at every size (its register allocation gives fewer waves per SIMD:
// at 1024 points 160 vs 128 VGPRs, 101 vs 89 us per launch of 26 probes).
template <int N, int V>
__global__ void __launch_bounds__((ColCfg<N, V>::THREADS), TOTEM_COLS_MINWAVES)
k_pass_cols(const ItemDev* __restrict__ items, const JobDev* __restrict__ jobs, int step, int groups,
            const float2* __restrict__ tw) {
  using Cfg = ColCfg<N, V>;
  using PL = typename Cfg::PL;
  using Lay = WaveLay<N, V>;
  constexpr int T = Cfg::T, C = Cfg::C, E = N / T, G = N / C, K = TOTEM_COLS_K;
  constexpr int NW = (Cfg::THREADS + 63) / 64;
  constexpr int LC = Cfg::LC;
  static_assert(Cfg::NBUF * LC * sizeof(float2) + NW * kMaxAper * sizeof(double) <= kLdsLimit,
                "columns-pass LDS above 64 KiB");
  __shared__ float2 lds[Cfg::NBUF * LC];
  __shared__ double red[NW * kMaxAper];
  float2* const lds1 = lds + (Cfg::NBUF - 1) * LC;
  const int c = threadIdx.x % C;
  const int t = threadIdx.x / C;
  const ColLay<C> lay{c};
  constexpr unsigned kBytes = (unsigned)N * N * sizeof(float2);
  constexpr int kStep = T * N * (int)sizeof(float2);
  float2 xa[E], xb[E];
  ColUnit ua = col_unit<N, V, K>(items, jobs, step, 0, groups);
  // byte offsets of (row t, column): in the wave (WaveLay) and in row-major arrays (H, exit wave);
  // both advance by T*N elements per v
  int voffa = Lay::at(t, ua.cg * C + c) * (int)sizeof(float2);
  int hoffa = (t * N + ua.cg * C + c) * (int)sizeof(float2);
  if (ua.active) {
#pragma unroll
    for (int v = 0; v < E; ++v) xa[v] = bld(ua.wr, voffa, v * kStep);
  }
#pragma unroll
  for (int u = 0; u < K; ++u) {
    ColUnit ub;
    int voffb = 0, hoffb = 0;
    if (u + 1 < K) {
      ub = col_unit<N, V, K>(items, jobs, step, u + 1, groups);
      voffb = Lay::at(t, ub.cg * C + c) * (int)sizeof(float2);
      hoffb = (t * N + ub.cg * C + c) * (int)sizeof(float2);
      if (ub.active) {
#pragma unroll
        for (int v = 0; v < E; ++v) xb[v] = bld(ub.wr, voffb, v * kStep);
      }
    }
    if (ua.active) {  // uniform over the workgroup
      const JobDev& jb = *ua.jb;
      const int col = ua.cg * C + c;
      const auto hr = buf_rsrc(jb.prop, kBytes);
      // the propagator in column-group order: a wave's 64 lanes read 512 contiguous bytes
      const bool grouped = jb.prop_cols == C;
      const int poff = grouped ? ((ua.cg * N + t) * C + c) * (int)sizeof(float2) : hoffa;
      const int pstep = grouped ? T * C * (int)sizeof(float2) : kStep;
      float2 hv[TOTEM_COLS_HPREFETCH ? E : 1];
      if constexpr (TOTEM_COLS_HPREFETCH) {
#pragma unroll
        for (int v = 0; v < E; ++v) hv[v] = bld(hr, poff, v * pstep);
      }
      block_fft<N, -1, Cfg::PP, 0, Cfg::SPLIT, PL>(xa, lds, lds1, lay, t, tw);
#pragma unroll
      for (int v = 0; v < E; ++v) xa[v] = cmul(xa[v], TOTEM_COLS_HPREFETCH ? hv[v] : bld(hr, poff, v * pstep));
      bool keep = true;
      {
        if (ua.mid >= 0) {
          for (int a0 = 0; a0 < jb.A; a0 += kMaxAper)
            detector_partial<N, E, T, C>(xa, jb.aper2 + (size_t)a0 * N * N, false, min(kMaxAper, jb.A - a0), t,
                                         col, red, ua.pbase + ((size_t)ua.mid * jb.A + a0) * G, G, ua.cg);
        }
        if (ua.last) {
          for (int a0 = 0; a0 < jb.A; a0 += kMaxAper)
            detector_partial<N, E, T, C>(xa, jb.aper + (size_t)a0 * N * N, true, min(kMaxAper, jb.A - a0), t, col,
                                         red, ua.pbase + ((size_t)jb.L * jb.A + a0) * G, G, ua.cg);
          keep = false;  // the exit face needs no inverse transform
          if (jb.exit) {
            const auto er = buf_rsrc(jb.exit + (size_t)ua.out * N * N, kBytes);
#pragma unroll
            for (int v = 0; v < E; ++v) bst(er, hoffa, v * kStep, xa[v]);
          }
        }
      }
      if (keep) {
        block_fft<N, +1, Cfg::PP, next_start<N, Cfg::PP, PL>(0), Cfg::SPLIT, PL>(xa, lds, lds1, lay, t, tw);
#pragma unroll
        for (int v = 0; v < E; ++v) bst(ua.wr, voffa, v * kStep, xa[v]);
      }
      if (u + 1 < K) __syncthreads();  // LDS reuse by the next unit
    }
    if (u + 1 < K) {
      ua = ub;
      voffa = voffb;
      hoffa = hoffb;
#pragma unroll
      for (int v = 0; v < E; ++v) xa[v] = xb[v];
    }
  }
}

// Fixed-order Kahan sum of the detector partials of the listed items.
__global__ void k_reduce_det(const ItemDev* __restrict__ items, const int* __restrict__ list, int n,
                             const JobDev* __restrict__ jobs, int G, int lmax1, int amax) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= n * lmax1 * amax) return;
  const int a = idx % amax;
  const int layer = (idx / amax) % lmax1;
  const int e = idx / (amax * lmax1);
  const ItemDev it = items[list ? list[e] : e];
  const JobDev& jb = jobs[it.job];
  if (a >= jb.A || layer > jb.L) return;
  double s = 0.0;
  if (layer == jb.L || jb.layer_hit[layer]) {
    const double* p = it.part + ((size_t)layer * jb.A + a) * G;
    double comp = 0.0;
    for (int g = 0; g < G; ++g) {
      int d = g - jb.band_g0;  // only the job's band groups are written (the rest is zero)
      if (d < 0) d += G;
      if (d >= jb.band_ng) continue;
      const double y = p[g] - comp;
      const double tt = s + y;
      comp = (tt - s) - y;
      s = tt;
    }
  }
  jb.det[((size_t)layer * jb.A + a) * jb.n_out + it.out] = s;
}

// --------------------------------------------------------------- plain 2-D FFT passes

template <int N, int SIGN>
__global__ void __launch_bounds__(FftRowCfg<N>::THREADS) k_fft_rows(float2* __restrict__ data, long rows, const float2* __restrict__ tw) {
  using Cfg = FftRowCfg<N>;
  constexpr int T = Cfg::T, RW = Cfg::RW, E = N / T;
  static_assert(RW * Cfg::LL * sizeof(float2) <= kLdsLimit, "row FFT LDS above 64 KiB");
  __shared__ float2 lds[RW * Cfg::LL];
  const int lr = threadIdx.x / T;
  const int t = threadIdx.x - lr * T;
  const long row = (long)blockIdx.x * RW + lr;
  float2* w = data + row * N;
  float2 x[E];
  const bool ok = row < rows;
#pragma unroll
  for (int v = 0; v < E; ++v) x[v] = ok ? w[t + T * v] : make_float2(0.f, 0.f);
  block_fft<N, SIGN, false, 0, Cfg::SPLIT>(x, lds, lds, RowLay{lr * Cfg::LL}, t, tw);
  if (ok) {
#pragma unroll
    for (int v = 0; v < E; ++v) w[t + T * v] = x[v];
  }
}

// epi: 0 store, 1 store exp(i*sigma*V) (transmission), 2 store scaled by 1/N^2
template <int N, int SIGN>
__global__ void __launch_bounds__(FftColCfg<N>::THREADS)
k_fft_cols(const float2* __restrict__ src, float2* __restrict__ dst, int epi, float sigma,
           const float2* __restrict__ tw) {
  using Cfg = FftColCfg<N>;
  constexpr int T = Cfg::T, C = Cfg::C, E = N / T, G = N / C;
  static_assert(Cfg::LC * sizeof(float2) <= kLdsLimit, "column FFT LDS above 64 KiB");
  __shared__ float2 lds[Cfg::LC];
  const int mat = blockIdx.x / G;
  const int cg = blockIdx.x - mat * G;
  const int c = threadIdx.x % C;
  const int t = threadIdx.x / C;
  const int col = cg * C + c;
  const float2* s = src + (size_t)mat * N * N + col;
  float2* d = dst + (size_t)mat * N * N + col;
  float2 x[E];
#pragma unroll
  for (int v = 0; v < E; ++v) x[v] = s[(size_t)(t + T * v) * N];
  block_fft<N, SIGN, false, 0, Cfg::SPLIT>(x, lds, lds, ColLay<C>{c}, t, tw);
#pragma unroll
  for (int v = 0; v < E; ++v) {
    float2 y = x[v];
    if (epi == 1) {
      // t = exp(i sigma V) = exp(-sigma Im V) (cos(sigma Re V) + i sin(sigma Re V))
      const float amp = expf(-sigma * y.y);
      float sn, cs;
      sincosf(sigma * y.x, &sn, &cs);
      y = make_float2(amp * cs, amp * sn);
    } else if (epi == 2) {
      y = cscale(y, 1.0f / (float)((long)N * N));
    }
    d[(size_t)(t + T * v) * N] = y;
  }
}

// ------------------------------------------------------------------- potential build

// ex[j][c] = w_j exp(-2 pi i kx1[c] x_j), ey[j][r] = exp(-2 pi i ky1[r] y_j); the
// product k.x is formed and reduced modulo 1 in FP64 before the (FP64) sincospi, and
// the term's weight is applied in FP64 before the rounding to FP32.  Only the first
// nc entries of each table are formed (the accumulation reads no more).
__global__ void k_phase(const double* __restrict__ tx, const double* __restrict__ ty,
                        const double* __restrict__ wt, int n_terms, const double* __restrict__ kx1,
                        const double* __restrict__ ky1, int np, int nc, float2* __restrict__ ex,
                        float2* __restrict__ ey) {
  const int c = blockIdx.y * blockDim.x + threadIdx.x;
  const int j = blockIdx.x;
  if (c >= nc || j >= n_terms) return;
  double fx = kx1[c] * tx[j];
  fx -= rint(fx);
  double fy = ky1[c] * ty[j];
  fy -= rint(fy);
  double sx, cx, sy, cy;
  sincospi(2.0 * fx, &sx, &cx);
  sincospi(2.0 * fy, &sy, &cy);
  const double w = wt[j];
  ex[(size_t)j * np + c] = make_float2((float)(w * cx), (float)(-w * sx));
  ey[(size_t)j * np + c] = make_float2((float)cy, (float)-sy);
}

// Terms with a zero coefficient are skipped: the one-term classes of erf-weighted
// slicing keep a single Gaussian (the branch is uniform: one class per wave).
__device__ __forceinline__ double peng5(const double* p, double s2) {
  double f = 0.0;
#pragma unroll
  for (int q = 0; q < 5; ++q)
    if (p[2 * q] != 0.0) f += p[2 * q] * exp(-p[2 * q + 1] * s2);
  return f;
}
__device__ __forceinline__ double lobato5(const double* p, double g2) {
  double f = 0.0;
#pragma unroll
  for (int q = 0; q < 5; ++q) {
    const double d = 1.0 + p[2 * q + 1] * g2;
    f += p[2 * q] * (2.0 + p[2 * q + 1] * g2) / (d * d);
  }
  return f;
}

// Form factor of a class at one reciprocal-space pixel (FP64, complex when absorptive).
__device__ __forceinline__ double2 form_factor(const PotClass& k, double s2, const double* __restrict__ corr,
                                               size_t pix, size_t np2) {
  double re = 0.0;
  switch (k.kind) {
    case kPeng: re = peng5(k.p1, s2); break;
    case kLobato: re = lobato5(k.p1, 4.0 * s2) * exp(-k.B * s2); break;
    case kPengIon: re = k.w1 * peng5(k.p1, s2) + k.w2 * peng5(k.p2, s2); break;
    case kLobatoIon: re = (k.w1 * lobato5(k.p1, 4.0 * s2) + k.w2 * lobato5(k.p2, 4.0 * s2)) * exp(-k.B * s2); break;
    case kCorr: re = corr[(size_t)k.i1 * np2 + pix] * exp(-k.B * s2); break;
    case kCorrIon:
      re = (k.w1 * corr[(size_t)k.i1 * np2 + pix] + k.w2 * corr[(size_t)k.i2 * np2 + pix]) * exp(-k.B * s2);
      break;
    default: break;
  }
  const double im = k.has_abs ? peng5(k.ab, s2) : 0.0;
  return make_double2(re, im);
}

// One slice: V_k(r,c) = PARAMETER * AperTrue * sum_classes f_cls(k) * sum_{j in cls} w_j ex_j[c] ey_j[r],
// written at the ifftshifted position (the reference's quadrant swap) as complex64.
//
// Per class the sum over terms is a complex matrix product, A[r][j] = ey_j[r] times
// B[j][c] = ex_j[c] (the weight w_j already in ex), on the matrix cores in FP32 (v_mfma_f32_16x16x4f32, four
// terms per product): lane l supplies A[l % 16][l / 16] and B[l / 16][l % 16], and
// accumulator element i of lane l is C[4 (l / 16) + i][l % 16].  The four real
// products are kept apart,
//   P1 = Ar Br,  P2 = Ai Bi,  P3 = Ar Bi,  P4 = Ai Br  (summed over terms),
// because the weights are real and the grid's frequencies come in pairs
// (ky[N - r] = -ky[r], kx[N - c] = -kx[c]): then ey_j[N - r] = conj(ey_j[r]) and
// ex_j[N - c] = conj(ex_j[c]), and one tile's products give four pixels,
//   S(r, c)         = (P1 - P2) + i (P3 + P4)
//   S(r, N - c)     = (P1 + P2) + i (P4 - P3)
//   S(N - r, c)     = (P1 + P2) + i (P3 - P4)
//   S(N - r, N - c) = (P1 - P2) - i (P3 + P4).
// With `sym` set the tiles cover rows and columns [0, N/2 + 16) (the Nyquist row
// and column, which have no pair, among them) and rows / columns r in
// [1, N/2 - 16] also write their pair N - r: every pixel is written once.
// Without it (a grid whose frequencies are not paired) the tiles cover the grid.
// W waves share a tile, each taking every W-th group of 16 terms of a class; the
// partial products are summed in LDS and each thread then weights one tile
// position's pixels by the class's FP64 form factor at each of them.
#ifndef TOTEM_ACC_MINWAVES
#define TOTEM_ACC_MINWAVES 1  // minimum waves per SIMD the accumulation is compiled for
#endif
typedef float floatx4 __attribute__((ext_vector_type(4)));
template <int W>
__global__ void __launch_bounds__(64 * W, TOTEM_ACC_MINWAVES)
k_accumulate_mfma16(const float2* __restrict__ ex, const float2* __restrict__ ey, const PotSeg* __restrict__ segs, int n_segs, const PotClass* __restrict__ classes,
                    const double* __restrict__ s2, const double* __restrict__ aper_true,
                    const double* __restrict__ corr, double parameter, int np, int sym,
                    float2* __restrict__ vshift) {
  constexpr int NT = 64 * W, NO = 1024 / NT;  // threads; output pixels per thread
  static_assert(NO * NT == 1024, "four pixels per tile position");
  __shared__ float red[W * 4 * 4 * 64];      // [wave][product][element][lane]
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int dt = sym ? np / 32 + 1 : np / 16;
  const int tr = blockIdx.x / dt, tc = blockIdx.x - tr * dt;
  const int r0 = tr * 16, c0 = tc * 16;
  const int li = lane & 15, lk = lane >> 4;
  const size_t np2 = (size_t)np * np;
  // this thread's tile position, its LDS source (lane, element) and output pixels
  const int pos = threadIdx.x & 255;
  const int pr = pos >> 4, pcl = pos & 15;
  const int src = (pr >> 2) * 16 + pcl + 64 * (pr & 3);  // source lane + 64 * accumulator element
  const int r = r0 + pr, c = c0 + pcl;
  const int pair_hi = np / 2 - 16;
  bool valid[NO];
  bool cj_r[NO], cj_c[NO];
  size_t pix[NO];
  double s2v[NO], vr[NO], vi[NO];
#pragma unroll
  for (int o = 0; o < NO; ++o) {
    const int quad = (threadIdx.x >> 8) + (NT >> 8) * o;  // bit 0: column pair, bit 1: row pair
    cj_c[o] = quad & 1;
    cj_r[o] = (quad >> 1) & 1;
    valid[o] = (!cj_r[o] || (sym && r >= 1 && r <= pair_hi)) && (!cj_c[o] || (sym && c >= 1 && c <= pair_hi));
    const int R = cj_r[o] ? np - r : r, C = cj_c[o] ? np - c : c;
    pix[o] = (size_t)R * np + C;
    s2v[o] = valid[o] ? s2[pix[o]] : 0.0;
    vr[o] = vi[o] = 0.0;
  }
  for (int sgi = 0; sgi < n_segs; ++sgi) {
    const PotSeg sg = segs[sgi];
    floatx4 p1 = {}, p2 = {}, p3 = {}, p4 = {};
    for (int j = sg.begin + 16 * wave; j < sg.end; j += 16 * W) {
      float2 a[4], b[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int jj = j + 4 * u + lk;
        a[u] = b[u] = make_float2(0.f, 0.f);
        if (jj < sg.end) {
          a[u] = ey[(size_t)jj * np + r0 + li];
          b[u] = ex[(size_t)jj * np + c0 + li];
        }
      }
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        p1 = __builtin_amdgcn_mfma_f32_16x16x4f32(a[u].x, b[u].x, p1, 0, 0, 0);
        p2 = __builtin_amdgcn_mfma_f32_16x16x4f32(a[u].y, b[u].y, p2, 0, 0, 0);
        p3 = __builtin_amdgcn_mfma_f32_16x16x4f32(a[u].x, b[u].y, p3, 0, 0, 0);
        p4 = __builtin_amdgcn_mfma_f32_16x16x4f32(a[u].y, b[u].x, p4, 0, 0, 0);
      }
    }
    float* const mine = red + wave * 1024;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      mine[0 * 256 + 64 * i + lane] = p1[i];
      mine[1 * 256 + 64 * i + lane] = p2[i];
      mine[2 * 256 + 64 * i + lane] = p3[i];
      mine[3 * 256 + 64 * i + lane] = p4[i];
    }
    __syncthreads();
    float q1 = 0.f, q2 = 0.f, q3 = 0.f, q4 = 0.f;
#pragma unroll
    for (int w = 0; w < W; ++w) {
      const float* o = red + w * 1024 + src;
      q1 += o[0];
      q2 += o[256];
      q3 += o[512];
      q4 += o[768];
    }
    __syncthreads();  // red is rewritten by the next class
    const PotClass& cls = classes[sg.cls];
#pragma unroll
    for (int o = 0; o < NO; ++o) {
      if (!valid[o]) continue;
      const double sre = cj_r[o] == cj_c[o] ? (double)(q1 - q2) : (double)(q1 + q2);
      const double sim = (double)(cj_c[o] ? -q3 : q3) + (double)(cj_r[o] ? -q4 : q4);
      const double2 f = form_factor(cls, s2v[o], corr, pix[o], np2);
      vr[o] += f.x * sre - f.y * sim;
      vi[o] += f.x * sim + f.y * sre;
    }
  }
  const int half = np / 2;
#pragma unroll
  for (int o = 0; o < NO; ++o) {
    if (!valid[o]) continue;
    const int R = (int)(pix[o] / np), C = (int)(pix[o] - (size_t)R * np);
    const double scale = parameter * aper_true[pix[o]];
    const int rs = (R + half) % np, cs = (C + half) % np;
    vshift[(size_t)rs * np + cs] = make_float2((float)(vr[o] * scale), (float)(vi[o] * scale));
  }
}

// prop [N][N] row-major -> column groups of C columns, rows in order: element
// (r, g*C + c) at (g*N + r)*C + c.
__global__ void k_prop_cols(const float2* __restrict__ prop, float2* __restrict__ out, int n, int C) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (long)n * n) return;
  const int r = (int)(i / n), col = (int)(i - (long)r * n);
  const int g = col / C, c = col - g * C;
  out[((long)g * n + r) * C + c] = prop[i];
}

// flags[c] = 0 for every column, then one thread per (column, chunk of n / kFlagChunks
// rows) sets it to 1 on a nonzero entry (every writer stores the same value);
// consecutive threads read consecutive entries of a row.
constexpr int kFlagChunks = 32;
__global__ void k_zero_i32(int* __restrict__ p, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
}
__global__ void k_column_flags(const float2* __restrict__ prop, int* __restrict__ flags, int n) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n) return;
  const int rows = (n + kFlagChunks - 1) / kFlagChunks;
  const int r0 = blockIdx.y * rows, r1 = min(n, r0 + rows);
  for (int r = r0; r < r1; ++r) {
    const float2 h = prop[(size_t)r * n + c];
    if (h.x != 0.f || h.y != 0.f) {
      flags[c] = 1;
      return;
    }
  }
}

__global__ void k_fill_c64(float2* __restrict__ p, long n, float re, float im) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = make_float2(re, im);
}
__global__ void k_fill_f64(double* __restrict__ p, long n, double v) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// --------------------------------------------------------------------- launchers

#define TOTEM_FOR_SIZES(X) X(64) X(128) X(256) X(320) X(512) X(768) X(1024) X(1536) X(2048)

// Twiddle tables of every plan in one static device array (no allocation on any
// path that launches the passes); each table starts at an even entry so that its
// float4 reads stay 16-byte aligned.
constexpr int tw_even(int n) { return (n + 1) & ~1; }
template <class Q, class... PLS> struct TwOffset {
  static constexpr int value = -(1 << 28);  // not registered
};
template <class Q, class P0, class... PLS> struct TwOffset<Q, P0, PLS...> {
  static constexpr int value =
      std::is_same<Q, P0>::value ? 0 : tw_even(plan_twiddle_len<P0>()) + TwOffset<Q, PLS...>::value;
};
template <class... PLS> struct PlanSet {
  static constexpr int total = (0 + ... + tw_even(plan_twiddle_len<PLS>()));
  template <class Q> static constexpr int offset() {
    constexpr int o = TwOffset<Q, PLS...>::value;
    static_assert(o >= 0, "plan without a twiddle table");
    return o;
  }
};
using AllPlans = PlanSet<Plan<64>, Plan<128>, Plan<256>, Plan<320>, Plan<512>, Plan<768>, Plan<1024>, Plan<1536>,
                         Plan<2048>, PlanB<768>, PlanB<1024>, PlanB<512>, PlanB<1536>>;
__device__ __align__(16) float2 g_twiddles[AllPlans::total];

// Stage (P, R) of a plan, P > 1: the factor exp(-2 pi i q k / (P R)) of input q of a
// butterfly with k = b mod P.  Laid out q-major so that one load instruction, whose
// lanes hold neighbouring k, reads neighbouring entries: for even R the pair
// (q, q + 1) = (2j, 2j + 1) of block k sits at [(j P + k) 2 + {0, 1}] (one 16-byte
// load per pair), for odd R entry q of block k at [q P + k].
template <class PL> void twiddle_host_plan(std::vector<float2>& out) {
  out.clear();
  int p = 1;
  for (int i = 0; i < PL::NRAD; ++i) {
    const int r = PL::radix(i);
    if (p > 1) {
      const size_t base = out.size();
      out.resize(base + (size_t)p * r);
      for (int k = 0; k < p; ++k)
        for (int q = 0; q < r; ++q) {
          const double a = -2.0 * M_PI * (double)q * (double)k / (double)(p * r);
          const size_t at = r % 2 == 0 ? ((size_t)(q / 2) * p + k) * 2 + q % 2 : (size_t)q * p + k;
          out[base + at] = unit_float2(std::cos(a), std::sin(a));
        }
    }
    p *= r;
  }
}
template <class PL> static void put_table(std::vector<float2>& all) {
  std::vector<float2> h;
  twiddle_host_plan<PL>(h);
  std::copy(h.begin(), h.end(), all.begin() + AllPlans::offset<PL>());
}
template <class... PLS> static void put_tables(std::vector<float2>& all, PlanSet<PLS...>*) {
  (put_table<PLS>(all), ...);
}

// Device address of g_twiddles on the current device, filled on first use.
static float2* twiddle_base() {
  int dev = 0;
  HIP_OK(hipGetDevice(&dev));
  static std::mutex mu;
  static std::map<int, float2*> base;
  std::lock_guard<std::mutex> lk(mu);
  auto it = base.find(dev);
  if (it == base.end()) {
    float2* p = nullptr;
    HIP_OK(hipGetSymbolAddress(reinterpret_cast<void**>(&p), HIP_SYMBOL(g_twiddles)));
    std::vector<float2> all(AllPlans::total, make_float2(0.f, 0.f));
    put_tables(all, static_cast<AllPlans*>(nullptr));
    HIP_OK(hipMemcpy(p, all.data(), all.size() * sizeof(float2), hipMemcpyHostToDevice));
    it = base.emplace(dev, p).first;
  }
  return it->second;
}
template <class PL> static const float2* tw_for() { return twiddle_base() + AllPlans::offset<PL>(); }

void twiddle_host(int n, std::vector<float2>& out) {
  out.clear();
#define X(S) if (n == S) { twiddle_host_plan<Plan<S>>(out); return; }
  TOTEM_FOR_SIZES(X)
#undef X
}

const float2* tw_table(int n) {
#define X(S) if (n == S) return tw_for<Plan<S>>();
  TOTEM_FOR_SIZES(X)
#undef X
  require(false, "no FFT plan for n=" + std::to_string(n), TOTEM_ERR_UNSUPPORTED);
  return nullptr;
}

template <int N, int V> void launch_pass_rows(hipStream_t st, const ItemDev* items, int n_items, const JobDev* jobs,
                                              int step) {
  using Cfg = RowCfg<N, V>;
  constexpr int K = Cfg::K;
  static_assert((N / Cfg::RW) % K == 0, "row groups per probe must be a multiple of K");
  const dim3 grid((unsigned)(n_items * (N / Cfg::RW) / K));
  hipLaunchKernelGGL((k_pass_rows<N, V>), grid, dim3(Cfg::THREADS), 0, st, items, jobs, step,
                     tw_for<typename Cfg::PL>());
}
template <int N, int V> void launch_pass_cols(hipStream_t st, const ItemDev* items, int n_items, const JobDev* jobs,
                                              int step, int groups) {
  using Cfg = ColCfg<N, V>;
  constexpr int K = TOTEM_COLS_K;
  static_assert((N / Cfg::C) % K == 0, "column groups per probe must be a multiple of K");
  if (groups <= 0 || groups > N / Cfg::C || groups % K != 0) throw std::runtime_error("columns pass: bad group count");
  const dim3 grid((unsigned)(n_items * groups / K));
  hipLaunchKernelGGL((k_pass_cols<N, V>), grid, dim3(Cfg::THREADS), 0, st, items, jobs, step, groups,
                     tw_for<typename Cfg::PL>());
}
template <int N> void launch_fft2d(hipStream_t st, float2* data, float2* dst, int batch, int sign, int epi,
                                   float sigma) {
  using R = FftRowCfg<N>;
  using Cc = FftColCfg<N>;
  const long rows = (long)batch * N;
  const dim3 rgrid((unsigned)((rows + R::RW - 1) / R::RW));
  const float2* tw = tw_for<Plan<N>>();
  if (sign < 0) hipLaunchKernelGGL((k_fft_rows<N, -1>), rgrid, dim3(R::THREADS), 0, st, data, rows, tw);
  else hipLaunchKernelGGL((k_fft_rows<N, +1>), rgrid, dim3(R::THREADS), 0, st, data, rows, tw);
  const dim3 cgrid((unsigned)(batch * (N / Cc::C)));
  if (sign < 0) hipLaunchKernelGGL((k_fft_cols<N, -1>), cgrid, dim3(Cc::THREADS), 0, st, data, dst, epi, sigma, tw);
  else hipLaunchKernelGGL((k_fft_cols<N, +1>), cgrid, dim3(Cc::THREADS), 0, st, data, dst, epi, sigma, tw);
}

template <class PL> static std::string plan_radices() {
  std::string r;
  for (int i = 0; i < PL::NRAD; ++i) r += (i ? "." : "") + std::to_string(PL::radix(i));
  return r;
}
template <int N, int V> static std::string variant_desc() {
  using R = RowCfg<N, V>;
  using C = ColCfg<N, V>;
  return "rows " + plan_radices<typename R::PL>() + " on " + std::to_string(R::T) + " lanes x " +
         std::to_string(R::RW) + " rows" + (R::SPLIT ? " (half-buffer exchanges)" : "") +
         (R::TPREFETCH ? "" : " (transmission loaded after the inverse transform)") + "; columns " +
         plan_radices<typename C::PL>() + " on " + std::to_string(C::T) + " lanes x " + std::to_string(C::C) +
         " columns" + (C::SPLIT ? " (half-buffer exchanges)" : "") + "; tiles " +
         std::to_string(WaveLay<N, V>::TR) + " x " + std::to_string(WaveLay<N, V>::TC);
}

bool size_supported(int n) {
#define X(S) if (n == S) return true;
  TOTEM_FOR_SIZES(X)
#undef X
  return false;
}

int pass_variant_count(int n) {
#define X(S) if (n == S) return pass_variants<S>();
  TOTEM_FOR_SIZES(X)
#undef X
  return 0;
}

std::string pass_variant_desc(int n, int v) {
#define X(S) \
  if (n == S) return on_variant<S>(v, [](auto vc) { return variant_desc<S, decltype(vc)::value>(); });
  TOTEM_FOR_SIZES(X)
#undef X
  return "";
}

int detector_groups(int n, int v) {
#define X(S) \
  if (n == S) return on_variant<S>(v, [](auto vc) { return S / ColCfg<S, decltype(vc)::value>::C; });
  TOTEM_FOR_SIZES(X)
#undef X
  return 0;
}

int detector_groups_max(int n) {
  int g = 0;
  for (int v = 0; v < pass_variant_count(n); ++v) g = std::max(g, detector_groups(n, v));
  return g;
}

void pass_rows(int n, int v, hipStream_t st, const ItemDev* items, int n_items, const JobDev* jobs, int step) {
#define X(S)                                                                                          \
  if (n == S) {                                                                                       \
    on_variant<S>(v, [&](auto vc) { launch_pass_rows<S, decltype(vc)::value>(st, items, n_items, jobs, step); }); \
    return;                                                                                           \
  }
  TOTEM_FOR_SIZES(X)
#undef X
}

void pass_cols(int n, int v, hipStream_t st, const ItemDev* items, int n_items, const JobDev* jobs, int step,
               int groups) {
#define X(S)                                                                                          \
  if (n == S) {                                                                                       \
    on_variant<S>(v, [&](auto vc) {                                                                   \
      launch_pass_cols<S, decltype(vc)::value>(st, items, n_items, jobs, step, groups);              \
    });                                                                                               \
    return;                                                                                           \
  }
  TOTEM_FOR_SIZES(X)
#undef X
}

void reduce_det(hipStream_t st, const ItemDev* items, const int* list, int n, const JobDev* jobs, int groups,
                int lmax1, int amax) {
  const int total = n * lmax1 * amax;
  if (total <= 0) return;
  hipLaunchKernelGGL(k_reduce_det, dim3((total + 255) / 256), dim3(256), 0, st, items, list, n, jobs, groups, lmax1,
                     amax);
}

int pass_cols_width(int n, int v) {
#define X(S) \
  if (n == S) return on_variant<S>(v, [](auto vc) { return ColCfg<S, decltype(vc)::value>::C; });
  TOTEM_FOR_SIZES(X)
#undef X
  return 0;
}

void prop_column_groups(hipStream_t st, const float2* prop, float2* out, int n, int C) {
  const long total = (long)n * n;
  hipLaunchKernelGGL(k_prop_cols, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, st, prop, out, n, C);
}

void column_flags(hipStream_t st, const float2* prop, int* flags, int n) {
  const unsigned nb = (unsigned)((n + 255) / 256);
  hipLaunchKernelGGL(k_zero_i32, dim3(nb), dim3(256), 0, st, flags, n);
  hipLaunchKernelGGL(k_column_flags, dim3(nb, kFlagChunks), dim3(256), 0, st, prop, flags, n);
}

void fill(hipStream_t st, float2* p, long n, float re, float im) {
  if (n > 0) hipLaunchKernelGGL(k_fill_c64, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, p, n, re, im);
}
void fill(hipStream_t st, double* p, long n, double v) {
  if (n > 0) hipLaunchKernelGGL(k_fill_f64, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, p, n, v);
}

void fft2d(int n, hipStream_t st, float2* data, float2* dst, int batch, int sign, int epi, float sigma) {
#define X(S) if (n == S) { launch_fft2d<S>(st, data, dst, batch, sign, epi, sigma); return; }
  TOTEM_FOR_SIZES(X)
#undef X
}

bool hermitian_pairs(const double* kx1, const double* ky1, int np) {
  if (np % 32 != 0) return false;
  for (int r = 1; r <= np / 2 - 16; ++r)
    if (ky1[np - r] != -ky1[r] || kx1[np - r] != -kx1[r]) return false;
  return true;
}

int phase_columns(int np, bool sym) { return sym ? np / 2 + 16 : np; }

void phase_tables(hipStream_t st, const double* tx, const double* ty, const double* wt, int n_terms,
                  const double* kx1, const double* ky1, int np, bool sym, float2* ex, float2* ey) {
  if (n_terms <= 0) return;
  const int nc = phase_columns(np, sym);
  hipLaunchKernelGGL(k_phase, dim3(n_terms, (nc + 255) / 256), dim3(256), 0, st, tx, ty, wt, n_terms, kx1, ky1, np,
                     nc, ex, ey);
}

#ifndef TOTEM_ACC_WAVES
#define TOTEM_ACC_WAVES 8  // waves sharing one potential tile
#endif

void accumulate(hipStream_t st, const float2* ex, const float2* ey, const PotSeg* segs, int n_segs,
                const PotClass* classes, const double* s2, const double* aper_true, const double* corr,
                double parameter, int np, bool sym, float2* vshift) {
#define X(S) static_assert(S % 16 == 0, "potential tiles need grid sides that are multiples of 16");
  TOTEM_FOR_SIZES(X)
#undef X
  const int dt = sym ? np / 32 + 1 : np / 16;
  constexpr int W = TOTEM_ACC_WAVES;
  hipLaunchKernelGGL((k_accumulate_mfma16<W>), dim3((unsigned)(dt * dt)), dim3(64 * W), 0, st, ex, ey, segs, n_segs,
                     classes, s2, aper_true, corr, parameter, np, sym ? 1 : 0, vshift);
}

}  // namespace totem
