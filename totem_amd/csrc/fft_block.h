// Workgroup-level Stockham FFT for CDNA4 (gfx950).
//
// A transform of length N is owned by T cooperating lanes; lane t keeps the
// E = N/T elements  t + T*v  (v = 0..E-1) in VGPRs.  Every Stockham stage of
// radix R runs E/R butterflies per lane on those registers and, except for the
// last stage, exchanges through LDS.  Because the register set {t + T*v} is the
// same at entry and exit, a forward transform can feed an inverse transform (or
// vice versa) without an LDS round trip, which is what lets the multislice
// passes fuse  inverse-FFT -> pointwise -> forward-FFT  in one kernel.
//
// Twiddles come from an FP32 table tw[m] = exp(-2 pi i m / N) that the host
// generates in FP64; the inverse direction conjugates them.
#pragma once
#include <hip/hip_runtime.h>

namespace totem {

// Global-memory accessors.  Pointers read from device-side structs are generic to
// the compiler; flat loads count against both the vector-memory and the LDS
// counters, so a wait for one serialises the other.  These assert the global
// address space (8-byte accesses; the float2 is reassembled in registers).
typedef __attribute__((address_space(1))) const unsigned long long* g64c_t;
typedef __attribute__((address_space(1))) unsigned long long* g64_t;
__device__ __forceinline__ float2 gld(const float2* p) {
  const unsigned long long v = *(g64c_t)(p);
  return make_float2(__uint_as_float((unsigned)(v & 0xffffffffull)), __uint_as_float((unsigned)(v >> 32)));
}
__device__ __forceinline__ void gst(float2* p, float2 v) {
  *(g64_t)(p) = ((unsigned long long)__float_as_uint(v.y) << 32) | (unsigned long long)__float_as_uint(v.x);
}
__device__ __forceinline__ double gldd(const double* p) { return *(__attribute__((address_space(1))) const double*)(p); }

// Raw buffer access: a 128-bit resource (base, byte range) in SGPRs plus a 32-bit
// per-lane offset and a per-access SGPR offset.  Column-strided accesses then need
// one VGPR per lane instead of one 64-bit address per element, and accesses
// outside the range read 0 / are dropped.
typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ __amdgpu_buffer_rsrc_t buf_rsrc(const void* base, unsigned bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), (short)0, (int)bytes, 0x00020000);
}
__device__ __forceinline__ float2 bld(__amdgpu_buffer_rsrc_t r, int voff, int soff) {
  const u32x2 v = __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, 0);
  return make_float2(__uint_as_float(v.x), __uint_as_float(v.y));
}
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
__device__ __forceinline__ float4 bld4(__amdgpu_buffer_rsrc_t r, int voff, int soff) {
  const u32x4 v = __builtin_amdgcn_raw_buffer_load_b128(r, voff, soff, 0);
  return make_float4(__uint_as_float(v.x), __uint_as_float(v.y), __uint_as_float(v.z), __uint_as_float(v.w));
}
__device__ __forceinline__ void bst(__amdgpu_buffer_rsrc_t r, int voff, int soff, float2 a) {
  const u32x2 v = {__float_as_uint(a.x), __float_as_uint(a.y)};
  __builtin_amdgcn_raw_buffer_store_b64(v, r, voff, soff, 0);
}

__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ float2 cscale(float2 a, float s) { return make_float2(a.x * s, a.y * s); }
// a * (SIGN * i)
template <int SIGN> __device__ __forceinline__ float2 mul_si(float2 a) {
  return SIGN < 0 ? make_float2(a.y, -a.x) : make_float2(-a.y, a.x);
}
// Butterfly constants as (hi, lo) float pairs: hi is the nearest float, lo the
// float nearest the remainder.  Multiplying by hi alone scales |w|^2 by 1 - O(1e-8)
// with the same sign every time (e.g. 2 * float(sqrt(1/2))^2 = 1 - 3.4e-8); the
// multislice loop applies each butterfly hundreds of times to similar data, so
// that bias would add up coherently (a 1.8e-5 intensity loss over 80 slices at
// 1024^2).  With the lo term folded in by FMA the products round once, unbiased.
struct K2 {
  float h, l;
};
constexpr K2 kR = {7.071067691e-01f, 1.210161749e-08f};     // sqrt(1/2)
constexpr K2 kC16 = {9.238795042e-01f, 2.830748969e-08f};   // cos(pi/8)
constexpr K2 kS16 = {3.826834261e-01f, 6.223350724e-09f};   // sin(pi/8)
constexpr K2 kS3 = {8.660253882e-01f, 1.554362505e-08f};    // sqrt(3)/2
constexpr K2 kC51 = {3.090170026e-01f, -8.207602598e-09f};  // cos(2 pi/5)
constexpr K2 kC52 = {-8.090170026e-01f, 8.207602598e-09f};  // cos(4 pi/5)
constexpr K2 kS51 = {9.510565400e-01f, -2.371720598e-08f};  // sin(2 pi/5)
constexpr K2 kS52 = {5.877852440e-01f, 8.304436250e-09f};   // sin(4 pi/5)

__device__ __forceinline__ float kmul(float v, K2 k) { return fmaf(v, k.h, v * k.l); }
__device__ __forceinline__ K2 kneg(K2 k) { return K2{-k.h, -k.l}; }

// a * (c + SIGN*i*s) with (hi, lo) constants
template <int SIGN> __device__ __forceinline__ float2 mul_cs(float2 a, K2 c, K2 s) {
  const K2 ss = SIGN < 0 ? kneg(s) : s;
  const float re = fmaf(a.x, c.h, fmaf(-a.y, ss.h, fmaf(a.x, c.l, -a.y * ss.l)));
  const float im = fmaf(a.x, ss.h, fmaf(a.y, c.h, fmaf(a.x, ss.l, a.y * c.l)));
  return make_float2(re, im);
}

// ------------------------------------------------------------------ small DFTs
// y_k = sum_n a_n exp(SIGN * 2 pi i n k / R), in place, natural order.

template <int SIGN> __device__ __forceinline__ void dft2(float2* a) {
  float2 t = a[0];
  a[0] = cadd(t, a[1]);
  a[1] = csub(t, a[1]);
}

template <int SIGN> __device__ __forceinline__ void dft3(float2* a) {
  float2 t1 = cadd(a[1], a[2]);
  float2 t2 = csub(a[1], a[2]);
  float2 y0 = cadd(a[0], t1);
  float2 m = make_float2(a[0].x - 0.5f * t1.x, a[0].y - 0.5f * t1.y);
  float2 n = mul_si<SIGN>(make_float2(kmul(t2.x, kS3), kmul(t2.y, kS3)));
  a[0] = y0;
  a[1] = cadd(m, n);
  a[2] = csub(m, n);
}

template <int SIGN> __device__ __forceinline__ void dft4(float2* a) {
  float2 t0 = cadd(a[0], a[2]);
  float2 t1 = csub(a[0], a[2]);
  float2 t2 = cadd(a[1], a[3]);
  float2 t3 = mul_si<SIGN>(csub(a[1], a[3]));
  a[0] = cadd(t0, t2);
  a[2] = csub(t0, t2);
  a[1] = cadd(t1, t3);
  a[3] = csub(t1, t3);
}

// c1*u + c2*v with (hi, lo) constants, one rounding per FMA
__device__ __forceinline__ float kdot(float u, K2 c1, float v, K2 c2, float base) {
  return fmaf(u, c1.h, fmaf(v, c2.h, base + (u * c1.l + v * c2.l)));
}

template <int SIGN> __device__ __forceinline__ void dft5(float2* a) {
  float2 t1 = cadd(a[1], a[4]), t2 = cadd(a[2], a[3]);
  float2 t3 = csub(a[1], a[4]), t4 = csub(a[2], a[3]);
  float2 y0 = cadd(a[0], cadd(t1, t2));
  float2 b1 = make_float2(kdot(t1.x, kC51, t2.x, kC52, a[0].x), kdot(t1.y, kC51, t2.y, kC52, a[0].y));
  float2 b2 = make_float2(kdot(t1.x, kC52, t2.x, kC51, a[0].x), kdot(t1.y, kC52, t2.y, kC51, a[0].y));
  float2 d1 = make_float2(kdot(t3.x, kS51, t4.x, kS52, 0.f), kdot(t3.y, kS51, t4.y, kS52, 0.f));
  float2 d2 = make_float2(kdot(t3.x, kS52, t4.x, kneg(kS51), 0.f), kdot(t3.y, kS52, t4.y, kneg(kS51), 0.f));
  float2 e1 = mul_si<SIGN>(d1), e2 = mul_si<SIGN>(d2);
  a[0] = y0;
  a[1] = cadd(b1, e1);
  a[4] = csub(b1, e1);
  a[2] = cadd(b2, e2);
  a[3] = csub(b2, e2);
}

// a * (1 + SIGN*i)/sqrt2 and a * (-1 + SIGN*i)/sqrt2
template <int SIGN> __device__ __forceinline__ float2 mul_w8(float2 a) {
  const float u = SIGN < 0 ? a.x + a.y : a.x - a.y;
  const float v = SIGN < 0 ? a.y - a.x : a.x + a.y;
  return make_float2(kmul(u, kR), kmul(v, kR));
}
template <int SIGN> __device__ __forceinline__ float2 mul_w83(float2 a) {
  const float u = SIGN < 0 ? a.y - a.x : -a.x - a.y;
  const float v = SIGN < 0 ? -a.x - a.y : a.x - a.y;
  return make_float2(kmul(u, kR), kmul(v, kR));
}

template <int SIGN> __device__ __forceinline__ void dft8(float2* a) {
  float2 e[4] = {a[0], a[2], a[4], a[6]};
  float2 o[4] = {a[1], a[3], a[5], a[7]};
  dft4<SIGN>(e);
  dft4<SIGN>(o);
  o[1] = mul_w8<SIGN>(o[1]);
  o[2] = mul_si<SIGN>(o[2]);
  o[3] = mul_w83<SIGN>(o[3]);
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    a[k] = cadd(e[k], o[k]);
    a[k + 4] = csub(e[k], o[k]);
  }
}

// 16 = 4 x 4: DFT4 over n1 of a[4 n1 + n2] for each n2, W16^(n2 k1), DFT4 over n2
// (184 FP32 operations against 196 for two 8-point halves).
template <int SIGN> __device__ __forceinline__ void dft16(float2* a) {
  float2 y[4][4];
#pragma unroll
  for (int n2 = 0; n2 < 4; ++n2) {
    float2 t[4] = {a[n2], a[4 + n2], a[8 + n2], a[12 + n2]};
    dft4<SIGN>(t);
#pragma unroll
    for (int k1 = 0; k1 < 4; ++k1) y[n2][k1] = t[k1];
  }
  y[1][1] = mul_cs<SIGN>(y[1][1], kC16, kS16);               // W^1
  y[1][2] = mul_w8<SIGN>(y[1][2]);                           // W^2
  y[1][3] = mul_cs<SIGN>(y[1][3], kS16, kC16);               // W^3
  y[2][1] = mul_w8<SIGN>(y[2][1]);                           // W^2
  y[2][2] = mul_si<SIGN>(y[2][2]);                           // W^4
  y[2][3] = mul_w83<SIGN>(y[2][3]);                          // W^6
  y[3][1] = mul_cs<SIGN>(y[3][1], kS16, kC16);               // W^3
  y[3][2] = mul_w83<SIGN>(y[3][2]);                          // W^6
  y[3][3] = mul_cs<SIGN>(y[3][3], kneg(kC16), kneg(kS16));   // W^9
#pragma unroll
  for (int k1 = 0; k1 < 4; ++k1) {
    float2 t[4] = {y[0][k1], y[1][k1], y[2][k1], y[3][k1]};
    dft4<SIGN>(t);
#pragma unroll
    for (int k2 = 0; k2 < 4; ++k2) a[k1 + 4 * k2] = t[k2];
  }
}

template <int R, int SIGN> __device__ __forceinline__ void dft(float2* a) {
  if constexpr (R == 2) dft2<SIGN>(a);
  else if constexpr (R == 3) dft3<SIGN>(a);
  else if constexpr (R == 4) dft4<SIGN>(a);
  else if constexpr (R == 5) dft5<SIGN>(a);
  else if constexpr (R == 8) dft8<SIGN>(a);
  else if constexpr (R == 16) dft16<SIGN>(a);
  else static_assert(R == 2, "unsupported radix");
}

// ------------------------------------------------------------------ Stockham stages
//
// Stage with radix R after radices of product P (P = 1 for the first stage):
// butterfly b in [0, N/R) reads x[b + q*N/R], multiplies input q by
// W^(q*k) with k = b mod P and W = exp(SIGN 2 pi i / (P R)), runs DFT_R and
// writes y[(b - k)*R + k + r*P].  Lane t runs butterflies b = t + T*u.
//
// Twiddles: the plan's table holds, for every stage after the first, the factors
// exp(-2 pi i q k / (P R)) for q < R and k < P, q-major (pairs of q for even R), so
// that one load instruction of a wave, whose lanes hold neighbouring k, reads
// neighbouring entries (16-byte loads for even R).
//
// LDS: with PP (ping-pong) two buffers are used alternately, so one barrier per
// exchange suffices (the next exchange writes the other buffer, and the one after
// that is ordered by the intervening barrier); without PP one buffer and two
// barriers per exchange (half the LDS, so more workgroups per CU).  Addresses are
// a per-lane base plus compile-time offsets, which the ds_read/ds_write immediate
// offsets absorb.

// Exchange barrier: the whole workgroup, or only the wave when the layout keeps
// every transform inside one wave (LDS operations of a wave complete in order,
// so a compiler-level fence is enough there).
template <bool WAVE> __device__ __forceinline__ void lds_sync() {
  if constexpr (WAVE) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  } else {
    __syncthreads();
  }
}

// SPLIT: every exchange runs in two rounds through an LDS buffer of N/2 elements
// per transform (elements [0, N/2) first, then [N/2, N)), for lengths whose full
// exchange buffer would not fit.  A lane reads elements t + T*v, which lie in the
// lower half exactly for v < E/2; the writes of a round are predicated per element.
template <int N, int T, int SIGN, class Lay, bool PP, int P, int PING, bool SPLIT, int... RS> struct Stages;

// Twiddles of the stage (P, R) for the U butterflies of lane t: block k = b mod P
// of the stage table, R entries (index u*R + q; entry 0 is 1).  A stage's loads
// are issued by the previous stage before its LDS exchange, so their latency
// overlaps the exchange instead of stalling the butterflies after the barrier.
// TOTEM_TW_LATE issues them after the previous stage's butterflies (its own
// twiddles are dead by then, so the two sets share registers) instead of before.
#ifndef TOTEM_TW_LATE
#define TOTEM_TW_LATE 1
#endif
template <int T, int P, int R, int U> struct StageTw {
  static constexpr int NW = (P > 1) ? U * R : 1;
  float2 w[NW];
  __device__ __forceinline__ void load(const float2* __restrict__ tw, int t) {
    if constexpr (P > 1) {
      // buffer loads: the lane's k in the vector offset, the q part in the scalar
      // offset (no 64-bit address per q)
      const auto rs = buf_rsrc(tw, (unsigned)(P * R * sizeof(float2)));
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int k = (t + T * u) % P;
        if constexpr (R % 2 == 0) {
#pragma unroll
          for (int j = 0; j < R / 2; ++j) {
            const float4 v = bld4(rs, k * 16, j * P * 16);
            w[u * R + 2 * j] = make_float2(v.x, v.y);
            w[u * R + 2 * j + 1] = make_float2(v.z, v.w);
          }
        } else {
#pragma unroll
          for (int q = 1; q < R; ++q) w[u * R + q] = bld(rs, k * 8, q * P * 8);
        }
      }
    }
  }
};

template <int N, int T, int SIGN, class Lay, bool PP, int P, int PING, bool SPLIT>
struct Stages<N, T, SIGN, Lay, PP, P, PING, SPLIT> {
  using Tw = StageTw<T, 1, 1, 1>;
  template <class W>
  static __device__ __forceinline__ void run(float2*, float2*, float2*, const Lay&, int, const float2*, const W&) {}
};

template <int N, int T, int SIGN, class Lay, bool PP, int P, int PING, bool SPLIT, int R, int... RS>
struct Stages<N, T, SIGN, Lay, PP, P, PING, SPLIT, R, RS...> {
  static constexpr int E = N / T;
  static constexpr int U = E / R;
  static_assert(E % R == 0, "radix must divide elements per lane");
  static_assert(T % 16 == 0, "lane count must be a multiple of 16");
  static_assert(!SPLIT || (!PP && E % 2 == 0 && (N / 2) % 16 == 0), "split exchange: one buffer, even E");
  static constexpr bool LAST = sizeof...(RS) == 0;
  using Tw = StageTw<T, P, R, U>;
  using Next = Stages<N, T, SIGN, Lay, PP, P * R, PP ? 1 - PING : 0, SPLIT, RS...>;

  // `tw` points at this stage's table; `cur` holds this stage's twiddles.
  static __device__ __forceinline__ void run(float2* x, float2* lds0, float2* lds1, const Lay& lay, int t,
                                             const float2* __restrict__ tw, const Tw& cur) {
    float2* buf = PING ? lds1 : lds0;
    const float2* tw_next = tw + (P > 1 ? P * R : 0);
    typename Next::Tw nxt;
    if constexpr (!LAST && !TOTEM_TW_LATE) nxt.load(tw_next, t);
    float2 out[E];
    int jw[U];  // exchange position of each butterfly's output 0 (split exchanges)
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int b = t + T * u;
      const int k = (P == 1) ? 0 : (b % P);
      float2 a[R];
#pragma unroll
      for (int q = 0; q < R; ++q) a[q] = x[u + q * U];
      if constexpr (P > 1) {
#pragma unroll
        for (int q = 1; q < R; ++q) {
          float2 w = cur.w[u * R + q];
          if (SIGN > 0) w.y = -w.y;
          a[q] = cmul(a[q], w);
        }
      }
      dft<R, SIGN>(a);
      if constexpr (LAST || SPLIT) {
        jw[u] = (b - k) * R + k;
#pragma unroll
        for (int r = 0; r < R; ++r) out[u + r * U] = a[r];
      } else {
        const int j = (b - k) * R + k;
        if constexpr (Lay::template linear_write<P, R>()) {
          float2* wp = buf + lay(j);
#pragma unroll
          for (int r = 0; r < R; ++r) wp[Lay::off(r * P)] = a[r];
        } else {
#pragma unroll
          for (int r = 0; r < R; ++r) buf[lay(j + r * P)] = a[r];
        }
      }
    }
    if constexpr (LAST) {
#pragma unroll
      for (int v = 0; v < E; ++v) x[v] = out[v];
    } else if constexpr (SPLIT) {
      if constexpr (TOTEM_TW_LATE) nxt.load(tw_next, t);
      constexpr int H = N / 2;
      // a butterfly's outputs jw + r P (r < R) lie in one aligned block of P R
      // elements, and P R divides N/2: each butterfly writes in one round only
      static_assert(H % (P * R) == 0, "split exchange: butterfly blocks must not straddle N/2");
#pragma unroll
      for (int h = 0; h < 2; ++h) {
#pragma unroll
        for (int u = 0; u < U; ++u) {
          if ((jw[u] >= H) == (h == 1)) {
            const int j = jw[u] - h * H;
            if constexpr (Lay::template linear_write<P, R>()) {
              float2* wp = buf + lay(j);
#pragma unroll
              for (int r = 0; r < R; ++r) wp[Lay::off(r * P)] = out[u + r * U];
            } else {
#pragma unroll
              for (int r = 0; r < R; ++r) buf[lay(j + r * P)] = out[u + r * U];
            }
          }
        }
        lds_sync<Lay::kWave>();
        const float2* rp = buf + lay(t);
#pragma unroll
        for (int v = h * (E / 2); v < (h + 1) * (E / 2); ++v) x[v] = rp[Lay::off(T * v - h * H)];
        lds_sync<Lay::kWave>();
      }
      Next::run(x, lds0, lds1, lay, t, tw_next, nxt);
    } else {
      if constexpr (TOTEM_TW_LATE) nxt.load(tw_next, t);
      lds_sync<Lay::kWave>();
      const float2* rp = buf + lay(t);
#pragma unroll
      for (int v = 0; v < E; ++v) x[v] = rp[Lay::off(T * v)];
      if constexpr (!PP) lds_sync<Lay::kWave>();
      Next::run(x, lds0, lds1, lay, t, tw_next, nxt);
    }
  }
};

// Radix plans: lane count T and radix list per supported length.
template <int N_, int T_, int... RS> struct PlanDef {
  static constexpr int N = N_;
  static constexpr int T = T_;
  static constexpr int NRAD = sizeof...(RS);
  static constexpr int EXCH = NRAD - 1;  // LDS exchanges per transform
  template <int S, class L, bool PP, int START, bool SPLIT = false>
  using St = Stages<N, T, S, L, PP, 1, START, SPLIT, RS...>;
  static constexpr int radix(int i) {
    constexpr int r[] = {RS...};
    return r[i];
  }
};
template <int N> struct Plan;
template <> struct Plan<64> : PlanDef<64, 16, 4, 4, 4> {};
template <> struct Plan<128> : PlanDef<128, 16, 8, 4, 4> {};
template <> struct Plan<256> : PlanDef<256, 16, 16, 16> {};
template <> struct Plan<320> : PlanDef<320, 16, 4, 4, 4, 5> {};
template <> struct Plan<512> : PlanDef<512, 32, 16, 16, 2> {};
template <> struct Plan<768> : PlanDef<768, 32, 8, 3, 8, 4> {};
template <> struct Plan<1024> : PlanDef<1024, 64, 16, 16, 4> {};
template <> struct Plan<1536> : PlanDef<1536, 64, 8, 3, 8, 8> {};
template <> struct Plan<2048> : PlanDef<2048, 128, 16, 16, 8> {};
// Alternative plans a pass variant may use instead of Plan<N>.
template <int N> struct PlanB;
template <> struct PlanB<768> : PlanDef<768, 64, 4, 4, 4, 4, 3> {};  // 12 elements per lane
template <> struct PlanB<1024> : PlanDef<1024, 128, 8, 8, 8, 2> {};  // 8 elements per lane
template <> struct PlanB<512> : PlanDef<512, 64, 8, 8, 8> {};          // 8 elements per lane
template <> struct PlanB<1536> : PlanDef<1536, 128, 4, 4, 4, 4, 2, 3> {};  // 12 elements per lane

// Number of twiddle entries of a plan's table.
template <class PL> __host__ __device__ constexpr int plan_twiddle_len() {
  int p = 1, len = 0;
  for (int i = 0; i < PL::NRAD; ++i) {
    if (p > 1) len += p * PL::radix(i);
    p *= PL::radix(i);
  }
  return len;
}
template <int N> __host__ __device__ constexpr int twiddle_len() { return plan_twiddle_len<Plan<N>>(); }

// x holds elements t + T*v of one transform; after the call it holds the
// corresponding outputs.  START selects the first LDS buffer: a transform that
// follows another must start on the buffer the previous one did not use last.
template <int N, int SIGN, bool PP, int START = 0, bool SPLIT = false, class PL = Plan<N>, class Lay>
__device__ __forceinline__ void block_fft(float2* x, float2* lds0, float2* lds1, const Lay& lay, int t,
                                          const float2* tw) {
  static_assert(PL::N == N, "plan length");
  using St = typename PL::template St<SIGN, Lay, PP, START, SPLIT>;
  typename St::Tw none;  // the first stage has no twiddles
  St::run(x, lds0, lds1, lay, t, tw, none);
}
// START for a transform that follows one started on buffer `prev`.
template <int N, bool PP, class PL = Plan<N>> constexpr int next_start(int prev) {
  return PP ? (prev + PL::EXCH) % 2 : 0;
}

// LDS layouts.  Row transforms: one padding slot per 16 elements so that the
// first stage's stride-R writes spread over the banks.  Column transforms:
// C interleaved columns, element e of column c at e*C + c (conflict-free
// because the C lanes of a group touch consecutive slots).
template <bool WAVE> struct RowLayT {
  static constexpr bool kWave = WAVE;  // each transform's lanes lie in one wave
  int base;
  __device__ __forceinline__ int operator()(int e) const { return base + e + (e >> 4); }
  // slot(e + k) - slot(e) for k a multiple of 16
  static __device__ __forceinline__ constexpr int off(int k) { return k + (k >> 4); }
  template <int P, int R> static constexpr bool linear_write() { return P % 16 == 0 || (P == 1 && R % 16 == 0); }
};
// Column transforms also pad one slot-row per 16 elements: without it the first
// stage's stride-R writes of lanes t, t+1, ... land R*C slots apart, i.e. on the
// same banks.
using RowLay = RowLayT<false>;
template <int C> struct ColLay {
  static constexpr bool kWave = false;
  int c;
  __device__ __forceinline__ int operator()(int e) const { return (e + (e >> 4)) * C + c; }
  static __device__ __forceinline__ constexpr int off(int k) { return (k + (k >> 4)) * C; }
  template <int P, int R> static constexpr bool linear_write() { return P % 16 == 0 || (P == 1 && R % 16 == 0); }
};
template <int N> __host__ __device__ constexpr int row_lds_len() { return N + N / 16; }
template <int N, int C> __host__ __device__ constexpr int col_lds_len() { return (N + N / 16) * C; }

}  // namespace totem
