// Specimen slicing and frozen-phonon displacement on the GPU.
//
// The reference displaces atoms on the host (randshiftpos.m:1-44: Gaussian
// per axis, redrawn while an atom leaves the slab or crosses a mid-slice face)
// and assigns them to slices on the host (CalAllEquilent20210223_lobato_peng.m:
// 23-57, 213-226; CalAllEquilent20210305_peng_corr.m:25-70).  Here both
// run per phonon configuration on the device, on terms that carry their atom
// and form-factor class: displaced positions -> (slice, class) keys -> stable
// radix sort -> term arrays grouped by slice and class, the layout the
// potential kernels read.  The sort is a stable LSD radix sort written here (8-bit
// digits: per-tile digit counts, an exclusive scan, an in-order scatter).  Random
// numbers: Philox4x32-10 (Salmon et al., SC'11),
// counter = (atom, configuration, attempt, lane), key = seed; normals by
// Box-Muller in double precision.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>

#include "launch.h"

namespace totem {

namespace {

__device__ __forceinline__ void philox_round(uint32_t c[4], const uint32_t k[2]) {
  const uint32_t m0 = 0xD2511F53u, m1 = 0xCD9E8D57u;
  const uint32_t hi0 = __umulhi(m0, c[0]), lo0 = m0 * c[0];
  const uint32_t hi1 = __umulhi(m1, c[2]), lo1 = m1 * c[2];
  const uint32_t o0 = hi1 ^ c[1] ^ k[0], o2 = hi0 ^ c[3] ^ k[1];
  c[0] = o0;
  c[1] = lo1;
  c[2] = o2;
  c[3] = lo0;
}

__device__ __forceinline__ void philox4x32_10(uint32_t c[4], uint64_t seed) {
  uint32_t k[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    philox_round(c, k);
    k[0] += 0x9E3779B9u;
    k[1] += 0xBB67AE85u;
  }
}

// uniform in (0, 1) from 53 random bits
__device__ __forceinline__ double u53(uint32_t a, uint32_t b) {
  return ((double)(a >> 5) * 67108864.0 + (double)(b >> 6) + 0.5) * (1.0 / 9007199254740992.0);
}

__device__ __forceinline__ void normals3(uint64_t seed, int atom, int config, int attempt, double n[3]) {
  uint32_t c[4] = {(uint32_t)atom, (uint32_t)config, (uint32_t)attempt, 0u};
  philox4x32_10(c, seed);
  const double r = sqrt(-2.0 * log(u53(c[0], c[1]))), th = 2.0 * M_PI * u53(c[2], c[3]);
  n[0] = r * cos(th);
  n[1] = r * sin(th);
  uint32_t d[4] = {(uint32_t)atom, (uint32_t)config, (uint32_t)attempt, 1u};
  philox4x32_10(d, seed);
  const double r2 = sqrt(-2.0 * log(u53(d[0], d[1]))), th2 = 2.0 * M_PI * u53(d[2], d[3]);
  n[2] = r2 * cos(th2);
}

constexpr int kMaxAttempts = 64;

__global__ void k_displace(int n, const double* __restrict__ xyz, const double* __restrict__ rms, uint64_t seed,
                           int config, double lo, double hi, const double* __restrict__ faces, int n_faces,
                           double* __restrict__ out) {
  const int a = blockIdx.x * blockDim.x + threadIdx.x;
  if (a >= n) return;
  const double x = xyz[3 * a], y = xyz[3 * a + 1], z = xyz[3 * a + 2];
  double px = x, py = y, pz = z;
  const double s = rms ? rms[a] : 0.0;
  if (s > 0.0) {
    for (int k = 0; k < kMaxAttempts; ++k) {
      double g[3];
      normals3(seed, a, config, k, g);
      const double cx = x + s * g[0], cy = y + s * g[1], cz = z + s * g[2];
      bool ok = cz >= lo && cz < hi;
      for (int f = 0; f < n_faces && ok; ++f) ok = (z < faces[f]) == (cz < faces[f]);
      if (ok) {
        px = cx;
        py = cy;
        pz = cz;
        break;
      }
    }
  }
  out[3 * a] = px;
  out[3 * a + 1] = py;
  out[3 * a + 2] = pz;
}

// Up to K key slots per term.  Slice i holds st[i] <= z < st[i] + dz with
// st[i] = origin + dz * i, the host's test, so an atom within rounding of a
// face can sit in two adjacent slices (as on the host) or in none.  Key =
// slice * Ct + class; an empty slot holds S * Ct and sorts last.  Per slot the
// weight (occupancy times the slice's share) goes to wslot.
//
// mode kSplitCorr (correction terms, CalAllEquilent20210305_peng_corr.m:41-58):
// with split set, a term within discon of its slice's lower (upper) face gives
// half its weight to the slice below (above), the outermost faces excepted.
// mode kErf5 / kErf10 (Peng rows, CalAllEquilent20210223_lobato_peng.m:23-57,
// 213-226) with m > 0: every slice whose range widened by m dz on both sides
// holds the atom gets it; where a face is nearer than m dz the Gaussian term q
// is weighted by |erf(2 pi z2 / sqrt(b_q)) - erf(2 pi z1 / sqrt(b_q))| / 2 and
// goes to the one-term class sub_base[class] + q.
constexpr int kPlain = 0, kErf5 = 1, kErf10 = 2, kSplitCorr = 3;

__device__ __forceinline__ double erf_share(double b, double z1, double z2) {
  const double sb = sqrt(b);
  return fabs(0.5 * erf(2.0 * M_PI * z2 / sb) - 0.5 * erf(2.0 * M_PI * z1 / sb));
}

__global__ void k_term_keys(int n, int K, const int* __restrict__ term_atom, const int* __restrict__ term_cls,
                            const double* __restrict__ term_w, const int* __restrict__ term_mode,
                            const PotClass* __restrict__ classes, const int* __restrict__ sub_base,
                            const double* __restrict__ xyz, int S, int Ct, double origin, double dz, double m,
                            int split, double discon, unsigned* __restrict__ keys, unsigned* __restrict__ vals,
                            double* __restrict__ wslot, int* __restrict__ counts) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double z = xyz[3 * term_atom[i] + 2];
  const int cls = term_cls[i], mode = term_mode[i];
  const double w = term_w[i];
  const unsigned none = (unsigned)S * (unsigned)Ct;
  const size_t base = (size_t)i * K;
  int slot = 0;
  auto emit = [&](int s, int c, double wt) {
    if (slot >= K) return;
    const unsigned key = (unsigned)s * (unsigned)Ct + (unsigned)c;
    keys[base + slot] = key;
    vals[base + slot] = (unsigned)(base + slot);
    wslot[base + slot] = wt;
    atomicAdd(&counts[key], 1);
    ++slot;
  };
  const double f = floor((z - origin) / dz);
  const bool erf_mode = (mode == kErf5 || mode == kErf10) && m > 0.0;
  if (erf_mode) {
    if (f > -m - 4.0 && f < (double)S + m + 4.0) {  // false for NaN
      const int nq = mode == kErf10 ? 10 : 5;
      const double md = m * dz;
      const int s_lo = max((int)floor(f - m) - 2, 0), s_hi = min((int)ceil(f + m) + 2, S - 1);
      for (int s = s_lo; s <= s_hi; ++s) {
        const double st = origin + dz * (double)s;
        if (!(z >= st - md && z < st + md + dz)) continue;
        double z1 = z - (st + dz), z2 = z - st;
        if (s == 0) z2 = INFINITY;
        if (s == S - 1) z1 = -INFINITY;
        if (fabs(z2) >= md) z2 = INFINITY;
        if (fabs(z1) >= md) z1 = -INFINITY;
        if (z2 == INFINITY && z1 == -INFINITY) {
          emit(s, cls, w);  // every share is 1: the atom's own class
          continue;
        }
        const PotClass& k = classes[cls];
        for (int q = 0; q < nq; ++q) {
          const double b = q < 5 ? k.p1[2 * q + 1] : k.p2[2 * (q - 5) + 1];
          emit(s, sub_base[cls] + q, w * erf_share(b, z1, z2));
        }
      }
    }
  } else if (f > -4.0 && f < (double)S + 4.0) {
    const int s0 = (int)f;
    int hit = 0;
    for (int s = max(s0 - 2, 0); s <= min(s0 + 2, S - 1) && hit < 2; ++s) {
      const double st = origin + dz * (double)s;
      if (!(z >= st && z < st + dz)) continue;
      ++hit;
      if (mode == kSplitCorr && split) {
        if (z - st < discon && s > 0) {
          emit(s, cls, w * 0.5);
          emit(s - 1, cls, w * 0.5);
          continue;
        }
        if (z - st - dz > -discon && s < S - 1) {
          emit(s, cls, w * 0.5);
          emit(s + 1, cls, w * 0.5);
          continue;
        }
      }
      emit(s, cls, w);
    }
  }
  for (; slot < K; ++slot) {
    keys[base + slot] = none;
    vals[base + slot] = (unsigned)(base + slot);
    wslot[base + slot] = 0.0;
  }
}

__global__ void k_fill_i32(int n, int v, int* __restrict__ p) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// ------------------------------------------------------------ stable radix sort
//
// Tiles of kSortTile = 1024 elements, one wave each; element e * 64 + lane of a
// tile is lane `lane`'s e-th, so a wave walks its tile in index order.  Per pass
// (digit d = bits [shift, shift + 8) of the key):
//   hist[d * G + tile] = elements of the tile with digit d
//   exclusive scan of hist in that (digit-major, tile-minor) order
//   scatter: an element goes to hist[d * G + tile] + (elements of its tile before
//            it with the same digit)
// which keeps equal digits in index order: the sort is stable, and LSD passes over
// the digits give the order of the whole key with ties in index order.
constexpr int kSortTile = 1024, kRadix = 256, kScanChunk = 1024;

__global__ void __launch_bounds__(64) k_radix_hist(const unsigned* __restrict__ keys, int n, int shift, int G,
                                                   unsigned* __restrict__ hist) {
  __shared__ unsigned bins[kRadix];
  const int lane = threadIdx.x, tile = blockIdx.x;
#pragma unroll
  for (int q = 0; q < kRadix / 64; ++q) bins[lane + 64 * q] = 0u;
  __syncthreads();
  for (int e = 0; e < kSortTile / 64; ++e) {
    const long i = (long)tile * kSortTile + e * 64 + lane;
    if (i < n) atomicAdd(&bins[(keys[i] >> shift) & (kRadix - 1)], 1u);
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < kRadix / 64; ++q) hist[(size_t)(lane + 64 * q) * G + tile] = bins[lane + 64 * q];
}

// Exclusive scan of 256 values, one per thread of a 256-thread block; `total` gets
// their sum.  lds: 4 words.
__device__ __forceinline__ unsigned block_exclusive_scan(unsigned v, unsigned* lds, unsigned& total) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  unsigned x = v;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const unsigned y = __shfl_up(x, o, 64);
    if (lane >= o) x += y;
  }
  if (lane == 63) lds[wave] = x;
  __syncthreads();
  unsigned off = 0;
  total = 0;
#pragma unroll
  for (int w = 0; w < 4; ++w) {
    const unsigned t = lds[w];
    if (w < wave) off += t;
    total += t;
  }
  __syncthreads();
  return off + x - v;
}

// sums[b] = sum of data[b * kScanChunk, (b + 1) * kScanChunk)
__global__ void __launch_bounds__(256) k_scan_sums(const unsigned* __restrict__ data, int m,
                                                   unsigned* __restrict__ sums) {
  __shared__ unsigned lds[4];
  const long base = (long)blockIdx.x * kScanChunk + 4 * threadIdx.x;
  unsigned v = 0;
#pragma unroll
  for (int k = 0; k < 4; ++k)
    if (base + k < m) v += data[base + k];
  unsigned total;
  (void)block_exclusive_scan(v, lds, total);
  if (threadIdx.x == 0) sums[blockIdx.x] = total;
}

// sums[0, nb) -> exclusive prefix sums, in place (one block, 1024 entries per round)
__global__ void __launch_bounds__(256) k_scan_top(unsigned* __restrict__ sums, int nb) {
  __shared__ unsigned lds[4];
  unsigned carry = 0;
  for (int base = 0; base < nb; base += 1024) {
    const int i0 = base + 4 * threadIdx.x;
    unsigned v[4], s = 0;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      v[k] = i0 + k < nb ? sums[i0 + k] : 0u;
      s += v[k];
    }
    unsigned total;
    unsigned run = carry + block_exclusive_scan(s, lds, total);
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      if (i0 + k < nb) sums[i0 + k] = run;
      run += v[k];
    }
    carry += total;
  }
}

// data -> exclusive prefix sums, each chunk offset by its scanned sums[b]
__global__ void __launch_bounds__(256) k_scan_apply(unsigned* __restrict__ data, int m,
                                                    const unsigned* __restrict__ sums) {
  __shared__ unsigned lds[4];
  const long base = (long)blockIdx.x * kScanChunk + 4 * threadIdx.x;
  unsigned v[4], s = 0;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    v[k] = base + k < m ? data[base + k] : 0u;
    s += v[k];
  }
  unsigned total;
  unsigned run = sums[blockIdx.x] + block_exclusive_scan(s, lds, total);
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    if (base + k < m) data[base + k] = run;
    run += v[k];
  }
}

__global__ void __launch_bounds__(64) k_radix_scatter(const unsigned* __restrict__ keys_in,
                                                      const unsigned* __restrict__ vals_in,
                                                      unsigned* __restrict__ keys_out, unsigned* __restrict__ vals_out,
                                                      int n, int shift, int G, const unsigned* __restrict__ offs) {
  __shared__ unsigned bins[kRadix];
  const int lane = threadIdx.x, tile = blockIdx.x;
#pragma unroll
  for (int q = 0; q < kRadix / 64; ++q) bins[lane + 64 * q] = offs[(size_t)(lane + 64 * q) * G + tile];
  __syncthreads();
  const unsigned long long below = (1ull << lane) - 1ull;
  for (int e = 0; e < kSortTile / 64; ++e) {
    const long i = (long)tile * kSortTile + e * 64 + lane;
    const bool ok = i < n;
    const unsigned key = ok ? keys_in[i] : 0u;
    const int d = (int)((key >> shift) & (kRadix - 1));
    unsigned long long todo = __ballot(ok);
    unsigned dst = 0;
    while (todo) {  // one digit value of this row of 64 elements per round
      const int leader = __ffsll((long long)todo) - 1;
      const int dl = __shfl(d, leader, 64);
      const unsigned long long same = __ballot(ok && d == dl);
      const unsigned base = bins[dl];
      if (ok && d == dl) dst = base + (unsigned)__popcll(same & below);
      if (lane == leader) bins[dl] = base + (unsigned)__popcll(same);
      todo &= ~same;
      __builtin_amdgcn_wave_barrier();
    }
    if (ok) {
      keys_out[dst] = key;
      vals_out[dst] = vals_in[i];
    }
  }
}

__global__ void k_term_gather(int n, int K, const unsigned* __restrict__ vals, const int* __restrict__ term_atom,
                              const double* __restrict__ wslot, const double* __restrict__ xyz,
                              double* __restrict__ tx, double* __restrict__ ty, double* __restrict__ wt) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const unsigned slot = vals[p];
  const int a = term_atom[slot / (unsigned)K];
  tx[p] = xyz[3 * a];
  ty[p] = xyz[3 * a + 1];
  wt[p] = wslot[slot];
}

}  // namespace

void displace(hipStream_t st, int n, const double* xyz, const double* rms, uint64_t seed, int config, double lo,
              double hi, const double* faces, int n_faces, double* out) {
  if (n > 0)
    hipLaunchKernelGGL(k_displace, dim3((n + 255) / 256), dim3(256), 0, st, n, xyz, rms, seed, config, lo, hi, faces,
                       n_faces, out);
}

void term_keys(hipStream_t st, const TermKeyArgs& a) {
  const int nc = a.S * a.Ct;
  hipLaunchKernelGGL(k_fill_i32, dim3((nc + 255) / 256), dim3(256), 0, st, nc, 0, a.counts);
  if (a.n > 0)
    hipLaunchKernelGGL(k_term_keys, dim3((a.n + 255) / 256), dim3(256), 0, st, a.n, a.K, a.term_atom, a.term_cls,
                       a.term_w, a.term_mode, a.classes, a.sub_base, a.xyz, a.S, a.Ct, a.origin, a.dz, a.m, a.split,
                       a.discon, a.keys, a.vals, a.wslot, a.counts);
}

size_t sort_pairs_temp_bytes(int n) {
  const size_t G = ((size_t)std::max(n, 1) + kSortTile - 1) / kSortTile;
  const size_t m = kRadix * G, nb = (m + kScanChunk - 1) / kScanChunk;
  return (m + nb + 64) * sizeof(unsigned);
}

unsigned* sort_pairs(hipStream_t st, void* temp, unsigned* keys, unsigned* keys2, unsigned* vals, unsigned* vals2,
                     int n, int bits) {
  if (n <= 0) return vals;
  const int G = (n + kSortTile - 1) / kSortTile;
  const int m = kRadix * G, nb = (m + kScanChunk - 1) / kScanChunk;
  unsigned* const hist = static_cast<unsigned*>(temp);
  unsigned* const sums = hist + m;
  unsigned *ks = keys, *kd = keys2, *vs = vals, *vd = vals2;
  for (int shift = 0; shift < bits; shift += 8) {
    hipLaunchKernelGGL(k_radix_hist, dim3(G), dim3(64), 0, st, ks, n, shift, G, hist);
    hipLaunchKernelGGL(k_scan_sums, dim3(nb), dim3(256), 0, st, hist, m, sums);
    hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(256), 0, st, sums, nb);
    hipLaunchKernelGGL(k_scan_apply, dim3(nb), dim3(256), 0, st, hist, m, sums);
    hipLaunchKernelGGL(k_radix_scatter, dim3(G), dim3(64), 0, st, ks, vs, kd, vd, n, shift, G, hist);
    std::swap(ks, kd);
    std::swap(vs, vd);
  }
  return vs;
}

void term_gather(hipStream_t st, int n, int K, const unsigned* vals, const int* term_atom, const double* wslot,
                 const double* xyz, double* tx, double* ty, double* wt) {
  if (n > 0)
    hipLaunchKernelGGL(k_term_gather, dim3((n + 255) / 256), dim3(256), 0, st, n, K, vals, term_atom, wslot, xyz, tx,
                       ty, wt);
}

}  // namespace totem
