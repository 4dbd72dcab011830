// gfx950 kernels for the guest-framework Lennard-Jones Vext path.
//
// Physics (restated from native/vext/kernel.cpp:13-44 and
// src/string_vext/neighbor.py:17-28 of the reference):
//   E(c, u) = sum_{s in {0,1}} sum_{images j, |d|^2 < rc^2} c12_j i6^2 - c6_j i6 + shift_j
//   d = c + o_s u - img_j,  i6 = (1 / max(|d|^2, floor2_j))^3.
//
// Work decomposition (one wave = one pose center x 64 orientations):
//   * pose centers are binned into spatial tiles; every tile owns a candidate
//     list of framework images within (tile radius + rc + max|o|), grouped by
//     LJ type and stored tile-relative (float4 + an FP64 copy);
//   * a wave streams its tile's list 64 entries at a time, keeps the entries
//     within rc + max|o| of its own center (ballot + mbcnt compaction into a
//     per-wave LDS buffer), then every lane (one orientation) sums both guest
//     sites against the compacted atoms, read from LDS by broadcast.
//   * with r = img - c (uniform over the wave) and t = r.u (per lane), the
//     two site distances are |r|^2 + o_s^2 - 2 o_s t: one dot product serves
//     both sites; for a symmetric guest the atoms inside the cutoff for every
//     orientation take both sites with one reciprocal (product form below).
//
// Mixed precision: the pair arithmetic is FP32 with the distance clamped to
// at most rc^2 (for shell atoms d^2/rc^2 is clamped to [0, 1] by the clamp bit
// of the fma that forms it; atoms an orientation could bring within the
// contact floor are always FP64).  Because the shifted potential vanishes at
// rc, a clamped out-of-range pair contributes exactly -shift, which the
// per-type count of kept pairs adds back in FP64; no cutoff branch or select
// is needed and a pair on the wrong side of rc by FP32 rounding changes the
// energy by ~1e-8 K.  FP32 partials of at most kFlush atoms are flushed into
// an FP64 accumulator.  Atoms closer to the center than a per-type near
// radius -- where a single pair term can exceed the configured bound (default
// 50 K) or the wall's slope exceeds 300 K/A, so that an FP32 distance error
// would matter -- are evaluated in FP64 with the reference formula, staged in
// LDS by the compacting lane; overlap poses keep FP64 relative accuracy.
//
// Matrix cores are not used: the distances are a 16x16x4 FP32 matrix product
// per 16 atoms x 16 orientations, but FP32 MFMA does not co-execute with the
// VALU on gfx950 and such a kernel measured slower than this one.
#include "sv_device.h"
#include <climits>

namespace sv {
namespace {

__device__ __forceinline__ int find_job(const long long* __restrict__ job_work, int n_jobs,
                                        long long w) {
  int lo = 0, hi = n_jobs - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (job_work[mid] <= w) lo = mid; else hi = mid - 1;
  }
  return lo;
}

__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
  return v;
}
__device__ __forceinline__ float wave_sum(float v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}
__device__ __forceinline__ int wave_exclusive_scan(int v, int lane, int* total) {
  int x = v;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const int y = __shfl_up(x, o, 64);
    if (lane >= o) x += y;
  }
  *total = __shfl(x, 63, 64);
  return x - v;
}

// Range of lattice shifts n_k whose image f + n lies within `radius` of the
// fractional point f0: |(f - f0 + n) H| <= radius implies
// |f_k - f0_k + n_k| <= radius * |column k of H^-1|.
struct ShiftBox {
  int lo[3], hi[3];
};
__device__ __forceinline__ ShiftBox shift_box(const double d[3], const double hn[3], double radius) {
  ShiftBox b;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const double h = radius * hn[k];
    b.lo[k] = (int)ceil(-d[k] - h);
    b.hi[k] = (int)floor(-d[k] + h);
  }
  return b;
}

// Cartesian offset of lattice image (d + n) and the class it falls in: 0 or 1
// within r2 (class 0 when its squared distance is below split2), -1 beyond.
// Written with explicit fma so the count and fill kernels (and both passes of
// the fill kernel) round identically and agree on every borderline image.
__device__ __forceinline__ int image_class(const double d[3], int na, int nb, int nc,
                                           const double* __restrict__ H, double r2,
                                           double split2, double* x, double* y, double* z) {
  const double ua = d[0] + na, ub = d[1] + nb, uc = d[2] + nc;
  *x = fma(ua, H[0], fma(ub, H[3], uc * H[6]));
  *y = fma(ua, H[1], fma(ub, H[4], uc * H[7]));
  *z = fma(ua, H[2], fma(ub, H[5], uc * H[8]));
  const double q = fma(*x, *x, fma(*y, *y, *z * *z));
  return q <= r2 ? (q < split2 ? 0 : 1) : -1;
}

// ---------------------------------------------------------------------------
// Tile candidate lists
// ---------------------------------------------------------------------------
// One block of kListThreads threads per tile; thread i takes atoms i,
// i + kListThreads, ... of each type.
constexpr int kListThreads = 256;

// Exclusive prefix over the block's threads in thread order; `total` gets the
// block sum.  Ends with a barrier, so it can be called again at once.
__device__ __forceinline__ int list_block_scan(int v, int* total) {
  __shared__ int s_w[kListThreads / 64];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  int total_w;
  const int x = wave_exclusive_scan(v, lane, &total_w);
  if (lane == 0) s_w[w] = total_w;
  __syncthreads();
  int before = 0, all = 0;
#pragma unroll
  for (int k = 0; k < kListThreads / 64; ++k) {
    before += k < w ? s_w[k] : 0;
    all += s_w[k];
  }
  __syncthreads();
  *total = all;
  return before + x;
}

// Segment of (tile g, type t, class c): seg_base[g] + t * n_cls + c.  Jobs with
// one class put every image of a type in class 0 (split2 = +inf).
__device__ __forceinline__ double class_split2(const Job& J, int t) {
  return J.n_cls > 1 ? J.split2[t] : INFINITY;
}

__global__ __launch_bounds__(kListThreads) void tile_count_kernel(const Job* __restrict__ jobs,
                                                                  const Tile* __restrict__ tiles,
                                                                  const long long* __restrict__ seg_base,
                                                                  int* __restrict__ counts) {
  const int g = blockIdx.x;
  const Tile& T = tiles[g];
  const Job& J = jobs[T.job];
  const double r2 = T.list_r * T.list_r;
  for (int t = 0; t < J.n_types; ++t) {
    const double split2 = class_split2(J, t);
    int count[2] = {0, 0};
    for (int a = J.type_start[t] + threadIdx.x; a < J.type_start[t + 1]; a += kListThreads) {
      double d[3];
#pragma unroll
      for (int k = 0; k < 3; ++k) d[k] = J.frac[3 * a + k] - T.frac[k];
      const ShiftBox b = shift_box(d, J.hn, T.list_r);
      for (int na = b.lo[0]; na <= b.hi[0]; ++na)
        for (int nb = b.lo[1]; nb <= b.hi[1]; ++nb)
          for (int nc = b.lo[2]; nc <= b.hi[2]; ++nc) {
            double x, y, z;
            const int c = image_class(d, na, nb, nc, J.H, r2, split2, &x, &y, &z);
            count[0] += c == 0;
            count[1] += c == 1;
          }
    }
    for (int c = 0; c < J.n_cls; ++c) {
      int total;
      (void)list_block_scan(count[c], &total);
      if (threadIdx.x == 0) counts[seg_base[g] + (long long)t * J.n_cls + c] = total;
    }
  }
}

// Exclusive scan of counts[0..n) into offsets[0..n], in three launches: every
// block sums one chunk of kScanChunk counts, one block scans the chunk sums,
// then every block scans its chunk from its chunk's offset.  Thread i of a
// block owns the kScanItems consecutive counts starting at i * kScanItems.
constexpr int kScanThreads = 1024, kScanItems = 8, kScanChunk = kScanThreads * kScanItems;

__device__ __forceinline__ long long scan_thread_items(const int* __restrict__ counts, long long n,
                                                       long long first, int v[kScanItems]) {
  long long sum = 0;
  if (first + kScanItems <= n) {
    const int4 a = *reinterpret_cast<const int4*>(counts + first);
    const int4 b = *reinterpret_cast<const int4*>(counts + first + 4);
    v[0] = a.x; v[1] = a.y; v[2] = a.z; v[3] = a.w;
    v[4] = b.x; v[5] = b.y; v[6] = b.z; v[7] = b.w;
  } else {
#pragma unroll
    for (int j = 0; j < kScanItems; ++j) v[j] = first + j < n ? counts[first + j] : 0;
  }
#pragma unroll
  for (int j = 0; j < kScanItems; ++j) sum += v[j];
  return sum;
}

// Exclusive prefix of `x` over the block's threads; `total` gets the block sum.
__device__ __forceinline__ long long block_exclusive_scan(long long x, long long* total) {
  __shared__ long long s_w[kScanThreads / 64];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  long long inc = x;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const long long y = __shfl_up(inc, o, 64);
    if (lane >= o) inc += y;
  }
  if (lane == 63) s_w[w] = inc;
  __syncthreads();
  long long before = 0, all = 0;
  for (int k = 0; k < kScanThreads / 64; ++k) {
    before += k < w ? s_w[k] : 0;
    all += s_w[k];
  }
  *total = all;
  return before + inc - x;
}

__global__ __launch_bounds__(kScanThreads) void scan_sums_kernel(const int* __restrict__ counts,
                                                                 long long n,
                                                                 long long* __restrict__ sums) {
  int v[kScanItems];
  const long long first = (long long)blockIdx.x * kScanChunk + (long long)threadIdx.x * kScanItems;
  long long total;
  (void)block_exclusive_scan(scan_thread_items(counts, n, first, v), &total);
  if (threadIdx.x == 0) sums[blockIdx.x] = total;
}

// One block: exclusive scan of the nb chunk sums in place; offsets[n] = total.
__global__ __launch_bounds__(kScanThreads) void scan_chunks_kernel(long long* __restrict__ sums,
                                                                   long long nb, long long n,
                                                                   long long* __restrict__ offsets) {
  const long long per = (nb + kScanThreads - 1) / kScanThreads;
  const long long beg = threadIdx.x * per < nb ? threadIdx.x * per : nb;
  const long long end = beg + per < nb ? beg + per : nb;
  long long mine = 0;
  for (long long i = beg; i < end; ++i) mine += sums[i];
  long long total;
  long long run = block_exclusive_scan(mine, &total);
  for (long long i = beg; i < end; ++i) {
    const long long c = sums[i];
    sums[i] = run;
    run += c;
  }
  if (threadIdx.x == 0) offsets[n] = total;
}

__global__ __launch_bounds__(kScanThreads) void scan_apply_kernel(const int* __restrict__ counts,
                                                                  long long n,
                                                                  const long long* __restrict__ sums,
                                                                  long long* __restrict__ offsets) {
  int v[kScanItems];
  const long long first = (long long)blockIdx.x * kScanChunk + (long long)threadIdx.x * kScanItems;
  long long total;
  long long run = (sums ? sums[blockIdx.x] : 0) +
                  block_exclusive_scan(scan_thread_items(counts, n, first, v), &total);
  // a single chunk (sums == nullptr) is scanned by this one launch
  if (!sums && threadIdx.x == 0) offsets[n] = total;
#pragma unroll
  for (int j = 0; j < kScanItems; ++j) {
    if (first + j < n) offsets[first + j] = run;
    run += v[j];
  }
}

__global__ __launch_bounds__(kListThreads) void tile_fill_kernel(const Job* __restrict__ jobs,
                                                                 const Tile* __restrict__ tiles,
                                                                 const long long* __restrict__ seg_base,
                                                                 const long long* __restrict__ offsets,
                                                                 float4* __restrict__ list32,
                                                                 double* __restrict__ list64) {
  const int g = blockIdx.x;
  const Tile& T = tiles[g];
  const Job& J = jobs[T.job];
  const double r2 = T.list_r * T.list_r, scale = J.list_scale;
  for (int t = 0; t < J.n_types; ++t) {
    const double split2 = class_split2(J, t);
    const long long seg = seg_base[g] + (long long)t * J.n_cls;
    long long base[2] = {offsets[seg], J.n_cls > 1 ? offsets[seg + 1] : 0};
    const long long seg_end[2] = {offsets[seg + 1], J.n_cls > 1 ? offsets[seg + 2] : 0};
    const int a_end = J.type_start[t + 1];
    for (int a0 = J.type_start[t]; a0 < a_end; a0 += kListThreads) {
      const int a = a0 + (int)threadIdx.x;
      const bool ok = a < a_end;
      double d[3] = {0.0, 0.0, 0.0};
      ShiftBox b = {{0, 0, 0}, {-1, -1, -1}};
      if (ok) {
#pragma unroll
        for (int k = 0; k < 3; ++k) d[k] = J.frac[3 * a + k] - T.frac[k];
        b = shift_box(d, J.hn, T.list_r);
      }
      // Pass 1: how many images of each class this thread emits.
      int mine[2] = {0, 0};
      for (int na = b.lo[0]; na <= b.hi[0]; ++na)
        for (int nb = b.lo[1]; nb <= b.hi[1]; ++nb)
          for (int nc = b.lo[2]; nc <= b.hi[2]; ++nc) {
            double x, y, z;
            const int c = image_class(d, na, nb, nc, J.H, r2, split2, &x, &y, &z);
            mine[0] += c == 0;
            mine[1] += c == 1;
          }
      long long slot[2];
      int total[2] = {0, 0};
      for (int c = 0; c < J.n_cls; ++c) slot[c] = base[c] + list_block_scan(mine[c], &total[c]);
      // Pass 2: write them in atom order (deterministic list order), never
      // past the end of their (tile, type, class) segment.
      for (int na = b.lo[0]; na <= b.hi[0]; ++na)
        for (int nb = b.lo[1]; nb <= b.hi[1]; ++nb)
          for (int nc = b.lo[2]; nc <= b.hi[2]; ++nc) {
            double x, y, z;
            const int c = image_class(d, na, nb, nc, J.H, r2, split2, &x, &y, &z);
            if (c < 0 || c >= J.n_cls || slot[c] >= seg_end[c]) continue;
            const long long q = slot[c]++;
            list32[q] = make_float4((float)(x * scale), (float)(y * scale), (float)(z * scale),
                                    __int_as_float(a));
            list64[3 * q + 0] = x;
            list64[3 * q + 1] = y;
            list64[3 * q + 2] = z;
          }
      for (int c = 0; c < J.n_cls; ++c) base[c] += total[c];
    }
  }
}

// ---------------------------------------------------------------------------
// Field kernels
// ---------------------------------------------------------------------------
struct WaveCtx {
  const Job* J;
  int slot;       // sorted center slot
  int tile;       // job-local tile
  int group;      // orientation group (64 axes per group)
  int k;          // this lane's orientation
  bool k_ok;
};

__device__ __forceinline__ void setup_item(const Job* __restrict__ jobs, int n_jobs,
                                           const long long* __restrict__ job_work, long long w,
                                           WaveCtx& c) {
  // Every value here is wave-uniform; readfirstlane lets the compiler keep the
  // job's fields in scalar registers also in the persistent kernel's loop.
  const int j = uniform(find_job(job_work, n_jobs, w));
  c.J = jobs + j;
  const long long local = w - c.J->work_begin;
  const int ng = c.J->n_groups;
  c.slot = uniform(ng == 1 ? (int)local : (int)(local / ng));
  c.group = uniform((int)(local - (long long)c.slot * ng));
  c.tile = ld_scalar(c.J->tile_of, c.slot);
  const int lane = threadIdx.x & 63;
  c.k = c.group * 64 + lane;
  c.k_ok = c.k < c.J->n_axes;
  if (!c.k_ok) c.k = c.J->n_axes - 1;
}

__device__ __forceinline__ bool setup_wave(const Job* __restrict__ jobs, int n_jobs,
                                           const long long* __restrict__ job_work,
                                           long long n_work, int wv, WaveCtx& c) {
  const long long w = xcd_block(blockIdx.x, gridDim.x) * 4 + wv;
  if (w >= n_work) return false;
  setup_item(jobs, n_jobs, job_work, w, c);
  return true;
}

// Work queue of the persistent launch: the item range is cut into 8
// contiguous partitions, one per XCD (blocks are dispatched round-robin over
// the XCDs, so block b starts on partition b % 8 and neighbouring items --
// same tile, same candidate list -- stay in one L2).  A wave takes items from
// its partition with one atomic per item and moves on to the next partition
// when it is drained; after all 8 are drained it returns -1, so every wave
// exits.
__device__ __forceinline__ long long next_item(unsigned long long* __restrict__ counters,
                                               long long n_work, int& part, int& visited) {
  const long long q = n_work / 8, r = n_work % 8;
  while (visited < 8) {
    const long long size = q + (part < r ? 1 : 0);
    const long long begin = part * q + (part < r ? part : r);
    unsigned long long got = 0;
    if ((threadIdx.x & 63) == 0) got = atomicAdd(&counters[part], 1ull);
    const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)got);
    const unsigned hi = __builtin_amdgcn_readfirstlane((unsigned)(got >> 32));
    const long long g = (long long)(((unsigned long long)hi << 32) | lo);
    if (g < size) return begin + g;
    part = (part + 1) & 7;
    ++visited;
  }
  return -1;
}

// Pose energies, fused marginal/grid, and the site field at the center
// (lane partials `site`, already including every shift term).  FAST (the mixed
// form) takes the Boltzmann terms exp(a - max) <= 1, their sum and its log in
// FP32: ~1e-7 relative on a sum of at most 64 terms, below the FP32 output's
// rounding.
template <bool FAST = false>
__device__ __forceinline__ void write_pose(const WaveCtx& c, double E, bool with_site, double site) {
  const Job& J = *c.J;
  const int cid = ld_scalar(J.center_id, c.slot);
  if (c.k_ok) {
    if (J.out64) J.out64[(long long)cid * J.n_axes + c.k] = E;
    if (J.out32) J.out32[(long long)c.k * J.n_centers + cid] = (float)E;
    if (!isfinite(E)) atomicOr(J.bad, 1);
  }
  if (J.marg32 && J.n_groups == 1) {
    // marginal = -T (logsumexp(-E/T) - ln K)   (reference physics.py:126-130)
    const double T = J.temperature;
    const double a = c.k_ok ? -E / T : -INFINITY;
    const double m = wave_max(a);
    double ls;
    if (FAST) {
      const float s = wave_sum(c.k_ok ? expf((float)(a - m)) : 0.0f);
      ls = (double)logf(s);
    } else {
      ls = log(wave_sum(c.k_ok ? exp(a - m) : 0.0));
    }
    if ((threadIdx.x & 63) == 0) {
      const double marg = -T * (m + ls - J.log_n_axes);
      const float mf = (float)marg;
      J.marg32[cid] = mf;
      if (J.grid32) J.grid32[cid] = asinhf(mf / 1000.0f);
      if (!isfinite(marg)) atomicOr(J.bad, 1);
    }
  }
  if (with_site) {
    const double v = wave_sum(site);
    if ((threadIdx.x & 63) == 0) {
      J.site32[cid] = (float)v;
      if (!isfinite(v)) atomicOr(J.bad, 1);
    }
  }
}

// FP64 pair term with the reference's exact cutoff/floor semantics
// (kernel.cpp:29-36).  d = o u - r where r = image - center.  RCP takes 1/d^2
// from v_rcp_f64 and one Newton step instead of the IEEE division.
template <bool RCP = false>
__device__ __forceinline__ double pair64(double o, double ux, double uy, double uz, double rx,
                                         double ry, double rz, double rc2, double c12, double c6,
                                         double sh, double fl2) {
  const double dx = o * ux - rx, dy = o * uy - ry, dz = o * uz - rz;
  const double d2 = dx * dx + dy * dy + dz * dz;
  double e = 0.0;
  if (d2 < rc2) {
    const double safe = fmax(d2, fl2);   // the floor: one v_max_f64
    const double inv = RCP ? inv_fast(safe) : 1.0 / safe;
    const double i6 = inv * inv * inv;
    e = c12 * i6 * i6 - c6 * i6 + sh;
  }
  return e;
}

// Same pair term from a precomputed squared distance (mixed near path).
__device__ __forceinline__ double pair64_d2(double d2, double rc2, double c12, double c6,
                                            double sh, double fl2) {
  const double safe = fmax(d2, fl2);   // the floor: one v_max_f64
  const double inv = inv_fast(safe);
  const double i6 = inv * inv * inv;
  const double e = fma(c12 * i6, i6, fma(-c6, i6, sh));
  return d2 < rc2 ? e : 0.0;
}

// The same term for a pair known to be inside the cutoff.
__device__ __forceinline__ double pair64_in(double d2, double c12, double c6, double sh,
                                            double fl2) {
  const double inv = inv_fast(fmax(d2, fl2));
  const double i6 = inv * inv * inv;
  return fma(c12 * i6, i6, fma(-c6, i6, sh));
}

// Atoms per FP32 partial sum in the mixed kernel.
constexpr int kFlush = 16;

struct MixedLds {
  float4 in[4][64];    // (rx, ry, rz, |r|^2 + o0^2); symmetric guests: (r, 4 w^2)
  float4 far[4][64];
  float4 out[4][64];   // outer shell (symmetric guests): (r, (|r|^2 + o^2) / rc^2)
  float in1[4][64];    // |r|^2 + o1^2 (asymmetric guests only)
  float far1[4][64];
  float2 inb[4][64];   // symmetric guests' inner entries: (128 w A, -96 w)
  double4 near[4][32]; // FP64 (qx, qy, qz, |q|^2 + o0^2) of near-contact atoms, 32 at a time
  // symmetric guests: (w^2, 8 w^3, -6 w, 1) with w = |q|^2 + o^2 for an atom
  // that takes the product form, (., ., ., 0) otherwise
  double4 nearb[4][32];
  int near_idx[4][64]; // queued near-contact list entries (offset in the type segment)
};

// Type t's 64-byte record through scalar loads (eight 8-byte words).
__device__ __forceinline__ FieldType ld_field_type(const FieldType* p, int t) {
  unsigned long long w[8];
#pragma unroll
  for (int i = 0; i < 8; ++i) w[i] = ld_scalar(reinterpret_cast<const unsigned long long*>(p + t), i);
  FieldType f;
  __builtin_memcpy(&f, w, sizeof(f));
  return f;
}

// One pose center x 64 orientations of the mixed-precision field.
template <bool SYM>
__device__ __forceinline__ void mixed_item(const WaveCtx& c, int wv, MixedLds& L,
                                           const long long* __restrict__ offsets,
                                           const float4* __restrict__ list32,
                                           const double* __restrict__ list64) {
  float4 (*s_in)[64] = L.in;
  float4 (*s_far)[64] = L.far;
  float4 (*s_out)[64] = L.out;
  const int lane = threadIdx.x & 63;
  const Job& J = *c.J;
  // The site field of a center is the same pair sum taken at the center
  // itself over the same candidates; the group-0 wave of the center
  // accumulates it lane-parallel while it compacts the candidate list.
  const bool do_site = J.site32 != nullptr && c.group == 0;

  const double ux64 = J.axes64[3 * c.k], uy64 = J.axes64[3 * c.k + 1], uz64 = J.axes64[3 * c.k + 2];
  // A symmetric guest's FP32 loops take t' = r . (4|o| u): the inner loop's
  // P = 4w^2 - (4|o| t)^2 is then one fma, and the shell constants absorb the
  // scale.  (o = 0 makes t' = 0, and the unscaled constants are 0 too.)
  const double k4d = SYM ? 4.0 * fabs(J.o0) : 1.0;
  const bool scale_u = SYM && k4d > 0.0;
  const double usc = SYM ? k4d : 1.0;
  const float ux = (float)(ux64 * usc), uy = (float)(uy64 * usc), uz = (float)(uz64 * usc);
  const double cx = ld_scalar(J.center_rel, 3LL * c.slot),
               cy = ld_scalar(J.center_rel, 3LL * c.slot + 1),
               cz = ld_scalar(J.center_rel, 3LL * c.slot + 2);
  const float cxf = uniform_f((float)cx), cyf = uniform_f((float)cy), czf = uniform_f((float)cz);
  const float rc2f = J.rc2f, keep2f = J.keep2f, m2o0 = J.m2o0f, m2o1 = J.m2o1f;
  const float o0sq = J.o0sqf, o1sq = J.o1sqf, in2 = J.in2f;

  // Outer shell (symmetric guests): with |r|^2 >= rc^2 - o^2 the two site
  // distances |r|^2 + o^2 -/+ 2|o| |t| cannot both be inside the cutoff, and
  // the farther one never is; only the closer site is evaluated.  The margin
  // covers the FP32 rounding of |r|^2.
  const float out2 = uniform_f(SYM ? (float)((J.rc2 - (double)J.o0 * J.o0) * (1.0 + 1e-6)) : 3.0e38f);
  const float m2abs = uniform_f(scale_u ? (float)(-0.5 / J.rc2) : -fabsf(J.m2o0sf));
  const double rc2 = J.rc2;
  const float inv_rc2f = J.inv_rc2f;
  const float m2o0s = uniform_f(scale_u ? (float)(-copysign(0.5, J.o0) / J.rc2) : J.m2o0sf);
  const float m2o1s = uniform_f(scale_u ? -m2o0s : J.m2o1sf);
  const double m2o0d = uniform_d(-2.0 * J.o0), m2o1d = uniform_d(-2.0 * J.o1);
  // near-contact product form: symmetric guest, every near-contact site inside
  const bool near_pf = SYM && J.near_inside;
  const double o0sqd = uniform_d(J.o0 * J.o0), o1sqd = uniform_d(J.o1 * J.o1);
  const double do1 = uniform_d(o1sqd - o0sqd);
  const long long seg0 = J.seg_base + (long long)c.tile * J.n_types;

  double E = 0.0, S = 0.0, S_shift = 0.0;
  for (int t = 0; t < J.n_types; ++t) {
    const long long beg = offsets[seg0 + t], end = offsets[seg0 + t + 1];
    const FieldType ft = ld_field_type(J.ftype, t);
    const float s6 = ft.s6f, fl2f = ft.floor2f;
    const float near2 = ft.near2f, lo2 = ft.lo2f;
    const double c6 = ft.c6, c12 = ft.c12, sh = ft.shift;
    const double fl2 = ft.floor2;
    const double m2c12 = uniform_d(-2.0 * c12);
    const double lo2d = (double)lo2;
    const double c6s = ft.c6s;   // shell loop scale
    const float k128s6 = uniform_f(128.0f * s6);
    const double c6s6 = uniform_d(c6 * (double)s6);
    const float s6s = ft.s6sf;
    long long kept_far = 0, kept_out = 0;
    float Sf = 0.0f;   // this lane's FP32 site terms of this type (one per chunk)
    // Near-contact atoms are queued by list offset and evaluated when the queue
    // would overflow and at the end of the type: the compacting lanes stage
    // their FP64 center-relative vectors (and site terms) lane-parallel, then
    // every orientation lane sums them in FP64.  Staging once per queue rather
    // than once per chunk keeps the FP64 code off most chunks.
    int n_near = 0;
    auto flush_near = [&]() {
      for (int h = 0; h < n_near; h += 32) {
        const int m = n_near - h < 32 ? n_near - h : 32;
        wave_sync();
        // Stage up to 32 atoms lane-parallel.  With a symmetric guest and every
        // near-contact site inside the cutoff, an atom that no orientation
        // brings within the contact floor (|q| above the floor reach) takes
        // both sites in one product-form term; the others take the per-site
        // form with the floor.
        if (lane < m) {
          const long long id = beg + L.near_idx[wv][h + lane];
          const double qx = list64[3 * id] - cx, qy = list64[3 * id + 1] - cy,
                       qz = list64[3 * id + 2] - cz;
          const double RR = qx * qx + qy * qy + qz * qz;
          // w = |q|^2 + o0^2; the second site's base is w + (o1^2 - o0^2)
          const double w = RR + o0sqd;
          L.near[wv][lane] = make_double4(qx, qy, qz, w);
          if (near_pf) {
            const double w2 = w * w;
            L.nearb[wv][lane] = make_double4(w2, 8.0 * w * w2, -6.0 * w, RR > lo2d ? 1.0 : 0.0);
          }
          if (do_site) S += pair64_d2(RR, rc2, c12, c6, sh, fl2);
        }
        wave_sync();
        for (int q = 0; q < m; ++q) {
          const double4 d = L.near[wv][q];
          if (near_pf && L.nearb[wv][q].w != 0.0) {
            // Product form (both sites): d0,1 = w -/+ t' with t' = 2 o q.u,
            // Q = d0 d1, a = 1/Q, S3 = 1/d0^3 + 1/d1^3 = a^2 (8 w^3 a - 6 w),
            // 1/(d0 d1)^3 = a^3, 1/d0^6 + 1/d1^6 = S3^2 - 2 a^3.
            const double4 b = L.nearb[wv][q];
            const double tt = (d.x * ux64 + d.y * uy64 + d.z * uz64) * m2o0d;
            const double a = inv_fast(fma(-tt, tt, b.x));
            const double a2 = a * a;
            const double S3 = a2 * fma(b.y, a, b.z);
            E = fma(S3, fma(c12, S3, -c6), E);
            E = fma(m2c12, a2 * a, fma(2.0, sh, E));
          } else if (J.near_inside) {
            // FP64 near-contact pairs: one dot product serves both sites.
            const double tt = d.x * ux64 + d.y * uy64 + d.z * uz64;
            E += pair64_in(fma(tt, m2o0d, d.w), c12, c6, sh, fl2);
            E += pair64_in(fma(tt, m2o1d, SYM ? d.w : d.w + do1), c12, c6, sh, fl2);
          } else {
            const double tt = d.x * ux64 + d.y * uy64 + d.z * uz64;
            E += pair64_d2(fma(tt, m2o0d, d.w), rc2, c12, c6, sh, fl2);
            E += pair64_d2(fma(tt, m2o1d, SYM ? d.w : d.w + do1), rc2, c12, c6, sh, fl2);
          }
        }
      }
      wave_sync();
      n_near = 0;
    };
    // The three FP32 classes queue their atoms in LDS across chunks, like the
    // near-contact class: a queue is evaluated when the next chunk would
    // overflow it and at the end of the type, so the orientation loops run
    // over long runs of atoms rather than a few per chunk.  FP32 partials cover
    // at most kFlush atoms before they are added into the FP64 accumulator, so
    // a partial never grows past a few hundred K and its rounding stays ~1e-5 K.
    int n_in = 0, n_far = 0, n_out = 0;
    auto run_in = [&]() {
      for (int q0 = 0; SYM && q0 < n_in; q0 += kFlush) {
        // Symmetric guest, both sites at once: d0,1 = w -/+ k t with k = 2|o|,
        // P = 4 d0 d1 = A - t'^2 with A = 4 w^2 and t' = 2 k t; a = 1/P, i_s = 1/d_s:
        //   S3 = i0^3 + i1^3 = 128 w (A - 3P/4) a^3 = a^2 (B a + C),
        //   i0^3 i1^3 = 64 a^3,  i0^6 + i1^6 = S3^2 - 128 a^3,
        // with B = 128 w A and C = -96 w stored per atom; both sites together are
        //   s6 (i0^6 + i1^6) - S3 = s6 S3^2 - a^2 (B a + C + 128 s6 a):
        // one reciprocal and ten VALU operations per atom.
        const int q1 = q0 + kFlush < n_in ? q0 + kFlush : n_in;
        float acc = 0.f, acc1 = 0.f;
#pragma unroll 4
        for (int q = q0; q < q1; ++q) {
          const float4 f = s_in[wv][q];
          const float2 g = L.inb[wv][q];
          const float t = f.x * ux + f.y * uy + f.z * uz;   // 4|o| r.u
          const float P = fmaf(-t, t, f.w);
          const float a = __builtin_amdgcn_rcpf(P);
          const float a2 = a * a;
          const float g1 = fmaf(g.x, a, g.y);
          const float S3 = a2 * g1;
          acc = fmaf(S3, S3, acc);                        // times c6 s6
          acc1 = fmaf(-a2, fmaf(k128s6, a, g1), acc1);   // times c6
        }
        E = fma(c6s6, (double)acc, E);
        E = fma(c6, (double)acc1, E);
      }
      for (int q0 = 0; !SYM && q0 < n_in; q0 += kFlush) {
        const int q1 = q0 + kFlush < n_in ? q0 + kFlush : n_in;
        float acc = 0.f, acc1 = 0.f;
#pragma unroll 4
        for (int q = q0; q < q1; ++q) {
          const float4 f = s_in[wv][q];
          const float tdot = f.x * ux + f.y * uy + f.z * uz;
          const float d0 = fmaf(tdot, m2o0, f.w);
          const float d1 = fmaf(tdot, m2o1, SYM ? f.w : L.in1[wv][q]);
          const float i0 = __builtin_amdgcn_rcpf(d0);
          const float i1 = __builtin_amdgcn_rcpf(d1);
          const float i06 = i0 * i0 * i0;
          const float i16 = i1 * i1 * i1;
          acc = fmaf(i06, fmaf(i06, s6, -1.0f), acc);
          acc1 = fmaf(i16, fmaf(i16, s6, -1.0f), acc1);
        }
        E = fma(c6, (double)acc, E);
        E = fma(c6, (double)acc1, E);
      }
      n_in = 0;
    };
    // Shell atoms work in units of rc^2: x = d^2 / rc^2 is clamped to [0, 1]
    // by the clamp bit of the fma that forms it (med3(., 0, 1) folds into it),
    // and c6 / rc^6, sigma^6 / rc^6 absorb the scale.  A pair beyond the cutoff
    // gives x = 1, i.e. exactly the clamped -shift of the unscaled form.
    auto run_far = [&]() {
      for (int q0 = 0; q0 < n_far; q0 += kFlush) {
        const int q1 = q0 + kFlush < n_far ? q0 + kFlush : n_far;
        float acc = 0.f, acc1 = 0.f;
#pragma unroll 4
        for (int q = q0; q < q1; ++q) {
          const float4 f = s_far[wv][q];
          const float tdot = f.x * ux + f.y * uy + f.z * uz;
          const float x0 = __builtin_amdgcn_fmed3f(fmaf(tdot, m2o0s, f.w), 0.0f, 1.0f);
          const float x1 =
              __builtin_amdgcn_fmed3f(fmaf(tdot, m2o1s, SYM ? f.w : L.far1[wv][q]), 0.0f, 1.0f);
          const float i0 = __builtin_amdgcn_rcpf(x0);
          const float i1 = __builtin_amdgcn_rcpf(x1);
          const float i06 = i0 * i0 * i0;
          const float i16 = i1 * i1 * i1;
          acc = fmaf(i06, fmaf(i06, s6s, -1.0f), acc);
          acc1 = fmaf(i16, fmaf(i16, s6s, -1.0f), acc1);
        }
        E = fma(c6s, (double)acc, E);
        E = fma(c6s, (double)acc1, E);
      }
      n_far = 0;
    };
    // Outer shell: the closer site only, |r|^2 + o^2 - 2|o||t| in units of
    // rc^2 (|t| is a free input modifier of the fma).
    auto run_out = [&]() {
      for (int q0 = 0; SYM && q0 < n_out; q0 += kFlush) {
        const int q1 = q0 + kFlush < n_out ? q0 + kFlush : n_out;
        float acc = 0.f;
#pragma unroll 4
        for (int q = q0; q < q1; ++q) {
          const float4 f = s_out[wv][q];
          const float tdot = f.x * ux + f.y * uy + f.z * uz;
          const float x = __builtin_amdgcn_fmed3f(fmaf(fabsf(tdot), m2abs, f.w), 0.0f, 1.0f);
          const float i = __builtin_amdgcn_rcpf(x);
          const float i3 = i * i * i;
          acc = fmaf(i3, fmaf(i3, s6s, -1.0f), acc);
        }
        E = fma(c6s, (double)acc, E);
      }
      n_out = 0;
    };
    for (long long base = beg; base < end; base += 64) {
      const long long idx = base + lane;
      const bool valid = idx < end;
      float rx = 0.f, ry = 0.f, rz = 0.f, rr = 3.0e38f;
      if (valid) {
        const float4 e = list32[idx];
        rx = e.x - cxf; ry = e.y - cyf; rz = e.z - czf;
        rr = rx * rx + ry * ry + rz * rz;
      }
      // (bitwise, not short-circuit, so that each class is a lane mask built
      // from the compare masks with scalar logic)
      const bool keep = rr < keep2f;
      // Atoms some orientation brings within the contact floor are always FP64,
      // so the FP32 classes never need the floor clamp.
      const bool is_near = keep & ((rr < near2) | (rr <= lo2));
      // inner: every orientation puts both sites strictly inside the cutoff, so
      // the pair needs no clamp
      const bool is_in = keep & !is_near & (rr < in2);
      const bool is_out = SYM & keep & !is_near & !is_in & (rr > out2);
      const bool is_far = keep & !is_near & !is_in & !is_out;
      // The class masks come from ballots of the plain compares and scalar
      // logic (a ballot of a combined predicate costs two VALU operations).
      const unsigned long long b_keep = __ballot(rr < keep2f);
      const unsigned long long b_near = __ballot(rr < near2) | __ballot(rr <= lo2);
      const unsigned long long b_in = __ballot(rr < in2);
      const unsigned long long b_out = SYM ? __ballot(rr > out2) : 0ull;
      const unsigned long long mnear = b_keep & b_near;
      const unsigned long long min_ = b_keep & ~b_near & b_in;
      const unsigned long long mout = b_keep & ~b_near & ~b_in & b_out;
      const unsigned long long mfar = b_keep & ~b_near & ~b_in & ~b_out;
      const int ni = __popcll(min_), nf = __popcll(mfar), no = SYM ? __popcll(mout) : 0;
      const int nn = __popcll(mnear);
      if (do_site) {
        // far candidates: clamped FP32 term (the shift is added per kept entry
        // below); near candidates get the FP64 reference term when flushed.
        const float ds = __builtin_amdgcn_fmed3f(rr, fl2f, rc2f);
        const float is = __builtin_amdgcn_rcpf(ds);
        const float is6 = is * is * is;
        const float sf = is6 * fmaf(is6, s6, -1.0f);
        Sf += is_far || is_in || is_out ? sf : 0.0f;
      }
      if (n_in + ni > 64) { run_in(); wave_sync(); }
      if (n_far + nf > 64) { run_far(); wave_sync(); }
      if (SYM && n_out + no > 64) { run_out(); wave_sync(); }
      if (n_near + nn > 64) flush_near();
      // Each site's |r|^2 + o_s^2 is rounded once, here; an asymmetric guest's
      // second base is stored beside the entry rather than derived from the first.
      if (is_in) {
        const int slot = n_in + lanes_below(min_);
        const float w = rr + o0sq;
        if (SYM) {
          // product form: (rx, ry, rz, A = 4 w^2) and (128 w A, -96 w)
          const float A = 4.0f * w * w, fw = 128.0f * w;
          s_in[wv][slot] = make_float4(rx, ry, rz, A);
          L.inb[wv][slot] = make_float2(fw * A, -0.75f * fw);
        } else {
          s_in[wv][slot] = make_float4(rx, ry, rz, w);
          L.in1[wv][slot] = rr + o1sq;
        }
      }
      if (is_far) {
        // shell entries in units of rc^2 (see the shell loop)
        const int slot = n_far + lanes_below(mfar);
        s_far[wv][slot] = make_float4(rx, ry, rz, (rr + o0sq) * inv_rc2f);
        if (!SYM) L.far1[wv][slot] = (rr + o1sq) * inv_rc2f;
      }
      if (SYM && is_out)
        s_out[wv][n_out + lanes_below(mout)] = make_float4(rx, ry, rz, (rr + o0sq) * inv_rc2f);
      if (is_near) L.near_idx[wv][n_near + lanes_below(mnear)] = (int)(idx - beg);
      n_in += ni;
      n_far += nf;
      n_out += no;
      n_near += nn;
      kept_far += ni + nf;
      kept_out += no;
      wave_sync();
    }
    if (n_in) run_in();
    if (n_far) run_far();
    if (SYM && n_out) run_out();
    if (n_near) flush_near();
    // Every kept far pair was clamped to rc^2 at worst, contributing -shift.
    // (one site per outer-shell atom: its other site is never inside)
    E = fma((double)(2 * kept_far + kept_out), sh, E);
    S_shift = fma((double)(kept_far + kept_out), sh, S_shift);
    S = fma(c6, (double)Sf, S);
  }
  write_pose<true>(c, E, do_site, lane == 0 ? S + S_shift : S);
}

// The block's 25 KB of LDS holds the CU to 6 waves per SIMD; the register
// target of 5 leaves the compiler's schedule room (1.4% faster than a
// target of 6, with the same occupancy and the same results).
template <bool SYM>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(5))) void field_mixed_kernel(
    const Job* __restrict__ jobs, int n_jobs, const long long* __restrict__ job_work,
    long long n_work, const long long* __restrict__ offsets, const float4* __restrict__ list32,
    const double* __restrict__ list64) {
  __shared__ MixedLds L;
  const int wv = uniform(threadIdx.x >> 6);
  WaveCtx c;
  if (!setup_wave(jobs, n_jobs, job_work, n_work, wv, c)) return;
  mixed_item<SYM>(c, wv, L, offsets, list32, list64);
}

// Persistent form: a fixed grid whose waves drain the XCD-partitioned queue.
template <bool SYM>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(5))) void field_mixed_persistent(const Job* __restrict__ jobs,
                                                              int n_jobs,
                                                              const long long* __restrict__ job_work,
                                                              long long n_work,
                                                              unsigned long long* __restrict__ counters,
                                                              const long long* __restrict__ offsets,
                                                              const float4* __restrict__ list32,
                                                              const double* __restrict__ list64) {
  __shared__ MixedLds L;
  const int wv = uniform(threadIdx.x >> 6);
  int part = (int)(blockIdx.x % 8), visited = 0;
  for (;;) {
    const long long w = next_item(counters, n_work, part, visited);
    if (w < 0) break;
    WaveCtx c;
    setup_item(jobs, n_jobs, job_work, w, c);
    mixed_item<SYM>(c, wv, L, offsets, list32, list64);
  }
}

template <bool RCP>
__global__ __launch_bounds__(256) void field_fp64_kernel(const Job* __restrict__ jobs, int n_jobs,
                                                         const long long* __restrict__ job_work,
                                                         long long n_work,
                                                         const long long* __restrict__ offsets,
                                                         const double* __restrict__ list64) {
  __shared__ double s_r[4][3][64];
  const int lane = threadIdx.x & 63;
  const int wv = uniform(threadIdx.x >> 6);
  WaveCtx c;
  if (!setup_wave(jobs, n_jobs, job_work, n_work, wv, c)) return;
  const Job& J = *c.J;
  const bool do_site = J.site32 != nullptr && c.group == 0;
  const double ux = J.axes64[3 * c.k], uy = J.axes64[3 * c.k + 1], uz = J.axes64[3 * c.k + 2];
  const double cx = ld_scalar(J.center_rel, 3LL * c.slot),
               cy = ld_scalar(J.center_rel, 3LL * c.slot + 1),
               cz = ld_scalar(J.center_rel, 3LL * c.slot + 2);
  const double keep2 = J.keep_r * J.keep_r, rc2 = J.rc2, o0 = J.o0, o1 = J.o1;
  const long long seg0 = J.seg_base + (long long)c.tile * J.n_types;
  double E = 0.0, S = 0.0;
  for (int t = 0; t < J.n_types; ++t) {
    const long long beg = offsets[seg0 + t], end = offsets[seg0 + t + 1];
    const double c12 = ld_scalar(J.c12, t), c6 = ld_scalar(J.c6, t), sh = ld_scalar(J.shift, t),
                 fl2 = ld_scalar(J.floor2, t);
    for (long long base = beg; base < end; base += 64) {
      const long long idx = base + lane;
      double rx = 0.0, ry = 0.0, rz = 0.0;
      bool keep = false;
      if (idx < end) {
        rx = list64[3 * idx] - cx; ry = list64[3 * idx + 1] - cy; rz = list64[3 * idx + 2] - cz;
        keep = rx * rx + ry * ry + rz * rz < keep2;
      }
      const unsigned long long m = __ballot(keep);
      if (keep) {
        const int p = lanes_below(m);
        s_r[wv][0][p] = rx; s_r[wv][1][p] = ry; s_r[wv][2][p] = rz;
        if (do_site) S += pair64<RCP>(0.0, ux, uy, uz, rx, ry, rz, rc2, c12, c6, sh, fl2);
      }
      wave_sync();
      const int n = __popcll(m);
      for (int q = 0; q < n; ++q) {
        const double ax = s_r[wv][0][q], ay = s_r[wv][1][q], az = s_r[wv][2][q];
        E += pair64<RCP>(o0, ux, uy, uz, ax, ay, az, rc2, c12, c6, sh, fl2);
        E += pair64<RCP>(o1, ux, uy, uz, ax, ay, az, rc2, c12, c6, sh, fl2);
      }
      wave_sync();
    }
  }
  write_pose(c, E, do_site, S);
}

// Exact count of in-cutoff (site, image) pairs: the pair-eval census that the
// throughput metric is quoted in.  Same traversal as field_fp64_kernel.
__global__ __launch_bounds__(256) void census_kernel(const Job* __restrict__ jobs, int n_jobs,
                                                     const long long* __restrict__ job_work,
                                                     long long n_work,
                                                     const long long* __restrict__ offsets,
                                                     const double* __restrict__ list64) {
  __shared__ double s_r[4][3][64];
  const int lane = threadIdx.x & 63;
  const int wv = uniform(threadIdx.x >> 6);
  WaveCtx c;
  if (!setup_wave(jobs, n_jobs, job_work, n_work, wv, c)) return;
  const Job& J = *c.J;
  const double ux = J.axes64[3 * c.k], uy = J.axes64[3 * c.k + 1], uz = J.axes64[3 * c.k + 2];
  const double cx = ld_scalar(J.center_rel, 3LL * c.slot),
               cy = ld_scalar(J.center_rel, 3LL * c.slot + 1),
               cz = ld_scalar(J.center_rel, 3LL * c.slot + 2);
  const double keep2 = J.keep_r * J.keep_r, rc2 = J.rc2, o0 = J.o0, o1 = J.o1;
  const long long seg0 = J.seg_base + (long long)c.tile * J.n_types;
  const long long beg = offsets[seg0], end = offsets[seg0 + J.n_types];
  unsigned long long pairs = 0;
  for (long long base = beg; base < end; base += 64) {
    const long long idx = base + lane;
    double rx = 0.0, ry = 0.0, rz = 0.0;
    bool keep = false;
    if (idx < end) {
      rx = list64[3 * idx] - cx; ry = list64[3 * idx + 1] - cy; rz = list64[3 * idx + 2] - cz;
      keep = rx * rx + ry * ry + rz * rz < keep2;
    }
    const unsigned long long m = __ballot(keep);
    if (keep) {
      const int p = lanes_below(m);
      s_r[wv][0][p] = rx; s_r[wv][1][p] = ry; s_r[wv][2][p] = rz;
    }
    wave_sync();
    const int n = __popcll(m);
    for (int q = 0; q < n; ++q) {
      const double ax = s_r[wv][0][q], ay = s_r[wv][1][q], az = s_r[wv][2][q];
      const double a0 = o0 * ux - ax, b0 = o0 * uy - ay, g0 = o0 * uz - az;
      const double a1 = o1 * ux - ax, b1 = o1 * uy - ay, g1 = o1 * uz - az;
      pairs += (a0 * a0 + b0 * b0 + g0 * g0 < rc2);
      pairs += (a1 * a1 + b1 * b1 + g1 * g1 < rc2);
    }
    wave_sync();
  }
  if (!c.k_ok) pairs = 0;
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) pairs += __shfl_xor(pairs, o, 64);
  if (lane == 0) atomicAdd(J.pair_count, pairs);
}

// Marginal over orientations for jobs with more than 64 axes (the field kernel
// fuses it when one wave holds all of a center's orientations).
__global__ __launch_bounds__(64) void marginal_kernel(const Job* __restrict__ jobs, int n_jobs,
                                                      const long long* __restrict__ job_first,
                                                      long long n_total) {
  const long long g = blockIdx.x;
  if (g >= n_total) return;
  const int j = find_job(job_first, n_jobs, g);
  const Job& J = jobs[j];
  // jobs of at most 64 axes got their marginal from the field kernel
  if (J.n_groups == 1 || !J.marg32) return;
  const int cid = (int)(g - job_first[j]);
  const int lane = threadIdx.x;
  const double T = J.temperature;
  double m = -INFINITY;
  for (int k = lane; k < J.n_axes; k += 64) m = fmax(m, -J.out64[(long long)cid * J.n_axes + k] / T);
  m = wave_max(m);
  double s = 0.0;
  for (int k = lane; k < J.n_axes; k += 64) s += exp(-J.out64[(long long)cid * J.n_axes + k] / T - m);
  s = wave_sum(s);
  if (lane == 0) {
    const double marg = -T * (m + log(s) - J.log_n_axes);
    const float mf = (float)marg;
    J.marg32[cid] = mf;
    if (J.grid32) J.grid32[cid] = asinhf(mf / 1000.0f);
    if (!isfinite(marg)) atomicOr(J.bad, 1);
  }
}

// ---------------------------------------------------------------------------
// Drop-in CSR kernel: evaluate_reused_neighbors (reference kernel.cpp:5-45).
// Same operation order, IEEE division and no FMA contraction, so a center's
// energies reproduce the reference's x86-64 results bit for bit on the same
// CSR.  One wave per (center, group of 64 axes).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void legacy_kernel(int n_centers, int n_axes, int n_groups,
                                                     double cutoff2,
                                                     const double* __restrict__ centers,
                                                     const double* __restrict__ axes,
                                                     const double* __restrict__ images,
                                                     const double* __restrict__ offsets,
                                                     const double* __restrict__ coeff12,
                                                     const double* __restrict__ coeff6,
                                                     const double* __restrict__ shift,
                                                     const double* __restrict__ floor2,
                                                     const long long* __restrict__ starts,
                                                     const int* __restrict__ indices,
                                                     double* __restrict__ output,
                                                     int* __restrict__ bad_center) {
#pragma clang fp contract(off)
  const int wv = uniform(threadIdx.x >> 6);
  const long long w = (long long)blockIdx.x * 4 + wv;
  if (w >= (long long)n_centers * n_groups) return;
  const int center = (int)(w / n_groups);
  const int k = (int)(w % n_groups) * 64 + (threadIdx.x & 63);
  const bool ok = k < n_axes;
  const int kk = ok ? k : n_axes - 1;
  const long long n0 = starts[center], n1 = starts[center + 1];
  double energy = 0.0;
  for (int site = 0; site < 2; ++site) {
    const double x = centers[3 * center] + offsets[site] * axes[3 * kk];
    const double y = centers[3 * center + 1] + offsets[site] * axes[3 * kk + 1];
    const double z = centers[3 * center + 2] + offsets[site] * axes[3 * kk + 2];
    for (long long n = n0; n < n1; ++n) {
      const int j = indices[n];
      const double dx = x - images[3 * j];
      const double dy = y - images[3 * j + 1];
      const double dz = z - images[3 * j + 2];
      const double d2 = dx * dx + dy * dy + dz * dz;
      if (d2 >= cutoff2) continue;
      const double safe = d2 < floor2[j] ? floor2[j] : d2;
      const double i2 = 1.0 / safe;
      const double i6 = i2 * i2 * i2;
      energy += coeff12[j] * i6 * i6 - coeff6[j] * i6 + shift[j];
    }
  }
  if (ok) {
    output[(long long)center * n_axes + k] = energy;
    if (!isfinite(energy)) atomicMin(bad_center, center);
  }
}

__global__ void zero_kernel(int* __restrict__ p, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
}

__global__ void zero_u64_kernel(unsigned long long* __restrict__ p, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0ull;
}

}  // namespace

// ---------------------------------------------------------------------------
// Launchers
// ---------------------------------------------------------------------------
// Clearing flags with a kernel rather than hipMemsetAsync keeps every step of
// a run capturable in a HIP graph.
hipError_t launch_zero(int* p, int n, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(zero_kernel, dim3((n + 255) / 256), dim3(256), 0, s, p, n);
  return hipGetLastError();
}

hipError_t launch_zero_u64(unsigned long long* p, int n, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(zero_u64_kernel, dim3((n + 255) / 256), dim3(256), 0, s, p, n);
  return hipGetLastError();
}

hipError_t launch_tile_count(const Job* d_jobs, const Tile* d_tiles, int n_tiles_total,
                             const long long* d_seg_base, int* d_counts, hipStream_t s) {
  if (n_tiles_total <= 0) return hipSuccess;
  hipLaunchKernelGGL(tile_count_kernel, dim3(n_tiles_total), dim3(kListThreads), 0, s, d_jobs, d_tiles,
                     d_seg_base, d_counts);
  return hipGetLastError();
}

long long scan_scratch_size(long long n) { return (n + kScanChunk - 1) / kScanChunk + 1; }

hipError_t launch_scan(const int* d_counts, long long n, long long* d_offsets, long long* d_sums,
                       hipStream_t s) {
  const long long nb = (n + kScanChunk - 1) / kScanChunk;
  if (nb <= 1) {
    hipLaunchKernelGGL(scan_apply_kernel, dim3(1), dim3(kScanThreads), 0, s, d_counts, n,
                       (const long long*)nullptr, d_offsets);
    return hipGetLastError();
  }
  hipLaunchKernelGGL(scan_sums_kernel, dim3((unsigned)nb), dim3(kScanThreads), 0, s, d_counts, n,
                     d_sums);
  hipLaunchKernelGGL(scan_chunks_kernel, dim3(1), dim3(kScanThreads), 0, s, d_sums, nb, n,
                     d_offsets);
  hipLaunchKernelGGL(scan_apply_kernel, dim3((unsigned)nb), dim3(kScanThreads), 0, s, d_counts, n,
                     d_sums, d_offsets);
  return hipGetLastError();
}

hipError_t launch_tile_fill(const Job* d_jobs, const Tile* d_tiles, int n_tiles_total,
                            const long long* d_seg_base, const long long* d_offsets,
                            float4* d_list32, double* d_list64, hipStream_t s) {
  if (n_tiles_total <= 0) return hipSuccess;
  hipLaunchKernelGGL(tile_fill_kernel, dim3(n_tiles_total), dim3(kListThreads), 0, s, d_jobs, d_tiles,
                     d_seg_base, d_offsets, d_list32, d_list64);
  return hipGetLastError();
}

hipError_t launch_field(int precision, const Job* d_jobs, int n_jobs, const long long* d_job_work,
                        long long n_work, const long long* d_offsets, const float4* d_list32,
                        const double* d_list64, hipStream_t s, bool all_sym) {
  if (n_work <= 0) return hipSuccess;
  const dim3 grid((unsigned)((n_work + 3) / 4)), block(256);
  if (precision == PREC_MIXED) {
    if (all_sym)
      hipLaunchKernelGGL(field_mixed_kernel<true>, grid, block, 0, s, d_jobs, n_jobs, d_job_work,
                         n_work, d_offsets, d_list32, d_list64);
    else
      hipLaunchKernelGGL(field_mixed_kernel<false>, grid, block, 0, s, d_jobs, n_jobs, d_job_work,
                         n_work, d_offsets, d_list32, d_list64);
  } else if (precision == PREC_FP64_RCP) {
    hipLaunchKernelGGL(field_fp64_kernel<true>, grid, block, 0, s, d_jobs, n_jobs, d_job_work,
                       n_work, d_offsets, d_list64);
  } else {
    hipLaunchKernelGGL(field_fp64_kernel<false>, grid, block, 0, s, d_jobs, n_jobs, d_job_work,
                       n_work, d_offsets, d_list64);
  }
  return hipGetLastError();
}

hipError_t launch_field_persistent(const Job* d_jobs, int n_jobs, const long long* d_job_work,
                                  long long n_work, unsigned long long* d_counters, int n_blocks,
                                  const long long* d_offsets, const float4* d_list32,
                                  const double* d_list64, hipStream_t s, bool all_sym) {
  if (n_work <= 0) return hipSuccess;
  hipError_t e = launch_zero_u64(d_counters, 8, s);
  if (e != hipSuccess) return e;
  const dim3 grid((unsigned)n_blocks), block(256);
  if (all_sym)
    hipLaunchKernelGGL(field_mixed_persistent<true>, grid, block, 0, s, d_jobs, n_jobs, d_job_work,
                       n_work, d_counters, d_offsets, d_list32, d_list64);
  else
    hipLaunchKernelGGL(field_mixed_persistent<false>, grid, block, 0, s, d_jobs, n_jobs, d_job_work,
                       n_work, d_counters, d_offsets, d_list32, d_list64);
  return hipGetLastError();
}

hipError_t launch_census(const Job* d_jobs, int n_jobs, const long long* d_job_work,
                         long long n_work, const long long* d_offsets, const double* d_list64,
                         hipStream_t s) {
  if (n_work <= 0) return hipSuccess;
  hipLaunchKernelGGL(census_kernel, dim3((unsigned)((n_work + 3) / 4)), dim3(256), 0, s, d_jobs,
                     n_jobs, d_job_work, n_work, d_offsets, d_list64);
  return hipGetLastError();
}

hipError_t launch_marginal(const Job* d_jobs, int n_jobs, const long long* d_job_first,
                           long long n_total, hipStream_t s) {
  if (n_total <= 0) return hipSuccess;
  hipLaunchKernelGGL(marginal_kernel, dim3((unsigned)n_total), dim3(64), 0, s, d_jobs, n_jobs,
                     d_job_first, n_total);
  return hipGetLastError();
}

hipError_t launch_legacy(int n_centers, int n_axes, double cutoff2, const double* centers,
                         const double* axes, const double* images, const double* offsets,
                         const double* coeff12, const double* coeff6, const double* shift,
                         const double* floor2, const long long* starts, const int* indices,
                         double* output, int* bad_center, hipStream_t s) {
  const int n_groups = (n_axes + 63) / 64;
  const long long waves = (long long)n_centers * n_groups;
  if (waves <= 0) return hipSuccess;
  hipLaunchKernelGGL(legacy_kernel, dim3((unsigned)((waves + 3) / 4)), dim3(256), 0, s, n_centers,
                     n_axes, n_groups, cutoff2, centers, axes, images, offsets, coeff12, coeff6,
                     shift, floor2, starts, indices, output, bad_center);
  return hipGetLastError();
}

}  // namespace sv
