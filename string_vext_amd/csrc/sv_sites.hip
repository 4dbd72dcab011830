// Independent poses on gfx950, evaluated site by site.
//
// A pose (center c, director u) of the two-site guest has the energy
//   E(c, u) = phi(c + o_0 u) + phi(c + o_1 u),
// where phi(p) is the single-site sum of the shifted, floored LJ term over
// every framework image within rc of p (the pair sum of
// native/vext/kernel.cpp:13-44 split by site), and its gradient is
//   dE/dc = g_0 + g_1,   dE/du = o_0 g_0 + o_1 g_1,   g_s = grad phi(p_s).
// With the poses independent, each lane takes one site position; its two
// sites land in different places, so nothing is shared between the sites of
// a pose and the pose sums are formed by a separate combine pass.
//
// Work decomposition:
//   * the site positions are binned on the GPU into a fixed grid of bins over
//     the cell; each bin is a tile whose candidate list (built once per grid)
//     holds every image within (bin circumradius + rc) of its centre, split per
//     LJ type into class 0 -- images a member site can bring within the
//     near-contact radius -- and class 1, the rest;
//   * within a bin the sites are ordered by the Morton index of their cell of
//     the bin's 4 x 4 x 4 sub-grid, and the cells are grouped greedily into
//     runs of whole cells holding at most 64 sites (a cell of more than 64 is
//     cut into equal parts), so a run covers a compact part of its bin;
//   * a block of kSiteWaves waves takes one run (one lane per site).  The
//     waves deal the list's 64-entry batches round robin over all (type,
//     class) segments; the lane that loads an entry keeps it only if it lies
//     within rc of the run's bounding box, and the kept entries are
//     compacted into LDS (ballot + mbcnt), where every lane reads them by
//     broadcast in unrolled groups; the waves' sums are added in wave order.
//
// Mixed precision: list and site coordinates are tile-relative Cartesian
// scaled by 1/rc, so x = |e - p|^2 is d^2 / rc^2 and the clamp bit of the fma
// that forms it bounds it to [0, 1]: a pair beyond the cutoff contributes the
// clamped term -shift, which the +shift added for every FP32 candidate
// cancels (class-1 pairs need no other test).  Class-0 candidates inside a
// lane's near-contact radius (where a term can exceed the plan's bound or the
// wall is steep, and always within reach of the contact floor) are forced to
// the clamp in FP32, marked in a per-lane bit mask and evaluated in FP64 with
// the reference formula after their group, each lane walking its own mask.
// FP32 partials go into FP64 accumulators after every group of kSiteGroup
// class-0 candidates and after every batch of at most 64 class-1 candidates.
#include "sv_device.h"

namespace sv {
namespace {

constexpr int kSiteWaves = 4;
constexpr int kSiteGroup = 16;    // candidates per unrolled FP32 group
constexpr int kSiteAcc = 4;     // doubles handed over per lane: E, g_x, g_y, g_z

// FP64 term of one (site, image) pair, q = image - site: energy and the
// gradient 2 dphi/d(d^2) (p - a) = -2 dphi/d(d^2) q (zero inside the floor,
// where the floored term is constant).  1/d^2 from v_rcp_f64 and one Newton
// step.
template <bool GRAD>
__device__ __forceinline__ void site_pair64(double qx, double qy, double qz, double rc2,
                                            double c12, double c6, double sh, double fl2,
                                            double& E, double G[3]) {
  const double d2 = qx * qx + qy * qy + qz * qz;
  if (d2 < rc2) {
    const double safe = fmax(d2, fl2);
    const double inv = inv_fast(safe);
    const double i6 = inv * inv * inv;
    E += c12 * i6 * i6 - c6 * i6 + sh;
    if (GRAD && d2 >= fl2) {
      const double g2 = -2.0 * (-6.0 * c12 * i6 * i6 + 3.0 * c6 * i6) * inv;
      G[0] = fma(g2, qx, G[0]);
      G[1] = fma(g2, qy, G[1]);
      G[2] = fma(g2, qz, G[2]);
    }
  }
}

// First batch of segment [beg, end) that dealer `who` of `dealers` takes when
// batches are dealt round robin in the order of all segments, `dealt` batches
// having gone to earlier segments; advances `dealt`.
__device__ __forceinline__ long long first_batch(long long beg, long long end, int who,
                                                 int dealers, long long& dealt) {
  const long long batches = (end - beg + 63) / 64;
  const int skip = (int)((who - dealt % dealers + dealers) % dealers);
  dealt += batches;
  return beg + 64LL * skip;
}

__device__ __forceinline__ float wave_min_f(float v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fminf(v, __shfl_xor(v, o, 64));
  return v;
}
__device__ __forceinline__ float wave_max_f(float v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o, 64));
  return v;
}

// FP32 partial sums of one lane over a run of staged candidates.
struct SitePartial {
  float e, g0, g1, g2;
};

// One staged candidate against the lane's site p (both in units of rc).
// NEAR (class 0): a candidate within the near-contact radius (x < nsx) takes
// the clamped FP32 term and returns true for its FP64 term.
template <bool NEAR, bool GRAD, bool ENERGY>
__device__ __forceinline__ bool site_term32(const float4 e, float pxf, float pyf, float pzf,
                                            float s6, float m2s6, float nsx, SitePartial& a) {
  const float rx = e.x - pxf, ry = e.y - pyf, rz = e.z - pzf;
  float x;
  bool near = false;
  if (NEAR) {
    const float xr = fmaf(rz, rz, fmaf(ry, ry, rx * rx));
    near = xr < nsx;
    x = near ? 1.0f : fminf(xr, 1.0f);
  } else {
    x = __builtin_amdgcn_fmed3f(fmaf(rz, rz, fmaf(ry, ry, rx * rx)), 0.0f, 1.0f);
  }
  const float i = __builtin_amdgcn_rcpf(x);
  const float i3 = i * i * i;
  if (ENERGY) a.e = fmaf(i3, fmaf(i3, s6, -1.0f), a.e);
  if (GRAD) {
    const float gg = x < 1.0f ? i3 * i * fmaf(m2s6, i3, 1.0f) : 0.0f;
    a.g0 = fmaf(gg, rx, a.g0);
    a.g1 = fmaf(gg, ry, a.g1);
    a.g2 = fmaf(gg, rz, a.g2);
  }
  return near;
}

// Per-type coefficients of the site form.
struct SiteType {
  double c12, c6, sh, fl2, c6s, kg;
  float s6, m2s6, nsx;
};

// All n staged candidates of one batch, in unrolled groups of kSiteGroup.
// Class 1 (NEAR false): one FP32 partial over the batch (its terms are small:
// every pair is beyond the near-contact radius), flushed into the FP64 sums
// once.  Class 0: partials flushed per group (its FP32 terms reach the
// near-contact bound), and the near-contact terms in FP64 after the batch,
// each lane walking the candidates its mask marks.
template <bool NEAR, bool GRAD, bool ENERGY>
__device__ __forceinline__ void site_batch32(const float4* __restrict__ s_e,
                                             const double (*__restrict__ s_d)[64], int n,
                                             float pxf, float pyf, float pzf, double px, double py,
                                             double pz, double rc2, const SiteType& ty, double& E,
                                             double G[3]) {
  SitePartial a = {0.f, 0.f, 0.f, 0.f};
  unsigned long long marks = 0ull;
  for (int q0 = 0; q0 < n; q0 += kSiteGroup) {
    unsigned m16 = 0u;
    if (q0 + kSiteGroup <= n) {
#pragma unroll
      for (int j = 0; j < kSiteGroup; ++j)
        if (site_term32<NEAR, GRAD, ENERGY>(s_e[q0 + j], pxf, pyf, pzf, ty.s6, ty.m2s6, ty.nsx, a))
          m16 |= 1u << j;
    } else {
      for (int j = 0; q0 + j < n; ++j)
        if (site_term32<NEAR, GRAD, ENERGY>(s_e[q0 + j], pxf, pyf, pzf, ty.s6, ty.m2s6, ty.nsx, a))
          m16 |= 1u << j;
    }
    if (NEAR) {
      marks |= (unsigned long long)m16 << q0;
      if (ENERGY) E = fma(ty.c6s, (double)a.e, E);
      if (GRAD) {
        G[0] = fma(ty.kg, (double)a.g0, G[0]);
        G[1] = fma(ty.kg, (double)a.g1, G[1]);
        G[2] = fma(ty.kg, (double)a.g2, G[2]);
      }
      a = SitePartial{0.f, 0.f, 0.f, 0.f};
    }
  }
  if (!NEAR) {
    if (ENERGY) E = fma(ty.c6s, (double)a.e, E);
    if (GRAD) {
      G[0] = fma(ty.kg, (double)a.g0, G[0]);
      G[1] = fma(ty.kg, (double)a.g1, G[1]);
      G[2] = fma(ty.kg, (double)a.g2, G[2]);
    }
  }
  if (NEAR) {
    while (marks) {
      const int q = __builtin_ctzll(marks);
      marks &= marks - 1ull;
      site_pair64<GRAD>(s_d[0][q] - px, s_d[1][q] - py, s_d[2][q] - pz, rc2, ty.c12, ty.c6,
                        ty.sh, ty.fl2, E, G);
    }
  }
}

template <bool MIXED, bool GRAD, bool ENERGY>
__global__ __launch_bounds__(64 * kSiteWaves) void site_kernel(const Job* __restrict__ jobs,
                                                               const SiteRun* __restrict__ runs,
                                                               const int* __restrict__ n_runs,
                                                               const long long* __restrict__ offsets,
                                                               const float4* __restrict__ list32,
                                                               const double* __restrict__ list64,
                                                               const SiteArrays sa) {
  __shared__ float4 s_e_all[kSiteWaves][64];
  __shared__ double s_d_all[kSiteWaves][3][64];
  __shared__ double s_acc[kSiteWaves - 1][kSiteAcc][64];
  // Block = (run, part): the run's batches are dealt over parts x kSiteWaves
  // waves, and each part writes its own partial sums.  The grid bounds the
  // run table built on the device: the first n_runs x parts blocks take the
  // work (remapped among themselves so that consecutive runs share an XCD)
  // and are dispatched first; the rest return at once.
  const int parts = sa.parts;
  const long long active = (long long)*n_runs * parts;
  if ((long long)blockIdx.x >= active) return;
  const long long logical = xcd_block(blockIdx.x, active);
  const SiteRun run = runs[logical / parts];
  const int part = (int)(logical % parts);
  const Job& J = jobs[0];
  const int lane = threadIdx.x & 63;
  const int wv = uniform(threadIdx.x >> 6);
  const int dealer = part * kSiteWaves + wv, dealers = parts * kSiteWaves;
  float4* s_e = s_e_all[wv];
  double(*s_d)[64] = s_d_all[wv];
  const bool ok = lane < run.count;
  const int slot = run.first + (ok ? lane : 0);
  const double px = sa.site_rel[3 * slot], py = sa.site_rel[3 * slot + 1],
               pz = sa.site_rel[3 * slot + 2];
  const double scale = J.list_scale;
  float pxf = (float)(px * scale), pyf = (float)(py * scale), pzf = (float)(pz * scale);
  // Bounding box of the run's sites (units of rc; lanes past the count repeat
  // the first site, NaN sites of invalid poses drop out of fmin/fmax): an
  // entry whose distance to the box exceeds 1 is within rc of no site (per
  // axis, the distance to the box is at most the distance to any site in
  // it).  The margin 1e-4 covers the FP32 rounding of the box.  A run of whole
  // cells has a fixed set of sites whatever slots the binning gave them, so
  // its box and every lane's result are fixed too; a part of an overfull cell
  // keeps the bin's whole list.
  const float lox = wave_min_f(pxf), hix = wave_max_f(pxf);
  const float loy = wave_min_f(pyf), hiy = wave_max_f(pyf);
  const float loz = wave_min_f(pzf), hiz = wave_max_f(pzf);
  const float cxf = uniform_f(0.5f * (lox + hix)), hx = uniform_f(0.5f * (hix - lox));
  const float cyf = uniform_f(0.5f * (loy + hiy)), hy = uniform_f(0.5f * (hiy - loy));
  const float czf = uniform_f(0.5f * (loz + hiz)), hz = uniform_f(0.5f * (hiz - loz));
  const float reach2 = run.whole ? (1.0f + 1e-4f) * (1.0f + 1e-4f) : INFINITY;
  // keep the FP32 site in registers (left alone, the compiler converts it again per batch)
  asm volatile("" : "+v"(pxf), "+v"(pyf), "+v"(pzf));
  const double rc2 = J.rc2;
  const long long seg0 = J.seg_base + (long long)run.bin * J.n_types * 2;
  double E = 0.0, G[3] = {0.0, 0.0, 0.0};
  long long dealt = 0;
  for (int t = 0; t < J.n_types; ++t) {
    SiteType ty;
    ty.c12 = J.c12[t];
    ty.c6 = J.c6[t];
    ty.sh = J.shift[t];
    ty.fl2 = J.floor2[t];
    // phi = c6s i^3 (s6s i^3 - 1), i = 1/x; grad_p phi = -(6 c6s / rc) i^4 (1 - 2 s6s i^3) r'
    ty.c6s = J.c6s[t];
    ty.kg = -6.0 * scale * ty.c6s;
    ty.s6 = J.s6sf[t];
    ty.m2s6 = -2.0f * ty.s6;
    ty.nsx = J.nsite2f[t];
    long long n_fp = 0;   // FP32 candidates of this type: +shift each
#pragma unroll
    for (int cls = 0; cls < 2; ++cls) {
      const long long beg = offsets[seg0 + 2 * t + cls], end = offsets[seg0 + 2 * t + cls + 1];
      // FP64 coordinates are staged for class 0 (near-contact terms) and for the FP64 form
      const bool need64 = !MIXED || cls == 0;
      for (long long base = first_batch(beg, end, dealer, dealers, dealt); base < end;
           base += 64LL * dealers) {
        const long long idx = base + lane;
        bool keep = false;
        float4 e = make_float4(0.f, 0.f, 0.f, 0.f);
        if (idx < end) {
          e = list32[idx];
          const float ex = fmaxf(fabsf(e.x - cxf) - hx, 0.0f);
          const float ey = fmaxf(fabsf(e.y - cyf) - hy, 0.0f);
          const float ez = fmaxf(fabsf(e.z - czf) - hz, 0.0f);
          keep = fmaf(ez, ez, fmaf(ey, ey, ex * ex)) <= reach2;
        }
        const unsigned long long kept = __ballot(keep);
        if (keep) {
          const int at = lanes_below(kept);
          s_e[at] = e;
          if (need64) {
            s_d[0][at] = list64[3 * idx];
            s_d[1][at] = list64[3 * idx + 1];
            s_d[2][at] = list64[3 * idx + 2];
          }
        }
        wave_sync();
        const int n = __popcll(kept);
        if (MIXED) {
          n_fp += n;
          if (cls == 0)
            site_batch32<true, GRAD, ENERGY>(s_e, s_d, n, pxf, pyf, pzf, px, py, pz, rc2, ty, E, G);
          else
            site_batch32<false, GRAD, ENERGY>(s_e, s_d, n, pxf, pyf, pzf, px, py, pz, rc2, ty, E, G);
        } else {
          for (int q = 0; q < n; ++q)
            site_pair64<GRAD>(s_d[0][q] - px, s_d[1][q] - py, s_d[2][q] - pz, rc2, ty.c12, ty.c6,
                              ty.sh, ty.fl2, E, G);
        }
        wave_sync();
      }
    }
    if (MIXED && ENERGY) E = fma((double)n_fp, ty.sh, E);
  }
  // Waves 1.. hand their sums to wave 0, which adds them in wave order.
  if (wv > 0) {
    double(*o)[64] = s_acc[wv - 1];
    if (ENERGY) o[0][lane] = E;
    if (GRAD) {
      o[1][lane] = G[0];
      o[2][lane] = G[1];
      o[3][lane] = G[2];
    }
  }
  __syncthreads();
  if (wv > 0 || !ok) return;
#pragma unroll
  for (int w = 0; w < kSiteWaves - 1; ++w) {
    if (ENERGY) E += s_acc[w][0][lane];
    if (GRAD) {
      G[0] += s_acc[w][1][lane];
      G[1] += s_acc[w][2][lane];
      G[2] += s_acc[w][3][lane];
    }
  }
  // an invalid pose (nonfinite center, zero or nonfinite director) was binned
  // with a NaN site position; its results are NaN
  if (!(isfinite(px) && isfinite(py) && isfinite(pz))) E = G[0] = G[1] = G[2] = __builtin_nan("");
  const long long k = sa.site_id[slot] + (long long)part * sa.part_stride;
  if (ENERGY) sa.site_e[k] = E;
  if (GRAD) {
    sa.site_g[3 * k] = G[0];
    sa.site_g[3 * k + 1] = G[1];
    sa.site_g[3 * k + 2] = G[2];
  }
}

// Pose sums from the site results, each the sum of its parts in part order:
// E = E_0 + E_1, dE/dc = g(G_0 + G_1), dE/du = g(o_0 G_0 + o_1 G_1) with g the
// upstream gradient (or 1).
__global__ __launch_bounds__(256) void site_combine_kernel(
    int n, int parts, long long part_stride, double o0, double o1,
    const double* __restrict__ site_e, const double* __restrict__ site_g,
    const double* __restrict__ grad_out, double* __restrict__ energy,
    double* __restrict__ grad_center, double* __restrict__ grad_axis, int* __restrict__ bad_flag) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  bool bad = false;
  if (energy) {
    double e0 = site_e[2 * (long long)i], e1 = site_e[2 * (long long)i + 1];
    for (int p = 1; p < parts; ++p) {
      e0 += site_e[p * part_stride + 2 * (long long)i];
      e1 += site_e[p * part_stride + 2 * (long long)i + 1];
    }
    const double e = e0 + e1;
    energy[i] = e;
    bad = !isfinite(e);
  }
  if (grad_center || grad_axis) {
    const double go = grad_out ? grad_out[i] : 1.0;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      double g0 = site_g[6 * (long long)i + k], g1 = site_g[6 * (long long)i + 3 + k];
      for (int p = 1; p < parts; ++p) {
        g0 += site_g[3 * p * part_stride + 6 * (long long)i + k];
        g1 += site_g[3 * p * part_stride + 6 * (long long)i + 3 + k];
      }
      const double gc = go * (g0 + g1);
      const double gu = go * (o0 * g0 + o1 * g1);
      if (grad_center) grad_center[3 * (long long)i + k] = gc;
      if (grad_axis) grad_axis[3 * (long long)i + k] = gu;
      bad = bad || !isfinite(gc) || !isfinite(gu);
    }
  }
  if (bad) atomicOr(bad_flag, 1);
}

// Site positions p = c + o_s u of pose i = k / 2, site s = k % 2, wrapped into
// the cell; false for a nonfinite center or a nonfinite or zero director.
__device__ __forceinline__ bool site_frac(const double* __restrict__ centers,
                                         const double* __restrict__ axes, double o0, double o1,
                                         const Mat3& Hinv, long long k, double f[3]) {
  const long long i = k >> 1;
  const double o = (k & 1) ? o1 : o0;
  const double u0 = axes[3 * i], u1 = axes[3 * i + 1], u2 = axes[3 * i + 2];
  const double nn = u0 * u0 + u1 * u1 + u2 * u2;
  const double p[3] = {centers[3 * i] + o * u0, centers[3 * i + 1] + o * u1,
                       centers[3 * i + 2] + o * u2};
  bool ok = isfinite(p[0]) && isfinite(p[1]) && isfinite(p[2]) && isfinite(nn) && nn > 0.0;
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const double x = p[0] * Hinv.m[a] + p[1] * Hinv.m[3 + a] + p[2] * Hinv.m[6 + a];
    f[a] = x - floor(x);
    ok = ok && f[a] >= 0.0 && f[a] <= 1.0;   // false for an overflowed (nonfinite) product
  }
  return ok;
}

// Bin of a wrapped fractional position and the Morton index of its cell of
// the bin's 4 x 4 x 4 sub-grid (axes order & 3, (order >> 2) & 3,
// (order >> 4) & 3 from the most significant bit down).
__device__ __forceinline__ int site_cell_of(const double f[3], int m0, int m1, int m2, int order) {
  const int m[3] = {m0, m1, m2};
  int idx[3], quarter[3];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const double t = f[a] * m[a];
    const int q = (int)t;
    idx[a] = q < m[a] - 1 ? q : m[a] - 1;
    const int c = (int)((t - idx[a]) * 4.0);
    quarter[a] = c < 3 ? c : 3;
  }
  const int u0 = quarter[order & 3], u1 = quarter[(order >> 2) & 3],
            u2 = quarter[(order >> 4) & 3];
  const int sub = ((u0 >> 1) << 5) | ((u1 >> 1) << 4) | ((u2 >> 1) << 3) | ((u0 & 1) << 2) |
                  ((u1 & 1) << 1) | (u2 & 1);
  return 64 * ((idx[0] * m1 + idx[1]) * m2 + idx[2]) + sub;
}

// Cell + rank of every site by one atomic per site; invalid poses go to cell
// 0 and set flag 2.  Also clears the running totals of sites and runs that
// the layout kernel, launched after this one, reserves its ranges from.
__global__ __launch_bounds__(256) void site_bin_kernel(long long n_sites,
                                                       const double* __restrict__ centers,
                                                       const double* __restrict__ axes, double o0,
                                                       double o1, Mat3 Hinv, int m0, int m1, int m2,
                                                       int order, int* __restrict__ cell,
                                                       int* __restrict__ rank,
                                                       int* __restrict__ cell_counts,
                                                       int* __restrict__ bad,
                                                       unsigned long long* __restrict__ totals) {
  const long long k = (long long)blockIdx.x * 256 + threadIdx.x;
  if (k == 0) *totals = 0ull;   // runs (low word) and slots (high word) taken
  if (k >= n_sites) return;
  double f[3];
  int c = 0;
  if (site_frac(centers, axes, o0, o1, Hinv, k, f)) {
    c = site_cell_of(f, m0, m1, m2, order);
  } else {
    atomicOr(bad, 2);
  }
  cell[k] = c;
  rank[k] = atomicAdd(&cell_counts[c], 1);
}

// Runs of one bin's 64 cells, in order: consecutive cells while their sites
// fit in 64 (greedy); a cell of more than 64 sites alone, cut into
// ceil(n / 64) runs of nearly equal size.  One wave per bin, lane q = cell q:
// with the bin's inclusive prefix sums `incl`, the run that starts at cell q
// ends before the first cell j with incl[j] - start(q) > 64 (found by binary
// search over the lanes), so the runs are the chain 0 -> next(0) -> ..., which
// a wave-uniform loop walks (a few steps per bin).  Returns the number of
// runs; with `out`, lane 0 writes them from slot first_slot on.
__device__ __forceinline__ int bin_runs(int c, int b, int first_slot, SiteRun* __restrict__ out) {
  const int lane = threadIdx.x & 63;
  int incl = c;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const int y = __shfl_up(incl, o, 64);
    if (lane >= o) incl += y;
  }
  const int start = incl - c;                    // slot of cell q within the bin
  const int total = __shfl(incl, 63, 64);
  // last cell whose inclusive sum stays within start + 64 (incl is nondecreasing)
  const int target = start + 64;
  int pos = -1;
#pragma unroll
  for (int step = 64; step > 0; step >>= 1) {
    const int cand = pos + step;
    const int v = __shfl(incl, cand < 64 ? cand : 63, 64);
    if (cand < 64 && v <= target) pos = cand;
  }
  const int next = pos + 1;                      // == lane: cell q alone overflows
  const bool writer = out && lane == 0;
  int n_runs = 0;
  for (int q = 0; q < 64;) {
    const int nq = __builtin_amdgcn_readlane(next, q);
    const int sq = __builtin_amdgcn_readlane(start, q);
    if (nq == q) {
      const int cq = __builtin_amdgcn_readlane(c, q);
      const int k = (cq + 63) >> 6;
      for (int j = 0; j < k; ++j) {
        const int lo = (int)((long long)cq * j / k), hi = (int)((long long)cq * (j + 1) / k);
        if (writer) out[n_runs] = SiteRun{b, first_slot + sq + lo, hi - lo, 0};
        ++n_runs;
      }
      q += 1;
      continue;
    }
    const int end = nq < 64 ? __builtin_amdgcn_readlane(start, nq) : total;
    if (end > sq) {
      if (writer) out[n_runs] = SiteRun{b, first_slot + sq, end - sq, 1};
      ++n_runs;
    }
    q = nq;
  }
  return n_runs;
}

// Slots and runs of every bin (one wave per bin, kLayoutWaves bins per
// block).  A block reserves the range of slots and the range of the run
// table for its bins with one 64-bit atomic on the packed running totals
// (runs in the low word, slots in the high word; cleared by the binning
// kernel), so the blocks take their ranges in no fixed order while the bins
// of a block stay in bin order; each bin's slots and runs are contiguous.  No
// result depends on that order: a site's sums depend only on the members of
// its run (which fix the run's box and culled list), and they are gathered
// by site index.  The low word ends as the length of the run table.
constexpr int kLayoutWaves = 16;

__global__ __launch_bounds__(64 * kLayoutWaves) void site_layout_kernel(
    const int* __restrict__ cell_counts, int n_bins, unsigned long long* __restrict__ totals,
    int* __restrict__ first, SiteRun* __restrict__ runs) {
  __shared__ int s_sites[kLayoutWaves], s_runs[kLayoutWaves];
  __shared__ unsigned long long s_base;
  const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int b = blockIdx.x * kLayoutWaves + wv;
  int c = 0, incl = 0, total = 0, k = 0;
  if (b < n_bins) {
    c = cell_counts[64 * (long long)b + lane];
    incl = c;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      const int y = __shfl_up(incl, o, 64);
      if (lane >= o) incl += y;
    }
    total = __shfl(incl, 63, 64);
    if (total > 0) k = bin_runs(c, b, 0, nullptr);
  }
  if (lane == 0) {
    s_sites[wv] = total;
    s_runs[wv] = k;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long sites = 0, nr = 0;
    for (int w = 0; w < kLayoutWaves; ++w) {
      sites += (unsigned)s_sites[w];
      nr += (unsigned)s_runs[w];
    }
    s_base = (sites | nr) ? atomicAdd(totals, (sites << 32) | nr) : 0ull;
  }
  __syncthreads();
  if (b >= n_bins || total == 0) return;   // wave-uniform
  int site0 = (int)(s_base >> 32), run0 = (int)(s_base & 0xffffffffull);
  for (int w = 0; w < wv; ++w) {
    site0 += s_sites[w];
    run0 += s_runs[w];
  }
  first[64 * (long long)b + lane] = site0 + incl - c;
  bin_runs(c, b, site0, runs + run0);
}

// Tile-sorted slots: Cartesian offset of the site from its bin centre, from
// fractional differences (NaN for invalid poses), and the site index.
__global__ __launch_bounds__(256) void site_scatter_kernel(
    long long n_sites, const double* __restrict__ centers, const double* __restrict__ axes,
    double o0, double o1, Mat3 H, Mat3 Hinv, int m0, int m1, int m2,
    const int* __restrict__ cell, const int* __restrict__ rank, const int* __restrict__ first,
    int* __restrict__ cell_counts, double* __restrict__ site_rel, int* __restrict__ site_id) {
  const long long k = (long long)blockIdx.x * 256 + threadIdx.x;
  if (k >= n_sites) return;
  double f[3];
  const bool ok = site_frac(centers, axes, o0, o1, Hinv, k, f);
  const int c = cell[k];
  const long long slot = (long long)first[c] + rank[k];
  // the counts have been scanned: clear them for the next binning (every
  // cell with a site is cleared by its sites, the others were never counted)
  cell_counts[c] = 0;
  const int b = c >> 6;
  const int idx[3] = {b / (m1 * m2), (b / m2) % m1, b % m2};
  const int m[3] = {m0, m1, m2};
  double d[3];
#pragma unroll
  for (int a = 0; a < 3; ++a) d[a] = f[a] - (idx[a] + 0.5) / m[a];
#pragma unroll
  for (int a = 0; a < 3; ++a)
    site_rel[3 * slot + a] = ok ? d[0] * H.m[a] + d[1] * H.m[3 + a] + d[2] * H.m[6 + a]
                                : __builtin_nan("");
  site_id[slot] = (int)k;
}

// Exact in-cutoff (site, image) pairs of the binned sites (FP64 test): the
// pair-eval count of the throughput metric.  One wave per run.
__global__ __launch_bounds__(64) void site_census_kernel(const Job* __restrict__ jobs,
                                                         const SiteRun* __restrict__ runs,
                                                         const int* __restrict__ n_runs,
                                                         const long long* __restrict__ offsets,
                                                         const double* __restrict__ list64,
                                                         const double* __restrict__ site_rel,
                                                         unsigned long long* __restrict__ total) {
  if ((int)blockIdx.x >= *n_runs) return;
  const SiteRun run = runs[blockIdx.x];
  const Job& J = jobs[0];
  const int lane = threadIdx.x;
  const bool ok = lane < run.count;
  const int slot = run.first + (ok ? lane : 0);
  const double px = site_rel[3 * slot], py = site_rel[3 * slot + 1], pz = site_rel[3 * slot + 2];
  const long long seg0 = J.seg_base + (long long)run.bin * J.n_types * 2;
  const long long beg = offsets[seg0], end = offsets[seg0 + 2 * J.n_types];
  unsigned long long pairs = 0;
  for (long long q = beg; q < end; ++q) {
    const double dx = list64[3 * q] - px, dy = list64[3 * q + 1] - py, dz = list64[3 * q + 2] - pz;
    pairs += dx * dx + dy * dy + dz * dz < J.rc2;
  }
  if (!ok) pairs = 0;
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) pairs += __shfl_xor(pairs, o, 64);
  if (lane == 0) atomicAdd(total, pairs);
}

}  // namespace

hipError_t launch_site_bin(long long n_sites, const double* centers, const double* axes,
                           double o0, double o1, const Mat3& Hinv, const int m[3], int order,
                           int* cell, int* rank, int* cell_counts, int* bad,
                           unsigned long long* totals, hipStream_t s) {
  if (n_sites <= 0) return hipSuccess;
  hipLaunchKernelGGL(site_bin_kernel, dim3((unsigned)((n_sites + 255) / 256)), dim3(256), 0, s,
                     n_sites, centers, axes, o0, o1, Hinv, m[0], m[1], m[2], order, cell, rank,
                     cell_counts, bad, totals);
  return hipGetLastError();
}

// Two consecutive runs of whole cells hold more than 64 sites together (the
// first was closed because the second's first cell did not fit); a cell of
// c > 64 sites makes ceil(c / 64) runs of more than 32 sites and may close
// one short run before it.  So a bin of n sites has at most 2 n / 64 + 1 runs
// of whole cells plus, per overfull cell (fewer than n / 64), its runs and
// one more: in all fewer than 3 n / 64 + 1.
long long site_runs_bound(int n_bins, long long n_sites) {
  return (long long)n_bins + 3 * ((n_sites + 63) / 64);
}

hipError_t launch_site_layout(const int* d_cell_counts, int n_bins, unsigned long long* d_totals,
                              int* d_first, SiteRun* d_runs, hipStream_t s) {
  if (n_bins <= 0) return hipSuccess;
  hipLaunchKernelGGL(site_layout_kernel, dim3((unsigned)((n_bins + kLayoutWaves - 1) / kLayoutWaves)),
                     dim3(64 * kLayoutWaves), 0, s, d_cell_counts, n_bins, d_totals, d_first,
                     d_runs);
  return hipGetLastError();
}

hipError_t launch_site_scatter(long long n_sites, const double* centers, const double* axes,
                               double o0, double o1, const Mat3& H, const Mat3& Hinv,
                               const int m[3], const int* cell, const int* rank, const int* first,
                               int* cell_counts, double* site_rel, int* site_id, hipStream_t s) {
  if (n_sites <= 0) return hipSuccess;
  hipLaunchKernelGGL(site_scatter_kernel, dim3((unsigned)((n_sites + 255) / 256)), dim3(256), 0,
                     s, n_sites, centers, axes, o0, o1, H, Hinv, m[0], m[1], m[2], cell, rank,
                     first, cell_counts, site_rel, site_id);
  return hipGetLastError();
}

hipError_t launch_site(int precision, bool energy, bool grad, const Job* d_jobs,
                       const SiteRun* d_runs, int max_runs, const int* d_n_runs,
                       const long long* d_offsets, const float4* d_list32, const double* d_list64,
                       const SiteArrays& a, hipStream_t s) {
  if (max_runs <= 0 || (!energy && !grad)) return hipSuccess;
  const dim3 g((unsigned)((long long)max_runs * a.parts)), b(64 * kSiteWaves);
#define SV_SITE(M, G, E)                                                                   \
  hipLaunchKernelGGL((site_kernel<M, G, E>), g, b, 0, s, d_jobs, d_runs, d_n_runs,        \
                     d_offsets, d_list32, d_list64, a)
  if (precision == PREC_MIXED) {
    if (!grad) SV_SITE(true, false, true);
    else if (!energy) SV_SITE(true, true, false);
    else SV_SITE(true, true, true);
  } else {
    if (!grad) SV_SITE(false, false, true);
    else if (!energy) SV_SITE(false, true, false);
    else SV_SITE(false, true, true);
  }
#undef SV_SITE
  return hipGetLastError();
}

hipError_t launch_site_combine(int n_poses, const SiteArrays& a, double o0, double o1,
                               const double* grad_out, double* energy, double* grad_center,
                               double* grad_axis, int* bad, hipStream_t s) {
  if (n_poses <= 0) return hipSuccess;
  hipLaunchKernelGGL(site_combine_kernel, dim3((unsigned)((n_poses + 255) / 256)), dim3(256), 0,
                     s, n_poses, a.parts, a.part_stride, o0, o1, a.site_e, a.site_g, grad_out,
                     energy, grad_center, grad_axis, bad);
  return hipGetLastError();
}

hipError_t launch_site_census(const Job* d_jobs, const SiteRun* d_runs, int max_runs,
                              const int* d_n_runs, const long long* d_offsets,
                              const double* d_list64, const double* site_rel,
                              unsigned long long* pairs, hipStream_t s) {
  if (max_runs <= 0) return hipSuccess;
  hipLaunchKernelGGL(site_census_kernel, dim3(max_runs), dim3(64), 0, s, d_jobs, d_runs, d_n_runs,
                     d_offsets, d_list64, site_rel, pairs);
  return hipGetLastError();
}

}  // namespace sv
