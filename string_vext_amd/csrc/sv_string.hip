// Batched String method on the GPU.
//
// Behaviour follows the reference String solver's iteration
// (native/string/source_code/RC_main.cu:1914-2393, kernels of global.h:
// 1281-2240), with its forward-difference forces replaced by the analytic
// gradient of the same pose energy:
//   1. force of image i: g_k = delta_k dV/dx_k on x = (a, b, c, alpha, beta,
//      gamma), delta = (move_frac x3, move_angle x3), each floored to a 1e-15
//      grid (global.h:1281-1324; a forward difference V(x + delta_k) - V(x)
//      is delta_k dV/dx_k to first order);
//   2. step with separate translational / rotational norms: s1 = s0 - delta g
//      when |g| < delta, else s0 - delta g/|g| (global.h:1328-1443);
//   3. endpoints fixed along the hop direction, the last image a periodic
//      replica of the first, interior images outside [0, 1] along the hop
//      direction replaced by the nearest in-range neighbour (global.h:1456-1647;
//      applied in image order here, the reference does it in parallel);
//   4. arc-length reparametrization of (a, b, c) by linear interpolation at
//      equal Cartesian arc-length fractions, angles unchanged (global.h:1653-1846);
//   5. angles of images whose energy rose are reverted, angles smoothed with
//      coefficient 1e-4 and the reference's periodic-unwrap branches
//      (global.h:1854-2085); in a mixed run the comparison is taken on FP64
//      energies when the mixed ones are too close to tell;
//   6. every 200 iterations the string stops once the summed per-image
//      coordinate and orientation changes fall below 30% of
//      move x n_images x sqrt(3) (RC_main.cu:1853-1854, 2361-2376);
//   7. s0 <- s2 with smoothed angles wrapped into [0, 2 pi] (global.h:2169-2240).
// The director of the body-x guest is (cos g cos b, sin g cos b, -sin b)
// (device_fxn.h:34-63), so alpha carries no force.
//
// Energies of string images are evaluated one wave per image: the image is
// located in a fixed grid of ~1 A bins over the cell whose candidate lists
// (tile-relative, FP32 + FP64) are built once per framework, and the lanes
// stride over its bin's list.  One iteration is four kernels (five in a
// mixed run, with the FP64 recheck); the runtime captures them in a HIP graph
// and replays it.
#include "sv_device.h"
#include "sv_objects.h"
#include "sv_vext.h"

#include <cmath>
#include <cstring>
#include <map>
#include <vector>

namespace sv {
namespace {

constexpr double kPi = 3.141592653589793;
constexpr double kRound = 1e-15;   // RC_main.cu:1733
constexpr double kSmooth = 1e-4;   // RC_main.cu:1734
constexpr int kCheckEvery = 200;   // RC_main.cu:2361
// Bins of the string images' candidate lists: edges of about
// max(kMinBinEdge, (V / kImageBins)^(1/3)).  Smaller bins mean shorter lists
// per image (a list covers the bin's circumradius + rc + max|o|): 1 A bins
// take 9% less time per FP64 iteration than 2 A bins on the 37 x 37 x 14 A cell;
// the bin target keeps a large cell's lists to a few hundred MB.
constexpr double kMinBinEdge = 1.0;   // A
constexpr double kImageBins = 16384.0;

struct StringParams {
  int n_strings;
  int n_images;
  double move_frac;
  double move_angle;
  double conv_c;   // coordinate threshold
  double conv_o;   // orientation threshold
};


__device__ __forceinline__ double wave_sum_d(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

__device__ __forceinline__ double grid_floor(double v) { return floor(v / kRound) * kRound; }

// Candidate lists of one framework for string images: a fixed grid of bins
// over the cell (bins of about 1 A, kMinBinEdge), each with the tile-relative list of every
// image within (bin radius + rc + max|o|) of its centre, built once.
struct ImageGrid {
  const Job* job;            // coefficients, cell, list layout (segment = bin * n_types + type)
  double Hinv[9];
  int m[3];
  int pad;
  const long long* offsets;
  const float4* list32;
  const double* list64;
};

// x - floor(x): the same value as fmod(x, 1) (+1 when negative) -- both
// subtractions are exact or round alike -- at one v_floor_f64.
__device__ __forceinline__ double wrap01_s(double x) { return x - floor(x); }

// The revert decision of the mixed run (V2 > V0) is taken on FP64 energies
// whenever the mixed energies differ by no more than their error bound: FP32
// pair terms are each below the near-contact energy (tens of K), so a sum of
// ~10^3 of them is off by well under kDecideAbs; the relative part covers the
// FP64 near-contact terms' share of large energies.
constexpr double kDecideAbs = 0.05;   // K
constexpr double kDecideRel = 1e-6;

__device__ __forceinline__ bool decision_in_doubt(double v0, double v2) {
  const double tol = kDecideAbs + kDecideRel * (fabs(v0) + fabs(v2));
  return !(fabs(v2 - v0) > tol);   // nonfinite values are in doubt too
}

// Type count up to which an image sweep runs as one stream over the bin's
// type segments (more types: one sweep per type).
constexpr int kStreamTypes = 16;

// FP64 re-evaluation of the images whose decision is in doubt: waves
// [0, n_poses) take x0, waves [n_poses, 2 n_poses) take x2; results in Vd.
struct Recheck {
  const double* x2;
  const double* t2;   // trig table of x2
  const double* V0;
  const double* V2;
};

// Trig table of positions x [P][6]: (sin beta, cos beta, sin gamma, cos
// gamma) per image [P][4], written wherever the positions are (string
// creation, the step kernel for x2, the finish kernel for x0), so that the
// image kernels, whose waves are latency-bound, do not each evaluate two FP64
// sincos at their start.
__device__ __forceinline__ void image_trig(const double* __restrict__ x6, double* __restrict__ t4) {
  double sb, cb, sg, cg;
  sincos(x6[4], &sb, &cb);
  sincos(x6[5], &sg, &cg);
  t4[0] = sb;
  t4[1] = cb;
  t4[2] = sg;
  t4[3] = cg;
}

__global__ __launch_bounds__(256) void string_trig_kernel(int P, const double* __restrict__ x,
                                                          double* __restrict__ t) {
  const int p = blockIdx.x * 256 + threadIdx.x;
  if (p < P) image_trig(x + 6LL * p, t + 4LL * p);
}

// Energy (and gradient) of one pose per wave.  x = (a, b, c, alpha, beta,
// gamma) with fractional center (a, b, c).  The image is located in its bin;
// lanes stride over the bin's candidate list.  G receives dV/dx.  RECHECK:
// the FP64 re-evaluation pass described at Recheck.
template <bool MIXED, bool GRAD, bool RECHECK = false>
__global__ __launch_bounds__(256) void image_kernel(const ImageGrid* __restrict__ grids,
                                                    const int* __restrict__ plan_of, int n_images,
                                                    int n_poses, const int* __restrict__ converged,
                                                    const double* __restrict__ x_in,
                                                    const double* __restrict__ trig_in,
                                                    double* __restrict__ V, double* __restrict__ G,
                                                    Recheck rc = Recheck{}) {
  __shared__ int s_near[4][128];   // queued near-contact candidates per wave
  // streamed sweeps: per-type FP64 coefficients (c12, c6, shift, floor2) and
  // the bin's type segment bounds, per wave
  __shared__ double4 s_coef[4][kStreamTypes];
  __shared__ long long s_bound[4][kStreamTypes + 1];
  // mixed sweep: per-type FP32 constants (sigma^6 s^3, 6 c6 s^4, near-contact
  // radius^2) and the FP64 scale c6 s^3 of the type's FP32 partials
  __shared__ float4 s_cf[4][kStreamTypes];
  __shared__ double s_c6x[4][kStreamTypes];
  const int lane = threadIdx.x & 63;
  const int wv = uniform(threadIdx.x >> 6);
  const int slot = blockIdx.x * 4 + wv;   // output index
  int p = slot;                           // image index
  const double* __restrict__ x = x_in;
  const double* __restrict__ trig = trig_in;
  if (RECHECK) {
    if (slot >= 2 * n_poses) return;
    if (slot >= n_poses) {
      p = slot - n_poses;
      x = rc.x2;
      trig = rc.t2;
    }
    if (!decision_in_doubt(rc.V0[p], rc.V2[p])) return;
  } else if (p >= n_poses) {
    return;
  }
  const int s = p / n_images;
  if (converged && converged[s]) return;
  const ImageGrid& IG = grids[plan_of[s]];
  const Job& J = *IG.job;
  const double sb = ld_scalar(trig, 4LL * p), cb = ld_scalar(trig, 4LL * p + 1),
               sg = ld_scalar(trig, 4LL * p + 2), cg = ld_scalar(trig, 4LL * p + 3);
  // the image's director and center are wave-uniform: scalar registers
  const double u[3] = {uniform_d(cg * cb), uniform_d(sg * cb), uniform_d(-sb)};
  const float uf[3] = {(float)u[0], (float)u[1], (float)u[2]};
  // bin of the center and its offset from the bin centre (fractional differences)
  int idx[3];
  double df[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    double f = wrap01_s(x[6 * p + k]);
    if (!(f >= 0.0 && f <= 1.0)) f = 0.0;   // nonfinite input: stay inside the grid
    const int q = (int)(f * IG.m[k]);
    idx[k] = q < IG.m[k] - 1 ? q : IG.m[k] - 1;
    df[k] = f - (idx[k] + 0.5) / IG.m[k];
  }
  double c[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) c[k] = uniform_d(df[0] * J.H[k] + df[1] * J.H[3 + k] + df[2] * J.H[6 + k]);
  const float cf[3] = {(float)c[0], (float)c[1], (float)c[2]};
  const long long seg0 = (long long)((idx[0] * IG.m[1] + idx[1]) * IG.m[2] + idx[2]) * J.n_types;
  const double o[2] = {J.o0, J.o1};
  const double keep2 = J.keep_r * J.keep_r, rc2 = J.rc2;
  double E = 0.0, gc[3] = {0.0, 0.0, 0.0}, gu[3] = {0.0, 0.0, 0.0};
  // FP64 pair terms of candidate i (reference semantics, both sites).
  auto pair64r = [&](const double* q, double c12, double c6, double sh, double fl2) {
    double r[3];
    for (int k = 0; k < 3; ++k) r[k] = q[k] - c[k];
    const double rr = r[0] * r[0] + r[1] * r[1] + r[2] * r[2];
    if (!(rr < keep2)) return;
    for (int si = 0; si < 2; ++si) {
      const double dv[3] = {o[si] * u[0] - r[0], o[si] * u[1] - r[1], o[si] * u[2] - r[2]};
      const double d2 = dv[0] * dv[0] + dv[1] * dv[1] + dv[2] * dv[2];
      if (d2 < rc2) {
        const bool floored = d2 < fl2;
        const double inv = inv_fast(fmax(d2, fl2));
        const double i6 = inv * inv * inv;
        E += c12 * i6 * i6 - c6 * i6 + sh;
        if (GRAD && !floored) {
          const double g = 2.0 * (-6.0 * c12 * i6 * i6 + 3.0 * c6 * i6) * inv;
          for (int k = 0; k < 3; ++k) {
            gc[k] += g * dv[k];
            gu[k] += g * o[si] * dv[k];
          }
        }
      }
    }
  };
  const int n_types = J.n_types;
  const bool stream = n_types <= kStreamTypes;
  if (stream && MIXED) {
    // The mixed sweep as one stream over the bin's type segments.  A lane's
    // FP32 partials cover a run of candidates of one type (its candidates
    // come in list order, so its type only grows); they go into the FP64 sums
    // with the type's FP64 scale when the type changes and at the end, as the
    // per-type sweep flushes them.  Queued near-contact candidates find their
    // type from the segment bounds when they are evaluated.
    if (lane < n_types) {
      s_coef[wv][lane] = make_double4(J.c12[lane], J.c6[lane], J.shift[lane], J.floor2[lane]);
      const float s6x = J.s6xf[lane];
      s_cf[wv][lane] = make_float4(s6x, (float)(6.0 * J.c6x[lane] * (double)J.inv_rc2f),
                                   fmaxf(J.near2f[lane], J.lo2f[lane]), 0.0f);
      s_c6x[wv][lane] = J.c6x[lane];
    }
    if (lane <= n_types) s_bound[wv][lane] = IG.offsets[seg0 + lane];
    wave_sync();
    const long long beg = s_bound[wv][0], end = s_bound[wv][n_types];
    const float sc = J.inv_rc2f;
    const float ox[2] = {J.o0sxf, J.o1sxf}, mx[2] = {J.m2o0xf, J.m2o1xf};
    const float of0 = (float)o[0], of1 = (float)o[1];
    int nq = 0;
    int ty = 0, run_ty = 0, slots = 0;   // the type of the lane's current partials
    float eacc = 0.f, gcf[3] = {0.f, 0.f, 0.f}, guf[3] = {0.f, 0.f, 0.f};
    auto flush = [&]() {
      E += fma(s_c6x[wv][run_ty], (double)eacc, 2.0 * s_coef[wv][run_ty].z * slots);
      eacc = 0.f;
      slots = 0;
      if (GRAD)
        for (int k = 0; k < 3; ++k) {
          gc[k] += (double)gcf[k];
          gu[k] += (double)guf[k];
          gcf[k] = guf[k] = 0.f;
        }
    };
    float4 next = make_float4(0.f, 0.f, 0.f, 0.f);
    if (beg + lane < end) next = IG.list32[beg + lane];
    for (long long b0 = beg; b0 < end; b0 += 64) {
      const long long i = b0 + lane;
      const float4 e = next;
      if (i + 64 < end) next = IG.list32[i + 64];
      float rf[3] = {0.f, 0.f, 0.f}, rrf = 3.0e38f;
      if (i < end) {
        rf[0] = e.x - cf[0];
        rf[1] = e.y - cf[1];
        rf[2] = e.z - cf[2];
        rrf = rf[0] * rf[0] + rf[1] * rf[1] + rf[2] * rf[2];
        while (i >= s_bound[wv][ty + 1]) ++ty;   // i < end = s_bound[n_types]: ty < n_types
      }
      if (ty != run_ty) {
        flush();
        run_ty = ty;
      }
      const float4 k4 = s_cf[wv][ty];   // (s6x, c6f6, near2x, -)
      const bool near = rrf < k4.z;     // false past the list's end (rrf = 3e38)
      const unsigned long long mnear = __ballot(near);
      if (near) s_near[wv][nq + lanes_below(mnear)] = (int)(i - beg);
      nq += __popcll(mnear);
      ++slots;
      const float rr = near ? 3.0e38f : rrf;
      const float tt = rf[0] * uf[0] + rf[1] * uf[1] + rf[2] * uf[2];
#pragma unroll
      for (int si = 0; si < 2; ++si) {
        const float x = __builtin_amdgcn_fmed3f(fmaf(tt, mx[si], fmaf(rr, sc, ox[si])), 0.0f, 1.0f);
        const float iv = __builtin_amdgcn_rcpf(x);
        const float i3 = iv * iv * iv;
        eacc = fmaf(i3, fmaf(i3, k4.x, -1.0f), eacc);
        if (GRAD) {
          // no gradient beyond the cutoff (x clamped to 1)
          const float g = x < 1.0f ? k4.y * (i3 * iv * fmaf(-2.0f * k4.x, i3, 1.0f)) : 0.0f;
          const float of = si ? of1 : of0;
#pragma unroll
          for (int k = 0; k < 3; ++k) {
            const float dk = fmaf(of, uf[k], -rf[k]);
            gcf[k] = fmaf(g, dk, gcf[k]);
            guf[k] = fmaf(g * of, dk, guf[k]);
          }
        }
      }
      if (nq >= 64 || b0 + 64 >= end) {
        wave_sync();
        for (int q0 = 0; q0 < nq; q0 += 64)
          if (q0 + lane < nq) {
            const long long id = beg + s_near[wv][q0 + lane];
            int qt = 0;
            while (id >= s_bound[wv][qt + 1]) ++qt;
            const double4 co = s_coef[wv][qt];
            const double qv[3] = {IG.list64[3 * id], IG.list64[3 * id + 1], IG.list64[3 * id + 2]};
            pair64r(qv, co.x, co.y, co.z, co.w);
          }
        wave_sync();
        nq = 0;
      }
    }
    flush();
  }
  if (stream && !MIXED) {
    // One sweep over the bin's type segments, which are contiguous: each lane
    // finds its candidate's type from the segment bounds and takes the type's
    // coefficients from LDS, so the sweep has no break per type and the next
    // candidate is always in flight.
    if (lane < n_types)
      s_coef[wv][lane] = make_double4(J.c12[lane], J.c6[lane], J.shift[lane], J.floor2[lane]);
    if (lane <= n_types) s_bound[wv][lane] = IG.offsets[seg0 + lane];
    wave_sync();
    const long long end = s_bound[wv][n_types];
    long long i = s_bound[wv][0] + lane;
    int ty = 0;
    double q[3] = {0.0, 0.0, 0.0};
    if (i < end)
      for (int k = 0; k < 3; ++k) q[k] = IG.list64[3 * i + k];
    for (; i < end; i += 64) {
      const double cur[3] = {q[0], q[1], q[2]};
      if (i + 64 < end)
        for (int k = 0; k < 3; ++k) q[k] = IG.list64[3 * (i + 64) + k];
      while (i >= s_bound[wv][ty + 1]) ++ty;   // i < end = s_bound[n_types]: ty < n_types
      const double4 co = s_coef[wv][ty];
      pair64r(cur, co.x, co.y, co.z, co.w);
    }
  }
  for (int t = 0; t < n_types && !stream; ++t) {
    const double c12 = J.c12[t], c6 = J.c6[t], sh = J.shift[t], fl2 = J.floor2[t];
    const float near2 = J.near2f[t];
    const long long beg = IG.offsets[seg0 + t], end = IG.offsets[seg0 + t + 1];
    if (!MIXED) {
      // the next candidate's coordinates are loaded while this one is evaluated
      long long i = beg + lane;
      double q[3] = {0.0, 0.0, 0.0};
      if (i < end)
        for (int k = 0; k < 3; ++k) q[k] = IG.list64[3 * i + k];
      for (; i < end; i += 64) {
        const double cur[3] = {q[0], q[1], q[2]};
        if (i + 64 < end)
          for (int k = 0; k < 3; ++k) q[k] = IG.list64[3 * (i + 64) + k];
        pair64r(cur, c12, c6, sh, fl2);
      }
      continue;
    }
    // Mixed: FP32 terms in the main sweep; near-contact candidates are queued
    // in LDS and evaluated in FP64 one per lane, so the sweep never branches
    // into the FP64 path.  The FP32 terms work in units of 1/s (s = FP32
    // 1/rc^2): x = s d^2 is clamped to [0, 1] by the fma clamp bit, so every
    // candidate slot of the sweep -- beyond the cutoff, past the list's end or
    // near-contact (forced) -- yields either its term or the clamped one,
    // which the +shift added per slot cancels (~3e-8 K per clamped pair from
    // the rounding of s).
    int nq = 0;
    // Lane partials of this type in FP32: e in units of c6 s^3, gradient in
    // K/A; flushed into the FP64 sums once per type.
    float eacc = 0.f, gcf[3] = {0.f, 0.f, 0.f}, guf[3] = {0.f, 0.f, 0.f};
    int slots = 0;
    const float sc = J.inv_rc2f, s6x = J.s6xf[t], m2s6x = -2.0f * s6x;
    const float ox[2] = {J.o0sxf, J.o1sxf}, mx[2] = {J.m2o0xf, J.m2o1xf};
    // 2 dphi/d(d^2) = 6 c6 s^4 i^4 (1 - 2 sigma^6 s^3 i^3), i = 1/x
    const float c6f6 = (float)(6.0 * J.c6x[t] * (double)sc);
    // candidates within the floor's reach are near-contact (FP64) too
    const float near2x = fmaxf(near2, J.lo2f[t]);
    const float of0 = (float)o[0], of1 = (float)o[1];
    // the next chunk's entry is loaded while this chunk is evaluated
    float4 next = make_float4(0.f, 0.f, 0.f, 0.f);
    if (beg + lane < end) next = IG.list32[beg + lane];
    for (long long b0 = beg; b0 < end; b0 += 64) {
      const long long i = b0 + lane;
      const float4 e = next;
      if (i + 64 < end) next = IG.list32[i + 64];
      float rf[3] = {0.f, 0.f, 0.f}, rrf = 3.0e38f;
      if (i < end) {
        rf[0] = e.x - cf[0];
        rf[1] = e.y - cf[1];
        rf[2] = e.z - cf[2];
        rrf = rf[0] * rf[0] + rf[1] * rf[1] + rf[2] * rf[2];
      }
      const bool near = rrf < near2x;   // false past the list's end (rrf = 3e38)
      const unsigned long long mnear = __ballot(near);
      if (near) s_near[wv][nq + lanes_below(mnear)] = (int)(i - beg);
      nq += __popcll(mnear);
      ++slots;
      const float rr = near ? 3.0e38f : rrf;
      const float tt = rf[0] * uf[0] + rf[1] * uf[1] + rf[2] * uf[2];
#pragma unroll
      for (int si = 0; si < 2; ++si) {
        const float x = __builtin_amdgcn_fmed3f(fmaf(tt, mx[si], fmaf(rr, sc, ox[si])), 0.0f, 1.0f);
        const float iv = __builtin_amdgcn_rcpf(x);
        const float i3 = iv * iv * iv;
        eacc = fmaf(i3, fmaf(i3, s6x, -1.0f), eacc);
        if (GRAD) {
          // no gradient beyond the cutoff (x clamped to 1)
          const float g = x < 1.0f ? c6f6 * (i3 * iv * fmaf(m2s6x, i3, 1.0f)) : 0.0f;
          const float of = si ? of1 : of0;
#pragma unroll
          for (int k = 0; k < 3; ++k) {
            const float dk = fmaf(of, uf[k], -rf[k]);
            gcf[k] = fmaf(g, dk, gcf[k]);
            guf[k] = fmaf(g * of, dk, guf[k]);
          }
        }
      }
      if (nq >= 64 || b0 + 64 >= end) {
        wave_sync();
        for (int q0 = 0; q0 < nq; q0 += 64)
          if (q0 + lane < nq) {
            const long long id = beg + s_near[wv][q0 + lane];
            const double qv[3] = {IG.list64[3 * id], IG.list64[3 * id + 1], IG.list64[3 * id + 2]};
            pair64r(qv, c12, c6, sh, fl2);
          }
        wave_sync();
        nq = 0;
      }
    }
    E += fma(J.c6x[t], (double)eacc, 2.0 * sh * slots);
    if (GRAD)
      for (int k = 0; k < 3; ++k) {
        gc[k] += (double)gcf[k];
        gu[k] += (double)guf[k];
      }
  }
  E = wave_sum_d(E);
  if (GRAD)
    for (int k = 0; k < 3; ++k) {
      gc[k] = wave_sum_d(gc[k]);
      gu[k] = wave_sum_d(gu[k]);
    }
  if (lane == 0) {
    V[slot] = E;
    if (GRAD) {
      // c = f H: dV/df_k = sum_i H[k][i] dV/dc_i
      for (int k = 0; k < 3; ++k)
        G[6 * p + k] = J.H[3 * k] * gc[0] + J.H[3 * k + 1] * gc[1] + J.H[3 * k + 2] * gc[2];
      G[6 * p + 3] = 0.0;
      G[6 * p + 4] = gu[0] * (-cg * sb) + gu[1] * (-sg * sb) + gu[2] * (-cb);
      G[6 * p + 5] = gu[0] * (-sg * cb) + gu[1] * (cg * cb);
    }
  }
}

// Steps 1-4: forces -> s1 -> endpoint fix -> reparametrized s2.
__global__ __launch_bounds__(256) void string_step_kernel(const ImageGrid* __restrict__ grids,
                                                          const int* __restrict__ plan_of,
                                                          const int* __restrict__ direction,
                                                          StringParams sp,
                                                          const int* __restrict__ converged,
                                                          const double* __restrict__ x0,
                                                          const double* __restrict__ G,
                                                          double* __restrict__ x2,
                                                          double* __restrict__ t2) {
  extern __shared__ double sh[];
  const int s = blockIdx.x;
  if (converged[s]) return;
  const int N = sp.n_images;
  double* s1 = sh;            // [N][6]
  double* cum = sh + 6 * N;   // [N] cumulative arc lengths, then their fractions of the total
  const Job& J = *grids[plan_of[s]].job;
  const int base = s * N;
  for (int i = threadIdx.x; i < N; i += blockDim.x) {
    const int p = base + i;
    double g[6];
    for (int k = 0; k < 6; ++k) g[k] = grid_floor((k < 3 ? sp.move_frac : sp.move_angle) * G[6 * p + k]);
    const double nt = grid_floor(sqrt(g[0] * g[0] + g[1] * g[1] + g[2] * g[2]));
    const double nr = grid_floor(sqrt(g[3] * g[3] + g[4] * g[4] + g[5] * g[5]));
    for (int k = 0; k < 6; ++k) {
      const double mv = k < 3 ? sp.move_frac : sp.move_angle;
      const double n = k < 3 ? nt : nr;
      const double step = n < mv ? g[k] : grid_floor(g[k] / n);
      s1[6 * i + k] = x0[6 * p + k] - mv * step;
    }
  }
  __syncthreads();
  const int d = direction[s] - 1;
  if (threadIdx.x == 0) {
    s1[d] = x0[6 * base + d];
    s1[6 * (N - 1) + d] = x0[6 * (base + N - 1) + d];
    for (int k = 0; k < 6; ++k)
      if (k != d) s1[6 * (N - 1) + k] = s1[k];
  }
  __syncthreads();
  // Interior images outside [0, 1] along the hop direction are rare; the
  // in-order replacement runs only when some image needs it.
  int outside = 0;
  for (int i = 1 + threadIdx.x; i < N - 1; i += blockDim.x)
    outside |= (s1[6 * i + d] < 0.0) | (s1[6 * i + d] > 1.0);
  if (__syncthreads_or(outside) && threadIdx.x == 0) {
    for (int i = 1; i < N - 1; ++i) {
      int from = -1;
      if (s1[6 * i + d] < 0.0) {
        for (int ii = i + 1; ii < N; ++ii)
          if (s1[6 * ii + d] > 0.0) { from = ii; break; }
      } else if (s1[6 * i + d] > 1.0) {
        for (int ii = i - 1; ii >= 0; --ii)
          if (s1[6 * ii + d] < 1.0) { from = ii; break; }
      }
      if (from >= 0)
        for (int k = 0; k < 6; ++k) s1[6 * i + k] = s1[6 * from + k];
    }
  }
  __syncthreads();
  // Cartesian segment lengths of (a, b, c) in parallel, then the cumulative
  // sum in image order.
  for (int i = 1 + threadIdx.x; i < N; i += blockDim.x) {
    double l2 = 0.0;
    for (int k = 0; k < 3; ++k) {
      const double prev = s1[6 * (i - 1)] * J.H[k] + s1[6 * (i - 1) + 1] * J.H[3 + k] +
                          s1[6 * (i - 1) + 2] * J.H[6 + k];
      const double cur = s1[6 * i] * J.H[k] + s1[6 * i + 1] * J.H[3 + k] + s1[6 * i + 2] * J.H[6 + k];
      l2 += (cur - prev) * (cur - prev);
    }
    cum[i] = sqrt(l2);
  }
  __syncthreads();
  // Inclusive prefix sum of the segment lengths by wave 0, 64 at a time
  // (a serial loop over the images was a third of this kernel's time).
  if (threadIdx.x < 64) {
    const int lane = threadIdx.x;
    double carry = 0.0;
    for (int c0 = 0; c0 < N; c0 += 64) {
      const int i = c0 + lane;
      double v = (i >= 1 && i < N) ? cum[i] : 0.0;
#pragma unroll
      for (int o = 1; o < 64; o <<= 1) {
        const double y = __shfl_up(v, o, 64);
        if (lane >= o) v += y;
      }
      v += carry;
      if (i < N) cum[i] = v;
      carry = __shfl(v, 63, 64);
    }
  }
  __syncthreads();
  const double total = cum[N - 1];
  __syncthreads();   // every thread has the total before cum turns into fractions
  double* frac = cum;
  for (int i = threadIdx.x; i < N; i += blockDim.x) frac[i] = total > 0.0 ? cum[i] / total : 0.0;
  __syncthreads();
  for (int i = threadIdx.x; i < N; i += blockDim.x) {
    const int p = base + i;
    const double target = (double)i / (double)(N - 1);
    int ii = 0;
    if (total > 0.0) {
      int lo = 0, hi = N - 2;   // last segment index whose start fraction <= target
      while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (frac[mid] <= target) lo = mid; else hi = mid - 1;
      }
      ii = lo;
    }
    const double f0 = frac[ii];
    const double f1 = frac[ii + 1];
    for (int k = 0; k < 3; ++k) {
      const double a0 = s1[6 * ii + k], a1 = s1[6 * (ii + 1) + k];
      x2[6 * p + k] = (f1 > f0) ? a0 + (target - f0) * ((a1 - a0) / (f1 - f0)) : a0;
    }
    for (int k = 3; k < 6; ++k) x2[6 * p + k] = s1[6 * i + k];
    image_trig(s1 + 6 * i, t2 + 4LL * p);
  }
}

// One periodic-aware smoothing of angle k of image i (global.h:1905-2085).
__device__ double smooth_one(const double* ang, int i, int N, int k) {
  const double c = ang[3 * i + k];
  if (i == 0 || i == N - 1) return c;
  const double l = ang[3 * (i - 1) + k], r = ang[3 * (i + 1) + k];
  const double dl = l - c, dr = r - c;
  double wraps;
  if (dl > kPi) wraps = dr > kPi ? -2.0 : -1.0;
  else if (dl < -kPi) wraps = dr < -kPi ? 2.0 : 0.0;
  else wraps = dr > kPi ? -1.0 : (dr < -kPi ? 1.0 : 0.0);
  return (1.0 - kSmooth) * c + 0.5 * kSmooth * (l + r + wraps * 2.0 * kPi);
}

// Steps 5-7: angle revert, smoothing, convergence test, commit.  Vd (mixed
// runs): FP64 energies at x0 [P] and x2 [P] of the images whose decision
// was in doubt; null in FP64 runs.
__global__ __launch_bounds__(256) void string_finish_kernel(StringParams sp,
                                                            int* __restrict__ converged,
                                                            int* __restrict__ iterations,
                                                            double* __restrict__ x0,
                                                            double* __restrict__ t0,
                                                            const double* __restrict__ x2,
                                                            const double* __restrict__ V0,
                                                            const double* __restrict__ V2,
                                                            const double* __restrict__ Vd) {
  extern __shared__ double sh[];
  const int s = blockIdx.x;
  if (converged[s]) return;
  const int N = sp.n_images;
  const int P = sp.n_strings * N;
  double* ang = sh;                 // [N][3] checked s2 angles
  double* red = sh + 3 * N;         // [2 * waves]
  const int base = s * N;
  double dc = 0.0, dor = 0.0;
  for (int i = threadIdx.x; i < N; i += blockDim.x) {
    const int p = base + i;
    const bool revert = Vd && decision_in_doubt(V0[p], V2[p]) ? Vd[P + p] > Vd[p] : V2[p] > V0[p];
    double sq = 0.0, so = 0.0;
    for (int k = 0; k < 3; ++k) {
      const double a = revert ? x0[6 * p + 3 + k] : x2[6 * p + 3 + k];
      ang[3 * i + k] = a;
      const double e = x0[6 * p + k] - x2[6 * p + k];
      const double q = x0[6 * p + 3 + k] - a;
      sq += e * e;
      so += q * q;
    }
    dc += sqrt(sq);
    dor += sqrt(so);
  }
  // wave sums, then the waves' sums in wave order
  dc = wave_sum_d(dc);
  dor = wave_sum_d(dor);
  const int wv = threadIdx.x >> 6, n_waves = blockDim.x >> 6;
  if ((threadIdx.x & 63) == 0) {
    red[wv] = dc;
    red[n_waves + wv] = dor;
  }
  __syncthreads();
  double tot_c = 0.0, tot_o = 0.0;
  for (int w = 0; w < n_waves; ++w) {
    tot_c += red[w];
    tot_o += red[n_waves + w];
  }
  const int it = iterations[s];
  const bool check = it % kCheckEvery == 0 && it >= kCheckEvery;
  const bool done = check && isfinite(tot_c) && isfinite(tot_o) && tot_c < sp.conv_c &&
                    tot_o < sp.conv_o;
  __syncthreads();
  if (done) {
    if (threadIdx.x == 0) converged[s] = 1;
    return;
  }
  for (int i = threadIdx.x; i < N; i += blockDim.x) {
    const int p = base + i;
    for (int k = 0; k < 3; ++k) x0[6 * p + k] = x2[6 * p + k];
    for (int k = 0; k < 3; ++k) {
      double a = smooth_one(ang, i, N, k);
      if (a > 2.0 * kPi) a -= 2.0 * kPi;
      else if (a < 0.0) a += 2.0 * kPi;
      x0[6 * p + 3 + k] = a;
    }
    image_trig(x0 + 6LL * p, t0 + 4LL * p);
  }
  if (threadIdx.x == 0) iterations[s] = it + 1;
}

// x: positions [P][6]; t: their trig table [P][4] (image_trig).
hipError_t launch_image(int precision, bool grad, const ImageGrid* plans, const int* plan_of, int N,
                        int P, const int* converged, const double* x, const double* t, double* V,
                        double* G, hipStream_t st) {
  const dim3 g((P + 3) / 4), b(256);
#define SV_IMAGE(M, GR)                                                                        \
  hipLaunchKernelGGL((image_kernel<M, GR>), g, b, 0, st, plans, plan_of, N, P, converged, x, t, \
                     V, G, Recheck{})
  if (precision == PREC_MIXED) {
    if (grad) SV_IMAGE(true, true);
    else SV_IMAGE(true, false);
  } else {
    if (grad) SV_IMAGE(false, true);
    else SV_IMAGE(false, false);
  }
#undef SV_IMAGE
  return hipGetLastError();
}

// FP64 energies of the images of a mixed run whose revert decision is in doubt.
hipError_t launch_recheck(const ImageGrid* plans, const int* plan_of, int N, int P,
                          const int* converged, const double* x0, const double* t0,
                          const double* x2, const double* t2, const double* V0, const double* V2,
                          double* Vd, hipStream_t st) {
  const dim3 g((2 * P + 3) / 4), b(256);
  hipLaunchKernelGGL((image_kernel<false, false, true>), g, b, 0, st, plans, plan_of, N, P,
                     converged, x0, t0, Vd, nullptr, Recheck{x2, t2, V0, V2});
  return hipGetLastError();
}

hipError_t launch_trig(int P, const double* x, double* t, hipStream_t st) {
  hipLaunchKernelGGL(string_trig_kernel, dim3((P + 255) / 256), dim3(256), 0, st, P, x, t);
  return hipGetLastError();
}

}  // namespace

struct StringSys {
  int device = 0;
  hipStream_t stream = nullptr;
  int S = 0, N = 0, P = 0, precision = 0;
  StringParams sp{};
  std::vector<PoseGrid*> grids;      // candidate lists per distinct framework
  DevBuf<ImageGrid> d_grids;
  DevBuf<int> d_plan_of, d_dir, d_conv, d_iter;
  DevBuf<double> d_x0, d_x2, d_G, d_V0, d_V2, d_Vd;   // d_Vd: mixed runs' FP64 rechecks [2P]
  DevBuf<double> d_t0, d_t2;                          // trig tables of x0 and x2 [P][4]
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  ~StringSys() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    if (stream) (void)hipStreamDestroy(stream);
    for (PoseGrid* g : grids) delete g;
  }
};

namespace {

#define SV_HIP_S(call)                                           \
  do {                                                           \
    hipError_t sv_e_ = (call);                                   \
    if (sv_e_ != hipSuccess) return hip_fail(sv_e_, #call);      \
  } while (0)

int iteration(StringSys* Y) {
  const hipStream_t st = Y->stream;
  const size_t step_shm = sizeof(double) * 7 * Y->N;
  const size_t fin_shm = sizeof(double) * (3 * Y->N + 2 * 256);
  SV_HIP_S(launch_image(Y->precision, true, Y->d_grids.p, Y->d_plan_of.p, Y->N, Y->P, Y->d_conv.p,
                        Y->d_x0.p, Y->d_t0.p, Y->d_V0.p, Y->d_G.p, st));
  hipLaunchKernelGGL(string_step_kernel, dim3(Y->S), dim3(256), step_shm, st, Y->d_grids.p,
                     Y->d_plan_of.p, Y->d_dir.p, Y->sp, Y->d_conv.p, Y->d_x0.p, Y->d_G.p,
                     Y->d_x2.p, Y->d_t2.p);
  SV_HIP_S(hipGetLastError());
  SV_HIP_S(launch_image(Y->precision, false, Y->d_grids.p, Y->d_plan_of.p, Y->N, Y->P, Y->d_conv.p,
                        Y->d_x2.p, Y->d_t2.p, Y->d_V2.p, nullptr, st));
  const bool mixed = Y->precision == PREC_MIXED;
  if (mixed)
    SV_HIP_S(launch_recheck(Y->d_grids.p, Y->d_plan_of.p, Y->N, Y->P, Y->d_conv.p, Y->d_x0.p,
                            Y->d_t0.p, Y->d_x2.p, Y->d_t2.p, Y->d_V0.p, Y->d_V2.p, Y->d_Vd.p, st));
  hipLaunchKernelGGL(string_finish_kernel, dim3(Y->S), dim3(256), fin_shm, st, Y->sp, Y->d_conv.p,
                     Y->d_iter.p, Y->d_x0.p, Y->d_t0.p, Y->d_x2.p, Y->d_V0.p, Y->d_V2.p,
                     mixed ? Y->d_Vd.p : nullptr);
  SV_HIP_S(hipGetLastError());
  return 0;
}

}  // namespace
}  // namespace sv

using namespace sv;

extern "C" {

int sv_strings_create(int n_strings, void* const* plans, int n_images, const double* init,
                      const int* direction, double move_frac, double move_angle_rad,
                      int precision, void** strings) {
  if (!plans || !init || !direction || !strings) return fail(1, "null argument");
  if (n_strings < 1 || n_images < 3 || n_images > 1024) return fail(1, "1+ strings of 3..1024 images required");
  if (precision != PREC_FP64 && precision != PREC_MIXED) return fail(1, "precision must be 0 or 1");
  if (!(move_frac > 0.0 && move_angle_rad > 0.0 && std::isfinite(move_frac) && std::isfinite(move_angle_rad)))
    return fail(1, "positive finite moves required");
  for (int s = 0; s < n_strings; ++s)
    if (direction[s] < 1 || direction[s] > 3 || !plans[s]) return fail(1, "direction must be 1, 2 or 3 and plans non-null");
  for (long long i = 0; i < 6LL * n_strings * n_images; ++i)
    if (!std::isfinite(init[i])) return fail(1, "nonfinite initial string");
  auto Y = new StringSys();
  Y->S = n_strings;
  Y->N = n_images;
  Y->P = n_strings * n_images;
  Y->precision = precision;
  Y->sp.n_strings = n_strings;
  Y->sp.n_images = n_images;
  Y->sp.move_frac = move_frac;
  Y->sp.move_angle = move_angle_rad;
  Y->sp.conv_c = 0.30 * move_frac * n_images * std::sqrt(3.0);
  Y->sp.conv_o = 0.30 * move_angle_rad * n_images * std::sqrt(3.0);
  std::map<const void*, int> index;
  std::vector<int> plan_of(n_strings);
  Y->device = static_cast<Plan*>(plans[0])->device;
  for (int s = 0; s < n_strings; ++s) {
    auto it = index.find(plans[s]);
    if (it == index.end()) {
      Plan* P = static_cast<Plan*>(plans[s]);
      if (P->device != Y->device) {
        delete Y;
        return fail(1, "all plans must be on one device");
      }
      // bins of about `edge` along every lattice direction
      const double edge = std::max(kMinBinEdge, std::cbrt(P->volume / kImageBins));
      int m[3];
      for (int k = 0; k < 3; ++k)
        m[k] = (int)std::min(256.0, std::max(1.0, std::floor(1.0 / (P->hn[k] * edge) + 0.5)));
      PoseGrid* g = nullptr;
      const int rc = pose_grid_build(P, m, false, &g);
      if (rc) {
        delete Y;
        return rc;
      }
      it = index.emplace(plans[s], (int)Y->grids.size()).first;
      Y->grids.push_back(g);
    }
    plan_of[s] = it->second;
  }
  std::vector<int> zeros(n_strings, 0);
  hipError_t e;
  std::vector<ImageGrid> ig(Y->grids.size());
  for (size_t i = 0; i < Y->grids.size(); ++i) {
    const PoseGrid* g = Y->grids[i];
    ig[i] = ImageGrid{};
    ig[i].job = g->B->d_jobs.p;
    std::memcpy(ig[i].Hinv, g->Hinv.m, sizeof(ig[i].Hinv));
    for (int k = 0; k < 3; ++k) ig[i].m[k] = g->m[k];
    ig[i].offsets = g->B->d_offsets.p;
    ig[i].list32 = g->B->d_list32.p;
    ig[i].list64 = g->B->d_list64.p;
  }
  if ((e = hipSetDevice(Y->device)) != hipSuccess || (e = hipStreamCreate(&Y->stream)) != hipSuccess ||
      (e = Y->d_grids.upload(ig.data(), ig.size(), Y->stream)) != hipSuccess ||
      (e = Y->d_plan_of.upload(plan_of.data(), n_strings, Y->stream)) != hipSuccess ||
      (e = Y->d_dir.upload(direction, n_strings, Y->stream)) != hipSuccess ||
      (e = Y->d_conv.upload(zeros.data(), n_strings, Y->stream)) != hipSuccess ||
      (e = Y->d_iter.upload(zeros.data(), n_strings, Y->stream)) != hipSuccess ||
      (e = Y->d_x0.upload(init, 6 * (size_t)Y->P, Y->stream)) != hipSuccess ||
      (e = Y->d_t0.reserve(4 * (size_t)Y->P)) != hipSuccess ||
      (e = Y->d_t2.reserve(4 * (size_t)Y->P)) != hipSuccess ||
      (e = launch_trig(Y->P, Y->d_x0.p, Y->d_t0.p, Y->stream)) != hipSuccess ||
      (e = Y->d_x2.reserve(6 * (size_t)Y->P)) != hipSuccess ||
      (e = Y->d_G.reserve(6 * (size_t)Y->P)) != hipSuccess ||
      (e = Y->d_V0.reserve(Y->P)) != hipSuccess || (e = Y->d_V2.reserve(Y->P)) != hipSuccess ||
      (precision == PREC_MIXED && (e = Y->d_Vd.reserve(2 * (size_t)Y->P)) != hipSuccess) ||
      (e = hipStreamSynchronize(Y->stream)) != hipSuccess) {
    delete Y;
    return hip_fail(e, "string setup");
  }
  *strings = Y;
  return 0;
}

int sv_strings_run(void* strings, int iterations, int use_graph) {
  if (!strings || iterations < 0) return fail(1, "invalid argument");
  StringSys* Y = static_cast<StringSys*>(strings);
  SV_HIP_S(hipSetDevice(Y->device));
  if (use_graph && !Y->exec) {
    SV_HIP_S(hipStreamBeginCapture(Y->stream, hipStreamCaptureModeThreadLocal));
    const int rc = iteration(Y);
    hipGraph_t g = nullptr;
    const hipError_t end = hipStreamEndCapture(Y->stream, &g);
    if (rc) {
      if (g) (void)hipGraphDestroy(g);
      return rc;
    }
    SV_HIP_S(end);
    Y->graph = g;
    SV_HIP_S(hipGraphInstantiate(&Y->exec, Y->graph, nullptr, nullptr, 0));
  }
  for (int i = 0; i < iterations; ++i) {
    if (use_graph) {
      SV_HIP_S(hipGraphLaunch(Y->exec, Y->stream));
    } else {
      const int rc = iteration(Y);
      if (rc) return rc;
    }
  }
  SV_HIP_S(hipStreamSynchronize(Y->stream));
  return 0;
}

int sv_strings_state(void* strings, double* path, double* energy, int* iterations, int* converged) {
  if (!strings) return fail(1, "null strings");
  StringSys* Y = static_cast<StringSys*>(strings);
  SV_HIP_S(hipSetDevice(Y->device));
  if (energy)
    SV_HIP_S(launch_image(Y->precision, false, Y->d_grids.p, Y->d_plan_of.p, Y->N, Y->P, nullptr,
                          Y->d_x0.p, Y->d_t0.p, Y->d_V0.p, nullptr, Y->stream));
  if (path) SV_HIP_S(hipMemcpyAsync(path, Y->d_x0.p, sizeof(double) * 6 * Y->P, hipMemcpyDeviceToHost, Y->stream));
  if (energy) SV_HIP_S(hipMemcpyAsync(energy, Y->d_V0.p, sizeof(double) * Y->P, hipMemcpyDeviceToHost, Y->stream));
  if (iterations) SV_HIP_S(hipMemcpyAsync(iterations, Y->d_iter.p, sizeof(int) * Y->S, hipMemcpyDeviceToHost, Y->stream));
  if (converged) SV_HIP_S(hipMemcpyAsync(converged, Y->d_conv.p, sizeof(int) * Y->S, hipMemcpyDeviceToHost, Y->stream));
  SV_HIP_S(hipStreamSynchronize(Y->stream));
  if (energy)
    for (int p = 0; p < Y->P; ++p)
      if (!std::isfinite(energy[p])) return fail(4, "nonfinite string energy");
  return 0;
}

int sv_strings_destroy(void* strings) {
  delete static_cast<StringSys*>(strings);
  return 0;
}

}  // extern "C"
