// Host runtime of libsv_vext.so: FP64 geometry setup, pose tiling, device
// tables, launch orchestration and the C ABI declared in include/sv_vext.h.
#include "sv_objects.h"
#include "sv_vext.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <utility>
#include <vector>

namespace sv {

hipError_t launch_zero(int* p, int n, hipStream_t s);
hipError_t launch_zero_u64(unsigned long long* p, int n, hipStream_t s);

thread_local std::string t_error;

int fail(int code, const std::string& msg) {
  t_error = msg;
  return code;
}

int hip_fail(hipError_t e, const char* what) {
  return fail(e == hipErrorOutOfMemory ? 6 : 5, std::string(what) + ": " + hipGetErrorString(e));
}

namespace {

#define SV_HIP(call)                                   \
  do {                                                 \
    hipError_t sv_e_ = (call);                         \
    if (sv_e_ != hipSuccess) return hip_fail(sv_e_, #call); \
  } while (0)

// numpy's np.mod(x, 1.0) for finite x (result in [0, 1], 1.0 only for tiny
// negative x), used by the reference to wrap fractional coordinates.
double wrap01(double x) {
  double r = std::fmod(x, 1.0);
  if (r != 0.0 && r < 0.0) r += 1.0;
  return r;
}

// u / |u| with |u| = sqrt((x^2 + y^2) + z^2), each operation rounded on its
// own: the normalization numpy applies to the directors of a pose set
// (axes / np.linalg.norm(axes, axis=1, keepdims=True), reference
// native.py:34).  False for a nonfinite or zero director.
bool unit_director(const double* u, double* out) {
#pragma clang fp contract(off)
  const double nn = (u[0] * u[0] + u[1] * u[1]) + u[2] * u[2];
  if (!std::isfinite(nn) || nn <= 0.0) return false;
  const double norm = std::sqrt(nn);
  for (int k = 0; k < 3; ++k) out[k] = u[k] / norm;
  return true;
}

bool invert3(const double* H, double* inv, double* det_out) {
  const double a = H[0], b = H[1], c = H[2], d = H[3], e = H[4], f = H[5], g = H[6], h = H[7],
               i = H[8];
  const double A = e * i - f * h, B = -(d * i - f * g), C = d * h - e * g;
  const double det = a * A + b * B + c * C;
  *det_out = det;
  if (!std::isfinite(det) || det == 0.0) return false;
  inv[0] = A / det;
  inv[3] = B / det;
  inv[6] = C / det;
  inv[1] = -(b * i - c * h) / det;
  inv[4] = (a * i - c * g) / det;
  inv[7] = -(a * h - b * g) / det;
  inv[2] = (b * f - c * e) / det;
  inv[5] = -(a * f - c * d) / det;
  inv[8] = (a * e - b * d) / det;
  return true;
}

}  // namespace

void plan_fill_job(const Plan& P, Job& J) {
  std::memcpy(J.H, P.H, sizeof(J.H));
  std::memcpy(J.hn, P.hn, sizeof(J.hn));
  J.frac = P.d_frac.p;
  J.type_start = P.d_type_start.p;
  J.n_atoms = P.n_atoms;
  J.n_types = P.n_types;
  J.o0 = P.o0;
  J.o1 = P.o1;
  J.rc2 = P.cutoff * P.cutoff;
  J.keep_r = P.keep_r;
  J.rc2f = (float)J.rc2;
  J.keep2f = (float)(P.keep_r * P.keep_r);
  J.m2o0f = (float)(-2.0 * P.o0);
  J.m2o1f = (float)(-2.0 * P.o1);
  J.o0sqf = (float)(P.o0 * P.o0);
  J.o1sqf = (float)(P.o1 * P.o1);
  const double in_r = P.cutoff - std::max(std::fabs(P.o0), std::fabs(P.o1)) - 1e-3;
  J.in2f = in_r > 0.0 ? (float)(in_r * in_r) : 0.0f;
  J.sym = (P.o1 == -P.o0);
  J.near_inside = P.near_inside ? 1 : 0;
  J.c12 = P.d_c12.p;
  J.c6 = P.d_c6.p;
  J.shift = P.d_shift.p;
  J.floor2 = P.d_floor2.p;
  J.sigma = P.d_sigma.p;
  J.eps = P.d_eps.p;
  J.s6f = P.d_s6f.p;
  J.floor2f = P.d_floor2f.p;
  J.near2f = P.d_near2f.p;
  J.lo2f = P.d_lo2f.p;
  J.c6s = P.d_c6s.p;
  J.s6sf = P.d_s6sf.p;
  J.ftype = P.d_ftype.p;
  J.c6x = P.d_c6x.p;
  J.s6xf = P.d_s6xf.p;
  J.o0sxf = (float)(P.o0 * P.o0 * (double)(float)(1.0 / J.rc2));
  J.o1sxf = (float)(P.o1 * P.o1 * (double)(float)(1.0 / J.rc2));
  J.m2o0xf = (float)(-2.0 * P.o0 * (double)(float)(1.0 / J.rc2));
  J.m2o1xf = (float)(-2.0 * P.o1 * (double)(float)(1.0 / J.rc2));
  J.inv_rc2f = (float)(1.0 / J.rc2);
  J.m2o0sf = (float)(-2.0 * P.o0 / J.rc2);
  J.m2o1sf = (float)(-2.0 * P.o1 / J.rc2);
  J.nsite2f = P.d_nsite2f.p;
  J.n_cls = 1;
  J.list_scale = 1.0;
}

namespace {

// Tile counts per lattice direction: tiles of roughly 64 centers, about as
// thick along every direction.  A regular grid of `grid` points per axis gets
// a divisor of `grid` so that tiles hold whole blocks of grid points.
void choose_tiles(const Plan& P, long long n_centers, int grid, int tile_centers, int m[3]) {
  const double rho = (double)n_centers / P.volume;
  const double edge = std::cbrt((double)tile_centers / std::max(rho, 1e-30));
  for (int k = 0; k < 3; ++k) {
    const double thick = 1.0 / P.hn[k];
    double want = thick / edge;
    if (want < 1.0) want = 1.0;
    if (grid > 0) {
      int best = 1;
      double best_err = 1e300;
      for (int d = 1; d <= grid; ++d) {
        if (grid % d) continue;
        const double err = std::fabs(std::log((double)d / want));
        if (err < best_err) best_err = err, best = d;
      }
      m[k] = best;
    } else {
      m[k] = (int)std::min(1024.0, std::max(1.0, std::floor(want + 0.5)));
    }
  }
}

// Bins of a site grid for about n_sites uniformly spread sites: the counts per
// lattice direction that minimize the expected candidate work per site,
//   sum_n Poisson(n; lambda) sum_{chunks j} (s_k r_b + rc)^3 / lambda,
// with lambda = n_sites / bins, k = ceil(n / 64) chunks per bin, r_b the bin's
// circumradius and s_k the mean bounding radius of a chunk of k in units of
// r_b (0.82, 0.72, 0.72, 0.62 for k = 1, 2, 3, >= 4: sampled over rhombic and
// skew cells when a bin was cut into k equal octant-ordered chunks).  With
// the site kernel's Morton runs and box culling, a CPU emulation of 65,536
// random poses in the 37 x 37 x 14 A cell still ranks the model's pick
// (11 x 12 x 4) first among nine nearby bin counts (1.70 lane slots per
// in-cutoff pair; 1.72-1.84 for the others).  Searched about the counts that
// give bins as thick as they are wide.
void choose_site_bins(const Plan& P, long long n_sites, int m_out[3]) {
  const double rho = (double)std::max(1LL, n_sites) / P.volume;
  const double rc = P.cutoff;
  int guess[3];
  const double edge = std::cbrt(96.0 / rho);
  for (int k = 0; k < 3; ++k) guess[k] = std::max(1, (int)std::lround(1.0 / P.hn[k] / edge));
  double best = INFINITY;
  for (int k = 0; k < 3; ++k) m_out[k] = guess[k];
  auto range = [](int g, int* lo, int* hi) {
    *lo = std::max(1, (int)std::floor(g * 0.55));
    *hi = std::min(1024, std::max(*lo, (int)std::ceil(g * 1.6)));
  };
  int lo[3], hi[3];
  for (int k = 0; k < 3; ++k) range(guess[k], &lo[k], &hi[k]);
  const double shrink[5] = {1.0, 0.82, 0.72, 0.72, 0.62};
  for (int a = lo[0]; a <= hi[0]; ++a)
    for (int b = lo[1]; b <= hi[1]; ++b)
      for (int c = lo[2]; c <= hi[2]; ++c) {
        const double bins = (double)a * b * c;
        if (bins > (1 << 18)) continue;
        const double lambda = (double)n_sites / bins;
        double rb = 0.0;
        for (int sb = -1; sb <= 1; sb += 2)
          for (int sc = -1; sc <= 1; sc += 2) {
            double r2 = 0.0;
            for (int k = 0; k < 3; ++k) {
              const double x = P.H[k] / a + sb * P.H[3 + k] / b + sc * P.H[6 + k] / c;
              r2 += x * x;
            }
            rb = std::max(rb, 0.5 * std::sqrt(r2));
          }
        // expected work over the Poisson count of one bin
        const int n_max = (int)(lambda + 8.0 * std::sqrt(lambda) + 64.0);
        double work = 0.0, logp = -lambda;   // log P(0)
        for (int n = 0; n <= n_max; ++n) {
          if (n > 0) logp += std::log(lambda) - std::log((double)n);
          const int chunks = (n + 63) / 64;
          if (chunks) {
            const double r = shrink[std::min(chunks, 4)] * rb + rc;
            work += std::exp(logp) * chunks * r * r * r;
          }
        }
        const double cost = work / lambda;
        if (cost < best) {
          best = cost;
          m_out[0] = a, m_out[1] = b, m_out[2] = c;
        }
      }
}

int build_poses(Plan* P, int n, const double* frac_in, int n_axes, const double* axes, int grid,
                int tile_centers, Poses** out) {
  if (n < 0 || n_axes < 1) return fail(1, "n_centers >= 0 and n_axes >= 1 required");
  if (tile_centers < 1 || tile_centers > 1 << 20) return fail(1, "tile_centers must be in 1..2^20");
  std::vector<double> f(3 * (size_t)n);
  for (size_t i = 0; i < 3 * (size_t)n; ++i) {
    if (!std::isfinite(frac_in[i])) return fail(1, "nonfinite center");
    f[i] = wrap01(frac_in[i]);
  }
  auto S = new Poses();
  S->plan = P;
  S->n_centers = n;
  S->n_axes = n_axes;
  S->n_groups = (n_axes + 63) / 64;
  S->axes.resize(3 * (size_t)n_axes);
  for (int k = 0; k < n_axes; ++k) {
    if (!unit_director(axes + 3 * k, &S->axes[3 * k])) {
      delete S;
      return fail(1, "finite nonzero directors required");
    }
  }
  int m[3];
  choose_tiles(*P, std::max(n, 1), grid, tile_centers, m);
  const long long nbins = (long long)m[0] * m[1] * m[2];
  std::vector<int> bin(n);
  std::vector<long long> cnt(nbins + 1, 0);
  for (int i = 0; i < n; ++i) {
    int idx[3];
    for (int k = 0; k < 3; ++k) idx[k] = std::min((int)(f[3 * i + k] * m[k]), m[k] - 1);
    bin[i] = (idx[0] * m[1] + idx[1]) * m[2] + idx[2];
    ++cnt[bin[i] + 1];
  }
  for (long long b = 0; b < nbins; ++b) cnt[b + 1] += cnt[b];
  std::vector<int> order(n);
  {
    std::vector<long long> pos(cnt.begin(), cnt.end() - 1);
    for (int i = 0; i < n; ++i) order[pos[bin[i]]++] = i;
  }
  std::vector<int> tile_of_bin(nbins, -1);
  for (long long b = 0; b < nbins; ++b) {
    if (cnt[b + 1] == cnt[b]) continue;
    Tile T{};
    const int ia = (int)(b / ((long long)m[1] * m[2]));
    const int ib = (int)((b / m[2]) % m[1]);
    const int ic = (int)(b % m[2]);
    T.frac[0] = (ia + 0.5) / m[0];
    T.frac[1] = (ib + 0.5) / m[1];
    T.frac[2] = (ic + 0.5) / m[2];
    T.first_member = (int)cnt[b];
    T.n_members = (int)(cnt[b + 1] - cnt[b]);
    tile_of_bin[b] = (int)S->tiles.size();
    S->tiles.push_back(T);
  }
  S->center_id = order;
  S->tile_of.resize(n);
  S->center_rel.resize(3 * (size_t)n);
  for (int s = 0; s < n; ++s) {
    const int i = order[s];
    const int t = tile_of_bin[bin[i]];
    S->tile_of[s] = t;
    Tile& T = S->tiles[t];
    const double da = f[3 * i] - T.frac[0], db = f[3 * i + 1] - T.frac[1],
                 dc = f[3 * i + 2] - T.frac[2];
    double r2 = 0.0;
    for (int k = 0; k < 3; ++k) {
      const double x = da * P->H[k] + db * P->H[3 + k] + dc * P->H[6 + k];
      S->center_rel[3 * (size_t)s + k] = x;
      r2 += x * x;
    }
    T.radius = std::max(T.radius, std::sqrt(r2));
  }
  for (Tile& T : S->tiles) T.list_r = T.radius + P->keep_r + 1e-3;

  hipError_t e;
  std::vector<float> a32(3 * (size_t)n_axes);
  for (size_t i = 0; i < a32.size(); ++i) a32[i] = (float)S->axes[i];
  if ((e = hipSetDevice(P->device)) != hipSuccess ||
      (e = S->d_center_rel.upload(S->center_rel.data(), S->center_rel.size(), P->stream)) != hipSuccess ||
      (e = S->d_center_id.upload(S->center_id.data(), n, P->stream)) != hipSuccess ||
      (e = S->d_tile_of.upload(S->tile_of.data(), n, P->stream)) != hipSuccess ||
      (e = S->d_axes64.upload(S->axes.data(), S->axes.size(), P->stream)) != hipSuccess ||
      (e = S->d_axes32.upload(a32.data(), a32.size(), P->stream)) != hipSuccess ||
      (e = hipStreamSynchronize(P->stream)) != hipSuccess) {
    delete S;
    return hip_fail(e, "pose upload");
  }
  *out = S;
  return 0;
}

int batch_build(int n_jobs, Plan* const* plans, Poses* const* poses, int precision,
                const double* temps, double* const* energy, float* const* orientation,
                float* const* marginal, float* const* grid, float* const* site, Batch** out) {
  if (n_jobs < 1) return fail(1, "at least one job required");
  if (precision != PREC_FP64 && precision != PREC_MIXED && precision != PREC_FP64_RCP)
    return fail(1, "precision must be 0, 1 or 2");
  const int device = plans[0]->device;
  auto B = new Batch();
  B->device = device;
  B->n_jobs = n_jobs;
  B->precision = precision;
  B->jobs.resize(n_jobs);
  B->job_work.resize(n_jobs + 1, 0);
  B->job_first.resize(n_jobs + 1, 0);
  long long segs = 0, work = 0, centers = 0;
  for (int j = 0; j < n_jobs; ++j) {
    Plan* P = plans[j];
    Poses* S = poses[j];
    if (!P || !S || S->plan != P || P->device != device) {
      delete B;
      return fail(1, "each job needs a plan and a pose set created from it, on one device");
    }
    Job& J = B->jobs[j];
    std::memset(&J, 0, sizeof(Job));
    plan_fill_job(*P, J);
    if (!J.sym) B->all_sym = false;
    J.center_rel = S->d_center_rel.p;
    J.center_id = S->d_center_id.p;
    J.tile_of = S->d_tile_of.p;
    J.axes64 = S->d_axes64.p;
    J.axes32 = S->d_axes32.p;
    J.n_cls = S->n_cls;
    J.list_scale = S->list_scale;
    J.split2 = S->split2;
    J.n_centers = S->n_centers;
    J.n_axes = S->n_axes;
    J.n_groups = S->n_groups;
    J.n_tiles = (int)S->tiles.size();
    J.tile_base = (int)B->tiles.size();
    J.seg_base = segs;
    J.work_begin = work;
    J.out64 = energy ? energy[j] : nullptr;
    J.out32 = orientation ? orientation[j] : nullptr;
    J.marg32 = marginal ? marginal[j] : nullptr;
    J.grid32 = grid ? grid[j] : nullptr;
    J.site32 = site ? site[j] : nullptr;
    J.temperature = temps ? temps[j] : 0.0;
    J.log_n_axes = std::log((double)J.n_axes);
    if ((J.marg32 || J.grid32) && !(temps && std::isfinite(temps[j]) && temps[j] > 0.0)) {
      delete B;
      return fail(1, "positive finite temperature required for the marginal");
    }
    if (J.grid32 && !J.marg32) {
      delete B;
      return fail(1, "grid output requires the marginal output");
    }
    if (J.marg32 && S->n_groups > 1) {
      if (!J.out64) {
        delete B;
        return fail(1, "marginal over more than 64 axes requires the FP64 energy output");
      }
      B->need_marginal = true;
    }
    if (J.site32) B->any_site = true;
    for (size_t t = 0; t < S->tiles.size(); ++t) {
      Tile T = S->tiles[t];
      T.job = j;
      B->tiles.push_back(T);
      B->seg_base.push_back(segs + (long long)t * P->n_types * S->n_cls);
      if (J.site32)
        for (int first = 0; first < T.n_members; first += 64)
          B->chunks.push_back({j, (int)t, T.first_member + first, std::min(64, T.n_members - first)});
    }
    segs += (long long)S->tiles.size() * P->n_types * S->n_cls;
    work += (long long)S->n_centers * S->n_groups;
    B->job_first[j] = centers;
    centers += S->n_centers;
    B->job_work[j + 1] = work;
    B->job_first[j + 1] = centers;
  }
  B->n_segs = segs;
  B->n_work = work;
  B->n_centers_total = centers;
  hipError_t e;
  Batch* b = B;
  if ((e = hipSetDevice(device)) != hipSuccess ||
      (e = hipStreamCreate(&b->stream)) != hipSuccess ||
      (e = b->d_jobs.upload(b->jobs.data(), b->jobs.size(), b->stream)) != hipSuccess ||
      (e = b->d_job_work.upload(b->job_work.data(), b->job_work.size(), b->stream)) != hipSuccess ||
      (e = b->d_job_first.upload(b->job_first.data(), b->job_first.size(), b->stream)) != hipSuccess ||
      (e = b->d_tiles.upload(b->tiles.data(), b->tiles.size(), b->stream)) != hipSuccess ||
      (e = b->d_seg_base.upload(b->seg_base.data(), b->seg_base.size(), b->stream)) != hipSuccess ||
      (e = b->d_chunks.upload(b->chunks.data(), b->chunks.size(), b->stream)) != hipSuccess ||
      (e = b->d_counts.reserve(std::max<long long>(segs, 1))) != hipSuccess ||
      (e = b->d_offsets.reserve(segs + 1)) != hipSuccess ||
      (e = b->d_scan.reserve(scan_scratch_size(segs))) != hipSuccess ||
      (e = b->d_bad.reserve(1)) != hipSuccess ||
      // the asynchronous pose-grid calls only ever OR into the flags and
      // clear them when they are read, so they must start cleared
      (e = hipMemsetAsync(b->d_bad.p, 0, sizeof(int), b->stream)) != hipSuccess ||
      (e = b->d_pairs.reserve(n_jobs)) != hipSuccess) {
    delete B;
    return hip_fail(e, "batch setup");
  }
  for (int j = 0; j < n_jobs; ++j) {
    b->jobs[j].bad = b->d_bad.p;
    b->jobs[j].pair_count = b->d_pairs.p + j;
  }
  if ((e = b->d_jobs.upload(b->jobs.data(), b->jobs.size(), b->stream)) != hipSuccess) {
    delete B;
    return hip_fail(e, "batch job table");
  }
  for (auto& ev : b->ev)
    if ((e = hipEventCreate(&ev)) != hipSuccess) {
      delete B;
      return hip_fail(e, "hipEventCreate");
    }
  if ((e = hipStreamSynchronize(b->stream)) != hipSuccess) {
    delete B;
    return hip_fail(e, "batch upload");
  }
  *out = B;
  return 0;
}

// Candidate lists: count -> scan -> fill.  The list size of a batch never
// changes, so only the first run reads it back (one host sync); later runs are
// fully asynchronous.
int batch_lists(Batch* B, bool timed) {
  const int nt = (int)B->tiles.size();
  if (timed) SV_HIP(hipEventRecord(B->ev[0], B->stream));
  SV_HIP(launch_tile_count(B->d_jobs.p, B->d_tiles.p, nt, B->d_seg_base.p, B->d_counts.p, B->stream));
  if (timed) SV_HIP(hipEventRecord(B->ev[1], B->stream));
  SV_HIP(launch_scan(B->d_counts.p, B->n_segs, B->d_offsets.p, B->d_scan.p, B->stream));
  if (timed) SV_HIP(hipEventRecord(B->ev[2], B->stream));
  if (B->list_total < 0) {
    long long total = 0;
    SV_HIP(hipMemcpyAsync(&total, B->d_offsets.p + B->n_segs, sizeof(total), hipMemcpyDeviceToHost,
                          B->stream));
    SV_HIP(hipStreamSynchronize(B->stream));
    if (total < 0) return fail(5, "negative candidate list size");
    SV_HIP(B->d_list32.reserve((size_t)std::max<long long>(total, 1)));
    SV_HIP(B->d_list64.reserve(3 * (size_t)std::max<long long>(total, 1)));
    B->list_total = total;
  }
  if (timed) SV_HIP(hipEventRecord(B->ev[3], B->stream));
  SV_HIP(launch_tile_fill(B->d_jobs.p, B->d_tiles.p, nt, B->d_seg_base.p, B->d_offsets.p,
                          B->d_list32.p, B->d_list64.p, B->stream));
  if (timed) SV_HIP(hipEventRecord(B->ev[4], B->stream));
  B->lists_built = true;
  return 0;
}

}  // namespace

// A grid of m[0] x m[1] x m[2] bins over the cell.  Every bin is a tile whose
// list radius covers the whole bin (the circumscribed sphere of the sub-cell
// a/m0, b/m1, c/m2 about its centre).  Site lists add rc, are split per type
// at rsite + bin radius (class 0: images a member site can bring within its
// near-contact radius) and hold FP32 coordinates in units of rc; center lists
// add rc + max|o|.  The lists are built here, once.
int pose_grid_build(Plan* P, const int m_in[3], bool site_lists, PoseGrid** out) {
  for (int k = 0; k < 3; ++k)
    if (m_in[k] < 1 || m_in[k] > 1024) return fail(1, "tile counts must be in 1..1024");
  const long long nb = (long long)m_in[0] * m_in[1] * m_in[2];
  if (nb > (1 << 18)) return fail(1, "pose grid would need more than 2^18 tiles");
  hipError_t e;
  if ((e = hipSetDevice(P->device)) != hipSuccess) return hip_fail(e, "hipSetDevice");
  auto G = new PoseGrid();
  G->P = P;
  for (int k = 0; k < 3; ++k) G->m[k] = m_in[k];
  G->n_bins = (int)nb;
  std::memcpy(G->H.m, P->H, sizeof(G->H.m));
  std::memcpy(G->Hinv.m, P->Hinv, sizeof(G->Hinv.m));
  double radius = 0.0;
  for (int sb = -1; sb <= 1; sb += 2)
    for (int sc = -1; sc <= 1; sc += 2) {
      double r2 = 0.0;
      for (int k = 0; k < 3; ++k) {
        const double x = P->H[k] / G->m[0] + sb * P->H[3 + k] / G->m[1] + sc * P->H[6 + k] / G->m[2];
        r2 += x * x;
      }
      radius = std::max(radius, 0.5 * std::sqrt(r2));
    }
  radius *= 1.0 + 1e-12;
  G->radius = radius;
  {
    // sub-grid bits along the lattice axes by bin edge length, longest first
    int ax[3] = {0, 1, 2};
    double len[3];
    for (int k = 0; k < 3; ++k)
      len[k] = std::sqrt(P->H[3 * k] * P->H[3 * k] + P->H[3 * k + 1] * P->H[3 * k + 1] +
                         P->H[3 * k + 2] * P->H[3 * k + 2]) / G->m[k];
    std::stable_sort(ax, ax + 3, [&](int a, int b) { return len[a] > len[b]; });
    G->order = ax[0] | ax[1] << 2 | ax[2] << 4;
  }
  std::vector<double> split2(P->n_types);
  for (int t = 0; t < P->n_types; ++t) {
    const double r = P->rsite[t] + radius + 1e-3;
    split2[t] = r * r;
  }
  if ((e = G->split2.upload(split2.data(), split2.size(), P->stream)) != hipSuccess ||
      (e = hipStreamSynchronize(P->stream)) != hipSuccess) {
    delete G;
    return hip_fail(e, "site grid tables");
  }
  auto S = new Poses();
  G->S = S;
  S->plan = P;
  S->n_axes = 1;
  S->n_groups = 1;
  S->per_pose = true;
  if (site_lists) {
    S->n_cls = 2;
    S->list_scale = 1.0 / P->cutoff;
    S->split2 = G->split2.p;
  }
  S->tiles.resize(G->n_bins);
  for (int b = 0; b < G->n_bins; ++b) {
    Tile& T = S->tiles[b];
    T = Tile{};
    const int idx[3] = {b / (G->m[1] * G->m[2]), (b / G->m[2]) % G->m[1], b % G->m[2]};
    for (int k = 0; k < 3; ++k) T.frac[k] = (idx[k] + 0.5) / G->m[k];
    T.radius = radius;
    T.list_r = radius + (site_lists ? P->cutoff : P->keep_r) + 1e-3;
  }
  Plan* pl = P;
  Poses* po = S;
  int rc = batch_build(1, &pl, &po, PREC_FP64, nullptr, nullptr, nullptr, nullptr, nullptr,
                       nullptr, &G->B);
  if (!rc) rc = batch_lists(G->B, false);
  if (!rc && ((e = G->counts.reserve(64 * (size_t)G->n_bins)) != hipSuccess ||
              (e = G->first.reserve(64 * (size_t)G->n_bins)) != hipSuccess ||
              (e = G->totals.reserve(1)) != hipSuccess ||
              // the binning pass expects cleared cell counts and clears them after use
              (e = hipMemsetAsync(G->counts.p, 0, 64 * sizeof(int) * (size_t)G->n_bins,
                                  G->B->stream)) != hipSuccess ||
              (e = hipStreamSynchronize(G->B->stream)) != hipSuccess))
    rc = hip_fail(e, "pose grid buffers");
  if (rc) {
    delete G;
    return rc;
  }
  *out = G;
  return 0;
}

namespace {

int batch_run(Batch* B, bool timed, bool sync) {
  SV_HIP(hipSetDevice(B->device));
  int rc = batch_lists(B, timed);
  if (rc) return rc;
  SV_HIP(launch_zero(B->d_bad.p, 1, B->stream));
  if (timed) SV_HIP(hipEventRecord(B->ev[5], B->stream));
  if (B->persistent_blocks > 0 && B->precision == PREC_MIXED)
    SV_HIP(launch_field_persistent(B->d_jobs.p, B->n_jobs, B->d_job_work.p, B->n_work,
                                   B->d_counters.p, B->persistent_blocks, B->d_offsets.p,
                                   B->d_list32.p, B->d_list64.p, B->stream, B->all_sym));
  else
    SV_HIP(launch_field(B->precision, B->d_jobs.p, B->n_jobs, B->d_job_work.p, B->n_work,
                        B->d_offsets.p, B->d_list32.p, B->d_list64.p, B->stream, B->all_sym));
  if (timed) SV_HIP(hipEventRecord(B->ev[6], B->stream));
  if (B->need_marginal)
    SV_HIP(launch_marginal(B->d_jobs.p, B->n_jobs, B->d_job_first.p, B->n_centers_total, B->stream));
  if (timed) SV_HIP(hipEventRecord(B->ev[7], B->stream));
  // (the site field is fused into the field kernel)
  if (timed) SV_HIP(hipEventRecord(B->ev[8], B->stream));
  B->timed_last = timed;
  if (!sync) return 0;
  int bad = 0;
  SV_HIP(hipMemcpyAsync(&bad, B->d_bad.p, sizeof(int), hipMemcpyDeviceToHost, B->stream));
  SV_HIP(hipStreamSynchronize(B->stream));
  if (bad) return fail(4, "nonfinite Vext energy");
  return 0;
}

// Bound on the repulsive-wall slope (K/A) of pairs left to FP32 arithmetic.
constexpr double kNearSlope = 300.0;

int plan_create(const double* cell, int n_atoms, const double* frac, const double* sigma,
                const double* epsilon, double gs, double ge, const double* offsets, double cutoff,
                double near_energy, Plan** out) {
  if (!cell || !frac || !sigma || !epsilon || !offsets || n_atoms < 1)
    return fail(1, "cell, atoms (n_atoms >= 1), LJ parameters and offsets required");
  for (int i = 0; i < 9; ++i)
    if (!std::isfinite(cell[i])) return fail(1, "nonfinite cell");
  if (!(std::isfinite(cutoff) && cutoff > 0.0)) return fail(1, "positive finite cutoff required");
  if (!(std::isfinite(gs) && gs > 0.0 && std::isfinite(ge) && ge > 0.0))
    return fail(1, "positive guest sigma/epsilon required");
  if (!(std::isfinite(offsets[0]) && std::isfinite(offsets[1])))
    return fail(1, "finite site offsets required");
  auto P = new Plan();
  std::memcpy(P->H, cell, sizeof(P->H));
  double det;
  if (!invert3(P->H, P->Hinv, &det) || det <= 0.0) {
    delete P;
    return fail(1, "cell matrix must have a positive determinant");
  }
  P->volume = det;
  for (int k = 0; k < 3; ++k)
    P->hn[k] = std::sqrt(P->Hinv[k] * P->Hinv[k] + P->Hinv[3 + k] * P->Hinv[3 + k] +
                         P->Hinv[6 + k] * P->Hinv[6 + k]);
  P->n_atoms = n_atoms;
  P->o0 = offsets[0];
  P->o1 = offsets[1];
  P->cutoff = cutoff;
  P->keep_r = cutoff + std::max(std::fabs(P->o0), std::fabs(P->o1)) + 1e-3;
  P->near_energy = near_energy;
  // Distinct (sigma, epsilon) pairs in order of first appearance.
  std::map<std::pair<double, double>, int> type_of;
  std::vector<std::pair<double, double>> types;
  std::vector<int> atype(n_atoms);
  for (int i = 0; i < n_atoms; ++i) {
    for (int k = 0; k < 3; ++k)
      if (!std::isfinite(frac[3 * i + k])) {
        delete P;
        return fail(1, "nonfinite framework coordinate");
      }
    if (!(std::isfinite(sigma[i]) && sigma[i] > 0.0 && std::isfinite(epsilon[i]) && epsilon[i] > 0.0)) {
      delete P;
      return fail(1, "invalid framework LJ parameters");
    }
    auto key = std::make_pair(sigma[i], epsilon[i]);
    auto it = type_of.find(key);
    if (it == type_of.end()) {
      it = type_of.emplace(key, (int)types.size()).first;
      types.push_back(key);
    }
    atype[i] = it->second;
  }
  P->n_types = (int)types.size();
  P->type_start.assign(P->n_types + 1, 0);
  for (int i = 0; i < n_atoms; ++i) ++P->type_start[atype[i] + 1];
  for (int t = 0; t < P->n_types; ++t) P->type_start[t + 1] += P->type_start[t];
  P->frac.resize(3 * (size_t)n_atoms);
  {
    std::vector<int> pos(P->type_start.begin(), P->type_start.end() - 1);
    for (int i = 0; i < n_atoms; ++i) {
      const int s = pos[atype[i]]++;
      for (int k = 0; k < 3; ++k) P->frac[3 * (size_t)s + k] = wrap01(frac[3 * i + k]);
    }
  }
  // Mixed coefficients, same operation order as the reference numpy code
  // (neighbor.py:21-28; physics.py:76-78).
  const double omax = std::max(std::fabs(P->o0), std::fabs(P->o1));
  for (int t = 0; t < P->n_types; ++t) {
    const double ms = (types[t].first + gs) / 2.0;
    const double me = std::sqrt(types[t].second * ge);
    const double s6 = std::pow(ms, 6.0);
    const double c12 = 4.0 * me * (s6 * s6);
    const double c6 = 4.0 * me * s6;
    const double cut6 = std::pow(ms / cutoff, 6.0);
    const double sh = -4.0 * me * (cut6 * cut6 - cut6);
    const double fl = (0.1 * ms) * (0.1 * ms);
    P->c12.push_back(c12);
    P->c6.push_back(c6);
    P->shift.push_back(sh);
    P->floor2.push_back(fl);
    P->sigma.push_back(ms);
    P->eps.push_back(me);
    P->s6f.push_back((float)(c12 / c6));
    P->floor2f.push_back((float)fl);
    // coefficients of the shell form, in units of rc^2
    const double rc6 = cutoff * cutoff * cutoff * cutoff * cutoff * cutoff;
    P->c6s.push_back(c6 / rc6);
    P->s6sf.push_back((float)(c12 / c6 / rc6));
    // ... and of the pose form, in units of 1/s with s the FP32 value of 1/rc^2
    // itself, so that only the clamp point (not the terms) carries its rounding
    const double s = (double)(float)(1.0 / (cutoff * cutoff));
    P->c6x.push_back(c6 * s * s * s);
    P->s6xf.push_back((float)(c12 / c6 * s * s * s));
    float near2 = 0.0f;
    // the site form of independent poses takes the same radii about one site
    double r_site = std::sqrt(fl) + 1e-3;
    if (near_energy > 0.0) {
      // Repulsive-branch distance where 4 eps (x^2 - x) = near_energy, x = (sigma/r)^6.
      const double x = 0.5 * (1.0 + std::sqrt(1.0 + near_energy / me));
      const double r_energy = ms * std::pow(x, -1.0 / 6.0);
      // ... and where the wall's slope |dphi/dr| = 24 eps/r (2 x^2 - x) drops to
      // kNearSlope: an FP32 distance error of ~2e-7 A then costs <= 6e-5 K per
      // term.  |dphi/dr| decreases monotonically on (0, 2^(1/6) sigma).
      double lo = 0.5 * ms, hi = std::pow(2.0, 1.0 / 6.0) * ms;
      for (int it = 0; it < 80; ++it) {
        const double r = 0.5 * (lo + hi);
        const double q = std::pow(ms / r, 6.0);
        const double slope = 24.0 * me / r * (2.0 * q * q - q);
        if (slope > kNearSlope) lo = r; else hi = r;
      }
      const double rn = std::max(r_energy, hi) + omax;
      near2 = (float)(rn * rn);
      r_site = std::max(r_site, std::max(r_energy, hi));
    }
    P->near2f.push_back(near2);
    P->rsite.push_back(r_site);
    P->nsite2f.push_back((float)((r_site / cutoff) * (r_site / cutoff)));
    const double lo = std::sqrt(fl) + omax + 1e-3;
    P->lo2f.push_back((float)(lo * lo));
  }
  // An atom is near-contact when its center distance is below
  // sqrt(max(near2f, lo2f)); both of its sites are then within that + omax.
  P->near_inside = true;
  for (int t = 0; t < P->n_types; ++t) {
    const double r = std::sqrt((double)std::max(P->near2f[t], P->lo2f[t]));
    if (!(r * (1.0 + 1e-6) + omax < cutoff)) P->near_inside = false;
  }
  std::vector<FieldType> ftype(P->n_types);
  for (int t = 0; t < P->n_types; ++t)
    ftype[t] = FieldType{P->c12[t], P->c6[t], P->shift[t], P->floor2[t], P->c6s[t], P->s6f[t],
                         P->floor2f[t], P->near2f[t], P->lo2f[t], P->s6sf[t], 0.0f};
  hipError_t e;
  if ((e = hipGetDevice(&P->device)) != hipSuccess || (e = hipStreamCreate(&P->stream)) != hipSuccess ||
      (e = P->d_ftype.upload(ftype.data(), ftype.size(), P->stream)) != hipSuccess ||
      (e = P->d_frac.upload(P->frac.data(), P->frac.size(), P->stream)) != hipSuccess ||
      (e = P->d_type_start.upload(P->type_start.data(), P->type_start.size(), P->stream)) != hipSuccess ||
      (e = P->d_c12.upload(P->c12.data(), P->n_types, P->stream)) != hipSuccess ||
      (e = P->d_c6.upload(P->c6.data(), P->n_types, P->stream)) != hipSuccess ||
      (e = P->d_shift.upload(P->shift.data(), P->n_types, P->stream)) != hipSuccess ||
      (e = P->d_floor2.upload(P->floor2.data(), P->n_types, P->stream)) != hipSuccess ||
      (e = P->d_sigma.upload(P->sigma.data(), P->n_types, P->stream)) != hipSuccess ||
      (e = P->d_eps.upload(P->eps.data(), P->n_types, P->stream)) != hipSuccess ||
      (e = P->d_s6f.upload(P->s6f.data(), P->n_types, P->stream)) != hipSuccess ||
      (e = P->d_floor2f.upload(P->floor2f.data(), P->n_types, P->stream)) != hipSuccess ||
      (e = P->d_near2f.upload(P->near2f.data(), P->n_types, P->stream)) != hipSuccess ||
      (e = P->d_lo2f.upload(P->lo2f.data(), P->n_types, P->stream)) != hipSuccess ||
      (e = P->d_c6s.upload(P->c6s.data(), P->n_types, P->stream)) != hipSuccess ||
      (e = P->d_s6sf.upload(P->s6sf.data(), P->n_types, P->stream)) != hipSuccess ||
      (e = P->d_c6x.upload(P->c6x.data(), P->n_types, P->stream)) != hipSuccess ||
      (e = P->d_s6xf.upload(P->s6xf.data(), P->n_types, P->stream)) != hipSuccess ||
      (e = P->d_nsite2f.upload(P->nsite2f.data(), P->n_types, P->stream)) != hipSuccess ||
      (e = hipStreamSynchronize(P->stream)) != hipSuccess) {
    delete P;
    return hip_fail(e, "plan upload");
  }
  *out = P;
  return 0;
}

// ---------------------------------------------------------------------------
// Drop-in CSR entry point state (shared by all callers, serialized).
// ---------------------------------------------------------------------------
struct LegacyCtx {
  int device = -1;
  hipStream_t stream = nullptr;
  DevBuf<double> centers, axes, images, offsets, c12, c6, shift, floor2, output;
  DevBuf<long long> starts;
  DevBuf<int> indices, bad;
};
std::mutex g_legacy_mutex;
LegacyCtx* g_legacy = nullptr;

// The reference kernel's input checks in its loop order (kernel.cpp:11-25):
// the first center whose starts decrease (2) or whose range holds an index
// out of [0, n_images) (3) stops the evaluation; the centers before it are
// computed.  Sets *stop (centers to evaluate) and *code.
void legacy_checks(int n_centers, int n_images, const int64_t* starts, const int32_t* indices,
                   int* stop, int* code) {
  *stop = n_centers;
  *code = 0;
  for (int c = 0; c < n_centers; ++c) {
    if (starts[c] > starts[c + 1] || starts[c] < 0) {
      *stop = c, *code = 2;
      return;
    }
    for (int64_t n = starts[c]; n < starts[c + 1]; ++n)
      if (indices[n] < 0 || indices[n] >= n_images) {
        *stop = c, *code = 3;
        return;
      }
  }
}

// ---------------------------------------------------------------------------
// Scratch of the allocation-free drop-in entry point: one copy per device,
// held until the process ends.  Host code reaches it only through the
// address hipGetSymbolAddress returns for the current device.
// ---------------------------------------------------------------------------
constexpr size_t kScratchBytes = size_t(64) << 20;
constexpr int kMaxDevices = 64;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];
unsigned char* g_scratch[kMaxDevices] = {};
std::mutex g_scratch_mutex;

// Byte offsets of the drop-in call's arrays in the scratch, each 256-aligned.
struct ScratchLayout {
  size_t centers, axes, images, offsets, c12, c6, shift, floor2, starts, indices, output, bad,
      total;
};

ScratchLayout scratch_layout(int n_centers, int n_axes, int n_images, int64_t nnz) {
  ScratchLayout L;
  size_t at = 0;
  auto take = [&](size_t bytes) {
    const size_t here = at;
    at += (bytes + 255) / 256 * 256;
    return here;
  };
  L.centers = take(3 * sizeof(double) * (size_t)n_centers);
  L.axes = take(3 * sizeof(double) * (size_t)n_axes);
  L.images = take(3 * sizeof(double) * (size_t)n_images);
  L.offsets = take(2 * sizeof(double));
  L.c12 = take(sizeof(double) * (size_t)n_images);
  L.c6 = take(sizeof(double) * (size_t)n_images);
  L.shift = take(sizeof(double) * (size_t)n_images);
  L.floor2 = take(sizeof(double) * (size_t)n_images);
  L.starts = take(sizeof(int64_t) * ((size_t)n_centers + 1));
  L.indices = take(sizeof(int32_t) * (size_t)std::max<int64_t>(nnz, 1));
  L.output = take(sizeof(double) * (size_t)n_centers * (size_t)n_axes);
  L.bad = take(sizeof(int));
  L.total = at;
  return L;
}

}  // namespace
}  // namespace sv

using namespace sv;

extern "C" {

const char* sv_last_error(void) { return t_error.c_str(); }

int sv_abi_version(void) { return 5; }

int sv_device_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) return 0;
  return n;
}

int evaluate_reused_neighbors(int n_centers, int n_axes, int n_images, double cutoff2,
                              const double* centers, const double* axes, const double* images,
                              const double* offsets, const double* coeff12, const double* coeff6,
                              const double* shift, const double* floor2, const int64_t* starts,
                              const int32_t* indices, double* output) {
  if (n_centers < 0 || n_axes < 1 || n_images < 1 || cutoff2 <= 0) return 1;
  if (n_centers == 0) return 0;
  int stop = n_centers, stop_code = 0;
  legacy_checks(n_centers, n_images, starts, indices, &stop, &stop_code);
  if (stop == 0) return stop_code;
  std::lock_guard<std::mutex> lock(g_legacy_mutex);
  if (!g_legacy) g_legacy = new LegacyCtx();
  LegacyCtx& L = *g_legacy;
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return fail(5, "no HIP device");
  if (L.device != dev) {
    if (L.stream) (void)hipStreamDestroy(L.stream);
    L.stream = nullptr;
    L.device = dev;
    if (hipStreamCreate(&L.stream) != hipSuccess) return fail(5, "hipStreamCreate failed");
  }
  const int64_t nnz = starts[stop];
  hipStream_t s = L.stream;
  SV_HIP(L.centers.upload(centers, 3 * (size_t)stop, s));
  SV_HIP(L.axes.upload(axes, 3 * (size_t)n_axes, s));
  SV_HIP(L.images.upload(images, 3 * (size_t)n_images, s));
  SV_HIP(L.offsets.upload(offsets, 2, s));
  SV_HIP(L.c12.upload(coeff12, n_images, s));
  SV_HIP(L.c6.upload(coeff6, n_images, s));
  SV_HIP(L.shift.upload(shift, n_images, s));
  SV_HIP(L.floor2.upload(floor2, n_images, s));
  SV_HIP(L.starts.upload((const long long*)starts, (size_t)stop + 1, s));
  SV_HIP(L.indices.upload(indices, (size_t)std::max<int64_t>(nnz, 1), s));
  SV_HIP(L.output.reserve((size_t)stop * n_axes, true));
  SV_HIP(L.bad.reserve(1));
  const int no_bad = 0x7fffffff;
  SV_HIP(hipMemcpyAsync(L.bad.p, &no_bad, sizeof(int), hipMemcpyHostToDevice, s));
  SV_HIP(launch_legacy(stop, n_axes, cutoff2, L.centers.p, L.axes.p, L.images.p, L.offsets.p,
                       L.c12.p, L.c6.p, L.shift.p, L.floor2.p, L.starts.p, L.indices.p,
                       L.output.p, L.bad.p, s));
  int bad_center = no_bad;
  SV_HIP(hipMemcpyAsync(output, L.output.p, (size_t)stop * n_axes * sizeof(double),
                        hipMemcpyDeviceToHost, s));
  SV_HIP(hipMemcpyAsync(&bad_center, L.bad.p, sizeof(int), hipMemcpyDeviceToHost, s));
  SV_HIP(hipStreamSynchronize(s));
  if (bad_center < stop) return 4;
  return stop_code;
}

int evaluate_reused_neighbors_noalloc(int n_centers, int n_axes, int n_images, double cutoff2,
                                      const double* centers, const double* axes,
                                      const double* images, const double* offsets,
                                      const double* coeff12, const double* coeff6,
                                      const double* shift, const double* floor2,
                                      const int64_t* starts, const int32_t* indices,
                                      double* output) {
  if (n_centers < 0 || n_axes < 1 || n_images < 1 || cutoff2 <= 0) return 1;
  if (n_centers == 0) return 0;
  int stop = n_centers, stop_code = 0;
  legacy_checks(n_centers, n_images, starts, indices, &stop, &stop_code);
  if (stop == 0) return stop_code;
  const ScratchLayout L = scratch_layout(stop, n_axes, n_images, starts[stop]);
  if (L.total > kScratchBytes)
    return fail(1, "evaluate_reused_neighbors_noalloc: " + std::to_string(L.total) +
                       " bytes of device staging exceed its " +
                       std::to_string(kScratchBytes >> 20) + " MiB scratch");
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return fail(5, "no HIP device");
  if (dev < 0 || dev >= kMaxDevices) return fail(5, "device ordinal beyond the scratch table");
  if (!g_scratch[dev]) {
    void* p = nullptr;
    SV_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    g_scratch[dev] = static_cast<unsigned char*>(p);
  }
  unsigned char* base = g_scratch[dev];
  auto at = [&](size_t off) { return base + off; };
  const hipStream_t s = nullptr;   // the device's null stream: no stream is created
  const int64_t nnz = starts[stop];
  SV_HIP(hipMemcpyAsync(at(L.centers), centers, 3 * sizeof(double) * (size_t)stop,
                        hipMemcpyHostToDevice, s));
  SV_HIP(hipMemcpyAsync(at(L.axes), axes, 3 * sizeof(double) * (size_t)n_axes,
                        hipMemcpyHostToDevice, s));
  SV_HIP(hipMemcpyAsync(at(L.images), images, 3 * sizeof(double) * (size_t)n_images,
                        hipMemcpyHostToDevice, s));
  SV_HIP(hipMemcpyAsync(at(L.offsets), offsets, 2 * sizeof(double), hipMemcpyHostToDevice, s));
  SV_HIP(hipMemcpyAsync(at(L.c12), coeff12, sizeof(double) * (size_t)n_images,
                        hipMemcpyHostToDevice, s));
  SV_HIP(hipMemcpyAsync(at(L.c6), coeff6, sizeof(double) * (size_t)n_images,
                        hipMemcpyHostToDevice, s));
  SV_HIP(hipMemcpyAsync(at(L.shift), shift, sizeof(double) * (size_t)n_images,
                        hipMemcpyHostToDevice, s));
  SV_HIP(hipMemcpyAsync(at(L.floor2), floor2, sizeof(double) * (size_t)n_images,
                        hipMemcpyHostToDevice, s));
  SV_HIP(hipMemcpyAsync(at(L.starts), starts, sizeof(int64_t) * ((size_t)stop + 1),
                        hipMemcpyHostToDevice, s));
  if (nnz > 0)
    SV_HIP(hipMemcpyAsync(at(L.indices), indices, sizeof(int32_t) * (size_t)nnz,
                          hipMemcpyHostToDevice, s));
  // the flag keeps the last call's value: it is set before every launch
  const int no_bad = 0x7fffffff;
  SV_HIP(hipMemcpyAsync(at(L.bad), &no_bad, sizeof(int), hipMemcpyHostToDevice, s));
  SV_HIP(launch_legacy(stop, n_axes, cutoff2, reinterpret_cast<const double*>(at(L.centers)),
                       reinterpret_cast<const double*>(at(L.axes)),
                       reinterpret_cast<const double*>(at(L.images)),
                       reinterpret_cast<const double*>(at(L.offsets)),
                       reinterpret_cast<const double*>(at(L.c12)),
                       reinterpret_cast<const double*>(at(L.c6)),
                       reinterpret_cast<const double*>(at(L.shift)),
                       reinterpret_cast<const double*>(at(L.floor2)),
                       reinterpret_cast<const long long*>(at(L.starts)),
                       reinterpret_cast<const int*>(at(L.indices)),
                       reinterpret_cast<double*>(at(L.output)),
                       reinterpret_cast<int*>(at(L.bad)), s));
  int bad_center = no_bad;
  SV_HIP(hipMemcpyAsync(output, at(L.output), sizeof(double) * (size_t)stop * n_axes,
                        hipMemcpyDeviceToHost, s));
  SV_HIP(hipMemcpyAsync(&bad_center, at(L.bad), sizeof(int), hipMemcpyDeviceToHost, s));
  SV_HIP(hipStreamSynchronize(s));
  if (bad_center < stop) return 4;
  return stop_code;
}

int sv_plan_create(const double* cell, int n_atoms, const double* frac, const double* sigma,
                   const double* epsilon, double guest_sigma, double guest_epsilon,
                   const double* offsets, double cutoff, double near_energy_K, void** plan) {
  if (!plan) return fail(1, "null output handle");
  Plan* P = nullptr;
  int rc = plan_create(cell, n_atoms, frac, sigma, epsilon, guest_sigma, guest_epsilon, offsets,
                       cutoff, near_energy_K, &P);
  *plan = P;
  return rc;
}

int sv_plan_destroy(void* plan) {
  delete static_cast<Plan*>(plan);
  return 0;
}

int sv_plan_info(void* plan, int* n_types, double* keep_radius) {
  if (!plan) return fail(1, "null plan");
  Plan* P = static_cast<Plan*>(plan);
  if (n_types) *n_types = P->n_types;
  if (keep_radius) *keep_radius = P->keep_r;
  return 0;
}

int sv_poses_create_ex(void* plan, int n_centers, const double* centers_frac, int n_axes,
                       const double* axes, int tile_centers, void** poses) {
  if (!plan || !poses || (n_centers > 0 && !centers_frac) || !axes) return fail(1, "null argument");
  Poses* S = nullptr;
  int rc = build_poses(static_cast<Plan*>(plan), n_centers, centers_frac, n_axes, axes, 0,
                       tile_centers, &S);
  *poses = S;
  return rc;
}

int sv_poses_create(void* plan, int n_centers, const double* centers_frac, int n_axes,
                    const double* axes, void** poses) {
  return sv_poses_create_ex(plan, n_centers, centers_frac, n_axes, axes, 64, poses);
}

int sv_poses_grid(void* plan, int size, int n_axes, const double* axes, void** poses) {
  return sv_poses_grid_ex(plan, size, n_axes, axes, 64, poses);
}

int sv_poses_grid_ex(void* plan, int size, int n_axes, const double* axes, int tile_centers,
                     void** poses) {
  if (!plan || !poses || !axes) return fail(1, "null argument");
  if (size < 1 || (long long)size * size * size > 0x7fffffffLL) return fail(1, "invalid grid size");
  const long long n = (long long)size * size * size;
  std::vector<double> f(3 * (size_t)n);
  for (int ia = 0; ia < size; ++ia)
    for (int ib = 0; ib < size; ++ib)
      for (int ic = 0; ic < size; ++ic) {
        const size_t i = ((size_t)ia * size + ib) * size + ic;
        f[3 * i] = (ia + 0.5) / size;
        f[3 * i + 1] = (ib + 0.5) / size;
        f[3 * i + 2] = (ic + 0.5) / size;
      }
  Poses* S = nullptr;
  int rc = build_poses(static_cast<Plan*>(plan), (int)n, f.data(), n_axes, axes, size,
                       tile_centers, &S);
  *poses = S;
  return rc;
}

int sv_batch_set_persistent(void* batch, int blocks_per_cu) {
  if (!batch || blocks_per_cu < 0 || blocks_per_cu > 64) return fail(1, "invalid argument");
  Batch* B = static_cast<Batch*>(batch);
  if (blocks_per_cu == 0) {
    B->persistent_blocks = 0;
    return 0;
  }
  int cus = 0;
  SV_HIP(hipSetDevice(B->device));
  SV_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, B->device));
  SV_HIP(B->d_counters.reserve(8));
  B->persistent_blocks = std::max(1, cus) * blocks_per_cu;
  return 0;
}

int sv_poses_destroy(void* poses) {
  delete static_cast<Poses*>(poses);
  return 0;
}

int sv_batch_create(int n_jobs, void* const* plans, void* const* poses, int precision,
                    const double* temperatures, double* const* energy, float* const* orientation,
                    float* const* marginal, float* const* grid, float* const* site, void** batch) {
  if (!plans || !poses || !batch) return fail(1, "null argument");
  Batch* B = nullptr;
  int rc = batch_build(n_jobs, reinterpret_cast<Plan* const*>(plans),
                       reinterpret_cast<Poses* const*>(poses), precision, temperatures, energy,
                       orientation, marginal, grid, site, &B);
  *batch = B;
  return rc;
}

int sv_batch_run(void* batch, int time_kernels, int sync) {
  if (!batch) return fail(1, "null batch");
  return batch_run(static_cast<Batch*>(batch), time_kernels != 0, sync != 0);
}

int sv_batch_sync(void* batch) {
  if (!batch) return fail(1, "null batch");
  Batch* B = static_cast<Batch*>(batch);
  SV_HIP(hipSetDevice(B->device));
  int bad = 0;
  SV_HIP(hipMemcpyAsync(&bad, B->d_bad.p, sizeof(int), hipMemcpyDeviceToHost, B->stream));
  SV_HIP(hipStreamSynchronize(B->stream));
  if (bad) return fail(4, "nonfinite Vext energy");
  return 0;
}

int sv_batch_timings(void* batch, double* ms7) {
  if (!batch || !ms7) return fail(1, "null argument");
  Batch* B = static_cast<Batch*>(batch);
  if (!B->timed_last) return fail(1, "last run was not timed");
  SV_HIP(hipEventSynchronize(B->ev[8]));
  const int pairs[6][2] = {{0, 1}, {1, 2}, {3, 4}, {5, 6}, {6, 7}, {7, 8}};
  for (int i = 0; i < 6; ++i) {
    float ms = 0.f;
    SV_HIP(hipEventElapsedTime(&ms, B->ev[pairs[i][0]], B->ev[pairs[i][1]]));
    ms7[i] = ms;
  }
  float all = 0.f;
  SV_HIP(hipEventElapsedTime(&all, B->ev[0], B->ev[8]));
  ms7[6] = all;
  return 0;
}

int sv_batch_census(void* batch, unsigned long long* pairs_per_job) {
  if (!batch || !pairs_per_job) return fail(1, "null argument");
  Batch* B = static_cast<Batch*>(batch);
  SV_HIP(hipSetDevice(B->device));
  int rc = batch_lists(B, false);
  if (rc) return rc;
  SV_HIP(launch_zero_u64(B->d_pairs.p, B->n_jobs, B->stream));
  SV_HIP(launch_census(B->d_jobs.p, B->n_jobs, B->d_job_work.p, B->n_work, B->d_offsets.p,
                       B->d_list64.p, B->stream));
  SV_HIP(hipMemcpyAsync(pairs_per_job, B->d_pairs.p, sizeof(unsigned long long) * B->n_jobs,
                        hipMemcpyDeviceToHost, B->stream));
  SV_HIP(hipStreamSynchronize(B->stream));
  return 0;
}

int sv_batch_destroy(void* batch) {
  delete static_cast<Batch*>(batch);
  return 0;
}

int sv_field_run(void* plan, void* poses, int precision, double temperature, double* energy,
                 float* orientation, float* marginal, float* grid, float* site) {
  if (!plan || !poses) return fail(1, "null argument");
  Plan* P = static_cast<Plan*>(plan);
  Poses* S = static_cast<Poses*>(poses);
  SV_HIP(hipSetDevice(P->device));
  const size_t nc = (size_t)S->n_centers, np = nc * S->n_axes;
  double* d_e = nullptr;
  float *d_o = nullptr, *d_m = nullptr, *d_g = nullptr, *d_s = nullptr;
  const bool need64 = energy || (marginal && S->n_groups > 1);
  if (need64) {
    SV_HIP(S->w_out64.reserve(np));
    d_e = S->w_out64.p;
  }
  if (orientation) {
    SV_HIP(S->w_out32.reserve(np));
    d_o = S->w_out32.p;
  }
  if (marginal || grid) {
    SV_HIP(S->w_marg.reserve(nc));
    d_m = S->w_marg.p;
  }
  if (grid) {
    SV_HIP(S->w_grid.reserve(nc));
    d_g = S->w_grid.p;
  }
  if (site) {
    SV_HIP(S->w_site.reserve(nc));
    d_s = S->w_site.p;
  }
  Batch* B = nullptr;
  void* pl = plan;
  void* po = poses;
  int rc = batch_build(1, reinterpret_cast<Plan* const*>(&pl), reinterpret_cast<Poses* const*>(&po),
                       precision, &temperature, &d_e, &d_o, &d_m, &d_g, &d_s, &B);
  if (rc) return rc;
  rc = batch_run(B, false, true);
  const hipStream_t s = B->stream;
  if (rc == 0 || rc == 4) {
    hipError_t e = hipSuccess;
    if (energy && e == hipSuccess) e = hipMemcpyAsync(energy, d_e, np * 8, hipMemcpyDeviceToHost, s);
    if (orientation && e == hipSuccess) e = hipMemcpyAsync(orientation, d_o, np * 4, hipMemcpyDeviceToHost, s);
    if (marginal && e == hipSuccess) e = hipMemcpyAsync(marginal, d_m, nc * 4, hipMemcpyDeviceToHost, s);
    if (grid && e == hipSuccess) e = hipMemcpyAsync(grid, d_g, nc * 4, hipMemcpyDeviceToHost, s);
    if (site && e == hipSuccess) e = hipMemcpyAsync(site, d_s, nc * 4, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) rc = hip_fail(e, "field download");
  }
  delete B;
  return rc;
}

int sv_field_census(void* plan, void* poses, unsigned long long* pairs) {
  if (!plan || !poses || !pairs) return fail(1, "null argument");
  Batch* B = nullptr;
  void* pl = plan;
  void* po = poses;
  int rc = batch_build(1, reinterpret_cast<Plan* const*>(&pl), reinterpret_cast<Poses* const*>(&po),
                       PREC_FP64, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, &B);
  if (rc) return rc;
  rc = sv_batch_census(B, pairs);
  delete B;
  return rc;
}

// ---------------------------------------------------------------------------
// Device-resident pose grids (PoseGrid in sv_objects.h) and host pose lists.
// ---------------------------------------------------------------------------
int sv_pose_grid_create(void* plan, int n_expected, int tile_centers, void** grid) {
  if (!plan || !grid) return fail(1, "null argument");
  if (n_expected < 1) return fail(1, "n_expected must be >= 1");
  if (tile_centers < 0 || tile_centers > 1 << 20) return fail(1, "tile_centers must be in 0..2^20");
  Plan* P = static_cast<Plan*>(plan);
  int m[3];
  // bins of about tile_centers site positions (two per pose), or the
  // automatic choice
  if (tile_centers > 0)
    choose_tiles(*P, 2LL * n_expected, 0, tile_centers, m);
  else
    choose_site_bins(*P, 2LL * n_expected, m);
  PoseGrid* G = nullptr;
  const int rc = pose_grid_build(P, m, true, &G);
  if (rc) return rc;
  *grid = G;
  return 0;
}

}  // extern "C"

namespace sv {

Poses::~Poses() { delete list_grid; }

namespace {

// The run count: the low 32 bits of the packed running totals (little endian).
const int* n_runs_of(const PoseGrid* G) { return reinterpret_cast<const int*>(G->totals.p); }

// Enqueue the site pass and the pose sums on stream st (events around them
// when the grid times its phases); the sites of the poses are already binned.
int site_pass(PoseGrid* G, hipStream_t st, int precision, const double* grad_out, double* energy,
              double* grad_center, double* grad_axis) {
  Batch* B = G->B;
  Poses* S = G->S;
  const int n = G->binned_n;
  const long long max_runs = site_runs_bound(G->n_bins, 2LL * n);
  const bool grad = grad_center || grad_axis;
  if (G->timing) SV_HIP(hipEventRecord(G->ev[1], st));
  const SiteArrays a = {S->d_center_rel.p, S->d_center_id.p, G->site_e.p, G->site_g.p,
                        2LL * G->capacity, G->parts};
  SV_HIP(launch_site(precision, energy != nullptr, grad, B->d_jobs.p, G->runs.p, (int)max_runs,
                     n_runs_of(G), B->d_offsets.p, B->d_list32.p, B->d_list64.p, a, st));
  if (G->timing) SV_HIP(hipEventRecord(G->ev[2], st));
  SV_HIP(launch_site_combine(n, a, G->P->o0, G->P->o1, grad_out, energy, grad_center, grad_axis,
                             B->d_bad.p, st));
  if (G->timing) SV_HIP(hipEventRecord(G->ev[3], st));
  return 0;
}

// Enqueue one evaluation of a pose grid on stream st; no host synchronization.
// Flags of invalid input (2) and nonfinite results (1) accumulate in the
// batch's device flag.
int pose_grid_enqueue(PoseGrid* G, hipStream_t st, int n, const double* centers_cart,
                      const double* axes, int precision, const double* grad_out, double* energy,
                      double* grad_center, double* grad_axis) {
  if (n < 1) return fail(1, "at least one pose required");
  if (n > (1 << 29)) return fail(1, "at most 2^29 poses per call");
  if (!centers_cart || !axes) return fail(1, "null pose arrays");
  if (precision != PREC_FP64 && precision != PREC_MIXED) return fail(1, "precision must be 0 or 1");
  Batch* B = G->B;
  Poses* S = G->S;
  Plan* P = G->P;
  G->binned_n = 0;
  SV_HIP(hipSetDevice(P->device));
  const long long n_sites = 2LL * n;
  const long long max_runs = site_runs_bound(G->n_bins, n_sites);
  if (max_runs * kMaxSiteParts > INT32_MAX) return fail(1, "too many poses");
  // Work of earlier calls may still use the working buffers: it is ordered on
  // its own stream, which is drained before a call on another stream or
  // before the buffers are reallocated.
  if (G->last_stream_set && G->last_stream != st) SV_HIP(hipStreamSynchronize(G->last_stream));
  if (n > G->capacity || (size_t)max_runs > G->runs.cap) {
    SV_HIP(hipStreamSynchronize(st));
    if (n > G->capacity) {
      const size_t c = std::min<size_t>((size_t)n + (size_t)n / 4, 1u << 29);
      SV_HIP(S->d_center_rel.reserve(6 * c));
      SV_HIP(S->d_center_id.reserve(2 * c));
      SV_HIP(G->bin.reserve(2 * c));
      SV_HIP(G->rank.reserve(2 * c));
      SV_HIP(G->site_e.reserve(2 * c * kMaxSiteParts));
      SV_HIP(G->site_g.reserve(6 * c * kMaxSiteParts));
      G->capacity = (int)c;
    }
    SV_HIP(G->runs.reserve((size_t)max_runs, true));
  }
  G->last_stream = st;
  G->last_stream_set = true;
  if (G->timing) SV_HIP(hipEventRecord(G->ev[0], st));
  SV_HIP(launch_site_bin(n_sites, centers_cart, axes, P->o0, P->o1, G->Hinv, G->m, G->order,
                         G->bin.p, G->rank.p, G->counts.p, B->d_bad.p, G->totals.p, st));
  SV_HIP(launch_site_layout(G->counts.p, G->n_bins, G->totals.p, G->first.p, G->runs.p, st));
  SV_HIP(launch_site_scatter(n_sites, centers_cart, axes, P->o0, P->o1, G->H, G->Hinv, G->m,
                             G->bin.p, G->rank.p, G->first.p, G->counts.p, S->d_center_rel.p,
                             S->d_center_id.p, st));
  G->binned_n = n;
  return site_pass(G, st, precision, grad_out, energy, grad_center, grad_axis);
}

// Evaluates the poses binned by the grid's last enqueue again (new outputs,
// precision or gradient choice), without binning them again.  The binned
// copies live in the grid, so the caller's pose arrays are not read.
int pose_grid_reenqueue(PoseGrid* G, hipStream_t st, int precision, const double* grad_out,
                        double* energy, double* grad_center, double* grad_axis) {
  if (G->binned_n < 1) return fail(1, "no poses binned on this grid yet");
  if (precision != PREC_FP64 && precision != PREC_MIXED) return fail(1, "precision must be 0 or 1");
  SV_HIP(hipSetDevice(G->P->device));
  if (G->last_stream_set && G->last_stream != st) SV_HIP(hipStreamSynchronize(G->last_stream));
  G->last_stream = st;
  G->last_stream_set = true;
  if (G->timing) SV_HIP(hipEventRecord(G->ev[0], st));
  return site_pass(G, st, precision, grad_out, energy, grad_center, grad_axis);
}

// Wait for the grid's last stream, report and clear its flags.
int pose_grid_flags(PoseGrid* G) {
  Batch* B = G->B;
  const hipStream_t st = G->last_stream_set ? G->last_stream : B->stream;
  int bad = 0;
  SV_HIP(hipMemcpyAsync(&bad, B->d_bad.p, sizeof(int), hipMemcpyDeviceToHost, st));
  SV_HIP(launch_zero(B->d_bad.p, 1, st));
  SV_HIP(hipStreamSynchronize(st));
  if (bad & 2) return fail(1, "nonfinite pose center or director, or zero director");
  if (bad) return fail(4, "nonfinite pose energy or gradient");
  return 0;
}

}  // namespace
}  // namespace sv

extern "C" {

int sv_pose_list_create(void* plan, int n_poses, const double* centers_cart, const double* axes,
                        void** poses) {
  if (!plan || !poses || (n_poses > 0 && (!centers_cart || !axes))) return fail(1, "null argument");
  if (n_poses < 1) return fail(1, "at least one pose required");
  Plan* P = static_cast<Plan*>(plan);
  for (int i = 0; i < n_poses; ++i) {
    const double* c = centers_cart + 3 * (size_t)i;
    const double* u = axes + 3 * (size_t)i;
    const double nn = u[0] * u[0] + u[1] * u[1] + u[2] * u[2];
    if (!(std::isfinite(c[0]) && std::isfinite(c[1]) && std::isfinite(c[2])))
      return fail(1, "nonfinite pose center");
    if (!std::isfinite(nn) || nn <= 0.0) return fail(1, "finite nonzero directors required");
  }
  void* grid = nullptr;
  int rc = sv_pose_grid_create(plan, n_poses, 0, &grid);
  if (rc) return rc;
  auto S = new Poses();
  S->plan = P;
  S->per_pose = true;
  S->list_grid = static_cast<PoseGrid*>(grid);
  S->list_n = n_poses;
  hipError_t e;
  if ((e = S->list_centers.upload(centers_cart, 3 * (size_t)n_poses, P->stream)) != hipSuccess ||
      (e = S->list_axes.upload(axes, 3 * (size_t)n_poses, P->stream)) != hipSuccess ||
      (e = hipStreamSynchronize(P->stream)) != hipSuccess) {
    delete S;
    return hip_fail(e, "pose upload");
  }
  *poses = S;
  return 0;
}

int sv_pose_list_eval(void* poses, int precision, const double* grad_out, double* energy,
                      double* grad_center, double* grad_axis) {
  if (!poses) return fail(1, "null pose list");
  Poses* S = static_cast<Poses*>(poses);
  if (!S->list_grid) return fail(1, "pose set was not created by sv_pose_list_create");
  PoseGrid* G = S->list_grid;
  const hipStream_t st = G->B->stream;
  // the list's own grid keeps its binned sites between calls
  const int rc = G->binned_n == S->list_n
                     ? pose_grid_reenqueue(G, st, precision, grad_out, energy, grad_center,
                                           grad_axis)
                     : pose_grid_enqueue(G, st, S->list_n, S->list_centers.p, S->list_axes.p,
                                         precision, grad_out, energy, grad_center, grad_axis);
  if (rc) return rc;
  return pose_grid_flags(G);
}

int sv_pose_grid_eval(void* grid, int n, const double* centers_cart, const double* axes,
                      int precision, const double* grad_out, double* energy, double* grad_center,
                      double* grad_axis) {
  if (!grid) return fail(1, "null pose grid");
  PoseGrid* G = static_cast<PoseGrid*>(grid);
  SV_HIP(hipSetDevice(G->P->device));
  // flags of earlier asynchronous calls are not this call's
  if (G->last_stream_set) SV_HIP(hipStreamSynchronize(G->last_stream));
  SV_HIP(launch_zero(G->B->d_bad.p, 1, G->B->stream));
  const int rc = pose_grid_enqueue(G, G->B->stream, n, centers_cart, axes, precision, grad_out,
                                   energy, grad_center, grad_axis);
  if (rc) return rc;
  return pose_grid_flags(G);
}

int sv_pose_grid_eval_async(void* grid, void* stream, int n, const double* centers_cart,
                            const double* axes, int precision, const double* grad_out,
                            double* energy, double* grad_center, double* grad_axis) {
  if (!grid) return fail(1, "null pose grid");
  return pose_grid_enqueue(static_cast<PoseGrid*>(grid), static_cast<hipStream_t>(stream), n,
                           centers_cart, axes, precision, grad_out, energy, grad_center,
                           grad_axis);
}

int sv_pose_grid_reeval_async(void* grid, void* stream, int precision, const double* grad_out,
                              double* energy, double* grad_center, double* grad_axis) {
  if (!grid) return fail(1, "null pose grid");
  return pose_grid_reenqueue(static_cast<PoseGrid*>(grid), static_cast<hipStream_t>(stream),
                             precision, grad_out, energy, grad_center, grad_axis);
}

int sv_pose_grid_check(void* grid) {
  if (!grid) return fail(1, "null pose grid");
  PoseGrid* G = static_cast<PoseGrid*>(grid);
  SV_HIP(hipSetDevice(G->P->device));
  return pose_grid_flags(G);
}

int sv_pose_grid_info(void* grid, int* n_tiles, int* capacity) {
  if (!grid) return fail(1, "null pose grid");
  PoseGrid* G = static_cast<PoseGrid*>(grid);
  if (n_tiles) *n_tiles = G->n_bins;
  if (capacity) *capacity = G->capacity;
  return 0;
}

int sv_pose_grid_census(void* grid, unsigned long long* pairs) {
  if (!grid || !pairs) return fail(1, "null argument");
  PoseGrid* G = static_cast<PoseGrid*>(grid);
  if (G->binned_n < 1) return fail(1, "no poses binned on this grid yet");
  Batch* B = G->B;
  SV_HIP(hipSetDevice(G->P->device));
  const hipStream_t st = G->last_stream_set ? G->last_stream : B->stream;
  const long long max_runs = site_runs_bound(G->n_bins, 2LL * G->binned_n);
  SV_HIP(launch_zero_u64(B->d_pairs.p, 1, st));
  SV_HIP(launch_site_census(B->d_jobs.p, G->runs.p, (int)max_runs, n_runs_of(G),
                            B->d_offsets.p, B->d_list64.p, G->S->d_center_rel.p, B->d_pairs.p,
                            st));
  unsigned long long total = 0;
  SV_HIP(hipMemcpyAsync(&total, B->d_pairs.p, sizeof(total), hipMemcpyDeviceToHost, st));
  SV_HIP(hipStreamSynchronize(st));
  *pairs = total;
  return 0;
}

int sv_pose_grid_set_parts(void* grid, int parts) {
  if (!grid) return fail(1, "null pose grid");
  if (parts < 1 || parts > kMaxSiteParts) return fail(1, "parts must be in 1..8");
  PoseGrid* G = static_cast<PoseGrid*>(grid);
  // earlier evaluations read the old layout of the per-site results
  if (G->last_stream_set) SV_HIP(hipStreamSynchronize(G->last_stream));
  G->parts = parts;
  return 0;
}

int sv_pose_grid_set_timing(void* grid, int on) {
  if (!grid) return fail(1, "null pose grid");
  PoseGrid* G = static_cast<PoseGrid*>(grid);
  SV_HIP(hipSetDevice(G->P->device));
  if (on)
    for (auto& ev : G->ev)
      if (!ev) SV_HIP(hipEventCreate(&ev));
  G->timing = on != 0;
  return 0;
}

int sv_pose_grid_timings(void* grid, double* ms3) {
  if (!grid || !ms3) return fail(1, "null argument");
  PoseGrid* G = static_cast<PoseGrid*>(grid);
  if (!G->timing || G->binned_n < 1) return fail(1, "no timed evaluation on this grid");
  SV_HIP(hipSetDevice(G->P->device));
  SV_HIP(hipEventSynchronize(G->ev[3]));
  for (int k = 0; k < 3; ++k) {
    float ms = 0.0f;
    SV_HIP(hipEventElapsedTime(&ms, G->ev[k], G->ev[k + 1]));
    ms3[k] = ms;
  }
  return 0;
}

int sv_pose_grid_destroy(void* grid) {
  delete static_cast<PoseGrid*>(grid);
  return 0;
}

int sv_device_malloc(void** ptr, uint64_t bytes) {
  if (!ptr) return fail(1, "null argument");
  SV_HIP(hipMalloc(ptr, std::max<uint64_t>(bytes, 1)));
  return 0;
}

int sv_device_free(void* ptr) {
  if (ptr) SV_HIP(hipFree(ptr));
  return 0;
}

int sv_memcpy_d2h(void* dst, const void* src, uint64_t bytes) {
  SV_HIP(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
  return 0;
}

int sv_memcpy_h2d(void* dst, const void* src, uint64_t bytes) {
  SV_HIP(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
  return 0;
}

}  // extern "C"
