// Host-side objects of libsv_vext.so shared by the runtime translation units.
#pragma once
#include "sv_internal.h"

#include <algorithm>
#include <string>
#include <vector>

namespace sv {

// Device allocation owned by one host object (grown, never shrunk).
template <class T>
struct DevBuf {
  T* p = nullptr;
  size_t cap = 0;
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  ~DevBuf() { release(); }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    cap = 0;
  }
  hipError_t reserve(size_t n, bool grow = false) {
    if (n <= cap && p) return hipSuccess;
    release();
    size_t want = std::max<size_t>(n, 1);
    if (grow) want += want / 4;
    hipError_t e = hipMalloc((void**)&p, want * sizeof(T));
    if (e == hipSuccess) cap = want;
    else p = nullptr;
    return e;
  }
  hipError_t upload(const T* h, size_t n, hipStream_t s) {
    hipError_t e = reserve(n);
    if (e != hipSuccess || n == 0) return e;
    return hipMemcpyAsync(p, h, n * sizeof(T), hipMemcpyHostToDevice, s);
  }
};

struct Plan {
  int device = 0;
  hipStream_t stream = nullptr;
  double H[9], Hinv[9], hn[3], volume;
  int n_atoms = 0, n_types = 0;
  std::vector<double> frac;       // sorted by type, wrapped
  std::vector<int> type_start;
  std::vector<double> c12, c6, shift, floor2, sigma, eps, c6s, c6x;
  std::vector<float> s6f, floor2f, near2f, lo2f, s6sf, s6xf;
  std::vector<double> rsite;      // near-contact radius about one site (A)
  std::vector<float> nsite2f;     // (rsite / rc)^2
  double o0, o1, cutoff, keep_r, near_energy;
  bool near_inside = false;   // every FP64 near-contact pair is inside the cutoff
  DevBuf<double> d_frac, d_c12, d_c6, d_shift, d_floor2, d_sigma, d_eps, d_c6s, d_c6x;
  DevBuf<int> d_type_start;
  DevBuf<float> d_s6f, d_floor2f, d_near2f, d_lo2f, d_s6sf, d_s6xf, d_nsite2f;
  DevBuf<FieldType> d_ftype;
  ~Plan() {
    if (stream) (void)hipStreamDestroy(stream);
  }
};

struct Poses {
  Plan* plan = nullptr;
  int n_centers = 0, n_axes = 0, n_groups = 0;
  std::vector<double> axes;
  std::vector<Tile> tiles;        // job field left 0; filled per batch
  std::vector<int> center_id, tile_of;
  std::vector<double> center_rel;
  DevBuf<double> d_center_rel, d_axes64;
  DevBuf<float> d_axes32;
  DevBuf<int> d_center_id, d_tile_of;
  DevBuf<double> w_out64;
  DevBuf<float> w_out32, w_marg, w_grid, w_site;
  // candidate list layout of the batch jobs built from this set
  int n_cls = 1;                  // classes per (tile, type) segment
  double list_scale = 1.0;        // list32 coordinate scale
  const double* split2 = nullptr; // [n_types] class-0 radius^2 on the device (n_cls == 2)
  // the tiles of a pose grid: members are binned on the device per call, so
  // no host chunk table is built
  bool per_pose = false;
  // a host pose list (sv_pose_list_create): its poses on the device and the
  // pose grid that evaluates them
  struct PoseGrid* list_grid = nullptr;
  DevBuf<double> list_centers, list_axes;
  int list_n = 0;
  ~Poses();
};

struct Batch {
  int device = 0;
  hipStream_t stream = nullptr;
  int n_jobs = 0, precision = 0;
  bool all_sym = true, need_marginal = false, any_site = false;
  std::vector<Job> jobs;
  std::vector<long long> job_work, job_first;
  std::vector<Tile> tiles;
  std::vector<long long> seg_base;
  std::vector<SiteChunk> chunks;
  long long n_segs = 0, n_work = 0, n_centers_total = 0, list_total = -1;
  DevBuf<Job> d_jobs;
  DevBuf<long long> d_job_work, d_job_first, d_seg_base, d_offsets, d_scan;
  DevBuf<Tile> d_tiles;
  DevBuf<SiteChunk> d_chunks;
  DevBuf<int> d_counts, d_bad;
  DevBuf<float4> d_list32;
  DevBuf<double> d_list64;
  DevBuf<unsigned long long> d_pairs;
  hipEvent_t ev[12] = {};
  bool timed_last = false;
  bool lists_built = false;
  int persistent_blocks = 0;          // > 0: persistent field launch of this many blocks
  DevBuf<unsigned long long> d_counters;
  ~Batch() {
    for (auto& e : ev)
      if (e) (void)hipEventDestroy(e);
    if (stream) (void)hipStreamDestroy(stream);
  }
};


// Device-resident independent poses, evaluated site by site: every pose
// (c, u) contributes its two site positions c + o_s u, which are binned on the
// GPU into a fixed grid of bins over the cell.  Every bin is a tile whose
// candidate list (radius: bin circumradius + rc, built once) is split per type
// into the images that can come within a member site's near-contact radius
// and the rest.  Per-site energies and gradients are summed per pose.
struct PoseGrid {
  Plan* P = nullptr;
  Poses* S = nullptr;        // the tile set (one tile per bin) and the slot arrays
  Batch* B = nullptr;        // job table, candidate lists, stream
  int m[3] = {1, 1, 1};
  int n_bins = 0;
  int order = 0;             // sub-grid bit order: lattice axes by bin edge, longest first
  int capacity = 0;          // poses the slot arrays hold (two sites each)
  int parts = 1;             // blocks per run of the site kernels
  int binned_n = 0;          // poses binned by the last enqueue (0: none)
  double radius = 0.0;       // circumscribed radius of one bin
  Mat3 H{}, Hinv{};
  hipStream_t last_stream = nullptr;   // stream of the last enqueued evaluation
  bool last_stream_set = false;
  DevBuf<int> bin, rank;              // per site: cell of the sub-grid, rank in it
  DevBuf<int> counts, first;          // per cell (64 per bin): sites, first slot
  DevBuf<SiteRun> runs;               // work units of the site kernels
  // runs (low 32 bits: the run count once laid out) and slots taken (high 32)
  DevBuf<unsigned long long> totals;
  DevBuf<double> split2;               // [n_types] class-0 radius^2
  DevBuf<double> site_e, site_g;       // per-site results [2 cap], [2 cap][3]
  bool timing = false;                 // record events around the phases
  hipEvent_t ev[4] = {};
  ~PoseGrid() {
    if (last_stream_set) (void)hipStreamSynchronize(last_stream);
    for (auto& e : ev)
      if (e) (void)hipEventDestroy(e);
    delete B;
    delete S;
  }
};

// Error reporting (thread-local message, C-ABI return code).
int fail(int code, const std::string& msg);
int hip_fail(hipError_t e, const char* what);

// Framework and guest fields of a job from a plan (outputs and pose fields untouched).
void plan_fill_job(const Plan& P, Job& J);

// Grid of m[0] x m[1] x m[2] bins over the cell with its candidate lists built
// (sv_runtime.hip).  site_lists: the site form of independent poses (list
// radius bin + rc, two classes per type, FP32 coordinates in units of rc);
// otherwise center lists (radius bin + rc + max|o|, one class, Angstrom), as
// the String image kernels read them.
int pose_grid_build(Plan* P, const int m[3], bool site_lists, PoseGrid** out);

}  // namespace sv
