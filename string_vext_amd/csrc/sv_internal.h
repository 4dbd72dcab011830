// Internal declarations shared by the host runtime and the gfx950 kernels of
// libsv_vext.so.  Nothing here is part of the public C ABI (include/sv_vext.h).
//
// Coordinates on the device are always *tile-relative*: framework images and
// pose centers are stored as Cartesian offsets from the centre of the spatial
// tile (bin of pose centers) they belong to, computed in FP64 from fractional
// differences.  That keeps FP32 copies of the same offsets accurate to
// ~ulp(15 A) instead of ~ulp(60 A).
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

#define SV_WAVE 64

namespace sv {

// 1/x in FP64 from v_rcp_f64 and one Newton step.  The step squares the
// seed's relative error (1 - x y1 = (1 - x y0)^2), so even a seed good to only
// 2^-20 gives ~1e-12.  For terms that need FP64-level relative accuracy but
// not IEEE rounding: the mixed kernels' near-contact terms and the String
// images.
__device__ __forceinline__ double inv_fast(double x) {
  const double y = __builtin_amdgcn_rcp(x);
  return fma(y, fma(-x, y, 1.0), y);
}


// The per-type constants one mixed-field wave reads for each LJ type, in one
// 64-byte record so that a type costs a single scalar block load.
struct alignas(64) FieldType {
  double c12, c6, shift, floor2;  // as the Job's per-type arrays
  double c6s;                     // c6 / rc^6 (shell form)
  float s6f, floor2f, near2f, lo2f, s6sf;
  float pad;
};
static_assert(sizeof(FieldType) == 64, "one 64-byte record per type");

// One field job = (framework, guest, pose set, axes, outputs).  A single
// Vext field is a batch of one job; a ragged batch is many jobs launched
// together.  Lives in device global memory and is read with uniform
// (scalar) loads.
struct Job {
  // ---- framework geometry (FP64) ------------------------------------------
  double H[9];          // rows are lattice vectors a, b, c (Angstrom)
  double hn[3];         // |column k of H^-1|: fractional half-width per Angstrom
  const double* frac;   // [n_atoms][3] wrapped fractional coords, sorted by type
  const int* type_start;// [n_types+1] atom ranges of each LJ type
  int n_atoms;
  int n_types;
  // ---- guest / cutoff ------------------------------------------------------
  double o0, o1;        // signed site offsets along the unit director
  double rc2;           // cutoff^2 (FP64 test, reference semantics)
  double keep_r;        // per-center candidate radius: rc + max|o| + margin
  float rc2f;           // cutoff^2 in FP32 (clamp bound of the mixed path)
  float keep2f;         // keep_r^2 in FP32
  float m2o0f, m2o1f;   // -2*o0, -2*o1
  float o0sqf;          // o0^2
  float o1sqf;          // o1^2
  float in2f;           // (rc - max|o| - margin)^2: closer atoms are inside the
                        // cutoff for every orientation
  float inv_rc2f;       // 1/rc^2: shell atoms of the mixed path work in units of rc^2
  float m2o0sf, m2o1sf; // -2*o0/rc^2, -2*o1/rc^2
  // the pose form works in units of 1/s, s = inv_rc2f exactly:
  float o0sxf, o1sxf;   // o^2 s
  float m2o0xf, m2o1xf; // -2 o s
  int sym;              // o1 == -o0 exactly
  int near_inside;      // both sites of every near-contact atom are inside the cutoff
  // ---- per-type coefficients (index = framework LJ type) ------------------
  const double* c12;    // 4 eps sigma^12          (mixed Lorentz-Berthelot)
  const double* c6;     // 4 eps sigma^6
  const double* shift;  // -4 eps ((sigma/rc)^12 - (sigma/rc)^6)
  const double* floor2; // (0.1 sigma)^2
  const double* sigma;  // mixed sigma   (site-field formula)
  const double* eps;    // mixed epsilon (site-field formula)
  const float* s6f;     // c12/c6 = sigma^6 in FP32
  const float* floor2f; // floor2 in FP32
  const float* near2f;  // (near radius)^2: center-atom distance below which the
                        // mixed kernel evaluates the pair in FP64
  const float* lo2f;    // (sqrt(floor2) + max|o| + margin)^2: beyond it no
                        // orientation reaches the contact floor
  const double* c6s;    // c6 / rc^6            (shell form of the mixed path)
  const float* s6sf;    // sigma^6 / rc^6 in FP32
  const FieldType* ftype; // the mixed field's per-type constants, one record per type
  const double* c6x;    // c6 s^3               (pose form, s = inv_rc2f)
  const float* s6xf;    // sigma^6 s^3 in FP32
  // ---- pose set -----------------------------------------------------------
  const double* center_rel;  // [n_centers][3] Cartesian offset from tile centre (sorted)
  const int* center_id;      // sorted slot -> caller's center index
  const int* tile_of;        // sorted slot -> job-local tile index
  const double* axes64;      // [n_axes][3] unit directors
  const float* axes32;
  int n_centers;
  int n_axes;
  int n_groups;              // ceil(n_axes / 64)
  int n_tiles;
  int tile_base;             // first global tile of this job
  int pad1;
  long long seg_base;        // index of (tile 0, type 0) in the global segment table
  long long work_begin;      // first work item (wave) of this job
  // ---- outputs (any may be null) -------------------------------------------
  double* out64;             // [n_centers][n_axes] FP64   (reference kernel layout)
  float* out32;              // [n_axes][n_centers] FP32   (orientation_K layout)
  float* marg32;             // [n_centers] marginal energy (needs n_axes <= 64)
  float* grid32;             // [n_centers] asinh(marginal/1000)
  float* site32;             // [n_centers] single-site field
  double temperature;
  double log_n_axes;         // ln(n_axes), for the marginal
  unsigned long long* pair_count;  // census output (pairs with d^2 < rc^2)
  int* bad;                  // set to 1 when a nonfinite energy is produced
  int pad2;
  // ---- candidate list layout ------------------------------------------------
  int n_cls;                 // classes per (tile, type) segment: 1, or 2 for site grids,
                             // whose class 0 holds the images within sqrt(split2[type])
                             // of the tile centre and class 1 the rest
  int pad3;
  double list_scale;         // list32 xyz = tile-relative Cartesian * list_scale
  const double* split2;      // [n_types] class-0 radius^2 (n_cls == 2)
  // ---- site decomposition of independent poses -----------------------------
  const float* nsite2f;      // [n_types] (site near-contact radius / rc)^2
};

// Global tile table entry (one per tile of every job).
struct Tile {
  double frac[3];     // fractional tile centre
  double radius;      // max distance of any member center from the centre
  double list_r;      // list radius = radius + keep_r + margin
  int job;
  int first_member;   // first sorted center slot of this tile (job-local)
  int n_members;
  int pad;
};

// A 3x3 matrix passed by value to kernels (row-major).
struct Mat3 {
  double m[9];
};

// Work unit of the site form of independent poses: up to 64 sites of one bin
// in slots [first, first + count).  whole: the run is made of whole cells of
// the bin's 4 x 4 x 4 sub-grid, so its set of sites is fixed by the cell
// counts; 0 when it is part of a single cell holding more than 64 sites.
struct SiteRun {
  int bin;
  int first;
  int count;
  int whole;
};

// Site-field work chunk: up to 64 member centers of one tile.
struct SiteChunk {
  int job;
  int tile;           // job-local tile
  int first;          // first sorted center slot
  int count;          // <= 64
};

// ---- kernel launchers (sv_kernels.hip) ------------------------------------
hipError_t launch_tile_count(const Job* d_jobs, const Tile* d_tiles, int n_tiles_total,
                             const long long* d_tile_seg_base, int* d_counts, hipStream_t s);
// Exclusive scan; d_sums holds scan_scratch_size(n) entries of scratch.
long long scan_scratch_size(long long n);
hipError_t launch_scan(const int* d_counts, long long n, long long* d_offsets, long long* d_sums,
                       hipStream_t s);
hipError_t launch_tile_fill(const Job* d_jobs, const Tile* d_tiles, int n_tiles_total,
                            const long long* d_tile_seg_base, const long long* d_offsets,
                            float4* d_list32, double* d_list64, hipStream_t s);
hipError_t launch_field(int precision, const Job* d_jobs, int n_jobs, const long long* d_job_work,
                        long long n_work, const long long* d_offsets, const float4* d_list32,
                        const double* d_list64, hipStream_t s, bool all_sym);
hipError_t launch_field_persistent(const Job* d_jobs, int n_jobs, const long long* d_job_work,
                                  long long n_work, unsigned long long* d_counters, int n_blocks,
                                  const long long* d_offsets, const float4* d_list32,
                                  const double* d_list64, hipStream_t s, bool all_sym);
hipError_t launch_zero_u64(unsigned long long* p, int n, hipStream_t s);
hipError_t launch_census(const Job* d_jobs, int n_jobs, const long long* d_job_work,
                         long long n_work, const long long* d_offsets, const double* d_list64,
                         hipStream_t s);
hipError_t launch_marginal(const Job* d_jobs, int n_jobs, const long long* d_job_first_center,
                           long long n_centers_total, hipStream_t s);
hipError_t launch_zero(int* p, int n, hipStream_t s);
// ---- independent poses evaluated site by site (sv_sites.hip) -----------------
// Site binning: every site goes to cell 64 b + sub, where b is its bin of
// the m[0] x m[1] x m[2] grid and sub the Morton index of its cell of the
// bin's 4 x 4 x 4 sub-grid, with the lattice axes order & 3, (order >> 2) & 3,
// (order >> 4) & 3 (longest bin edge first) as the most significant.
// totals: the packed running totals (runs in the low 32 bits, slots in the
// high 32) that the bin kernel clears for launch_site_layout.
hipError_t launch_site_bin(long long n_sites, const double* centers, const double* axes,
                           double o0, double o1, const Mat3& Hinv, const int m[3], int order,
                           int* cell, int* rank, int* cell_counts, int* bad,
                           unsigned long long* totals, hipStream_t s);
// First slot of every cell and the run table, after launch_site_bin on the
// same stream: the cells of each bin are taken in order into runs of at most
// 64 sites (a cell of more than 64 sites gets ceil(count / 64) runs of its
// own); each bin's slots and runs are contiguous, the bins in no fixed order.
// The table's length ends in the low 32 bits of *d_totals (the int at its
// address); it holds at most site_runs_bound(...).
long long site_runs_bound(int n_bins, long long n_sites);
hipError_t launch_site_layout(const int* d_cell_counts, int n_bins, unsigned long long* d_totals,
                              int* d_first, SiteRun* d_runs, hipStream_t s);
// Slots of the binned sites (offset from the bin centre, site id) and the
// cell counts cleared for the next binning.
hipError_t launch_site_scatter(long long n_sites, const double* centers, const double* axes,
                               double o0, double o1, const Mat3& H, const Mat3& Hinv,
                               const int m[3], const int* cell, const int* rank, const int* first,
                               int* cell_counts, double* site_rel, int* site_id, hipStream_t s);
// Per-site energies site_e [2 n] and/or gradients site_g [2 n][3] of the
// binned sites (site_rel, site_id by slot); max_runs bounds the grid and
// *d_n_runs (device) holds the run count.
struct SiteArrays {
  const double* site_rel;
  const int* site_id;
  double* site_e;          // [parts][part_stride]: part p of site k at p * part_stride + k
  double* site_g;          // [parts][part_stride][3]
  long long part_stride;   // >= 2 n
  int parts;               // blocks per run (1..kMaxSiteParts), each with every parts-th batch
};
constexpr int kMaxSiteParts = 8;
hipError_t launch_site(int precision, bool energy, bool grad, const Job* d_jobs,
                       const SiteRun* d_runs, int max_runs, const int* d_n_runs,
                       const long long* d_offsets, const float4* d_list32, const double* d_list64,
                       const SiteArrays& a, hipStream_t s);
// Pose sums (each site's parts added in part order): energy[i] = E_2i +
// E_2i+1, grad_center = g (G_2i + G_2i+1), grad_axis = g (o0 G_2i + o1 G_2i+1),
// g = grad_out[i] or 1 (outputs may be null); nonfinite results set flag 1.
hipError_t launch_site_combine(int n_poses, const SiteArrays& a, double o0, double o1,
                               const double* grad_out, double* energy, double* grad_center,
                               double* grad_axis, int* bad, hipStream_t s);
hipError_t launch_site_census(const Job* d_jobs, const SiteRun* d_runs, int max_runs,
                              const int* d_n_runs, const long long* d_offsets,
                              const double* d_list64, const double* site_rel,
                              unsigned long long* pairs, hipStream_t s);
hipError_t launch_legacy(int n_centers, int n_axes, double cutoff2, const double* centers,
                         const double* axes, const double* images, const double* offsets,
                         const double* coeff12, const double* coeff6, const double* shift,
                         const double* floor2, const long long* starts, const int* indices,
                         double* output, int* bad_center, hipStream_t s);

// Precision selectors for launch_field.  PREC_FP64_RCP (fields only) is the
// FP64 form with 1/d^2 from v_rcp_f64 and one Newton step instead of IEEE
// division.
enum { PREC_FP64 = 0, PREC_MIXED = 1, PREC_FP64_RCP = 2 };

}  // namespace sv
