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
  // ---- independent poses (one director per center) ------------------------
  const double* pose_axes64; // [n_centers][3] director of each sorted slot
  const float* pose_axes32;
  const double* grad_out;    // [n_centers] upstream gradient by caller index (null = 1)
  double* pose_e;            // [n_centers] FP64 energy by caller index
  double* grad_c;            // [n_centers][3] dE/d(center), Cartesian
  double* grad_u;            // [n_centers][3] dE/d(director), director used as given
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
// n_chunks bounds the grid; d_n_chunks (device, may be null) holds the count.
hipError_t launch_pose(int precision, bool grad, const Job* d_jobs, const SiteChunk* d_chunks,
                       int n_chunks, const int* d_n_chunks, const long long* d_offsets,
                       const float4* d_list32, const double* d_list64, hipStream_t s);
hipError_t launch_pose_bin(int n, const double* d_centers, const double* d_axes, const Mat3& Hinv,
                           const int m[3], int* d_bin, int* d_rank, int* d_counts, int* d_bad,
                           hipStream_t s);
hipError_t launch_pose_tile_scan(const int* d_counts, int n_bins, int job, int* d_first,
                                 SiteChunk* d_chunks, int* d_n_chunks, hipStream_t s);
hipError_t launch_pose_scatter(int n, const double* d_centers, const double* d_axes, const Mat3& H,
                               const Mat3& Hinv, const int m[3], const int* d_bin,
                               const int* d_rank, const int* d_first, double* d_center_rel,
                               int* d_center_id, int* d_tile_of, double* d_axes64, float* d_axes32,
                               hipStream_t s);
hipError_t launch_zero(int* p, int n, hipStream_t s);
// Sets the per-call pose pointers of job 0 from kernel arguments (stream-ordered).
// Patches the pose job's arrays and clears zero[0..n_zero) in one launch.
hipError_t launch_job_patch(Job* d_job, const double* center_rel, const int* center_id,
                            const int* tile_of, const double* axes64, const float* axes32, int n,
                            const double* grad_out, double* pose_e, double* grad_c,
                            double* grad_u, int* zero, int n_zero, hipStream_t s);
hipError_t launch_legacy(int n_centers, int n_axes, double cutoff2, const double* centers,
                         const double* axes, const double* images, const double* offsets,
                         const double* coeff12, const double* coeff6, const double* shift,
                         const double* floor2, const long long* starts, const int* indices,
                         double* output, int* bad_center, hipStream_t s);

// Precision selectors for launch_field.
enum { PREC_FP64 = 0, PREC_MIXED = 1 };

}  // namespace sv
