/*
 * libsv_vext.so - MI355X (gfx950) guest-framework Lennard-Jones Vext.
 *
 * C ABI only: plain pointers, sizes and opaque handles.  All pointers are
 * host pointers unless a function says otherwise.  Every function returns 0
 * on success and a nonzero code on failure; sv_last_error() returns a
 * thread-local description of the last failure.
 *
 * Return codes shared by every entry point:
 *   0  success
 *   1  invalid size/argument      (reference native/vext/kernel.cpp:11-12)
 *   2  non-monotone CSR starts    (kernel.cpp:14-15)
 *   3  neighbor index out of range (kernel.cpp:24-25)
 *   4  non-finite energy          (kernel.cpp:39-40)
 *   5  HIP runtime error (no device, launch failure, ...)
 *   6  device memory exhausted
 */
#ifndef SV_VEXT_H
#define SV_VEXT_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------------ */
/* Drop-in entry point.                                                      */
/* Replaces evaluate_reused_neighbors of the reference CPU kernel            */
/* (native/vext/kernel.cpp:5-10), loaded by NativeReusePotential through     */
/* ctypes (src/string_vext/native.py:18-22) and selectable with              */
/* `string-vext --kernel`.  Same arguments, same return codes:               */
/*   centers [n_centers][3] Cartesian, axes [n_axes][3] unit directors,      */
/*   images [n_images][3] Cartesian, offsets[2] site offsets along the axis, */
/*   coeff12/coeff6/shift/floor2 [n_images], CSR starts [n_centers+1] and    */
/*   indices [starts[n_centers]] of candidate images per center,             */
/*   output [n_centers][n_axes] (FP64).                                       */
/* Evaluated on the GPU in FP64 with the reference's operation order, so a   */
/* center's energies equal the reference CPU results on the same CSR.        */
/* Synchronous and thread-safe (calls are serialized).  On error, rows of    */
/* centers before the failing one hold their energies.                       */
/* ------------------------------------------------------------------------ */
int evaluate_reused_neighbors(int n_centers, int n_axes, int n_images, double cutoff2,
                              const double *centers, const double *axes, const double *images,
                              const double *offsets, const double *coeff12,
                              const double *coeff6, const double *shift, const double *floor2,
                              const int64_t *starts, const int32_t *indices, double *output);

/* Allocation-free form of evaluate_reused_neighbors, with the same         */
/* arguments, results and return codes.  No counterpart in the reference:   */
/* it is for callers that cannot allocate device memory (serving loops and   */
/* harnesses that forbid hipMalloc).  Its device staging -- the inputs, the  */
/* CSR, the output and a flag, each rounded up to 256 bytes -- lives in a    */
/* static 64 MiB array per device, held until the process ends; a call whose */
/* staging would exceed it returns 1 with a message in sv_last_error.  The   */
/* call makes no hipMalloc/hipFree and creates no stream (it runs on the     */
/* null stream).  Calls in one process are serialized; the array keeps the   */
/* last call's data, and every array a call reads is written by that call.   */
int evaluate_reused_neighbors_noalloc(int n_centers, int n_axes, int n_images, double cutoff2,
                                      const double *centers, const double *axes,
                                      const double *images, const double *offsets,
                                      const double *coeff12, const double *coeff6,
                                      const double *shift, const double *floor2,
                                      const int64_t *starts, const int32_t *indices,
                                      double *output);

/* ------------------------------------------------------------------------ */
/* Library information                                                       */
/* ------------------------------------------------------------------------ */
const char *sv_last_error(void);
/* ABI version of this header (increases when entry points are added). */
int sv_abi_version(void);
/* Number of visible HIP devices (0 when none). */
int sv_device_count(void);

/* ------------------------------------------------------------------------ */
/* Framework plan: one framework, one two-site guest, one cutoff.            */
/* Replaces the setup of ReusePotential / TriclinicSitePotential             */
/* (src/string_vext/neighbor.py:9-30, physics.py:42-57): cell matrix rows    */
/* are the lattice vectors (cell.py:24-26), framework coordinates are        */
/* fractional, Lorentz-Berthelot mixing with the guest, shifted cutoff at rc,*/
/* contact floor (0.1 sigma)^2.  Device tables are built once per plan.      */
/*   cell[9]            row-major 3x3, det > 0                                */
/*   frac [n_atoms][3], sigma/epsilon [n_atoms] (> 0)                         */
/*   guest_sigma/epsilon  the two identical guest sites                      */
/*   offsets[2]         signed site offsets along the director               */
/*                      (physics.py:88-90)                                   */
/*   near_energy_K      bound on a single FP32 pair term in the mixed        */
/*                      precision kernel; closer atoms are evaluated in FP64. */
/*                      <= 0 keeps FP64 only for atoms within reach of the  */
/*                      contact floor (0.1 sigma + max|offset|).             */
/* ------------------------------------------------------------------------ */
int sv_plan_create(const double *cell, int n_atoms, const double *frac, const double *sigma,
                   const double *epsilon, double guest_sigma, double guest_epsilon,
                   const double *offsets, double cutoff, double near_energy_K, void **plan);
int sv_plan_destroy(void *plan);
/* Number of distinct framework LJ types of the plan. */
int sv_plan_info(void *plan, int *n_types, double *keep_radius);

/* ------------------------------------------------------------------------ */
/* Pose sets: every (center, axis) pair of n_centers fractional centers and  */
/* n_axes directors (the product evaluated by NativeReusePotential.          */
/* evaluate, native.py:26-68).  Centers are wrapped into [0,1) like          */
/* np.mod(centers, 1) (native.py:35) and directors are normalized to u/|u|   */
/* like native.py:34 (finite nonzero directors required, else code 1).       */
/* sv_poses_grid uses the regular pose grid centers (i+0.5)/size on an       */
/* ij-indexed meshgrid (orientation.py:7-16).                                */
/* ------------------------------------------------------------------------ */
int sv_poses_create(void *plan, int n_centers, const double *centers_frac, int n_axes,
                    const double *axes, void **poses);
int sv_poses_grid(void *plan, int size, int n_axes, const double *axes, void **poses);
int sv_poses_destroy(void *poses);
/* As above with the tile target: pose centers are binned into spatial tiles  */
/* of about `tile_centers` centers that share one candidate list (the plain  */
/* entry points use 64; string_vext_amd.autotune picks it per workload).     */
int sv_poses_create_ex(void *plan, int n_centers, const double *centers_frac, int n_axes,
                       const double *axes, int tile_centers, void **poses);
int sv_poses_grid_ex(void *plan, int size, int n_axes, const double *axes, int tile_centers,
                     void **poses);

/* ------------------------------------------------------------------------ */
/* Whole-field evaluation, host outputs (synchronous).                       */
/*   precision   0 = FP64 (reference arithmetic), 1 = mixed FP32 pair math   */
/*               with FP64 accumulation and FP64 near-contact terms,         */
/*               2 = FP64 with 1/d^2 from the hardware reciprocal and one    */
/*               Newton step instead of IEEE division (fields and batches;   */
/*               the pose and string entry points take 0 or 1)               */
/*   energy      [n_centers][n_axes] FP64   (reference kernel layout) or NULL */
/*   orientation [n_axes][n_centers] FP32   (vext.py:31 orientation_K) or NULL*/
/*   marginal    [n_centers] FP32  -T(logsumexp(-E/T) - ln K) (physics.py:126)*/
/*   grid        [n_centers] FP32  asinh(marginal/1000)        (vext.py:40)   */
/*   site        [n_centers] FP32  single-site field            (vext.py:27-29)*/
/*   temperature  needed when marginal or grid is requested                  */
/* Returns 4 when any requested output is non-finite.                        */
/* ------------------------------------------------------------------------ */
int sv_field_run(void *plan, void *poses, int precision, double temperature, double *energy,
                 float *orientation, float *marginal, float *grid, float *site);

/* Exact number of in-cutoff (site, image) pairs of the pose set: the        */
/* "pair-eval" count of the throughput metric (FP64 cutoff test).            */
int sv_field_census(void *plan, void *poses, unsigned long long *pairs);

/* ------------------------------------------------------------------------ */
/* Batches: many (plan, poses) fields evaluated by one set of launches.      */
/* Outputs are DEVICE pointers (arrays of n_jobs pointers, entries may be    */
/* NULL).  sv_batch_run is asynchronous on the batch's stream unless         */
/* `sync` is nonzero; sv_batch_sync waits and reports non-finite outputs.    */
/* With time_kernels nonzero, HIP events bracket each phase on the stream    */
/* and sv_batch_timings returns the last run's milliseconds per phase:       */
/*   [0] list count, [1] scan, [2] list fill, [3] field, [4] marginal,        */
/*   [5] site field, [6] whole run.                                           */
/* ------------------------------------------------------------------------ */
int sv_batch_create(int n_jobs, void *const *plans, void *const *poses, int precision,
                    const double *temperatures, double *const *energy,
                    float *const *orientation, float *const *marginal, float *const *grid,
                    float *const *site, void **batch);
int sv_batch_run(void *batch, int time_kernels, int sync);
int sv_batch_sync(void *batch);
int sv_batch_timings(void *batch, double *ms7);
int sv_batch_census(void *batch, unsigned long long *pairs_per_job);
int sv_batch_destroy(void *batch);
/* Mixed-precision batches only: launch the field as one persistent grid of   */
/* blocks_per_cu x (number of CUs) blocks whose waves drain an XCD-partitioned */
/* work queue (0 restores one block per 4 work items).                        */
int sv_batch_set_persistent(void *batch, int blocks_per_cu);

/* ------------------------------------------------------------------------ */
/* Independent poses: pose i is (center_i, director_i); used for energies at */
/* arbitrary poses and their gradients (torch.autograd, string images).  No  */
/* counterpart in the reference, whose String forces are forward differences */
/* of the energy (native/string/source_code/global.h:1281-1324).             */
/*   centers_cart [n][3] Cartesian (wrapped internally), axes [n][3] used as  */
/*   given (callers normalize): the sites are c + o_s u in both precisions.  */
/* Each pose is evaluated as its two site positions, E = phi(p_0) + phi(p_1),*/
/* dE/dc = g_0 + g_1, dE/du = o_0 g_0 + o_1 g_1 with g_s = grad phi(p_s).    */
/* A pose list is destroyed with sv_poses_destroy.                           */
/* sv_pose_list_eval writes DEVICE arrays (each may be NULL):                */
/*   energy [n] FP64; grad_center [n][3] = g_i dE_i/dc_i;                     */
/*   grad_axis [n][3] = g_i dE_i/du_i, with g = grad_out [n] (device) or 1.    */
/* The backward pass recomputes every pair term (nothing is stored).         */
/* Synchronous; returns 4 on non-finite results.                             */
/* ------------------------------------------------------------------------ */
int sv_pose_list_create(void *plan, int n_poses, const double *centers_cart, const double *axes,
                        void **poses);
int sv_pose_list_eval(void *poses, int precision, const double *grad_out, double *energy,
                      double *grad_center, double *grad_axis);

/* ------------------------------------------------------------------------ */
/* Device-resident independent poses (no counterpart in the reference; for   */
/* callers whose poses already live on the GPU, e.g. torch.autograd).        */
/* sv_pose_grid_create fixes a grid of spatial tiles over the cell, sized    */
/* for the 2 n_expected site positions of about n_expected poses at           */
/* tile_centers sites per tile (0: the bin counts that minimize the expected */
/* candidate work for uniformly spread poses), and builds every tile's       */
/* candidate list once.  sv_pose_grid_eval takes DEVICE arrays               */
/* centers_cart [n][3] and axes [n][3] (FP64), bins the poses into the tiles */
/* on the GPU and evaluates them exactly like sv_pose_list_eval (same        */
/* outputs, all DEVICE pointers, each may be NULL).  Any n >= 1 is accepted  */
/* (buffers grow).  Synchronous; returns 1 for a nonfinite center or a       */
/* nonfinite/zero director, 4 for nonfinite results.                         */
/* sv_pose_grid_info reports the tile count and the current pose capacity.   */
/* ------------------------------------------------------------------------ */
int sv_pose_grid_create(void *plan, int n_expected, int tile_centers, void **grid);
int sv_pose_grid_eval(void *grid, int n, const double *centers_cart, const double *axes,
                      int precision, const double *grad_out, double *energy,
                      double *grad_center, double *grad_axis);
int sv_pose_grid_info(void *grid, int *n_tiles, int *capacity);
int sv_pose_grid_destroy(void *grid);
/* Asynchronous form: enqueues the same evaluation on `stream` (a hipStream_t; */
/* NULL = the default stream) and returns without waiting, so the outputs    */
/* are ordered with the caller's other work on that stream.  Calls on one    */
/* grid are serialized: a call on a different stream first waits for the    */
/* previous one.  Invalid input and nonfinite results are not reported by   */
/* the call; they accumulate in the grid's flags, which sv_pose_grid_check   */
/* returns (1 or 4, as above; after waiting for the grid's last stream) and  */
/* clears.  Results of invalid poses are NaN.                                 */
int sv_pose_grid_eval_async(void *grid, void *stream, int n, const double *centers_cart,
                            const double *axes, int precision, const double *grad_out,
                            double *energy, double *grad_center, double *grad_axis);
int sv_pose_grid_check(void *grid);
/* sv_pose_grid_reeval_async evaluates the poses of the grid's last            */
/* sv_pose_grid_eval(_async) call again -- e.g. the gradient pass after an    */
/* energy pass -- from the binned copies the grid keeps, skipping the         */
/* binning (the caller's pose arrays are not read).  Outputs, precision and   */
/* grad_out are this call's.  Returns 1 when no poses have been binned.       */
int sv_pose_grid_reeval_async(void *grid, void *stream, int precision, const double *grad_out,
                              double *energy, double *grad_center, double *grad_axis);
/* Exact number of in-cutoff (site, image) pairs (FP64 test) of the poses of  */
/* the grid's last evaluation: the pair-eval count of the throughput metric. */
/* Synchronous; returns 1 when no poses have been binned.                    */
int sv_pose_grid_census(void *grid, unsigned long long *pairs);
/* With timing on, HIP events on the evaluation's stream bracket its phases; */
/* sv_pose_grid_timings waits for the last timed evaluation and returns its  */
/* milliseconds: [0] binning (0 for a re-evaluation), [1] site kernel,        */
/* [2] pose sums.                                                             */
int sv_pose_grid_set_timing(void *grid, int on);
/* Blocks that share each run of up to 64 sites (1..8; default 1): each      */
/* takes every parts-th batch of the run's candidate list and writes its own */
/* partial sums, which the pose sums add in part order.  More parts spread   */
/* the work of a call over more blocks (a shorter last wave of blocks) at a  */
/* fixed cost per block.                                                      */
int sv_pose_grid_set_parts(void *grid, int parts);
int sv_pose_grid_timings(void *grid, double *ms3);

/* ------------------------------------------------------------------------ */
/* Batched String method: n_strings independent strings of n_images images. */
/* Image state x = (a, b, c, alpha, beta, gamma): fractional center and      */
/* Euler angles (rad) of the body-x guest, director (cos g cos b,            */
/* sin g cos b, -sin b) (native/string/source_code/device_fxn.h:34-63).      */
/* Each iteration follows the reference iteration (RC_main.cu:1914-2393):    */
/* normalized steepest descent with move_frac / move_angle_rad and the       */
/* analytic gradient in place of forward differences, endpoint fix along     */
/* `direction` (1 = a, 2 = b, 3 = c), arc-length reparametrization, angle    */
/* revert and smoothing, convergence test every 200 iterations (a converged  */
/* string stops).  plans[s] is the framework plan of string s.               */
/*   init [n_strings][n_images][6]; n_images in 3..1024.                     */
/* sv_strings_run runs `iterations` more iterations; with use_graph one       */
/* iteration is captured once in a HIP graph and replayed.                   */
/* sv_strings_state returns the current strings (host arrays, may be NULL):  */
/*   path [S][N][6], energy [S][N] (evaluated now), iterations done [S],     */
/*   converged [S].                                                           */
/* No counterpart in the reference's C interface (its String solver is a     */
/* CUDA executable, src/string_vext/string.py:33-39).                         */
/* ------------------------------------------------------------------------ */
int sv_strings_create(int n_strings, void *const *plans, int n_images, const double *init,
                      const int *direction, double move_frac, double move_angle_rad,
                      int precision, void **strings);
int sv_strings_run(void *strings, int iterations, int use_graph);
int sv_strings_state(void *strings, double *path, double *energy, int *iterations,
                     int *converged);
int sv_strings_destroy(void *strings);

/* Device memory helpers for callers without their own allocator. */
int sv_device_malloc(void **ptr, uint64_t bytes);
int sv_device_free(void *ptr);
int sv_memcpy_d2h(void *dst, const void *src, uint64_t bytes);
int sv_memcpy_h2d(void *dst, const void *src, uint64_t bytes);

#ifdef __cplusplus
}
#endif
#endif /* SV_VEXT_H */
