/*
 * FP64 CPU oracle for the guest-framework LJ Vext path.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py load this library, and only as the checker or
 * the timed CPU baseline; the product (libsv_vext.so) never calls it.
 *
 * Built with -ffp-contract=off (and no -march), so every product and sum is
 * rounded separately, as in the reference build (scripts/build_vext.sh:11-12,
 * g++ -O3 for baseline x86-64, which has no FMA).
 *
 * Pinning (tests/test_oracle.py, tests/test_oracle_pins.py); no function here
 * is left unpinned:
 *   oracle_reuse  == reference kernel.cpp compiled from source (oracle/_ref),
 *                    bit for bit, and == a brute-force all-images sum; a
 *                    planted dropped shift term fails;
 *   oracle_site   == the DIRECT formula checked against oracle_reuse at the
 *                    tolerances of tests/unit/test_public_api.py:60-61;
 *   oracle_pose_grad == central differences of its own energy, which equals
 *                    oracle_reuse;
 *   oracle_string_pair == the reference's triclinic_pair.h through its own
 *                    host shim (oracle/_ref), bit for bit, and the brute-force
 *                    invariants of test_native_string_triclinic.py:59-90;
 *   oracle_count_pairs == a brute-force count of d^2 < rc^2 over an oversized
 *                    block of images; a pair exactly at the cutoff is not
 *                    counted (a planted <= would count it).
 */
#include <math.h>
#include <stdint.h>

/* Pose energies of every (center, axis) pair over a CSR of candidate images.
 * Follows native/vext/kernel.cpp:13-44: loop center -> axis -> site ->
 * candidate, skip pairs with d^2 >= rc^2, floor d^2 at floor2, accumulate
 * c12 i6^2 - c6 i6 + shift in one FP64 sum per pose; return codes 1-4 as in
 * kernel.cpp:11-12, 14-15, 24-25, 39-40. */
int oracle_reuse(int n_centers, int n_axes, int n_images, double rc2, const double *centers,
                 const double *axes, const double *images, const double *offsets,
                 const double *c12, const double *c6, const double *shift,
                 const double *floor2, const int64_t *starts, const int32_t *indices,
                 double *out) {
  if (n_centers < 0 || n_axes < 1 || n_images < 1 || rc2 <= 0) return 1;
  for (int c = 0; c < n_centers; ++c) {
    const int64_t first = starts[c], last = starts[c + 1];
    if (first > last) return 2;
    const double *p = centers + 3 * (int64_t)c;
    for (int k = 0; k < n_axes; ++k) {
      const double *u = axes + 3 * (int64_t)k;
      double e = 0.0;
      for (int s = 0; s < 2; ++s) {
        const double sx = p[0] + offsets[s] * u[0];
        const double sy = p[1] + offsets[s] * u[1];
        const double sz = p[2] + offsets[s] * u[2];
        for (int64_t n = first; n < last; ++n) {
          const int32_t j = indices[n];
          if (j < 0 || j >= n_images) return 3;
          const double *a = images + 3 * (int64_t)j;
          const double dx = sx - a[0], dy = sy - a[1], dz = sz - a[2];
          const double d2 = dx * dx + dy * dy + dz * dz;
          if (d2 >= rc2) continue;
          const double q = d2 < floor2[j] ? floor2[j] : d2;
          const double inv = 1.0 / q;
          const double i6 = inv * inv * inv;
          e += c12[j] * i6 * i6 - c6[j] * i6 + shift[j];
        }
      }
      if (!isfinite(e)) return 4;
      out[(int64_t)c * n_axes + k] = e;
    }
  }
  return 0;
}

/* Single-site energy at query points over a CSR of images within rc.
 * Follows src/string_vext/physics.py:72-85: r = |d|, ratio =
 * (sigma / max(r, 0.1 sigma))^6 (pow, as numpy ** 6), 4 eps (ratio^2 - ratio)
 * + shifted for r < rc, per-image mixed sigma/eps/shift, summed in candidate
 * order (np.bincount). */
int oracle_site(int n_points, const double *points, const double *images, const double *sigma,
                const double *eps, const double *shifted, double cutoff, const int64_t *starts,
                const int32_t *indices, double *out) {
  for (int i = 0; i < n_points; ++i) {
    const double *p = points + 3 * (int64_t)i;
    double e = 0.0;
    for (int64_t n = starts[i]; n < starts[i + 1]; ++n) {
      const int32_t j = indices[n];
      const double *a = images + 3 * (int64_t)j;
      const double dx = p[0] - a[0], dy = p[1] - a[1], dz = p[2] - a[2];
      const double r = sqrt(dx * dx + dy * dy + dz * dz);
      if (r >= cutoff) continue;
      const double lo = 0.1 * sigma[j];
      const double ratio = pow(sigma[j] / (r > lo ? r : lo), 6.0);
      e += 4.0 * eps[j] * (ratio * ratio - ratio) + shifted[j];
    }
    out[i] = e;
  }
  return 0;
}

/* Energy and analytic gradient of independent poses (center_i, axis_i):
 * E = sum_s sum_j phi(|c + o_s u - a_j|^2) with the reference pair term,
 * dE/dc = sum phi'(d^2) 2 d,  dE/du = sum phi'(d^2) 2 o_s d, where u is the
 * unit director as given (no renormalisation).  phi' = 0 beyond rc and inside
 * the contact floor (the floored term is constant there). */
int oracle_pose_grad(int n_poses, const double *centers, const double *axes,
                     const double *images, const double *offsets, const double *c12,
                     const double *c6, const double *shift, const double *floor2, double rc2,
                     const int64_t *starts, const int32_t *indices, double *energy,
                     double *grad_center, double *grad_axis) {
  for (int i = 0; i < n_poses; ++i) {
    const double *p = centers + 3 * (int64_t)i, *u = axes + 3 * (int64_t)i;
    double e = 0.0, gc[3] = {0.0, 0.0, 0.0}, gu[3] = {0.0, 0.0, 0.0};
    for (int s = 0; s < 2; ++s) {
      const double o = offsets[s];
      const double sx = p[0] + o * u[0], sy = p[1] + o * u[1], sz = p[2] + o * u[2];
      for (int64_t n = starts[i]; n < starts[i + 1]; ++n) {
        const int32_t j = indices[n];
        const double *a = images + 3 * (int64_t)j;
        const double d[3] = {sx - a[0], sy - a[1], sz - a[2]};
        const double d2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
        if (!(d2 < rc2)) continue;
        const int floored = d2 < floor2[j];
        const double q = floored ? floor2[j] : d2;
        const double inv = 1.0 / q;
        const double i6 = inv * inv * inv;
        e += c12[j] * i6 * i6 - c6[j] * i6 + shift[j];
        if (floored) continue;
        /* d phi / d(d^2) = (-6 c12 i6^2 + 3 c6 i6) / d^2 */
        const double g = (-6.0 * c12[j] * i6 * i6 + 3.0 * c6[j] * i6) * inv;
        for (int k = 0; k < 3; ++k) {
          gc[k] += 2.0 * g * d[k];
          gu[k] += 2.0 * g * o * d[k];
        }
      }
    }
    energy[i] = e;
    for (int k = 0; k < 3; ++k) {
      grad_center[3 * (int64_t)i + k] = gc[k];
      grad_axis[3 * (int64_t)i + k] = gu[k];
    }
  }
  return 0;
}

/* The String solver's pair energy (native/string/source_code/triclinic_pair.h:
 * 14-68) restated: one guest site (Cartesian) and one framework atom
 * (fractional) in a cell given by Cartesian rows A, B, C of the
 * fractional-to-Cartesian matrix; every lattice image inside the cutoff,
 * distance floored at 0.01 sigma^2 before the cutoff test, shifted LJ. */
double oracle_string_pair(double x, double y, double z, double fa, double fb, double fc,
                          const double *A, const double *B, const double *C, double eps_guest,
                          double sigma_guest, double eps_frame, double sigma_frame,
                          double cutoff) {
  const double det = A[0] * (B[1] * C[2] - B[2] * C[1]) - A[1] * (B[0] * C[2] - B[2] * C[0]) +
                     A[2] * (B[0] * C[1] - B[1] * C[0]);
  if (!isfinite(det) || fabs(det) < 1e-12 || cutoff <= 0.0) return NAN;
  /* rows of the inverse: fractional coordinate k = inv[k] . (x, y, z) */
  const double inv[3][3] = {
      {(B[1] * C[2] - B[2] * C[1]) / det, (A[2] * C[1] - A[1] * C[2]) / det,
       (A[1] * B[2] - A[2] * B[1]) / det},
      {(B[2] * C[0] - B[0] * C[2]) / det, (A[0] * C[2] - A[2] * C[0]) / det,
       (A[2] * B[0] - A[0] * B[2]) / det},
      {(B[0] * C[1] - B[1] * C[0]) / det, (A[1] * C[0] - A[0] * C[1]) / det,
       (A[0] * B[1] - A[1] * B[0]) / det}};
  const double atom[3] = {fa, fb, fc};
  double frac[3], reach[3];
  for (int k = 0; k < 3; ++k) {
    frac[k] = inv[k][0] * x + inv[k][1] * y + inv[k][2] * z - atom[k];
    frac[k] -= floor(frac[k]);
    reach[k] = cutoff * sqrt(inv[k][0] * inv[k][0] + inv[k][1] * inv[k][1] + inv[k][2] * inv[k][2]);
  }
  int lo[3], hi[3];
  for (int k = 0; k < 3; ++k) {
    lo[k] = (int)ceil(frac[k] - reach[k] - 1e-12);
    hi[k] = (int)floor(frac[k] + reach[k] + 1e-12);
  }
  const double sigma = (sigma_guest + sigma_frame) * 0.5;
  const double epsilon = sqrt(eps_guest * eps_frame);
  const double s2 = sigma * sigma, rc2 = cutoff * cutoff;
  const double q = s2 / rc2, q6 = q * q * q;
  double total = 0.0;
  for (int i = lo[0]; i <= hi[0]; ++i)
    for (int j = lo[1]; j <= hi[1]; ++j)
      for (int k = lo[2]; k <= hi[2]; ++k) {
        const double u = frac[0] - i, v = frac[1] - j, w = frac[2] - k;
        const double dx = A[0] * u + A[1] * v + A[2] * w;
        const double dy = B[0] * u + B[1] * v + B[2] * w;
        const double dz = C[0] * u + C[1] * v + C[2] * w;
        double d2 = dx * dx + dy * dy + dz * dz;
        if (d2 < 0.01 * s2) d2 = 0.01 * s2;
        if (d2 < rc2) {
          const double r2 = s2 / d2, r6 = r2 * r2 * r2;
          total += 4.0 * epsilon * (r6 * r6 - r6 - q6 * q6 + q6);
        }
      }
  return total;
}

/* Number of (center, axis, site, candidate) pairs with d^2 < rc^2: the
 * pair-eval count of the throughput metric, with the loop of oracle_reuse. */
int64_t oracle_count_pairs(int n_centers, int n_axes, double rc2, const double *centers,
                           const double *axes, const double *images, const double *offsets,
                           const int64_t *starts, const int32_t *indices) {
  int64_t count = 0;
  for (int c = 0; c < n_centers; ++c) {
    const double *p = centers + 3 * (int64_t)c;
    for (int k = 0; k < n_axes; ++k) {
      const double *u = axes + 3 * (int64_t)k;
      for (int s = 0; s < 2; ++s) {
        const double sx = p[0] + offsets[s] * u[0];
        const double sy = p[1] + offsets[s] * u[1];
        const double sz = p[2] + offsets[s] * u[2];
        for (int64_t n = starts[c]; n < starts[c + 1]; ++n) {
          const double *a = images + 3 * (int64_t)indices[n];
          const double dx = sx - a[0], dy = sy - a[1], dz = sz - a[2];
          count += (dx * dx + dy * dy + dz * dz < rc2);
        }
      }
    }
  }
  return count;
}
