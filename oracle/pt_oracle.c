/*
 * CPU oracle for the trajectory and lag-statistics path of particle_transport.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the parity checker: only tests/,
 * __graft_entry__.smoke() and bench.py's CPU-baseline leg may load it.  The
 * product (libpt_hip.so) never links or calls it.
 *
 * It restates the reference's Julia algorithms in plain C, FP64, in the
 * reference's operation order, built with -ffp-contract=off so no FMA is
 * formed behind the source's back.  Arrays follow the reference's Julia
 * conventions: column-major (x fastest), and each function cites the
 * reference file:line it follows.  Indices passed in and out are 0-based;
 * where the reference is 1-based the translation is stated.
 *
 * Parity pinning (which check pins which function) is listed in the header
 * comment of each function; the tests live in tests/test_oracle.py.
 * Parity unpinned: Julia's MersenneTwister streams (injection draws, random
 * particle selection; the tests hand both sides identical arrays instead) and
 * the bit pattern of Julia's pairwise mean in oracle_mean_abs_b (its known
 * answer pins the value; other fields compare at 1e-15).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define C_LIGHT 2.99792458e8
#define Q_E 1.602176634e-19
#define M_P 1.67262192369e-27
#define GEV_TO_J (1e9 * 1.602176634e-19)

/* Julia's floor(Int, x) then clamp(., lo, hi): phase_space_gpu_runner.jl:252-254 */
static inline int64_t clamp_floor(double f, int64_t lo, int64_t hi) {
    double fl = floor(f);
    int64_t i;
    if (fl < (double)lo) return lo;
    if (fl > (double)hi) return hi;
    i = (int64_t)fl;
    return i;
}

/*
 * trilinear_sample: phase_space_gpu_runner.jl:247-276 (identical copy in
 * compute_delta_mu2_curve.jl:62-91).  arr is (nx,ny,nz) column-major.
 * Pinned by: exactness on trilinear polynomials and at grid nodes.
 */
double oracle_trilinear(const double *arr, double x, double y, double z,
                        double xmin, double ymin, double zmin,
                        double dx, double dy, double dz,
                        int64_t nx, int64_t ny, int64_t nz) {
    double fx = (x - xmin) / dx + 1.0;
    double fy = (y - ymin) / dy + 1.0;
    double fz = (z - zmin) / dz + 1.0;
    int64_t i0 = clamp_floor(fx, 1, nx - 1);
    int64_t j0 = clamp_floor(fy, 1, ny - 1);
    int64_t k0 = clamp_floor(fz, 1, nz - 1);
    double tx = fx - (double)i0;
    double ty = fy - (double)j0;
    double tz = fz - (double)k0;
    /* 1-based (i,j,k) -> 0-based linear offset */
#define A(i, j, k) arr[((i) - 1) + nx * (((j) - 1) + ny * ((k) - 1))]
    double c000 = A(i0, j0, k0);
    double c100 = A(i0 + 1, j0, k0);
    double c010 = A(i0, j0 + 1, k0);
    double c110 = A(i0 + 1, j0 + 1, k0);
    double c001 = A(i0, j0, k0 + 1);
    double c101 = A(i0 + 1, j0, k0 + 1);
    double c011 = A(i0, j0 + 1, k0 + 1);
    double c111 = A(i0 + 1, j0 + 1, k0 + 1);
#undef A
    double c00 = c000 * (1 - tx) + c100 * tx;
    double c10 = c010 * (1 - tx) + c110 * tx;
    double c01 = c001 * (1 - tx) + c101 * tx;
    double c11 = c011 * (1 - tx) + c111 * tx;
    double c0 = c00 * (1 - ty) + c10 * ty;
    double c1 = c01 * (1 - ty) + c11 * ty;
    return c0 * (1 - tz) + c1 * tz;
}

/* cross3: phase_space_gpu_runner.jl:278-280 */
static inline void cross3(double ax, double ay, double az, double bx, double by, double bz,
                          double *cx, double *cy, double *cz) {
    *cx = ay * bz - az * by;
    *cy = az * bx - ax * bz;
    *cz = ax * by - ay * bx;
}

/*
 * boris_push: phase_space_gpu_runner.jl:282-323, proton constants :8-10.
 * out = (pxn, pyn, pzn, xn, yn, zn).
 * Pinned by: |p| conservation and rotation angle 2*atan(|t|) in uniform B,
 * E x B drift velocity in crossed uniform fields.
 */
void oracle_boris_push(double px, double py, double pz, double x, double y, double z,
                       double ex, double ey, double ez, double bx, double by, double bz,
                       double dt, double *out) {
    double half = dt / 2.0;
    double qm = Q_E, mp = M_P, cl = C_LIGHT;
    double pmx = px + qm * ex * half;
    double pmy = py + qm * ey * half;
    double pmz = pz + qm * ez * half;
    double mp2 = mp * mp, cl2 = cl * cl;
    double gamma_minus = sqrt(1.0 + (pmx * pmx + pmy * pmy + pmz * pmz) / (mp2 * cl2));
    double tx = qm * bx * dt / (2.0 * mp * gamma_minus);
    double ty = qm * by * dt / (2.0 * mp * gamma_minus);
    double tz = qm * bz * dt / (2.0 * mp * gamma_minus);
    double t2 = tx * tx + ty * ty + tz * tz;
    double sx = 2.0 * tx / (1.0 + t2);
    double sy = 2.0 * ty / (1.0 + t2);
    double sz = 2.0 * tz / (1.0 + t2);
    double c1x, c1y, c1z, c2x, c2y, c2z;
    cross3(pmx, pmy, pmz, tx, ty, tz, &c1x, &c1y, &c1z);
    double ppx = pmx + c1x, ppy = pmy + c1y, ppz = pmz + c1z;
    cross3(ppx, ppy, ppz, sx, sy, sz, &c2x, &c2y, &c2z);
    double ppx2 = pmx + c2x, ppy2 = pmy + c2y, ppz2 = pmz + c2z;
    double pxn = ppx2 + qm * ex * half;
    double pyn = ppy2 + qm * ey * half;
    double pzn = ppz2 + qm * ez * half;
    double gamma_new = sqrt(1.0 + (pxn * pxn + pyn * pyn + pzn * pzn) / (mp2 * cl2));
    double vxn = pxn / (gamma_new * mp);
    double vyn = pyn / (gamma_new * mp);
    double vzn = pzn / (gamma_new * mp);
    out[0] = pxn;
    out[1] = pyn;
    out[2] = pzn;
    out[3] = x + vxn * dt;
    out[4] = y + vyn * dt;
    out[5] = z + vzn * dt;
}

/* wrap_periodic_coordinate: phase_space_gpu_runner.jl:325-328 */
double oracle_wrap(double x, double xmin, double xmax) {
    double width = xmax - xmin;
    return x - floor((x - xmin) / width) * width;
}

/* Geometry of build_uniform_coords (phase_space_gpu_runner.jl:125-130):
 * x_i = range(0, L, length=n); the kernel uses xmin=x[1], xmax=x[end],
 * dx = x[2]-x[1] (:522-527).  The caller passes the Julia-computed values. */
typedef struct {
    int64_t nx, ny, nz;
    double xmin, ymin, zmin, xmax, ymax, zmax, dx, dy, dz;
} oracle_geom;

/*
 * One launch of advance_particles_kernel! (phase_space_gpu_runner.jl:330-391)
 * over all n particles.  step arrays (x*_step, p*_step) are written as the
 * reference writes them.  Returns the number of particles alive afterwards
 * (the reference's sum(dalive), :576).
 * Pinned by: the uniform-field closed forms through oracle_boris_push, and
 * by the wrap/kill rules checked directly in the tests.
 */
int64_t oracle_advance_step(double *const xs[3], double *const ps[3], double *const xstep[3],
                            double *const pstep[3], uint8_t *alive,
                            const double *const F[6], const oracle_geom *g, int64_t n,
                            double dt, int do_push, int periodic) {
    int64_t alive_count = 0;
    for (int64_t p = 0; p < n; ++p) {
        if (!alive[p]) {
            for (int c = 0; c < 3; ++c) {
                xstep[c][p] = NAN;
                pstep[c][p] = NAN;
            }
            continue;
        }
        double xx = xs[0][p], yy = xs[1][p], zz = xs[2][p];
        xstep[0][p] = xx;
        xstep[1][p] = yy;
        xstep[2][p] = zz;
        pstep[0][p] = ps[0][p];
        pstep[1][p] = ps[1][p];
        pstep[2][p] = ps[2][p];
        if (do_push) {
            double v[6];
            for (int c = 0; c < 6; ++c)
                v[c] = oracle_trilinear(F[c], xx, yy, zz, g->xmin, g->ymin, g->zmin, g->dx, g->dy,
                                        g->dz, g->nx, g->ny, g->nz);
            double ex, ey, ez;
            cross3(v[3], v[4], v[5], v[0], v[1], v[2], &ex, &ey, &ez);
            ex = -ex;
            ey = -ey;
            ez = -ez;
            double o[6];
            oracle_boris_push(ps[0][p], ps[1][p], ps[2][p], xx, yy, zz, ex, ey, ez, v[0], v[1], v[2],
                              dt, o);
            if (periodic) {
                o[3] = oracle_wrap(o[3], g->xmin, g->xmax);
                o[4] = oracle_wrap(o[4], g->ymin, g->ymax);
                o[5] = oracle_wrap(o[5], g->zmin, g->zmax);
            }
            ps[0][p] = o[0];
            ps[1][p] = o[1];
            ps[2][p] = o[2];
            xs[0][p] = o[3];
            xs[1][p] = o[4];
            xs[2][p] = o[5];
            if (!periodic && (o[3] < g->xmin || o[3] > g->xmax || o[4] < g->ymin ||
                              o[4] > g->ymax || o[5] < g->zmin || o[5] > g->zmax))
                alive[p] = 0;
        }
        if (alive[p]) ++alive_count;
    }
    return alive_count;
}

/*
 * pitch_mu_from_pB: compute_delta_mu2_curve.jl:93-110.
 * Pinned by: closed form on hand-built vectors (parallel, antiparallel,
 * perpendicular, zero norm, NaN).
 */
double oracle_mu_from_pB(double px, double py, double pz, double bx, double by, double bz) {
    double pn = sqrt(px * px + py * py + pz * pz);
    double bn = sqrt(bx * bx + by * by + bz * bz);
    if (!isfinite(pn) || !isfinite(bn) || pn == 0.0 || bn == 0.0) return NAN;
    double mu = (px * bx + py * by + pz * bz) / (pn * bn);
    if (mu < -1.0) mu = -1.0;
    if (mu > 1.0) mu = 1.0;
    return mu;
}

/* reconstruct_mu_step_kernel! (multimode_multienergy_cache_pipeline.jl:1349-1374)
 * applied to the *_step arrays of one save step. */
void oracle_mu_step(double *mu, double *const xstep[3], double *const pstep[3],
                    const double *const F[3], const oracle_geom *g, int64_t n) {
    for (int64_t p = 0; p < n; ++p) {
        double xx = xstep[0][p], yy = xstep[1][p], zz = xstep[2][p];
        double px = pstep[0][p], py = pstep[1][p], pz = pstep[2][p];
        double value = NAN;
        if (!(isnan(xx) || isnan(yy) || isnan(zz) || isnan(px) || isnan(py) || isnan(pz))) {
            if (g->xmin <= xx && xx <= g->xmax && g->ymin <= yy && yy <= g->ymax &&
                g->zmin <= zz && zz <= g->zmax) {
                double b[3];
                for (int c = 0; c < 3; ++c)
                    b[c] = oracle_trilinear(F[c], xx, yy, zz, g->xmin, g->ymin, g->zmin, g->dx,
                                            g->dy, g->dz, g->nx, g->ny, g->nz);
                value = oracle_mu_from_pB(px, py, pz, b[0], b[1], b[2]);
            }
        }
        mu[p] = value;
    }
}

/*
 * Step loop of run_gpu_ensemble / run_mu_cache_ensemble
 * (phase_space_gpu_runner.jl:562-591; multimode_multienergy_cache_pipeline.jl:1467-1498).
 * Runs steps si = first_step .. first_step+n_steps-1 (1-based, as the
 * reference's si), with do_push = si < total_steps and a save at every si
 * with (si-1) % stride == 0 (sampled_step_indices = 1:stride:nsteps, :393-396).
 *   out_kind 0: nothing saved; 1: phase space into pos/mom laid out as Julia
 *   (n,3,nsave) i.e. C [save][3][n]; 2: mu into mu_out [save][n].
 * alive_counts[k] receives sum(dalive) after step first_step+k.
 * Returns the number of saves written.
 */
int64_t oracle_run(double *const xs[3], double *const ps[3], uint8_t *alive,
                   const double *const F[6], const oracle_geom *g, int64_t n, double dt,
                   int64_t first_step, int64_t n_steps, int64_t total_steps, int64_t stride,
                   int periodic, int out_kind, double *pos, double *mom, double *mu_out,
                   int64_t *alive_counts) {
    double *xstep[3], *pstep[3];
    for (int c = 0; c < 3; ++c) {
        xstep[c] = (double *)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
        pstep[c] = (double *)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
        for (int64_t p = 0; p < n; ++p) xstep[c][p] = pstep[c][p] = NAN;
    }
    int64_t nsaved = 0;
    for (int64_t k = 0; k < n_steps; ++k) {
        int64_t si = first_step + k;
        int do_push = si < total_steps;
        int64_t cnt = oracle_advance_step(xs, ps, xstep, pstep, alive, F, g, n, dt, do_push, periodic);
        if (alive_counts) alive_counts[k] = cnt;
        if ((si - 1) % stride == 0) {
            if (out_kind == 1) {
                for (int c = 0; c < 3; ++c)
                    for (int64_t p = 0; p < n; ++p) {
                        pos[(nsaved * 3 + c) * n + p] = xstep[c][p];
                        mom[(nsaved * 3 + c) * n + p] = pstep[c][p];
                    }
            } else if (out_kind == 2) {
                oracle_mu_step(mu_out + nsaved * n, xstep, pstep, F, g, n);
            }
            ++nsaved;
        }
    }
    for (int c = 0; c < 3; ++c) {
        free(xstep[c]);
        free(pstep[c]);
    }
    return nsaved;
}

/*
 * reconstruct_mu_chunk_cpu: compute_delta_mu2_curve.jl:112-149.
 * positions/momenta are Julia (n,3,nsave) (C [s][3][n]); mu out is Julia
 * (nsave, n) i.e. time fastest: mu[s + nsave*p].
 */
void oracle_reconstruct_mu_chunk(const double *positions, const double *momenta,
                                 const double *const F[3], const oracle_geom *g, int64_t n,
                                 int64_t nsave, double *mu) {
    for (int64_t p = 0; p < n; ++p) {
        for (int64_t s = 0; s < nsave; ++s) {
            double xx = positions[(s * 3 + 0) * n + p], yy = positions[(s * 3 + 1) * n + p],
                   zz = positions[(s * 3 + 2) * n + p];
            double px = momenta[(s * 3 + 0) * n + p], py = momenta[(s * 3 + 1) * n + p],
                   pz = momenta[(s * 3 + 2) * n + p];
            mu[s + nsave * p] = NAN;
            if (isnan(xx) || isnan(yy) || isnan(zz) || isnan(px) || isnan(py) || isnan(pz)) continue;
            if (xx < g->xmin || xx > g->xmax || yy < g->ymin || yy > g->ymax || zz < g->zmin ||
                zz > g->zmax)
                continue;
            double b[3];
            for (int c = 0; c < 3; ++c)
                b[c] = oracle_trilinear(F[c], xx, yy, zz, g->xmin, g->ymin, g->zmin, g->dx, g->dy,
                                        g->dz, g->nx, g->ny, g->nz);
            mu[s + nsave * p] = oracle_mu_from_pB(px, py, pz, b[0], b[1], b[2]);
        }
    }
}

/*
 * mu_bin_index_full: compute_delta_mu2_dmumu_full.jl:308-318 (searchsortedlast
 * with a right-closed last bin).  Returns 1..nbins, or 0 for dropped.
 */
int64_t oracle_mu_bin_index(double v, const double *edges, int64_t nedges) {
    if (!isfinite(v)) return 0;
    if (v < edges[0] || v > edges[nedges - 1]) return 0;
    int64_t lo = 0, hi = nedges; /* largest i (1-based) with edges[i] <= v */
    while (lo < hi) {
        int64_t mid = (lo + hi) / 2;
        if (edges[mid] <= v)
            lo = mid + 1;
        else
            hi = mid;
    }
    int64_t idx = lo; /* 1-based searchsortedlast */
    if (idx == nedges) return nedges - 1;
    return idx < 1 ? 0 : idx;
}

/* combine_welford_stats!: compute_delta_mu2_dmumu_full.jl:439-455 */
void oracle_combine_welford(int64_t *counts, double *means, double *m2s, int64_t li,
                            int64_t local_count, double local_mean, double local_m2) {
    if (local_count <= 0) return;
    int64_t global_count = counts[li];
    if (global_count == 0) {
        counts[li] = local_count;
        means[li] = local_mean;
        m2s[li] = local_m2;
        return;
    }
    int64_t combined = global_count + local_count;
    double delta = local_mean - means[li];
    means[li] += delta * (double)local_count / (double)combined;
    m2s[li] += local_m2 + delta * delta * (double)global_count * (double)local_count / (double)combined;
    counts[li] = combined;
}

/*
 * process_mu_chunk_full_pairs_cpu! (compute_delta_mu2_dmumu_full.jl:457-560)
 * and process_mu_chunk_injection_pairs_cpu! (:562-654), with one thread
 * (Threads.maxthreadid() == 1), so per-thread locals are combined once per
 * lag exactly as the reference does.
 *   mu is Julia (nsave, chunk): element (s, p) at mu[s + nsave*p] (0-based).
 *   lag_steps are the reference's 1-based offsets (>= 1).
 *   Accumulators use the reference's shapes: per lag (nlags), and
 *   (n_bins, n_lags) column-major for the bin arrays.
 * Pinned by: brute force over hand-built series in pure Python, Welford vs
 * two-pass variance, and injection == sliding when nsave == 2.
 */
void oracle_dmumu(const float *mu, int64_t nsave, int64_t chunk, const int64_t *lag_steps,
                  int64_t nlags, const double *edges, int64_t nedges, int injection, int bin_abs,
                  int64_t *particle_counts, double *particle_means, double *particle_m2s,
                  double *pair_sum_squares, int64_t *pair_counts, int64_t *dmumu_counts,
                  double *dmumu_sum_delta, double *dmumu_sum_delta2) {
    int64_t nbins = nedges - 1;
    int64_t *lc = (int64_t *)calloc((size_t)nbins, sizeof(int64_t));
    double *ls1 = (double *)calloc((size_t)nbins, sizeof(double));
    double *ls2 = (double *)calloc((size_t)nbins, sizeof(double));
    for (int64_t li = 0; li < nlags; ++li) {
        int64_t lag = lag_steps[li];
        if (lag >= nsave) continue;
        memset(lc, 0, sizeof(int64_t) * (size_t)nbins);
        memset(ls1, 0, sizeof(double) * (size_t)nbins);
        memset(ls2, 0, sizeof(double) * (size_t)nbins);
        int64_t lpc = 0, lpairc = 0;
        double lmean = 0.0, lm2 = 0.0, lpss = 0.0;
        for (int64_t p = 0; p < chunk; ++p) {
            const float *m = mu + nsave * p;
            if (injection) {
                double a = (double)m[0], b = (double)m[lag];
                if (!(isfinite(a) && isfinite(b))) continue;
                double d = b - a, d2 = d * d;
                int64_t bi = oracle_mu_bin_index(bin_abs ? fabs(a) : a, edges, nedges);
                if (1 <= bi && bi <= nbins) {
                    lc[bi - 1] += 1;
                    ls1[bi - 1] += d;
                    ls2[bi - 1] += d2;
                }
                lpc += 1;
                double delta = d2 - lmean;
                lmean += delta / (double)lpc;
                double delta2 = d2 - lmean;
                lm2 += delta * delta2;
                lpss += d2;
                lpairc += 1;
            } else {
                double psum = 0.0;
                int64_t valid = 0;
                int64_t last_start = nsave - lag;
                for (int64_t s = 0; s < last_start; ++s) {
                    double a = (double)m[s], b = (double)m[s + lag];
                    if (!(isfinite(a) && isfinite(b))) continue;
                    double d = b - a, d2 = d * d;
                    psum += d2;
                    valid += 1;
                    int64_t bi = oracle_mu_bin_index(bin_abs ? fabs(a) : a, edges, nedges);
                    if (1 <= bi && bi <= nbins) {
                        lc[bi - 1] += 1;
                        ls1[bi - 1] += d;
                        ls2[bi - 1] += d2;
                    }
                }
                if (valid > 0) {
                    double pv = psum / (double)valid;
                    lpc += 1;
                    double delta = pv - lmean;
                    lmean += delta / (double)lpc;
                    double delta2 = pv - lmean;
                    lm2 += delta * delta2;
                    lpss += psum;
                    lpairc += valid;
                }
            }
        }
        for (int64_t b = 0; b < nbins; ++b) {
            dmumu_counts[b + nbins * li] += lc[b];
            dmumu_sum_delta[b + nbins * li] += ls1[b];
            dmumu_sum_delta2[b + nbins * li] += ls2[b];
        }
        oracle_combine_welford(particle_counts, particle_means, particle_m2s, li, lpc, lmean, lm2);
        pair_sum_squares[li] += lpss;
        pair_counts[li] += lpairc;
    }
    free(lc);
    free(ls1);
    free(ls2);
}

/* momentum_magnitude: compute_dpp_full.jl:232-234 */
static inline double pmag(float px, float py, float pz) {
    return sqrt((double)px * (double)px + (double)py * (double)py + (double)pz * (double)pz);
}

/*
 * process_momenta_chunk_dpp_cpu!: compute_dpp_full.jl:247-288 with one
 * thread.  momenta is Julia (chunk,3,nsave) i.e. C [s][3][chunk].
 * Pinned by: brute force over hand-built momentum series in pure Python.
 */
void oracle_dpp(const float *momenta, int64_t chunk, int64_t nsave, const int64_t *lag_steps,
                int64_t nlags, double p0, int64_t *counts, double *sum_dp, double *sum_dp2,
                double *sum_dpn, double *sum_dpn2) {
    for (int64_t li = 0; li < nlags; ++li) {
        int64_t lag = lag_steps[li];
        if (lag >= nsave) continue;
        int64_t last_start = nsave - lag;
        int64_t c = 0;
        double s1 = 0, s2 = 0, s3 = 0, s4 = 0;
        for (int64_t p = 0; p < chunk; ++p) {
            for (int64_t s = 0; s < last_start; ++s) {
                const float *a = momenta + (s * 3) * chunk + p;
                const float *b = momenta + ((s + lag) * 3) * chunk + p;
                double ps = pmag(a[0], a[chunk], a[2 * chunk]);
                double pe = pmag(b[0], b[chunk], b[2 * chunk]);
                if (!(isfinite(ps) && isfinite(pe))) continue;
                double d = pe - ps;
                double dn = d / p0;
                c += 1;
                s1 += d;
                s2 += d * d;
                s3 += dn;
                s4 += dn * dn;
            }
        }
        counts[li] += c;
        sum_dp[li] += s1;
        sum_dp2[li] += s2;
        sum_dpn[li] += s3;
        sum_dpn2[li] += s4;
    }
}

/* kinetic_energy_gev_from_pmag: compute_dpp_full.jl:236-240 */
double oracle_kinetic_energy_gev(double pm) {
    if (!isfinite(pm)) return NAN;
    double r = pm / (M_P * C_LIGHT);
    double gamma = sqrt(1.0 + r * r);
    return (gamma - 1.0) * M_P * C_LIGHT * C_LIGHT / GEV_TO_J;
}

/* fill_energy_snapshots!: compute_dpp_full.jl:290-299.  snapshot_idx are
 * 0-based save indices; out is Julia (k, total) column-major with this chunk
 * written at column offset col0. */
void oracle_energy_snapshots(const float *momenta, int64_t chunk, const int64_t *snapshot_idx,
                             int64_t k, double *out, int64_t col0) {
    for (int64_t j = 0; j < k; ++j) {
        int64_t s = snapshot_idx[j];
        for (int64_t p = 0; p < chunk; ++p) {
            const float *a = momenta + (s * 3) * chunk + p;
            out[j + k * (col0 + p)] = oracle_kinetic_energy_gev(pmag(a[0], a[chunk], a[2 * chunk]));
        }
    }
}

/*
 * reference_B0_T: time_units.jl:18-20, mean(sqrt.(Bx.^2 .+ By.^2 .+ Bz.^2)).
 * Julia's mean is a pairwise SIMD sum whose order is machine dependent; the
 * oracle sums in long double so it is accurate to well below 1e-15 relative.
 */
double oracle_mean_abs_b(const double *bx, const double *by, const double *bz, int64_t n) {
    long double acc = 0.0L;
    for (int64_t i = 0; i < n; ++i) acc += (long double)sqrt(bx[i] * bx[i] + by[i] * by[i] + bz[i] * bz[i]);
    return (double)(acc / (long double)n);
}

/*
 * Counter-based injection draws.  The reference draws from Julia's
 * MersenneTwister(seed), per particle, in the order of init_particles
 * (phase_space_gpu_runner.jl:196-245).  That stream is sequential, so the
 * device injection draws from Philox4x64-10 (Salmon, Moraes, Dror and Shaw,
 * "Parallel random numbers: as easy as 1, 2, 3", SC 2011) instead: key
 * (seed, 0), counter (particle index, block, 0, 0), four 64-bit words per
 * block.  Draw j of particle i is word j % 4 of block j / 4, mapped to
 * [0, 1) as (w >> 11) * 2^-53.
 * Pinned by: numpy.random.Philox, an independent implementation of the same
 * generator, on arbitrary keys and counters, and its random() doubles.
 */
static inline uint64_t mulhi64(uint64_t a, uint64_t b) {
    return (uint64_t)(((unsigned __int128)a * (unsigned __int128)b) >> 64);
}

void oracle_philox4x64_10(const uint64_t ctr_in[4], const uint64_t key_in[2], uint64_t out[4]) {
    const uint64_t M0 = 0xD2E7470EE14C6C93ULL, M1 = 0xCA5A826395121157ULL;
    const uint64_t W0 = 0x9E3779B97F4A7C15ULL, W1 = 0xBB67AE8584CAA73BULL;
    uint64_t c0 = ctr_in[0], c1 = ctr_in[1], c2 = ctr_in[2], c3 = ctr_in[3];
    uint64_t k0 = key_in[0], k1 = key_in[1];
    for (int r = 0; r < 10; ++r) {
        if (r > 0) {
            k0 += W0;
            k1 += W1;
        }
        uint64_t hi0 = mulhi64(M0, c0), lo0 = M0 * c0;
        uint64_t hi1 = mulhi64(M1, c2), lo1 = M1 * c2;
        uint64_t n0 = hi1 ^ c1 ^ k0, n2 = hi0 ^ c3 ^ k1;
        c0 = n0;
        c1 = lo1;
        c2 = n2;
        c3 = lo0;
    }
    out[0] = c0;
    out[1] = c1;
    out[2] = c2;
    out[3] = c3;
}

double oracle_inject_draw(uint64_t seed, int64_t i, int j) {
    uint64_t ctr[4] = {(uint64_t)i, (uint64_t)(j / 4), 0, 0}, key[2] = {seed, 0}, w[4];
    oracle_philox4x64_10(ctr, key, w);
    return (double)(w[j % 4] >> 11) * (1.0 / 9007199254740992.0);
}

/*
 * fixed_mu_unit_vector: phase_space_gpu_runner.jl:170-194 in FP64 with the
 * gyrophase phi already drawn.  Returns -1 where |B| is zero or non-finite
 * (the reference raises an error there), else 0.
 */
static int fixed_mu_unit(double bx, double by, double bz, double mu0, double phi, double u[3]) {
    double bn = sqrt(bx * bx + by * by + bz * bz);
    if (!isfinite(bn) || bn == 0.0) return -1;
    double bhx = bx / bn, bhy = by / bn, bhz = bz / bn;
    double e1x, e1y, e1z;
    if (fabs(bhx) < 0.9) {
        e1x = 0.0;
        e1y = bhz;
        e1z = -bhy;
    } else {
        e1x = -bhz;
        e1y = 0.0;
        e1z = bhx;
    }
    double e1n = sqrt(e1x * e1x + e1y * e1y + e1z * e1z);
    e1x = e1x / e1n;
    e1y = e1y / e1n;
    e1z = e1z / e1n;
    double e2x = bhy * e1z - bhz * e1y;
    double e2y = bhz * e1x - bhx * e1z;
    double e2z = bhx * e1y - bhy * e1x;
    double perp = sqrt(fmax(0.0, 1.0 - mu0 * mu0));
    double cphi = cos(phi), sphi = sin(phi);
    u[0] = mu0 * bhx + perp * (cphi * e1x + sphi * e2x);
    u[1] = mu0 * bhy + perp * (cphi * e1y + sphi * e2y);
    u[2] = mu0 * bhz + perp * (cphi * e1z + sphi * e2z);
    return 0;
}

/*
 * init_particles: phase_space_gpu_runner.jl:196-245 in FP64, with the draws
 * of oracle_inject_draw, for global particles first .. first+count-1.  Per
 * particle, in the reference's order: three position draws
 * rand * (xmax - xmin) + xmin (skipped for a fixed position), then either
 * z = 2u - 1 and phi = 2 pi u (rand_unit_vector, :61-66) or, for fixed mu,
 * trilinear B at the position and phi = 2 pi u.  p = pscale * unit vector.
 * out is [x1 x2 x3 p1 p2 p3][count].  B (three column-major (nx,ny,nz)
 * arrays) is read only for fixed mu.  Returns -1 if a fixed-mu particle sits
 * where |B| is zero or non-finite, else 0.
 * Pinned by: positions equal numpy Philox doubles mapped onto the box;
 * fixed-mu momenta have p.b / (|p||b|) == mu0 and |p| == pscale; isotropic
 * directions give <u_z> = 0, <u_z^2> = 1/3 and a uniform azimuth.
 */
int oracle_inject(uint64_t seed, int64_t first, int64_t count, int fixed_position, const double pos[3],
                  int fixed_mu, double mu0, double pscale, const double *const B[3],
                  const oracle_geom *g, double *out) {
    const double two_pi = 6.283185307179586; /* Julia's 2 * pi as Float64 */
    const double lo[3] = {g->xmin, g->ymin, g->zmin}, hi[3] = {g->xmax, g->ymax, g->zmax};
    int status = 0;
    for (int64_t q = 0; q < count; ++q) {
        const int64_t i = first + q;
        int j = 0;
        double x[3], u[3];
        for (int c = 0; c < 3; ++c) {
            if (fixed_position) {
                x[c] = pos[c];
            } else {
                x[c] = oracle_inject_draw(seed, i, j++) * (hi[c] - lo[c]) + lo[c];
            }
        }
        if (fixed_mu) {
            double b[3];
            for (int c = 0; c < 3; ++c)
                b[c] = oracle_trilinear(B[c], x[0], x[1], x[2], g->xmin, g->ymin, g->zmin, g->dx, g->dy,
                                        g->dz, g->nx, g->ny, g->nz);
            double phi = two_pi * oracle_inject_draw(seed, i, j++);
            if (fixed_mu_unit(b[0], b[1], b[2], mu0, phi, u) != 0) {
                status = -1;
                u[0] = u[1] = u[2] = 0.0;
            }
        } else {
            double z = 2 * oracle_inject_draw(seed, i, j++) - 1;
            double phi = two_pi * oracle_inject_draw(seed, i, j++);
            double r = sqrt(fmax(0.0, 1.0 - z * z));
            u[0] = r * cos(phi);
            u[1] = r * sin(phi);
            u[2] = z;
        }
        for (int c = 0; c < 3; ++c) {
            out[c * count + q] = x[c];
            out[(3 + c) * count + q] = pscale * u[c];
        }
    }
    return status;
}
