/*
 * cosims_oracle.h -- FP64 CPU restatement of the CoSIMS He trajectory path.
 *
 * TEST INFRASTRUCTURE ONLY.  This oracle is the checker for the HIP engine in
 * cosims_amd/: only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline
 * leg may load it.  The engine never links, calls or falls back to it.
 *
 * Every function restates the reference algorithm (ChristopherAMyers/CoSIMS,
 * src/ paths cited per function in cosims_oracle.c) in plain IEEE double,
 * sequential, in the reference's own expression order, built with
 * -ffp-contract=off so results match an x86-64 build of the reference bit for
 * bit where the reference is deterministic.
 *
 * Pinning (what checks each function; see tests/):
 *   or_mol_properties, or_inertia, or_veldist, or_boundary, or_ellipsoid,
 *   or_enlarge, or_pcm_clusters, or_cluster_moments, or_accel, or_potential,
 *   or_verlet                  -> golden vectors written by the compiled
 *                                 reference (oracle/_ref) + example logs +
 *                                 closed forms (finite differences, limits)
 *   or_sample_ic               -> statistical moments of the sampled density
 *                                 and the geometry identities of the entry
 *                                 point (on the ellipsoid, direction = -v0)
 *   or_reduce                  -> closed-form reduction on hand-made lists
 *   or_philox4x32              -> Random123 known-answer vectors
 */
#ifndef COSIMS_ORACLE_H
#define COSIMS_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Prepared molecule: what the hot path reads (reference molecule2::data after
 * setMolProperties + setInertia, plus forceAccel's cluster tables). */
typedef struct or_mol {
    int n;
    const double *x, *y, *z;      /* rotated, centred coordinates (A) */
    const double *q;              /* partial charges (e) */
    const double *sig6;           /* sigma^6 */
    const double *eps24m;         /* 24*eps'/mu (atom 0: 24*eps', quirk) */
    const double *eps4;           /* 4*eps' */
    int charged;                  /* charge_ext */
    double dipole_scalar;         /* 28276.01363 / mu */
    int force_type;               /* 1..6 (forceAccel::setClusters) */
    /* clusters (types 1,2,3,5) */
    int ncl;
    const int *cl_start;          /* ncl+1 offsets into cl_atoms */
    const int *cl_atoms;          /* member atom indices, cluster order */
    const double *cl_pos;         /* 3*ncl cluster positions */
    const double *cl_rad;         /* ncl cluster radii */
    const double *cl_q;           /* ncl total charges */
    const double *cl_dip;         /* 3*ncl dipoles */
    const double *cl_quad;        /* 9*ncl quadrupoles, row-major [i][j] */
    const int *cl_replace;        /* ncl: centre replaced by centre of charge */
    double multipole_radius;
    double dispersion_radius;
    int multipole_order;
} or_mol;

/* ---- setup ------------------------------------------------------------ */
/* setMolProperties numeric part.  In: n, xyz[3n] (file order), mass, sigma,
 * eps (kcal/mol), charge (may be NULL => 0), charged flag.  Out: xyz centred
 * in place, sig6, eps24m, eps4, *mu, *total_mass. */
void or_mol_properties(int n, double *xyz, const double *mass,
                       const double *sigma, const double *eps,
                       double *sig6, double *eps24m, double *eps4,
                       double *mu, double *total_mass);
/* Default cut-off radius from the largest sigma in the type table. */
double or_default_dispersion_radius(double sigma_max);
/* setInertia(true): rotates xyz in place, returns 0 on success, -1 when the
 * reference would take its undefined-behaviour fallback branch.  rot[9]
 * receives the rotation matrix, maxxyz[3] the signed extents. */
int or_inertia(int n, double *xyz, const double *mass, double rot[9],
               double maxxyz[3]);

typedef struct or_veldist {
    double a;          /* beta/2 */
    double mean, mode, sigma;
    double gmin, gmax;
    double norm;       /* after renormalize */
    double max_val;    /* velDistr(mode) */
    double cdf_frac;   /* cdf(gmax)-cdf(gmin) */
    double ccs_scale;  /* const1*const2 / bMax^2 (= const1) */
    double min_kin;    /* 0.5*mu*(mean2-2*sigma2)^2, boundary energy */
} or_veldist;
void or_veldist_init(double mu, double temp, or_veldist *vd);
double or_veldist_pdf(const or_veldist *vd, double g);

/* impactParam::setBoundry.  Writes up to cap points (3 doubles each) and
 * returns the count (or -1 if cap too small). */
int or_boundary(const or_mol *m, const double maxxyz[3], double min_kin,
                double *pts, int cap);
typedef struct or_ellipsoid {
    double a, b, c;
    double a_unc, b_unc, c_unc, rmsd;
    double cov[6];       /* aa bb cc ab ac bc */
    double avg, avg_unc, area, area_unc;
    double enlarge;
} or_ellipsoid;
void or_ellipsoid_fit(const double *pts, int npts, or_ellipsoid *e);
void or_enlarge(const double *pts, int npts, or_ellipsoid *e);

/* PCM_Cluster::formClusters.  Outputs in reference cluster order: ncl,
 * start[ncl+1], atoms[n], center[3*ncl], maxdist[ncl].  Capacity for start /
 * center / maxdist is n+1 clusters. Returns ncl or -1. */
int or_pcm_clusters(int n, const double *xyz, int max_cluster_size,
                    int *start, int *atoms, double *center, double *maxdist);
/* forceAccel::addMultipole + addClusterPos for every cluster. */
void or_cluster_moments(int n, const double *xyz, const double *q, int ncl,
                        const int *start, const int *atoms,
                        const double *center, double *cl_q, double *cl_dip,
                        double *cl_quad, double *cl_pos, int *cl_replace);
int or_force_type(int charged, int multipole, int disp_cutoff);

/* ---- hot path ---------------------------------------------------------- */
void or_accel(const or_mol *m, const double pos[3], double acc[3]);
double or_potential(const or_mol *m, const double pos[3]);

typedef struct or_traj_out {
    double chi;
    int ok;            /* verletVel return value */
    int steps;         /* p.stepCount as the reference leaves it */
    double ergy;       /* energy % deviation of the accepted attempt */
    int64_t total_steps; /* all integration steps over all attempts */
    int attempts;      /* force-loop restarts (1 + energy retries) */
} or_traj_out;
void or_verlet(const or_mol *m, double dt, const double p0[3],
               const double v0[3], double a, double b, double c,
               int projection, double mu, or_traj_out *out);
void or_verlet_batch(const or_mol *m, double dt, int64_t n, const double *p0,
                     const double *v0, double a, double b, double c,
                     int projection, double mu, double *chi, int32_t *ok,
                     int32_t *steps, double *ergy, int64_t *total_steps,
                     int nthreads);
void or_accel_batch(const or_mol *m, int64_t n, const double *pos,
                    double *acc, int nthreads);
void or_potential_batch(const or_mol *m, int64_t n, const double *pos,
                        double *u, int nthreads);

/* ---- initial conditions (counter-based) and reduction ------------------ */
void or_philox4x32(const uint32_t ctr[4], const uint32_t key[2],
                   uint32_t out[4]);
typedef struct or_ic_cfg {
    uint64_t seed; uint32_t job; uint32_t iteration;
    double gmin, gmax, max_val;   /* velocity window and f(mode) */
    double vd_a, vd_norm;         /* density parameters */
    double bmax;                  /* impact parameter range */
    double ea, eb, ec;            /* ellipsoid axes */
    int projection;
} or_ic_cfg;
/* Draw trajectory `traj`: writes p0, v0, g; returns 1 on hit, 0 on miss. */
int or_sample_ic(const or_ic_cfg *cfg, uint32_t traj, double p0[3],
                 double v0[3], double *g, int *tries);

typedef struct or_ccs {
    double omega, err;       /* per-iteration CCS and standard error */
    double total, total2;    /* sums of x and x^2 */
    int64_t denom;           /* currSuccTraj */
} or_ccs;
/* monteTraj::runThreads/updateCounts/runTraj for one iteration, one thread.
 * hit[i]=0 => miss (chi=0, success flag stale). */
void or_reduce_iteration(int64_t n, const double *chi, const int32_t *ok,
                         const int32_t *hit, double ccs_const,
                         int64_t update_cond, or_ccs *out);
void or_reduce_final(int iters, const double *omega, const double *err,
                     double *mean, double *sd, double *wmean, double *werr);

#ifdef __cplusplus
}
#endif
#endif
