/*
 * cosims_engine.h -- C ABI of the MI355X (gfx950) He trajectory engine.
 *
 * Drop-in boundary for the CoSIMS trajectory path (ChristopherAMyers/CoSIMS).
 * The reference has no FFI; its boundary is the in-process C++ surface
 * listed beside each entry point below.  Conventions:
 *   - plain pointers and sizes; caller-owned inputs are copied, outputs are
 *     written to caller arrays; the engine owns its device buffers;
 *   - every call returns 0 on success and a negative COSIMS_E_* code on
 *     failure; cosims_strerror() gives text, cosims_last_error() details;
 *   - a cosims_mol handle is immutable after creation; calls on one handle
 *     may run concurrently (each call uses its own HIP stream);
 *   - units are the reference's: A, ps, u, e; energies in u A^2 ps^-2.
 */
#ifndef COSIMS_ENGINE_H
#define COSIMS_ENGINE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define COSIMS_OK 0
#define COSIMS_E_INVALID (-1)     /* bad argument */
#define COSIMS_E_HIP (-2)         /* HIP runtime / device error */
#define COSIMS_E_NOMEM (-3)       /* host or device allocation failed */
#define COSIMS_E_SETUP (-4)       /* molecule setup failed (see last_error) */
#define COSIMS_E_IO (-5)          /* file could not be read */
#define COSIMS_E_UNSUPPORTED (-6) /* input the reference handles only by undefined behaviour */
#define COSIMS_E_NODEVICE (-7)    /* no usable gfx950 device */
#define COSIMS_E_PENDING (-8)     /* request not finished within the wait */

#define COSIMS_RNG_PHILOX 0    /* counter-based Philox4x32-10 stream (default) */
#define COSIMS_RNG_REFERENCE 1 /* the reference's own single-thread stream */

#define COSIMS_PREC_FP64 0  /* all-FP64 force sum, reference expression order */
#define COSIMS_PREC_MIXED 1 /* FP32 force sum, FP64 Verlet / energy check / CCS */

/* Run parameters: the keys of the reference input file
 * (fileParams::setDefaultParams / importParams, src/fileParams.cpp:203-510;
 * globals src/myStructs.h:28-74).  `traj` is a raw count here (the input
 * file gives it in thousands). */
typedef struct cosims_params {
    double temp;              /* temp_ext (K), default 298 */
    double dt;                /* dt_ext (ps), default 0.01 */
    int64_t traj;             /* trajNum_ext, default 50000 */
    int32_t iter;             /* iterations_ext, default 10 */
    int32_t charge;           /* charge_ext: polarisation terms on */
    uint64_t seed;            /* seed_ext */
    int32_t multipole;        /* multipole_ext, default 1 */
    int32_t dispersion_cutoff;/* dispersionCutoff_ext, default 1 */
    double dispersion_radius; /* 0 => sigma_max * 300^(1/3) (src/molecule2.cpp:881-887) */
    double multipole_radius;  /* 0 => 20 A (src/molecule2.cpp:888-895) */
    int32_t multipole_order;  /* multipole_order_ext, default 1 */
    int32_t max_cluster_size; /* maxClusterSize_ext (int), default 4 */
    int32_t projection;       /* projection_ext, default 1 */
    int32_t print_rate;       /* printRate_ext; 0 => traj/5 as with an input file */
    double total_charge;      /* totalCharge_ext (informational) */
    int32_t threads;          /* host threads for setup; 0 => all */
    char name[128];           /* name_ext */
    int32_t setup_gpu;        /* 1: the boundary search's potential marches run on the GPU
                                 (same points as the host loop); 0: host threads */
} cosims_params;

/* fileParams::setDefaultParams  src/fileParams.cpp:203-259 */
void cosims_params_default(cosims_params *p);
/* fileParams::importParams  src/fileParams.cpp:261-510: applies an input
 * file on top of *p (unknown keys ignored, last value wins, `traj` x1000,
 * `charge` switches charge_ext).  print_rate defaults to traj/5. */
int cosims_params_read(const char *input_path, cosims_params *p);

/* A molecule file read through the reference's readers and type table:
 * molecule2::assignDefaultFF/importFile/importMFJ/importPDB
 * (src/molecule2.cpp:43-192, 442-768), types resolved as in
 * setMolProperties (:782-839). */
typedef struct cosims_atom_type {   /* one row of molecule2::fieldData (src/molecule2.h) */
    char name[32];        /* lower case when it came from atoms.dat */
    char symbol[8];
    double mass, sigma, epsilon;  /* amu, A, kcal/mol */
    int32_t int_mass;
    int32_t count;        /* atoms of this type in the molecule (fieldData::atomCount) */
} cosims_atom_type;

typedef struct cosims_molfile {
    int32_t n;
    double *xyz;          /* 3n, file order */
    double *mass;         /* n */
    double *sigma;        /* n, A */
    double *epsilon;      /* n, kcal/mol */
    double *charge;       /* n */
    int32_t charged;      /* charge_ext as the reader sets it */
    double total_charge;  /* totalCharge_ext as the reader sets it */
    double sigma_max;     /* largest sigma of the whole type table */
    char name[128];       /* MFJ title line (empty for PDB) */
    int32_t ntypes;       /* the type table after atoms.dat, in table order */
    cosims_atom_type *types;
} cosims_molfile;
int cosims_molfile_read(const char *mol_path, const char *atoms_dat_path, cosims_molfile **out);
void cosims_molfile_free(cosims_molfile *f);

/* Raw atoms for setup (what setMolProperties receives). */
typedef struct cosims_atoms {
    int32_t n;
    const double *xyz;     /* 3n, file order, A */
    const double *mass;    /* n */
    const double *sigma;   /* n, A */
    const double *epsilon; /* n, kcal/mol */
    const double *charge;  /* n, or NULL for a neutral molecule */
    double sigma_max;      /* largest sigma of the type table */
} cosims_atoms;

/* A prepared molecule as the reference leaves it after setup, for injecting
 * setup results computed elsewhere (parity harnesses). */
typedef struct cosims_prepared {
    int32_t n;
    const double *xyz;     /* 3n rotated, centred coordinates */
    const double *charge;  /* n */
    const double *sig6;    /* n, sigma^6 */
    const double *eps24m;  /* n, 24 eps'/mu (atom 0 unscaled, as the reference) */
    const double *eps4;    /* n, 4 eps' */
    double mu;             /* reduced mass */
    double a, b, c;        /* ellipsoid axes after enlargement */
    int32_t ncl;           /* PCM clusters (0 when neither cut-off is on) */
    const int32_t *cl_start;  /* ncl+1 */
    const int32_t *cl_atoms;  /* n */
    const double *cl_center;  /* 3 ncl geometric centres */
    const double *cl_maxdist; /* ncl */
    double dispersion_radius;
    double multipole_radius;
} cosims_prepared;

typedef struct cosims_mol cosims_mol;

/* Setup = molecule2::setMolProperties + setInertia(true), impactParam
 * setBoundry/findEllipsoidBound/enlargeEllipBoundry, VelocityDistribution,
 * forceAccel::setClusters (PCM_Cluster + multipoles); replaces
 * `new monteTraj(mol, shared)` (src/main.cpp:119-138, src/monteTraj.cpp:26-142). */
int cosims_mol_create(const cosims_params *p, const cosims_atoms *atoms, cosims_mol **out);
/* Same, from injected setup results. */
int cosims_mol_create_prepared(const cosims_params *p, const cosims_prepared *prep, cosims_mol **out);
void cosims_mol_destroy(cosims_mol *m);
/* Creation is host-only; device buffers are made on the first GPU call on
 * the handle (thread-safe, once).  This call does it eagerly. */
int cosims_mol_upload(const cosims_mol *m);

typedef struct cosims_mol_info {
    int32_t n, ncl, force_type, charged, projection, multipole_order;
    double mu, total_mass, dipole_scalar;
    double a, b, c, bmax, enlarge;
    double gmin, gmax, gmean, gmode, gsigma, vd_a, vd_norm, max_val, cdf_frac;
    double ccs_const;          /* const1*const2 (src/monteTraj.cpp:91-94) */
    double min_kin;            /* boundary energy before the 1% factor */
    double dispersion_radius, multipole_radius;
    double rot[9], max_xyz[3];
    double fit_abc[3], fit_rmsd; /* least-squares axes before enlargement */
    int32_t n_boundary;
    double setup_ms;
    /* wall time of setup phases within setup_ms: the boundary search's
     * potential marches, PCM clustering (params.threads host threads), the
     * cluster multipoles */
    double setup_boundary_ms, setup_cluster_ms, setup_moments_ms;
} cosims_mol_info;
int cosims_mol_info_get(const cosims_mol *m, cosims_mol_info *out);
/* Prepared atoms (n each, xyz 3n) and clusters; any pointer may be NULL. */
int cosims_mol_atoms(const cosims_mol *m, double *xyz, double *charge, double *sig6, double *eps24m,
                     double *eps4);
int cosims_mol_clusters(const cosims_mol *m, int32_t *start, int32_t *atoms, double *center,
                        double *maxdist, double *pos, double *q, double *dip, double *quad,
                        int32_t *replace);

/* forceAccel::getReducedVec (src/forceAccel.cpp:887-908) for n positions. */
int cosims_accel(const cosims_mol *m, int32_t prec, int64_t n, const double *pos, double *acc);
/* forceAccel::getPotential (src/forceAccel.cpp:943-982), FP64, n positions. */
int cosims_potential(const cosims_mol *m, int64_t n, const double *pos, double *u);

/* Initial conditions of trajectories first..first+n-1 of (seed, job,
 * iteration): rejection sampling of monteTraj::setUpTraj
 * (src/monteTraj.cpp:174-278) driven by Philox4x32-10 instead of
 * mt19937_64 + rand().  hit=0 marks a line that misses the ellipsoid. */
int cosims_sample(const cosims_mol *m, uint64_t seed, uint32_t job, uint32_t iteration, int64_t first,
                  int64_t n, double *p0, double *v0, double *g, int32_t *hit);

/* Initial conditions exactly as a single-thread reference run draws them
 * (monteTraj::runTraj / setUpTraj, src/monteTraj.cpp:174-278, 301-374):
 * per CCS integral a fresh std::mt19937_64(seed) discarded by the previous
 * integral's draw count, one rand() stream seeded 1 for the whole run.
 * Host-side; iters*traj entries each. */
int cosims_sample_reference(const cosims_mol *m, uint64_t seed, int32_t iters, int64_t traj, double *p0, double *v0,
                            double *g, int32_t *hit);

/* diffEqSolver::verletVel (src/diffEqSolver.cpp:81-211) for n trajectories,
 * one GPU lane each.  Outputs: chi (rad), ok (return value), steps
 * (p.stepCount), ergy (% energy deviation of the accepted attempt),
 * total_steps (all steps over all attempts; may be NULL). */
int cosims_verlet(const cosims_mol *m, int32_t prec, int64_t n, const double *p0, const double *v0,
                  double *chi, int32_t *ok, int32_t *steps, double *ergy, int64_t *total_steps);

/* monteTraj::runTraj (src/monteTraj.cpp:281-454): iter CCS integrals of
 * traj trajectories, all in one persistent launch per attempt round. */
typedef struct cosims_run_cfg {
    int32_t prec;          /* COSIMS_PREC_* */
    uint32_t job;          /* Philox key extension (request id) */
    int32_t iter;          /* 0 => params.iter */
    int64_t traj;          /* 0 => params.traj */
    int32_t lanes;         /* lanes per trajectory: 0 tuned, 1 lane-per-trajectory
                              (atom-order sums, bit-identical FP64), 4..64 group kernel */
    int32_t wg_size;       /* 0 => tuned default */
    int32_t blocks_per_cu; /* 0 => occupancy limit */
    int32_t rng;           /* COSIMS_RNG_*: initial-condition stream */
    /* optional per-trajectory outputs (iter*traj entries each, may be NULL) */
    double *chi;
    int32_t *status;       /* 0 miss, 1 success, 2 failure */
    double *g;
} cosims_run_cfg;

#define COSIMS_MAX_ROUNDS 4
typedef struct cosims_ccs {
    double mean, sd, wmean, werr, max_ergy;   /* src/monteTraj.cpp:412-442 */
    int64_t n_traj, n_hit, n_fail, n_retry;
    int64_t total_steps;                      /* integration steps, all attempts */
    int64_t pair_evals;                       /* force-loop atom pairs */
    int64_t pot_evals;                        /* getPotential atom pairs */
    double run_ms;                            /* trajectory phase wall time */
    int32_t n_rounds;
    double traj_kernel_ms[COSIMS_MAX_ROUNDS]; /* HIP-event time of each trajectory launch */
    int64_t round_steps[COSIMS_MAX_ROUNDS];   /* integration steps of each launch */
    int64_t round_evals[COSIMS_MAX_ROUNDS];   /* force evaluations of each launch */
    int64_t round_items[COSIMS_MAX_ROUNDS];   /* trajectories queued for each launch */
    double check_kernel_ms, sample_kernel_ms;
    int32_t lanes, wg_size;                   /* launch shape used */
    /* Clustered force types (group kernels), over every force evaluation:
     * clusters examined, far clusters (multipole terms), far clusters with
     * dipole terms, member atoms with charges and LJ, with charges only,
     * with LJ only. */
    int64_t cl_counts[6];
    /* Reference FLOPs of the trajectory launches, counted per branch as the
     * reference writes them: type 4 58 a pair + 21 an evaluation
     * (src/forceAccel.cpp:764-802), type 6 22 a pair (:865-880); type 1/2
     * 12 a near cluster's header, 116 a far cluster with dipole terms (49
     * without, +89 at order 2), 62 an atom with charges and LJ, 48 with
     * charges only, 22 with LJ only, 27 an evaluation (:244-442); type 3 10
     * a header, 62 / 48 an atom, 21 an evaluation (:655-746); type 5 10 a
     * header, 22 an atom (:819-851); plus 50 an integration step
     * (src/diffEqSolver.cpp:131-174) and 24 (charged) or 15 an atom of
     * getPotential (:954-972).  The lane-per-trajectory kernel counts
     * neither potential pairs nor branches: its figure leaves out the
     * potential, and is 0 for the clustered types. */
    double ref_flops;
} cosims_ccs;
/* omega/omega_err: iter entries each. */
int cosims_run(const cosims_mol *m, const cosims_params *p, const cosims_run_cfg *cfg, double *omega,
               double *omega_err, cosims_ccs *out);

/* cosims_run without device allocation.  No counterpart in the reference:
 * it is for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  The
 * library calls no hipMalloc, hipFree, stream or event creation in it: the
 * run's device buffers are slices of a static device array (27.3 MB on every
 * device the library runs on, held until the process ends), its work goes to
 * the calling thread's default stream (hipStreamPerThread), and the
 * per-launch times (traj_kernel_ms, sample_kernel_ms) stay 0.  Limits: at
 * most COSIMS_NOALLOC_MAX_TRAJ trajectories (iter x traj) in at most
 * COSIMS_NOALLOC_MAX_ITER integrals (COSIMS_E_INVALID beyond), Philox
 * initial conditions, the group kernels (lanes 4..64 and wg_size given, or
 * lanes 0 for the shape cosims_run has tuned for runs of this size), a
 * molecule already on the current device (cosims_mol_upload).  Results equal
 * cosims_run's with the same configuration bit for bit.  Calls on one device
 * share its array: they run one at a time (a lock per device); the array
 * keeps the last call's contents, and every buffer a call reads it writes
 * first. */
#define COSIMS_NOALLOC_MAX_TRAJ 262144
#define COSIMS_NOALLOC_MAX_ITER 256
int cosims_run_noalloc(const cosims_mol *m, const cosims_params *p, const cosims_run_cfg *cfg, double *omega,
                       double *omega_err, cosims_ccs *out);

/* Many CCS runs in one call.  The reference runs one molecule per process
 * (src/main.cpp:119-140, one monteTraj::runTraj each); here all jobs are
 * sampled into one work list, ordered by launch shape (force type x lanes
 * per trajectory x workgroup size), larger molecules first.  By default
 * the jobs of a launch shape whose molecule fits one workgroup's
 * LDS share one persistent launch: each workgroup stages the molecule of the
 * job it works on in LDS, refills its groups from that job's trajectories
 * (longest expected first), and when they are spent moves to the job with
 * the most trajectories left, restaging; larger molecules of the exact types
 * run tiled launches of their own and the clustered types share launches
 * that read the molecules through L2.  With cfg.own_launches, a job of an
 * exact type with enough work runs a launch of its own (LDS-staged or
 * tiled) and only the rest share launches.  The launches of a batch run concurrently on
 * separate streams.  Each job's results are those cosims_run gives for the
 * same molecule, seed, job id and launch shape, in either mode.
 *
 * Lanes per trajectory follow the molecule's size: size class k holds the
 * exact-type molecules of n atoms with 48 * 4 * 2^(k-1) <= n < 48 * 4 * 2^k
 * (class 0 below 192, class 4 from 1536 up); the size rule gives class k
 * 4 * 2^k lanes, the tuner times each class's choices.  Molecules of the
 * clustered types (1, 2, 3, 5) are in class 4 and run 64 lanes (one group a
 * wavefront, the cluster-list kernel) unless `lanes` fixes every job: with
 * fewer lanes they run 4-14x slower at 300-5,000 atoms.  Only the class's
 * workgroup size applies to them. */
#define COSIMS_SIZE_CLASSES 5

typedef struct cosims_batch_job {
    const cosims_mol *mol;
    uint64_t seed;            /* Philox key (seed_ext) */
    uint32_t job;             /* Philox key extension */
    int32_t iter;             /* CCS integrals; 0 => the molecule's params */
    int64_t traj;             /* trajectories per integral; 0 => the molecule's params */
    double *omega, *omega_err;/* iter entries each (caller-owned) */
    cosims_ccs *out;          /* per-job statistics (lanes / wg_size: the job's launch shape) */
} cosims_batch_job;

typedef struct cosims_batch_cfg {
    int32_t prec;             /* COSIMS_PREC_* */
    int32_t lanes;            /* > 0 (4..64): every job; 0: the size rule; -1: per size class,
                                 from class_lanes / class_wg, or tuned on the batch when all zero */
    int32_t wg_size;          /* workgroup size where no class value applies; 0 => 512 */
    int32_t blocks_per_cu;    /* 0 => occupancy limit */
    int32_t class_lanes[COSIMS_SIZE_CLASSES];
    int32_t class_wg[COSIMS_SIZE_CLASSES];
    int32_t own_launches;     /* 0 (default): the jobs of a launch shape within LDS share one persistent
                                 launch, each workgroup staging the molecule of the job it works on;
                                 1: a job of an exact type whose work is large enough runs a launch
                                 of its own (its molecule staged in LDS, or tiled through LDS) */
} cosims_batch_cfg;

typedef struct cosims_batch_stats {
    int32_t lanes, wg_size, launches;
    int32_t class_lanes[COSIMS_SIZE_CLASSES], class_wg[COSIMS_SIZE_CLASSES];
    double kernel_ms;         /* first launch to last launch end (HIP events) */
    double sample_ms, run_ms; /* sampling launches; whole call */
    int64_t n_traj, n_hit, total_steps, force_evals, pot_evals;
    int64_t pair_evals;       /* atom pairs of the force evaluations (exact types: every atom;
                                 clustered types: the member atoms their branches evaluate) */
    double ref_flops;         /* reference FLOPs of all launches (as cosims_ccs.ref_flops) */
} cosims_batch_stats;

int cosims_run_batch(cosims_batch_job *jobs, int32_t njobs, const cosims_batch_cfg *cfg,
                     cosims_batch_stats *stats);
/* Per size class, the lanes and workgroup size the tuner picks for these
 * jobs in the default (shared-launch) mode (0 for classes without jobs). */
int cosims_tune_batch(const cosims_batch_job *jobs, int32_t njobs, int32_t prec, int32_t *class_lanes,
                      int32_t *class_wg);

/* ------------------------------------------------------------ serving
 * A CCS request server; the reference runs one molecule per process
 * (src/main.cpp:119-140) and has no counterpart.  A request (atoms, run
 * parameters, target relative standard error of the weighted-mean CCS) is set
 * up on a host worker (boundary marches on the GPU when params.setup_gpu),
 * admitted while an active slot and the device-memory budget allow, and run
 * in blocks: one CCS integral of block_traj trajectories each, Philox
 * (params.seed, job) with iteration = block index.  Each slice is one
 * persistent launch per launch shape over blocks of every active request;
 * requests arriving meanwhile join the next slice.  Each block's CCS and the
 * stopping rule are computed on the device as the slice completes: a request
 * retires at the first block, from min_blocks on, where werr / wmean <=
 * target (or at max_blocks); its result is that of cosims_run with iter = blocks,
 * traj = block_traj and the request's launch shape (fixed lanes, or the size
 * rule of cosims_run_batch), whatever shared the launches.  Per-block sums
 * stay on the device until read. */
typedef struct cosims_server cosims_server;

typedef struct cosims_server_cfg {
    int32_t prec;             /* COSIMS_PREC_* */
    int32_t lanes;            /* lanes per trajectory for every request; 0 => the size rule */
    int32_t wg_size;          /* 0 => 512 */
    int64_t hbm_budget;       /* device bytes for admitted requests; 0 => 80 % of free memory */
    int32_t max_active;       /* admitted requests at once; 0 => 64 */
    int32_t min_blocks;       /* blocks before the stopping rule applies; 0 => 3 */
    int32_t max_blocks;       /* blocks at most per request; 0 => 1000 */
    int32_t setup_threads;    /* host setup workers; 0 => 8 */
    int64_t block_traj;       /* trajectories per block; 0 => 10000 */
    int64_t slice_traj;       /* trajectories per slice at most; 0 => 2^20 */
    double slice_ms;          /* kernel time per slice to aim for; 0 => 150 */
    int32_t policy;           /* COSIMS_SERVE_* */
    int64_t min_slice_traj;   /* shortest-first slices are topped up to this many; 0 => 65536 */
    int32_t pipeline_depth;   /* slices in flight at once (1..8); 0 => 3 */
    double issue_fraction;    /* once min_blocks are in: share of the estimated remaining blocks
                                 issued per slice (0..1]; 0 => 1 (all of the estimate) */
    double prior_rel_sd;      /* before any block is in, issue the blocks a relative spread of the
                                 per-trajectory CCS terms this large would need for the target
                                 ((prior_rel_sd / target)^2 trajectories, at least min_blocks);
                                 0 => min_blocks only */
    int32_t class_lanes[COSIMS_SIZE_CLASSES]; /* with lanes 0: lanes of each size class (e.g. from
                                                 cosims_tune_batch); 0 => the size rule */
    int32_t class_wg[COSIMS_SIZE_CLASSES];    /* workgroup size of each class; 0 => wg_size */
    int32_t own_launches;     /* 0: a slice's requests of a launch shape share one persistent launch with
                                 their molecules staged in LDS (as cosims_run_batch's default); 1: a
                                 launch of its own per request with enough work */
} cosims_server_cfg;

/* Slice composition: shortest remaining work first (requests ordered by
 * blocks still wanted x measured cost per block fill the slice's work
 * budget), or an equal share of blocks for every active request. */
#define COSIMS_SERVE_SHORTEST_FIRST 0
#define COSIMS_SERVE_EQUAL_SHARE 1

typedef struct cosims_request {
    cosims_params params;     /* as for cosims_mol_create; seed keys the request's stream */
    cosims_atoms atoms;       /* copied at submit */
    uint32_t job;             /* Philox key extension */
    double target_rel_err;    /* 0 => 0.005 */
} cosims_request;

typedef struct cosims_result {
    int32_t status;           /* 0 done, 1 pending, < 0 error code */
    int32_t blocks;           /* blocks in the result */
    cosims_ccs ccs;           /* statistics over those blocks */
    double rel_err;           /* werr / wmean */
    double t_submit_ms, t_ready_ms, t_admit_ms, t_done_ms;  /* server clock */
    double setup_ms;
    int64_t block_traj;
    /* setup phases within setup_ms (as cosims_mol_info): boundary marches,
     * PCM clustering, cluster multipoles; then the molecule's device upload */
    double setup_boundary_ms, setup_cluster_ms, setup_moments_ms, setup_upload_ms;
} cosims_result;

typedef struct cosims_server_stats {
    int64_t submitted, done, failed, queued, active;
    int64_t slices, launches, traj, pair_evals, force_evals, total_steps;
    double kernel_ms, busy_ms, now_ms;
    int64_t bytes_in_use, bytes_peak, budget;
    double ref_flops;         /* reference FLOPs of every launch (as cosims_ccs.ref_flops) */
} cosims_server_stats;

int cosims_server_create(const cosims_server_cfg *cfg, cosims_server **out);
int cosims_server_submit(cosims_server *s, const cosims_request *r, int64_t *ticket);
/* Waits up to timeout_ms (< 0: no limit); COSIMS_E_PENDING if not done.
 * omega/omega_err (may be NULL) receive up to cap per-block values. */
int cosims_server_wait(cosims_server *s, int64_t ticket, double timeout_ms, cosims_result *out, double *omega,
                       double *omega_err, int32_t cap);
int cosims_server_stats_get(cosims_server *s, cosims_server_stats *out);
/* Finishes every submitted request, then frees the server. */
void cosims_server_destroy(cosims_server *s);

const char *cosims_strerror(int code);
const char *cosims_last_error(void);
/* Device name and compute-unit count of the current device. */
int cosims_device_info(char *name, size_t cap, int32_t *cus);
/* Measured FP32 FMA rate of the current device for code built as this
 * library is (no packed FP32; all operands in VGPRs, which gfx950 issues at
 * 2 cycles a wavefront), in TFLOP/s. */
int cosims_device_fma_rate(double *tflops);
/* Measured rate of the force sum alone on this molecule (exact force
 * types 4 and 6, molecule within one workgroup's LDS): the trajectory
 * kernels' force call (per-pair arithmetic, group sums over `lanes` lanes,
 * epilogue) with the molecule staged in LDS, a full-occupancy grid of
 * wg_size-thread workgroups, every group evaluating the same number of
 * forces -- no Verlet bookkeeping, no energy checks, no launch drain.  The
 * ceiling of the trajectory kernels' instruction mix, in reference TFLOP/s
 * (58 FLOP a pair + 21 an evaluation for type 4, src/forceAccel.cpp:764-802;
 * 22 a pair for type 6, :865-880), atom pairs per second and the launch's
 * time (best of three).  No counterpart in the reference (a benchmark
 * query); pairs_per_s and kernel_ms may be NULL. */
int cosims_force_rate(const cosims_mol *m, int32_t prec, int32_t lanes, int32_t wg_size, double *tflops,
                      double *pairs_per_s, double *kernel_ms);
/* Waits for all work on the current device (benchmark brackets). */
int cosims_device_sync(void);

#ifdef __cplusplus
}
#endif
#endif
