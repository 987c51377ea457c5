// prep.hpp -- host-side molecule setup of the trajectory engine.
//
// Everything the persistent trajectory kernel reads is produced here, in
// IEEE double with the reference's operand order (the file is compiled with
// -ffp-contract=off), so that a prepared molecule is bit-identical to what
// the reference's molecule2 / impactParam / PCM_Cluster / forceAccel leave
// in memory.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace cosims {

constexpr double kPi = 3.14159265358979323846;    // src/vector3D.h:250
constexpr double kPi2 = 1.57079632679489661923;   // src/vector3D.h:251
constexpr double kBoltz = 0.831446214563;         // src/maxBoltz.h:29
constexpr double kHeMass = 4.002602;              // src/molecule2.h:72
constexpr double kAlphaKe = 28276.01363;          // src/forceAccel.cpp:35

struct TypeRow {
    std::string name, symbol;
    double mass, sigma, eps;
    int intmass;
};

std::vector<TypeRow> default_types();                                    // assignDefaultFF
bool apply_atoms_file(std::vector<TypeRow> &t, const std::string &path); // importFile

struct MolFile {
    std::vector<double> xyz, charge;
    std::vector<std::string> symbols;
    bool charged = false;
    double total_charge = 0;
    std::string name;
};
bool read_mfj(const std::string &path, const std::vector<TypeRow> &t, MolFile &out, std::string &err);
bool read_pdb(const std::string &path, MolFile &out, std::string &err);

struct Params {
    double temp = 298, dt = 0.01;
    int64_t traj = 50000;
    int iter = 10;
    bool charge = false;
    uint64_t seed = 0;
    bool multipole = true, dispersion_cutoff = true;
    double dispersion_radius = 0, multipole_radius = 0;
    int multipole_order = 1, max_cluster_size = 4;
    bool projection = true;
    int print_rate = 20000;
    double total_charge = 0;
    int threads = 0;
};

struct VelDist {
    double a = 0, mean = 0, mode = 0, sigma = 0, gmin = 0, gmax = 0, norm = 0, max_val = 0,
           cdf_frac = 0, const1 = 0, min_kin = 0;
};
VelDist velocity_distribution(double mu, double temp);
double veldist_pdf(const VelDist &v, double g);

// Read-only data of the force evaluator (forceAccel after setMol/setClusters).
struct Prepared {
    int n = 0;
    std::vector<double> xyz, q, sig6, eps24m, eps4, mass;
    double mu = 0, total_mass = 0, dipole_scalar = 0;
    bool charged = false;
    int force_type = 6, multipole_order = 1;
    bool projection = true;
    double dispersion_radius = 0, multipole_radius = 0;
    double rot[9] = {0}, max_xyz[3] = {0};
    double a = 0, b = 0, c = 0, fit_a = 0, fit_b = 0, fit_c = 0, fit_rmsd = 0, enlarge = 0;
    int n_boundary = 0;
    VelDist vd;
    double bmax = 0, ccs_const = 0;
    // clusters (reference order)
    int ncl = 0;
    std::vector<int32_t> cl_start, cl_atoms, cl_replace;
    std::vector<double> cl_center, cl_maxdist, cl_pos, cl_q, cl_dip, cl_quad;
    // wall time of the setup phases (prepare): boundary search, PCM
    // clustering, cluster moments
    double boundary_ms = 0, cluster_ms = 0, moments_ms = 0;
};

int force_type_of(bool charged, bool multipole, bool cutoff);

// One potential march of the boundary search (impactParam::setBoundry,
// src/impactParam.cpp:602-736): points p_k = p_{k-1} + v (FP64, component
// by component) from p_0 = start, with v = dir * sl and sl = -1 when
// U(start) + mk < 0, else 1.  Points are tested from p_first (0 or 1): the
// walk stops at the first p_k with (U(p_k) + mk) * sl < 0 (a crossing), or,
// when axis is set, with the ray past the z axis (the test of :684-686).
struct March {
    double start[3], dir[3];
    double prev0[3];   // the point before p_first (p_{first-1}; zero for first = 0)
    int first, axis;
};
// Result: found 1 (crossing at p, previous point prev), 0 (left past the
// axis), -1 (no crossing within the step limit).
struct MarchHit {
    int found;
    double p[3], prev[3];
};
constexpr int64_t kMaxMarchPoints = int64_t(1) << 22;
// The z-axis walk (src/impactParam.cpp:640-655 loops until it crosses) ends
// once it is kZWalkBeyond A past the far side of the molecule, z < -(z0 +
// kZWalkBeyond) with z0 = |maxZ| + 20: a walk that finds no crossing there
// is a failed setup.  Every crossing of a real molecule lies within a few
// hundred points of the start; without the bound a walk that never crosses
// would run 2^22 points, one wavefront for tens of seconds on a molecule
// of thousands of atoms.  Rays end at the z axis (<= posmag / 0.1 + 2
// points) and keep kMaxMarchPoints.
constexpr double kZWalkBeyond = 1.0e4;
inline int64_t zwalk_points(double z0) { return (int64_t)((2 * z0 + kZWalkBeyond) / 0.5) + 2; }

// Inputs of the boundary search fixed before any march: the z-axis walk's
// start, and what the rays' geometry needs once zMax is known
// (src/impactParam.cpp:602-665).
constexpr int kBoundaryRot = 20;
struct BoundaryPlan {
    double z0;                          // z-axis walk start, |maxZ| + 20
    double posmag0;                     // |max(|maxX|, |maxY|)|
    double maxz;                        // |maxZ|
    double cos_t[kBoundaryRot], sin_t[kBoundaryRot];      // cos, sin of n 2 pi / 20
    double dir_x[kBoundaryRot], dir_y[kBoundaryRot];      // -0.1 cos, -0.1 sin
};
// Ray geometry from the z-axis crossing (src/impactParam.cpp:632-665).
struct RayGeometry {
    double zmax, posmag, zinc;
    int levels;
};
RayGeometry ray_geometry(const BoundaryPlan &bp, const MarchHit &zhit);
March ray_march(const BoundaryPlan &bp, const RayGeometry &g, int level, int n);

// Evaluates the boundary marches somewhere other than the host loop (the
// GPU).  Must give the same points as the host loop: the same FP64
// additions and the reference-order getPotential.  `boundary` runs the
// z-axis walk and every ray (levels + 1) x 20 of the geometry it implies.
struct BoundaryMarcher {
    virtual ~BoundaryMarcher() = default;
    virtual int boundary(const Prepared &m, double mk, const BoundaryPlan &bp, MarchHit &zhit,
                         std::vector<MarchHit> &rays, std::string &err) = 0;
};

// Full setup from raw atoms.  Returns 0, or a negative engine error code with
// a message in err.  With a marcher the boundary marches run there.
int prepare(const Params &p, int n, const double *xyz, const double *mass, const double *sigma,
            const double *eps, const double *charge, double sigma_max, Prepared &out, std::string &err,
            BoundaryMarcher *bm = nullptr);

// Completes an injected prepared molecule (cluster moments, velocity window,
// CCS constant).
int complete_prepared(const Params &p, Prepared &m, std::string &err);

// Initial conditions in the reference's own random stream for a single
// OpenMP thread (monteTraj::runTraj / setUpTraj, src/monteTraj.cpp:174-278,
// 301-369): per CCS integral a fresh std::mt19937_64(seed) discarded by the
// previous integral's draw count (restartGen, :344-349, :374), and one
// process-wide rand() stream seeded 1.  Writes iters*traj entries.
void sample_compat(const Prepared &P, uint64_t seed, int iters, int64_t traj, double *p0, double *v0, double *g,
                   int32_t *hit);

// Records the message returned by cosims_last_error() on this thread and
// returns code (defined in engine.hip).
int set_last_error(int code, const std::string &msg);

// getPotential on the host (boundary search).
double potential_host(const Prepared &m, const double pos[3]);

}  // namespace cosims
