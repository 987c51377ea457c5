// ref_harness.cpp -- C entry points over the compiled reference sources.
//
// TEST INFRASTRUCTURE ONLY.  Built by oracle/Makefile into
// oracle/_ref/libcosims_ref.so from the unmodified reference translation
// units (everything in src/ except main.cpp) plus this file.  It is used to
// write the golden vectors under tests/golden/ and to pin the oracle; it is
// never part of the engine.
//
// The globals below are the ones src/main.cpp:39-82 defines; the setup
// sequence follows src/main.cpp:96-138 and monteTraj::monteTraj
// (src/monteTraj.cpp:26-95), calling the reference's own classes.
#include <cstdlib>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "diffEqSolver.h"
#include "fileParams.h"
#include "impactParam.h"
#include "maxBoltz.h"
#include "molecule2.h"
#include "monteTraj.h"
#include "myStructs.h"
#include "PCM_Cluster.h"
#include "VelocityDistribution.h"

ofstream fileStream_ext;
string omegaFilePath_ext;
string fileLocaiton_ext;
string molFilePath_ext;
string molFileName_ext;
string atomFilePath_ext;
string dataFilePath_ext;
string xyzFilePath_ext;
string iterFilePath_ext;
string name_ext;
string subDir_ext;
string printMolName_ext;
double temp_ext;
double dt_ext;
double mass_ext;
double mu;
int trajNum_ext;
int iterations_ext;
int numThreads_ext;
long long int seed_ext;
int printRate_ext;
int maxClusterSize_ext;
int multipole_order_ext;
double cutOffRadius_ext;
double gMax_ext;
double gMin_ext;
double totalCharge_ext;
bool displayProg_ext;
double dispersion_radius_ext;
double multipole_radius_ext;
bool multipole_ext;
bool dispersionCutoff_ext;
bool charge_ext;
bool projection_ext;
bool shorten_ext;
bool printTraj_ext;
bool printTrajOnFail_ext;
bool printData_ext;
bool printDataOnFail_ext;
bool printMol_ext;
bool printToColsole_ext = false;

namespace {
molecule2 *g_mol = nullptr;
impactParam *g_ip = nullptr;
diffEqSolver *g_solver = nullptr;
PCM_Cluster *g_pcm = nullptr;
VelocityDistribution g_vd;
double g_scal[32];
}

extern "C" {

// Runs the reference's command-line interpretation and setup.
// argv-style: "-m <mol> [-i <input>] [-a <atoms>] -s <seed>", cwd = out_dir.
int ref_setup(int argc, char **argv, const char *out_dir)
{
    fileParams files;
    std::vector<char> cwd(out_dir, out_dir + std::strlen(out_dir) + 1);
    if (!files.interpretCMD(argc, argv, cwd.data())) return -1;
    void *mem = std::calloc(1, sizeof(molecule2)); // zeroed: the reference ctor leaves totals unset
    g_mol = new (mem) molecule2();
    if (files.getMolFileExt().compare("pdb") == 0) g_mol->importPDB();
    else g_mol->importMFJ(files.getNameChanged(), files.getChargeChanged());
    if (g_mol->numAtoms == 0) return -2;
    files.printInfo();
    g_mol->printMolStats();

    // monteTraj::monteTraj, src/monteTraj.cpp:31-95
    double heliumMass = 4.002602;
    mass_ext = mu = g_mol->totalMass * heliumMass / (g_mol->totalMass + heliumMass);
    double beta = mu / (K * temp_ext);
    g_vd.setProperties(5, beta / 2.0);
    double gMin = g_vd.mean() - 2 * g_vd.stdDev();
    double gMax = g_vd.mean() + 4 * g_vd.stdDev();
    g_vd.renormalize(gMin, gMax);
    double maxVal = g_vd(g_vd.mode());
    g_ip = new impactParam();
    g_ip->setMol(*g_mol);
    double minKin = 0.5 * mass_ext * pow(g_vd.mean(2, beta / 2.0) - 2 * g_vd.stdDev(2, beta / 2.0), 2);
    g_ip->setBoundry(minKin);
    g_ip->findEllipsoidBound();
    g_ip->printBoundryResults();
    double bMax = g_ip->getMax();
    g_solver = new diffEqSolver();
    g_solver->setMol(*g_mol);
    double const1 = pow(beta, 3) / (64 * M_PI);
    const1 = const1 * 8 * M_PI * M_PI;
    const1 = const1 / g_vd.normFactor();

    g_pcm = new PCM_Cluster();
    for (int i = 0; i < g_mol->numAtoms; i++) g_pcm->addCoord(g_mol->data[i].coords);
    g_pcm->formClusters();

    double s[32] = {0};
    s[0] = mu; s[1] = 28276.01363 / mass_ext; s[2] = g_ip->a; s[3] = g_ip->b; s[4] = g_ip->c;
    s[5] = bMax; s[6] = gMin; s[7] = gMax; s[8] = g_vd.mean(); s[9] = g_vd.mode(); s[10] = g_vd.stdDev();
    s[11] = g_vd.normFactor(); s[12] = maxVal; s[13] = g_vd.cdf(gMax) - g_vd.cdf(gMin); s[14] = const1;
    s[15] = minKin; s[16] = dispersion_radius_ext; s[17] = multipole_radius_ext; s[18] = g_mol->totalMass;
    s[19] = charge_ext; s[20] = multipole_ext; s[21] = dispersionCutoff_ext; s[22] = multipole_order_ext;
    s[23] = maxClusterSize_ext; s[24] = projection_ext; s[25] = dt_ext; s[26] = temp_ext;
    s[27] = g_mol->maxX; s[28] = g_mol->maxY; s[29] = g_mol->maxZ; s[30] = totalCharge_ext;
    s[31] = g_pcm->getNumClusters();
    std::memcpy(g_scal, s, sizeof(s));
    return g_mol->numAtoms;
}

void ref_scalars(double *out) { std::memcpy(out, g_scal, sizeof(g_scal)); }

void ref_atoms(double *xyz, double *q, double *sig6, double *eps24m, double *eps4, double *mass)
{
    for (int i = 0; i < g_mol->numAtoms; i++) {
        const dataList &d = g_mol->data[i];
        xyz[3 * i] = d.coords.X; xyz[3 * i + 1] = d.coords.Y; xyz[3 * i + 2] = d.coords.Z;
        q[i] = d.charge; sig6[i] = d.sigAvg6; eps24m[i] = d.eps24_mass; eps4[i] = d.eps4; mass[i] = d.mass;
    }
}

int ref_clusters(int *start, int *atoms, double *center, double *maxdist)
{
    int nc = g_pcm->getNumClusters(), off = 0;
    start[0] = 0;
    for (int i = 0; i < nc; i++) {
        std::vector<int> c = g_pcm->getCluster(i);
        for (size_t k = 0; k < c.size(); k++) atoms[off++] = c[k];
        start[i + 1] = off;
        vector3D<double> ce = g_pcm->getCenter(i);
        center[3 * i] = ce.X; center[3 * i + 1] = ce.Y; center[3 * i + 2] = ce.Z;
        maxdist[i] = g_pcm->getMaxDist(i);
    }
    return nc;
}

// Rebuild the force evaluator with other force-type switches
// (src/forceAccel.cpp:125-140 reads these globals in setClusters).
void ref_set_flags(int charge, int multipole, int cutoff, int order)
{
    charge_ext = charge != 0;
    multipole_ext = multipole != 0;
    dispersionCutoff_ext = cutoff != 0;
    multipole_order_ext = order;
    delete g_solver;
    g_solver = new diffEqSolver();
    g_solver->setMol(*g_mol);
}

void ref_accel(int n, const double *pos, double *acc)
{
    for (int i = 0; i < n; i++) {
        vector3D<double> p(pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]);
        vector3D<double> a = g_solver->forceFunct.getReducedVec(p);
        acc[3 * i] = a.X; acc[3 * i + 1] = a.Y; acc[3 * i + 2] = a.Z;
    }
}

void ref_potential(int n, const double *pos, double *u)
{
    for (int i = 0; i < n; i++) {
        vector3D<double> p(pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]);
        u[i] = g_solver->forceFunct.getPotential(p);
    }
}

void ref_verlet(int n, const double *p0, const double *v0, double a, double b, double c, double *chi, int *ok,
                int *steps, double *ergy)
{
    for (int i = 0; i < n; i++) {
        vector3D<double> p(p0[3 * i], p0[3 * i + 1], p0[3 * i + 2]);
        vector3D<double> v(v0[3 * i], v0[3 * i + 1], v0[3 * i + 2]);
        localThreadParams lp;
        lp.stepCount = 0;
        lp.maxErgyPctDev = 0;
        double x = 0;
        ok[i] = g_solver->verletVel(dt_ext, p, v, a, b, c, x, lp) ? 1 : 0;
        chi[i] = x;
        steps[i] = lp.stepCount;
        ergy[i] = lp.maxErgyPctDev;
    }
}

// Initial conditions from the reference's own setUpTraj, drawn the way a
// single-thread runTraj draws them (src/monteTraj.cpp:301-374).
int ref_sample(long long seed, int iters, int traj, double *p0, double *v0, double *g, int *hit, int *tries)
{
    static threadSharedParams shared;
    static monteTraj *mt = nullptr;
    if (!mt) mt = new monteTraj(*g_mol, shared);
    srand(1);
    unsigned restart = 0;
    for (int it = 0; it < iters; it++) {
        mt19937_64 generator(seed);
        generator.discard(it == 0 ? 0 : restart);
        uniform_real_distribution<double> distr(0.0, 1.0);
        localThreadParams p;
        p.ranNumCount = 0;
        p.missCount = 0;
        for (int n = 0; n < traj; n++) {
            vector3D<double> ip, iv;
            const int o = it * traj + n;
            int before = p.ranNumCount;
            bool h = mt->setUpTraj(0, n, ip, iv, p, generator, distr);
            hit[o] = h ? 1 : 0;
            g[o] = p.g;
            tries[o] = (p.ranNumCount - before) / 8;
            p0[3 * o] = ip.X; p0[3 * o + 1] = ip.Y; p0[3 * o + 2] = ip.Z;
            v0[3 * o] = iv.X; v0[3 * o + 1] = iv.Y; v0[3 * o + 2] = iv.Z;
        }
        restart = p.ranNumCount;
    }
    return 0;
}

} // extern "C"
