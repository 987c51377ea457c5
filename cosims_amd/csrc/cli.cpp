// cli.cpp -- the command-line program `cosims`, a drop-in for the
// reference's executable (src/main.cpp:94-160): same flags, same input file,
// same molecule files, and the same omega_<name>.txt log, with the trajectory
// phase run by libcosims_engine on the GPU.  It uses the C ABI only.
//
//   cosims -m mol.{pdb,mfj} [-i input] [-a atoms.dat] [-n threads] [-s seed]
//          [-p mixed|fp64] [-r philox|reference] [-l lanes] [-g 1|0]
//
// -m -i -a -n -s are fileParams::interpretCMD's (src/fileParams.cpp:78-201);
// -n sets the host threads of the molecule setup.  The engine's own flags:
// -p precision (default mixed: FP32 force sum, FP64 integration and CCS),
// -r initial-condition stream (default philox; `reference` draws the
// reference's single-thread mt19937_64 stream), -l lanes per trajectory
// (default 0, tuned; 1 = one trajectory per lane, the reference's operation
// order), -g boundary search on the GPU (default 1) or host threads (0).
//
// Log lines follow the reference's formats: fileParams::printInfo
// (src/fileParams.cpp:513-636), molecule2::printMolStats
// (src/molecule2.cpp:905-935), impactParam::printBoundryResults
// (src/impactParam.cpp:997-1056), the cluster summary of forceAccel
// (src/forceAccel.cpp:91-120), the velocity statistics of monteTraj
// (src/monteTraj.cpp:97-108), and runTraj / dispResults
// (src/monteTraj.cpp:288-452, 520-538).  Progress lines are the ones a
// one-thread reference run prints, recomputed from the per-trajectory
// deflection angles in trajectory order.  The whole run is one engine call,
// so the per-integral times are the trajectory phase's wall time shared out
// evenly over the integrals.
#include <unistd.h>

#include <algorithm>
#include <cctype>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <ctime>
#include <fstream>
#include <random>
#include <sstream>
#include <string>
#include <vector>

#include "cosims_engine.h"

namespace {

std::string fmt(const char *f, ...)
{
    char buf[4096];
    va_list ap;
    va_start(ap, f);
    std::vsnprintf(buf, sizeof buf, f, ap);
    va_end(ap);
    return buf;
}

// Every line goes to stdout and to omega_<name>.txt (fileStream_ext).
struct Log {
    FILE *f = nullptr;
    void open(const std::string &path) { f = std::fopen(path.c_str(), "w"); }
    void put(const std::string &s, bool to_file = true, bool to_console = true)
    {
        if (to_console) {
            std::fputs(s.c_str(), stdout);
            std::fflush(stdout);
        }
        if (to_file && f) {
            std::fputs(s.c_str(), f);
            std::fflush(f);
        }
    }
    ~Log()
    {
        if (f) std::fclose(f);
    }
};

bool is_int(const char *s)  // fileParams::isInt (src/fileParams.cpp:53-59)
{
    for (; *s; s++)
        if (!std::isdigit((unsigned char)*s)) return false;
    return true;
}

int length_of_int(long long v)  // fileParams::length_of_int (src/fileParams.cpp:36-41)
{
    int n = 0;
    for (; v != 0; v /= 10) n++;
    return n;
}

std::string now_text()
{
    std::time_t t = std::time(nullptr);
    return std::ctime(&t);  // ends in '\n', as the reference prints it
}

struct Cmd {
    std::string mol, input, atoms;
    int threads = -1;
    bool seed_set = false;
    long long seed = 0;
    int prec = COSIMS_PREC_MIXED;
    int rng = COSIMS_RNG_PHILOX;
    int lanes = 0;
    int setup_gpu = 1;
};

// fileParams::interpretCMD's flag loop (src/fileParams.cpp:98-150): pairs
// "-x value" with a one-letter flag and a value that does not start with
// '-'; anything else is an error.
bool parse_cmd(int argc, char **argv, Cmd &c)
{
    for (int i = 1; i < argc; i += 2) {
        if (i + 1 >= argc) return false;
        const char *o = argv[i], *v = argv[i + 1];
        if (!(o[0] == '-' && v[0] != '-' && std::strlen(o) == 2)) return false;
        switch (o[1]) {
        case 'i': c.input = v; break;
        case 'm': c.mol = v; break;
        case 'a': c.atoms = v; break;
        case 'n':
            if (!is_int(v)) return false;
            c.threads = std::atoi(v);
            break;
        case 's':
            if (!is_int(v)) return false;
            c.seed = std::atoll(v);
            c.seed_set = true;
            break;
        case 'p':
            if (!std::strcmp(v, "mixed")) c.prec = COSIMS_PREC_MIXED;
            else if (!std::strcmp(v, "fp64")) c.prec = COSIMS_PREC_FP64;
            else return false;
            break;
        case 'r':
            if (!std::strcmp(v, "philox")) c.rng = COSIMS_RNG_PHILOX;
            else if (!std::strcmp(v, "reference")) c.rng = COSIMS_RNG_REFERENCE;
            else return false;
            break;
        case 'l':
            if (!is_int(v)) return false;
            c.lanes = std::atoi(v);
            break;
        case 'g':
            if (!is_int(v)) return false;
            c.setup_gpu = std::atoi(v) != 0;
            break;
        default: return false;
        }
    }
    return true;
}

// The project name fileParams::setDefaultParams derives from the molecule
// path: the file name up to its first '.' (src/fileParams.cpp:203-224).
std::string default_name(const std::string &mol)
{
    std::string t;
    for (int i = (int)mol.size() - 1; i >= 0; i--) {
        if (mol[i] == '/' || mol[i] == '\\') break;
        if (mol[i] == '.') t.clear();
        else t += mol[i];
    }
    std::reverse(t.begin(), t.end());
    return t;
}

std::string abs_path(const std::string &cwd, const std::string &p)
{
    return (!p.empty() && p[0] == '/') ? p : cwd + "/" + p;
}

// What the CLI needs of the input file beyond cosims_params_read: the keys
// that steer names, paths and messages, and importParams' per-line notes
// (src/fileParams.cpp:296-484).
struct InputScan {
    bool name = false, dir = false, seed = false, print_rate = false, charge = false;
    std::string name_v, dir_v;
    double print_rate_v = 0;
    std::string notes;
};

InputScan scan_input(const std::string &path, const std::string &shown)
{
    static const char *known[] = {"name", "dir", "dt", "temp", "traj", "iter", "threads", "seed",
                                  "gmax", "gmin", "print_mol", "proj", "print_traj", "print_traj_fail",
                                  "print_data", "print_data_fail", "dispersion_cutoff", "dispersion_radius",
                                  "multipole_radius", "multipole_order", "print_rate", "multipole",
                                  "max_cluster_size", "charge"};
    InputScan s;
    std::ifstream f(path);
    std::string line;
    std::vector<std::string> seen;
    int ln = 1;
    while (std::getline(f, line)) {
        if (!line.empty()) {
            std::istringstream ss(line);
            std::string opt, val;
            ss >> opt >> val;
            if (val.empty()) {
                s.notes += fmt("\t ERROR: Line number %i in %s is incomplete;\n"
                               "\t This line wil be ignored.\n\n",
                               ln, shown.c_str());
            } else if (opt[0] != '#') {
                bool ok = false;
                for (const char *k : known) ok |= opt == k;
                if (!ok) {
                    s.notes += fmt("\tERROR: Can not interpret option '%s' on \n\tline number %i in %s;\n"
                                   "\tThis line wil be ignored.\n\n",
                                   opt.c_str(), ln, shown.c_str());
                } else {
                    if (std::find(seen.begin(), seen.end(), opt) != seen.end()) {
                        std::string up = opt;
                        std::transform(up.begin(), up.end(), up.begin(), ::toupper);
                        s.notes += fmt("\tWARNING: Parameter %s already specified at line number %i;\n"
                                       "\tMost recent value will be applied.\n\n",
                                       up.c_str(), ln);
                    } else {
                        seen.push_back(opt);
                    }
                    if (opt == "name") { s.name = true; s.name_v = val; }
                    else if (opt == "dir") { s.dir = true; s.dir_v = val; }
                    else if (opt == "seed") s.seed = true;
                    else if (opt == "charge") s.charge = true;
                    else if (opt == "print_rate") { s.print_rate = true; s.print_rate_v = std::atof(val.c_str()); }
                }
            }
        }
        ln++;
    }
    return s;
}

int fail_engine(Log &log, const char *what, int rc)
{
    log.put(fmt("ERROR: %s: %s (%s)\nProgram will now exit...\n", what, cosims_strerror(rc), cosims_last_error()));
    return 1;
}

}  // namespace

int main(int argc, char **argv)
{
    Log log;
    Cmd cmd;
    if (!parse_cmd(argc, argv, cmd)) {
        std::printf("ERROR: Check command line arguments.\nProgram will now exit...\n");
        return 0;
    }
    char cwd_buf[4096];
    std::string cwd = getcwd(cwd_buf, sizeof cwd_buf) ? cwd_buf : ".";
    const auto t_start = std::chrono::steady_clock::now();

    // parameters: setDefaultParams, then the input file (importParams)
    cosims_params p;
    cosims_params_default(&p);
    std::string name = default_name(cmd.mol);
    std::string subdir;
    bool display = true;
    std::string if_notes;
    InputScan in;
    if (!cmd.input.empty()) {
        const std::string ip = abs_path(cwd, cmd.input);
        if_notes = fmt(" Reading input file: %s\n", ip.c_str());
        in = scan_input(ip, cmd.input);
        if_notes += in.notes;
        std::ifstream probe(ip);
        if (probe) {
            int rc = cosims_params_read(ip.c_str(), &p);
            if (rc) {
                std::printf("ERROR: %s\n", cosims_last_error());
                return 0;
            }
        }
        if (in.name) name = in.name_v;
        if (in.dir) subdir = in.dir_v;
        if (in.print_rate && ((int)in.print_rate_v > p.traj || (int)in.print_rate_v == 0)) display = false;
        if (in.seed) cmd.seed_set = true, cmd.seed = (long long)p.seed;
        if (!subdir.empty() && subdir.back() != '/') subdir += '/';
        name += "_" + std::to_string((int)p.total_charge);  // src/fileParams.cpp:506-509
    }
    if (!cmd.seed_set) cmd.seed = (long long)std::random_device{}();
    p.seed = (uint64_t)cmd.seed;
    if (cmd.threads > 0) p.threads = cmd.threads;
    p.setup_gpu = cmd.setup_gpu;

    // output paths (fileParams::setFileInfo, src/fileParams.cpp:638-676)
    const std::string mol_path = abs_path(cwd, cmd.mol);
    const std::string atoms_path = cmd.atoms.empty() ? std::string() : abs_path(cwd, cmd.atoms);
    std::string out_dir = subdir.empty() ? cwd + "/" : (subdir[0] == '/' ? subdir : cwd + "/" + subdir);
    const std::string omega_path = out_dir + "omega_" + name + ".txt";
    log.open(omega_path);

    std::string head = " Starting Program \n Current Date/Time: " + now_text() + "\n" +
                       " Importing and assigning molecule information...\n";
    log.put(head, false, true);
    log.put("\n" + head, true, false);
    log.put(if_notes);
    log.put(" Importing files complete.\n");

    if (cmd.mol.empty()) {
        log.put("ERROR: No molecule file was given. Check command-line aguments\n");
        std::printf("ERROR: Check command line arguments.\nProgram will now exit...\n");
        return 0;
    }
    std::string ext = mol_path.size() >= 4 ? mol_path.substr(mol_path.size() - 3) : "";
    std::transform(ext.begin(), ext.end(), ext.begin(), ::tolower);
    if (mol_path.size() < 4 || mol_path[mol_path.size() - 4] != '.' || (ext != "pdb" && ext != "mfj")) {
        std::printf("ERROR: Check command line arguments.\nProgram will now exit...\n");
        return 0;
    }

    // molecule (molecule2::importMFJ / importPDB)
    if (!atoms_path.empty()) log.put("\nAtom parameters are located at: \n" + atoms_path + "\n");
    cosims_molfile *mf = nullptr;
    if (cosims_molfile_read(mol_path.c_str(), atoms_path.empty() ? nullptr : atoms_path.c_str(), &mf)) {
        log.put(fmt("ERROR: Cannot read molecule file.\n %s\nProgram will now exit...\n", cosims_last_error()));
        return 0;
    }
    std::vector<double> charges;
    if (ext == "mfj") {
        std::string warn;
        if (in.name) warn += " Warning: Project name specified in both MFJ and input file.\n"
                             " Program will use MFJ file name\n";
        if (in.charge) warn += " Warning: Molecule charge specified in both MFJ and input file.\n"
                               " Program will use MFJ molecule charge\n";
        if (mf->name[0]) name = mf->name;
        else warn += fmt(" Warning: Project name in MFJ file is empty\n Using default name of %s\n", name.c_str());
        if (!warn.empty()) log.put("\n" + warn + "\n");
        p.charge = mf->charged;
        p.total_charge = mf->total_charge;
        if (mf->charged) charges.assign(mf->charge, mf->charge + mf->n);
    } else if (p.charge) {  // uniform charges from the input file (src/molecule2.cpp:807-816)
        charges.assign(mf->n, p.total_charge / (double)mf->n);
    }
    std::snprintf(p.name, sizeof p.name, "%s", name.c_str());
    if (p.dispersion_cutoff && p.dispersion_radius == 0)  // src/molecule2.cpp:881-887
        log.put(fmt(" Dispersion cut-off radius not specified in input file: \n\tUsing a cut-off radius of %6.2f\n"
                    "\t Determined by largest VDW radius of %5.3f\n",
                    mf->sigma_max * std::pow(150.0 / 0.5, 1.0 / 3.0), mf->sigma_max));

    // fileParams::printInfo
    {
        long long width = std::max((long long)p.traj, cmd.seed);
        width = std::max((long long)p.iter, width);
        int len = std::max(std::max((int)name.size(), length_of_int(width)), 8);
        std::string o = "\n OmegaCal will use the following settings:\n";
        o += fmt("\t Project name:                                              %*s\n", len, name.c_str());
        o += fmt("\t Number of CCS integrals:                                   %*i\n", len, p.iter);
        o += fmt("\t Number of trajectories per CCS integral:                   %*d\n", len, (int)p.traj);
        o += fmt("\t Trajectory time step (pico sec):                           %*.5f\n", len, p.dt);
        o += fmt("\t Temperature of the molecule system (Kelvin):               %*.2f\n", len, p.temp);
        o += fmt("\t Random number seed for 64-bit Mersenne-Twister:            %*lli\n", len, cmd.seed);
        o += fmt("\t Total charge of molecule (e):                              %*.3g\n", len, p.total_charge);
        if (p.charge)
            o += p.multipole ? (p.dispersion_cutoff ? "\t Multipole expansion and dispersion cut-off will be used.\n"
                                                    : "\t Multipole expansion and full dispersion interactions will be used.\n")
                             : (p.dispersion_cutoff ? "\t Exact electrostatics and dispersion cut-off will be used.\n"
                                                    : "\t Exact electrostatics and full dispersion interactions will be used.\n");
        else
            o += p.dispersion_cutoff ? "\t No-charges present: molecule: only dispersion cutoff interactions will be used.\n"
                                     : "\t No-charges present: only (full) dispersion interactions will be used.\n";
        o += "\n Program input and output:\n\t Program will print all outputs to:\n\t " + omega_path +
             "\n\t Molecule will be imported from:\n\t " + mol_path + "\n\n";
        log.put(o);
    }

    // setup: PCM clusters, boundary search, velocity window
    cosims_atoms a = {mf->n, mf->xyz, mf->mass, mf->sigma, mf->epsilon,
                      charges.empty() ? nullptr : charges.data(), mf->sigma_max};
    cosims_mol *m = nullptr;
    if (int rc = cosims_mol_create(&p, &a, &m)) {
        cosims_molfile_free(mf);
        return fail_engine(log, "molecule setup", rc);
    }
    cosims_mol_info info;
    cosims_mol_info_get(m, &info);

    // molecule2::printMolStats
    {
        std::string o = " Atom type breakdown:\n\t Name        Symbol    Mass    Sigma    Epsilon  IntMass  Count\n";
        for (int k = 0; k < mf->ntypes; k++) {
            const cosims_atom_type &t = mf->types[k];
            if (t.count)
                o += fmt("\t %-12s  %2s  %8.2f  %8.4f  %8.5f    %3i    %5i\n", t.name, t.symbol, t.mass, t.sigma,
                         t.epsilon, t.int_mass, t.count);
        }
        o += fmt("\t Total number of atoms:   %d\n", mf->n);
        o += fmt("\t Total mass of molecule:  %.5g a.u.\n", info.total_mass);
        o += fmt("\t Reduced mass:            %.5g a.u.\n", info.mu);
        log.put(o);
    }
    // impactParam::printBoundryResults: the fitted axes, their RMSD and the enlargement
    {
        std::string o = "\n Creating potential energy surface...\n\n Impact parameter properties:\n";
        o += "                      X         Y         Z\n -----------------------------------------------\n";
        o += fmt(" Axis Length    : %9.5G %9.5G %9.5G\n", info.fit_abc[0], info.fit_abc[1], info.fit_abc[2]);
        o += " -----------------------------------------------\n";
        o += fmt("\t RMSD: %10.5G\n", info.fit_rmsd);
        o += fmt(" Ellipsoid axes further expanded by %6.2f Ang.\n\n", info.enlarge);
        log.put(o);
    }
    // forceAccel: cluster formation summary
    if (info.ncl > 0) {
        std::vector<double> maxdist(info.ncl);
        std::vector<int32_t> rep(info.ncl);
        cosims_mol_clusters(m, nullptr, nullptr, nullptr, maxdist.data(), nullptr, nullptr, nullptr, nullptr,
                            rep.data());
        double rad = 0;
        long replaced = 0;
        for (int k = 0; k < info.ncl; k++) {
            rad += maxdist[k];
            replaced += rep[k] != 0;
        }
        std::string o = " Cluster formation summary:\n";
        if (p.multipole) o += fmt("\t Number of centers replaced\n\t by center of charge:          %6li\n", replaced);
        o += fmt("\t Number of clusters formed:    %6i\n", info.ncl);
        o += fmt("\t Average cluster size:         %6.2f atoms\n", (double)info.n / info.ncl);
        o += fmt("\t Average cluster radius:       %6.2f Ang.\n", rad / info.ncl);
        if (p.multipole) o += fmt("\t Multipole cut-off radius:     %6.1f Ang.\n", info.multipole_radius);
        if (p.dispersion_cutoff) o += fmt("\t Dispersion cut-off radius:    %6.1f Ang.\n", info.dispersion_radius);
        log.put(o + "\n");
    }
    // monteTraj: velocity distribution statistics
    log.put(fmt(" Velocity distribution statistics: \n\t Maximum velocity:  %6.3f ang/ps\n"
                "\t Minimum velocity:  %6.3f ang/ps\n\t Mean velocity:     %6.3f ang/ps\n"
                "\t Mode velocity:     %6.3f ang/ps\n\t Percent of velocity \n\t distribution used: %6.3f%%\n\n",
                info.gmax, info.gmin, info.gmean, info.gmode, 100 * info.cdf_frac));

    // trajectory phase: one engine call for every integral
    log.put("\n Starting Trajectory Calculations:\n _________________________________\n\n");
    const int iters = p.iter;
    const int64_t T = p.traj;
    const int64_t n = (int64_t)iters * T;
    std::vector<double> omega(std::max(iters, 1)), err(std::max(iters, 1)), chi(std::max<int64_t>(n, 1));
    std::vector<int32_t> status(std::max<int64_t>(n, 1));
    cosims_run_cfg cfg = {};
    cfg.prec = cmd.prec;
    cfg.rng = cmd.rng;
    cfg.lanes = cmd.lanes;
    cfg.chi = chi.data();
    cfg.status = status.data();
    cosims_ccs out;
    if (int rc = cosims_run(m, &p, &cfg, omega.data(), err.data(), &out)) {
        cosims_mol_destroy(m);
        cosims_molfile_free(mf);
        return fail_engine(log, "trajectory run", rc);
    }
    char dev[256] = "GPU";
    int32_t cus = 0;
    cosims_device_info(dev, sizeof dev, &cus);
    const double run_s = out.run_ms / 1000.0;

    // monteTraj::runThreads / updateCounts / dispResults as one thread runs
    // them (src/monteTraj.cpp:455-538): counters merge every print_rate
    // trajectories; the denominator is the merged count minus the failures
    // since the previous merge, and a miss keeps the previous attempt's
    // success flag.
    const int64_t rate = p.print_rate;
    const int width = T > 0 ? (int)std::log10((double)T) : 0;
    for (int k = 0; k < iters; k++) {
        std::string o = fmt("\n Computing CCS integral %i / %i on %s\n", k + 1, iters, dev);
        if (display && rate > 0) {
            double total = 0, total2 = 0, prev_total = 0, prev_total2 = 0, cur_total = 0, cur_total2 = 0;
            int64_t count = 0, prev_count = 0, orbit = 0, cur_count = 0, cur_succ = 0;
            bool succ = false;
            for (int64_t j = 0; j < T; j++) {
                const int32_t s = status[(int64_t)k * T + j];
                double c = 0;
                if (s != 0) {
                    succ = s == 1;
                    c = chi[(int64_t)k * T + j];
                }
                count++;
                if (!succ) orbit++;
                const double x = info.ccs_const * (1 - std::cos(std::fabs(c)));
                total += x;
                total2 += x * x;
                if (count % rate == 0) {
                    const int64_t prev_mod = cur_count % rate;
                    cur_count += count - prev_count;
                    cur_succ = cur_count - orbit;
                    cur_total += total - prev_total;
                    cur_total2 += total2 - prev_total2;
                    orbit = 0;
                    prev_total = total;
                    prev_total2 = total2;
                    prev_count = count;
                    if (cur_count % rate <= prev_mod) {
                        const double ns = (double)cur_succ;
                        const double e = std::sqrt((cur_total2 / ns - (cur_total / ns) * (cur_total / ns)) / ns);
                        o += fmt(" progress: Traj %*i of %i (%4.1f %%); CCS = %#9.5g +- %9.5g (%4.2f %%)\n", width,
                                 (int)cur_count, (int)T, (double)(100 * cur_count) / (double)T, cur_total / ns, e,
                                 e * 100 * ns / cur_total);
                    }
                }
            }
        }
        const double om = omega[k], e = err[k];
        const int w1 = e > 0 ? (int)std::log10(e) + 3 : 3;
        const int w2 = (e > 0 && om > 0) ? (int)std::log10(100 * e / om) + 3 : 3;
        o += fmt(" Number of completed trajectories: %i\n", (int)T);
        o += fmt(" Omega: %11.2f +/- %-*.2f (%*.2f %%)\n", om, w1, e, w2, 100 * e / om);
        const double tt = run_s * (k + 1) / iters;
        o += fmt(" Total time: %5.3f; estimated time remaining: %5.3f\n", tt, tt * (iters - k - 1) / (k + 1));
        log.put(o);
    }

    // final statistics (src/monteTraj.cpp:404-446)
    {
        std::string o = "\n\n ___________________________________\n";
        o += fmt("   Mean Cross-section:   %9.3f\n", out.mean);
        o += fmt("   Standard Dev:         %9.3f\n", out.sd);
        o += fmt("   StdDev Percent:       %9.3f\n\n", out.sd * 100 / out.mean);
        o += fmt("   Weighted CCS Mean:    %9.3f\n", out.wmean);
        o += fmt("   Weighted Error:       %9.3f\n", out.werr);
        o += fmt("   Weighted Pct. Error:  %9.4f\n", out.werr * 100 / out.wmean);
        o += fmt("   Max Energy Deviation: %9.4f\n", out.max_ergy);
        o += " ___________________________________\n";
        o += fmt(" Engine: %s precision, %s initial conditions, %d lane(s) per trajectory, workgroup %d, %s (%d CUs)\n",
                 cmd.prec == COSIMS_PREC_FP64 ? "FP64" : "mixed", cmd.rng == COSIMS_RNG_REFERENCE ? "reference" : "Philox",
                 out.lanes, out.wg_size, dev, cus);
        log.put(o);
    }
    const double total_s =
        std::chrono::duration<double>(std::chrono::steady_clock::now() - t_start).count();
    log.put(" Ending Program \n Current Time: " + now_text() + "\n" + fmt(" Total Time: %g\n", total_s));
    cosims_mol_destroy(m);
    cosims_molfile_free(mf);
    return 0;
}
