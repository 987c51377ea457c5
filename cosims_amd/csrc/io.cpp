// io.cpp -- input-file parameters and molecule files (C ABI).
//
//   cosims_params_read    fileParams::importParams  src/fileParams.cpp:261-510
//   cosims_molfile_read   molecule2::assignDefaultFF / importFile /
//                         importMFJ / importPDB     src/molecule2.cpp:43-192, 442-768
#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>

#include "cosims_engine.h"
#include "prep.hpp"

namespace {

bool is_int(const std::string &s)
{
    for (char c : s)
        if (!std::isdigit((unsigned char)c)) return false;
    return true;
}

// fileParams::getBool (src/fileParams.cpp:62-76): integers via noboolalpha,
// otherwise "true"/"false" (case-insensitive); anything else reads as false.
bool get_bool(const std::string &v)
{
    std::string s = v;
    std::transform(s.begin(), s.end(), s.begin(), ::tolower);
    if (is_int(v)) return std::atoi(v.c_str()) != 0;
    return s == "true";
}

}  // namespace

extern "C" {

int cosims_params_read(const char *path, cosims_params *p)
{
    if (!path || !p) return cosims::set_last_error(COSIMS_E_INVALID, "cosims_params_read: null argument");
    std::ifstream f(path);
    if (!f) return cosims::set_last_error(COSIMS_E_IO, std::string("cannot open ") + path);
    bool print_rate_set = false;
    std::string line;
    while (std::getline(f, line)) {
        if (line.empty()) continue;
        std::istringstream ss(line);
        std::string opt, val;
        ss >> opt >> val;
        if (val.empty() || opt.empty() || opt[0] == '#') continue;
        const char *v = val.c_str();
        if (opt == "name") std::snprintf(p->name, sizeof(p->name), "%s", v);
        else if (opt == "dt") p->dt = std::atof(v);
        else if (opt == "temp") p->temp = std::atof(v);
        else if (opt == "traj") p->traj = (int64_t)(int)(std::atof(v) * 1000);
        else if (opt == "iter") p->iter = (int)std::atof(v);
        else if (opt == "threads") p->threads = (int)std::atof(v);
        else if (opt == "seed") p->seed = (uint64_t)std::atoll(v);
        else if (opt == "proj") p->projection = (int)std::atof(v) != 0;
        else if (opt == "dispersion_cutoff") p->dispersion_cutoff = get_bool(val);
        else if (opt == "dispersion_radius") p->dispersion_radius = std::atof(v);
        else if (opt == "multipole_radius") p->multipole_radius = std::atof(v);
        else if (opt == "multipole_order") p->multipole_order = std::atoi(v);
        else if (opt == "print_rate") { p->print_rate = (int)std::atof(v); print_rate_set = true; }
        else if (opt == "multipole") p->multipole = get_bool(val);
        else if (opt == "max_cluster_size") p->max_cluster_size = std::atoi(v);
        else if (opt == "charge") {
            p->total_charge = std::atof(v);
            p->charge = p->total_charge != 0;
        }
        // dir, gmax, gmin, print_* are accepted and have no effect on results
    }
    if (!print_rate_set) p->print_rate = (int)(p->traj / 5.0);
    else if (p->print_rate > p->traj || p->print_rate == 0) p->print_rate = (int)p->traj;
    return COSIMS_OK;
}

int cosims_molfile_read(const char *mol_path, const char *atoms_path, cosims_molfile **out)
{
    if (!mol_path || !out) return cosims::set_last_error(COSIMS_E_INVALID, "cosims_molfile_read: null argument");
    std::vector<cosims::TypeRow> table = cosims::default_types();
    if (atoms_path && *atoms_path && !cosims::apply_atoms_file(table, atoms_path))
        return cosims::set_last_error(COSIMS_E_IO, std::string("cannot open ") + atoms_path);
    std::string path(mol_path), err;
    if (path.size() < 4 || path[path.size() - 4] != '.')
        return cosims::set_last_error(COSIMS_E_INVALID, "molecule file needs a .pdb or .mfj extension");
    std::string ext = path.substr(path.size() - 3);
    std::transform(ext.begin(), ext.end(), ext.begin(), ::tolower);
    cosims::MolFile mf;
    bool ok;
    if (ext == "pdb") ok = cosims::read_pdb(path, mf, err);
    else if (ext == "mfj") ok = cosims::read_mfj(path, table, mf, err);
    else return cosims::set_last_error(COSIMS_E_INVALID, "molecule file needs a .pdb or .mfj extension");
    if (!ok) return cosims::set_last_error(COSIMS_E_IO, err);
    const int n = (int)mf.symbols.size();
    auto *f = new cosims_molfile();
    std::memset(f, 0, sizeof(*f));
    f->n = n;
    f->xyz = new double[3 * n];
    f->mass = new double[n];
    f->sigma = new double[n];
    f->epsilon = new double[n];
    f->charge = new double[n];
    std::memcpy(f->xyz, mf.xyz.data(), sizeof(double) * 3 * n);
    for (int i = 0; i < n; i++) {
        const cosims::TypeRow *row = nullptr;
        for (const auto &r : table)
            if (r.symbol == mf.symbols[i]) {
                row = &r;
                break;
            }
        if (!row) {  // the reference continues with uninitialised parameters (src/molecule2.cpp:826-835)
            cosims_molfile_free(f);
            return cosims::set_last_error(COSIMS_E_UNSUPPORTED, "atom " + std::to_string(i + 1) +
                                                                    ": type not in the force field");
        }
        f->mass[i] = row->mass;
        f->sigma[i] = row->sigma;
        f->epsilon[i] = row->eps;
        f->charge[i] = mf.charge[i];
    }
    f->charged = mf.charged ? 1 : 0;
    f->total_charge = mf.total_charge;
    double smax = 0;
    for (const auto &r : table) smax = std::max(smax, r.sigma);
    f->sigma_max = smax;
    std::snprintf(f->name, sizeof(f->name), "%s", mf.name.c_str());
    // fieldData rows with their atom counts (src/molecule2.cpp:820, 920-927)
    f->ntypes = (int32_t)table.size();
    f->types = new cosims_atom_type[table.size()];
    for (size_t k = 0; k < table.size(); k++) {
        cosims_atom_type &t = f->types[k];
        std::memset(&t, 0, sizeof(t));
        std::snprintf(t.name, sizeof(t.name), "%s", table[k].name.c_str());
        std::snprintf(t.symbol, sizeof(t.symbol), "%s", table[k].symbol.c_str());
        t.mass = table[k].mass;
        t.sigma = table[k].sigma;
        t.epsilon = table[k].eps;
        t.int_mass = table[k].intmass;
    }
    for (const auto &s : mf.symbols)  // first matching row, as the assignment loop
        for (size_t k = 0; k < table.size(); k++)
            if (s == table[k].symbol) {
                f->types[k].count++;
                break;
            }
    *out = f;
    return COSIMS_OK;
}

void cosims_molfile_free(cosims_molfile *f)
{
    if (!f) return;
    delete[] f->xyz;
    delete[] f->mass;
    delete[] f->sigma;
    delete[] f->epsilon;
    delete[] f->charge;
    delete[] f->types;
    delete f;
}

}  // extern "C"
