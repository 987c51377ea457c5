// prep.cpp -- host-side molecule setup (see prep.hpp).
//
// Restates, in the reference's arithmetic order:
//   molecule2 type table / readers      src/molecule2.cpp:43-192, 442-768
//   setMolProperties                    src/molecule2.cpp:770-903
//   setInertia(true)                    src/molecule2.cpp:1131-1391
//   VelocityDistribution + monteTraj    src/VelocityDistribution.cpp:46-151,
//                                       src/monteTraj.cpp:48-95
//   impactParam boundary/fit/enlarge    src/impactParam.cpp:602-736, 853-995
//   PCM_Cluster + eigen3D               src/PCM_Cluster.cpp:38-249, src/eigen3D.cpp:34-113
//   forceAccel::setClusters moments     src/forceAccel.cpp:53-210
#include "prep.hpp"

#include <algorithm>
#include <cctype>
#include <chrono>
#include <cfloat>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <atomic>
#include <random>
#include <stdlib.h>
#include <thread>

#include "cosims_engine.h"

namespace cosims {
namespace {

struct V3 {
    double x = 0, y = 0, z = 0;
    V3() = default;
    V3(double a, double b, double c) : x(a), y(b), z(c) {}
    V3 operator+(const V3 &o) const { return {x + o.x, y + o.y, z + o.z}; }
    V3 operator-(const V3 &o) const { return {x - o.x, y - o.y, z - o.z}; }
    V3 operator*(double s) const { return {x * s, y * s, z * s}; }
    V3 operator/(double s) const { return {x / s, y / s, z / s}; }
    double dot(const V3 &o) const { return (x * o.x) + (y * o.y) + (z * o.z); }
    double mag2() const { return (x * x) + (y * y) + (z * z); }
    double mag() const { return std::sqrt((x * x) + (y * y) + (z * z)); }
    double at(int i) const { return i == 0 ? x : (i == 1 ? y : z); }
};

V3 atom(const std::vector<double> &xyz, int i) { return {xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]}; }

std::string strip_ws(std::string s)
{
    s.erase(std::remove_if(s.begin(), s.end(), [](unsigned char c) { return std::isspace(c); }), s.end());
    return s;
}

bool all_digits(const std::string &s, bool allow_dot)
{
    for (char c : s)
        if (!std::isdigit((unsigned char)c) && !(allow_dot && c == '.')) return false;
    return true;
}

}  // namespace

// ---------------------------------------------------------------- types --
std::vector<TypeRow> default_types()
{
    return {
        {"Hydrogen", "H", 1.0079, 2.38, 0.01498936, 1},
        {"Carbon", "C", 12.010, 3.043, 0.03090114, 12},
        {"Nitrogen", "N", 14.006, 3.043, 0.03090114, 14},
        {"Oxygen", "O", 15.999, 3.043, 0.03090114, 16},
        {"Sodium", "Na", 22.98977, 3.043, 0.03090114, 23},
        {"Silicon", "Si", 28.086, 3.5, 0.03113175, 28},
        {"Phosphorus", "P", 30.97396, 3.5, 0.03113175, 31},
        {"Sulfur", "S", 32.07, 3.5, 0.03113175, 32},
        {"Iron", "Fe", 55.84, 3.5, 0.03113175, 56},
    };
}

bool apply_atoms_file(std::vector<TypeRow> &t, const std::string &path)
{
    std::ifstream f(path);
    if (!f) return false;
    std::string name, symbol, mass, sigma, eps, imass;
    while (f >> name >> symbol >> mass >> sigma >> eps >> imass) {
        std::transform(name.begin(), name.end(), name.begin(), ::tolower);
        if (name.at(0) == '#') continue;
        size_t idx = t.size();
        for (size_t i = 0; i < t.size(); i++)
            if (symbol == t[i].symbol || t[i].intmass == std::atoi(imass.c_str())) {
                idx = i;
                break;
            }
        if (idx == t.size()) t.push_back(TypeRow());
        t[idx] = {name, symbol, std::atof(mass.c_str()), std::atof(sigma.c_str()), std::atof(eps.c_str()),
                  std::atoi(imass.c_str())};
    }
    return true;
}

bool read_mfj(const std::string &path, const std::vector<TypeRow> &t, MolFile &out, std::string &err)
{
    std::ifstream f(path);
    if (!f) {
        err = "cannot open " + path;
        return false;
    }
    std::string line[6];
    for (auto &l : line) {
        std::getline(f, l);
        l = strip_ws(l);
    }
    out.name = line[0];
    int natoms = std::atoi(line[2].c_str());
    std::string unit = line[3];
    std::transform(unit.begin(), unit.end(), unit.begin(), ::tolower);
    bool uniform = false, calc = false;
    const std::string &ct = line[4];
    out.charged = false;
    out.total_charge = 0;
    if (ct.empty() || ct == "none") {
        out.charged = false;
    } else if (ct == "equal") {
        out.charged = true; out.total_charge = 1.0; uniform = true;
    } else if (all_digits(ct, false)) {
        out.charged = true; out.total_charge = std::atoi(ct.c_str()); uniform = true;
    } else if (ct == "calc") {
        calc = true; out.charged = true;
    }
    double scale = 1.0;
    if (!line[5].empty() && (all_digits(line[5], false) || all_digits(line[5], true)))
        scale = std::atof(line[5].c_str());
    double conv = (unit == "au") ? 0.529177210 : 1.0;
    std::string sx, sy, sz, sm, sq;
    double total = 0;
    while (f >> sx >> sy >> sz >> sm >> sq) {
        out.xyz.push_back(std::atof(sx.c_str()) * scale * conv);
        out.xyz.push_back(std::atof(sy.c_str()) * scale * conv);
        out.xyz.push_back(std::atof(sz.c_str()) * scale * conv);
        double c = 0.0;
        if (out.charged) {
            c = uniform ? out.total_charge / (double)natoms : std::atof(sq.c_str());
            total += c;
        }
        out.charge.push_back(c);
        int im = std::atoi(sm.c_str());
        std::string sym;
        for (const auto &r : t)
            if (r.intmass == im) {
                sym = r.symbol;
                break;
            }
        out.symbols.push_back(sym);
    }
    if (calc) out.total_charge = total;
    if (out.symbols.empty()) {
        err = "no atoms in " + path;
        return false;
    }
    return true;
}

bool read_pdb(const std::string &path, MolFile &out, std::string &err)
{
    std::ifstream f(path);
    if (!f) {
        err = "cannot open " + path;
        return false;
    }
    std::string raw;
    while (std::getline(f, raw)) {
        std::string l = raw.substr(0, 799);
        if (l.compare(0, 4, "ATOM") != 0 && l.compare(0, 6, "HETATM") != 0) continue;
        std::string at;
        size_t namelen = 1;
        if (l.size() >= 78) {
            at = l.substr(76, 2);
            if (at[0] == ' ' && at[1] == ' ') at = l.substr(12, 5);
            else namelen = 2;
        } else {
            at = l.size() > 12 ? l.substr(12, 5) : std::string();
        }
        std::string sym;
        for (char ch : at)
            if (std::isalpha((unsigned char)ch)) {
                sym.push_back(ch);
                if (sym.size() == namelen) break;
            }
        out.symbols.push_back(sym);
        auto field = [&](size_t s) { return l.size() > s ? std::atof(l.substr(s, 8).c_str()) : 0.0; };
        out.xyz.push_back(field(30));
        out.xyz.push_back(field(38));
        out.xyz.push_back(field(46));
        out.charge.push_back(0.0);
    }
    out.charged = false;
    if (out.symbols.empty()) {
        err = "no ATOM/HETATM records in " + path;
        return false;
    }
    return true;
}

// ------------------------------------------------------- velocity window --
namespace {
double fact(int x) { return x <= 1 ? 1 : x * fact(x - 1); }
double dfact(int x) { return (x == 1 || x == 0) ? 1 : x * dfact(x - 2); }

double vd_cdf(int n, double a, double x)
{
    double sum = 0;
    if (n % 2 == 0) {
        for (int k = 0; k <= (n / 2.0 - 1); k++) sum += std::pow(2 * x * x * a, k) / dfact(2 * k + 1);
        return std::erf(std::sqrt(a) * x) - std::sqrt(a / kPi) * std::exp(-a * x * x) * 2 * x * sum;
    }
    for (int k = 0; k <= n; k++) sum += std::pow(x * x * a, k) / fact(k);  // sums to n (reference)
    return 1 - std::exp(-a * x * x) * sum;
}
double vd_mean(int n, double a) { return std::tgamma((n + 2) / 2.0) / (std::tgamma((n + 1) / 2.0) * std::sqrt(a)); }
double vd_var(int n, double a)
{
    return ((n + 1) / 2.0 - std::pow(std::tgamma((n + 2) / 2.0) / std::tgamma((n + 1) / 2.0), 2)) / a;
}
}  // namespace

double veldist_pdf(const VelDist &v, double g) { return std::pow(g, 5) * std::exp(-v.a * g * g) * v.norm; }

VelDist velocity_distribution(double mu, double temp)
{
    const int n = 5;
    VelDist v;
    double beta = mu / (kBoltz * temp);
    v.a = beta / 2.0;
    v.norm = 2 * std::pow(v.a, (n + 1) / 2.0) / std::tgamma((n + 1) / 2.0);
    v.mode = std::sqrt(n / (2 * v.a));
    v.mean = vd_mean(n, v.a);
    v.sigma = std::sqrt(vd_var(n, v.a));
    v.gmin = v.mean - 2 * v.sigma;
    v.gmax = v.mean + 4 * v.sigma;
    v.cdf_frac = vd_cdf(n, v.a, v.gmax) - vd_cdf(n, v.a, v.gmin);
    v.norm = v.norm / (vd_cdf(n, v.a, v.gmax) - vd_cdf(n, v.a, v.gmin));
    v.max_val = veldist_pdf(v, v.mode);
    double c1 = std::pow(beta, 3) / (64 * kPi);
    c1 = c1 * 8 * kPi * kPi;
    v.const1 = c1 / v.norm;
    double m2 = vd_mean(2, beta / 2.0), s2 = std::sqrt(vd_var(2, beta / 2.0));
    v.min_kin = 0.5 * mu * std::pow(m2 - 2 * s2, 2);
    return v;
}

int force_type_of(bool charged, bool multipole, bool cutoff)
{
    if (charged) return multipole ? (cutoff ? 1 : 2) : (cutoff ? 3 : 4);
    return cutoff ? 5 : 6;
}

// ------------------------------------------------------------ potential --
double potential_host(const Prepared &m, const double P[3])
{
    double total = 0, sX = 0, sY = 0, sZ = 0;
    for (int n = 0; n < m.n; n++) {
        double xD = P[0] - m.xyz[3 * n], yD = P[1] - m.xyz[3 * n + 1], zD = P[2] - m.xyz[3 * n + 2];
        double r2 = xD * xD + yD * yD + zD * zD;
        double r6 = r2 * r2 * r2;
        double t2 = m.sig6[n] / r6;
        double t1 = t2 * t2;
        total += m.eps4[n] * (t1 - t2);
        if (m.charged) {
            double r_3 = m.q[n] / (r2 * std::sqrt(r2));
            sX += xD * r_3; sY += yD * r_3; sZ += zD * r_3;
        }
    }
    if (m.charged) total -= m.dipole_scalar * 0.5 * (sX * sX + sY * sY + sZ * sZ);
    return total;
}

namespace {

// ------------------------------------------------------------- inertia --
void eig_vals_inertia(double A[3][3], double ev[3])
{
    double a0 = A[0][0] * A[1][1] * A[2][2] - A[0][0] * A[1][2] * A[2][1] - A[1][1] * A[0][2] * A[2][0] -
                A[2][2] * A[0][1] * A[1][0] + A[0][1] * A[1][2] * A[2][0] + A[0][2] * A[2][1] * A[1][0];
    double a1 = A[0][1] * A[1][0] + A[0][2] * A[2][0] + A[1][2] * A[2][1] - A[0][0] * A[1][1] -
                A[1][1] * A[2][2] - A[2][2] * A[0][0];
    double a2 = A[0][0] + A[1][1] + A[2][2];
    a0 = -a0; a1 = -a1; a2 = -a2;
    double p = (3 * a1 - a2 * a2) / 3;
    double q = (9 * a1 * a2 - 27 * a0 - 2 * a2 * a2 * a2) / 27;
    double Q = p / 3, R = q / 2, D = Q * Q * Q + R * R, a, b;
    if (D < 0) { a = R; b = std::sqrt(-D); }
    else { a = R + std::sqrt(D); b = 0; }
    double ph = std::atan2(b, a), mg = std::pow(a * a + b * b, 1.0 / 6.0);
    ev[0] = -(a2 / 3) + mg * 2 * std::cos(ph / 3.0);
    ev[1] = -(a2 / 3) - mg * (std::cos(ph / 3.0) + std::sqrt(3.0) * std::sin(ph / 3.0));
    ev[2] = -(a2 / 3) - mg * (std::cos(ph / 3.0) - std::sqrt(3.0) * std::sin(ph / 3.0));
}

void eig_vecs_inertia(double A[3][3], const double l[3], double E[3][3])
{
    double det = l[0] * l[0] - l[0] * A[1][1] - A[1][2] * A[1][2] - l[0] * A[2][2] + A[1][1] * A[2][2];
    double y = (l[0] * A[1][0] - A[2][2] * A[1][0] + A[1][2] * A[2][0]) / det;
    double z = (A[2][1] * A[1][0] + l[0] * A[2][0] - A[1][1] * A[2][0]) / det;
    double nf = std::sqrt(1 + y * y + z * z);
    E[0][0] = 1.0 / nf; E[0][1] = y / nf; E[0][2] = z / nf;
    det = l[1] * l[1] - l[1] * A[0][0] - A[0][2] * A[0][2] - l[1] * A[2][2] + A[0][0] * A[2][2];
    double x = (l[1] * A[0][1] + A[0][2] * A[2][1] - A[0][1] * A[2][2]) / det;
    z = (A[0][1] * A[0][2] + (l[1] - A[0][0]) * A[2][1]) / det;
    nf = std::sqrt(1 + x * x + z * z);
    E[1][0] = x / nf; E[1][1] = 1 / nf; E[1][2] = z / nf;
    E[2][0] = E[0][1] * E[1][2] - E[0][2] * E[1][1];
    E[2][1] = E[0][2] * E[1][0] - E[0][0] * E[1][2];
    E[2][2] = E[0][0] * E[1][1] - E[0][1] * E[1][0];
}

int set_inertia(Prepared &m)
{
    const int n = m.n;
    double I[3][3] = {{0}}, ev[3], E[3][3];
    for (int i = 0; i < n; i++) {
        const double *c = &m.xyz[3 * i];
        double r2 = c[0] * c[0] + c[1] * c[1] + c[2] * c[2];
        for (int a = 0; a < 3; a++)
            for (int b = 0; b < 3; b++) I[a][b] += m.mass[i] * (r2 * (a == b ? 1.0 : 0.0) - c[a] * c[b]);
    }
    eig_vals_inertia(I, ev);
    eig_vecs_inertia(I, ev, E);
    auto rowdot = [&](int i, int j) { return E[i][0] * E[j][0] + E[i][1] * E[j][1] + E[i][2] * E[j][2]; };
    double dp = std::fabs(rowdot(0, 1));
    dp = std::max(std::fabs(rowdot(0, 2)), dp);
    dp = std::max(std::fabs(rowdot(1, 2)), dp);
    if (dp > 1E-6) return COSIMS_E_UNSUPPORTED;  // reference fallback reads past the atom array
    for (int a = 0; a < 3; a++)
        for (int b = 0; b < 3; b++) m.rot[3 * a + b] = E[a][b];
    for (int i = 0; i < n; i++) {
        double x = m.xyz[3 * i], y = m.xyz[3 * i + 1], z = m.xyz[3 * i + 2];
        m.xyz[3 * i] = x * E[0][0] + y * E[0][1] + z * E[0][2];
        m.xyz[3 * i + 1] = x * E[1][0] + y * E[1][1] + z * E[1][2];
        m.xyz[3 * i + 2] = x * E[2][0] + y * E[2][1] + z * E[2][2];
    }
    double mx = 0, my = 0, mz = 0;
    for (int i = 0; i < n; i++) {
        if (std::fabs(m.xyz[3 * i]) > std::fabs(mx)) mx = std::fabs(m.xyz[3 * i]);
        if (std::fabs(m.xyz[3 * i + 1]) > std::fabs(my)) my = std::fabs(m.xyz[3 * i + 1]);
        if (std::fabs(m.xyz[3 * i + 2]) > std::fabs(mz)) mz = std::fabs(m.xyz[3 * i + 2]);
    }
    if (mx > my && mx > mz) {
        for (int i = 0; i < n; i++) {
            double t = m.xyz[3 * i];
            m.xyz[3 * i] = -m.xyz[3 * i + 2];
            m.xyz[3 * i + 2] = t;
        }
    } else if (my > mx && my > mz) {
        for (int i = 0; i < n; i++) {
            double t = m.xyz[3 * i + 1];
            m.xyz[3 * i + 1] = m.xyz[3 * i + 2];
            m.xyz[3 * i + 2] = -t;
        }
    }
    mx = my = mz = 0;
    for (int i = 0; i < n; i++) {
        if (std::fabs(m.xyz[3 * i]) > std::fabs(mx)) mx = m.xyz[3 * i];
        if (std::fabs(m.xyz[3 * i + 1]) > std::fabs(my)) my = m.xyz[3 * i + 1];
        if (std::fabs(m.xyz[3 * i + 2]) > std::fabs(mz)) mz = m.xyz[3 * i + 2];
    }
    m.max_xyz[0] = mx; m.max_xyz[1] = my; m.max_xyz[2] = mz;
    return 0;
}

// ------------------------------------------------------------ boundary --
// One march on the host, the loop of setBoundry (src/impactParam.cpp:666-690)
// and of its z-axis search (:616-631).
MarchHit march_host(const Prepared &m, double mk, const March &w, int64_t max_points)
{
    MarchHit h{};
    V3 p(w.start[0], w.start[1], w.start[2]);
    const double e0 = potential_host(m, w.start);
    const double sl = (e0 + mk < 0) ? -1.0 : 1.0;
    const V3 v = V3(w.dir[0], w.dir[1], w.dir[2]) * sl;
    V3 pv(w.prev0[0], w.prev0[1], w.prev0[2]);
    if (w.first) p = p + v;
    const double sx = w.start[0], sy = w.start[1];
    for (int64_t k = w.first; k < max_points; k++) {
        const double q[3] = {p.x, p.y, p.z};
        const double en = potential_host(m, q);
        if ((en + mk) * sl < 0) {
            h.found = 1;
            h.p[0] = p.x; h.p[1] = p.y; h.p[2] = p.z;
            h.prev[0] = pv.x; h.prev[1] = pv.y; h.prev[2] = pv.z;
            return h;
        }
        if (w.axis && ((sx == 0 && sy / p.y <= 0) || (sy == 0 && sx / p.x <= 0) || (sx / p.x <= 0 && sy / p.y <= 0))) {
            h.found = 0;
            return h;
        }
        pv = p;
        p = p + v;
    }
    h.found = -1;
    return h;
}

void march_all(const Prepared &m, double mk, const std::vector<March> &w, std::vector<MarchHit> &out, int threads)
{
    out.assign(w.size(), MarchHit{});
    int nthr = threads > 0 ? threads : (int)std::max(1u, std::thread::hardware_concurrency());
    nthr = std::max(1, std::min<int>(nthr, (int)w.size()));
    std::vector<std::thread> pool;
    std::atomic_int next{0};
    for (int t = 0; t < nthr; t++)
        pool.emplace_back([&]() {
            for (int k; (k = next.fetch_add(1)) < (int)w.size();) out[k] = march_host(m, mk, w[k], kMaxMarchPoints);
        });
    for (auto &th : pool) th.join();
}

}  // namespace

RayGeometry ray_geometry(const BoundaryPlan &bp, const MarchHit &zhit)
{
    RayGeometry g;
    double zMax = std::max(zhit.p[2], zhit.prev[2]);
    zMax *= 1.3;
    g.zmax = zMax;
    g.posmag = bp.posmag0 + std::fabs(zMax - bp.maxz);
    g.levels = (int)std::ceil((double)std::max(25, (int)std::ceil(zMax)));
    g.zinc = 2 * zMax / (g.levels);
    return g;
}

March ray_march(const BoundaryPlan &bp, const RayGeometry &g, int level, int n)
{
    March w;
    w.start[0] = g.posmag * bp.cos_t[n];
    w.start[1] = g.posmag * bp.sin_t[n];
    w.start[2] = -g.zmax + level * g.zinc;
    w.dir[0] = bp.dir_x[n]; w.dir[1] = bp.dir_y[n]; w.dir[2] = 0.0;
    w.prev0[0] = w.prev0[1] = w.prev0[2] = 0;
    w.first = 0;
    w.axis = 1;
    return w;
}

namespace {

// Boundary points of impactParam::setBoundry (src/impactParam.cpp:602-736):
// a march down the z axis in 0.5 A steps finds zMax; then (levels+1) x 20
// rays at heights -zMax..zMax march towards the z axis in 0.1 A steps, each
// keeping the point of the crossing farther from the centre.
int boundary_points(const Prepared &m, double min_kin, int threads, BoundaryMarcher *bm, std::vector<V3> &out,
                    std::string &err)
{
    const double velMag = 0.1;
    const double mk = 0.01 * min_kin;
    BoundaryPlan bp;
    bp.z0 = std::fabs(m.max_xyz[2]) + 20;
    bp.posmag0 = std::fabs(std::max(std::fabs(m.max_xyz[0]), std::fabs(m.max_xyz[1])));
    bp.maxz = std::fabs(m.max_xyz[2]);
    for (int n = 0; n < kBoundaryRot; n++) {
        double theta = n * 2 * kPi / ((double)kBoundaryRot);
        bp.cos_t[n] = std::cos(theta);
        bp.sin_t[n] = std::sin(theta);
        bp.dir_x[n] = -velMag * std::cos(theta);
        bp.dir_y[n] = -velMag * std::sin(theta);
    }
    MarchHit zh{};
    std::vector<MarchHit> hits;
    if (bm) {
        if (int rc = bm->boundary(m, mk, bp, zh, hits, err)) return rc;
    } else {
        // z axis: pos.z -= 0.5 * slope, i.e. the step (0, 0, -0.5) * slope
        March zw{{0, 0, bp.z0}, {0, 0, -0.5}, {0, 0, bp.z0}, 1, 0};
        zh = march_host(m, mk, zw, zwalk_points(bp.z0));
        if (zh.found == 1) {
            const RayGeometry g = ray_geometry(bp, zh);
            std::vector<March> rays((size_t)(g.levels + 1) * kBoundaryRot);
            for (int i = 0; i <= g.levels; i++)
                for (int n = 0; n < kBoundaryRot; n++) rays[(size_t)i * kBoundaryRot + n] = ray_march(bp, g, i, n);
            march_all(m, mk, rays, hits, threads);
        }
    }
    if (zh.found != 1) {
        err = "boundary search: no potential crossing on the z axis";
        return COSIMS_E_SETUP;
    }
    out.clear();
    for (const MarchHit &h : hits) {
        if (h.found < 0) {
            err = "boundary search: a ray found no potential crossing";
            return COSIMS_E_SETUP;
        }
        if (h.found == 0) continue;
        const V3 p(h.p[0], h.p[1], h.p[2]), pv(h.prev[0], h.prev[1], h.prev[2]);
        const V3 q = (pv.mag() > p.mag()) ? pv : p;
        if (!(q.x == 0 && q.y == 0 && q.z == 0)) out.push_back(q);
    }
    return 0;
}

void fit_ellipsoid(const std::vector<V3> &P, Prepared &m)
{
    double mu = 0, nu = 0, ep = 0, de = 0, si = 0, rho = 0, al = 0, be = 0, ga = 0;
    for (const V3 &p : P) {
        double x2 = p.x * p.x, y2 = p.y * p.y, z2 = p.z * p.z;
        double x4 = x2 * x2, y4 = y2 * y2, z4 = z2 * z2;
        mu = mu + x4; nu = nu + y4; ep = ep + z4;
        de = de + x2 * y2; si = si + x2 * z2; rho = rho + y2 * z2;
        al = al + x2; be = be + y2; ga = ga + z2;
    }
    double det = ep * mu * nu - ep * de * de - mu * rho * rho + 2 * de * rho * si - nu * si * si;
    double A = (al * (nu * ep - rho * rho) + de * (rho * ga - be * ep) + si * (be * rho - nu * ga)) / det;
    double B = (mu * (be * ep - ga * rho) + al * (rho * si - de * ep) + si * (de * ga - be * si)) / det;
    double C = (mu * (nu * ga - be * rho) + de * (be * si - de * ga) + al * (de * rho - nu * si)) / det;
    m.fit_a = 1 / std::sqrt(A);
    m.fit_b = 1 / std::sqrt(B);
    m.fit_c = 1 / std::sqrt(C);
    size_t np = P.size();
    m.fit_rmsd = std::sqrt((A * A * mu + B * B * nu + C * C * ep +
                            2 * (A * B * de + B * C * rho + A * C * si - A * al - B * be - C * ga) + np) /
                           (np - 1));
}

void enlarge(const std::vector<V3> &P, Prepared &m)
{
    double aOld = m.fit_a, a = m.fit_a, b = m.fit_b, c = m.fit_c;
    const int np = (int)P.size();
    bool accept = false;
    do {
        for (int i = 0; i < np + 1; i++) {
            if (i == np || a > 3 * aOld) {
                accept = true;
                break;
            }
            double mag = P[i].x * P[i].x / (a * a) + P[i].y * P[i].y / (b * b) + P[i].z * P[i].z / (c * c);
            if (mag > 1) {
                a = a + 0.01; b = b + 0.01; c = c + 0.01;
                break;
            }
        }
    } while (!accept);
    m.a = a; m.b = b; m.c = c;
    m.enlarge = a - aOld;
}

double ms_since(std::chrono::steady_clock::time_point t0)
{
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

// ---------------------------------------------------------------- PCM --
void eig_vals_pcm(double A[3][3], double ev[3])
{
    double a0 = A[0][0] * A[1][1] * A[2][2] - A[0][0] * A[1][2] * A[2][1] - A[1][1] * A[0][2] * A[2][0] -
                A[2][2] * A[0][1] * A[1][0] + A[0][1] * A[1][2] * A[2][0] + A[0][2] * A[2][1] * A[1][0];
    double a1 = A[0][1] * A[1][0] + A[0][2] * A[2][0] + A[1][2] * A[2][1] - A[0][0] * A[1][1] -
                A[1][1] * A[2][2] - A[2][2] * A[0][0];
    double a2 = A[0][0] + A[1][1] + A[2][2];
    a0 = -a0; a1 = -a1; a2 = -a2;
    double p = (3 * a1 - a2 * a2) / 3;
    double q = (9 * a1 * a2 - 27 * a0 - 2 * a2 * a2 * a2) / 27;
    double Q = p / 3, R = q / 2, D = Q * Q * Q + R * R, a, b = 0;  // b: unset by src/eigen3D.cpp:65-68 when D >= 0
    if (D < 0) { a = R; b = std::sqrt(-D); }
    else a = R + std::sqrt(D);
    double ph = std::atan2(b, a), mg = std::pow(a * a + b * b, 1.0 / 6.0);
    ev[0] = -(a2 / 3) + mg * 2 * std::cos(ph / 3.0);
    ev[1] = -(a2 / 3) - mg * (std::cos(ph / 3.0) + std::sqrt(3.0) * std::sin(ph / 3.0));
    ev[2] = -(a2 / 3) - mg * (std::cos(ph / 3.0) - std::sqrt(3.0) * std::sin(ph / 3.0));
}

void eig_vecs_pcm(double A[3][3], const double l[3], double E[3][3])
{
    for (int n = 0; n < 3; n++) {
        double det = l[n] * A[0][2] - A[0][2] * A[1][1] + A[0][1] * A[1][2];
        double y = (A[0][2] * A[1][0] + l[n] * A[1][2] - A[0][0] * A[1][2]) / det;
        double z = (l[n] * l[n] - l[n] * (A[0][0] + A[1][1]) - A[0][1] * A[1][0] + A[0][0] * A[1][1]) / det;
        double nf = std::sqrt(1 + y * y + z * z);
        E[n][0] = 1.0 / nf; E[n][1] = y / nf; E[n][2] = z / nf;
    }
}

// Geometric centre and largest centre distance of one cluster
// (src/PCM_Cluster.cpp:180-205): the first member only seeds minDist.
void cluster_prop(const std::vector<double> &xyz, const std::vector<int> &G, V3 &center, double &maxd)
{
    V3 c;
    for (int idx : G) c = c + atom(xyz, idx);
    c = c / (double)G.size();
    center = c;
    V3 md;
    bool first = true;
    for (int idx : G) {
        V3 d = c - atom(xyz, idx);
        double dm = d.mag();
        if (first) first = false;
        else if (dm > md.mag()) md = d;
    }
    maxd = md.mag();
}

// One PCM split of cluster G (src/PCM_Cluster.cpp:38-176): covariance about
// the mean, eigenvalues and eigenvectors by eigen3D's closed forms, members
// on either side of the plane normal to the dominant eigenvector (angle to it
// at most pi/2: c1), both halves' centres and extents.  ok = false when a
// half is empty.
struct Split {
    std::vector<int> c1, c2;
    V3 cen1, cen2;
    double md1 = 0, md2 = 0;
    bool ok = true;
};

void split_cluster(const std::vector<double> &xyz, const std::vector<int> &G, Split &S)
{
    double mean[3] = {0, 0, 0}, cov[3][3], ev[3], E[3][3];
    for (int idx : G)
        for (int j = 0; j < 3; j++) mean[j] += xyz[3 * idx + j];
    for (int j = 0; j < 3; j++) mean[j] = mean[j] / ((double)G.size());
    for (int a = 0; a < 3; a++)
        for (int b = a; b < 3; b++) {
            double s = 0;
            for (int idx : G) s += (xyz[3 * idx + a] - mean[a]) * (xyz[3 * idx + b] - mean[b]);
            cov[a][b] = s;
            if (a != b) cov[b][a] = s;
        }
    eig_vals_pcm(cov, ev);
    eig_vecs_pcm(cov, ev, E);
    int emax = 0;
    for (int k = 0; k < 3; k++)
        if (std::fabs(ev[k]) > ev[emax]) emax = k;
    V3 c;
    for (int idx : G) c = c + atom(xyz, idx);
    c = c / ((double)G.size());
    S.c1.clear();
    S.c2.clear();
    for (int idx : G) {
        V3 d = atom(xyz, idx) - c;
        double t1 = d.x * E[emax][0] + d.y * E[emax][1] + d.z * E[emax][2];
        double t2 = d.mag() * std::sqrt(E[emax][0] * E[emax][0] + E[emax][1] * E[emax][1] + E[emax][2] * E[emax][2]);
        double t3 = t1 / t2, ang;
        if (t3 > 1) ang = 0;
        else if (t3 < -1) ang = kPi;
        else ang = std::acos(t1 / t2);
        (ang <= kPi2 ? S.c1 : S.c2).push_back(idx);
    }
    S.ok = !S.c1.empty() && !S.c2.empty();
    if (!S.ok) return;
    cluster_prop(xyz, S.c1, S.cen1, S.md1);
    cluster_prop(xyz, S.c2, S.cen2, S.md2);
}

// f(i) for i in [0, n) on up to `threads` host threads (0: all), taking
// indices from a shared counter.
template <class F>
void parallel_for(int n, int threads, F &&f)
{
    const int hw = (int)std::max(1u, std::thread::hardware_concurrency());
    const int nt = std::max(1, std::min(threads > 0 ? threads : hw, n));
    if (nt == 1) {
        for (int i = 0; i < n; i++) f(i);
        return;
    }
    std::atomic<int> next{0};
    auto work = [&] {
        for (int i; (i = next.fetch_add(1)) < n;) f(i);
    };
    std::vector<std::thread> th;
    for (int t = 1; t < nt; t++) th.emplace_back(work);
    work();
    for (auto &t : th) t.join();
}

// PCM_Cluster::clusterAtoms (src/PCM_Cluster.cpp:38-249): passes over the
// cluster list split every cluster wider than mcs (its first half in place,
// its second appended, in list order) until none is; then orderClusters'
// selection sort on centre z.  The splits of one pass are independent and
// run on host threads; the list is assembled in the reference's order, so
// membership and order do not depend on the thread count.  Clusters keep
// the centre and extent computed when they were formed (the reference
// recomputes them every pass, with the same result).
int form_clusters(Prepared &m, int mcs, int threads, std::string &err)
{
    std::vector<std::vector<int>> g(1);
    for (int i = 0; i < m.n; i++) g[0].push_back(i);
    std::vector<V3> center(1);
    std::vector<double> maxd(1);
    cluster_prop(m.xyz, g[0], center[0], maxd[0]);
    const int thr = m.n >= 2048 ? threads : 1;  // small molecules: thread start-up outweighs the work
    std::vector<int> wide;
    std::vector<Split> sp;
    for (;;) {
        wide.clear();
        for (size_t i = 0; i < g.size(); i++)
            if (maxd[i] > mcs) wide.push_back((int)i);
        if (wide.empty()) break;
        sp.assign(wide.size(), Split());
        parallel_for((int)wide.size(), thr, [&](int k) { split_cluster(m.xyz, g[wide[k]], sp[k]); });
        for (size_t k = 0; k < wide.size(); k++) {
            Split &S = sp[k];
            if (!S.ok) {
                err = "PCM split produced an empty cluster (the reference stalls here)";
                return COSIMS_E_UNSUPPORTED;
            }
            const int i = wide[k];
            g[i].swap(S.c1);
            center[i] = S.cen1;
            maxd[i] = S.md1;
            g.push_back(std::move(S.c2));
            center.push_back(S.cen2);
            maxd.push_back(S.md2);
        }
    }
    const int ng = (int)g.size();
    // orderClusters: selection sort on centre z (first minimum, swapped
    // forward).  With all centre z distinct its result is the one ascending
    // order, which an index sort gives in O(ng log ng); ties keep the
    // selection sort, whose order of equal keys depends on its swaps.
    std::vector<int> ord(ng);
    for (int k = 0; k < ng; k++) ord[k] = k;
    std::sort(ord.begin(), ord.end(), [&](int a, int b) { return center[a].z < center[b].z; });
    bool distinct = true;
    for (int k = 1; k < ng && distinct; k++) distinct = center[ord[k - 1]].z < center[ord[k]].z;
    if (!distinct) {
        for (int k = 0; k < ng; k++) ord[k] = k;
        for (int k = 0; k < ng - 1; k++) {
            int mi = k;
            double mv = center[ord[k]].z;
            for (int i = k + 1; i < ng; i++)
                if (center[ord[i]].z < mv) {
                    mv = center[ord[i]].z;
                    mi = i;
                }
            std::swap(ord[k], ord[mi]);
        }
    }
    m.ncl = ng;
    m.cl_start.assign(1, 0);
    m.cl_atoms.clear();
    m.cl_center.clear();
    m.cl_maxdist.clear();
    for (int k = 0; k < ng; k++) {
        const int i = ord[k];
        m.cl_atoms.insert(m.cl_atoms.end(), g[i].begin(), g[i].end());
        m.cl_start.push_back((int32_t)m.cl_atoms.size());
        m.cl_center.push_back(center[i].x);
        m.cl_center.push_back(center[i].y);
        m.cl_center.push_back(center[i].z);
        m.cl_maxdist.push_back(maxd[i]);
    }
    return 0;
}

void cluster_moments(Prepared &m)
{
    const int ncl = m.ncl;
    m.cl_q.assign(ncl, 0);
    m.cl_dip.assign(3 * ncl, 0);
    m.cl_quad.assign(9 * ncl, 0);
    m.cl_pos.assign(3 * ncl, 0);
    m.cl_replace.assign(ncl, 0);
    for (int c = 0; c < ncl; c++) {
        V3 cc(m.cl_center[3 * c], m.cl_center[3 * c + 1], m.cl_center[3 * c + 2]);
        double sumq = 0, dip[3] = {0, 0, 0}, Q[9] = {0};
        for (int k = m.cl_start[c]; k < m.cl_start[c + 1]; k++) {
            int idx = m.cl_atoms[k];
            double charge = m.q[idx];
            sumq += charge;
            double p[3] = {m.xyz[3 * idx] - cc.x, m.xyz[3 * idx + 1] - cc.y, m.xyz[3 * idx + 2] - cc.z};
            double m2 = p[0] * p[0] + p[1] * p[1] + p[2] * p[2];
            for (int i = 0; i < 3; i++) {
                dip[i] += charge * p[i];
                for (int j = 0; j < 3; j++) Q[3 * i + j] += charge * (3 * p[i] * p[j] - m2 * (i == j ? 1.0 : 0.0));
            }
        }
        m.cl_q[c] = sumq;
        std::memcpy(&m.cl_dip[3 * c], dip, sizeof(dip));
        std::memcpy(&m.cl_quad[9 * c], Q, sizeof(Q));
        V3 coc;
        for (int k = m.cl_start[c]; k < m.cl_start[c + 1]; k++) {
            int idx = m.cl_atoms[k];
            coc = coc + atom(m.xyz, idx) * m.q[idx];
        }
        coc = coc / sumq;
        double pd = 100 * std::fabs((coc.mag() - cc.mag()) / coc.mag());
        V3 use = (pd <= 5) ? coc : cc;
        m.cl_replace[c] = (pd <= 5) ? 1 : 0;
        m.cl_pos[3 * c] = use.x; m.cl_pos[3 * c + 1] = use.y; m.cl_pos[3 * c + 2] = use.z;
    }
}

}  // namespace

int complete_prepared(const Params &p, Prepared &m, std::string &err)
{
    (void)err;
    m.dipole_scalar = kAlphaKe / m.mu;
    m.charged = p.charge;
    m.force_type = force_type_of(p.charge, p.multipole, p.dispersion_cutoff);
    m.multipole_order = p.multipole_order;
    m.projection = p.projection;
    m.vd = velocity_distribution(m.mu, p.temp);
    m.bmax = std::max(m.a, std::max(m.b, m.c));
    m.ccs_const = m.vd.const1 * (m.bmax * m.bmax);
    const auto t0 = std::chrono::steady_clock::now();
    if (m.ncl > 0) cluster_moments(m);
    m.moments_ms = ms_since(t0);
    return 0;
}

int prepare(const Params &p, int n, const double *xyz, const double *mass, const double *sigma, const double *eps,
            const double *charge, double sigma_max, Prepared &m, std::string &err, BoundaryMarcher *bm)
{
    if (n <= 0 || !xyz || !mass || !sigma || !eps) {
        err = "empty molecule";
        return COSIMS_E_INVALID;
    }
    m = Prepared();
    m.n = n;
    m.xyz.assign(xyz, xyz + 3 * n);
    m.mass.assign(mass, mass + n);
    m.q.assign(n, 0.0);
    if (p.charge && charge) m.q.assign(charge, charge + n);
    m.sig6.resize(n);
    m.eps24m.resize(n);
    m.eps4.resize(n);
    // setMolProperties
    const double conv = 4.184 / (6.02214199e+23 * 1.67262158e-27 * 10.0);
    double cx = 0, cy = 0, cz = 0, tm = 0;
    for (int i = 0; i < n; i++) {
        double e = eps[i];
        e *= conv;
        m.sig6[i] = std::pow(sigma[i], 6);
        m.eps24m[i] = e * 24;
        m.eps4[i] = e * 4;
        cx += mass[i] * m.xyz[3 * i];
        cy += mass[i] * m.xyz[3 * i + 1];
        cz += mass[i] * m.xyz[3 * i + 2];
        tm += mass[i];
    }
    m.total_mass = tm;
    m.mu = (tm * kHeMass) / (tm + kHeMass);
    cx /= tm; cy /= tm; cz /= tm;
    for (int i = 1; i < n; i++) m.eps24m[i] /= m.mu;  // atom 0 keeps 24 eps (src/molecule2.cpp:862)
    for (int i = 0; i < n; i++) {
        m.xyz[3 * i] -= cx; m.xyz[3 * i + 1] -= cy; m.xyz[3 * i + 2] -= cz;
    }
    m.dispersion_radius = p.dispersion_radius;
    m.multipole_radius = p.multipole_radius;
    if (p.dispersion_cutoff && m.dispersion_radius == 0) m.dispersion_radius = sigma_max * std::pow(150.0 / 0.5, 1.0 / 3.0);
    if (p.multipole && m.multipole_radius == 0) m.multipole_radius = 20.0;
    int rc = set_inertia(m);
    if (rc) {
        err = "inertia eigenvectors not orthogonal: the reference's fallback branch is undefined behaviour";
        return rc;
    }
    m.charged = p.charge;
    m.dipole_scalar = kAlphaKe / m.mu;
    m.vd = velocity_distribution(m.mu, p.temp);
    std::vector<V3> pts;
    auto t0 = std::chrono::steady_clock::now();
    rc = boundary_points(m, m.vd.min_kin, p.threads, bm, pts, err);
    m.boundary_ms = ms_since(t0);
    if (rc) return rc;
    if (pts.size() < 4) {
        err = "potential boundary search found too few points";
        return COSIMS_E_SETUP;
    }
    m.n_boundary = (int)pts.size();
    fit_ellipsoid(pts, m);
    enlarge(pts, m);
    if (p.multipole || p.dispersion_cutoff) {
        t0 = std::chrono::steady_clock::now();
        rc = form_clusters(m, p.max_cluster_size, p.threads, err);
        m.cluster_ms = ms_since(t0);
        if (rc) return rc;
    }
    return complete_prepared(p, m, err);
}

void sample_compat(const Prepared &P, uint64_t seed, int iters, int64_t traj, double *p0, double *v0, double *g,
                   int32_t *hit)
{
    // rand() of a fresh process: glibc's additive generator in its default
    // 128-byte state seeded with 1, kept private here instead of touching
    // the caller's global rand() state.
    struct random_data rd;
    std::memset(&rd, 0, sizeof(rd));
    static thread_local char state[128];
    initstate_r(1, state, sizeof(state), &rd);
    const double gMin = P.vd.gmin, gMax = P.vd.gmax, maxVal = P.vd.max_val, bMax = P.bmax;
    const double maxDist = (float)bMax;  // monteTraj::maxDist is a float (src/monteTraj.h:67)
    unsigned restart = 0;
    for (int it = 0; it < iters; it++) {
        std::mt19937_64 gen((unsigned long long)(long long)seed);
        gen.discard(it == 0 ? 0ULL : (unsigned long long)restart);
        std::uniform_real_distribution<double> distr(0.0, 1.0);
        int count = 0;
        for (int64_t n = 0; n < traj; n++) {
            const int64_t o = (int64_t)it * traj + n;
            double gg, phi, theta, gamma, b;
            for (;;) {
                count += 8;
                gg = distr(gen) * (gMax - gMin) + gMin;
                int32_t r = 0;
                random_r(&rd, &r);
                double gCheck = ((double)r / (double)RAND_MAX) * maxVal;
                phi = distr(gen) * 2.0 * kPi;
                theta = distr(gen) * kPi;
                double thetaCheck = 0.5 * distr(gen);
                gamma = distr(gen) * 2.0 * kPi;
                b = distr(gen) * bMax;
                double bCheck = distr(gen) * (2 / bMax);
                if (thetaCheck <= std::sin(theta) / 2 && bCheck <= 2 * b / (bMax * bMax) &&
                    gCheck <= veldist_pdf(P.vd, gg))
                    break;
            }
            double X = b * std::cos(gamma) * std::cos(phi) * std::cos(theta) - b * std::sin(gamma) * std::sin(phi) +
                       maxDist * std::cos(phi) * std::sin(theta);
            double Y = b * std::cos(gamma) * std::sin(phi) * std::cos(theta) + b * std::sin(gamma) * std::cos(phi) +
                       maxDist * std::sin(phi) * std::sin(theta);
            double Z = maxDist * std::cos(theta) - b * std::cos(gamma) * std::sin(theta);
            double VX = -gg * std::cos(phi) * std::sin(theta), VY = -gg * std::sin(phi) * std::sin(theta),
                   VZ = -gg * std::cos(theta);
            bool h = true;
            if (P.projection) {
                V3 v(VX / P.a, VY / P.b, VZ / P.c), p(X / P.a, Y / P.b, Z / P.c);
                V3 cr((p.y * v.z) - (p.z * v.y), (p.z * v.x) - (p.x * v.z), (p.x * v.y) - (p.y * v.x));
                double disc = -cr.dot(cr) + v.dot(v);
                if (disc < 0) {
                    h = false;
                } else {
                    double pp = -(p.dot(v) / (v.dot(v))) - std::sqrt(disc) / v.dot(v);
                    X = X + VX * pp; Y = Y + VY * pp; Z = Z + VZ * pp;
                }
            }
            p0[3 * o] = X; p0[3 * o + 1] = Y; p0[3 * o + 2] = Z;
            v0[3 * o] = VX; v0[3 * o + 1] = VY; v0[3 * o + 2] = VZ;
            g[o] = gg;
            hit[o] = h ? 1 : 0;
        }
        restart = (unsigned)count;
    }
}

}  // namespace cosims
