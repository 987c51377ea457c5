/*
 * cosims_oracle.c -- FP64 CPU restatement of the CoSIMS He trajectory path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker).  Loaded by tests/, smoke() and the
 * CPU-baseline leg of bench.py; never by the engine in cosims_amd/.
 *
 * Each routine cites the reference file:line it restates.  Expressions keep
 * the reference's operand order because the x86-64 reference build performs
 * plain sequential IEEE double arithmetic (no FMA, no reassociation); this
 * file is compiled with -ffp-contract=off for the same reason.
 */
#include "cosims_oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define OR_PI 3.14159265358979323846       /* src/vector3D.h:250 */
#define OR_PI_2 1.57079632679489661923     /* src/vector3D.h:251 */
#define OR_K 0.831446214563                /* src/maxBoltz.h:29 */
#define OR_HE 4.002602                     /* src/molecule2.h:72 */

static double mag3(const double v[3]) { return sqrt((v[0] * v[0]) + (v[1] * v[1]) + (v[2] * v[2])); }
static double mag2_3(const double v[3]) { return (v[0] * v[0]) + (v[1] * v[1]) + (v[2] * v[2]); }
static double dot3(const double u[3], const double v[3]) { return (u[0] * v[0]) + (u[1] * v[1]) + (u[2] * v[2]); }

/* vector3D::Angle(vector3D&)  src/vector3D.h:134-143 */
static double angle_vec(const double u[3], const double v[3])
{
    double term1 = dot3(u, v);
    double term2 = mag3(u) * mag3(v);
    double term3 = term1 / term2;
    if (term3 > 1) return 0;
    else if (term3 < -1) return OR_PI;
    return acos(term1 / term2);
}

/* ====================================================================== */
/* molecule2::setMolProperties  src/molecule2.cpp:770-865                  */
/* ====================================================================== */
void or_mol_properties(int n, double *xyz, const double *mass,
                       const double *sigma, const double *eps,
                       double *sig6, double *eps24m, double *eps4,
                       double *mu, double *total_mass)
{
    double cx = 0, cy = 0, cz = 0, tm = 0;
    const double conv = 4.184 / (6.02214199e+23 * 1.67262158e-27 * 10.0); /* :800 */
    for (int i = 0; i < n; i++) {
        double epsAvg = eps[i];
        epsAvg *= conv;
        sig6[i] = pow(sigma[i], 6);                 /* :801 */
        eps24m[i] = epsAvg * 24;                    /* :803 */
        eps4[i] = epsAvg * 4;                       /* :805 */
        cx += mass[i] * xyz[3 * i + 0];             /* :851-854 */
        cy += mass[i] * xyz[3 * i + 1];
        cz += mass[i] * xyz[3 * i + 2];
        tm += mass[i];
    }
    double m = (tm * OR_HE) / (tm + OR_HE);         /* :859 */
    cx /= tm; cy /= tm; cz /= tm;                   /* :861 */
    for (int i = 1; i < n; i++) eps24m[i] /= m;     /* :862-863 (quirk: from 1) */
    for (int i = 0; i < n; i++) {                   /* setCenter :194-209 */
        xyz[3 * i + 0] -= cx;
        xyz[3 * i + 1] -= cy;
        xyz[3 * i + 2] -= cz;
    }
    *mu = m;
    *total_mass = tm;
}

/* src/molecule2.cpp:867-895 */
double or_default_dispersion_radius(double sigma_max)
{
    return sigma_max * pow(150.0 / 0.5, 1.0 / 3.0);
}

/* molecule2::getEigenvalues  src/molecule2.cpp:1315-1358 */
static void eig_values_mol(double A[3][3], double ev[3])
{
    double a0, a1, a2, p, q, R, Q, D, a, b, phase, mag;
    a0 = A[0][0] * A[1][1] * A[2][2] - A[0][0] * A[1][2] * A[2][1] - A[1][1] * A[0][2] * A[2][0] -
         A[2][2] * A[0][1] * A[1][0] + A[0][1] * A[1][2] * A[2][0] + A[0][2] * A[2][1] * A[1][0];
    a1 = A[0][1] * A[1][0] + A[0][2] * A[2][0] + A[1][2] * A[2][1] - A[0][0] * A[1][1] - A[1][1] * A[2][2] -
         A[2][2] * A[0][0];
    a2 = A[0][0] + A[1][1] + A[2][2];
    a0 = -a0; a1 = -a1; a2 = -a2;
    p = (3 * a1 - a2 * a2) / 3;
    q = (9 * a1 * a2 - 27 * a0 - 2 * a2 * a2 * a2) / 27;
    Q = p / 3; R = q / 2; D = Q * Q * Q + R * R;
    if (D < 0) { a = R; b = sqrt(-D); }
    else { a = R + sqrt(D); b = 0; }
    phase = atan2(b, a);
    mag = pow(a * a + b * b, 1.0 / 6.0);
    ev[0] = -(a2 / 3) + mag * 2 * cos(phase / 3.0);
    ev[1] = -(a2 / 3) - mag * (cos(phase / 3.0) + sqrt(3.0) * sin(phase / 3.0));
    ev[2] = -(a2 / 3) - mag * (cos(phase / 3.0) - sqrt(3.0) * sin(phase / 3.0));
}

/* molecule2::getEigenvectors  src/molecule2.cpp:1360-1391 */
static void eig_vectors_mol(double A[3][3], const double l[3], double E[3][3])
{
    double det, x, y, z, nf;
    det = l[0] * l[0] - l[0] * A[1][1] - A[1][2] * A[1][2] - l[0] * A[2][2] + A[1][1] * A[2][2];
    y = (l[0] * A[1][0] - A[2][2] * A[1][0] + A[1][2] * A[2][0]) / det;
    z = (A[2][1] * A[1][0] + l[0] * A[2][0] - A[1][1] * A[2][0]) / det;
    nf = sqrt(1 + y * y + z * z);
    E[0][0] = 1.0 / nf; E[0][1] = y / nf; E[0][2] = z / nf;
    det = l[1] * l[1] - l[1] * A[0][0] - A[0][2] * A[0][2] - l[1] * A[2][2] + A[0][0] * A[2][2];
    x = (l[1] * A[0][1] + A[0][2] * A[2][1] - A[0][1] * A[2][2]) / det;
    z = (A[0][1] * A[0][2] + (l[1] - A[0][0]) * A[2][1]) / det;
    nf = sqrt(1 + x * x + z * z);
    E[1][0] = x / nf; E[1][1] = 1 / nf; E[1][2] = z / nf;
    E[2][0] = E[0][1] * E[1][2] - E[0][2] * E[1][1];
    E[2][1] = E[0][2] * E[1][0] - E[0][0] * E[1][2];
    E[2][2] = E[0][0] * E[1][1] - E[0][1] * E[1][0];
}

static double dmax(double a, double b) { return (a < b) ? b : a; } /* std::max */

/* molecule2::setInertia(true)  src/molecule2.cpp:1131-1308 */
int or_inertia(int n, double *xyz, const double *mass, double rot[9], double maxxyz[3])
{
    double I[3][3] = {{0}}, ev[3], E[3][3];
    for (int i = 0; i < n; i++) {
        const double *c = xyz + 3 * i;
        double r2 = c[0] * c[0] + c[1] * c[1] + c[2] * c[2];
        for (int a = 0; a < 3; a++)
            for (int b = 0; b < 3; b++)
                I[a][b] += mass[i] * (r2 * (a == b ? 1.0 : 0.0) - c[a] * c[b]);
    }
    eig_values_mol(I, ev);
    eig_vectors_mol(I, ev, E);
    double dp = fabs(dot3(E[0], E[1]));
    dp = dmax(fabs(dot3(E[0], E[2])), dp);
    dp = dmax(fabs(dot3(E[1], E[2])), dp);
    if (dp > 1E-6)
        return -1; /* :1181-1228 reads data[numAtoms] (out of bounds); not restated */
    for (int a = 0; a < 3; a++)
        for (int b = 0; b < 3; b++) rot[3 * a + b] = E[a][b];
    for (int i = 0; i < n; i++) { /* rotate2 src/vector3D.h:250-260 */
        double x = xyz[3 * i], y = xyz[3 * i + 1], z = xyz[3 * i + 2];
        xyz[3 * i + 0] = x * E[0][0] + y * E[0][1] + z * E[0][2];
        xyz[3 * i + 1] = x * E[1][0] + y * E[1][1] + z * E[1][2];
        xyz[3 * i + 2] = x * E[2][0] + y * E[2][1] + z * E[2][2];
    }
    double mx = 0, my = 0, mz = 0;
    for (int i = 0; i < n; i++) {
        if (fabs(xyz[3 * i]) > fabs(mx)) mx = fabs(xyz[3 * i]);
        if (fabs(xyz[3 * i + 1]) > fabs(my)) my = fabs(xyz[3 * i + 1]);
        if (fabs(xyz[3 * i + 2]) > fabs(mz)) mz = fabs(xyz[3 * i + 2]);
    }
    if (mx > my && mx > mz) {
        for (int i = 0; i < n; i++) {
            double t = xyz[3 * i];
            xyz[3 * i] = -xyz[3 * i + 2];
            xyz[3 * i + 2] = t;
        }
    } else if (my > mx && my > mz) {
        for (int i = 0; i < n; i++) {
            double t = xyz[3 * i + 1];
            xyz[3 * i + 1] = xyz[3 * i + 2];
            xyz[3 * i + 2] = -t;
        }
    }
    mx = 0; my = 0; mz = 0;
    for (int i = 0; i < n; i++) {
        if (fabs(xyz[3 * i]) > fabs(mx)) mx = xyz[3 * i];
        if (fabs(xyz[3 * i + 1]) > fabs(my)) my = xyz[3 * i + 1];
        if (fabs(xyz[3 * i + 2]) > fabs(mz)) mz = xyz[3 * i + 2];
    }
    maxxyz[0] = mx; maxxyz[1] = my; maxxyz[2] = mz;
    return 0;
}

/* ====================================================================== */
/* VelocityDistribution  src/VelocityDistribution.cpp:46-151,             */
/* used by monteTraj::monteTraj  src/monteTraj.cpp:48-57,69,91-93         */
/* ====================================================================== */
static double vd_fact(int x) { return x <= 1 ? 1 : x * vd_fact(x - 1); }
static double vd_dfact(int x) { return (x == 1 || x == 0) ? 1 : x * vd_dfact(x - 2); }

static double vd_cdf(int power, double param, double x) /* :128-146 */
{
    double sum = 0;
    if (power % 2 == 0) {
        for (int k = 0; k <= (power / 2.0 - 1); k++)
            sum += pow(2 * x * x * param, k) / vd_dfact(2 * k + 1);
        return erf(sqrt(param) * x) - sqrt(param / OR_PI) * exp(-param * x * x) * 2 * x * sum;
    }
    for (int k = 0; k <= power; k++) /* quirk: sums to power */
        sum += pow(x * x * param, k) / vd_fact(k);
    return 1 - exp(-param * x * x) * sum;
}
static double vd_mean(int p, double a) { return tgamma((p + 2) / 2.0) / (tgamma((p + 1) / 2.0) * sqrt(a)); }
static double vd_var(int p, double a)
{
    return ((p + 1) / 2.0 - pow(tgamma((p + 2) / 2.0) / tgamma((p + 1) / 2.0), 2)) / a;
}

void or_veldist_init(double mu, double temp, or_veldist *vd)
{
    const int n = 5;
    double beta = mu / (OR_K * temp);               /* src/monteTraj.cpp:49 */
    double a = beta / 2.0;
    vd->a = a;
    vd->norm = 2 * pow(a, (n + 1) / 2.0) / tgamma((n + 1) / 2.0); /* :68-71 */
    vd->mode = sqrt(n / (2 * a));
    vd->mean = vd_mean(n, a);
    vd->sigma = sqrt(vd_var(n, a));
    vd->gmin = vd->mean - 2 * vd->sigma;            /* src/monteTraj.cpp:54-55 */
    vd->gmax = vd->mean + 4 * vd->sigma;
    vd->norm = vd->norm / (vd_cdf(n, a, vd->gmax) - vd_cdf(n, a, vd->gmin)); /* :63-66 */
    vd->max_val = or_veldist_pdf(vd, vd->mode);
    vd->cdf_frac = vd_cdf(n, a, vd->gmax) - vd_cdf(n, a, vd->gmin);
    double c1 = pow(beta, 3) / (64 * OR_PI);        /* src/monteTraj.cpp:91-93 */
    c1 = c1 * 8 * OR_PI * OR_PI;
    c1 = c1 / vd->norm;
    vd->ccs_scale = c1;
    double m2 = tgamma((2 + 2) / 2.0) / (tgamma((2 + 1) / 2.0) * sqrt(beta / 2.0));
    double s2 = sqrt(vd_var(2, beta / 2.0));
    vd->min_kin = 0.5 * mu * pow(m2 - 2 * s2, 2); /* src/monteTraj.cpp:69 */
}

double or_veldist_pdf(const or_veldist *vd, double g) /* :148-151 */
{
    return pow(g, 5) * exp(-vd->a * g * g) * vd->norm;
}

/* ====================================================================== */
/* forceAccel  src/forceAccel.cpp                                         */
/* ====================================================================== */

/* chargeDipoleLJ (type 4)  src/forceAccel.cpp:751-805 */
static void acc_type4(const or_mol *m, const double *P, double *out)
{
    double sumXX = 0.0, sumYY = 0.0, sumZZ = 0.0, sumXY = 0.0, sumXZ = 0.0, sumYZ = 0.0;
    double dipX = 0.0, dipY = 0.0, dipZ = 0.0, tx = 0, ty = 0, tz = 0;
    for (int n = 0; n < m->n; n++) {
        double xD = P[0] - m->x[n], yD = P[1] - m->y[n], zD = P[2] - m->z[n];
        double x2 = xD * xD, y2 = yD * yD, z2 = zD * zD;
        double r2 = x2 + y2 + z2;
        double r = sqrt(r2);
        double r3 = r2 * r;
        double r_3 = 1.0 / r3;
        double r_5 = r_3 / r2;
        double r6 = r2 * r2 * r2;
        r_3 *= m->q[n];
        r_5 *= m->q[n];
        double term2 = m->sig6[n] / r6;
        double term1 = 2 * term2 * term2;
        double sc = m->eps24m[n] * (term1 - term2) / r2;
        tx += sc * xD; ty += sc * yD; tz += sc * zD;
        dipX += xD * r_3; dipY += yD * r_3; dipZ += zD * r_3;
        sumXX += (3.0 * x2 - r2) * r_5;
        sumYY += (3.0 * y2 - r2) * r_5;
        sumZZ += (3.0 * z2 - r2) * r_5;
        sumXY += 3 * xD * yD * r_5;
        sumXZ += 3 * xD * zD * r_5;
        sumYZ += 3 * yD * zD * r_5;
    }
    tx -= m->dipole_scalar * (sumXX * dipX + sumXY * dipY + sumXZ * dipZ);
    ty -= m->dipole_scalar * (sumXY * dipX + sumYY * dipY + sumYZ * dipZ);
    tz -= m->dipole_scalar * (sumXZ * dipX + sumYZ * dipY + sumZZ * dipZ);
    out[0] = tx; out[1] = ty; out[2] = tz;
}

/* dispersionFull (type 6)  src/forceAccel.cpp:856-884 */
static void acc_type6(const or_mol *m, const double *P, double *out)
{
    double tx = 0, ty = 0, tz = 0;
    for (int n = 0; n < m->n; n++) {
        double xD = P[0] - m->x[n], yD = P[1] - m->y[n], zD = P[2] - m->z[n];
        double r2 = xD * xD + yD * yD + zD * zD;
        double r6 = r2 * r2 * r2;
        double term2 = m->sig6[n] / r6;
        double term1 = 2 * term2 * term2;
        double sc = m->eps24m[n] * (term1 - term2) / r2;
        tx += sc * xD; ty += sc * yD; tz += sc * zD;
    }
    out[0] = tx; out[1] = ty; out[2] = tz;
}

/* accumulators shared by the clustered variants */
typedef struct {
    double fX, fY, fZ, sXY, sXZ, sYZ, aXY, aXZ, aYZ, dXX, dYY, dZZ, sumX, sumY, sumZ;
} cl_acc;

/* exact charge (+ optional LJ) over a cluster's atoms,
 * src/forceAccel.cpp:366-403 (with LJ) and :408-436 (without) */
static void cl_atoms_exact(const or_mol *m, int c, const double *P, int with_lj, cl_acc *A)
{
    for (int k = m->cl_start[c]; k < m->cl_start[c + 1]; k++) {
        int idx = m->cl_atoms[k];
        double charge = m->q[idx];
        double d0 = P[0] - m->x[idx], d1 = P[1] - m->y[idx], d2 = P[2] - m->z[idx];
        double r2 = d0 * d0 + d1 * d1 + d2 * d2;
        double r_5 = 1.0 / (r2 * r2 * sqrt(r2));
        double r_3 = r_5 * r2;
        r_3 *= charge;
        r_5 *= charge;
        A->fX += d0 * r_3; A->fY += d1 * r_3; A->fZ += d2 * r_3;
        A->dXX += (r2 - 3 * d0 * d0) * r_5;
        A->dYY += (r2 - 3 * d1 * d1) * r_5;
        A->dZZ += (r2 - 3 * d2 * d2) * r_5;
        A->sXY -= 3 * d0 * d1 * r_5;
        A->sXZ -= 3 * d0 * d2 * r_5;
        A->sYZ -= 3 * d1 * d2 * r_5;
        if (with_lj) {
            double r6 = r2 * r2 * r2;
            double term2 = m->sig6[idx] / r6;
            double term1 = 2 * term2 * term2;
            double sc = m->eps24m[idx] * (term1 - term2) / r2;
            A->sumX += sc * d0; A->sumY += sc * d1; A->sumZ += sc * d2;
        }
    }
}

/* LJ over a cluster's atoms with parameters of atom `pidx`,
 * src/forceAccel.cpp:341-360 (pidx = cluster index, quirk) and :563-582 */
static void cl_atoms_lj(const or_mol *m, int c, const double *P, int quirk, cl_acc *A)
{
    for (int k = m->cl_start[c]; k < m->cl_start[c + 1]; k++) {
        int idx = m->cl_atoms[k];
        int pidx = quirk ? c : idx;
        double d0 = P[0] - m->x[idx], d1 = P[1] - m->y[idx], d2 = P[2] - m->z[idx];
        double r2 = d0 * d0 + d1 * d1 + d2 * d2;
        double r6 = r2 * r2 * r2;
        double term2 = m->sig6[pidx] / r6;
        double term1 = 2 * term2 * term2;
        double sc = m->eps24m[pidx] * (term1 - term2) / r2;
        A->sumX += sc * d0; A->sumY += sc * d1; A->sumZ += sc * d2;
    }
}

/* far-field cluster header, src/forceAccel.cpp:258-337 (== :485-561) */
static void cl_far(const or_mol *m, int n, const double dx[3], double r2, cl_acc *A)
{
    const int order = m->multipole_order;
    const int rep = m->cl_replace[n];
    const double *dip = m->cl_dip + 3 * n;
    const double *Q = m->cl_quad + 9 * n;
    double xx = dx[0] * dx[0], yy = dx[1] * dx[1], zz = dx[2] * dx[2];
    double r_2 = 1.0 / r2;
    double r_1 = 1.0 / sqrt(r2);
    double r_3 = r_2 * r_1;
    double r_5 = r_3 * r_2;
    double r_7 = 0, const1 = 0, const2 = 0, tM[3] = {0, 0, 0};
    double xy = dx[0] * dx[1], xz = dx[0] * dx[2], yz = dx[1] * dx[2];
    double cR5 = m->cl_q[n] * r_5;
    double cR3 = cR5 * r2;
    if (order >= 1 && !rep) {
        r_7 = r_5 * r_2;
        const1 = 3 * (dip[0] * dx[0] + dip[1] * dx[1] + dip[2] * dx[2]);
    }
    if (order >= 2) {
        r_7 = r_5 * r_2;
        const2 = 0;
        tM[0] = (Q[0] * dx[0] + Q[3] * dx[1] + Q[6] * dx[2]) * r_7;
        tM[1] = (Q[1] * dx[0] + Q[4] * dx[1] + Q[7] * dx[2]) * r_7;
        tM[2] = (Q[2] * dx[0] + Q[5] * dx[1] + Q[8] * dx[2]) * r_7;
        const2 = 5 * (tM[0] * dx[0] + tM[1] * dx[1] + tM[2] * dx[2]) * r_2 * 0.5;
    }
    A->fX += cR3 * dx[0]; A->fY += cR3 * dx[1]; A->fZ += cR3 * dx[2];
    A->dXX += (r2 - 3 * xx) * cR5;
    A->dYY += (r2 - 3 * yy) * cR5;
    A->dZZ += (r2 - 3 * zz) * cR5;
    A->sXY -= 3 * xy * cR5;
    A->sXZ -= 3 * xz * cR5;
    A->sYZ -= 3 * yz * cR5;
    if (order >= 1 && !rep) {
        A->fX += (const1 * dx[0] - r2 * dip[0]) * r_5;
        A->fY += (const1 * dx[1] - r2 * dip[1]) * r_5;
        A->fZ += (const1 * dx[2] - r2 * dip[2]) * r_5;
        A->dXX += (r2 - 5 * xx) * r_7 * const1;
        A->dYY += (r2 - 5 * yy) * r_7 * const1;
        A->dZZ += (r2 - 5 * zz) * r_7 * const1;
        A->sXY -= 5 * xy * r_7 * const1;
        A->sXZ -= 5 * xz * r_7 * const1;
        A->sYZ -= 5 * yz * r_7 * const1;
        A->aXY += 3 * (dip[1] * dx[0] - dip[0] * dx[1]) * r_5;
        A->aXZ += 3 * (dip[2] * dx[0] - dip[0] * dx[2]) * r_5;
        A->aYZ += 3 * (dip[2] * dx[1] - dip[1] * dx[2]) * r_5;
    }
    if (order >= 2) {
        A->fX += (const2 * dx[0] - tM[0]) * r2;
        A->fY += (const2 * dx[1] - tM[1]) * r2;
        A->fZ += (const2 * dx[2] - tM[2]) * r2;
        A->dXX += ((r2 - 7 * xx) * const2 - Q[0] * r_5);
        A->dYY += ((r2 - 7 * yy) * const2 - Q[4] * r_5);
        A->dZZ += ((r2 - 7 * zz) * const2 - Q[8] * r_5);
        A->sXY -= (7 * xy * const2 + Q[1] * r_5);
        A->sXZ -= (7 * xz * const2 + Q[2] * r_5);
        A->sYZ -= (7 * yz * const2 + Q[5] * r_5);
        A->aXY += 5 * (tM[1] * dx[0] - tM[0] * dx[1]);
        A->aXZ += 5 * (tM[2] * dx[0] - tM[0] * dx[2]);
        A->aYZ += 5 * (tM[2] * dx[1] - tM[1] * dx[2]);
    }
}

/* types 1 and 2: src/forceAccel.cpp:221-445 and :447-633 */
static void acc_multipole(const or_mol *m, const double *P, int cutoff, double *out)
{
    cl_acc A;
    memset(&A, 0, sizeof(A));
    for (int n = 0; n < m->ncl; n++) {
        double dx[3];
        dx[0] = P[0] - m->cl_pos[3 * n + 0];
        dx[1] = P[1] - m->cl_pos[3 * n + 1];
        dx[2] = P[2] - m->cl_pos[3 * n + 2];
        double xx = dx[0] * dx[0], yy = dx[1] * dx[1], zz = dx[2] * dx[2];
        double r2 = xx + yy + zz;
        double mpr = m->multipole_radius + m->cl_rad[n];
        mpr *= mpr;
        double dpr = m->dispersion_radius + m->cl_rad[n];
        dpr *= dpr;
        if (r2 > mpr) {
            cl_far(m, n, dx, r2, &A);
            if (cutoff) {
                if (!(r2 > dpr)) cl_atoms_lj(m, n, P, 1, &A);   /* :339-361 */
            } else {
                cl_atoms_lj(m, n, P, 0, &A);                    /* :563-582 */
            }
        } else if (!cutoff || !(r2 > dpr)) {
            cl_atoms_exact(m, n, P, 1, &A);                     /* :364-404, :585-625 */
        } else {
            cl_atoms_exact(m, n, P, 0, &A);                     /* :406-437 */
        }
    }
    double ds = m->dipole_scalar;
    A.sumX += ds * (A.dXX * A.fX + (A.sXY + A.aXY) * A.fY + (A.sXZ + A.aXZ) * A.fZ);
    A.sumY += ds * ((A.sXY - A.aXY) * A.fX + A.dYY * A.fY + (A.sYZ + A.aYZ) * A.fZ);
    A.sumZ += ds * ((A.sXZ - A.aXZ) * A.fX + (A.sYZ - A.aYZ) * A.fY + A.dZZ * A.fZ);
    out[0] = A.sumX; out[1] = A.sumY; out[2] = A.sumZ;
}

/* dispersionOnly (type 3)  src/forceAccel.cpp:635-749 */
static void acc_type3(const or_mol *m, const double *P, double *out)
{
    cl_acc A;
    memset(&A, 0, sizeof(A));
    for (int n = 0; n < m->ncl; n++) {
        double dx0 = P[0] - m->cl_pos[3 * n + 0];
        double dx1 = P[1] - m->cl_pos[3 * n + 1];
        double dx2 = P[2] - m->cl_pos[3 * n + 2];
        double r2 = dx0 * dx0 + dx1 * dx1 + dx2 * dx2;
        double dpr = m->dispersion_radius + m->cl_rad[n];
        dpr *= dpr;
        cl_atoms_exact(m, n, P, !(r2 > dpr), &A);
    }
    double ds = m->dipole_scalar;
    A.sumX += ds * (A.dXX * A.fX + A.sXY * A.fY + A.sXZ * A.fZ);
    A.sumY += ds * (A.sXY * A.fX + A.dYY * A.fY + A.sYZ * A.fZ);
    A.sumZ += ds * (A.sXZ * A.fX + A.sYZ * A.fY + A.dZZ * A.fZ);
    out[0] = A.sumX; out[1] = A.sumY; out[2] = A.sumZ;
}

/* dispersionNoCharge (type 5)  src/forceAccel.cpp:807-854 */
static void acc_type5(const or_mol *m, const double *P, double *out)
{
    cl_acc A;
    memset(&A, 0, sizeof(A));
    for (int n = 0; n < m->ncl; n++) {
        double dx0 = P[0] - m->cl_pos[3 * n + 0];
        double dx1 = P[1] - m->cl_pos[3 * n + 1];
        double dx2 = P[2] - m->cl_pos[3 * n + 2];
        double r2 = dx0 * dx0 + dx1 * dx1 + dx2 * dx2;
        double dpr = m->dispersion_radius + m->cl_rad[n];
        dpr *= dpr;
        if (!(r2 > dpr)) cl_atoms_lj(m, n, P, 0, &A);
    }
    out[0] = A.sumX; out[1] = A.sumY; out[2] = A.sumZ;
}

/* getReducedVec  src/forceAccel.cpp:887-908 */
void or_accel(const or_mol *m, const double pos[3], double acc[3])
{
    switch (m->force_type) {
    case 1: acc_multipole(m, pos, 1, acc); break;
    case 2: acc_multipole(m, pos, 0, acc); break;
    case 3: acc_type3(m, pos, acc); break;
    case 4: acc_type4(m, pos, acc); break;
    case 5: acc_type5(m, pos, acc); break;
    default: acc_type6(m, pos, acc); break;
    }
}

/* getPotential  src/forceAccel.cpp:943-982 */
double or_potential(const or_mol *m, const double P[3])
{
    double total = 0, sX = 0, sY = 0, sZ = 0;
    for (int n = 0; n < m->n; n++) {
        double xD = P[0] - m->x[n], yD = P[1] - m->y[n], zD = P[2] - m->z[n];
        double r2 = xD * xD + yD * yD + zD * zD;
        double r6 = r2 * r2 * r2;
        double term2 = m->sig6[n] / r6;
        double term1 = term2 * term2;
        total += m->eps4[n] * (term1 - term2);
        if (m->charged) {
            double r_3 = m->q[n] / (r2 * sqrt(r2));
            sX += xD * r_3; sY += yD * r_3; sZ += zD * r_3;
        }
    }
    if (m->charged) total -= m->dipole_scalar * 0.5 * (sX * sX + sY * sY + sZ * sZ);
    return total;
}

/* forceAccel::setClusters forceType selection  src/forceAccel.cpp:125-140 */
int or_force_type(int charged, int multipole, int disp_cutoff)
{
    if (charged) return multipole ? (disp_cutoff ? 1 : 2) : (disp_cutoff ? 3 : 4);
    return disp_cutoff ? 5 : 6;
}

/* ====================================================================== */
/* diffEqSolver::verletVel  src/diffEqSolver.cpp:81-211                   */
/* ====================================================================== */
void or_verlet(const or_mol *m, double time, const double p0[3], const double v0[3],
               double a, double b, double c, int projection, double mu, or_traj_out *out)
{
    double maxRad = dmax(a, dmax(b, c));
    double maxTries = 8, trajTries = 0;
    double dt = time, rDot_new = 0, rDot_old = 0, product, orbits = 0;
    double startEnergy, endEnergy, pct;
    double pos[3], vel[3], acc[3];
    int stepCount = 0;
    int64_t total = 0;
    int attempts = 0;
    out->ergy = 0;
    startEnergy = or_potential(m, p0) + 0.5 * mu * mag2_3(v0);
    while (trajTries < maxTries && maxTries < 10) {
        if (stepCount == 0) {
            or_accel(m, p0, acc);
            memcpy(pos, p0, sizeof(pos));
            memcpy(vel, v0, sizeof(vel));
            rDot_old = 0; rDot_new = 0; orbits = 0;
            rDot_new = dot3(pos, vel);
            attempts++;
        }
        for (int k = 0; k < 3; k++) pos[k] = pos[k] + vel[k] * dt + acc[k] * dt * dt * 0.5;
        for (int k = 0; k < 3; k++) vel[k] = vel[k] + acc[k] * dt * 0.5;
        or_accel(m, pos, acc);
        for (int k = 0; k < 3; k++) vel[k] = vel[k] + acc[k] * dt * 0.5;
        stepCount++;
        total++;
        int track = 1;
        rDot_old = rDot_new;
        rDot_new = dot3(pos, vel);
        product = rDot_new * rDot_old;
        if (stepCount > 2 && product < 0 && !(rDot_new > rDot_old)) {
            orbits++;
            if (orbits > 8) track = 0;
        }
        if (!track) {
            trajTries++;
            dt = time / ((trajTries + 1));
        } else {
            double t[3] = {pos[0] / a, pos[1] / b, pos[2] / c};
            double tm2 = mag2_3(t);
            if (((tm2 > 1 && projection) || (tm2 > maxRad && !projection)) && stepCount > 10) {
                endEnergy = or_potential(m, pos) + 0.5 * mu * mag2_3(vel);
                pct = 100 * fabs(startEnergy - endEnergy) / startEnergy;
                out->chi = angle_vec(v0, vel);
                out->steps = stepCount;
                if (pct > 0.5) {
                    trajTries++;
                    maxTries++;
                    dt = time / ((trajTries + 1));
                    stepCount = 0;
                } else {
                    out->ergy = pct;
                    out->ok = 1;
                    out->total_steps = total;
                    out->attempts = attempts;
                    return;
                }
            }
        }
    }
    out->chi = OR_PI_2;
    out->steps = stepCount;
    out->ok = 0;
    out->total_steps = total;
    out->attempts = attempts;
}

void or_verlet_batch(const or_mol *m, double dt, int64_t n, const double *p0, const double *v0,
                     double a, double b, double c, int projection, double mu, double *chi,
                     int32_t *ok, int32_t *steps, double *ergy, int64_t *total_steps, int nthreads)
{
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 16)
#endif
    for (int64_t i = 0; i < n; i++) {
        or_traj_out o;
        or_verlet(m, dt, p0 + 3 * i, v0 + 3 * i, a, b, c, projection, mu, &o);
        chi[i] = o.chi;
        ok[i] = o.ok;
        steps[i] = o.steps;
        ergy[i] = o.ergy;
        if (total_steps) total_steps[i] = o.total_steps;
    }
}

void or_accel_batch(const or_mol *m, int64_t n, const double *pos, double *acc, int nthreads)
{
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(static)
#endif
    for (int64_t i = 0; i < n; i++) or_accel(m, pos + 3 * i, acc + 3 * i);
}

void or_potential_batch(const or_mol *m, int64_t n, const double *pos, double *u, int nthreads)
{
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(static)
#endif
    for (int64_t i = 0; i < n; i++) u[i] = or_potential(m, pos + 3 * i);
}

/* ====================================================================== */
/* impactParam  src/impactParam.cpp                                       */
/* ====================================================================== */
typedef struct { double p[3]; int set; } bpt;

/* setBoundry  src/impactParam.cpp:602-736 */
int or_boundary(const or_mol *m, const double maxxyz[3], double min_kin, double *pts, int cap)
{
    const int numRot = 20;
    const double velMag = 0.1;
    double posMag = fabs(dmax(fabs(maxxyz[0]), fabs(maxxyz[1])));
    double pos[3] = {0, 0, fabs(maxxyz[2]) + 20}, prev[3];
    double energy = or_potential(m, pos);
    double slope = 1.0, zMax = 0;
    min_kin = 0.01 * min_kin;
    if (energy + min_kin < 0) slope = -1;
    for (;;) {
        memcpy(prev, pos, sizeof(pos));
        pos[2] -= 0.5 * slope;
        energy = or_potential(m, pos);
        if ((energy + min_kin) * slope < 0) { zMax = dmax(pos[2], prev[2]); break; }
    }
    zMax *= 1.3;
    posMag += fabs(zMax - fabs(maxxyz[2]));
    int ci = (int)ceil(zMax);
    int maxZLevels = (int)ceil((double)(25 > ci ? 25 : ci));
    double zInc = 2 * zMax / (maxZLevels);
    int nslots = (maxZLevels + 1) * numRot;
    bpt *tmp = (bpt *)calloc((size_t)nslots, sizeof(bpt));
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 1)
#endif
    for (int i = 0; i < maxZLevels + 1; i++) {
        double zPos = -zMax + i * zInc;
        double prevPos[3] = {0, 0, 0};
        for (int n = 0; n < numRot; n++) {
            double sl = 1;
            double theta = n * 2 * OR_PI / ((double)numRot);
            double startX = posMag * cos(theta), startY = posMag * sin(theta);
            double p[3] = {startX, startY, zPos};
            double v[3] = {-velMag * cos(theta), -velMag * sin(theta), 0.0};
            int hit = 0;
            double e = or_potential(m, p);
            if (e + min_kin < 0) sl = -1;
            for (int k = 0; k < 3; k++) v[k] = v[k] * sl;
            while (!hit) {
                e = or_potential(m, p);
                if ((e + min_kin) * sl < 0) {
                    hit = 1;
                    bpt *dst = &tmp[i * numRot + n];
                    if (mag3(prevPos) > mag3(p)) memcpy(dst->p, prevPos, sizeof(p));
                    else memcpy(dst->p, p, sizeof(p));
                    dst->set = 1;
                } else if ((startX == 0 && startY / p[1] <= 0) || (startY == 0 && startX / p[0] <= 0) ||
                           (startX / p[0] <= 0 && startY / p[1] <= 0)) {
                    hit = 1;
                } else {
                    memcpy(prevPos, p, sizeof(p));
                    for (int k = 0; k < 3; k++) p[k] = p[k] + v[k];
                }
            }
        }
    }
    int count = 0;
    for (int i = 0; i < nslots; i++) {
        const double *q = tmp[i].p;
        if (!(q[0] == 0 && q[1] == 0 && q[2] == 0)) { /* :727-735 */
            if (count >= cap) { free(tmp); return -1; }
            memcpy(pts + 3 * count, q, 3 * sizeof(double));
            count++;
        }
    }
    free(tmp);
    return count;
}

/* findEllipsoidBound  src/impactParam.cpp:853-939 */
void or_ellipsoid_fit(const double *P, int np, or_ellipsoid *e)
{
    double mu = 0, nu = 0, ep = 0, de = 0, si = 0, rho = 0, al = 0, be = 0, ga = 0;
    for (int i = 0; i < np; i++) {
        double x2 = P[3 * i] * P[3 * i], y2 = P[3 * i + 1] * P[3 * i + 1], z2 = P[3 * i + 2] * P[3 * i + 2];
        double x4 = x2 * x2, y4 = y2 * y2, z4 = z2 * z2;
        mu = mu + x4; nu = nu + y4; ep = ep + z4;
        de = de + x2 * y2; si = si + x2 * z2; rho = rho + y2 * z2;
        al = al + x2; be = be + y2; ga = ga + z2;
    }
    double det = ep * mu * nu - ep * de * de - mu * rho * rho + 2 * de * rho * si - nu * si * si;
    double A = (al * (nu * ep - rho * rho) + de * (rho * ga - be * ep) + si * (be * rho - nu * ga)) / det;
    double B = (mu * (be * ep - ga * rho) + al * (rho * si - de * ep) + si * (de * ga - be * si)) / det;
    double C = (mu * (nu * ga - be * rho) + de * (be * si - de * ga) + al * (de * rho - nu * si)) / det;
    double a = 1 / sqrt(A), b = 1 / sqrt(B), c = 1 / sqrt(C);
    size_t npts = (size_t)np;
    double rmsd = sqrt((A * A * mu + B * B * nu + C * C * ep +
                        2 * (A * B * de + B * C * rho + A * C * si - A * al - B * be - C * ga) + npts) /
                       (npts - 1));
    e->a = a; e->b = b; e->c = c; e->rmsd = rmsd;
    double caa = 2 * rmsd * rmsd * (ep * nu - rho * rho) * a * a / (4 * A * A * det);
    double cbb = 2 * rmsd * rmsd * (ep * mu - si * si) * b * b / (4 * B * B * det);
    double ccc = 2 * rmsd * rmsd * (mu * nu - de * de) * c * c / (4 * C * C * det);
    double cab = 2 * rmsd * rmsd * (rho * si - de * ep) * a * b / (4 * A * B * det);
    double cac = 2 * rmsd * rmsd * (de * rho - nu * si) * a * c / (4 * A * C * det);
    double cbc = 2 * rmsd * rmsd * (de * si - mu * rho) * b * c / (4 * B * C * det);
    e->cov[0] = caa; e->cov[1] = cbb; e->cov[2] = ccc; e->cov[3] = cab; e->cov[4] = cac; e->cov[5] = cbc;
    e->a_unc = (a / A) * sqrt((ep * nu - rho * rho) * 2 / det) * rmsd / 2;
    e->b_unc = (b / B) * sqrt((ep * mu - si * si) * 2 / det) * rmsd / 2;
    e->c_unc = (c / C) * sqrt((mu * nu - de * de) * 2 / det) * rmsd / 2;
    e->avg = (a + b + c) / 3.0;
    e->avg_unc = sqrt(caa + cbb + ccc + 2 * (cab + cac + cbc)) / 3.0;
    double p = 1.6075;
    double f = (pow(a * b, p) + pow(a * c, p) + pow(b * c, p)) / 3;
    e->area = 4 * OR_PI * pow(f, 1.0 / p);
    double da = e->area * pow(a, p - 1) * (pow(b, p) + pow(c, p)) / (3 * f);
    double db = e->area * pow(b, p - 1) * (pow(a, p) + pow(c, p)) / (3 * f);
    double dc = e->area * pow(c, p - 1) * (pow(a, p) + pow(b, p)) / (3 * f);
    e->area_unc = sqrt(da * da * caa + db * db * cbb + dc * dc * ccc + 2 * (da * db * cab + da * dc * cac + db * dc * cbc));
    e->enlarge = 0;
}

/* enlargeEllipBoundry  src/impactParam.cpp:950-995 */
void or_enlarge(const double *P, int np, or_ellipsoid *e)
{
    double aOld = e->a, a = e->a, b = e->b, c = e->c;
    int accept = 0;
    do {
        for (int i = 0; i < np + 1; i++) {
            if (i == np || a > 3 * aOld) { accept = 1; break; }
            double mag = P[3 * i] * P[3 * i] / (a * a) + P[3 * i + 1] * P[3 * i + 1] / (b * b) +
                         P[3 * i + 2] * P[3 * i + 2] / (c * c);
            if (mag > 1) { a = a + 0.01; b = b + 0.01; c = c + 0.01; break; }
        }
    } while (!accept);
    e->a = a; e->b = b; e->c = c;
    e->enlarge = a - aOld;
}

/* ====================================================================== */
/* PCM_Cluster + eigen3D  src/PCM_Cluster.cpp, src/eigen3D.cpp            */
/* ====================================================================== */

/* eigen3D::getEigenvalues  src/eigen3D.cpp:34-79.  The reference leaves
 * `b` uninitialised when D >= 0 (:65-68); it is taken as 0 here. */
static void eig_values_pcm(double A[3][3], double ev[3])
{
    double a0, a1, a2, p, q, R, Q, D, a, b = 0, phase, mag;
    a0 = A[0][0] * A[1][1] * A[2][2] - A[0][0] * A[1][2] * A[2][1] - A[1][1] * A[0][2] * A[2][0] -
         A[2][2] * A[0][1] * A[1][0] + A[0][1] * A[1][2] * A[2][0] + A[0][2] * A[2][1] * A[1][0];
    a1 = A[0][1] * A[1][0] + A[0][2] * A[2][0] + A[1][2] * A[2][1] - A[0][0] * A[1][1] - A[1][1] * A[2][2] -
         A[2][2] * A[0][0];
    a2 = A[0][0] + A[1][1] + A[2][2];
    a0 = -a0; a1 = -a1; a2 = -a2;
    p = (3 * a1 - a2 * a2) / 3;
    q = (9 * a1 * a2 - 27 * a0 - 2 * a2 * a2 * a2) / 27;
    Q = p / 3; R = q / 2; D = Q * Q * Q + R * R;
    if (D < 0) { a = R; b = sqrt(-D); }
    else { a = R + sqrt(D); }
    phase = atan2(b, a);
    mag = pow(a * a + b * b, 1.0 / 6.0);
    ev[0] = -(a2 / 3) + mag * 2 * cos(phase / 3.0);
    ev[1] = -(a2 / 3) - mag * (cos(phase / 3.0) + sqrt(3.0) * sin(phase / 3.0));
    ev[2] = -(a2 / 3) - mag * (cos(phase / 3.0) - sqrt(3.0) * sin(phase / 3.0));
}

/* eigen3D::getEigenvectors  src/eigen3D.cpp:81-113 (rows = vectors) */
static void eig_vectors_pcm(double A[3][3], const double l[3], double E[3][3])
{
    for (int n = 0; n < 3; n++) {
        double det = l[n] * A[0][2] - A[0][2] * A[1][1] + A[0][1] * A[1][2];
        double y = (A[0][2] * A[1][0] + l[n] * A[1][2] - A[0][0] * A[1][2]) / det;
        double z = (l[n] * l[n] - l[n] * (A[0][0] + A[1][1]) - A[0][1] * A[1][0] + A[0][0] * A[1][1]) / det;
        double nf = sqrt(1 + y * y + z * z);
        E[n][0] = 1.0 / nf; E[n][1] = y / nf; E[n][2] = z / nf;
    }
}

typedef struct { int *idx; int n; } grp;

/* setCenterVecs + setProperties  src/PCM_Cluster.cpp:163-225 */
static void pcm_props(const double *xyz, const grp *g, int ng, double *center, double *maxd)
{
    for (int i = 0; i < ng; i++) {
        double c[3] = {0, 0, 0};
        for (int j = 0; j < g[i].n; j++)
            for (int k = 0; k < 3; k++) c[k] = c[k] + xyz[3 * g[i].idx[j] + k];
        for (int k = 0; k < 3; k++) c[k] = c[k] / (double)g[i].n;
        memcpy(center + 3 * i, c, sizeof(c));
        double md[3] = {0, 0, 0};
        int first = 1;
        for (int j = 0; j < g[i].n; j++) {
            double d[3];
            for (int k = 0; k < 3; k++) d[k] = c[k] - xyz[3 * g[i].idx[j] + k];
            double dm = mag3(d);
            if (first) first = 0;            /* :193-196 first atom only seeds minDist */
            else if (dm > mag3(md)) memcpy(md, d, sizeof(d));
        }
        maxd[i] = mag3(md);
    }
}

/* formClusters  src/PCM_Cluster.cpp:38-132, getCovMat :134-161,
 * orderClusters :227-249 */
int or_pcm_clusters(int n, const double *xyz, int mcs, int *start, int *atoms, double *center, double *maxdist)
{
    int cap = n + 1, ng = 1;
    grp *g = (grp *)calloc((size_t)cap, sizeof(grp));
    double *cen = (double *)malloc(sizeof(double) * 3 * cap);
    double *md = (double *)malloc(sizeof(double) * cap);
    g[0].idx = (int *)malloc(sizeof(int) * (n > 0 ? n : 1));
    g[0].n = n;
    for (int i = 0; i < n; i++) g[0].idx[i] = i;
    pcm_props(xyz, g, ng, cen, md);
    int accept, rc = 0;
    do {
        accept = 1;
        int size = ng;
        for (int i = 0; i < size; i++) {
            if (md[i] > mcs) {
                double cov[3][3], mean[3] = {0, 0, 0}, ev[3], E[3][3];
                for (int k = 0; k < g[i].n; k++)
                    for (int j = 0; j < 3; j++) mean[j] += xyz[3 * g[i].idx[k] + j];
                for (int j = 0; j < 3; j++) mean[j] = mean[j] / ((double)g[i].n);
                for (int a = 0; a < 3; a++)
                    for (int b = a; b < 3; b++) {
                        double s = 0;
                        for (int k = 0; k < g[i].n; k++)
                            s += (xyz[3 * g[i].idx[k] + a] - mean[a]) * (xyz[3 * g[i].idx[k] + b] - mean[b]);
                        cov[a][b] = s;
                        if (a != b) cov[b][a] = s;
                    }
                eig_values_pcm(cov, ev);
                eig_vectors_pcm(cov, ev, E);
                int emax = 0;
                for (int k = 0; k < 3; k++)
                    if (fabs(ev[k]) > ev[emax]) emax = k;
                double c[3] = {0, 0, 0};
                for (int k = 0; k < g[i].n; k++)
                    for (int j = 0; j < 3; j++) c[j] = c[j] + xyz[3 * g[i].idx[k] + j];
                for (int j = 0; j < 3; j++) c[j] = c[j] / ((double)g[i].n);
                int *c1 = (int *)malloc(sizeof(int) * g[i].n), n1 = 0;
                int *c2 = (int *)malloc(sizeof(int) * g[i].n), n2 = 0;
                for (int k = 0; k < g[i].n; k++) {
                    double d[3];
                    for (int j = 0; j < 3; j++) d[j] = xyz[3 * g[i].idx[k] + j] - c[j];
                    /* vector3D::Angle(type(&)[3])  src/vector3D.h:146-159 */
                    double t1 = d[0] * E[emax][0] + d[1] * E[emax][1] + d[2] * E[emax][2];
                    double t2 = mag3(d) * sqrt(E[emax][0] * E[emax][0] + E[emax][1] * E[emax][1] + E[emax][2] * E[emax][2]);
                    double t3 = t1 / t2, ang;
                    if (t3 > 1) ang = 0;
                    else if (t3 < -1) ang = OR_PI;
                    else ang = acos(t1 / t2);
                    if (ang <= OR_PI_2) c1[n1++] = g[i].idx[k];
                    else c2[n2++] = g[i].idx[k];
                }
                if (n1 == 0 || n2 == 0 || ng >= cap) { /* reference stalls on an empty cluster (:214-218) */
                    free(c1); free(c2); rc = -1; goto done;
                }
                free(g[i].idx);
                g[i].idx = c1; g[i].n = n1;
                g[ng].idx = c2; g[ng].n = n2;
                ng++;
            }
        }
        double *ncen = (double *)realloc(cen, sizeof(double) * 3 * cap);
        cen = ncen;
        pcm_props(xyz, g, ng, cen, md);
        for (int i = 0; i < ng; i++)
            if (md[i] > mcs) { accept = 0; break; }
    } while (!accept);
    pcm_props(xyz, g, ng, cen, md);
    for (int k = 0; k < ng - 1; k++) { /* orderClusters: selection sort on centre z */
        int mi = k;
        double mv = cen[3 * k + 2];
        for (int i = k + 1; i < ng; i++)
            if (cen[3 * i + 2] < mv) { mv = cen[3 * i + 2]; mi = i; }
        double t = md[k]; md[k] = md[mi]; md[mi] = t;
        for (int j = 0; j < 3; j++) { t = cen[3 * k + j]; cen[3 * k + j] = cen[3 * mi + j]; cen[3 * mi + j] = t; }
        grp tg = g[k]; g[k] = g[mi]; g[mi] = tg;
    }
    start[0] = 0;
    for (int i = 0; i < ng; i++) {
        memcpy(atoms + start[i], g[i].idx, sizeof(int) * g[i].n);
        start[i + 1] = start[i] + g[i].n;
        memcpy(center + 3 * i, cen + 3 * i, 3 * sizeof(double));
        maxdist[i] = md[i];
    }
    rc = ng;
done:
    for (int i = 0; i < ng; i++) free(g[i].idx);
    free(g); free(cen); free(md);
    return rc;
}

/* forceAccel::addMultipole :143-185 and addClusterPos :187-210 */
void or_cluster_moments(int n, const double *xyz, const double *q, int ncl, const int *start, const int *atoms,
                        const double *center, double *cl_q, double *cl_dip, double *cl_quad, double *cl_pos,
                        int *cl_replace)
{
    (void)n;
    for (int c = 0; c < ncl; c++) {
        const double *cc = center + 3 * c;
        double sumq = 0, dip[3] = {0, 0, 0}, Q[9] = {0};
        for (int k = start[c]; k < start[c + 1]; k++) {
            int idx = atoms[k];
            double charge = q[idx], p[3];
            sumq += charge;
            for (int j = 0; j < 3; j++) p[j] = xyz[3 * idx + j] - cc[j];
            double m2 = p[0] * p[0] + p[1] * p[1] + p[2] * p[2];
            for (int i = 0; i < 3; i++) {
                dip[i] += charge * p[i];
                for (int j = 0; j < 3; j++) Q[3 * i + j] += charge * (3 * p[i] * p[j] - m2 * (i == j ? 1.0 : 0.0));
            }
        }
        cl_q[c] = sumq;
        memcpy(cl_dip + 3 * c, dip, sizeof(dip));
        memcpy(cl_quad + 9 * c, Q, sizeof(Q));
        double coc[3] = {0, 0, 0};
        for (int k = start[c]; k < start[c + 1]; k++) {
            int idx = atoms[k];
            for (int j = 0; j < 3; j++) coc[j] = coc[j] + xyz[3 * idx + j] * q[idx];
        }
        for (int j = 0; j < 3; j++) coc[j] = coc[j] / sumq;
        double pd = 100 * fabs((mag3(coc) - mag3(cc)) / mag3(coc));
        if (pd <= 5) {
            cl_replace[c] = 1;
            memcpy(cl_pos + 3 * c, coc, sizeof(coc));
        } else {
            cl_replace[c] = 0;
            memcpy(cl_pos + 3 * c, cc, 3 * sizeof(double));
        }
    }
}

/* ====================================================================== */
/* Initial conditions: Philox4x32-10 (Salmon et al., SC'11) keyed by the  */
/* seed, counter (trajectory, iteration, job, draw block); rejection       */
/* sampling of monteTraj::setUpTraj  src/monteTraj.cpp:174-278            */
/* ====================================================================== */
void or_philox4x32(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4])
{
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3], k0 = key[0], k1 = key[1];
    for (int r = 0; r < 10; r++) {
        uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
        uint32_t n1 = (uint32_t)p1;
        uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
        uint32_t n3 = (uint32_t)p0;
        c0 = n0; c1 = n1; c2 = n2; c3 = n3;
        k0 += 0x9E3779B9u; k1 += 0xBB67AE85u;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

static double u01(uint32_t hi, uint32_t lo)
{
    uint64_t x = ((uint64_t)hi << 32) | lo;
    return (double)(x >> 11) * (1.0 / 9007199254740992.0);
}

int or_sample_ic(const or_ic_cfg *cfg, uint32_t traj, double p0[3], double v0[3], double *gout, int *tries)
{
    const uint32_t key[2] = {(uint32_t)cfg->seed, (uint32_t)(cfg->seed >> 32)};
    double g, gCheck, phi, theta, thetaCheck, gamma, b, bCheck;
    double bMax = cfg->bmax;
    double maxDist = (float)bMax; /* monteTraj::maxDist is a float (src/monteTraj.h:67) */
    int t = 0;
    for (;; t++) {
        uint32_t w[16];
        for (int blk = 0; blk < 4; blk++) {
            uint32_t ctr[4] = {traj, cfg->iteration, cfg->job, (uint32_t)(4 * t + blk)};
            or_philox4x32(ctr, key, w + 4 * blk);
        }
        double u[8];
        for (int k = 0; k < 8; k++) u[k] = u01(w[2 * k], w[2 * k + 1]);
        g = u[0] * (cfg->gmax - cfg->gmin) + cfg->gmin;
        gCheck = u[1] * cfg->max_val;
        phi = u[2] * 2.0 * OR_PI;
        theta = u[3] * OR_PI;
        thetaCheck = 0.5 * u[4];
        gamma = u[5] * 2.0 * OR_PI;
        b = u[6] * bMax;
        bCheck = u[7] * (2 / bMax);
        if (thetaCheck <= sin(theta) / 2 && bCheck <= 2 * b / (bMax * bMax) &&
            gCheck <= pow(g, 5) * exp(-cfg->vd_a * g * g) * cfg->vd_norm)
            break;
        if (t >= 4095) { *tries = t + 1; *gout = 0; return 0; }
    }
    *tries = t + 1;
    double P[3], V[3];
    P[0] = b * cos(gamma) * cos(phi) * cos(theta) - b * sin(gamma) * sin(phi) + maxDist * cos(phi) * sin(theta);
    P[1] = b * cos(gamma) * sin(phi) * cos(theta) + b * sin(gamma) * cos(phi) + maxDist * sin(phi) * sin(theta);
    P[2] = maxDist * cos(theta) - b * cos(gamma) * sin(theta);
    V[0] = -g * cos(phi) * sin(theta);
    V[1] = -g * sin(phi) * sin(theta);
    V[2] = -g * cos(theta);
    *gout = g;
    if (cfg->projection) {
        double v[3] = {V[0] / cfg->ea, V[1] / cfg->eb, V[2] / cfg->ec};
        double p[3] = {P[0] / cfg->ea, P[1] / cfg->eb, P[2] / cfg->ec};
        double cr[3] = {(p[1] * v[2]) - (p[2] * v[1]), (p[2] * v[0]) - (p[0] * v[2]), (p[0] * v[1]) - (p[1] * v[0])};
        double disc = -dot3(cr, cr) + dot3(v, v);
        if (disc < 0) {
            memcpy(p0, P, sizeof(P));
            memcpy(v0, V, sizeof(V));
            return 0;
        }
        double pp = -(dot3(p, v) / (dot3(v, v))) - sqrt(disc) / dot3(v, v);
        for (int k = 0; k < 3; k++) P[k] = P[k] + V[k] * pp;
    }
    memcpy(p0, P, sizeof(P));
    memcpy(v0, V, sizeof(V));
    return 1;
}

/* monteTraj::runThreads :456-502, updateCounts :504-521, runTraj :380-387
 * for one OpenMP thread. */
void or_reduce_iteration(int64_t n, const double *chi, const int32_t *ok, const int32_t *hit, double ccs_const,
                         int64_t update_cond, or_ccs *out)
{
    int succ = 0;
    int64_t orbit = 0, cnt = 0;
    double total = 0, total2 = 0;
    for (int64_t i = 0; i < n; i++) {
        double c = 0;
        if (hit[i]) { succ = ok[i] != 0; c = chi[i]; }
        cnt++;
        if (!succ) orbit++;
        double x = ccs_const * (1 - cos(fabs(c)));
        total += x;
        total2 += x * x;
        if (update_cond > 0 && cnt % update_cond == 0) orbit = 0;
    }
    int64_t d = n - orbit;
    out->total = total;
    out->total2 = total2;
    out->denom = d;
    out->omega = total / d;
    out->err = sqrt((total2 / d - (total / d) * (total / d)) / d);
}

/* monteTraj::runTraj final statistics  src/monteTraj.cpp:412-425 */
void or_reduce_final(int iters, const double *omega, const double *err, double *mean, double *sd, double *wmean,
                     double *werr)
{
    double avg = 0, s = 0, we = 0, wm = 0;
    for (int i = 0; i < iters; i++) avg += omega[i];
    avg = avg / (double)iters;
    for (int i = 0; i < iters; i++) {
        s += (omega[i] - avg) * (omega[i] - avg);
        we += 1.0 / (err[i] * err[i]);
        wm += omega[i] / (err[i] * err[i]);
    }
    wm /= we;
    s = sqrt(s / (double)iters);
    we = 1.0 / sqrt(we);
    *mean = avg; *sd = s; *wmean = wm; *werr = we;
}
