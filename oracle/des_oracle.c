/*
 * des_oracle.c -- CPU restatement of DynEarthSol's explicit time step.
 *
 * TEST INFRASTRUCTURE ONLY.  This is the oracle: the tests, the smoke check and
 * the benchmark's CPU-baseline leg call it as the checker of the HIP library.
 * The product library never links, loads or calls it.
 *
 * It follows the reference's arithmetic expression by expression (same operand
 * order, no FMA contraction: build with -ffp-contract=off) so that, except for
 * libm transcendentals, it reproduces the reference's serial results.  Node
 * sums that the reference scatters element by element (utils.hpp:18-44) are
 * gathered here through the node->element CSR in ascending element order,
 * which is the reference's serial summation order, so the sums are identical.
 *
 * Compiled twice, like the reference (Makefile:233-238): -DNDIMS=3 gives the
 * o3_* symbols, -DNDIMS=2 the o2_* symbols.
 */
#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "des_amd.h"

#ifndef NDIMS
#error "compile with -DNDIMS=2 or -DNDIMS=3"
#endif

#if NDIMS == 3
#define NPE 4
#define NSTR 6
#define FN(x) o3_##x
#else
#define NPE 3
#define NSTR 3
#define FN(x) o2_##x
#endif
#define NPF NDIMS

/* boundary flag bits (constants.hpp:38-51) */
#define BOUNDX0 1u
#define BOUNDX1 2u
#define BOUNDY0 4u
#define BOUNDY1 8u
#define BOUNDZ0 16u
#define BOUNDZ1 32u
#define BOUND_ANY 1023u
#define IBOUNDZ0 4
#define IBOUNDZ1 5
#define IBOUNDN0 6
#define IBOUNDN3 9

#define DEG2RAD (M_PI / 180)

/* local node numbers of a facet (constants.hpp:61-71) */
#if NDIMS == 3
static const int NODE_OF_FACET[4][3] = {{1, 2, 3}, {0, 3, 2}, {0, 1, 3}, {0, 2, 1}};
#else
static const int NODE_OF_FACET[3][2] = {{1, 2}, {2, 0}, {0, 1}};
#endif

/*
 * Model state: the reference's Variables (parameters.hpp:229-279) with every
 * array in its AoS layout.  Mirrored field by field in oracle/oracle.py.
 */
typedef struct ostate {
    int nnode, nelem, nmat, markers_per_element;
    const int *conn;            /* nelem x NPE */
    const unsigned *bcflag;     /* nnode */
    const int *elemmarkers;     /* nelem x nmat or NULL */
    const int *csr_off;         /* nnode+1 */
    const int *csr_ent;         /* e*NPE+i, ascending e per node */
    int nbfacet;
    const int *bf_bdry, *bf_elem, *bf_face;
    const double *bnormals;     /* 10 x NDIMS */
    const double *edge_vectors; /* 10 x 10 x NDIMS */
    const int *edge_valid;      /* 10 x 10 */
    double *coord, *coord0, *vel, *force, *temperature, *mass, *tmass, *volume_n, *ntmp;
    double *volume, *volume_old, *edvoldt, *plstrain, *delta_plstrain, *dpressure, *stressyy;
    double *strain_rate, *strain, *stress, *shpdx, *shpdy, *shpdz;
    const double *elem_visc, *elem_cohesion;
    double *ework;              /* scratch: nelem x NPE x NDIMS */
    double dt, time, compensation_pressure;
    int steps;
    int status;
} ostate;

static double dmin(double a, double b) { return (b < a) ? b : a; }   /* std::min */
static double dmax(double a, double b) { return (a < b) ? b : a; }   /* std::max */

/* ---------------------------------------------------------------- utils.hpp */

static double trace(const double *s)                    /* utils.hpp:100-107 */
{
#if NDIMS == 3
    return s[0] + s[1] + s[2];
#else
    return s[0] + s[1];
#endif
}

static double second_invariant2(const double *t)        /* utils.hpp:110-119 */
{
#if NDIMS == 3
    double a = (t[0] + t[1] + t[2]) / 3;
    return (0.5 * ((t[0] - a) * (t[0] - a) + (t[1] - a) * (t[1] - a) + (t[2] - a) * (t[2] - a))
            + t[3] * t[3] + t[4] * t[4] + t[5] * t[5]);
#else
    return 0.25 * (t[0] - t[1]) * (t[0] - t[1]) + t[2] * t[2];
#endif
}

static double second_invariant(const double *t)         /* utils.hpp:122-128 */
{
    return sqrt(second_invariant2(t));
}

/* ------------------------------------------------------------- matprops.cxx */

static int count_of(const ostate *S, int e, int m)
{
    if (S->elemmarkers) return S->elemmarkers[e * S->nmat + m];
    return m == 0 ? S->markers_per_element : 0;
}

/* harmonic_mean / arithmetic_mean with the one-value shortcut (matprops.cxx:102-128) */
static double harmonic_mean(const ostate *S, const double *s, int e)
{
    if (S->nmat == 1) return s[0];
    double result = 0;
    int m = 0;
    for (int i = 0; i < S->nmat; i++) {
        int k = count_of(S, e, i);
        result += k / s[i];
        m += k;
    }
    return m / result;
}

static double arithmetic_mean(const ostate *S, const double *s, int e)
{
    if (S->nmat == 1) return s[0];
    double result = 0;
    int m = 0;
    for (int i = 0; i < S->nmat; i++) {
        int k = count_of(S, e, i);
        result += k * s[i];
        m += k;
    }
    return result / m;
}

static double mat_bulkm(const ostate *S, const des_params *P, int e)    /* matprops.cxx:263-266 */
{
    return harmonic_mean(S, P->bulk_modulus, e);
}

static double mat_shearm(const ostate *S, const des_params *P, int e)   /* matprops.cxx:269-272 */
{
    return harmonic_mean(S, P->shear_modulus, e);
}

static double elem_mean_T(const ostate *S, int e)
{
    double T = 0;
    const int *conn = S->conn + e * NPE;
    for (int i = 0; i < NPE; ++i) T += S->temperature[conn[i]];
    T /= NPE;
    return T;
}

/* MatProps::visc (matprops.cxx:275-311); a per-element viscosity field, when
   given, replaces the power law and keeps the [visc_min, visc_max] clamp. */
static double mat_visc(const ostate *S, const des_params *P, int e)
{
    double visc;
    if (S->elem_visc) {
        visc = S->elem_visc[e];
    } else {
        const double gas_constant = 8.3144;
        const double min_strain_rate = 1e-30;
        double T = elem_mean_T(S, e);
        double edot = second_invariant(S->strain_rate + e * NSTR);
        edot = dmax(edot, min_strain_rate);
        double result = 0;
        int n = 0;
        for (int m = 0; m < S->nmat; m++) {
            double pw = 1 / P->visc_exponent[m] - 1;
            double pw1 = -1 / P->visc_exponent[m];
            double visc0 = 0.25 * pow(edot, pw) * pow(0.75 * P->visc_coefficient[m], pw1)
                * exp(P->visc_activation_energy[m] / (P->visc_exponent[m] * gas_constant * T)) * 1e6;
            int k = count_of(S, e, m);
            result += k / visc0;
            n += k;
        }
        visc = n / result;
    }
    visc = dmin(dmax(visc, P->visc_min), P->visc_max);
    return visc;
}

/* MatProps::plastic_weakening + plastic_props (matprops.cxx:314-372) */
static void mat_plastic_props(const ostate *S, const des_params *P, int e, double pls,
                              double *amc, double *anphi, double *anpsi,
                              double *hardn, double *ten_max)
{
    double c, f, d, h;
    c = f = d = h = 0;
    int n = 0;
    for (int m = 0; m < S->nmat; m++) {
        int k = count_of(S, e, m);
        if (k == 0) continue;
        n += k;
        double c0 = S->elem_cohesion ? S->elem_cohesion[e] : P->cohesion0[m];
        double c1 = P->cohesion1[m];
        if (pls < P->pls0[m]) {
            c += c0 * k;
            f += P->friction_angle0[m] * k;
            d += P->dilation_angle0[m] * k;
            h += 0;
        } else if (pls < P->pls1[m]) {
            double p = (pls - P->pls0[m]) / (P->pls1[m] - P->pls0[m]);
            c += (c0 + p * (c1 - c0)) * k;
            f += (P->friction_angle0[m] + p * (P->friction_angle1[m] - P->friction_angle0[m])) * k;
            d += (P->dilation_angle0[m] + p * (P->dilation_angle1[m] - P->dilation_angle0[m])) * k;
            h += (c1 - c0) / (P->pls1[m] - P->pls0[m]) * k;
        } else {
            c += c1 * k;
            f += P->friction_angle1[m] * k;
            d += P->dilation_angle1[m] * k;
            h += 0;
        }
    }
    double cohesion = c / n;
    double phi = f / n;
    double psi = d / n;
    *hardn = h / n;

    double sphi = sin(phi * DEG2RAD);
    double spsi = sin(psi * DEG2RAD);
    *anphi = (1 + sphi) / (1 - sphi);
    *anpsi = (1 + spsi) / (1 - spsi);
    *amc = 2 * cohesion * sqrt(*anphi);
    *ten_max = (phi == 0) ? P->tension_max : dmin(P->tension_max, cohesion / tan(phi * DEG2RAD));
}

static double mat_rho(const ostate *S, const des_params *P, int e)      /* matprops.cxx:375-396 */
{
    const double celsius0 = 273;
    double T = elem_mean_T(S, e);
    double TinCelsius = T - celsius0;
    double result = 0;
    int n = 0;
    for (int m = 0; m < S->nmat; m++) {
        int k = count_of(S, e, m);
        result += P->rho0[m] * (1 - P->alpha[m] * TinCelsius) * k;
        n += k;
    }
    return result / n;
}

static double mat_cp(const ostate *S, const des_params *P, int e) { return arithmetic_mean(S, P->heat_capacity, e); }
static double mat_hr(const ostate *S, const des_params *P, int e) { return arithmetic_mean(S, P->radiogenic_heating, e); }
static double mat_k(const ostate *S, const des_params *P, int e) { return arithmetic_mean(S, P->therm_cond, e); }

/* PREM pressure tables (matprops.cxx:11-99): reference data, depth in m, pressure in Pa */
static const double prem_depth[46] = {
    0e3, 3e3, 15e3, 24.4e3, 40e3, 60e3, 80e3, 115e3, 150e3, 185e3, 220e3, 265e3, 310e3, 355e3,
    400e3, 450e3, 500e3, 550e3, 600e3, 635e3, 670e3, 721e3, 771e3, 871e3, 971e3, 1071e3,
    1171e3, 1271e3, 1371e3, 1471e3, 1571e3, 1671e3, 1771e3, 1871e3, 1971e3, 2071e3, 2171e3,
    2271e3, 2371e3, 2471e3, 2571e3, 2671e3, 2741e3, 2771e3, 2871e3, 2891e3};
static const double prem_p_iso[46] = {
    0e8, 0.3e8, 3.3e8, 6.0e8, 11.2e8, 17.8e8, 24.5e8, 36.1e8, 47.8e8, 59.4e8, 71.1e8, 86.4e8,
    102.0e8, 117.7e8, 133.5e8, 152.2e8, 171.3e8, 190.7e8, 210.4e8, 224.3e8, 238.3e8, 260.7e8,
    282.9e8, 327.6e8, 372.8e8, 418.6e8, 464.8e8, 511.6e8, 558.9e8, 606.8e8, 655.2e8, 704.1e8,
    753.5e8, 803.6e8, 854.3e8, 905.6e8, 957.6e8, 1010.3e8, 1063.8e8, 1118.2e8, 1173.4e8,
    1229.7e8, 1269.7e8, 1287.0e8, 1345.6e8, 1357.5e8};
static const double prem_p_cont[46] = {
    0e8, 0.82e8, 4.1e8, 6.7e8, 11.2e8, 17.8e8, 24.5e8, 36.1e8, 47.8e8, 59.4e8, 71.1e8, 86.4e8,
    102.0e8, 117.7e8, 133.5e8, 152.2e8, 171.3e8, 190.7e8, 210.4e8, 224.3e8, 238.3e8, 260.7e8,
    282.9e8, 327.6e8, 372.8e8, 418.6e8, 464.8e8, 511.6e8, 558.9e8, 606.8e8, 655.2e8, 704.1e8,
    753.5e8, 803.6e8, 854.3e8, 905.6e8, 957.6e8, 1010.3e8, 1063.8e8, 1118.2e8, 1173.4e8,
    1229.7e8, 1269.7e8, 1287.0e8, 1345.6e8, 1357.5e8};

static double prem_pressure(const double *tab, double depth)   /* matprops.cxx:11-53 */
{
    if (depth <= 0) return 0;
    int n;
    for (n = 1; n < 46; n++)
        if (depth <= prem_depth[n]) break;
    return tab[n - 1] + (tab[n] - tab[n - 1]) * (depth - prem_depth[n - 1])
        / (prem_depth[n] - prem_depth[n - 1]);
}

static double ref_pressure(const des_params *P, double z)       /* matprops.cxx:189-201 */
{
    double depth = -z;
    double p = 0;
    if (P->ref_pressure_option == 0)
        p = P->rho0[0] * P->gravity * depth;
    else if (P->ref_pressure_option == 1)
        p = prem_pressure(prem_p_iso, depth);
    else if (P->ref_pressure_option == 2)
        p = prem_pressure(prem_p_cont, depth);
    return p;
}

/* ------------------------------------------------------------ geometry.cxx */

static double dist2(const double *a, const double *b)   /* geometry.cxx:16-24 */
{
    double sum = 0;
    for (int i = 0; i < NDIMS; ++i) {
        double d = b[i] - a[i];
        sum += d * d;
    }
    return sum;
}

#if NDIMS == 3
static double tetrahedron_volume(const double *d0, const double *d1,
                                 const double *d2, const double *d3)   /* geometry.cxx:29-49 */
{
    double x01 = d0[0] - d1[0], x12 = d1[0] - d2[0], x23 = d2[0] - d3[0];
    double y01 = d0[1] - d1[1], y12 = d1[1] - d2[1], y23 = d2[1] - d3[1];
    double z01 = d0[2] - d1[2], z12 = d1[2] - d2[2], z23 = d2[2] - d3[2];
    return (x01 * (y23 * z12 - y12 * z23) +
            x12 * (y01 * z23 - y23 * z01) +
            x23 * (y12 * z01 - y01 * z12)) / 6;
}
#endif

static double triangle_area(const double *a, const double *b, const double *c) /* geometry.cxx:53-83 */
{
    double ab0 = b[0] - a[0], ab1 = b[1] - a[1];
    double ac0 = c[0] - a[0], ac1 = c[1] - a[1];
#if NDIMS == 2
    return fabs(ab0 * ac1 - ab1 * ac0) / 2;
#else
    double ab2 = b[2] - a[2], ac2 = c[2] - a[2];
    double d0 = ab1 * ac2 - ab2 * ac1;
    double d1 = ab2 * ac0 - ab0 * ac2;
    double d2 = ab0 * ac1 - ab1 * ac0;
    return sqrt(d0 * d0 + d1 * d1 + d2 * d2) / 2;
#endif
}

static double elem_volume(const ostate *S, int e)
{
    const int *c = S->conn + e * NPE;
    const double *x = S->coord;
#if NDIMS == 3
    return tetrahedron_volume(x + 3 * c[0], x + 3 * c[1], x + 3 * c[2], x + 3 * c[3]);
#else
    return triangle_area(x + 2 * c[0], x + 2 * c[1], x + 2 * c[2]);
#endif
}

void FN(compute_volume)(ostate *S)                        /* geometry.cxx:86-108 */
{
    #pragma omp parallel for
    for (int e = 0; e < S->nelem; ++e) S->volume[e] = elem_volume(S, e);
}

void FN(compute_shape_fn)(ostate *S)                      /* geometry.cxx:368-455 */
{
    #pragma omp parallel for
    for (int e = 0; e < S->nelem; ++e) {
        const int *c = S->conn + e * NPE;
        const double *d0 = S->coord + NDIMS * c[0];
        const double *d1 = S->coord + NDIMS * c[1];
        const double *d2 = S->coord + NDIMS * c[2];
        double *sx = S->shpdx + e * NPE, *sz = S->shpdz + e * NPE;
#if NDIMS == 3
        double *sy = S->shpdy + e * NPE;
        const double *d3 = S->coord + NDIMS * c[3];
        double iv = 1 / (6 * S->volume[e]);
        double x01 = d0[0] - d1[0], x02 = d0[0] - d2[0], x03 = d0[0] - d3[0];
        double x12 = d1[0] - d2[0], x13 = d1[0] - d3[0], x23 = d2[0] - d3[0];
        double y01 = d0[1] - d1[1], y02 = d0[1] - d2[1], y03 = d0[1] - d3[1];
        double y12 = d1[1] - d2[1], y13 = d1[1] - d3[1], y23 = d2[1] - d3[1];
        double z01 = d0[2] - d1[2], z02 = d0[2] - d2[2], z03 = d0[2] - d3[2];
        double z12 = d1[2] - d2[2], z13 = d1[2] - d3[2], z23 = d2[2] - d3[2];
        sx[0] = iv * (y13 * z12 - y12 * z13);
        sx[1] = iv * (y02 * z23 - y23 * z02);
        sx[2] = iv * (y13 * z03 - y03 * z13);
        sx[3] = iv * (y01 * z02 - y02 * z01);
        sy[0] = iv * (z13 * x12 - z12 * x13);
        sy[1] = iv * (z02 * x23 - z23 * x02);
        sy[2] = iv * (z13 * x03 - z03 * x13);
        sy[3] = iv * (z01 * x02 - z02 * x01);
        sz[0] = iv * (x13 * y12 - x12 * y13);
        sz[1] = iv * (x02 * y23 - x23 * y02);
        sz[2] = iv * (x13 * y03 - x03 * y13);
        sz[3] = iv * (x01 * y02 - x02 * y01);
#else
        double iv = 1 / (2 * S->volume[e]);
        sx[0] = iv * (d1[1] - d2[1]);
        sx[1] = iv * (d2[1] - d0[1]);
        sx[2] = iv * (d0[1] - d1[1]);
        sz[0] = iv * (d2[0] - d1[0]);
        sz[1] = iv * (d0[0] - d2[0]);
        sz[2] = iv * (d1[0] - d0[0]);
#endif
    }
}

/* compute_mass (geometry.cxx:313-365): element values, then CSR gather */
void FN(compute_mass)(ostate *S, const des_params *P)
{
    const double pseudo_speed = P->max_vbc_val * P->inertial_scaling;
    double *w = S->ework;   /* [e][0]=m, [e][1]=tm */
    #pragma omp parallel for
    for (int e = 0; e < S->nelem; ++e) {
        double rho = P->is_quasi_static ?
            mat_bulkm(S, P, e) / (pseudo_speed * pseudo_speed) : mat_rho(S, P, e);
        w[2 * e] = rho * S->volume[e] / NPE;
        w[2 * e + 1] = mat_rho(S, P, e) * mat_cp(S, P, e) * S->volume[e] / NPE;
    }
    #pragma omp parallel for
    for (int n = 0; n < S->nnode; ++n) {
        double vn = 0, m = 0, tm = 0;
        for (int p = S->csr_off[n]; p < S->csr_off[n + 1]; ++p) {
            int e = S->csr_ent[p] / NPE;
            vn += S->volume[e];
            m += w[2 * e];
            if (P->has_thermal_diffusion) tm += w[2 * e + 1];
        }
        S->volume_n[n] = vn;
        S->mass[n] = m;
        S->tmass[n] = tm;
    }
}

void FN(compute_dvoldt)(ostate *S)                        /* geometry.cxx:111-155 */
{
    #pragma omp parallel for
    for (int n = 0; n < S->nnode; ++n) {
        double s = 0;
        for (int p = S->csr_off[n]; p < S->csr_off[n + 1]; ++p) {
            int e = S->csr_ent[p] / NPE;
            double dj = trace(S->strain_rate + e * NSTR);
            s += dj * S->volume[e];
        }
        S->ntmp[n] = s;
    }
    #pragma omp parallel for
    for (int n = 0; n < S->nnode; ++n) S->ntmp[n] /= S->volume_n[n];
}

void FN(compute_edvoldt)(ostate *S)                       /* geometry.cxx:158-179 */
{
    #pragma omp parallel for
    for (int e = 0; e < S->nelem; ++e) {
        const int *c = S->conn + e * NPE;
        double dj = 0;
        for (int i = 0; i < NPE; ++i) dj += S->ntmp[c[i]];
        S->edvoldt[e] = dj / NPE;
    }
}

void FN(nmd_stress)(ostate *S)                            /* geometry.cxx:182-236 */
{
    #pragma omp parallel for
    for (int n = 0; n < S->nnode; ++n) {
        double s = 0;
        for (int p = S->csr_off[n]; p < S->csr_off[n + 1]; ++p) {
            int e = S->csr_ent[p] / NPE;
            s += S->dpressure[e] * S->volume[e];
        }
        S->ntmp[n] = s;
    }
    #pragma omp parallel for
    for (int n = 0; n < S->nnode; ++n) S->ntmp[n] /= S->volume_n[n];
    #pragma omp parallel for
    for (int e = 0; e < S->nelem; ++e) {
        const int *c = S->conn + e * NPE;
        double dp = 0;
        for (int i = 0; i < NPE; ++i) dp += S->ntmp[c[i]];
        double dp_el = dp / NPE;
        double *s = S->stress + e * NSTR;
        double dp_orig = S->dpressure[e];
        for (int i = 0; i < NDIMS; ++i) s[i] += (-dp_orig + dp_el) / NDIMS;
    }
}

/* compute_dt (geometry.cxx:239-310); returns 0 and sets *status=11 when dt <= 0 */
double FN(compute_dt)(ostate *S, const des_params *P)
{
    if (P->fixed_dt != 0) return P->fixed_dt;
    double dt_maxwell = DBL_MAX, dt_diffusion = DBL_MAX, minl = DBL_MAX;
    for (int e = 0; e < S->nelem; ++e) {
        const int *c = S->conn + e * NPE;
        const double *a = S->coord + NDIMS * c[0];
        const double *b = S->coord + NDIMS * c[1];
        const double *cc = S->coord + NDIMS * c[2];
        double minh;
#if NDIMS == 3
        const double *d = S->coord + NDIMS * c[3];
        double maxa = dmax(dmax(triangle_area(a, b, cc), triangle_area(a, b, d)),
                           dmax(triangle_area(cc, d, a), triangle_area(cc, d, b)));
        minh = 3 * S->volume[e] / maxa;
#else
        double maxl = sqrt(dmax(dmax(dist2(a, b), dist2(b, cc)), dist2(a, cc)));
        minh = 2 * S->volume[e] / maxl;
#endif
        dt_maxwell = dmin(dt_maxwell, 0.5 * P->visc_min / (1e-40 + mat_shearm(S, P, e)));
        if (P->has_thermal_diffusion)
            dt_diffusion = dmin(dt_diffusion, 0.5 * minh * minh / P->therm_diff_max);
        minl = dmin(minl, minh);
    }
    double dt_advection = 0.5 * minl / P->max_vbc_val;
    double dt_elastic = P->is_quasi_static ?
        0.5 * minl / (P->max_vbc_val * P->inertial_scaling) :
        0.5 * minl / sqrt(mat_bulkm(S, P, 0) / mat_rho(S, P, 0));
    double dt = dmin(dmin(dt_elastic, dt_maxwell), dmin(dt_advection, dt_diffusion)) * P->dt_fraction;
    if (!(dt > 0)) S->status = DES_ERR_RUNTIME;
    return dt;
}

double FN(elem_quality)(ostate *S, int e)                 /* geometry.cxx:458-496 */
{
    double vol = S->volume[e];
    const int *c = S->conn + e * NPE;
    const double *a = S->coord + NDIMS * c[0];
    const double *b = S->coord + NDIMS * c[1];
    const double *cc = S->coord + NDIMS * c[2];
#if NDIMS == 3
    const double *d = S->coord + NDIMS * c[3];
    double normalization_factor = 216 * sqrt(3);
    double area_sum = (triangle_area(a, b, cc) + triangle_area(a, b, d) +
                       triangle_area(cc, d, a) + triangle_area(cc, d, b));
    return normalization_factor * vol * vol / (area_sum * area_sum * area_sum);
#else
    double normalization_factor = 4 * sqrt(3);
    double dist2_sum = dist2(a, b) + dist2(b, cc) + dist2(a, cc);
    return normalization_factor * vol / dist2_sum;
#endif
}

double FN(worst_elem_quality)(ostate *S, int *worst_elem) /* geometry.cxx:499-512 */
{
    double q = 1;
    *worst_elem = 0;
    for (int e = 0; e < S->nelem; e++) {
        double quality = FN(elem_quality)(S, e);
        if (quality < q) {
            q = quality;
            *worst_elem = e;
        }
    }
    return q;
}

/* bad_mesh_quality (remeshing.cxx:1882-1934): returns 3, 2, 1 or 0 */
int FN(bad_mesh_quality)(ostate *S, const des_params *P, int remeshing_option,
                         double max_boundary_distortion, int *index)
{
#if NDIMS == 3
    const double sizefactor = 0.118;
#else
    const double sizefactor = 0.433;
#endif
    const double smallest_vol = P->smallest_size * sizefactor * pow(P->resolution, NDIMS);
    for (int e = 0; e < S->nelem; e++) {
        if (S->volume[e] < smallest_vol) {
            *index = e;
            return 3;
        }
    }
    if (remeshing_option == 1 || remeshing_option == 2 || remeshing_option == 11) {
        double bottom = -P->zlength;
        const double dist = max_boundary_distortion * P->resolution;
        for (int i = 0; i < S->nnode; ++i) {
            if (S->bcflag[i] & BOUNDZ0) {
                double z = S->coord[NDIMS * i + NDIMS - 1];
                if (fabs(z - bottom) > dist) {
                    *index = i;
                    return 2;
                }
            }
        }
    }
    int worst_elem;
    double q = FN(worst_elem_quality)(S, &worst_elem);
#if NDIMS == 3
    q = pow(q, 1.0 / 3);
#endif
    if (q < P->min_quality) {
        *index = worst_elem;
        return 1;
    }
    return 0;
}

/* -------------------------------------------------------------- fields.cxx */

void FN(update_temperature)(ostate *S, const des_params *P)   /* fields.cxx:96-163 */
{
    double *w = S->ework;   /* [e][i] contribution to tdot[conn[i]] */
    #pragma omp parallel for
    for (int e = 0; e < S->nelem; ++e) {
        double D[NPE][NPE];
        const int *conn = S->conn + e * NPE;
        double kv = mat_k(S, P, e) * S->volume[e];
        double rh = (mat_hr(S, P, e) / mat_cp(S, P, e)) * S->volume[e];
        const double *sx = S->shpdx + e * NPE, *sz = S->shpdz + e * NPE;
#if NDIMS == 3
        const double *sy = S->shpdy + e * NPE;
#endif
        for (int i = 0; i < NPE; ++i)
            for (int j = 0; j < NPE; ++j) {
#if NDIMS == 3
                D[i][j] = (sx[i] * sx[j] + sy[i] * sy[j] + sz[i] * sz[j]);
#else
                D[i][j] = (sx[i] * sx[j] + sz[i] * sz[j]);
#endif
            }
        for (int i = 0; i < NPE; ++i) {
            double diffusion = 0;
            for (int j = 0; j < NPE; ++j) diffusion += D[i][j] * S->temperature[conn[j]];
            w[e * NPE + i] = diffusion * kv - rh;
        }
    }
    #pragma omp parallel for
    for (int n = 0; n < S->nnode; ++n) {
        double tdot = 0;
        for (int p = S->csr_off[n]; p < S->csr_off[n + 1]; ++p) tdot += w[S->csr_ent[p]];
        S->ntmp[n] = tdot;
    }
    #pragma omp parallel for
    for (int n = 0; n < S->nnode; ++n) {
        if (S->bcflag[n] & BOUNDZ1)
            S->temperature[n] = P->surface_temperature;
        else if ((S->bcflag[n] & BOUNDZ0) && !P->has_fixed_bottom_temperature)
            S->temperature[n] = P->mantle_temperature;
        else
            S->temperature[n] -= S->ntmp[n] * S->dt / S->tmass[n];
    }
}

void FN(update_strain_rate)(ostate *S)                    /* fields.cxx:166-232 */
{
    #pragma omp parallel for
    for (int e = 0; e < S->nelem; ++e) {
        const int *conn = S->conn + e * NPE;
        const double *sx = S->shpdx + e * NPE, *sz = S->shpdz + e * NPE;
        double *s = S->strain_rate + e * NSTR;
        const double *v[NPE];
        for (int i = 0; i < NPE; ++i) v[i] = S->vel + NDIMS * conn[i];
        int n = 0;
        s[n] = 0;
        for (int i = 0; i < NPE; ++i) s[n] += v[i][0] * sx[i];
#if NDIMS == 3
        const double *sy = S->shpdy + e * NPE;
        n = 1;
        s[n] = 0;
        for (int i = 0; i < NPE; ++i) s[n] += v[i][1] * sy[i];
        n = 2;
#else
        n = 1;
#endif
        s[n] = 0;
        for (int i = 0; i < NPE; ++i) s[n] += v[i][NDIMS - 1] * sz[i];
#if NDIMS == 3
        n = 3;
        s[n] = 0;
        for (int i = 0; i < NPE; ++i) s[n] += 0.5 * (v[i][0] * sy[i] + v[i][1] * sx[i]);
        n = 4;
#else
        n = 2;
#endif
        s[n] = 0;
        for (int i = 0; i < NPE; ++i) s[n] += 0.5 * (v[i][0] * sz[i] + v[i][NDIMS - 1] * sx[i]);
#if NDIMS == 3
        n = 5;
        s[n] = 0;
        for (int i = 0; i < NPE; ++i) s[n] += 0.5 * (v[i][1] * sz[i] + v[i][2] * sy[i]);
#endif
    }
}

/* normal_vector_of_facet (bc.cxx:12-47) */
static void normal_vector_of_facet(const ostate *S, int f, const int *conn,
                                   double *normal, double *zcenter)
{
    int n0 = conn[NODE_OF_FACET[f][0]];
    int n1 = conn[NODE_OF_FACET[f][1]];
    const double *x = S->coord;
#if NDIMS == 3
    int n2 = conn[NODE_OF_FACET[f][2]];
    double v01[3], v02[3];
    for (int i = 0; i < 3; ++i) {
        v01[i] = x[3 * n1 + i] - x[3 * n0 + i];
        v02[i] = x[3 * n2 + i] - x[3 * n0 + i];
    }
    normal[0] = (v01[1] * v02[2] - v01[2] * v02[1]) / 2;
    normal[1] = (v01[2] * v02[0] - v01[0] * v02[2]) / 2;
    normal[2] = (v01[0] * v02[1] - v01[1] * v02[0]) / 2;
    *zcenter = (x[3 * n0 + 2] + x[3 * n1 + 2] + x[3 * n2 + 2]) / NPF;
#else
    double v01[2];
    for (int i = 0; i < 2; ++i) v01[i] = x[2 * n1 + i] - x[2 * n0 + i];
    normal[0] = v01[1];
    normal[1] = -v01[0];
    *zcenter = (x[2 * n0 + 1] + x[2 * n1 + 1]) / NPF;
#endif
}

/* apply_stress_bcs (bc.cxx:493-566) */
static void apply_stress_bcs(ostate *S, const des_params *P, double *force)
{
    if (P->gravity == 0) return;
    for (int i = 0; i < DES_NBDRY; i++) {
        int t = P->vbc_types[i];
        if (t != 0 && t != 2 && t != 4) continue;
        if (i == IBOUNDZ0 && !P->has_winkler_foundation) continue;
        if (i == IBOUNDZ1 && !P->has_water_loading) continue;
        for (int k = 0; k < S->nbfacet; ++k) {
            if (S->bf_bdry[k] != i) continue;
            int e = S->bf_elem[k];
            int f = S->bf_face[k];
            const int *conn = S->conn + e * NPE;
            double normal[NDIMS], zcenter;
            normal_vector_of_facet(S, f, conn, normal, &zcenter);
            double p;
            if (i == IBOUNDZ0 && P->has_winkler_foundation) {
                p = S->compensation_pressure -
                    (mat_rho(S, P, e) + P->winkler_delta_rho) * P->gravity * (zcenter + P->zlength);
            } else if (i == IBOUNDZ1 && P->has_water_loading) {
                const double sea_level = 0;
                p = 0;
                if (zcenter < sea_level) {
                    const double sea_water_density = 1030;
                    p = sea_water_density * P->gravity * (sea_level - zcenter);
                }
            } else {
                p = ref_pressure(P, zcenter);
            }
            for (int j = 0; j < NPF; ++j) {
                int n = conn[NODE_OF_FACET[f][j]];
                for (int d = 0; d < NDIMS; ++d) force[NDIMS * n + d] -= p * normal[d] / NPF;
            }
        }
    }
    if (P->has_elastic_foundation) {
        for (int n = 0; n < S->nnode; ++n) {
            if (!(S->bcflag[n] & BOUNDZ0)) continue;
            force[NDIMS * n + NDIMS - 1] -= P->elastic_foundation_constant *
                (S->coord[NDIMS * n + NDIMS - 1] - S->coord0[NDIMS * n + NDIMS - 1]);
        }
    }
}

/* apply_damping (fields.cxx:235-288) */
static void apply_damping(ostate *S, const des_params *P, double *ff)
{
    const double *v = S->vel;
    const double small_vel = 1e-13;
    switch (P->damping_option) {
    case 0:
        break;
    case 1:
        #pragma omp parallel for
        for (int i = 0; i < S->nnode * NDIMS; ++i)
            if (fabs(v[i]) > small_vel) ff[i] -= P->damping_factor * copysign(ff[i], v[i]);
        break;
    case 2:
        #pragma omp parallel for
        for (int i = 0; i < S->nnode * NDIMS; ++i) ff[i] -= P->damping_factor * ff[i];
        break;
    case 3:
        #pragma omp parallel for
        for (int i = 0; i < S->nnode * NDIMS; ++i) {
            if ((ff[i] < 0) == (v[i] < 0))
                ff[i] -= P->damping_factor * ff[i];
            else
                ff[i] += (1 - P->damping_factor) * ff[i];
        }
        break;
    case 4:
        break;
    default:
        S->status = DES_ERR_INPUT;
    }
}

void FN(update_force)(ostate *S, const des_params *P)     /* fields.cxx:291-340 */
{
    double *w = S->ework;   /* [e][i][d] */
    const double gravity = P->gravity;
    #pragma omp parallel for
    for (int e = 0; e < S->nelem; ++e) {
        const double *sx = S->shpdx + e * NPE, *sz = S->shpdz + e * NPE;
#if NDIMS == 3
        const double *sy = S->shpdy + e * NPE;
#endif
        const double *s = S->stress + e * NSTR;
        double vol = S->volume[e];
        double buoy = 0;
        if (gravity != 0) buoy = mat_rho(S, P, e) * gravity / NPE;
        for (int i = 0; i < NPE; ++i) {
            double *f = w + (e * NPE + i) * NDIMS;
#if NDIMS == 3
            f[0] = (s[0] * sx[i] + s[3] * sy[i] + s[4] * sz[i]) * vol;
            f[1] = (s[3] * sx[i] + s[1] * sy[i] + s[5] * sz[i]) * vol;
            f[2] = (s[4] * sx[i] + s[5] * sy[i] + s[2] * sz[i] + buoy) * vol;
#else
            f[0] = (s[0] * sx[i] + s[2] * sz[i]) * vol;
            f[1] = (s[2] * sx[i] + s[1] * sz[i] + buoy) * vol;
#endif
        }
    }
    #pragma omp parallel for
    for (int n = 0; n < S->nnode; ++n) {
        double f[NDIMS];
        for (int d = 0; d < NDIMS; ++d) f[d] = 0;
        for (int p = S->csr_off[n]; p < S->csr_off[n + 1]; ++p) {
            const double *c = w + S->csr_ent[p] * NDIMS;
            for (int d = 0; d < NDIMS; ++d) f[d] -= c[d];
        }
        for (int d = 0; d < NDIMS; ++d) S->force[NDIMS * n + d] = f[d];
    }
    apply_stress_bcs(S, P, S->force);
    if (P->is_quasi_static) apply_damping(S, P, S->force);
}

void FN(update_velocity)(ostate *S)                       /* fields.cxx:343-355 */
{
    #pragma omp parallel for
    for (int i = 0; i < S->nnode * NDIMS; ++i) {
        int n = i / NDIMS;
        S->vel[i] += S->dt * S->force[i] / S->mass[n];
    }
}

void FN(update_coordinate)(ostate *S)                     /* fields.cxx:358-368 */
{
    #pragma omp parallel for
    for (int i = 0; i < S->nnode * NDIMS; ++i) S->coord[i] += S->vel[i] * S->dt;
}

#if NDIMS == 3
static void jaumann_rate_3d(double *s, double dt, double w3, double w4, double w5) /* fields.cxx:375-389 */
{
    double s_inc[NSTR];
    s_inc[0] = -2.0 * s[3] * w3 - 2.0 * s[4] * w4;
    s_inc[1] = 2.0 * s[3] * w3 - 2.0 * s[5] * w5;
    s_inc[2] = 2.0 * s[4] * w4 + 2.0 * s[5] * w5;
    s_inc[3] = s[0] * w3 - s[1] * w3 - s[4] * w5 - s[5] * w4;
    s_inc[4] = s[0] * w4 - s[2] * w4 + s[3] * w5 - s[5] * w3;
    s_inc[5] = s[1] * w5 - s[2] * w5 + s[3] * w4 + s[4] * w3;
    for (int i = 0; i < NSTR; ++i) s[i] += dt * s_inc[i];
}
#else
static void jaumann_rate_2d(double *s, double dt, double w2)  /* fields.cxx:393-404 */
{
    double s_inc[NSTR];
    s_inc[0] = -2.0 * s[2] * w2;
    s_inc[1] = 2.0 * s[2] * w2;
    s_inc[2] = s[0] * w2 - s[1] * w2;
    for (int i = 0; i < NSTR; ++i) s[i] += dt * s_inc[i];
}
#endif

void FN(rotate_stress)(ostate *S)                         /* fields.cxx:411-483 */
{
    #pragma omp parallel for
    for (int e = 0; e < S->nelem; ++e) {
        const int *conn = S->conn + e * NPE;
        const double *sx = S->shpdx + e * NPE, *sz = S->shpdz + e * NPE;
        const double *v[NPE];
        for (int i = 0; i < NPE; ++i) v[i] = S->vel + NDIMS * conn[i];
#if NDIMS == 3
        const double *sy = S->shpdy + e * NPE;
        double w3 = 0, w4 = 0, w5 = 0;
        for (int i = 0; i < NPE; ++i) w3 += 0.5 * (v[i][0] * sy[i] - v[i][1] * sx[i]);
        for (int i = 0; i < NPE; ++i) w4 += 0.5 * (v[i][0] * sz[i] - v[i][NDIMS - 1] * sx[i]);
        for (int i = 0; i < NPE; ++i) w5 += 0.5 * (v[i][1] * sz[i] - v[i][NDIMS - 1] * sy[i]);
        jaumann_rate_3d(S->stress + e * NSTR, S->dt, w3, w4, w5);
        jaumann_rate_3d(S->strain + e * NSTR, S->dt, w3, w4, w5);
#else
        double w2 = 0;
        for (int i = 0; i < NPE; ++i) w2 += 0.5 * (v[i][NDIMS - 1] * sx[i] - v[i][0] * sz[i]);
        jaumann_rate_2d(S->stress + e * NSTR, S->dt, w2);
        jaumann_rate_2d(S->strain + e * NSTR, S->dt, w2);
#endif
    }
}

/* ------------------------------------------------------------------ bc.cxx */

static const double *edge_vec(const ostate *S, int ic, int ib)
{
    if (!S->edge_valid || !S->edge_valid[ic * DES_NBDRY + ib]) return NULL;
    return S->edge_vectors + (ic * DES_NBDRY + ib) * NDIMS;
}

void FN(apply_vbcs)(ostate *S, const des_params *P)       /* bc.cxx:179-490 */
{
    const int *T = P->vbc_types;
    const double *V = P->vbc_values;
    int err = 0;
    #pragma omp parallel for reduction(|:err)
    for (int i = 0; i < S->nnode; ++i) {
        unsigned flag = S->bcflag[i];
        if (!(flag & BOUND_ANY)) continue;
        double *v = S->vel + NDIMS * i;
        if (flag & BOUNDX0) {
            switch (T[0]) {
            case 1: v[0] = V[0]; break;
            case 2: v[1] = 0;
#if NDIMS == 3
                v[2] = 0;
#endif
                break;
            case 3: v[0] = V[0]; v[1] = 0;
#if NDIMS == 3
                v[2] = 0;
#endif
                break;
#if NDIMS == 3
            case 4: v[1] = V[0]; v[2] = 0; break;
            case 5: v[0] = 0; v[1] = V[0]; v[2] = 0; break;
            case 7: v[0] = V[0]; v[1] = 0; break;
#endif
            }
        }
        if (flag & BOUNDX1) {
            switch (T[1]) {
            case 1: v[0] = V[1]; break;
            case 2: v[1] = 0;
#if NDIMS == 3
                v[2] = 0;
#endif
                break;
            case 3: v[0] = V[1]; v[1] = 0;
#if NDIMS == 3
                v[2] = 0;
#endif
                break;
#if NDIMS == 3
            case 4: v[1] = V[1]; v[2] = 0; break;
            case 5: v[0] = 0; v[1] = V[1]; v[2] = 0; break;
            case 7: v[0] = V[1]; v[1] = 0; break;
#endif
            }
        }
#if NDIMS == 3
        if (flag & BOUNDY0) {
            switch (T[2]) {
            case 1: v[1] = V[2]; break;
            case 2: v[0] = 0; v[2] = 0; break;
            case 3: v[0] = 0; v[1] = V[2]; v[2] = 0; break;
            case 4: v[0] = V[2]; v[2] = 0; break;
            case 5: v[0] = V[2]; v[1] = 0; v[2] = 0; break;
            case 7: v[0] = 0; v[1] = V[2]; break;
            }
        }
        if (flag & BOUNDY1) {
            switch (T[3]) {
            case 1: v[1] = V[3]; break;
            case 2: v[0] = 0; v[2] = 0; break;
            /* bc.cxx:325-329: case 3 of the y1 wall sets v[0], as the reference does */
            case 3: v[0] = V[3]; v[1] = 0; v[2] = 0; break;
            case 4: v[0] = V[3]; v[2] = 0; break;
            case 5: v[0] = V[3]; v[1] = 0; v[2] = 0; break;
            case 7: v[0] = 0; v[1] = V[3]; break;
            }
        }
#endif
        for (int ib = IBOUNDN0; ib <= IBOUNDN3; ib++) {
            const double *n = S->bnormals + ib * NDIMS;
            if (!(flag & (1u << ib))) continue;
            double fac = 0;
            switch (T[ib]) {
            case 1:
                if (flag == (1u << ib)) {
                    double vn = 0;
                    for (int d = 0; d < NDIMS; d++) vn += v[d] * n[d];
                    for (int d = 0; d < NDIMS; d++) v[d] += (V[ib] - vn) * n[d];
                } else {
                    for (int ic = 0; ic < ib; ic++) {
                        if (!(flag & (1u << ic))) continue;
                        if (T[ic] == 0) {
                            double vn = 0;
                            for (int d = 0; d < NDIMS; d++) vn += v[d] * n[d];
                            for (int d = 0; d < NDIMS; d++) v[d] += (V[ib] - vn) * n[d];
                        } else if (T[ic] == 1) {
                            const double *edge = edge_vec(S, ic, ib);
                            if (!edge) { err |= 1; continue; }
                            double ve = 0;
                            for (int d = 0; d < NDIMS; d++) ve += v[d] * edge[d];
                            for (int d = 0; d < NDIMS; d++) v[d] = ve * edge[d];
                        }
                    }
                }
                break;
            case 3:
                for (int d = 0; d < NDIMS; d++) v[d] = V[ib] * n[d];
                break;
            case 11:
                fac = 1 / sqrt(1 - n[NDIMS - 1] * n[NDIMS - 1]);
                if (flag == (1u << ib)) {
                    double vn = 0;
                    for (int d = 0; d < NDIMS - 1; d++) vn += v[d] * n[d];
                    for (int d = 0; d < NDIMS - 1; d++) v[d] += (V[ib] * fac - vn) * n[d];
                } else {
                    for (int ic = 0; ic < ib; ic++) {
                        if (!(flag & (1u << ic))) continue;
                        if (T[ic] == 0) {
                            double vn = 0;
                            for (int d = 0; d < NDIMS - 1; d++) vn += v[d] * n[d];
                            for (int d = 0; d < NDIMS - 1; d++) v[d] += (V[ib] * fac - vn) * n[d];
                        } else if (T[ic] == 1) {
                            const double *edge = edge_vec(S, ic, ib);
                            if (!edge) { err |= 1; continue; }
                            double ve = 0;
                            for (int d = 0; d < NDIMS; d++) ve += v[d] * edge[d];
                            for (int d = 0; d < NDIMS; d++) v[d] = ve * edge[d];
                        }
                    }
                }
                break;
            case 13:
                fac = 1 / sqrt(1 - n[NDIMS - 1] * n[NDIMS - 1]);
                for (int d = 0; d < NDIMS - 1; d++) v[d] = V[ib] * fac * n[d];
                v[NDIMS - 1] = 0;
                break;
            }
        }
        if (T[4] == 0 && T[5] == 0) continue;
        if (flag & BOUNDZ0) {
            switch (T[4]) {
            case 1: v[NDIMS - 1] = V[4]; break;
            case 2: v[0] = 0;
#if NDIMS == 3
                v[1] = 0;
#endif
                break;
            case 3: v[0] = 0;
#if NDIMS == 3
                v[1] = 0;
#endif
                v[NDIMS - 1] = V[4];
                break;
            }
        }
        if (flag & BOUNDZ1) {
            switch (T[5]) {
            case 1: v[NDIMS - 1] = V[5]; break;
            case 2: v[0] = 0;
#if NDIMS == 3
                v[1] = 0;
#endif
                break;
            case 3: v[0] = 0;
#if NDIMS == 3
                v[1] = 0;
#endif
                v[NDIMS - 1] = V[5];
                break;
            }
        }
    }
    if (err) S->status = DES_ERR_ASSERT;
}

/* ------------------------------------------------------------------- 3x3-C */
/* The routines of this section restate, step for step, Joachim Kopp's
 * "Numerical diagonalization of 3x3 matrices" (3x3-C: dsyevc3, dsytrd3,
 * dsyevq3, dsyevh3), Copyright (C) 2006 Joachim Kopp, distributed under the
 * GNU Lesser General Public License, version 2.1 or later (3x3-C/COPYING of
 * the reference); this restatement is covered by the same licence.
 * J. Kopp, Int. J. Mod. Phys. C 19 (2008) 523-548. */
#if NDIMS == 3
#define SQR(x) ((x) * (x))

/* Cardano eigenvalues of a symmetric 3x3 (3x3-C/dsyevc3.c:31-81) */
static void dsyevc3(double A[3][3], double w[3])
{
    double m, c1, c0;
    double de = A[0][1] * A[1][2];
    double dd = SQR(A[0][1]);
    double ee = SQR(A[1][2]);
    double ff = SQR(A[0][2]);
    m = A[0][0] + A[1][1] + A[2][2];
    c1 = (A[0][0] * A[1][1] + A[0][0] * A[2][2] + A[1][1] * A[2][2]) - (dd + ee + ff);
    c0 = A[2][2] * dd + A[0][0] * ee + A[1][1] * ff - A[0][0] * A[1][1] * A[2][2]
        - 2.0 * A[0][2] * de;
    double p, sqrt_p, q, c, s, phi;
    p = SQR(m) - 3.0 * c1;
    q = m * (p - (3.0 / 2.0) * c1) - (27.0 / 2.0) * c0;
    sqrt_p = sqrt(fabs(p));
    phi = 27.0 * (0.25 * SQR(c1) * (p - c1) + c0 * (q + 27.0 / 4.0 * c0));
    phi = (1.0 / 3.0) * atan2(sqrt(fabs(phi)), q);
    c = sqrt_p * cos(phi);
    s = (1.0 / 1.73205080756887729352744634151) * sqrt_p * sin(phi);
    w[1] = (1.0 / 3.0) * (m - c);
    w[2] = w[1] + s;
    w[0] = w[1] + c;
    w[1] -= s;
}

/* Householder reduction to tridiagonal form (3x3-C/dsytrd3.c:28-104) */
static void dsytrd3(double A[3][3], double Q[3][3], double d[3], double e[2])
{
    const int n = 3;
    double u[3], q[3];
    double omega, f, K, h, g;
    for (int i = 0; i < n; i++) {
        Q[i][i] = 1.0;
        for (int j = 0; j < i; j++) Q[i][j] = Q[j][i] = 0.0;
    }
    h = SQR(A[0][1]) + SQR(A[0][2]);
    if (A[0][1] > 0)
        g = -sqrt(h);
    else
        g = sqrt(h);
    e[0] = g;
    f = g * A[0][1];
    u[1] = A[0][1] - g;
    u[2] = A[0][2];
    omega = h - f;
    if (omega > 0.0) {
        omega = 1.0 / omega;
        K = 0.0;
        for (int i = 1; i < n; i++) {
            f = A[1][i] * u[1] + A[i][2] * u[2];
            q[i] = omega * f;
            K += u[i] * f;
        }
        K *= 0.5 * SQR(omega);
        for (int i = 1; i < n; i++) q[i] = q[i] - K * u[i];
        d[0] = A[0][0];
        d[1] = A[1][1] - 2.0 * q[1] * u[1];
        d[2] = A[2][2] - 2.0 * q[2] * u[2];
        for (int j = 1; j < n; j++) {
            f = omega * u[j];
            for (int i = 1; i < n; i++) Q[i][j] = Q[i][j] - f * u[i];
        }
        e[1] = A[1][2] - q[1] * u[2] - u[1] * q[2];
    } else {
        for (int i = 0; i < n; i++) d[i] = A[i][i];
        e[1] = A[1][2];
    }
}

/* QL with implicit shifts (3x3-C/dsyevq3.c:29-134) */
static int dsyevq3(double A[3][3], double Q[3][3], double w[3])
{
    const int n = 3;
    double e[3];
    double g, r, p, f, b, s, c, t;
    int nIter, m;
    dsytrd3(A, Q, w, e);
    for (int l = 0; l < n - 1; l++) {
        nIter = 0;
        while (1) {
            for (m = l; m <= n - 2; m++) {
                g = fabs(w[m]) + fabs(w[m + 1]);
                if (fabs(e[m]) + g == g) break;
            }
            if (m == l) break;
            if (nIter++ >= 30) return -1;
            g = (w[l + 1] - w[l]) / (e[l] + e[l]);
            r = sqrt(SQR(g) + 1.0);
            if (g > 0)
                g = w[m] - w[l] + e[l] / (g + r);
            else
                g = w[m] - w[l] + e[l] / (g - r);
            s = c = 1.0;
            p = 0.0;
            for (int i = m - 1; i >= l; i--) {
                f = s * e[i];
                b = c * e[i];
                if (fabs(f) > fabs(g)) {
                    c = g / f;
                    r = sqrt(SQR(c) + 1.0);
                    e[i + 1] = f * r;
                    c *= (s = 1.0 / r);
                } else {
                    s = f / g;
                    r = sqrt(SQR(s) + 1.0);
                    e[i + 1] = g * r;
                    s *= (c = 1.0 / r);
                }
                g = w[i + 1] - p;
                r = (w[i] - g) * s + 2.0 * c * b;
                p = s * r;
                w[i + 1] = g + p;
                g = c * r - b;
                for (int k = 0; k < n; k++) {
                    t = Q[k][i + 1];
                    Q[k][i + 1] = s * Q[k][i] + c * t;
                    Q[k][i] = c * Q[k][i] - s * t;
                }
            }
            w[l] -= p;
            e[l] = g;
            e[m] = 0.0;
        }
    }
    return 0;
}

/* Cardano + cross-product eigenvectors with QL fallback (3x3-C/dsyevh3.c:31-134) */
static int dsyevh3(double A[3][3], double Q[3][3], double w[3])
{
    double norm, error, t, u;
    dsyevc3(A, w);
    t = fabs(w[0]);
    if ((u = fabs(w[1])) > t) t = u;
    if ((u = fabs(w[2])) > t) t = u;
    if (t < 1.0)
        u = t;
    else
        u = SQR(t);
    error = 256.0 * DBL_EPSILON * SQR(u);
    Q[0][1] = A[0][1] * A[1][2] - A[0][2] * A[1][1];
    Q[1][1] = A[0][2] * A[0][1] - A[1][2] * A[0][0];
    Q[2][1] = SQR(A[0][1]);
    Q[0][0] = Q[0][1] + A[0][2] * w[0];
    Q[1][0] = Q[1][1] + A[1][2] * w[0];
    Q[2][0] = (A[0][0] - w[0]) * (A[1][1] - w[0]) - Q[2][1];
    norm = SQR(Q[0][0]) + SQR(Q[1][0]) + SQR(Q[2][0]);
    if (norm <= error) return dsyevq3(A, Q, w);
    norm = sqrt(1.0 / norm);
    for (int j = 0; j < 3; j++) Q[j][0] = Q[j][0] * norm;
    Q[0][1] = Q[0][1] + A[0][2] * w[1];
    Q[1][1] = Q[1][1] + A[1][2] * w[1];
    Q[2][1] = (A[0][0] - w[1]) * (A[1][1] - w[1]) - Q[2][1];
    norm = SQR(Q[0][1]) + SQR(Q[1][1]) + SQR(Q[2][1]);
    if (norm <= error) return dsyevq3(A, Q, w);
    norm = sqrt(1.0 / norm);
    for (int j = 0; j < 3; j++) Q[j][1] = Q[j][1] * norm;
    Q[0][2] = Q[1][0] * Q[2][1] - Q[2][0] * Q[1][1];
    Q[1][2] = Q[2][0] * Q[0][1] - Q[0][0] * Q[2][1];
    Q[2][2] = Q[0][0] * Q[1][1] - Q[1][0] * Q[0][1];
    return 0;
}

/* principal_stresses3 (rheology.cxx:13-68): eigenpairs sorted ascending */
static void principal_stresses3(const double *s, double p[3], double v[3][3])
{
    double a[3][3];
    a[0][0] = s[0]; a[1][1] = s[1]; a[2][2] = s[2];
    a[0][1] = s[3]; a[0][2] = s[4]; a[1][2] = s[5];
    a[1][0] = a[2][0] = a[2][1] = 0;
    dsyevh3(a, v, p);
    for (int pass = 0; pass < 3; ++pass) {
        int i = (pass == 1) ? 1 : 0;   /* swap order (0,1), (1,2), (0,1) */
        if (p[i] > p[i + 1]) {
            double tmp = p[i];
            p[i] = p[i + 1];
            p[i + 1] = tmp;
            for (int k = 0; k < 3; ++k) {
                double b = v[k][i];
                v[k][i] = v[k][i + 1];
                v[k][i + 1] = b;
            }
        }
    }
}

/* test hooks: expose the eigen solver for known-answer tests */
int o3_dsyevh3(const double *a6, double *w, double *q9)
{
    double A[3][3] = {{a6[0], a6[3], a6[4]}, {0, a6[1], a6[5]}, {0, 0, a6[2]}};
    double Q[3][3];
    int r = dsyevh3(A, Q, w);
    memcpy(q9, Q, sizeof Q);
    return r;
}

void o3_principal_stresses3(const double *s, double *p, double *v9)
{
    double v[3][3];
    principal_stresses3(s, p, v);
    memcpy(v9, v, sizeof v);
}
#else
/* principal_stresses2 (rheology.cxx:71-102) */
static void principal_stresses2(const double *s, double p[2], double *cos2t, double *sin2t)
{
    double s0 = 0.5 * (s[0] + s[1]);
    double rad = second_invariant(s);
    p[0] = s0 - rad;
    p[1] = s0 + rad;
    const double eps = 1e-15;
    double a = 0.5 * (s[0] - s[1]);
    double b = -rad;
    if (b < -eps) {
        *cos2t = a / b;
        *sin2t = s[2] / b;
    } else {
        *cos2t = 1;
        *sin2t = 0;
    }
}
#endif

/* ----------------------------------------------------------- rheology.cxx */

static void elastic(double bulkm, double shearm, const double *de, double *s)  /* rheology.cxx:105-115 */
{
    double lambda = bulkm - 2. / 3 * shearm;
    double dev = trace(de);
    for (int i = 0; i < NDIMS; ++i) s[i] += 2 * shearm * de[i] + lambda * dev;
    for (int i = NDIMS; i < NSTR; ++i) s[i] += 2 * shearm * de[i];
}

static void maxwell(double bulkm, double shearm, double viscosity, double dt,
                    double dv, const double *de, double *s)        /* rheology.cxx:118-134 */
{
    double tmp = 0.5 * dt * shearm / viscosity;
    double f1 = 1 - tmp;
    double f2 = 1 / (1 + tmp);
    double dev = trace(de) / NDIMS;
    double s0 = trace(s) / NDIMS;
    for (int i = 0; i < NDIMS; ++i)
        s[i] = ((s[i] - s0) * f1 + 2 * shearm * (de[i] - dev)) * f2 + s0 + bulkm * dv;
    for (int i = NDIMS; i < NSTR; ++i)
        s[i] = (s[i] * f1 + 2 * shearm * de[i]) * f2;
}

static void viscous(double bulkm, double viscosity, double total_dv,
                    const double *edot, double *s)                 /* rheology.cxx:137-148 */
{
    double dev = trace(edot) / NDIMS;
    for (int i = 0; i < NDIMS; ++i) s[i] = 2 * viscosity * (edot[i] - dev) + bulkm * total_dv;
    for (int i = NDIMS; i < NSTR; ++i) s[i] = 2 * viscosity * edot[i];
}

/* elasto_plastic, Mohr-Coulomb return in principal space (rheology.cxx:151-299) */
static void elasto_plastic(double bulkm, double shearm, double amc, double anphi, double anpsi,
                           double hardn, double ten_max, const double *de, double *depls,
                           double *s, int *failure_mode)
{
    elastic(bulkm, shearm, de, s);
    *depls = 0;
    *failure_mode = 0;
    double p[NDIMS];
#if NDIMS == 3
    double v[3][3];
    principal_stresses3(s, p, v);
#else
    double cos2t, sin2t;
    principal_stresses2(s, p, &cos2t, &sin2t);
#endif
    double fs = p[0] - p[NDIMS - 1] * anphi + amc;
    double ft = p[NDIMS - 1] - ten_max;
    if (fs > 0 && ft < 0) return;
    double pa = sqrt(1 + anphi * anphi) + anphi;
    double ps = ten_max * anphi - amc;
    double h = p[NDIMS - 1] - ten_max + pa * (p[0] - ps);
    double a1 = bulkm + 4. / 3 * shearm;
    double a2 = bulkm - 2. / 3 * shearm;
    double alam;
    if (h < 0) {
        *failure_mode = 10;
        alam = fs / (a1 - a2 * anpsi + a1 * anphi * anpsi - a2 * anphi + 2 * sqrt(anphi) * hardn);
        p[0] -= alam * (a1 - a2 * anpsi);
#if NDIMS == 3
        p[1] -= alam * (a2 - a2 * anpsi);
#endif
        p[NDIMS - 1] -= alam * (a2 - a1 * anpsi);
#if NDIMS == 3
        *depls = fabs(alam) * sqrt((7 + 4 * anpsi + 7 * anpsi * anpsi) / 18);
#else
        *depls = fabs(alam) * sqrt((3 + 2 * anpsi + 3 * anpsi * anpsi) / 8);
#endif
    } else {
        *failure_mode = 1;
        alam = ft / a1;
        p[0] -= alam * a2;
#if NDIMS == 3
        p[1] -= alam * a2;
#endif
        p[NDIMS - 1] -= alam * a1;
#if NDIMS == 3
        *depls = fabs(alam) * sqrt(7. / 18);
#else
        *depls = fabs(alam) * sqrt(3. / 8);
#endif
    }
#if NDIMS == 3
    double ss[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
    for (int m = 0; m < 3; m++)
        for (int n = m; n < 3; n++)
            for (int k = 0; k < 3; k++) ss[m][n] += v[m][k] * v[n][k] * p[k];
    s[0] = ss[0][0];
    s[1] = ss[1][1];
    s[2] = ss[2][2];
    s[3] = ss[0][1];
    s[4] = ss[0][2];
    s[5] = ss[1][2];
#else
    double dc2 = (p[0] - p[1]) * cos2t;
    double dss = p[0] + p[1];
    s[0] = 0.5 * (dss + dc2);
    s[1] = 0.5 * (dss - dc2);
    s[2] = 0.5 * (p[0] - p[1]) * sin2t;
#endif
}

#if NDIMS == 2
/* geoFLAC-style plane-strain Mohr-Coulomb (rheology.cxx:302-507) */
static void elasto_plastic2d(double bulkm, double shearm, double amc, double anphi, double anpsi,
                             double hardn, double ten_max, const double *de, double *depls,
                             double *s, double *syy, int *failure_mode)
{
    *depls = 0;
    *failure_mode = 0;
    double a1 = bulkm + 4. / 3 * shearm;
    double a2 = bulkm - 2. / 3 * shearm;
    double sxx = s[0] + de[1] * a2 + de[0] * a1;
    double szz = s[1] + de[0] * a2 + de[1] * a1;
    double sxz = s[2] + de[2] * 2 * shearm;
    *syy += (de[0] + de[1]) * a2;
    double p[3];
    double cos2t, sin2t;
    int n1, n2, n3;
    {
        double s0 = 0.5 * (sxx + szz);
        double rad = 0.5 * sqrt((sxx - szz) * (sxx - szz) + 4 * sxz * sxz);
        double si = s0 - rad;
        double sii = s0 + rad;
        const double eps = 1e-15;
        if (rad > eps) {
            cos2t = 0.5 * (szz - sxx) / rad;
            sin2t = -sxz / rad;
        } else {
            cos2t = 1;
            sin2t = 0;
        }
        if (*syy > sii) {
            n1 = 0; n2 = 1; n3 = 2;
            p[0] = si; p[1] = sii; p[2] = *syy;
        } else if (*syy < si) {
            n1 = 1; n2 = 2; n3 = 0;
            p[0] = *syy; p[1] = si; p[2] = sii;
        } else {
            n1 = 0; n2 = 2; n3 = 1;
            p[0] = si; p[1] = *syy; p[2] = sii;
        }
    }
    if (p[0] >= ten_max) {
        s[0] = s[1] = *syy = ten_max;
        s[2] = 0.0;
        *failure_mode = 1;
        return;
    }
    if (p[1] >= ten_max) {
        p[1] = p[2] = ten_max;
        *failure_mode = 2;
    } else if (p[2] >= ten_max) {
        p[2] = ten_max;
        *failure_mode = 3;
    }
    double fs = p[0] - p[2] * anphi + amc;
    if (fs >= 0.0) {
        s[0] = sxx;
        s[1] = szz;
        s[2] = sxz;
        return;
    }
    *failure_mode += 10;
    const double alams = fs / (a1 - a2 * anpsi + a1 * anphi * anpsi - a2 * anphi + hardn);
    p[0] -= alams * (a1 - a2 * anpsi);
    p[1] -= alams * (a2 - a2 * anpsi);
    p[2] -= alams * (a2 - a1 * anpsi);
    *depls = 0.5 * fabs(alams + alams * anpsi);
    if (p[0] >= ten_max) {
        s[0] = s[1] = *syy = ten_max;
        s[2] = 0.0;
        *failure_mode += 20;
        return;
    }
    if (p[1] >= ten_max) {
        p[1] = p[2] = ten_max;
        *failure_mode += 20;
    } else if (p[2] >= ten_max) {
        p[2] = ten_max;
        *failure_mode += 20;
    }
    double dc2 = (p[n1] - p[n2]) * cos2t;
    double dss = p[n1] + p[n2];
    s[0] = 0.5 * (dss + dc2);
    s[1] = 0.5 * (dss - dc2);
    s[2] = 0.5 * (p[n1] - p[n2]) * sin2t;
    *syy = p[n3];
}
#endif

static int plane_strain_return(double bulkm, double shearm, double amc, double anphi,
                               double anpsi, double hardn, double ten_max, const double *de,
                               double *depls, double *s, double *syy, int is_plane_strain)
{
    int failure_mode;
#if NDIMS == 2
    if (is_plane_strain) {
        elasto_plastic2d(bulkm, shearm, amc, anphi, anpsi, hardn, ten_max, de, depls, s, syy,
                         &failure_mode);
        return failure_mode;
    }
#else
    (void)syy;
    (void)is_plane_strain;
#endif
    elasto_plastic(bulkm, shearm, amc, anphi, anpsi, hardn, ten_max, de, depls, s, &failure_mode);
    return failure_mode;
}

/* update_stress (rheology.cxx:510-646) */
void FN(update_stress)(ostate *S, const des_params *P)
{
    const int rheol_type = P->rheol_type;
    int bad = 0;
    #pragma omp parallel for reduction(|:bad)
    for (int e = 0; e < S->nelem; ++e) {
        double *s = S->stress + e * NSTR;
        double *syy = S->stressyy + e;
        double *es = S->strain + e * NSTR;
        double *edot = S->strain_rate + e * NSTR;
        double old_s = trace(s);
        {
            double div = trace(edot);
            for (int i = 0; i < NDIMS; ++i) edot[i] += (S->edvoldt[e] - div) / NDIMS;
        }
        for (int i = 0; i < NSTR; ++i) es[i] += edot[i] * S->dt;
        double de[NSTR];
        for (int i = 0; i < NSTR; ++i) de[i] = edot[i] * S->dt;
        switch (rheol_type) {
        case DES_RH_ELASTIC:
            elastic(mat_bulkm(S, P, e), mat_shearm(S, P, e), de, s);
            break;
        case DES_RH_VISCOUS: {
            double bulkm = mat_bulkm(S, P, e);
            double viscosity = mat_visc(S, P, e);
            double total_dv = trace(es);
            viscous(bulkm, viscosity, total_dv, edot, s);
            break;
        }
        case DES_RH_MAXWELL: {
            double bulkm = mat_bulkm(S, P, e);
            double shearm = mat_shearm(S, P, e);
            double viscosity = mat_visc(S, P, e);
            double dv = S->volume[e] / S->volume_old[e] - 1;
            maxwell(bulkm, shearm, viscosity, S->dt, dv, de, s);
            break;
        }
        case DES_RH_EP: {
            double depls = 0;
            double bulkm = mat_bulkm(S, P, e);
            double shearm = mat_shearm(S, P, e);
            double amc, anphi, anpsi, hardn, ten_max;
            mat_plastic_props(S, P, e, S->plstrain[e], &amc, &anphi, &anpsi, &hardn, &ten_max);
            plane_strain_return(bulkm, shearm, amc, anphi, anpsi, hardn, ten_max, de, &depls, s,
                                syy, P->is_plane_strain);
            S->plstrain[e] += depls;
            S->delta_plstrain[e] = depls;
            break;
        }
        case DES_RH_EVP: {
            double depls = 0;
            double bulkm = mat_bulkm(S, P, e);
            double shearm = mat_shearm(S, P, e);
            double viscosity = mat_visc(S, P, e);
            double dv = S->volume[e] / S->volume_old[e] - 1;
            double sv[NSTR];
            for (int i = 0; i < NSTR; ++i) sv[i] = s[i];
            maxwell(bulkm, shearm, viscosity, S->dt, dv, de, sv);
            double svII = second_invariant2(sv);
            double amc, anphi, anpsi, hardn, ten_max;
            mat_plastic_props(S, P, e, S->plstrain[e], &amc, &anphi, &anpsi, &hardn, &ten_max);
            double sp[NSTR], spyy = *syy;
            for (int i = 0; i < NSTR; ++i) sp[i] = s[i];
            plane_strain_return(bulkm, shearm, amc, anphi, anpsi, hardn, ten_max, de, &depls, sp,
                                &spyy, P->is_plane_strain);
            double spII = second_invariant2(sp);
            if (svII < spII) {
                for (int i = 0; i < NSTR; ++i) s[i] = sv[i];
            } else {
                for (int i = 0; i < NSTR; ++i) s[i] = sp[i];
                S->plstrain[e] += depls;
                S->delta_plstrain[e] = depls;
                /* rheology.cxx:632 copies spyy, which only plane strain defines */
                if (P->is_plane_strain) *syy = spyy;
            }
            break;
        }
        default:
            bad = 1;
        }
        S->dpressure[e] = trace(s) - old_s;
    }
    if (bad) S->status = DES_ERR_INPUT;
}

/* --------------------------------------------------------- dynearthsol.cxx */

/* init()'s derived-state block (dynearthsol.cxx:86-91) */
void FN(init_derived)(ostate *S, const des_params *P)
{
    FN(compute_volume)(S);
    memcpy(S->volume_old, S->volume, sizeof(double) * S->nelem);
    FN(compute_mass)(S, P);
    FN(compute_shape_fn)(S);
}

/* update_mesh (dynearthsol.cxx:221-232), surface_process_option 0 only */
void FN(update_mesh)(ostate *S, const des_params *P)
{
    FN(update_coordinate)(S);
    if (P->surface_process_option != 0) S->status = DES_ERR_INPUT;
    double *t = S->volume;
    S->volume = S->volume_old;
    S->volume_old = t;
    FN(compute_volume)(S);
    FN(compute_mass)(S, P);
    FN(compute_shape_fn)(S);
}

/* One pass of the main loop body (dynearthsol.cxx:325-361) without phase
   changes, markers, output and remeshing.  Returns the status. */
int FN(step)(ostate *S, const des_params *P)
{
    S->steps++;
    S->time += S->dt;
    if (P->has_thermal_diffusion) FN(update_temperature)(S, P);
    FN(update_strain_rate)(S);
    FN(compute_dvoldt)(S);
    FN(compute_edvoldt)(S);
    FN(update_stress)(S, P);
    FN(nmd_stress)(S);
    FN(update_force)(S, P);
    FN(update_velocity)(S);
    FN(apply_vbcs)(S, P);
    FN(update_mesh)(S, P);
    if (P->rheol_type & DES_RH_ELASTIC) FN(rotate_stress)(S);
    if (S->steps % 10 == 0) S->dt = FN(compute_dt)(S, P);
    return S->status;
}

int FN(run)(ostate *S, const des_params *P, int nsteps)
{
    for (int k = 0; k < nsteps; ++k)
        if (FN(step)(S, P)) break;
    return S->status;
}

/* create_support (mesh.cxx:1707-1721) as a CSR: ascending elements per node */
int FN(build_support)(int nnode, int nelem, const int *conn, int *off, int *ent)
{
    memset(off, 0, sizeof(int) * (nnode + 1));
    for (int k = 0; k < nelem * NPE; ++k) {
        int n = conn[k];
        if (n < 0 || n >= nnode) return DES_ERR_MESH;
        off[n + 1]++;
    }
    for (int n = 0; n < nnode; ++n) off[n + 1] += off[n];
    int *fill = (int *)malloc(sizeof(int) * nnode);
    memcpy(fill, off, sizeof(int) * nnode);
    for (int e = 0; e < nelem; ++e)
        for (int i = 0; i < NPE; ++i) ent[fill[conn[e * NPE + i]]++] = e * NPE + i;
    free(fill);
    return 0;
}

/* test hooks for the material functions */
double FN(mat_visc)(ostate *S, const des_params *P, int e) { return mat_visc(S, P, e); }
double FN(mat_rho)(ostate *S, const des_params *P, int e) { return mat_rho(S, P, e); }
void FN(mat_plastic_props)(ostate *S, const des_params *P, int e, double pls, double *out5)
{
    mat_plastic_props(S, P, e, pls, out5, out5 + 1, out5 + 2, out5 + 3, out5 + 4);
}
double FN(ref_pressure)(const des_params *P, double z) { return ref_pressure(P, z); }
int FN(sizeof_state)(void) { return (int)sizeof(ostate); }
