/*
 * des_adjoint.c -- FP64 CPU reference gradient of the 3D explicit time step.
 *
 * TEST INFRASTRUCTURE ONLY: the checker of the HIP adjoint.  The reference has
 * no adjoint; this file differentiates the oracle's step (des_oracle.c, which
 * it includes for the material and eigen helpers) by hand, in reverse mode.
 * Its own pin is a finite-difference check against the oracle's forward step
 * (tests/test_adjoint_cpu.py).
 *
 * The map that is differentiated is one step of dynearthsol.cxx:325-361 in
 * the reference's order, for the configurations the adjoint supports
 * (adj3_supported): 3D, one material, constant density (alpha = 0, so
 * temperature does not feed back into the mechanics), per-element viscosity
 * (elem_visc), fixed dt, damping option 0 or 1, no traction-loaded boundary
 * facets, no elastic foundation, axis-aligned velocity boundaries, rheology
 * elastic / maxwell / elasto-plastic / elasto-visco-plastic.
 *
 * State carried from step to step: coordinates X, velocity V, stress S (after
 * the Jaumann rotation), plastic strain P and volume_old VO = vol(X of the
 * previous step).  The parameters are the per-element viscosity values and
 * the per-element cohesion (elem_cohesion).
 */
#define NDIMS 3
#include "des_oracle.c"

#include <stdio.h>

static void cross3(const double *a, const double *b, double *c)
{
    c[0] = a[1] * b[2] - a[2] * b[1];
    c[1] = a[2] * b[0] - a[0] * b[2];
    c[2] = a[0] * b[1] - a[1] * b[0];
}

/* G[i] = d vol / d x_i for vol = det[x1-x0, x2-x0, x3-x0] / 6, which is the
   reference's tetrahedron_volume (geometry.cxx:29-49); the reference's shape
   function derivatives are G[i] / vol (geometry.cxx:400-436). */
static void tet_grad(const double x[4][3], double G[4][3])
{
    double e1[3], e2[3], e3[3];
    for (int d = 0; d < 3; ++d) {
        e1[d] = x[1][d] - x[0][d];
        e2[d] = x[2][d] - x[0][d];
        e3[d] = x[3][d] - x[0][d];
    }
    cross3(e2, e3, G[1]);
    cross3(e3, e1, G[2]);
    cross3(e1, e2, G[3]);
    for (int d = 0; d < 3; ++d) {
        G[1][d] /= 6;
        G[2][d] /= 6;
        G[3][d] /= 6;
        G[0][d] = -(G[1][d] + G[2][d] + G[3][d]);
    }
}

/* xb += d/dx of sum_i Gb[i] . G[i](x)  (the Hessian of vol applied to Gb) */
static void tet_grad_adj(const double x[4][3], const double Gb[4][3], double xb[4][3])
{
    double e1[3], e2[3], e3[3], g1[3], g2[3], g3[3];
    for (int d = 0; d < 3; ++d) {
        e1[d] = x[1][d] - x[0][d];
        e2[d] = x[2][d] - x[0][d];
        e3[d] = x[3][d] - x[0][d];
        g1[d] = Gb[1][d] - Gb[0][d];
        g2[d] = Gb[2][d] - Gb[0][d];
        g3[d] = Gb[3][d] - Gb[0][d];
    }
    double a[3], b[3], eb[3][3];
    cross3(g2, e3, a);
    cross3(e2, g3, b);
    for (int d = 0; d < 3; ++d) eb[0][d] = (a[d] + b[d]) / 6;
    cross3(e3, g1, a);
    cross3(g3, e1, b);
    for (int d = 0; d < 3; ++d) eb[1][d] = (a[d] + b[d]) / 6;
    cross3(g1, e2, a);
    cross3(e1, g2, b);
    for (int d = 0; d < 3; ++d) eb[2][d] = (a[d] + b[d]) / 6;
    for (int d = 0; d < 3; ++d) {
        xb[1][d] += eb[0][d];
        xb[2][d] += eb[1][d];
        xb[3][d] += eb[2][d];
        xb[0][d] -= eb[0][d] + eb[1][d] + eb[2][d];
    }
}

static double tet_vol(const double x[4][3])
{
    double G[4][3];
    tet_grad(x, G);
    /* vol = G[1] . (x1 - x0) / 3 ... use the determinant directly */
    double e1[3], e2[3], e3[3], c[3];
    for (int d = 0; d < 3; ++d) {
        e1[d] = x[1][d] - x[0][d];
        e2[d] = x[2][d] - x[0][d];
        e3[d] = x[3][d] - x[0][d];
    }
    cross3(e2, e3, c);
    (void)G;
    return (e1[0] * c[0] + e1[1] * c[1] + e1[2] * c[2]) / 6;
}

static void gather4(const double *a, const int *c, double x[4][3])
{
    for (int i = 0; i < 4; ++i)
        for (int d = 0; d < 3; ++d) x[i][d] = a[3 * c[i] + d];
}

/* ------------------------------------------------------------ support check */

int adj3_supported(const ostate *S, const des_params *P)
{
    if (P->ndims != 3 || P->nmat != 1 || S->nmat != 1) return 0;
    if (P->alpha[0] != 0) return 0;
    if (P->fixed_dt == 0) return 0;
    if (P->damping_option != 0 && P->damping_option != 1) return 0;
    if (P->has_winkler_foundation || P->has_water_loading || P->has_elastic_foundation) return 0;
    for (int i = 0; i < 4; ++i) {                          /* lithostatic side walls */
        int t = P->vbc_types[i];
        if (P->gravity != 0 && (t == 0 || t == 2 || t == 4)) return 0;
    }
    for (int n = 0; n < S->nnode; ++n)                      /* slant boundaries */
        if (S->bcflag[n] & (64u | 128u | 256u | 512u)) return 0;
    int rh = P->rheol_type;
    if (rh != DES_RH_ELASTIC && rh != DES_RH_MAXWELL && rh != DES_RH_EP && rh != DES_RH_EVP) return 0;
    if ((rh == DES_RH_MAXWELL || rh == DES_RH_EVP) && !S->elem_visc) return 0;
    return 1;
}

/* ------------------------------------------------------ rheology adjoints */

/* Maxwell (rheology.cxx:118-134): s -> out.  Given ob = d/d out, adds
   d/d s, d/d de, d/d dv and d/d viscosity. */
static void maxwell_adj(double bulkm, double shearm, double visc, double dt, const double *de,
                        const double *s, const double *ob, double *sb, double *deb, double *dvb,
                        double *viscb)
{
    double tmp = 0.5 * dt * shearm / visc;
    double f1 = 1 - tmp, f2 = 1 / (1 + tmp);
    double dev = (de[0] + de[1] + de[2]) / 3, s0 = (s[0] + s[1] + s[2]) / 3;
    double A = (ob[0] + ob[1] + ob[2]) / 3;
    for (int j = 0; j < 3; ++j) {
        sb[j] += f1 * f2 * (ob[j] - A) + A;
        deb[j] += 2 * shearm * f2 * (ob[j] - A);
    }
    for (int j = 3; j < 6; ++j) {
        sb[j] += f1 * f2 * ob[j];
        deb[j] += 2 * shearm * f2 * ob[j];
    }
    *dvb += bulkm * 3 * A;
    double f1b = 0, f2b = 0;
    for (int i = 0; i < 3; ++i) {
        f1b += ob[i] * (s[i] - s0) * f2;
        f2b += ob[i] * ((s[i] - s0) * f1 + 2 * shearm * (de[i] - dev));
    }
    for (int i = 3; i < 6; ++i) {
        f1b += ob[i] * s[i] * f2;
        f2b += ob[i] * (s[i] * f1 + 2 * shearm * de[i]);
    }
    double tmpb = -f1b - f2b * f2 * f2;
    *viscb += tmpb * (-tmp / visc);
}

/* elastic (rheology.cxx:105-115) */
static void elastic_adj(double bulkm, double shearm, const double *ob, double *sb, double *deb)
{
    double lambda = bulkm - 2. / 3 * shearm;
    double tr = ob[0] + ob[1] + ob[2];
    for (int i = 0; i < 3; ++i) {
        sb[i] += ob[i];
        deb[i] += 2 * shearm * ob[i] + lambda * tr;
    }
    for (int i = 3; i < 6; ++i) {
        sb[i] += ob[i];
        deb[i] += 2 * shearm * ob[i];
    }
}

/* adjoints of the plastic properties (amc, anphi, anpsi, hardn, ten_max) */
typedef struct { double amc, anphi, anpsi, hardn, ten_max; } pprops;

/* mat_plastic_props for one material and its derivatives: d/d pls and d/d c0 */
static void plastic_props_d(const des_params *P, double c0, double pls, pprops *v, pprops *dpls,
                            pprops *dc0)
{
    double c1 = P->cohesion1[0], p0 = P->pls0[0], p1 = P->pls1[0];
    double C, phi, psi, h, Cp = 0, phip = 0, psip = 0, Cc = 0, hc = 0;
    if (pls < p0) {
        C = c0; phi = P->friction_angle0[0]; psi = P->dilation_angle0[0]; h = 0;
        Cc = 1;
    } else if (pls < p1) {
        double p = (pls - p0) / (p1 - p0);
        C = c0 + p * (c1 - c0);
        phi = P->friction_angle0[0] + p * (P->friction_angle1[0] - P->friction_angle0[0]);
        psi = P->dilation_angle0[0] + p * (P->dilation_angle1[0] - P->dilation_angle0[0]);
        h = (c1 - c0) / (p1 - p0);
        Cp = (c1 - c0) / (p1 - p0);
        phip = (P->friction_angle1[0] - P->friction_angle0[0]) / (p1 - p0);
        psip = (P->dilation_angle1[0] - P->dilation_angle0[0]) / (p1 - p0);
        Cc = 1 - p;
        hc = -1 / (p1 - p0);
    } else {
        C = c1; phi = P->friction_angle1[0]; psi = P->dilation_angle1[0]; h = 0;
    }
    double sphi = sin(phi * DEG2RAD), spsi = sin(psi * DEG2RAD);
    double anphi = (1 + sphi) / (1 - sphi), anpsi = (1 + spsi) / (1 - spsi);
    double danphi = 2 * cos(phi * DEG2RAD) * DEG2RAD / ((1 - sphi) * (1 - sphi));   /* d/d phi */
    double danpsi = 2 * cos(psi * DEG2RAD) * DEG2RAD / ((1 - spsi) * (1 - spsi));
    v->anphi = anphi;
    v->anpsi = anpsi;
    v->hardn = h;
    v->amc = 2 * C * sqrt(anphi);
    double tC = 0, tphi = 0;
    if (phi == 0) {
        v->ten_max = P->tension_max;
    } else {
        double t = C / tan(phi * DEG2RAD);
        if (P->tension_max < t) {
            v->ten_max = P->tension_max;
        } else {
            v->ten_max = t;
            tC = 1 / tan(phi * DEG2RAD);
            double sn = sin(phi * DEG2RAD);
            tphi = -C * DEG2RAD / (sn * sn);
        }
    }
    /* chain: d amc = 2 sqrt(anphi) dC + C / sqrt(anphi) d anphi */
    dpls->anphi = danphi * phip;
    dpls->anpsi = danpsi * psip;
    dpls->hardn = 0;
    dpls->amc = 2 * sqrt(anphi) * Cp + C / sqrt(anphi) * danphi * phip;
    dpls->ten_max = tC * Cp + tphi * phip;
    dc0->anphi = 0;
    dc0->anpsi = 0;
    dc0->hardn = hc;
    dc0->amc = 2 * sqrt(anphi) * Cc;
    dc0->ten_max = tC * Cc;
}

/* symmetric-matrix adjoint <-> stored components {xx,yy,zz,xy,xz,yz} */
static void comp_to_mat_adj(const double *sb, double A[3][3])
{
    A[0][0] = sb[0]; A[1][1] = sb[1]; A[2][2] = sb[2];
    A[0][1] = A[1][0] = 0.5 * sb[3];
    A[0][2] = A[2][0] = 0.5 * sb[4];
    A[1][2] = A[2][1] = 0.5 * sb[5];
}

static void mat_to_comp_adj(const double A[3][3], double *sb)
{
    sb[0] += A[0][0]; sb[1] += A[1][1]; sb[2] += A[2][2];
    sb[3] += A[0][1] + A[1][0];
    sb[4] += A[0][2] + A[2][0];
    sb[5] += A[1][2] + A[2][1];
}

/* elasto_plastic (3D, rheology.cxx:151-299) with adjoint.  Forward from s and
   de; ob = d/d out, plb = d/d depls; adds d/d s, d/d de and d/d props. */
static void elasto_plastic_adj(double bulkm, double shearm, const pprops *pp, const double *de,
                               const double *s, const double *ob, double plb, double *sb, double *deb,
                               pprops *ppb)
{
    double str[6];
    for (int i = 0; i < 6; ++i) str[i] = s[i];
    elastic(bulkm, shearm, de, str);
    double p[3], v[3][3];
    principal_stresses3(str, p, v);
    const double anphi = pp->anphi, anpsi = pp->anpsi, amc = pp->amc, hardn = pp->hardn;
    const double ten_max = pp->ten_max;
    double fs = p[0] - p[2] * anphi + amc;
    double ft = p[2] - ten_max;
    double strb[6] = {0, 0, 0, 0, 0, 0};
    if (fs > 0 && ft < 0) {
        for (int i = 0; i < 6; ++i) strb[i] = ob[i];
    } else {
        double pa = sqrt(1 + anphi * anphi) + anphi;
        double ps = ten_max * anphi - amc;
        double h = p[2] - ten_max + pa * (p[0] - ps);
        double a1 = bulkm + 4. / 3 * shearm, a2 = bulkm - 2. / 3 * shearm;
        double w[3], alam, den = 0, c3 = 0;
        int shear = h < 0;
        if (shear) {
            den = a1 - a2 * anpsi + a1 * anphi * anpsi - a2 * anphi + 2 * sqrt(anphi) * hardn;
            alam = fs / den;
            w[0] = a1 - a2 * anpsi; w[1] = a2 - a2 * anpsi; w[2] = a2 - a1 * anpsi;
            c3 = sqrt((7 + 4 * anpsi + 7 * anpsi * anpsi) / 18);
        } else {
            alam = ft / a1;
            w[0] = a2; w[1] = a2; w[2] = a1;
            c3 = sqrt(7. / 18);
        }
        double pn[3];
        for (int k = 0; k < 3; ++k) pn[k] = p[k] - alam * w[k];
        /* out = V diag(pn) V^T: B = V^T A V */
        double A[3][3], B[3][3], T[3][3];
        comp_to_mat_adj(ob, A);
        for (int i = 0; i < 3; ++i)
            for (int k = 0; k < 3; ++k) {
                T[i][k] = 0;
                for (int j = 0; j < 3; ++j) T[i][k] += A[i][j] * v[j][k];
            }
        for (int k = 0; k < 3; ++k)
            for (int l = 0; l < 3; ++l) {
                B[k][l] = 0;
                for (int i = 0; i < 3; ++i) B[k][l] += v[i][k] * T[i][l];
            }
        double pnb[3] = {B[0][0], B[1][1], B[2][2]};
        /* eigenvector sensitivity: C_kl = B_kl (pn_l - pn_k) / (p_l - p_k) */
        double Cm[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
        const double scale = fabs(p[0]) + fabs(p[1]) + fabs(p[2]);
        for (int k = 0; k < 3; ++k)
            for (int l = 0; l < 3; ++l) {
                if (k == l) continue;
                double dp = p[l] - p[k];
                if (fabs(dp) > 1e-14 * scale) Cm[k][l] = B[k][l] * (pn[l] - pn[k]) / dp;
            }
        /* pn = p - alam w */
        double pb[3] = {pnb[0], pnb[1], pnb[2]};
        double alamb = -(pnb[0] * w[0] + pnb[1] * w[1] + pnb[2] * w[2]);
        double wb[3] = {-alam * pnb[0], -alam * pnb[1], -alam * pnb[2]};
        alamb += plb * (alam >= 0 ? 1 : -1) * c3;
        double c3b = plb * fabs(alam);
        if (shear) {
            double fsb = alamb / den, denb = -alamb * alam / den;
            pb[0] += fsb;
            pb[2] -= fsb * anphi;
            ppb->anphi += -fsb * p[2] + denb * (a1 * anpsi - a2 + hardn / sqrt(anphi));
            ppb->amc += fsb;
            ppb->anpsi += denb * (-a2 + a1 * anphi) - a2 * wb[0] - a2 * wb[1] - a1 * wb[2];
            ppb->hardn += denb * 2 * sqrt(anphi);
            ppb->anpsi += c3b * (4 + 14 * anpsi) / (36 * c3);
        } else {
            pb[2] += alamb / a1;
            ppb->ten_max -= alamb / a1;
        }
        /* d p_k = v_k^T dS v_k ; eigenvector part: V C V^T (symmetrised) */
        double M[3][3];
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) {
                double a = 0;
                for (int k = 0; k < 3; ++k) {
                    a += v[i][k] * v[j][k] * pb[k];
                    for (int l = 0; l < 3; ++l) a += v[i][k] * 0.5 * (Cm[k][l] + Cm[l][k]) * v[j][l];
                }
                M[i][j] = a;
            }
        mat_to_comp_adj(M, strb);
    }
    elastic_adj(bulkm, shearm, strb, sb, deb);
}

/* update_stress of one element (rheology.cxx:510-646) with adjoint.
   Inputs: s (start stress), de, dv, visc, c0, pls; ob = d/d out stress,
   plb = d/d new plastic strain.  Adds d/d s, de, dv, visc, c0, pls. */
static void stress_adj(const des_params *P, double bulkm, double shearm, double visc, double c0,
                       double pls, double dt, const double *s, const double *de, double dv,
                       const double *ob, double plb, double *sb, double *deb, double *dvb,
                       double *viscb, double *c0b, double *plsb)
{
    *plsb += plb;   /* pls' = pls + depls */
    switch (P->rheol_type) {
    case DES_RH_ELASTIC:
        elastic_adj(bulkm, shearm, ob, sb, deb);
        break;
    case DES_RH_MAXWELL:
        maxwell_adj(bulkm, shearm, visc, dt, de, s, ob, sb, deb, dvb, viscb);
        break;
    case DES_RH_EP:
    case DES_RH_EVP: {
        pprops pp, dp, dc;
        plastic_props_d(P, c0, pls, &pp, &dp, &dc);
        int plastic = 1;
        if (P->rheol_type == DES_RH_EVP) {
            double sv[6], sp[6], depls;
            int fm;
            for (int i = 0; i < 6; ++i) sv[i] = sp[i] = s[i];
            maxwell(bulkm, shearm, visc, dt, dv, de, sv);
            elasto_plastic(bulkm, shearm, pp.amc, pp.anphi, pp.anpsi, pp.hardn, pp.ten_max, de, &depls, sp, &fm);
            plastic = !(second_invariant2(sv) < second_invariant2(sp));
        }
        if (!plastic) {
            maxwell_adj(bulkm, shearm, visc, dt, de, s, ob, sb, deb, dvb, viscb);
        } else {
            pprops ppb = {0, 0, 0, 0, 0};
            elasto_plastic_adj(bulkm, shearm, &pp, de, s, ob, plb, sb, deb, &ppb);
            double t = ppb.amc * dp.amc + ppb.anphi * dp.anphi + ppb.anpsi * dp.anpsi +
                       ppb.hardn * dp.hardn + ppb.ten_max * dp.ten_max;
            *plsb += t;
            *c0b += ppb.amc * dc.amc + ppb.hardn * dc.hardn + ppb.ten_max * dc.ten_max;
        }
        break;
    }
    }
}

/* --------------------------------------------------------------- records */

typedef struct {
    double *X, *V, *S, *P, *VO;    /* state at the start of a step */
} adj_rec;

static void rec_alloc(adj_rec *r, int N, int E)
{
    r->X = malloc(sizeof(double) * 3 * N);
    r->V = malloc(sizeof(double) * 3 * N);
    r->S = malloc(sizeof(double) * 6 * E);
    r->P = malloc(sizeof(double) * E);
    r->VO = malloc(sizeof(double) * E);
}

static void rec_free(adj_rec *r)
{
    free(r->X); free(r->V); free(r->S); free(r->P); free(r->VO);
}

static void rec_save(adj_rec *r, const ostate *S)
{
    memcpy(r->X, S->coord, sizeof(double) * 3 * S->nnode);
    memcpy(r->V, S->vel, sizeof(double) * 3 * S->nnode);
    memcpy(r->S, S->stress, sizeof(double) * 6 * S->nelem);
    memcpy(r->P, S->plstrain, sizeof(double) * S->nelem);
    memcpy(r->VO, S->volume_old, sizeof(double) * S->nelem);
}

/* ---------------------------------------------------------- reverse step */

typedef struct {
    double *Xb, *Vb, *Sb, *Pb, *VOb;   /* adjoints of the state after the step */
} adj_bar;

/* Reverse of one step from its start state r; b holds the adjoints of the
   state after the step on entry and of the state before it on return.
   viscb / cohb accumulate d/d elem_visc and d/d elem_cohesion. */
static void reverse_step(const ostate *S0, const des_params *P, const adj_rec *r, adj_bar *b,
                         double *viscb, double *cohb)
{
    const int N = S0->nnode, E = S0->nelem;
    const int *conn = S0->conn;
    const double dt = P->fixed_dt;
    const double bulkm = P->bulk_modulus[0], shearm = P->shear_modulus[0];
    const double rho = P->rho0[0];
    const double pseudo_speed = P->max_vbc_val * P->inertial_scaling;
    const double rho_eff = P->is_quasi_static ? bulkm / (pseudo_speed * pseudo_speed) : rho;
    const double buoy = P->gravity != 0 ? rho * P->gravity / NPE : 0;

    /* ---- forward recompute of the step ---- */
    double *vol = calloc(E, sizeof(double)), *G = calloc((size_t)E * 12, sizeof(double));
    double *vn = calloc(N, sizeof(double)), *m = calloc(N, sizeof(double));
    double *D = calloc((size_t)E * 6, sizeof(double)), *dvn = calloc(N, sizeof(double));
    double *Sp = calloc((size_t)E * 6, sizeof(double)), *dp = calloc(E, sizeof(double));
    double *dpn = calloc(N, sizeof(double)), *Spp = calloc((size_t)E * 6, sizeof(double));
    double *F = calloc((size_t)N * 3, sizeof(double)), *Ft = calloc((size_t)N * 3, sizeof(double));
    double *V2 = calloc((size_t)N * 3, sizeof(double)), *X1 = calloc((size_t)N * 3, sizeof(double));
    double *de_all = calloc((size_t)E * 6, sizeof(double)), *dv_all = calloc(E, sizeof(double));
    for (int e = 0; e < E; ++e) {
        double x[4][3], Ge[4][3];
        gather4(r->X, conn + 4 * e, x);
        vol[e] = tet_vol(x);
        tet_grad(x, Ge);
        memcpy(G + 12 * e, Ge, sizeof Ge);
        for (int i = 0; i < 4; ++i) {
            vn[conn[4 * e + i]] += vol[e];
            m[conn[4 * e + i]] += rho_eff * vol[e] / NPE;
        }
        double v[4][3];
        gather4(r->V, conn + 4 * e, v);
        double *d = D + 6 * e;
        for (int i = 0; i < 4; ++i) {
            double g[3] = {Ge[i][0] / vol[e], Ge[i][1] / vol[e], Ge[i][2] / vol[e]};
            d[0] += v[i][0] * g[0];
            d[1] += v[i][1] * g[1];
            d[2] += v[i][2] * g[2];
            d[3] += 0.5 * (v[i][0] * g[1] + v[i][1] * g[0]);
            d[4] += 0.5 * (v[i][0] * g[2] + v[i][2] * g[0]);
            d[5] += 0.5 * (v[i][1] * g[2] + v[i][2] * g[1]);
        }
        for (int i = 0; i < 4; ++i) dvn[conn[4 * e + i]] += (d[0] + d[1] + d[2]) * vol[e];
    }
    for (int n = 0; n < N; ++n) dvn[n] /= vn[n];
    for (int e = 0; e < E; ++e) {
        const int *c = conn + 4 * e;
        double ed = (dvn[c[0]] + dvn[c[1]] + dvn[c[2]] + dvn[c[3]]) / 4;
        double *d = D + 6 * e, *de = de_all + 6 * e;
        double tr = d[0] + d[1] + d[2];
        for (int i = 0; i < 6; ++i) de[i] = d[i] * dt;
        for (int i = 0; i < 3; ++i) de[i] = (d[i] + (ed - tr) / 3) * dt;
        dv_all[e] = vol[e] / r->VO[e] - 1;
        double s[6];
        memcpy(s, r->S + 6 * e, sizeof s);
        double old = s[0] + s[1] + s[2];
        double visc = S0->elem_visc ? dmin(dmax(S0->elem_visc[e], P->visc_min), P->visc_max) : P->visc_min;
        double c0 = S0->elem_cohesion ? S0->elem_cohesion[e] : P->cohesion0[0];
        switch (P->rheol_type) {
        case DES_RH_ELASTIC: elastic(bulkm, shearm, de, s); break;
        case DES_RH_MAXWELL: maxwell(bulkm, shearm, visc, dt, dv_all[e], de, s); break;
        default: {
            pprops pp, d1, d2;
            plastic_props_d(P, c0, r->P[e], &pp, &d1, &d2);
            double sp[6], depls;
            int fm;
            memcpy(sp, s, sizeof sp);
            elasto_plastic(bulkm, shearm, pp.amc, pp.anphi, pp.anpsi, pp.hardn, pp.ten_max, de, &depls, sp, &fm);
            if (P->rheol_type == DES_RH_EVP) {
                double sv[6];
                memcpy(sv, s, sizeof sv);
                maxwell(bulkm, shearm, visc, dt, dv_all[e], de, sv);
                if (second_invariant2(sv) < second_invariant2(sp)) memcpy(sp, sv, sizeof sp);
            }
            memcpy(s, sp, sizeof s);
        }
        }
        memcpy(Sp + 6 * e, s, sizeof s);
        dp[e] = s[0] + s[1] + s[2] - old;
        for (int i = 0; i < 4; ++i) dpn[c[i]] += dp[e] * vol[e];
    }
    for (int n = 0; n < N; ++n) dpn[n] /= vn[n];
    for (int e = 0; e < E; ++e) {
        const int *c = conn + 4 * e;
        double dpe = (dpn[c[0]] + dpn[c[1]] + dpn[c[2]] + dpn[c[3]]) / 4;
        double *s = Spp + 6 * e;
        memcpy(s, Sp + 6 * e, sizeof(double) * 6);
        for (int i = 0; i < 3; ++i) s[i] += (dpe - dp[e]) / 3;
        for (int i = 0; i < 4; ++i) {
            const double *g = G + 12 * e + 3 * i;   /* G_i = g_i vol */
            double *f = F + 3 * c[i];
            f[0] -= s[0] * g[0] + s[3] * g[1] + s[4] * g[2];
            f[1] -= s[3] * g[0] + s[1] * g[1] + s[5] * g[2];
            f[2] -= s[4] * g[0] + s[5] * g[1] + s[2] * g[2] + buoy * vol[e];
        }
    }
    for (int k = 0; k < 3 * N; ++k) {
        Ft[k] = F[k];
        if (P->is_quasi_static && P->damping_option == 1 && fabs(r->V[k]) > 1e-13)
            Ft[k] -= P->damping_factor * copysign(F[k], r->V[k]);
        V2[k] = r->V[k] + dt * Ft[k] / m[k / 3];
    }
    {   /* boundary velocities */
        ostate T = *S0;
        T.vel = V2;
        o3_apply_vbcs(&T, P);
    }
    for (int k = 0; k < 3 * N; ++k) X1[k] = r->X[k] + dt * V2[k];

    /* ---- reverse ---- */
    double *Vb = calloc((size_t)N * 3, sizeof(double));
    double *Xb = calloc((size_t)N * 3, sizeof(double));
    double *Sb = calloc((size_t)E * 6, sizeof(double));
    double *Pb = calloc(E, sizeof(double)), *VOb = calloc(E, sizeof(double));
    double *V2b = calloc((size_t)N * 3, sizeof(double)), *X1b = calloc((size_t)N * 3, sizeof(double));
    double *Sppb = calloc((size_t)E * 6, sizeof(double));
    double *volb = calloc(E, sizeof(double)), *Gb = calloc((size_t)E * 12, sizeof(double));
    double *vnb = calloc(N, sizeof(double)), *mb = calloc(N, sizeof(double));
    double *Fb = calloc((size_t)N * 3, sizeof(double));
    memcpy(X1b, b->Xb, sizeof(double) * 3 * N);
    memcpy(V2b, b->Vb, sizeof(double) * 3 * N);

    /* rotate_stress (fields.cxx:411-483) with the new velocity and geometry */
    if (P->rheol_type & DES_RH_ELASTIC) {
        for (int e = 0; e < E; ++e) {
            const int *c = conn + 4 * e;
            double x[4][3], G1[4][3], v[4][3], g[4][3];
            gather4(X1, c, x);
            gather4(V2, c, v);
            double vol1 = tet_vol(x);
            tet_grad(x, G1);
            for (int i = 0; i < 4; ++i)
                for (int d = 0; d < 3; ++d) g[i][d] = G1[i][d] / vol1;
            double w3 = 0, w4 = 0, w5 = 0;
            for (int i = 0; i < 4; ++i) {
                w3 += 0.5 * (v[i][0] * g[i][1] - v[i][1] * g[i][0]);
                w4 += 0.5 * (v[i][0] * g[i][2] - v[i][2] * g[i][0]);
                w5 += 0.5 * (v[i][1] * g[i][2] - v[i][2] * g[i][1]);
            }
            const double *s = Spp + 6 * e, *o = b->Sb + 6 * e;
            double *sb = Sppb + 6 * e;
            for (int i = 0; i < 6; ++i) sb[i] += o[i];
            sb[3] += dt * (-2 * w3 * o[0] + 2 * w3 * o[1] + w5 * o[4] + w4 * o[5]);
            sb[4] += dt * (-2 * w4 * o[0] + 2 * w4 * o[2] - w5 * o[3] + w3 * o[5]);
            sb[5] += dt * (-2 * w5 * o[1] + 2 * w5 * o[2] - w4 * o[3] - w3 * o[4]);
            sb[0] += dt * (w3 * o[3] + w4 * o[4]);
            sb[1] += dt * (-w3 * o[3] + w5 * o[5]);
            sb[2] += dt * (-w4 * o[4] - w5 * o[5]);
            double w3b = dt * (-2 * s[3] * o[0] + 2 * s[3] * o[1] + (s[0] - s[1]) * o[3] - s[5] * o[4] + s[4] * o[5]);
            double w4b = dt * (-2 * s[4] * o[0] + 2 * s[4] * o[2] - s[5] * o[3] + (s[0] - s[2]) * o[4] + s[3] * o[5]);
            double w5b = dt * (-2 * s[5] * o[1] + 2 * s[5] * o[2] - s[4] * o[3] + s[3] * o[4] + (s[1] - s[2]) * o[5]);
            double gb[4][3], G1b[4][3], xb[4][3];
            memset(xb, 0, sizeof xb);
            double vol1b = 0;
            for (int i = 0; i < 4; ++i) {
                double *vb = V2b + 3 * c[i];
                vb[0] += 0.5 * (w3b * g[i][1] + w4b * g[i][2]);
                vb[1] += 0.5 * (-w3b * g[i][0] + w5b * g[i][2]);
                vb[2] += 0.5 * (-w4b * g[i][0] - w5b * g[i][1]);
                gb[i][0] = 0.5 * (-w3b * v[i][1] - w4b * v[i][2]);
                gb[i][1] = 0.5 * (w3b * v[i][0] - w5b * v[i][2]);
                gb[i][2] = 0.5 * (w4b * v[i][0] + w5b * v[i][1]);
                for (int d = 0; d < 3; ++d) {
                    G1b[i][d] = gb[i][d] / vol1;
                    vol1b -= gb[i][d] * g[i][d] / vol1;
                }
            }
            for (int i = 0; i < 4; ++i)
                for (int d = 0; d < 3; ++d) xb[i][d] += vol1b * G1[i][d];
            tet_grad_adj(x, G1b, xb);
            for (int i = 0; i < 4; ++i)
                for (int d = 0; d < 3; ++d) X1b[3 * c[i] + d] += xb[i][d];
        }
    } else {
        memcpy(Sppb, b->Sb, sizeof(double) * 6 * E);
    }
    /* update_coordinate: X1 = X + dt V2 */
    for (int k = 0; k < 3 * N; ++k) {
        Xb[k] += X1b[k];
        V2b[k] += dt * X1b[k];
    }
    /* apply_vbcs: a projection that sets components, so its adjoint zeroes them */
    {
        des_params Q = *P;
        for (int i = 0; i < DES_NBDRY; ++i) Q.vbc_values[i] = 0;
        ostate T = *S0;
        T.vel = V2b;
        o3_apply_vbcs(&T, &Q);
    }
    /* update_velocity and damping */
    for (int k = 0; k < 3 * N; ++k) {
        Vb[k] += V2b[k];
        double ftb = dt * V2b[k] / m[k / 3];
        mb[k / 3] -= dt * V2b[k] * Ft[k] / (m[k / 3] * m[k / 3]);
        double fac = 1;
        if (P->is_quasi_static && P->damping_option == 1 && fabs(r->V[k]) > 1e-13)
            fac = 1 - P->damping_factor * ((F[k] < 0) == (r->V[k] < 0) ? 1 : -1);
        Fb[k] = ftb * fac;
    }
    /* update_force: F_n = -sum (S'' g_i vol + buoy vol z) ; mass */
    for (int e = 0; e < E; ++e) {
        const int *c = conn + 4 * e;
        const double *s = Spp + 6 * e;
        double *sb = Sppb + 6 * e;
        for (int i = 0; i < 4; ++i) {
            const double *fb = Fb + 3 * c[i];
            const double *g = G + 12 * e + 3 * i;
            double *gb = Gb + 12 * e + 3 * i;
            sb[0] -= fb[0] * g[0];
            sb[1] -= fb[1] * g[1];
            sb[2] -= fb[2] * g[2];
            sb[3] -= fb[0] * g[1] + fb[1] * g[0];
            sb[4] -= fb[0] * g[2] + fb[2] * g[0];
            sb[5] -= fb[1] * g[2] + fb[2] * g[1];
            gb[0] -= fb[0] * s[0] + fb[1] * s[3] + fb[2] * s[4];
            gb[1] -= fb[0] * s[3] + fb[1] * s[1] + fb[2] * s[5];
            gb[2] -= fb[0] * s[4] + fb[1] * s[5] + fb[2] * s[2];
            volb[e] -= fb[2] * buoy;
            volb[e] += mb[c[i]] * rho_eff / NPE;
        }
    }
    /* NMD_stress: S'' = S' + (mean dpn - dp) / 3 on the diagonal */
    double *dpb = calloc(E, sizeof(double)), *dpnb = calloc(N, sizeof(double));
    for (int e = 0; e < E; ++e) {
        const double *sb = Sppb + 6 * e;
        double q = (sb[0] + sb[1] + sb[2]) / 3;
        dpb[e] -= q;
        for (int i = 0; i < 4; ++i) dpnb[conn[4 * e + i]] += q / 4;
    }
    for (int n = 0; n < N; ++n) {       /* dpn = num / vn */
        vnb[n] -= dpnb[n] * dpn[n] / vn[n];
        dpnb[n] /= vn[n];               /* now d/d num */
    }
    for (int e = 0; e < E; ++e)
        for (int i = 0; i < 4; ++i) {
            int n = conn[4 * e + i];
            dpb[e] += dpnb[n] * vol[e];
            volb[e] += dpnb[n] * dp[e];
        }
    /* update_stress */
    double *Db = calloc((size_t)E * 6, sizeof(double));
    double *edb = calloc(E, sizeof(double));
    for (int e = 0; e < E; ++e) {
        double ob[6];
        for (int i = 0; i < 6; ++i) ob[i] = Sppb[6 * e + i];
        double *sb = Sb + 6 * e;
        for (int i = 0; i < 3; ++i) {   /* dp = tr(S') - tr(S) */
            ob[i] += dpb[e];
            sb[i] -= dpb[e];
        }
        double deb[6] = {0, 0, 0, 0, 0, 0}, dvb = 0, vb_ = 0, cb = 0;
        double visc = S0->elem_visc ? dmin(dmax(S0->elem_visc[e], P->visc_min), P->visc_max) : P->visc_min;
        double c0 = S0->elem_cohesion ? S0->elem_cohesion[e] : P->cohesion0[0];
        stress_adj(P, bulkm, shearm, visc, c0, r->P[e], dt, r->S + 6 * e, de_all + 6 * e, dv_all[e],
                   ob, b->Pb[e], sb, deb, &dvb, &vb_, &cb, Pb + e);
        if (viscb && S0->elem_visc && S0->elem_visc[e] > P->visc_min && S0->elem_visc[e] < P->visc_max)
            viscb[e] += vb_;
        if (cohb && S0->elem_cohesion) cohb[e] += cb;
        /* dv = vol / VO - 1 */
        volb[e] += dvb / r->VO[e];
        VOb[e] -= dvb * vol[e] / (r->VO[e] * r->VO[e]);
        /* de = (D + (ed - tr D)/3 I) dt */
        double a = (deb[0] + deb[1] + deb[2]) * dt / 3;
        for (int i = 0; i < 6; ++i) Db[6 * e + i] = deb[i] * dt;
        for (int i = 0; i < 3; ++i) Db[6 * e + i] -= a;
        edb[e] = a;
    }
    /* the next step's volume_old is this step's volume */
    for (int e = 0; e < E; ++e) volb[e] += b->VOb[e];
    /* edvoldt = mean of dvoldt; dvoldt = num / vn; num = sum tr(D) vol */
    double *dvnb = calloc(N, sizeof(double));
    for (int e = 0; e < E; ++e)
        for (int i = 0; i < 4; ++i) dvnb[conn[4 * e + i]] += edb[e] / 4;
    for (int n = 0; n < N; ++n) {
        vnb[n] -= dvnb[n] * dvn[n] / vn[n];
        dvnb[n] /= vn[n];
    }
    for (int e = 0; e < E; ++e) {
        const double *d = D + 6 * e;
        double tr = d[0] + d[1] + d[2];
        for (int i = 0; i < 4; ++i) {
            int n = conn[4 * e + i];
            for (int j = 0; j < 3; ++j) Db[6 * e + j] += dvnb[n] * vol[e];
            volb[e] += dvnb[n] * tr + vnb[n];
        }
    }
    /* strain rate, then geometry */
    for (int e = 0; e < E; ++e) {
        const int *c = conn + 4 * e;
        double x[4][3], v[4][3], xb[4][3], G1b[4][3];
        gather4(r->X, c, x);
        gather4(r->V, c, v);
        memset(xb, 0, sizeof xb);
        const double *db = Db + 6 * e;
        const double *Ge = G + 12 * e;
        double vole = vol[e], vb_ = volb[e];
        for (int i = 0; i < 4; ++i) {
            double g[3] = {Ge[3 * i] / vole, Ge[3 * i + 1] / vole, Ge[3 * i + 2] / vole};
            double *vb = Vb + 3 * c[i];
            vb[0] += db[0] * g[0] + 0.5 * db[3] * g[1] + 0.5 * db[4] * g[2];
            vb[1] += db[1] * g[1] + 0.5 * db[3] * g[0] + 0.5 * db[5] * g[2];
            vb[2] += db[2] * g[2] + 0.5 * db[4] * g[0] + 0.5 * db[5] * g[1];
            double gb[3];
            gb[0] = db[0] * v[i][0] + 0.5 * db[3] * v[i][1] + 0.5 * db[4] * v[i][2];
            gb[1] = db[1] * v[i][1] + 0.5 * db[3] * v[i][0] + 0.5 * db[5] * v[i][2];
            gb[2] = db[2] * v[i][2] + 0.5 * db[4] * v[i][0] + 0.5 * db[5] * v[i][1];
            for (int d = 0; d < 3; ++d) {
                G1b[i][d] = Gb[12 * e + 3 * i + d] + gb[d] / vole;
                vb_ -= gb[d] * g[d] / vole;
            }
        }
        for (int i = 0; i < 4; ++i)
            for (int d = 0; d < 3; ++d) xb[i][d] += vb_ * Ge[3 * i + d];
        tet_grad_adj(x, G1b, xb);
        for (int i = 0; i < 4; ++i)
            for (int d = 0; d < 3; ++d) Xb[3 * c[i] + d] += xb[i][d];
    }
    /* the start stress, plastic strain and volume_old adjoints */
    memcpy(b->Xb, Xb, sizeof(double) * 3 * N);
    memcpy(b->Vb, Vb, sizeof(double) * 3 * N);
    memcpy(b->Sb, Sb, sizeof(double) * 6 * E);
    memcpy(b->Pb, Pb, sizeof(double) * E);
    memcpy(b->VOb, VOb, sizeof(double) * E);
    free(vol); free(G); free(vn); free(m); free(D); free(dvn); free(Sp); free(dp); free(dpn);
    free(Spp); free(F); free(Ft); free(V2); free(X1); free(de_all); free(dv_all);
    free(Vb); free(Xb); free(Sb); free(Pb); free(VOb); free(V2b); free(X1b); free(Sppb);
    free(volb); free(Gb); free(vnb); free(mb); free(Fb); free(dpb); free(dpnb); free(Db);
    free(edb); free(dvnb);
}

/* nsteps oracle steps from the current state, then the gradient of
   L = sum coord_bar . X_final with respect to elem_visc (viscb) and
   elem_cohesion (cohb), both accumulated (+=).  Returns 0, or 1 for an
   unsupported configuration. */
int adj3_gradient(ostate *S, const des_params *P, int nsteps, const double *coord_bar,
                  double *viscb, double *cohb)
{
    if (!adj3_supported(S, P)) return 1;
    const int N = S->nnode, E = S->nelem;
    adj_rec *rec = calloc(nsteps, sizeof(adj_rec));
    for (int k = 0; k < nsteps; ++k) {
        rec_alloc(rec + k, N, E);
        rec_save(rec + k, S);
        o3_step(S, P);
    }
    adj_bar b;
    b.Xb = calloc((size_t)N * 3, sizeof(double));
    b.Vb = calloc((size_t)N * 3, sizeof(double));
    b.Sb = calloc((size_t)E * 6, sizeof(double));
    b.Pb = calloc(E, sizeof(double));
    b.VOb = calloc(E, sizeof(double));
    memcpy(b.Xb, coord_bar, sizeof(double) * 3 * N);
    for (int k = nsteps - 1; k >= 0; --k) reverse_step(S, P, rec + k, &b, viscb, cohb);
    for (int k = 0; k < nsteps; ++k) rec_free(rec + k);
    free(rec);
    free(b.Xb); free(b.Vb); free(b.Sb); free(b.Pb); free(b.VOb);
    return 0;
}

/* test hooks */
void adj3_tet_grad(const double *x12, double *G12) { tet_grad((const double (*)[3])x12, (double (*)[3])G12); }
void adj3_tet_grad_adj(const double *x12, const double *Gb12, double *xb12)
{
    tet_grad_adj((const double (*)[3])x12, (const double (*)[3])Gb12, (double (*)[3])xb12);
}
