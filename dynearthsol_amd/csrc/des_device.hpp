// des_device.hpp -- per-element and per-node math of the DynEarthSol explicit
// step as device functions for gfx950.
//
// Each function restates one reference routine (file:line in its comment) with
// the same operand order; the library is compiled with -ffp-contract=off so
// no multiply-add is fused and FP64 results match the reference's x86-64
// build except where libm transcendentals round differently.
#pragma once

#include <hip/hip_runtime.h>
#include <cfloat>
#include <cstdint>

#include "des_amd.h"

namespace des {

template <int ND> struct Dim;
template <> struct Dim<3> { static constexpr int NPE = 4, NSTR = 6; };
template <> struct Dim<2> { static constexpr int NPE = 3, NSTR = 3; };

// boundary flag bits (constants.hpp:38-51)
constexpr unsigned BOUNDX0 = 1u, BOUNDX1 = 2u, BOUNDY0 = 4u, BOUNDY1 = 8u;
constexpr unsigned BOUNDZ0 = 16u, BOUNDZ1 = 32u, BOUND_ANY = 1023u;
constexpr int IBOUNDN0 = 6, IBOUNDN3 = 9;

constexpr double DEG2RAD = 3.1415926535897932384626433832795 / 180;

// Scalars the step changes on the device (var.dt, var.time, ...).
struct DevScalars {
    double dt;         // var.dt
    double dt_prev;    // dt of the previous step (deferred Jaumann rotation)
    double time;       // var.time
    double compensation_pressure;
    int steps;
    int err;           // first device-detected error (des_status)
    int rot_pending;   // previous step's rotate_stress still to apply
    int pad;
    double time0;      // Output::time0: var.time at the averaging interval's first step
};

// What a mixed-precision 3D element whose elastic trial is not clearly inside
// the Mohr-Coulomb surface leaves for the return kernel: the FP64 trial, the
// pressures of the elastic trial (pp), of the Maxwell stress (pv) and before
// the update (p_old), and the FP32 deviator increments of both candidates.
struct MixedRet {
    double s64[6], pp, pv, p_old;
    float inc[6], incv[6], svII, pad;
};
static_assert(sizeof(MixedRet) == 128, "one 128-byte line per record");

// A pointer into LDS with its address space in the type: loads through it
// stay LDS loads, and the compiler cannot merge them with global loads of the
// same shape on another path into generic (flat) loads.
template <class T> using lds_ptr = const __attribute__((address_space(3))) T *;

// Every device array of one model.  Node vectors and element tensors are
// component-major (SoA); connectivity stays AoS as in the reference.
struct DevState {
    int N, E, nbf, mpe;
    const des_params *P;
    const int *conn;            // E x NPE
    const unsigned *bcflag;     // N
    const int *elemmarkers;     // E x nmat, or null
    const int *ehost, *nhost;   // device element / node -> host index (model-local in a group)
    const int *csr_off;         // N+1
    const int *csr_ent;         // e*NPE+i, ascending host element index per node
    const int *ff_off;          // N+1: node -> traction-loaded facets
    const int *ff_ent;          // facet index, ascending
    const int *bf_bdry, *bf_elem, *bf_face;
    const double *bnormals;     // 10 x ND
    const double *edge_vectors; // 10 x 10 x ND
    const int *edge_valid;      // 10 x 10
    double *coord, *coord0, *vel, *force;         // ND x N
    double *T, *mass, *tmass, *volume_n, *ntmp;   // N
    double *volume, *volume_old, *edvoldt, *plstrain, *delta_plstrain, *dpressure, *stressyy;
    double *strain_rate, *strain, *stress;        // NSTR x E
    double *shp;                                  // ND x NPE x E
    const double *elem_visc, *elem_cohesion;      // E or null
    double *ework;                                // E x NPE scalar contributions (temperature)
    double *fwork;                                // force contributions, ND per CSR slot
    double *frec;                                 // force contributions, element-major records (FrecLayout)
    const int *slot;                              // E x NPE: CSR position of entry e*NPE+i
    const unsigned *srank;                        // E: byte i = position of entry e*NPE+i in its
                                                  // node's list (null: a list longer than 256)
    double *erec;                                 // E x 8 packed records (fused pass A -> B)
    double *escal;                                // 4 x E element scalars
    double *ftrac;                                // nbf x ND facet tractions
    DevScalars *sc;
    double *red;                                  // reduction partials
    int *redi;
    // mixed-precision state (des_run with DES_MIXED): stress = sdev + pres * I
    double *pres;                                 // E, mean of the stress diagonal
    float *sdev;                                  // NSTR x E, stress minus pres * I
    float *strain32;                              // NSTR x E
    float *dpl32;                                 // E (the plastic strain itself stays FP64)
    const float *visc32, *coh32;                  // E or null
    // batched models (null for a single model): elements and nodes of model m
    // fill whole blocks; block b of an element (node) kernel belongs to model
    // eblk_model[b] (nblk_model[b]) and holds real elements below eblk_lim[b]
    const int *eblk_model, *eblk_lim, *nblk_model, *nblk_lim;
    const des_params *Pb;                         // per-model parameters
    DevScalars *scb;                              // per-model scalars
    const int *fmodel;                            // model of each traction-loaded facet slot
    const double *bnb, *evb;                      // per-model boundary normals / edge vectors
    const int *evvb;                              // per-model edge_valid
    const int *mblk_off;                          // first element block of each model (+ total)
    const int *e0dev;                             // per model: device index of host element 0
    int nmodels;
    // 1: each XCD walks its contiguous range of blocks from the top down
    // (the fused schedule's node passes; see xcd_block)
    int rev;
    // elements whose Mohr-Coulomb return is left to the return kernel of the
    // fused schedule: count, element ids, and per entry the trial stress and
    // the stress trace before the update (8 doubles)
    int *ret_count, *ret_elem;
    double *ret_val;
    MixedRet *ret_mix;                            // the mixed-precision schedule's records
    // node records of the fused schedule (NREC doubles per node): x[ND],
    // v[ND], T, aux (dvoldt after pass B, dp_nd after pass D), one cache
    // line per node for the element passes' gathers
    double *nrec;
    // per block of BLOCK elements (the element passes' workgroups): {offset,
    // count} of its distinct nodes in stage_node (count -1: more than
    // STAGE_MAX, the block gathers from nrec directly), and per element its
    // nodes' positions in that list, 8 bits each
    const int2 *stage;
    const int *stage_node;
    const unsigned *stage_lidx;
    // Output's averaged fields (output.cxx:211-251), allocated when averaging
    // is on: coord and strain at the interval's first step, the running sums
    // of stress and delta_plstrain
    double *avg_coord0, *avg_strain0, *avg_stress, *avg_dpl;
};

constexpr int NREC = 8;
constexpr int STAGE_MAX = 256;   // node records an element block stages in LDS

// slots of the node record
template <int ND> struct NrecSlot {
    static constexpr int X = 0, V = ND, T = 2 * ND, AUX = 2 * ND + 1;
};

// the records of an element's nodes: coordinates, velocities, T and aux
template <int ND> __device__ __forceinline__ void nrec_unpack(const double2 *r, double x[ND], double v[ND],
                                                            double &T, double &aux)
{
    if constexpr (ND == 3) {
        const double2 a = r[0], b = r[1], d = r[2], f = r[3];
        x[0] = a.x; x[1] = a.y; x[2] = b.x;
        v[0] = b.y; v[1] = d.x; v[2] = d.y;
        T = f.x; aux = f.y;
    } else {
        const double2 a = r[0], b = r[1], d = r[2];
        x[0] = a.x; x[1] = a.y;
        v[0] = b.x; v[1] = b.y;
        T = d.x; aux = d.y;
    }
}

// the same record staged in LDS
template <int ND> __device__ __forceinline__ void nrec_unpack(lds_ptr<double2> r, double x[ND], double v[ND],
                                                            double &T, double &aux)
{
    const lds_ptr<double> d = (lds_ptr<double>)r;
#pragma unroll
    for (int k = 0; k < ND; ++k) {
        x[k] = d[k];
        v[k] = d[ND + k];
    }
    T = d[2 * ND];
    aux = d[2 * ND + 1];
}

template <int ND> __device__ __forceinline__ void load_nrec(const double *nrec, const int *c,
                                                          double x[Dim<ND>::NPE][ND],
                                                          double v[Dim<ND>::NPE][ND],
                                                          double T[Dim<ND>::NPE], double aux[Dim<ND>::NPE])
{
#pragma unroll
    for (int i = 0; i < Dim<ND>::NPE; ++i)
        nrec_unpack<ND>(reinterpret_cast<const double2 *>(nrec + (long)NREC * c[i]), x[i], v[i], T[i], aux[i]);
}

// mean over an element's nodes, in node order (elem_mean's summation)
template <int ND> __device__ __forceinline__ double node_mean(const double a[Dim<ND>::NPE])
{
    double t = 0;
#pragma unroll
    for (int i = 0; i < Dim<ND>::NPE; ++i) t += a[i];
    return t / Dim<ND>::NPE;
}

__device__ __forceinline__ double dmin(double a, double b) { return (b < a) ? b : a; }  // std::min
__device__ __forceinline__ double dmax(double a, double b) { return (a < b) ? b : a; }  // std::max

// ------------------------------------------------------------- utils.hpp

template <int ND> __device__ __forceinline__ double trace(const double *s)   // utils.hpp:100-107
{
    if constexpr (ND == 3) return s[0] + s[1] + s[2];
    else return s[0] + s[1];
}

template <int ND> __device__ __forceinline__ double second_invariant2(const double *t)  // utils.hpp:110-119
{
    if constexpr (ND == 3) {
        double a = (t[0] + t[1] + t[2]) / 3;
        return (0.5 * ((t[0] - a) * (t[0] - a) + (t[1] - a) * (t[1] - a) + (t[2] - a) * (t[2] - a))
                + t[3] * t[3] + t[4] * t[4] + t[5] * t[5]);
    } else {
        return 0.25 * (t[0] - t[1]) * (t[0] - t[1]) + t[2] * t[2];
    }
}

template <int ND> __device__ __forceinline__ double second_invariant(const double *t)  // utils.hpp:122-128
{
    return sqrt(second_invariant2<ND>(t));
}

// --------------------------------------------------------- element loads

template <int ND> __device__ __forceinline__ void load_conn(const DevState &s, int e, int c[Dim<ND>::NPE])
{
    if constexpr (ND == 3) {
        int4 v = reinterpret_cast<const int4 *>(s.conn)[e];
        c[0] = v.x; c[1] = v.y; c[2] = v.z; c[3] = v.w;
    } else {
        const int *p = s.conn + 3 * e;
        c[0] = p[0]; c[1] = p[1]; c[2] = p[2];
    }
}

template <int ND> __device__ __forceinline__ void gather_vec(const double *a, int N, const int *c,
                                                           double x[Dim<ND>::NPE][ND])
{
#pragma unroll
    for (int i = 0; i < Dim<ND>::NPE; ++i)
#pragma unroll
        for (int d = 0; d < ND; ++d) x[i][d] = a[d * N + c[i]];
}

// ---------------------------------------------------------- geometry.cxx

__device__ __forceinline__ double tetrahedron_volume(const double *d0, const double *d1,
                                                     const double *d2, const double *d3)  // geometry.cxx:29-49
{
    double x01 = d0[0] - d1[0], x12 = d1[0] - d2[0], x23 = d2[0] - d3[0];
    double y01 = d0[1] - d1[1], y12 = d1[1] - d2[1], y23 = d2[1] - d3[1];
    double z01 = d0[2] - d1[2], z12 = d1[2] - d2[2], z23 = d2[2] - d3[2];
    return (x01 * (y23 * z12 - y12 * z23) +
            x12 * (y01 * z23 - y23 * z01) +
            x23 * (y12 * z01 - y01 * z12)) / 6;
}

template <int ND> __device__ __forceinline__ double triangle_area(const double *a, const double *b,
                                                                const double *c)  // geometry.cxx:53-83
{
    double ab0 = b[0] - a[0], ab1 = b[1] - a[1];
    double ac0 = c[0] - a[0], ac1 = c[1] - a[1];
    if constexpr (ND == 2) {
        return fabs(ab0 * ac1 - ab1 * ac0) / 2;
    } else {
        double ab2 = b[2] - a[2], ac2 = c[2] - a[2];
        double d0 = ab1 * ac2 - ab2 * ac1;
        double d1 = ab2 * ac0 - ab0 * ac2;
        double d2 = ab0 * ac1 - ab1 * ac0;
        return sqrt(d0 * d0 + d1 * d1 + d2 * d2) / 2;
    }
}

template <int ND> __device__ __forceinline__ double dist2(const double *a, const double *b)  // geometry.cxx:16-24
{
    double sum = 0;
#pragma unroll
    for (int i = 0; i < ND; ++i) {
        double d = b[i] - a[i];
        sum += d * d;
    }
    return sum;
}

template <int ND> __device__ __forceinline__ double elem_volume(const double x[Dim<ND>::NPE][ND])
{
    if constexpr (ND == 3) return tetrahedron_volume(x[0], x[1], x[2], x[3]);
    else return triangle_area<2>(x[0], x[1], x[2]);
}

// compute_shape_fn (geometry.cxx:368-455): sh[d][i] = dN_i/dx_d
template <int ND> __device__ __forceinline__ void shape_fn(const double x[Dim<ND>::NPE][ND], double vol,
                                                         double sh[ND][Dim<ND>::NPE])
{
    if constexpr (ND == 3) {
        const double *d0 = x[0], *d1 = x[1], *d2 = x[2], *d3 = x[3];
        double iv = 1 / (6 * vol);
        double x01 = d0[0] - d1[0], x02 = d0[0] - d2[0], x03 = d0[0] - d3[0];
        double x12 = d1[0] - d2[0], x13 = d1[0] - d3[0], x23 = d2[0] - d3[0];
        double y01 = d0[1] - d1[1], y02 = d0[1] - d2[1], y03 = d0[1] - d3[1];
        double y12 = d1[1] - d2[1], y13 = d1[1] - d3[1], y23 = d2[1] - d3[1];
        double z01 = d0[2] - d1[2], z02 = d0[2] - d2[2], z03 = d0[2] - d3[2];
        double z12 = d1[2] - d2[2], z13 = d1[2] - d3[2], z23 = d2[2] - d3[2];
        sh[0][0] = iv * (y13 * z12 - y12 * z13);
        sh[0][1] = iv * (y02 * z23 - y23 * z02);
        sh[0][2] = iv * (y13 * z03 - y03 * z13);
        sh[0][3] = iv * (y01 * z02 - y02 * z01);
        sh[1][0] = iv * (z13 * x12 - z12 * x13);
        sh[1][1] = iv * (z02 * x23 - z23 * x02);
        sh[1][2] = iv * (z13 * x03 - z03 * x13);
        sh[1][3] = iv * (z01 * x02 - z02 * x01);
        sh[2][0] = iv * (x13 * y12 - x12 * y13);
        sh[2][1] = iv * (x02 * y23 - x23 * y02);
        sh[2][2] = iv * (x13 * y03 - x03 * y13);
        sh[2][3] = iv * (x01 * y02 - x02 * y01);
    } else {
        const double *d0 = x[0], *d1 = x[1], *d2 = x[2];
        double iv = 1 / (2 * vol);
        sh[0][0] = iv * (d1[1] - d2[1]);
        sh[0][1] = iv * (d2[1] - d0[1]);
        sh[0][2] = iv * (d0[1] - d1[1]);
        sh[1][0] = iv * (d2[0] - d1[0]);
        sh[1][1] = iv * (d0[0] - d2[0]);
        sh[1][2] = iv * (d1[0] - d0[0]);
    }
}

// ------------------------------------------------------------ fields.cxx

// update_strain_rate (fields.cxx:166-232), sh[0]=dx, sh[ND-1]=dz
template <int ND> __device__ __forceinline__ void strain_rate(const double v[Dim<ND>::NPE][ND],
                                                            const double sh[ND][Dim<ND>::NPE],
                                                            double s[Dim<ND>::NSTR])
{
    constexpr int NPE = Dim<ND>::NPE;
    const double *sx = sh[0], *sz = sh[ND - 1];
    int n = 0;
    s[n] = 0;
#pragma unroll
    for (int i = 0; i < NPE; ++i) s[n] += v[i][0] * sx[i];
    if constexpr (ND == 3) {
        const double *sy = sh[1];
        s[1] = 0;
#pragma unroll
        for (int i = 0; i < NPE; ++i) s[1] += v[i][1] * sy[i];
        n = 2;
    } else {
        n = 1;
    }
    s[n] = 0;
#pragma unroll
    for (int i = 0; i < NPE; ++i) s[n] += v[i][ND - 1] * sz[i];
    if constexpr (ND == 3) {
        const double *sy = sh[1];
        s[3] = 0;
#pragma unroll
        for (int i = 0; i < NPE; ++i) s[3] += 0.5 * (v[i][0] * sy[i] + v[i][1] * sx[i]);
        s[4] = 0;
#pragma unroll
        for (int i = 0; i < NPE; ++i) s[4] += 0.5 * (v[i][0] * sz[i] + v[i][ND - 1] * sx[i]);
        s[5] = 0;
#pragma unroll
        for (int i = 0; i < NPE; ++i) s[5] += 0.5 * (v[i][1] * sz[i] + v[i][2] * sy[i]);
    } else {
        s[2] = 0;
#pragma unroll
        for (int i = 0; i < NPE; ++i) s[2] += 0.5 * (v[i][0] * sz[i] + v[i][ND - 1] * sx[i]);
    }
}

// strain_rate and, when with_spin, the spin of rotate_stress from the same
// products, each sum in the order of its own function (bitwise equal to
// calling both)
template <int ND> __device__ __forceinline__ void strain_rate_spin(const double v[Dim<ND>::NPE][ND],
                                                                 const double sh[ND][Dim<ND>::NPE],
                                                                 double s[Dim<ND>::NSTR], double w[3],
                                                                 bool with_spin)
{
    constexpr int NPE = Dim<ND>::NPE;
    const double *sx = sh[0], *sz = sh[ND - 1];
    if constexpr (ND == 3) {
        const double *sy = sh[1];
        double s0 = 0, s1 = 0, s2 = 0, s3 = 0, s4 = 0, s5 = 0, w3 = 0, w4 = 0, w5 = 0;
#pragma unroll
        for (int i = 0; i < NPE; ++i) {
            const double xy = v[i][0] * sy[i], yx = v[i][1] * sx[i];
            const double xz = v[i][0] * sz[i], zx = v[i][2] * sx[i];
            const double yz = v[i][1] * sz[i], zy = v[i][2] * sy[i];
            s0 += v[i][0] * sx[i];
            s1 += v[i][1] * sy[i];
            s2 += v[i][2] * sz[i];
            s3 += 0.5 * (xy + yx);
            s4 += 0.5 * (xz + zx);
            s5 += 0.5 * (yz + zy);
            w3 += 0.5 * (xy - yx);
            w4 += 0.5 * (xz - zx);
            w5 += 0.5 * (yz - zy);
        }
        s[0] = s0; s[1] = s1; s[2] = s2; s[3] = s3; s[4] = s4; s[5] = s5;
        w[0] = with_spin ? w3 : 0; w[1] = with_spin ? w4 : 0; w[2] = with_spin ? w5 : 0;
    } else {
        double s0 = 0, s1 = 0, s2 = 0, w2 = 0;
#pragma unroll
        for (int i = 0; i < NPE; ++i) {
            const double xz = v[i][0] * sz[i], zx = v[i][1] * sx[i];
            s0 += v[i][0] * sx[i];
            s1 += v[i][1] * sz[i];
            s2 += 0.5 * (xz + zx);
            w2 += 0.5 * (zx - xz);
        }
        s[0] = s0; s[1] = s1; s[2] = s2;
        w[0] = with_spin ? w2 : 0; w[1] = 0; w[2] = 0;
    }
}

// spin of rotate_stress (fields.cxx:433-456 / 466-475): w[0..2] = w3,w4,w5 or w[0] = w2
template <int ND> __device__ __forceinline__ void spin(const double v[Dim<ND>::NPE][ND],
                                                     const double sh[ND][Dim<ND>::NPE], double w[3])
{
    constexpr int NPE = Dim<ND>::NPE;
    const double *sx = sh[0], *sz = sh[ND - 1];
    if constexpr (ND == 3) {
        const double *sy = sh[1];
        double w3 = 0, w4 = 0, w5 = 0;
#pragma unroll
        for (int i = 0; i < NPE; ++i) w3 += 0.5 * (v[i][0] * sy[i] - v[i][1] * sx[i]);
#pragma unroll
        for (int i = 0; i < NPE; ++i) w4 += 0.5 * (v[i][0] * sz[i] - v[i][ND - 1] * sx[i]);
#pragma unroll
        for (int i = 0; i < NPE; ++i) w5 += 0.5 * (v[i][1] * sz[i] - v[i][ND - 1] * sy[i]);
        w[0] = w3; w[1] = w4; w[2] = w5;
    } else {
        double w2 = 0;
#pragma unroll
        for (int i = 0; i < NPE; ++i) w2 += 0.5 * (v[i][ND - 1] * sx[i] - v[i][0] * sz[i]);
        w[0] = w2; w[1] = 0; w[2] = 0;
    }
}

// jaumann_rate_3d / jaumann_rate_2d (fields.cxx:375-404)
template <int ND> __device__ __forceinline__ void jaumann(double *s, double dt, const double w[3])
{
    if constexpr (ND == 3) {
        const double w3 = w[0], w4 = w[1], w5 = w[2];
        double s_inc[6];
        s_inc[0] = -2.0 * s[3] * w3 - 2.0 * s[4] * w4;
        s_inc[1] = 2.0 * s[3] * w3 - 2.0 * s[5] * w5;
        s_inc[2] = 2.0 * s[4] * w4 + 2.0 * s[5] * w5;
        s_inc[3] = s[0] * w3 - s[1] * w3 - s[4] * w5 - s[5] * w4;
        s_inc[4] = s[0] * w4 - s[2] * w4 + s[3] * w5 - s[5] * w3;
        s_inc[5] = s[1] * w5 - s[2] * w5 + s[3] * w4 + s[4] * w3;
#pragma unroll
        for (int i = 0; i < 6; ++i) s[i] += dt * s_inc[i];
    } else {
        const double w2 = w[0];
        double s_inc[3];
        s_inc[0] = -2.0 * s[2] * w2;
        s_inc[1] = 2.0 * s[2] * w2;
        s_inc[2] = s[0] * w2 - s[1] * w2;
#pragma unroll
        for (int i = 0; i < 3; ++i) s[i] += dt * s_inc[i];
    }
}

// ---------------------------------------------------------- matprops.cxx

__device__ __forceinline__ int count_of(const DevState &s, int e, int m)
{
    if (s.elemmarkers) return s.elemmarkers[e * s.P->nmat + m];
    return m == 0 ? s.mpe : 0;
}

// harmonic_mean / arithmetic_mean with the one-value shortcut (matprops.cxx:102-128)
__device__ __forceinline__ double harmonic_mean(const DevState &s, const double *v, int e)
{
    const int nmat = s.P->nmat;
    if (nmat == 1) return v[0];
    double result = 0;
    int m = 0;
    for (int i = 0; i < nmat; i++) {
        int k = count_of(s, e, i);
        result += k / v[i];
        m += k;
    }
    return m / result;
}

__device__ __forceinline__ double arithmetic_mean(const DevState &s, const double *v, int e)
{
    const int nmat = s.P->nmat;
    if (nmat == 1) return v[0];
    double result = 0;
    int m = 0;
    for (int i = 0; i < nmat; i++) {
        int k = count_of(s, e, i);
        result += k * v[i];
        m += k;
    }
    return result / m;
}

template <int ND> __device__ __forceinline__ double elem_mean(const double *T, const int *c)
{
    double t = 0;
#pragma unroll
    for (int i = 0; i < Dim<ND>::NPE; ++i) t += T[c[i]];
    t /= Dim<ND>::NPE;
    return t;
}

// MatProps::rho (matprops.cxx:375-396) from the element's mean temperature
__device__ __forceinline__ double mat_rho(const DevState &s, int e, double Tmean)
{
    const des_params *P = s.P;
    const double celsius0 = 273;
    double TinCelsius = Tmean - celsius0;
    double result = 0;
    int n = 0;
    for (int m = 0; m < P->nmat; m++) {
        int k = count_of(s, e, m);
        result += P->rho0[m] * (1 - P->alpha[m] * TinCelsius) * k;
        n += k;
    }
    return result / n;
}

// MatProps::visc (matprops.cxx:275-311); a per-element field replaces the power law
template <int ND> __device__ __forceinline__ double mat_visc(const DevState &s, int e, double Tmean,
                                                           const double *edot)
{
    const des_params *P = s.P;
    double visc;
    if (s.elem_visc) {
        visc = s.elem_visc[e];
    } else if (P->visc_min == P->visc_max) {
        // the clamp below maps every non-NaN power-law value to this one
        return P->visc_min;
    } else {
        const double gas_constant = 8.3144;
        const double min_strain_rate = 1e-30;
        double ed = second_invariant<ND>(edot);
        ed = dmax(ed, min_strain_rate);
        double result = 0;
        int n = 0;
        for (int m = 0; m < P->nmat; m++) {
            double pw = 1 / P->visc_exponent[m] - 1;
            double pw1 = -1 / P->visc_exponent[m];
            double visc0 = 0.25 * pow(ed, pw) * pow(0.75 * P->visc_coefficient[m], pw1)
                * exp(P->visc_activation_energy[m] / (P->visc_exponent[m] * gas_constant * Tmean)) * 1e6;
            int k = count_of(s, e, m);
            result += k / visc0;
            n += k;
        }
        visc = n / result;
    }
    return dmin(dmax(visc, P->visc_min), P->visc_max);
}

// MatProps::plastic_weakening + plastic_props (matprops.cxx:314-372)
__device__ __forceinline__ void mat_plastic_props(const DevState &s, int e, double pls,
                                                  double &amc, double &anphi, double &anpsi,
                                                  double &hardn, double &ten_max)
{
    const des_params *P = s.P;
    double c = 0, f = 0, d = 0, h = 0;
    int n = 0;
    for (int m = 0; m < P->nmat; m++) {
        int k = count_of(s, e, m);
        if (k == 0) continue;
        n += k;
        double c0 = s.elem_cohesion ? s.elem_cohesion[e] : P->cohesion0[m];
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
    hardn = h / n;
    double sphi = sin(phi * DEG2RAD);
    double spsi = sin(psi * DEG2RAD);
    anphi = (1 + sphi) / (1 - sphi);
    anpsi = (1 + spsi) / (1 - spsi);
    amc = 2 * cohesion * sqrt(anphi);
    ten_max = (phi == 0) ? P->tension_max : dmin(P->tension_max, cohesion / tan(phi * DEG2RAD));
}

// --------------------------------------------------------------- 3x3-C
// The four routines below restate, step for step, Joachim Kopp's "Numerical
// diagonalization of 3x3 matrices" (3x3-C: dsyevc3, dsytrd3, dsyevq3,
// dsyevh3), Copyright (C) 2006 Joachim Kopp, distributed under the GNU
// Lesser General Public License, version 2.1 or later (3x3-C/COPYING of the
// reference).  The return mapping depends on that code's exact eigenvalue and
// eigenvector choices, so its steps are kept; this restatement is covered by
// the same licence.  Published description: J. Kopp, "Efficient numerical
// diagonalization of hermitian 3x3 matrices", Int. J. Mod. Phys. C 19 (2008)
// 523-548.
#define DES_SQR(x) ((x) * (x))

// Cardano eigenvalues (3x3-C/dsyevc3.c:31-81); A given by its upper triangle
__device__ __forceinline__ void dsyevc3(const double A[3][3], double w[3])
{
    double m, c1, c0;
    double de = A[0][1] * A[1][2];
    double dd = DES_SQR(A[0][1]);
    double ee = DES_SQR(A[1][2]);
    double ff = DES_SQR(A[0][2]);
    m = A[0][0] + A[1][1] + A[2][2];
    c1 = (A[0][0] * A[1][1] + A[0][0] * A[2][2] + A[1][1] * A[2][2]) - (dd + ee + ff);
    c0 = A[2][2] * dd + A[0][0] * ee + A[1][1] * ff - A[0][0] * A[1][1] * A[2][2]
        - 2.0 * A[0][2] * de;
    double p, sqrt_p, q, c, s, phi;
    p = DES_SQR(m) - 3.0 * c1;
    q = m * (p - (3.0 / 2.0) * c1) - (27.0 / 2.0) * c0;
    sqrt_p = sqrt(fabs(p));
    phi = 27.0 * (0.25 * DES_SQR(c1) * (p - c1) + c0 * (q + 27.0 / 4.0 * c0));
    phi = (1.0 / 3.0) * atan2(sqrt(fabs(phi)), q);
    c = sqrt_p * cos(phi);
    s = (1.0 / 1.73205080756887729352744634151) * sqrt_p * sin(phi);
    w[1] = (1.0 / 3.0) * (m - c);
    w[2] = w[1] + s;
    w[0] = w[1] + c;
    w[1] -= s;
}

// Householder tridiagonalisation (3x3-C/dsytrd3.c:28-104)
__device__ __noinline__ void dsytrd3(const double A[3][3], double Q[3][3], double d[3], double e[2])
{
    const int n = 3;
    double u[3], q[3];
    double omega, f, K, h, g;
    for (int i = 0; i < n; i++) {
        Q[i][i] = 1.0;
        for (int j = 0; j < i; j++) Q[i][j] = Q[j][i] = 0.0;
    }
    h = DES_SQR(A[0][1]) + DES_SQR(A[0][2]);
    if (A[0][1] > 0) g = -sqrt(h);
    else g = sqrt(h);
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
        K *= 0.5 * DES_SQR(omega);
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

// QL with implicit shifts, at most 30 sweeps per eigenvalue (3x3-C/dsyevq3.c:29-134)
__device__ __noinline__ int dsyevq3(const double A[3][3], double Q[3][3], double w[3])
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
            r = sqrt(DES_SQR(g) + 1.0);
            if (g > 0) g = w[m] - w[l] + e[l] / (g + r);
            else g = w[m] - w[l] + e[l] / (g - r);
            s = c = 1.0;
            p = 0.0;
            for (int i = m - 1; i >= l; i--) {
                f = s * e[i];
                b = c * e[i];
                if (fabs(f) > fabs(g)) {
                    c = g / f;
                    r = sqrt(DES_SQR(c) + 1.0);
                    e[i + 1] = f * r;
                    c *= (s = 1.0 / r);
                } else {
                    s = f / g;
                    r = sqrt(DES_SQR(s) + 1.0);
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

// Cardano + cross-product eigenvectors, QL fallback (3x3-C/dsyevh3.c:31-134)
__device__ __forceinline__ int dsyevh3(const double A[3][3], double Q[3][3], double w[3])
{
    double norm, error, t, u;
    dsyevc3(A, w);
    t = fabs(w[0]);
    if ((u = fabs(w[1])) > t) t = u;
    if ((u = fabs(w[2])) > t) t = u;
    if (t < 1.0) u = t;
    else u = DES_SQR(t);
    error = 256.0 * DBL_EPSILON * DES_SQR(u);
    Q[0][1] = A[0][1] * A[1][2] - A[0][2] * A[1][1];
    Q[1][1] = A[0][2] * A[0][1] - A[1][2] * A[0][0];
    Q[2][1] = DES_SQR(A[0][1]);
    Q[0][0] = Q[0][1] + A[0][2] * w[0];
    Q[1][0] = Q[1][1] + A[1][2] * w[0];
    Q[2][0] = (A[0][0] - w[0]) * (A[1][1] - w[0]) - Q[2][1];
    norm = DES_SQR(Q[0][0]) + DES_SQR(Q[1][0]) + DES_SQR(Q[2][0]);
    if (norm <= error) return dsyevq3(A, Q, w);
    norm = sqrt(1.0 / norm);
    for (int j = 0; j < 3; j++) Q[j][0] = Q[j][0] * norm;
    Q[0][1] = Q[0][1] + A[0][2] * w[1];
    Q[1][1] = Q[1][1] + A[1][2] * w[1];
    Q[2][1] = (A[0][0] - w[1]) * (A[1][1] - w[1]) - Q[2][1];
    norm = DES_SQR(Q[0][1]) + DES_SQR(Q[1][1]) + DES_SQR(Q[2][1]);
    if (norm <= error) return dsyevq3(A, Q, w);
    norm = sqrt(1.0 / norm);
    for (int j = 0; j < 3; j++) Q[j][1] = Q[j][1] * norm;
    Q[0][2] = Q[1][0] * Q[2][1] - Q[2][0] * Q[1][1];
    Q[1][2] = Q[2][0] * Q[0][1] - Q[0][0] * Q[2][1];
    Q[2][2] = Q[0][0] * Q[1][1] - Q[1][0] * Q[0][1];
    return 0;
}

// principal_stresses3 (rheology.cxx:13-68): eigenpairs in ascending order
__device__ __forceinline__ void principal_stresses3(const double *s, double p[3], double v[3][3])
{
    double a[3][3];
    a[0][0] = s[0]; a[1][1] = s[1]; a[2][2] = s[2];
    a[0][1] = s[3]; a[0][2] = s[4]; a[1][2] = s[5];
    a[1][0] = 0; a[2][0] = 0; a[2][1] = 0;
    dsyevh3(a, v, p);
#pragma unroll
    for (int pass = 0; pass < 3; ++pass) {
        const int i = (pass == 1) ? 1 : 0;
        if (p[i] > p[i + 1]) {
            double tmp = p[i];
            p[i] = p[i + 1];
            p[i + 1] = tmp;
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                double b = v[k][i];
                v[k][i] = v[k][i + 1];
                v[k][i + 1] = b;
            }
        }
    }
}

// principal_stresses2 (rheology.cxx:71-102)
__device__ __forceinline__ void principal_stresses2(const double *s, double p[2], double &cos2t,
                                                    double &sin2t)
{
    double s0 = 0.5 * (s[0] + s[1]);
    double rad = second_invariant<2>(s);
    p[0] = s0 - rad;
    p[1] = s0 + rad;
    const double eps = 1e-15;
    double a = 0.5 * (s[0] - s[1]);
    double b = -rad;
    if (b < -eps) {
        cos2t = a / b;
        sin2t = s[2] / b;
    } else {
        cos2t = 1;
        sin2t = 0;
    }
}

// ---------------------------------------------------------- rheology.cxx

template <int ND> __device__ __forceinline__ void elastic(double bulkm, double shearm, const double *de,
                                                         double *s)  // rheology.cxx:105-115
{
    double lambda = bulkm - 2. / 3 * shearm;
    double dev = trace<ND>(de);
#pragma unroll
    for (int i = 0; i < ND; ++i) s[i] += 2 * shearm * de[i] + lambda * dev;
#pragma unroll
    for (int i = ND; i < Dim<ND>::NSTR; ++i) s[i] += 2 * shearm * de[i];
}

template <int ND> __device__ __forceinline__ void maxwell(double bulkm, double shearm, double viscosity,
                                                         double dt, double dv, const double *de,
                                                         double *s)  // rheology.cxx:118-134
{
    double tmp = 0.5 * dt * shearm / viscosity;
    double f1 = 1 - tmp;
    double f2 = 1 / (1 + tmp);
    double dev = trace<ND>(de) / ND;
    double s0 = trace<ND>(s) / ND;
#pragma unroll
    for (int i = 0; i < ND; ++i)
        s[i] = ((s[i] - s0) * f1 + 2 * shearm * (de[i] - dev)) * f2 + s0 + bulkm * dv;
#pragma unroll
    for (int i = ND; i < Dim<ND>::NSTR; ++i) s[i] = (s[i] * f1 + 2 * shearm * de[i]) * f2;
}

template <int ND> __device__ __forceinline__ void viscous(double bulkm, double viscosity, double total_dv,
                                                         const double *edot, double *s)  // rheology.cxx:137-148
{
    double dev = trace<ND>(edot) / ND;
#pragma unroll
    for (int i = 0; i < ND; ++i) s[i] = 2 * viscosity * (edot[i] - dev) + bulkm * total_dv;
#pragma unroll
    for (int i = ND; i < Dim<ND>::NSTR; ++i) s[i] = 2 * viscosity * edot[i];
}

// Mohr-Coulomb return in principal space of an elastic trial stress s
// (rheology.cxx:167-298)
// Whether the 3D trial stress s is inside the Mohr-Coulomb surface with a
// margin that no rounding of the eigen solve can close.  The eigenvalues of s
// lie within m -+ sqrt(2/3) |s - m I|_F (m = mean stress), so
// fs >= (m - r) - (m + r) anphi + amc and ft <= m + r - ten_max.  The margin,
// 1e-7 of the stress scale, is far above the eigenvalue error of the
// analytic solver (at worst about sqrt(eps) of the scale).  When this holds
// the return leaves s unchanged, exactly as after the solve.
__device__ __forceinline__ bool clearly_elastic3(const double *s, double amc, double anphi, double ten_max)
{
    const double m = (s[0] + s[1] + s[2]) / 3;
    const double d0 = s[0] - m, d1 = s[1] - m, d2 = s[2] - m;
    const double r = sqrt((2.0 / 3) * (d0 * d0 + d1 * d1 + d2 * d2 +
                                       2 * (s[3] * s[3] + s[4] * s[4] + s[5] * s[5])));
    const double tol = 1e-7 * (fabs(m) + r + fabs(amc) + fabs(ten_max)) * (1 + anphi);
    return (m - r) - (m + r) * anphi + amc > tol && (m + r) - ten_max < -tol;
}

template <int ND> __device__ __forceinline__ void mc_return(double bulkm, double shearm, double amc,
                                                           double anphi, double anpsi, double hardn,
                                                           double ten_max, double &depls, double *s)
{
    depls = 0;
    if constexpr (ND == 3)
        if (clearly_elastic3(s, amc, anphi, ten_max)) return;
    double p[ND];
    double v[3][3];
    double cos2t = 0, sin2t = 0;
    if constexpr (ND == 3) principal_stresses3(s, p, v);
    else principal_stresses2(s, p, cos2t, sin2t);
    double fs = p[0] - p[ND - 1] * anphi + amc;
    double ft = p[ND - 1] - ten_max;
    if (fs > 0 && ft < 0) return;
    double pa = sqrt(1 + anphi * anphi) + anphi;
    double ps = ten_max * anphi - amc;
    double h = p[ND - 1] - ten_max + pa * (p[0] - ps);
    double a1 = bulkm + 4. / 3 * shearm;
    double a2 = bulkm - 2. / 3 * shearm;
    double alam;
    if (h < 0) {
        alam = fs / (a1 - a2 * anpsi + a1 * anphi * anpsi - a2 * anphi + 2 * sqrt(anphi) * hardn);
        p[0] -= alam * (a1 - a2 * anpsi);
        if constexpr (ND == 3) p[1] -= alam * (a2 - a2 * anpsi);
        p[ND - 1] -= alam * (a2 - a1 * anpsi);
        if constexpr (ND == 3) depls = fabs(alam) * sqrt((7 + 4 * anpsi + 7 * anpsi * anpsi) / 18);
        else depls = fabs(alam) * sqrt((3 + 2 * anpsi + 3 * anpsi * anpsi) / 8);
    } else {
        alam = ft / a1;
        p[0] -= alam * a2;
        if constexpr (ND == 3) p[1] -= alam * a2;
        p[ND - 1] -= alam * a1;
        if constexpr (ND == 3) depls = fabs(alam) * sqrt(7. / 18);
        else depls = fabs(alam) * sqrt(3. / 8);
    }
    if constexpr (ND == 3) {
        double ss[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
#pragma unroll
        for (int m = 0; m < 3; m++)
#pragma unroll
            for (int n = m; n < 3; n++)
#pragma unroll
                for (int k = 0; k < 3; k++) ss[m][n] += v[m][k] * v[n][k] * p[k];
        s[0] = ss[0][0];
        s[1] = ss[1][1];
        s[2] = ss[2][2];
        s[3] = ss[0][1];
        s[4] = ss[0][2];
        s[5] = ss[1][2];
    } else {
        double dc2 = (p[0] - p[1]) * cos2t;
        double dss = p[0] + p[1];
        s[0] = 0.5 * (dss + dc2);
        s[1] = 0.5 * (dss - dc2);
        s[2] = 0.5 * (p[0] - p[1]) * sin2t;
    }
}

// elasto_plastic: elastic trial then Mohr-Coulomb return (rheology.cxx:151-299)
template <int ND> __device__ __forceinline__ void elasto_plastic(double bulkm, double shearm, double amc,
                                                                double anphi, double anpsi, double hardn,
                                                                double ten_max, const double *de,
                                                                double &depls, double *s)
{
    elastic<ND>(bulkm, shearm, de, s);
    mc_return<ND>(bulkm, shearm, amc, anphi, anpsi, hardn, ten_max, depls, s);
}

// geoFLAC-style plane-strain Mohr-Coulomb (rheology.cxx:302-507)
__device__ __noinline__ void elasto_plastic2d(double bulkm, double shearm, double amc, double anphi,
                                              double anpsi, double hardn, double ten_max,
                                              const double *de, double &depls, double *s, double &syy)
{
    depls = 0;
    double a1 = bulkm + 4. / 3 * shearm;
    double a2 = bulkm - 2. / 3 * shearm;
    double sxx = s[0] + de[1] * a2 + de[0] * a1;
    double szz = s[1] + de[0] * a2 + de[1] * a1;
    double sxz = s[2] + de[2] * 2 * shearm;
    syy += (de[0] + de[1]) * a2;
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
        if (syy > sii) {
            n1 = 0; n2 = 1; n3 = 2;
            p[0] = si; p[1] = sii; p[2] = syy;
        } else if (syy < si) {
            n1 = 1; n2 = 2; n3 = 0;
            p[0] = syy; p[1] = si; p[2] = sii;
        } else {
            n1 = 0; n2 = 2; n3 = 1;
            p[0] = si; p[1] = syy; p[2] = sii;
        }
    }
    if (p[0] >= ten_max) {
        s[0] = s[1] = syy = ten_max;
        s[2] = 0.0;
        return;
    }
    if (p[1] >= ten_max) {
        p[1] = p[2] = ten_max;
    } else if (p[2] >= ten_max) {
        p[2] = ten_max;
    }
    double fs = p[0] - p[2] * anphi + amc;
    if (fs >= 0.0) {
        s[0] = sxx;
        s[1] = szz;
        s[2] = sxz;
        return;
    }
    const double alams = fs / (a1 - a2 * anpsi + a1 * anphi * anpsi - a2 * anphi + hardn);
    p[0] -= alams * (a1 - a2 * anpsi);
    p[1] -= alams * (a2 - a2 * anpsi);
    p[2] -= alams * (a2 - a1 * anpsi);
    depls = 0.5 * fabs(alams + alams * anpsi);
    if (p[0] >= ten_max) {
        s[0] = s[1] = syy = ten_max;
        s[2] = 0.0;
        return;
    }
    if (p[1] >= ten_max) {
        p[1] = p[2] = ten_max;
    } else if (p[2] >= ten_max) {
        p[2] = ten_max;
    }
    double dc2 = (p[n1] - p[n2]) * cos2t;
    double dss = p[n1] + p[n2];
    s[0] = 0.5 * (dss + dc2);
    s[1] = 0.5 * (dss - dc2);
    s[2] = 0.5 * (p[n1] - p[n2]) * sin2t;
    syy = p[n3];
}

template <int ND> __device__ __forceinline__ void plastic_return(double bulkm, double shearm, double amc,
                                                                double anphi, double anpsi, double hardn,
                                                                double ten_max, const double *de,
                                                                double &depls, double *s, double &syy,
                                                                int is_plane_strain)
{
    if constexpr (ND == 2) {
        if (is_plane_strain) {
            elasto_plastic2d(bulkm, shearm, amc, anphi, anpsi, hardn, ten_max, de, depls, s, syy);
            return;
        }
    }
    elasto_plastic<ND>(bulkm, shearm, amc, anphi, anpsi, hardn, ten_max, de, depls, s);
}

// The rheology part of update_stress (rheology.cxx:542-641) for one element
// whose strain rate is already anti-locked and whose strain is already
// updated; total_dv = trace(strain) for the viscous rheology.  Returns false
// for an unknown rheology.  PLASTIC = false leaves the elasto-plastic and
// elasto-visco-plastic branches out (returning false), for a caller that
// handles those rheologies itself: the Mohr-Coulomb return and its eigen
// solve then do not set the kernel's register budget.
template <int ND, bool PLASTIC = true> __device__ __forceinline__ bool rheology_after_strain(const DevState &st, int e, double dt,
                                                                       double total_dv, double Tmean,
                                                                       double vol, double vol_old,
                                                                       double *s, double &syy,
                                                                       const double *edot, double &plstrain,
                                                                       double &delta_plstrain,
                                                                       double &dpressure)
{
    constexpr int NSTR = Dim<ND>::NSTR;
    const des_params *P = st.P;
    double old_s = trace<ND>(s);
    double de[NSTR];
#pragma unroll
    for (int i = 0; i < NSTR; ++i) de[i] = edot[i] * dt;
    bool ok = true;
    switch (P->rheol_type) {
    case DES_RH_ELASTIC:
        elastic<ND>(harmonic_mean(st, P->bulk_modulus, e), harmonic_mean(st, P->shear_modulus, e), de, s);
        break;
    case DES_RH_VISCOUS: {
        double bulkm = harmonic_mean(st, P->bulk_modulus, e);
        double viscosity = mat_visc<ND>(st, e, Tmean, edot);
        viscous<ND>(bulkm, viscosity, total_dv, edot, s);
        break;
    }
    case DES_RH_MAXWELL: {
        double bulkm = harmonic_mean(st, P->bulk_modulus, e);
        double shearm = harmonic_mean(st, P->shear_modulus, e);
        double viscosity = mat_visc<ND>(st, e, Tmean, edot);
        double dv = vol / vol_old - 1;
        maxwell<ND>(bulkm, shearm, viscosity, dt, dv, de, s);
        break;
    }
    case DES_RH_EP: {
        if constexpr (!PLASTIC) {
            ok = false;
            break;
        }
        double depls = 0;
        double bulkm = harmonic_mean(st, P->bulk_modulus, e);
        double shearm = harmonic_mean(st, P->shear_modulus, e);
        double amc, anphi, anpsi, hardn, ten_max;
        mat_plastic_props(st, e, plstrain, amc, anphi, anpsi, hardn, ten_max);
        plastic_return<ND>(bulkm, shearm, amc, anphi, anpsi, hardn, ten_max, de, depls, s, syy,
                           P->is_plane_strain);
        plstrain += depls;
        delta_plstrain = depls;
        break;
    }
    case DES_RH_EVP: {
        if constexpr (!PLASTIC) {
            ok = false;
            break;
        }
        double depls = 0;
        double bulkm = harmonic_mean(st, P->bulk_modulus, e);
        double shearm = harmonic_mean(st, P->shear_modulus, e);
        double viscosity = mat_visc<ND>(st, e, Tmean, edot);
        double dv = vol / vol_old - 1;
        double sv[NSTR];
#pragma unroll
        for (int i = 0; i < NSTR; ++i) sv[i] = s[i];
        maxwell<ND>(bulkm, shearm, viscosity, dt, dv, de, sv);
        double svII = second_invariant2<ND>(sv);
        double amc, anphi, anpsi, hardn, ten_max;
        mat_plastic_props(st, e, plstrain, amc, anphi, anpsi, hardn, ten_max);
        double sp[NSTR], spyy = syy;
#pragma unroll
        for (int i = 0; i < NSTR; ++i) sp[i] = s[i];
        plastic_return<ND>(bulkm, shearm, amc, anphi, anpsi, hardn, ten_max, de, depls, sp, spyy,
                           P->is_plane_strain);
        double spII = second_invariant2<ND>(sp);
        if (svII < spII) {
#pragma unroll
            for (int i = 0; i < NSTR; ++i) s[i] = sv[i];
        } else {
#pragma unroll
            for (int i = 0; i < NSTR; ++i) s[i] = sp[i];
            plstrain += depls;
            delta_plstrain = depls;
            if (P->is_plane_strain) syy = spyy;   // rheology.cxx:632; only plane strain defines spyy
        }
        break;
    }
    default:
        ok = false;
    }
    dpressure = trace<ND>(s) - old_s;
    return ok;
}

// The elasto-plastic / elasto-visco-plastic branch of update_stress
// (rheology.cxx:573-635) for one 3D element, with the Mohr-Coulomb return
// left to a second kernel when the elastic trial is not clearly elastic.
// Returns true for such an element: s then holds the Maxwell stress (EVP) or
// the trial stress (EP), trial and old_s what the return needs, and plstrain
// and delta_plstrain are untouched.  Otherwise the element is complete with
// the values rheology_after_strain gives it: the plastic strain unchanged
// (plstrain += 0) and, when the elastic trial is taken, delta_plstrain = 0
// (reset_dpl; the Maxwell stress of an EVP element leaves it as it is).
__device__ __forceinline__ bool plastic_deferred3(const DevState &st, int e, double dt, double Tmean,
                                                  double vol, double vol_old, double *s, const double *edot,
                                                  double plstrain, bool &reset_dpl,
                                                  double &dpressure, double *trial, double &old_s)
{
    reset_dpl = false;
    const des_params *P = st.P;
    old_s = trace<3>(s);
    double de[6];
#pragma unroll
    for (int i = 0; i < 6; ++i) de[i] = edot[i] * dt;
    const double bulkm = harmonic_mean(st, P->bulk_modulus, e);
    const double shearm = harmonic_mean(st, P->shear_modulus, e);
    const bool evp = P->rheol_type == DES_RH_EVP;
    double sv[6], svII = 0;
    if (evp) {
        const double viscosity = mat_visc<3>(st, e, Tmean, edot);
#pragma unroll
        for (int i = 0; i < 6; ++i) sv[i] = s[i];
        maxwell<3>(bulkm, shearm, viscosity, dt, vol / vol_old - 1, de, sv);
        svII = second_invariant2<3>(sv);
    }
    double amc, anphi, anpsi, hardn, ten_max;
    mat_plastic_props(st, e, plstrain, amc, anphi, anpsi, hardn, ten_max);
#pragma unroll
    for (int i = 0; i < 6; ++i) trial[i] = s[i];
    elastic<3>(bulkm, shearm, de, trial);
    const bool defer = !clearly_elastic3(trial, amc, anphi, ten_max);
    if (!defer && !(evp && svII < second_invariant2<3>(trial))) {
#pragma unroll
        for (int i = 0; i < 6; ++i) s[i] = trial[i];
        reset_dpl = true;
    } else if (evp) {
#pragma unroll
        for (int i = 0; i < 6; ++i) s[i] = sv[i];
    } else {
#pragma unroll
        for (int i = 0; i < 6; ++i) s[i] = trial[i];
    }
    dpressure = trace<3>(s) - old_s;
    return defer;
}

// Body of update_stress for one element (rheology.cxx:520-641).  edot is
// anti-locked in place; returns false for an unknown rheology.
template <int ND> __device__ __forceinline__ bool stress_update(const DevState &st, int e, double dt,
                                                               double edvoldt, double Tmean,
                                                               double vol, double vol_old,
                                                               double *s, double &syy, double *es,
                                                               double *edot, double &plstrain,
                                                               double &delta_plstrain, double &dpressure)
{
    constexpr int NSTR = Dim<ND>::NSTR;
    {
        double div = trace<ND>(edot);
#pragma unroll
        for (int i = 0; i < ND; ++i) edot[i] += (edvoldt - div) / ND;
    }
#pragma unroll
    for (int i = 0; i < NSTR; ++i) es[i] += edot[i] * dt;
    return rheology_after_strain<ND>(st, e, dt, trace<ND>(es), Tmean, vol, vol_old, s, syy, edot, plstrain,
                                     delta_plstrain, dpressure);
}

// ===================================================== mixed precision
//
// The stress is carried as an FP64 mean p of its diagonal plus an FP32
// remainder sd (the deviator); strain, plastic strain and strain rate are
// FP32.  The rheology arithmetic on sd, the strain and the material
// properties runs in FP32; the pressure, the volume change and the
// Mohr-Coulomb eigen-return run in FP64.  At ~3e8 Pa of lithostatic pressure
// an FP32 total stress would round away the ~1e2-1e3 Pa per-step increments.

template <int ND> __device__ __forceinline__ float trace_f(const float *s)
{
    if constexpr (ND == 3) return s[0] + s[1] + s[2];
    else return s[0] + s[1];
}

template <int ND> __device__ __forceinline__ float second_invariant2_f(const float *t)
{
    if constexpr (ND == 3) {
        float a = (t[0] + t[1] + t[2]) / 3;
        return (0.5f * ((t[0] - a) * (t[0] - a) + (t[1] - a) * (t[1] - a) + (t[2] - a) * (t[2] - a))
                + t[3] * t[3] + t[4] * t[4] + t[5] * t[5]);
    } else {
        return 0.25f * (t[0] - t[1]) * (t[0] - t[1]) + t[2] * t[2];
    }
}

// Jaumann increment dt (S W - W S) of a symmetric tensor (fields.cxx:375-404)
template <int ND, typename T> __device__ __forceinline__ void jaumann_inc(const T *s, T dt, const T w[3],
                                                                         T *inc)
{
    if constexpr (ND == 3) {
        const T w3 = w[0], w4 = w[1], w5 = w[2];
        inc[0] = dt * (T(-2) * s[3] * w3 - T(2) * s[4] * w4);
        inc[1] = dt * (T(2) * s[3] * w3 - T(2) * s[5] * w5);
        inc[2] = dt * (T(2) * s[4] * w4 + T(2) * s[5] * w5);
        inc[3] = dt * (s[0] * w3 - s[1] * w3 - s[4] * w5 - s[5] * w4);
        inc[4] = dt * (s[0] * w4 - s[2] * w4 + s[3] * w5 - s[5] * w3);
        inc[5] = dt * (s[1] * w5 - s[2] * w5 + s[3] * w4 + s[4] * w3);
    } else {
        const T w2 = w[0];
        inc[0] = dt * (T(-2) * s[2] * w2);
        inc[1] = dt * (T(2) * s[2] * w2);
        inc[2] = dt * (s[0] * w2 - s[1] * w2);
    }
}

// MatProps::visc in FP32 (power law of matprops.cxx:275-311, same clamp)
template <int ND> __device__ __forceinline__ float mat_visc_f(const DevState &s, int e, float Tmean,
                                                            const float *edot)
{
    const des_params *P = s.P;
    float visc;
    if (s.visc32) {
        visc = s.visc32[e];
    } else if (P->visc_min == P->visc_max) {
        visc = (float)P->visc_min;
    } else {
        float ed = sqrtf(second_invariant2_f<ND>(edot));
        ed = fmaxf(ed, 1e-30f);
        float result = 0;
        int n = 0;
        for (int m = 0; m < P->nmat; m++) {
            const float ne = (float)P->visc_exponent[m];
            float visc0 = 0.25f * powf(ed, 1 / ne - 1) * powf(0.75f * (float)P->visc_coefficient[m], -1 / ne)
                * expf((float)P->visc_activation_energy[m] / (ne * 8.3144f * Tmean)) * 1e6f;
            int k = count_of(s, e, m);
            result += k / visc0;
            n += k;
        }
        visc = n / result;
    }
    const float vmin = (float)P->visc_min, vmax = (float)P->visc_max;
    return fminf(fmaxf(visc, vmin), vmax);
}

// plastic_weakening + plastic_props in FP32 (matprops.cxx:314-372)
__device__ __forceinline__ void mat_plastic_props_f(const DevState &s, int e, float pls, double &amc,
                                                    double &anphi, double &anpsi, double &hardn,
                                                    double &ten_max)
{
    const des_params *P = s.P;
    float c = 0, f = 0, d = 0, h = 0;
    int n = 0;
    for (int m = 0; m < P->nmat; m++) {
        int k = count_of(s, e, m);
        if (k == 0) continue;
        n += k;
        const float c0 = s.coh32 ? s.coh32[e] : (float)P->cohesion0[m];
        const float c1 = (float)P->cohesion1[m];
        const float p0 = (float)P->pls0[m], p1 = (float)P->pls1[m];
        const float f0 = (float)P->friction_angle0[m], f1 = (float)P->friction_angle1[m];
        const float d0 = (float)P->dilation_angle0[m], d1 = (float)P->dilation_angle1[m];
        if (pls < p0) {
            c += c0 * k; f += f0 * k; d += d0 * k;
        } else if (pls < p1) {
            float p = (pls - p0) / (p1 - p0);
            c += (c0 + p * (c1 - c0)) * k;
            f += (f0 + p * (f1 - f0)) * k;
            d += (d0 + p * (d1 - d0)) * k;
            h += (c1 - c0) / (p1 - p0) * k;
        } else {
            c += c1 * k; f += f1 * k; d += d1 * k;
        }
    }
    float cohesion = c / n, phi = f / n, psi = d / n;
    hardn = h / n;
    const float deg = (float)DEG2RAD;
    float sphi = sinf(phi * deg), spsi = sinf(psi * deg);
    float aphi = (1 + sphi) / (1 - sphi), apsi = (1 + spsi) / (1 - spsi);
    anphi = aphi;
    anpsi = apsi;
    amc = 2 * cohesion * sqrtf(aphi);
    ten_max = (phi == 0) ? P->tension_max : dmin(P->tension_max, (double)(cohesion / tanf(phi * deg)));
}

// The FP32 deviator is only ever changed by adding one increment per step:
// the Jaumann rotation rs (~1e-8 of the stress, below an FP32 ulp on its own)
// and the rheological increment are summed first and rounded once.

// elastic increment (rheology.cxx:105-115) of (p, sd): the diagonal mean
// takes (lambda + 2G/ND) tr(de), the deviator 2G (de - tr(de)/ND)
template <int ND> __device__ __forceinline__ void elastic_incr(float K, float G, const float *de,
                                                              const float *rs, double &p, float *inc)
{
    const float lambda = K - 2.f / 3 * G;
    const float tr = trace_f<ND>(de), mean = tr / ND;
#pragma unroll
    for (int i = 0; i < ND; ++i) inc[i] = rs[i] + 2 * G * (de[i] - mean);
#pragma unroll
    for (int i = ND; i < Dim<ND>::NSTR; ++i) inc[i] = rs[i] + 2 * G * de[i];
    p += (double)((lambda + 2 * G / ND) * tr);
}

// maxwell (rheology.cxx:118-134) in increment form: with t = dt G / (2 eta),
// s' = (s (1 - t) + 2G de') / (1 + t) = s + (2G de' - 2t s) / (1 + t); the
// relaxation 2t s keeps FP32's relative precision where 1 - t would round t
// to ~1e-3 of itself.  The diagonal mean takes only the bulk term.
template <int ND> __device__ __forceinline__ void maxwell_incr(float K, float G, float eta, float dt,
                                                              double dv, const float *de, const float *sd,
                                                              const float *rs, double &p, float *inc)
{
    const float t = 0.5f * dt * G / eta;
    const float f2 = 1 / (1 + t);
    const float dev = trace_f<ND>(de) / ND;
#pragma unroll
    for (int i = 0; i < ND; ++i) inc[i] = rs[i] + (2 * G * (de[i] - dev) - 2 * t * sd[i] - 2 * t * rs[i]) * f2;
#pragma unroll
    for (int i = ND; i < Dim<ND>::NSTR; ++i) inc[i] = rs[i] + (2 * G * de[i] - 2 * t * sd[i] - 2 * t * rs[i]) * f2;
    p += (double)K * dv;
}

template <int ND> __device__ __forceinline__ void join_stress(double p, const float *sd, double *s)
{
#pragma unroll
    for (int i = 0; i < ND; ++i) s[i] = (double)sd[i] + p;
#pragma unroll
    for (int i = ND; i < Dim<ND>::NSTR; ++i) s[i] = (double)sd[i];
}

template <int ND> __device__ __forceinline__ void split_stress(const double *s, double &p, float *sd)
{
    p = trace<ND>(s) / ND;
#pragma unroll
    for (int i = 0; i < ND; ++i) sd[i] = (float)(s[i] - p);
#pragma unroll
    for (int i = ND; i < Dim<ND>::NSTR; ++i) sd[i] = (float)s[i];
}

// update_stress for one element in mixed precision (plane strain excluded).
// wf, dtp, rot: the spin and dt of the previous step's Jaumann rotation,
// owed when rot is set.  Its increments of the strain (re) and of the deviator
// (rs) are folded into this step's single rounding, each formed where it is
// used, and the updated strain es is stored to strain32 as soon as it is
// formed: carried from the caller instead, they kept ~45 more registers live
// through the rheology (4 waves per SIMD instead of 3).
//
// DEFER (3D): an elasto-(visco-)plastic element whose elastic trial is not
// clearly elastic (clearly_elastic3) fills rec and sets deferred; p, sd, pl
// and dpl are then untouched and the return kernel completes the element.
// reset_dpl: dpl was zeroed (the elastic trial taken); in 3D nothing else
// here changes pl or dpl.
// Every other element gets exactly what the inline return gives it, since
// the return leaves a clearly elastic trial unchanged.
template <int ND, bool DEFER = false>
__device__ __forceinline__ bool stress_update_mixed(const DevState &st, int e, float dtf, float edvoldt,
                                                    float Tmean, double vol, double vol_old, double &p,
                                                    float *sd, float *es, const float *wf, float dtp, bool rot,
                                                    float *edot, double &pl, float &dpl, bool &reset_dpl,
                                                    double &dpressure, bool &deferred, MixedRet &rec)
{
    deferred = false;
    reset_dpl = false;
    constexpr int NSTR = Dim<ND>::NSTR;
    const des_params *P = st.P;
    const double p_old = p;
    {
        float div = trace_f<ND>(edot);
#pragma unroll
        for (int i = 0; i < ND; ++i) edot[i] += (edvoldt - div) / ND;
    }
    float de[NSTR];
    {
        float re[NSTR];
        jaumann_inc<ND, float>(es, dtp, wf, re);
#pragma unroll
        for (int i = 0; i < NSTR; ++i) {
            de[i] = edot[i] * dtf;
            es[i] += (rot ? re[i] : 0.f) + de[i];
        }
    }
    const float tr_es = trace_f<ND>(es);
#pragma unroll
    for (int i = 0; i < NSTR; ++i) __builtin_nontemporal_store(es[i], st.strain32 + i * (long)st.E + e);
    const float K = (float)harmonic_mean(st, P->bulk_modulus, e);
    const float G = (float)harmonic_mean(st, P->shear_modulus, e);
    float inc[NSTR];
    float rs[NSTR];
    jaumann_inc<ND, float>(sd, dtp, wf, rs);
#pragma unroll
    for (int i = 0; i < NSTR; ++i) rs[i] = rot ? rs[i] : 0.f;
    bool ok = true;
    switch (P->rheol_type) {
    case DES_RH_ELASTIC:
        elastic_incr<ND>(K, G, de, rs, p, inc);
#pragma unroll
        for (int i = 0; i < NSTR; ++i) sd[i] += inc[i];
        break;
    case DES_RH_VISCOUS: {
        float eta = mat_visc_f<ND>(st, e, Tmean, edot);
        float dev = trace_f<ND>(edot) / ND;
#pragma unroll
        for (int i = 0; i < ND; ++i) sd[i] = 2 * eta * (edot[i] - dev);
#pragma unroll
        for (int i = ND; i < NSTR; ++i) sd[i] = 2 * eta * edot[i];
        p = (double)K * (double)tr_es;
        break;
    }
    case DES_RH_MAXWELL: {
        float eta = mat_visc_f<ND>(st, e, Tmean, edot);
        maxwell_incr<ND>(K, G, eta, dtf, vol / vol_old - 1, de, sd, rs, p, inc);
#pragma unroll
        for (int i = 0; i < NSTR; ++i) sd[i] += inc[i];
        break;
    }
    case DES_RH_EP:
    case DES_RH_EVP: {
        double pv = p;
        float incv[NSTR];
#pragma unroll
        for (int i = 0; i < NSTR; ++i) incv[i] = 0;
        float svII = 0;
        if (P->rheol_type == DES_RH_EVP) {
            float eta = mat_visc_f<ND>(st, e, Tmean, edot);
            maxwell_incr<ND>(K, G, eta, dtf, vol / vol_old - 1, de, sd, rs, pv, incv);
            float sv[NSTR];
#pragma unroll
            for (int i = 0; i < NSTR; ++i) sv[i] = sd[i] + incv[i];
            svII = second_invariant2_f<ND>(sv);
        }
        double amc, anphi, anpsi, hardn, ten_max;
        mat_plastic_props_f(st, e, (float)pl, amc, anphi, anpsi, hardn, ten_max);
        double pp = p;
        elastic_incr<ND>(K, G, de, rs, pp, inc);
        double s64[NSTR];
#pragma unroll
        for (int i = 0; i < ND; ++i) s64[i] = (double)sd[i] + (double)inc[i] + pp;
#pragma unroll
        for (int i = ND; i < NSTR; ++i) s64[i] = (double)sd[i] + (double)inc[i];
        if constexpr (DEFER && ND == 3) {
            if (!clearly_elastic3(s64, amc, anphi, ten_max)) {
                deferred = true;
#pragma unroll
                for (int i = 0; i < 6; ++i) {
                    rec.s64[i] = s64[i];
                    rec.inc[i] = inc[i];
                    rec.incv[i] = incv[i];
                }
                rec.pp = pp;
                rec.pv = pv;
                rec.p_old = p_old;
                rec.svII = svII;
                dpressure = 0;
                return ok;
            }
            // clearly elastic: the return would leave s64 as it is, depls = 0
        }
        double depls = 0;
        if constexpr (!(DEFER && ND == 3))
            mc_return<ND>((double)K, (double)G, amc, anphi, anpsi, hardn, ten_max, depls, s64);
        if (P->rheol_type == DES_RH_EVP && (double)svII < second_invariant2<ND>(s64)) {
            p = pv;
#pragma unroll
            for (int i = 0; i < NSTR; ++i) sd[i] += incv[i];
        } else if (depls == 0) {
            // elastic trial accepted: keep the single-rounding increment
            p = pp;
#pragma unroll
            for (int i = 0; i < NSTR; ++i) sd[i] += inc[i];
            dpl = 0;
            reset_dpl = true;
        } else {
            split_stress<ND>(s64, p, sd);
            pl += depls;                 // accumulated in FP64: increments below an FP32 ulp count
            dpl = (float)depls;
        }
        break;
    }
    default:
        ok = false;
    }
    dpressure = ND * (p - p_old);
    return ok;
}

}  // namespace des
