// des_amd.hip -- HIP kernels for gfx950 and the C-ABI (include/des_amd.h) of
// the DynEarthSol explicit time step.
//
// Two ways through a step:
//   * per-function kernels, one per reference function, in the order of the
//     reference's main loop (des_step, dynearthsol.cxx:325-361);
//   * the fused six-pass schedule (des_run): three element passes and three
//     node passes per step, shape functions and strain rates recomputed in
//     registers instead of stored, mass / volume_n / tmass built inside the
//     node pass that first needs them, and the Jaumann rotation of step k
//     applied when step k+1 first reads the stress.
// Element->node sums are gathers over the node->element CSR in ascending
// element order -- the reference's serial summation order -- so no atomics and
// bitwise run-to-run determinism.
#include <hip/hip_runtime.h>
#include <hip/hip_ext.h>
#include <hip/hip_cooperative_groups.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <initializer_list>
#include <mutex>
#include <thread>
#include <vector>

#include "des_amd.h"
#include "des_device.hpp"

using namespace des;

namespace {

constexpr int BLOCK = 256;
constexpr int RED_BLOCKS = 1024;

inline int nblocks(long n) { return (int)((n + BLOCK - 1) / BLOCK); }

// ---------------------------------------------------------------- kernels

// Workgroups are dispatched round-robin over the 8 XCDs, each with its own
// L2, in ascending workgroup order.  Renumbering them so that XCD x runs one
// contiguous range of logical blocks gives every XCD the same contiguous slab
// of the (renumbered) mesh in every pass: what one pass writes is re-read
// from the same L2 by the next.  With rev set, each XCD walks its range from
// the top down: a pass that follows one walking upwards starts where that one
// ended, on the lines it wrote last, which are still in the L2 and the 256 MB
// Infinity Cache when the pass's data are larger than both.
__device__ __forceinline__ int xcd_block(int rev = 0)
{
    const int nb = gridDim.x, b = blockIdx.x;
    const int q = nb >> 3, r = nb & 7, x = b & 7;
    int j = b >> 3;
    if (rev) j = q + (x < r ? 1 : 0) - 1 - j;    // XCD x has q (+1 if x < r) blocks
    return x * q + min(x, r) + j;
}

// Element / node index of this thread.  For batched models the block's model
// supplies the parameters and scalars (the kernel's by-value DevState is
// rebound), and padding lanes past the model's last element return.
#define BIND_ELEM_AT(s, e, BLK)                        \
    const int blk_ = (BLK);                            \
    const int e = blk_ * BLOCK + threadIdx.x;          \
    if (s.eblk_model) {                                \
        if (e >= s.eblk_lim[blk_]) return;             \
        const int m_ = s.eblk_model[blk_];             \
        s.P = s.Pb + m_;                               \
        s.sc = s.scb + m_;                             \
    } else if (e >= s.E) {                             \
        return;                                        \
    }
#define BIND_ELEM(s, e) BIND_ELEM_AT(s, e, xcd_block(s.rev))

#define BIND_NODE(s, n)                                \
    const int blk_ = xcd_block(s.rev);                 \
    const int n = blk_ * BLOCK + threadIdx.x;          \
    if (s.nblk_model) {                                \
        if (n >= s.nblk_lim[blk_]) return;             \
        const int m_ = s.nblk_model[blk_];             \
        bind_model<ND>(s, m_);                         \
    } else if (n >= s.N) {                             \
        return;                                        \
    }

// The element passes stage their block's node records in LDS: the block's
// distinct nodes (a compact patch of the curve-ordered mesh) are loaded once,
// four lanes per 64-byte record, instead of four scattered records per
// element.  Every lane of the block calls this before any lane returns.
__device__ __forceinline__ bool stage_block(const DevState &s, int blk, double2 *lds, int *loff = nullptr)
{
    if (!s.stage) return false;
    const int2 st = s.stage[blk];
    if (st.y < 0) return false;
    const double2 *src = reinterpret_cast<const double2 *>(s.nrec);
    for (int k = threadIdx.x; k < (NREC / 2) * st.y; k += BLOCK)
        lds[k] = src[(long)s.stage_node[st.x + k / (NREC / 2)] * (NREC / 2) + k % (NREC / 2)];
    if (loff && s.srank)       // and where each staged node's list starts
        for (int k = threadIdx.x; k < st.y; k += BLOCK) loff[k] = s.csr_off[s.stage_node[st.x + k]];
    __syncthreads();
    return true;
}

// the records of element e's nodes, from the staged block (pli: the
// element's stage_lidx entry when the caller loaded it) or from nrec
template <int ND> __device__ __forceinline__ void load_elem_nodes(const DevState &s, int e, bool staged,
                                                                const double2 *lds,
                                                                double x[Dim<ND>::NPE][ND],
                                                                double v[Dim<ND>::NPE][ND],
                                                                double T[Dim<ND>::NPE],
                                                                double aux[Dim<ND>::NPE], const unsigned *pli = nullptr)
{
    if (staged) {
        const unsigned li = pli ? *pli : s.stage_lidx[e];
        const lds_ptr<double2> l = (lds_ptr<double2>)lds;
#pragma unroll
        for (int i = 0; i < Dim<ND>::NPE; ++i)
            nrec_unpack<ND>(l + (NREC / 2) * ((li >> (8 * i)) & 255u), x[i], v[i], T[i], aux[i]);
    } else {
        int c[Dim<ND>::NPE];
        load_conn<ND>(s, e, c);
        load_nrec<ND>(s.nrec, c, x, v, T, aux);
    }
}

// Stores of per-lane records (an element's erec, a node's nrec) at a stride
// of R 16-byte chunks cost one L2 write request per lane and chunk: 16 bytes
// per request.  A full wave instead passes its 64 records through LDS and
// writes them back so that store k's lane l writes the wave's chunk
// 64k + l: consecutive lanes write consecutive chunks and a store
// instruction fills whole 64-byte segments.  Chunks C..R-1 of each record
// are left as they are.  True when the wave was full and the records are
// written; false (nothing written) for a partial wave, whose lanes write
// their own records.  wbuf: this wave's 64*C chunks of LDS.
__device__ __forceinline__ bool wave_full() { return __builtin_amdgcn_read_exec() == ~0ull; }

template <int C, int R>
__device__ __forceinline__ bool wave_store_records(double2 *dst, long r0, const double2 (&v)[C], double2 *wbuf)
{
    if (!wave_full()) return false;
    const int l = __lane_id();
#pragma unroll
    for (int k = 0; k < C; ++k) wbuf[l * C + k] = v[k];
    // the lanes read what other lanes wrote: LDS serves one wave's
    // instructions in order, and the fences keep the compiler from moving
    // the reads above the writes
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
#pragma unroll
    for (int k = 0; k < C; ++k) {
        const int j = k * 64 + l;
        dst[(r0 + j / C) * R + j % C] = wbuf[j];
    }
    return true;
}

// Node lane n of a kernel launched in blocks of NB threads (NB divides
// BLOCK); a batched group's block tables are kept per BLOCK nodes, entry tb_.
#define BIND_NODE_NB(s, n, NB) BIND_NODE_NB_AT(s, n, NB, xcd_block(s.rev))
#define BIND_NODE_NB_AT(s, n, NB, BLK)                 \
    const int n = (BLK) * (NB) + threadIdx.x;          \
    const int tb_ = n / BLOCK;                         \
    if (s.nblk_model) {                                \
        if (n >= s.nblk_lim[tb_]) return;              \
        bind_model<ND>(s, s.nblk_model[tb_]);          \
    } else if (n >= s.N) {                             \
        return;                                        \
    }

// parameters, scalars and boundary geometry of batched model m
template <int ND> __device__ __forceinline__ void bind_model(DevState &s, int m)
{
    s.P = s.Pb + m;
    s.sc = s.scb + m;
    s.bnormals = s.bnb + m * (DES_NBDRY * ND);
    s.edge_vectors = s.evb + m * (DES_NBDRY * DES_NBDRY * ND);
    s.edge_valid = s.evvb + m * (DES_NBDRY * DES_NBDRY);
}

// compute_volume (geometry.cxx:86-108)
template <int ND> __global__ __launch_bounds__(BLOCK) void k_volume(DevState s)
{
    BIND_ELEM(s, e);
    constexpr int NPE = Dim<ND>::NPE;
    int c[NPE];
    load_conn<ND>(s, e, c);
    double x[NPE][ND];
    gather_vec<ND>(s.coord, s.N, c, x);
    s.volume[e] = elem_volume<ND>(x);
}

// compute_shape_fn (geometry.cxx:368-455)
template <int ND> __global__ __launch_bounds__(BLOCK) void k_shape_fn(DevState s)
{
    BIND_ELEM(s, e);
    constexpr int NPE = Dim<ND>::NPE;
    int c[NPE];
    load_conn<ND>(s, e, c);
    double x[NPE][ND], sh[ND][NPE];
    gather_vec<ND>(s.coord, s.N, c, x);
    shape_fn<ND>(x, s.volume[e], sh);
#pragma unroll
    for (int d = 0; d < ND; ++d)
#pragma unroll
        for (int i = 0; i < NPE; ++i) s.shp[(d * NPE + i) * (long)s.E + e] = sh[d][i];
}

template <int ND> __device__ __forceinline__ void load_shp(const DevState &s, int e,
                                                         double sh[ND][Dim<ND>::NPE])
{
    constexpr int NPE = Dim<ND>::NPE;
#pragma unroll
    for (int d = 0; d < ND; ++d)
#pragma unroll
        for (int i = 0; i < NPE; ++i) sh[d][i] = s.shp[(d * NPE + i) * (long)s.E + e];
}

// element part of compute_mass (geometry.cxx:346-352): m and tm of element e
template <int ND> __device__ __forceinline__ void mass_of(const DevState &s, int e, double vol,
                                                        double Tmean, double &m, double &tm)
{
    constexpr int NPE = Dim<ND>::NPE;
    const des_params *P = s.P;
    const double pseudo_speed = P->max_vbc_val * P->inertial_scaling;
    double rho_e = mat_rho(s, e, Tmean);
    double rho = P->is_quasi_static ? harmonic_mean(s, P->bulk_modulus, e) / (pseudo_speed * pseudo_speed)
                                    : rho_e;
    m = rho * vol / NPE;
    tm = rho_e * arithmetic_mean(s, P->heat_capacity, e) * vol / NPE;
}

template <int ND> __global__ __launch_bounds__(BLOCK) void k_mass_elem(DevState s)
{
    BIND_ELEM(s, e);
    int c[Dim<ND>::NPE];
    load_conn<ND>(s, e, c);
    double m, tm;
    mass_of<ND>(s, e, s.volume[e], elem_mean<ND>(s.T, c), m, tm);
    s.escal[1 * (long)s.E + e] = m;
    s.escal[2 * (long)s.E + e] = tm;
}

// node part of compute_mass (geometry.cxx:353-359)
template <int ND> __global__ __launch_bounds__(BLOCK) void k_mass_node(DevState s)
{
    BIND_NODE(s, n);
    constexpr int NPE = Dim<ND>::NPE;
    const int thermal = s.P->has_thermal_diffusion;
    double vn = 0, m = 0, tm = 0;
    for (int p = s.csr_off[n]; p < s.csr_off[n + 1]; ++p) {
        int e = s.csr_ent[p] / NPE;
        vn += s.volume[e];
        m += s.escal[1 * (long)s.E + e];
        if (thermal) tm += s.escal[2 * (long)s.E + e];
    }
    s.volume_n[n] = vn;
    s.mass[n] = m;
    s.tmass[n] = tm;
}

// thermal contribution of element e to tdot at its nodes (fields.cxx:110-142)
template <int ND> __device__ __forceinline__ void temperature_contrib(const DevState &s, int e, double vol,
                                                                    const double sh[ND][Dim<ND>::NPE],
                                                                    const double Tn[Dim<ND>::NPE],
                                                                    double out[Dim<ND>::NPE])
{
    constexpr int NPE = Dim<ND>::NPE;
    const des_params *P = s.P;
    double kv = arithmetic_mean(s, P->therm_cond, e) * vol;
    double rh = (arithmetic_mean(s, P->radiogenic_heating, e) / arithmetic_mean(s, P->heat_capacity, e)) * vol;
    double D[NPE][NPE];
#pragma unroll
    for (int i = 0; i < NPE; ++i)
#pragma unroll
        for (int j = 0; j < NPE; ++j) {
            if constexpr (ND == 3) D[i][j] = (sh[0][i] * sh[0][j] + sh[1][i] * sh[1][j] + sh[2][i] * sh[2][j]);
            else D[i][j] = (sh[0][i] * sh[0][j] + sh[1][i] * sh[1][j]);
        }
#pragma unroll
    for (int i = 0; i < NPE; ++i) {
        double diffusion = 0;
#pragma unroll
        for (int j = 0; j < NPE; ++j) diffusion += D[i][j] * Tn[j];
        out[i] = diffusion * kv - rh;
    }
}

template <int ND> __global__ __launch_bounds__(BLOCK) void k_temp_elem(DevState s)
{
    BIND_ELEM(s, e);
    constexpr int NPE = Dim<ND>::NPE;
    int c[NPE];
    load_conn<ND>(s, e, c);
    double sh[ND][NPE], Tn[NPE], out[NPE];
    load_shp<ND>(s, e, sh);
#pragma unroll
    for (int i = 0; i < NPE; ++i) Tn[i] = s.T[c[i]];
    temperature_contrib<ND>(s, e, s.volume[e], sh, Tn, out);
#pragma unroll
    for (int i = 0; i < NPE; ++i) s.ework[e * NPE + i] = out[i];
}

// temperature update with its boundary conditions (fields.cxx:153-162)
__device__ __forceinline__ double temperature_new_f(const DevState &s, unsigned f, double T, double tdot,
                                                    double dt, double tmass)
{
    const des_params *P = s.P;
    if (f & BOUNDZ1) return P->surface_temperature;
    if ((f & BOUNDZ0) && !P->has_fixed_bottom_temperature) return P->mantle_temperature;
    return T - tdot * dt / tmass;
}

__device__ __forceinline__ double temperature_new(const DevState &s, int n, double T, double tdot,
                                                  double dt, double tmass)
{
    return temperature_new_f(s, s.bcflag[n], T, tdot, dt, tmass);
}

template <int ND> __global__ __launch_bounds__(BLOCK) void k_temp_node(DevState s)
{
    BIND_NODE(s, n);
    double tdot = 0;
    for (int p = s.csr_off[n]; p < s.csr_off[n + 1]; ++p) tdot += s.ework[s.csr_ent[p]];
    s.ntmp[n] = tdot;
    s.T[n] = temperature_new(s, n, s.T[n], tdot, s.sc->dt, s.tmass[n]);
}

template <int ND> __global__ __launch_bounds__(BLOCK) void k_strain_rate(DevState s)
{
    BIND_ELEM(s, e);
    constexpr int NPE = Dim<ND>::NPE, NSTR = Dim<ND>::NSTR;
    int c[NPE];
    load_conn<ND>(s, e, c);
    double v[NPE][ND], sh[ND][NPE], sr[NSTR];
    gather_vec<ND>(s.vel, s.N, c, v);
    load_shp<ND>(s, e, sh);
    strain_rate<ND>(v, sh, sr);
#pragma unroll
    for (int k = 0; k < NSTR; ++k) s.strain_rate[k * (long)s.E + e] = sr[k];
}

// compute_dvoldt (geometry.cxx:111-155), into ntmp
template <int ND> __global__ __launch_bounds__(BLOCK) void k_dvoldt_node(DevState s)
{
    BIND_NODE(s, n);
    constexpr int NPE = Dim<ND>::NPE;
    double sum = 0;
    for (int p = s.csr_off[n]; p < s.csr_off[n + 1]; ++p) {
        int e = s.csr_ent[p] / NPE;
        double sr[ND];
#pragma unroll
        for (int k = 0; k < ND; ++k) sr[k] = s.strain_rate[k * (long)s.E + e];
        double dj = trace<ND>(sr);
        sum += dj * s.volume[e];
    }
    s.ntmp[n] = sum / s.volume_n[n];
}

// compute_edvoldt (geometry.cxx:158-179)
template <int ND> __global__ __launch_bounds__(BLOCK) void k_edvoldt(DevState s)
{
    BIND_ELEM(s, e);
    int c[Dim<ND>::NPE];
    load_conn<ND>(s, e, c);
    double dj = 0;
#pragma unroll
    for (int i = 0; i < Dim<ND>::NPE; ++i) dj += s.ntmp[c[i]];
    s.edvoldt[e] = dj / Dim<ND>::NPE;
}

// Non-temporal variants for arrays a pass streams once per step and no pass
// reads again before the next step's same pass (strain, plastic strain, the
// SoA velocity and coordinates): they keep the L2 for the node records and
// contributions that the following passes gather.
template <int ND> __device__ __forceinline__ void load_tensor_nt(const double *a, long E, int e,
                                                               double t[Dim<ND>::NSTR])
{
#pragma unroll
    for (int k = 0; k < Dim<ND>::NSTR; ++k) t[k] = __builtin_nontemporal_load(a + k * E + e);
}

template <int ND> __device__ __forceinline__ void store_tensor_nt(double *a, long E, int e,
                                                                const double t[Dim<ND>::NSTR])
{
#pragma unroll
    for (int k = 0; k < Dim<ND>::NSTR; ++k) __builtin_nontemporal_store(t[k], a + k * E + e);
}

template <int ND> __device__ __forceinline__ void load_tensor(const double *a, long E, int e,
                                                            double t[Dim<ND>::NSTR])
{
#pragma unroll
    for (int k = 0; k < Dim<ND>::NSTR; ++k) t[k] = a[k * E + e];
}

template <int ND> __device__ __forceinline__ void store_tensor(double *a, long E, int e,
                                                             const double t[Dim<ND>::NSTR])
{
#pragma unroll
    for (int k = 0; k < Dim<ND>::NSTR; ++k) a[k * E + e] = t[k];
}

__device__ __forceinline__ void set_error(DevScalars *sc, int code)
{
    atomicCAS(&sc->err, 0, code);
}

// update_stress (rheology.cxx:510-646)
template <int ND> __global__ __launch_bounds__(BLOCK) void k_update_stress(DevState s)
{
    BIND_ELEM(s, e);
    constexpr int NPE = Dim<ND>::NPE, NSTR = Dim<ND>::NSTR;
    const long E = s.E;
    int c[NPE];
    load_conn<ND>(s, e, c);
    double st[NSTR], es[NSTR], edot[NSTR];
    load_tensor<ND>(s.stress, E, e, st);
    load_tensor<ND>(s.strain, E, e, es);
    load_tensor<ND>(s.strain_rate, E, e, edot);
    double syy = s.stressyy[e], pl = s.plstrain[e], dpl = s.delta_plstrain[e], dp;
    double Tm = elem_mean<ND>(s.T, c);
    bool ok = stress_update<ND>(s, e, s.sc->dt, s.edvoldt[e], Tm, s.volume[e], s.volume_old[e],
                                st, syy, es, edot, pl, dpl, dp);
    if (!ok) set_error(s.sc, DES_ERR_INPUT);
    store_tensor<ND>(s.stress, E, e, st);
    store_tensor<ND>(s.strain, E, e, es);
    store_tensor<ND>(s.strain_rate, E, e, edot);
    s.stressyy[e] = syy;
    s.plstrain[e] = pl;
    s.delta_plstrain[e] = dpl;
    s.dpressure[e] = dp;
}

// NMD_stress (geometry.cxx:182-236): node average of dp into ntmp ...
template <int ND> __global__ __launch_bounds__(BLOCK) void k_nmd_node(DevState s)
{
    BIND_NODE(s, n);
    constexpr int NPE = Dim<ND>::NPE;
    double sum = 0;
    for (int p = s.csr_off[n]; p < s.csr_off[n + 1]; ++p) {
        int e = s.csr_ent[p] / NPE;
        sum += s.dpressure[e] * s.volume[e];
    }
    s.ntmp[n] = sum / s.volume_n[n];
}

// ... and the element correction of the stress diagonal
template <int ND> __device__ __forceinline__ void nmd_correct(const DevState &s, int e, const int *c,
                                                            double *st)
{
    double dp = 0;
#pragma unroll
    for (int i = 0; i < Dim<ND>::NPE; ++i) dp += s.ntmp[c[i]];
    double dp_el = dp / Dim<ND>::NPE;
    double dp_orig = s.dpressure[e];
#pragma unroll
    for (int i = 0; i < ND; ++i) st[i] += (-dp_orig + dp_el) / ND;
}

template <int ND> __global__ __launch_bounds__(BLOCK) void k_nmd_elem(DevState s)
{
    BIND_ELEM(s, e);
    int c[Dim<ND>::NPE];
    load_conn<ND>(s, e, c);
    double st[ND];
#pragma unroll
    for (int i = 0; i < ND; ++i) st[i] = s.stress[i * (long)s.E + e];
    nmd_correct<ND>(s, e, c, st);
#pragma unroll
    for (int i = 0; i < ND; ++i) s.stress[i * (long)s.E + e] = st[i];
}

// element part of update_force (fields.cxx:304-330): what element e takes
// away from the force of its nodes
template <int ND> __device__ __forceinline__ void force_contrib(const DevState &s, int e, double vol,
                                                              const double sh[ND][Dim<ND>::NPE],
                                                              const double *st, double Tmean,
                                                              double out[Dim<ND>::NPE][ND])
{
    constexpr int NPE = Dim<ND>::NPE;
    const double gravity = s.P->gravity;
    double buoy = 0;
    if (gravity != 0) buoy = mat_rho(s, e, Tmean) * gravity / NPE;
#pragma unroll
    for (int i = 0; i < NPE; ++i) {
        if constexpr (ND == 3) {
            out[i][0] = (st[0] * sh[0][i] + st[3] * sh[1][i] + st[4] * sh[2][i]) * vol;
            out[i][1] = (st[3] * sh[0][i] + st[1] * sh[1][i] + st[5] * sh[2][i]) * vol;
            out[i][2] = (st[4] * sh[0][i] + st[5] * sh[1][i] + st[2] * sh[2][i] + buoy) * vol;
        } else {
            out[i][0] = (st[0] * sh[0][i] + st[2] * sh[1][i]) * vol;
            out[i][1] = (st[2] * sh[0][i] + st[1] * sh[1][i] + buoy) * vol;
        }
    }
}

// Node-major contributions: element e's contribution to its i-th node goes to
// that node's CSR slot slot[e*NPE+i], so the node gather reads one contiguous
// run per node.  With the mesh in curve order an element's slots are close,
// and the scattered stores cost the element pass less than the contiguous
// gather saves (measured against 96-byte element-major records).
// The CSR positions (force slots) of element e's NPE entries: the start of
// each node's list plus the entry's position in it, one byte each (srank),
// with the list starts from the staged block's LDS (loff) or from csr_off;
// the full 4-byte positions (slot) when some list is longer than 256.  pli:
// the element's stage_lidx and srank entries when the caller loaded them.
template <int ND>
__device__ __forceinline__ void force_slots(const DevState &s, int e, bool staged, const int *loff,
                                            int pos[Dim<ND>::NPE], const unsigned *pli = nullptr)
{
    constexpr int NPE = Dim<ND>::NPE;
    if (s.srank) {
        const unsigned r = pli ? pli[1] : s.srank[e];
        if (staged) {
            const unsigned li = pli ? pli[0] : s.stage_lidx[e];
            const lds_ptr<int> l = (lds_ptr<int>)loff;
#pragma unroll
            for (int i = 0; i < NPE; ++i) pos[i] = l[(li >> (8 * i)) & 255u] + (int)((r >> (8 * i)) & 255u);
        } else {
            int c[NPE];
            load_conn<ND>(s, e, c);
#pragma unroll
            for (int i = 0; i < NPE; ++i) pos[i] = s.csr_off[c[i]] + (int)((r >> (8 * i)) & 255u);
        }
    } else {
#pragma unroll
        for (int i = 0; i < NPE; ++i) pos[i] = s.slot[e * NPE + i];
    }
}

template <int ND> __device__ __forceinline__ void store_force_contrib(const DevState &s, const int pos[Dim<ND>::NPE],
                                                                    const double out[Dim<ND>::NPE][ND])
{
    constexpr int NPE = Dim<ND>::NPE;
#pragma unroll
    for (int i = 0; i < NPE; ++i) {
        double *w = s.fwork + (long)pos[i] * ND;
#pragma unroll
        for (int d = 0; d < ND; ++d) w[d] = out[i][d];
    }
}

// The mixed-precision schedule's contributions: the same slots holding FP32
// values (12 bytes in 3D), summed in FP64 by the node pass.
template <int ND> __device__ __forceinline__ void store_force_contrib_f32(const DevState &s,
                                                                        const int pos[Dim<ND>::NPE],
                                                                        const double out[Dim<ND>::NPE][ND])
{
    constexpr int NPE = Dim<ND>::NPE;
    float *fw = reinterpret_cast<float *>(s.fwork);
#pragma unroll
    for (int i = 0; i < NPE; ++i) {
        float *w = fw + (long)pos[i] * ND;
#pragma unroll
        for (int d = 0; d < ND; ++d) w[d] = (float)out[i][d];
    }
}

// Element-major contributions (the alternative layout of the fused FP64
// schedule): element e's record holds its contributions to its NPE nodes,
// node i's in 16-byte chunks CH*i .. CH*i+CH-1 (3D: fx fy | fz 0, 128 bytes
// per element; 2D: fx fy, 64 bytes with one unused chunk).  Pass E writes
// whole records with coalesced stores (a wave's records through LDS); pass F
// gathers node i's chunks of each element in the node's CSR order, so the
// sums run in the same order as with node-major slots.
template <int ND> struct FrecLayout {
    static constexpr int CH = ND == 3 ? 2 : 1;        // chunks per node
    static constexpr int R = ND == 3 ? 8 : 4;         // chunks per element record
};

// wbuf: this wave's LDS for the transposed store (64 x 4 chunks)
template <int ND> __device__ __forceinline__ void store_force_rec(const DevState &s, int e,
                                                                 const double out[Dim<ND>::NPE][ND],
                                                                 double2 *wbuf)
{
    constexpr int NPE = Dim<ND>::NPE, R = FrecLayout<ND>::R;
    double2 c[R];
#pragma unroll
    for (int i = 0; i < NPE; ++i) {
        if constexpr (ND == 3) {
            c[2 * i] = make_double2(out[i][0], out[i][1]);
            c[2 * i + 1] = make_double2(out[i][2], 0.0);
        } else {
            c[i] = make_double2(out[i][0], out[i][1]);
        }
    }
    if constexpr (ND == 2) c[3] = make_double2(0.0, 0.0);
    double2 *dst = reinterpret_cast<double2 *>(s.frec);
    const long r0 = e - (int)__lane_id();
    bool done;
    if constexpr (ND == 3) {
        const double2 lo[4] = {c[0], c[1], c[2], c[3]}, hi[4] = {c[4], c[5], c[6], c[7]};
        done = wave_store_records<4, R>(dst, r0, lo, wbuf);
        if (done) wave_store_records<4, R>(dst + 4, r0, hi, wbuf);
    } else {
        done = wave_store_records<4, R>(dst, r0, c, wbuf);
    }
    if (!done) {
#pragma unroll
        for (int k = 0; k < R; ++k) dst[(long)e * R + k] = c[k];
    }
}

template <int ND> __global__ __launch_bounds__(BLOCK) void k_force_elem(DevState s)
{
    BIND_ELEM(s, e);
    constexpr int NPE = Dim<ND>::NPE, NSTR = Dim<ND>::NSTR;
    int c[NPE];
    load_conn<ND>(s, e, c);
    double sh[ND][NPE], st[NSTR], out[NPE][ND];
    load_shp<ND>(s, e, sh);
    load_tensor<ND>(s.stress, s.E, e, st);
    force_contrib<ND>(s, e, s.volume[e], sh, st, elem_mean<ND>(s.T, c), out);
    int pos[NPE];
    force_slots<ND>(s, e, false, nullptr, pos);
    store_force_contrib<ND>(s, pos, out);
}

// PREM pressure tables (matprops.cxx:11-99)
__constant__ double c_prem_depth[46] = {
    0e3, 3e3, 15e3, 24.4e3, 40e3, 60e3, 80e3, 115e3, 150e3, 185e3, 220e3, 265e3, 310e3, 355e3,
    400e3, 450e3, 500e3, 550e3, 600e3, 635e3, 670e3, 721e3, 771e3, 871e3, 971e3, 1071e3,
    1171e3, 1271e3, 1371e3, 1471e3, 1571e3, 1671e3, 1771e3, 1871e3, 1971e3, 2071e3, 2171e3,
    2271e3, 2371e3, 2471e3, 2571e3, 2671e3, 2741e3, 2771e3, 2871e3, 2891e3};
__constant__ double c_prem_p[2][46] = {
    {0e8, 0.3e8, 3.3e8, 6.0e8, 11.2e8, 17.8e8, 24.5e8, 36.1e8, 47.8e8, 59.4e8, 71.1e8, 86.4e8,
     102.0e8, 117.7e8, 133.5e8, 152.2e8, 171.3e8, 190.7e8, 210.4e8, 224.3e8, 238.3e8, 260.7e8,
     282.9e8, 327.6e8, 372.8e8, 418.6e8, 464.8e8, 511.6e8, 558.9e8, 606.8e8, 655.2e8, 704.1e8,
     753.5e8, 803.6e8, 854.3e8, 905.6e8, 957.6e8, 1010.3e8, 1063.8e8, 1118.2e8, 1173.4e8,
     1229.7e8, 1269.7e8, 1287.0e8, 1345.6e8, 1357.5e8},
    {0e8, 0.82e8, 4.1e8, 6.7e8, 11.2e8, 17.8e8, 24.5e8, 36.1e8, 47.8e8, 59.4e8, 71.1e8, 86.4e8,
     102.0e8, 117.7e8, 133.5e8, 152.2e8, 171.3e8, 190.7e8, 210.4e8, 224.3e8, 238.3e8, 260.7e8,
     282.9e8, 327.6e8, 372.8e8, 418.6e8, 464.8e8, 511.6e8, 558.9e8, 606.8e8, 655.2e8, 704.1e8,
     753.5e8, 803.6e8, 854.3e8, 905.6e8, 957.6e8, 1010.3e8, 1063.8e8, 1118.2e8, 1173.4e8,
     1229.7e8, 1269.7e8, 1287.0e8, 1345.6e8, 1357.5e8}};

__device__ double ref_pressure(const des_params *P, double z)   // matprops.cxx:189-201
{
    double depth = -z;
    if (P->ref_pressure_option == 0) return P->rho0[0] * P->gravity * depth;
    if (P->ref_pressure_option == 1 || P->ref_pressure_option == 2) {
        const double *tab = c_prem_p[P->ref_pressure_option - 1];
        if (depth <= 0) return 0;
        int n;
        for (n = 1; n < 46; n++)
            if (depth <= c_prem_depth[n]) break;
        return tab[n - 1] + (tab[n] - tab[n - 1]) * (depth - c_prem_depth[n - 1])
            / (c_prem_depth[n] - c_prem_depth[n - 1]);
    }
    return 0;
}

// traction of one loaded boundary facet (apply_stress_bcs, bc.cxx:502-556):
// ftrac[k] = p * normal / NODES_PER_FACET
template <int ND> __device__ __forceinline__ void facet_traction_item(DevState s, const int *facets, int t)
{
    int k = facets[t];
    if (s.fmodel) bind_model<ND>(s, s.fmodel[t]);
    const des_params *P = s.P;
    int i = s.bf_bdry[k], e = s.bf_elem[k], f = s.bf_face[k];
    constexpr int NPE = Dim<ND>::NPE;
    int c[NPE];
    load_conn<ND>(s, e, c);
    static constexpr int NOF3[4][3] = {{1, 2, 3}, {0, 3, 2}, {0, 1, 3}, {0, 2, 1}};
    static constexpr int NOF2[3][2] = {{1, 2}, {2, 0}, {0, 1}};
    double normal[ND], zcenter;
    if constexpr (ND == 3) {
        int n0 = c[NOF3[f][0]], n1 = c[NOF3[f][1]], n2 = c[NOF3[f][2]];
        double v01[3], v02[3];
        for (int d = 0; d < 3; ++d) {
            v01[d] = s.coord[d * s.N + n1] - s.coord[d * s.N + n0];
            v02[d] = s.coord[d * s.N + n2] - s.coord[d * s.N + n0];
        }
        normal[0] = (v01[1] * v02[2] - v01[2] * v02[1]) / 2;
        normal[1] = (v01[2] * v02[0] - v01[0] * v02[2]) / 2;
        normal[2] = (v01[0] * v02[1] - v01[1] * v02[0]) / 2;
        zcenter = (s.coord[2 * s.N + n0] + s.coord[2 * s.N + n1] + s.coord[2 * s.N + n2]) / 3;
    } else {
        int n0 = c[NOF2[f][0]], n1 = c[NOF2[f][1]];
        double v01[2];
        for (int d = 0; d < 2; ++d) v01[d] = s.coord[d * s.N + n1] - s.coord[d * s.N + n0];
        normal[0] = v01[1];
        normal[1] = -v01[0];
        zcenter = (s.coord[s.N + n0] + s.coord[s.N + n1]) / 2;
    }
    double p;
    if (i == 4 && P->has_winkler_foundation) {
        p = s.sc->compensation_pressure -
            (mat_rho(s, e, elem_mean<ND>(s.T, c)) + P->winkler_delta_rho) * P->gravity * (zcenter + P->zlength);
    } else if (i == 5 && P->has_water_loading) {
        const double sea_level = 0;
        p = 0;
        if (zcenter < sea_level) {
            const double sea_water_density = 1030;
            p = sea_water_density * P->gravity * (sea_level - zcenter);
        }
    } else {
        p = ref_pressure(P, zcenter);
    }
#pragma unroll
    for (int d = 0; d < ND; ++d) s.ftrac[k * ND + d] = p * normal[d] / ND;
}

template <int ND> __global__ __launch_bounds__(BLOCK) void k_facet_traction(DevState s, const int *facets,
                                                                          int nf)
{
    const int t = blockIdx.x * BLOCK + threadIdx.x;
    if (t < nf) facet_traction_item<ND>(s, facets, t);
}

// node part of update_force: gather, stress BCs, elastic foundation,
// damping (fields.cxx:291-340, bc.cxx:493-566, fields.cxx:235-288)
template <int ND, bool F32 = false, int DEPTH = 4, bool EREC = false>
__device__ __forceinline__ void node_force(const DevState &s, int n, double f[ND], const double v[ND])
{
    constexpr int NPE = Dim<ND>::NPE;
    const des_params *P = s.P;
#pragma unroll
    for (int d = 0; d < ND; ++d) f[d] = 0;
    if constexpr (EREC) {
        // element-major records (FrecLayout), gathered in CSR (ascending host
        // element) order
        constexpr int CH = FrecLayout<ND>::CH, R = FrecLayout<ND>::R;
        const double2 *fr = reinterpret_cast<const double2 *>(s.frec);
        const int p1 = s.csr_off[n + 1];
        int p = s.csr_off[n];
        for (; p + DEPTH <= p1; p += DEPTH) {
            int en[DEPTH];
#pragma unroll
            for (int j = 0; j < DEPTH; ++j) en[j] = s.csr_ent[p + j];
            double c[DEPTH][ND];
#pragma unroll
            for (int j = 0; j < DEPTH; ++j) {
                const long b = (long)(en[j] / NPE) * R + CH * (en[j] % NPE);
                const double2 a = fr[b];
                c[j][0] = a.x;
                c[j][1] = a.y;
                if constexpr (ND == 3) c[j][2] = reinterpret_cast<const double *>(fr + b + 1)[0];
            }
#pragma unroll
            for (int j = 0; j < DEPTH; ++j)
#pragma unroll
                for (int d = 0; d < ND; ++d) f[d] -= c[j][d];
        }
        for (; p < p1; ++p) {
            const int ent = s.csr_ent[p];
            const long b = (long)(ent / NPE) * R + CH * (ent % NPE);
            const double2 a = fr[b];
            f[0] -= a.x;
            f[1] -= a.y;
            if constexpr (ND == 3) f[2] -= reinterpret_cast<const double *>(fr + b + 1)[0];
        }
    } else if constexpr (F32) {
        // FP32 contributions (mixed precision), accumulated in FP64 in slot order
        const float *fw = reinterpret_cast<const float *>(s.fwork);
        const int p1 = s.csr_off[n + 1];
        int p = s.csr_off[n];
        for (; p + DEPTH <= p1; p += DEPTH) {
            float c[DEPTH][ND];
#pragma unroll
            for (int j = 0; j < DEPTH; ++j)
#pragma unroll
                for (int d = 0; d < ND; ++d) c[j][d] = fw[(long)(p + j) * ND + d];
#pragma unroll
            for (int j = 0; j < DEPTH; ++j)
#pragma unroll
                for (int d = 0; d < ND; ++d) f[d] -= (double)c[j][d];
        }
        for (; p < p1; ++p) {
            const float *c = fw + (long)p * ND;
#pragma unroll
            for (int d = 0; d < ND; ++d) f[d] -= (double)c[d];
        }
    } else {
        const int p1 = s.csr_off[n + 1];
        int p = s.csr_off[n];
        for (; p + DEPTH <= p1; p += DEPTH) {   // DEPTH entries in flight, summed in order
            double c[DEPTH][ND];
#pragma unroll
            for (int j = 0; j < DEPTH; ++j)
#pragma unroll
                for (int d = 0; d < ND; ++d) c[j][d] = s.fwork[(long)(p + j) * ND + d];
#pragma unroll
            for (int j = 0; j < DEPTH; ++j)
#pragma unroll
                for (int d = 0; d < ND; ++d) f[d] -= c[j][d];
        }
        for (; p < p1; ++p) {
            const double *c = s.fwork + (long)p * ND;
#pragma unroll
            for (int d = 0; d < ND; ++d) f[d] -= c[d];
        }
    }
    if (s.ff_off) {
        for (int p = s.ff_off[n]; p < s.ff_off[n + 1]; ++p) {
            const double *t = s.ftrac + (long)s.ff_ent[p] * ND;
#pragma unroll
            for (int d = 0; d < ND; ++d) f[d] -= t[d];
        }
    }
    if (P->gravity != 0 && P->has_elastic_foundation && (s.bcflag[n] & BOUNDZ0)) {
        f[ND - 1] -= P->elastic_foundation_constant *
            (s.coord[(ND - 1) * s.N + n] - s.coord0[(ND - 1) * s.N + n]);
    }
    if (P->is_quasi_static) {
        const double small_vel = 1e-13;
        const double df = P->damping_factor;
        switch (P->damping_option) {
        case 1:
#pragma unroll
            for (int d = 0; d < ND; ++d)
                if (fabs(v[d]) > small_vel) f[d] -= df * copysign(f[d], v[d]);
            break;
        case 2:
#pragma unroll
            for (int d = 0; d < ND; ++d) f[d] -= df * f[d];
            break;
        case 3:
#pragma unroll
            for (int d = 0; d < ND; ++d) {
                if ((f[d] < 0) == (v[d] < 0)) f[d] -= df * f[d];
                else f[d] += (1 - df) * f[d];
            }
            break;
        default:
            break;
        }
    }
}

template <int ND> __global__ __launch_bounds__(BLOCK) void k_force_node(DevState s)
{
    BIND_NODE(s, n);
    double f[ND], v[ND];
#pragma unroll
    for (int d = 0; d < ND; ++d) v[d] = s.vel[d * s.N + n];
    node_force<ND>(s, n, f, v);
#pragma unroll
    for (int d = 0; d < ND; ++d) s.force[d * s.N + n] = f[d];
}

template <int ND> __global__ __launch_bounds__(BLOCK) void k_update_velocity(DevState s)
{
    BIND_NODE(s, n);
    const double dt = s.sc->dt, m = s.mass[n];
#pragma unroll
    for (int d = 0; d < ND; ++d) s.vel[d * s.N + n] += dt * s.force[d * s.N + n] / m;
}

// the slant boundaries n0..n3 of apply_vbcs (bc.cxx:350-436); every index
// into v is static, so the node's velocity stays in registers
template <int ND> __device__ __forceinline__ bool vbc_slant(const DevState &s, unsigned flag, double v[ND])
{
    const int *T = s.P->vbc_types;
    const double *V = s.P->vbc_values;
    bool ok = true;
    for (int ib = IBOUNDN0; ib <= IBOUNDN3; ib++) {
        if (!(flag & (1u << ib))) continue;
        const double *nv = s.bnormals + ib * ND;
        double fac = 0;
        switch (T[ib]) {
        case 1:
        case 11: {
            const int nh = (T[ib] == 1) ? ND : ND - 1;   // type 11: horizontal components only
            if (T[ib] == 11) fac = 1 / sqrt(1 - nv[ND - 1] * nv[ND - 1]);
            const double target = (T[ib] == 1) ? V[ib] : V[ib] * fac;
            if (flag == (1u << ib)) {
                double vn = 0;
#pragma unroll
                for (int d = 0; d < ND; d++)
                    if (d < nh) vn += v[d] * nv[d];
#pragma unroll
                for (int d = 0; d < ND; d++)
                    if (d < nh) v[d] += (target - vn) * nv[d];
            } else {
                for (int ic = 0; ic < ib; ic++) {
                    if (!(flag & (1u << ic))) continue;
                    if (T[ic] == 0) {
                        double vn = 0;
#pragma unroll
                        for (int d = 0; d < ND; d++)
                            if (d < nh) vn += v[d] * nv[d];
#pragma unroll
                        for (int d = 0; d < ND; d++)
                            if (d < nh) v[d] += (target - vn) * nv[d];
                    } else if (T[ic] == 1) {
                        if (!s.edge_valid[ic * DES_NBDRY + ib]) { ok = false; continue; }
                        const double *edge = s.edge_vectors + (ic * DES_NBDRY + ib) * ND;
                        double ve = 0;
#pragma unroll
                        for (int d = 0; d < ND; d++) ve += v[d] * edge[d];
#pragma unroll
                        for (int d = 0; d < ND; d++) v[d] = ve * edge[d];
                    }
                }
            }
            break;
        }
        case 3:
#pragma unroll
            for (int d = 0; d < ND; d++) v[d] = V[ib] * nv[d];
            break;
        case 13:
            fac = 1 / sqrt(1 - nv[ND - 1] * nv[ND - 1]);
#pragma unroll
            for (int d = 0; d < ND - 1; d++) v[d] = V[ib] * fac * nv[d];
            v[ND - 1] = 0;
            break;
        }
    }
    return ok;
}

// apply_vbcs for one node (bc.cxx:179-490); returns false if an edge vector
// the reference's map lacks is needed
template <int ND> __device__ __forceinline__ bool vbc_node(const DevState &s, unsigned flag, double v[ND])
{
    if (!(flag & BOUND_ANY)) return true;
    const int *T = s.P->vbc_types;
    const double *V = s.P->vbc_values;
    bool ok = true;
    if (flag & BOUNDX0) {
        switch (T[0]) {
        case 1: v[0] = V[0]; break;
        case 2: v[1] = 0; if constexpr (ND == 3) v[2] = 0; break;
        case 3: v[0] = V[0]; v[1] = 0; if constexpr (ND == 3) v[2] = 0; break;
        case 4: if constexpr (ND == 3) { v[1] = V[0]; v[2] = 0; } break;
        case 5: if constexpr (ND == 3) { v[0] = 0; v[1] = V[0]; v[2] = 0; } break;
        case 7: if constexpr (ND == 3) { v[0] = V[0]; v[1] = 0; } break;
        }
    }
    if (flag & BOUNDX1) {
        switch (T[1]) {
        case 1: v[0] = V[1]; break;
        case 2: v[1] = 0; if constexpr (ND == 3) v[2] = 0; break;
        case 3: v[0] = V[1]; v[1] = 0; if constexpr (ND == 3) v[2] = 0; break;
        case 4: if constexpr (ND == 3) { v[1] = V[1]; v[2] = 0; } break;
        case 5: if constexpr (ND == 3) { v[0] = 0; v[1] = V[1]; v[2] = 0; } break;
        case 7: if constexpr (ND == 3) { v[0] = V[1]; v[1] = 0; } break;
        }
    }
    if constexpr (ND == 3) {
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
            case 3: v[0] = V[3]; v[1] = 0; v[2] = 0; break;   // as bc.cxx:325-329
            case 4: v[0] = V[3]; v[2] = 0; break;
            case 5: v[0] = V[3]; v[1] = 0; v[2] = 0; break;
            case 7: v[0] = 0; v[1] = V[3]; break;
            }
        }
    }
    if (flag & 0x3C0u)       // BOUNDN0..BOUNDN3
        if (!vbc_slant<ND>(s, flag, v)) ok = false;
    if (T[4] == 0 && T[5] == 0) return ok;
    if (flag & BOUNDZ0) {
        switch (T[4]) {
        case 1: v[ND - 1] = V[4]; break;
        case 2: v[0] = 0; if constexpr (ND == 3) v[1] = 0; break;
        case 3: v[0] = 0; if constexpr (ND == 3) v[1] = 0; v[ND - 1] = V[4]; break;
        }
    }
    if (flag & BOUNDZ1) {
        switch (T[5]) {
        case 1: v[ND - 1] = V[5]; break;
        case 2: v[0] = 0; if constexpr (ND == 3) v[1] = 0; break;
        case 3: v[0] = 0; if constexpr (ND == 3) v[1] = 0; v[ND - 1] = V[5]; break;
        }
    }
    return ok;
}

template <int ND> __global__ __launch_bounds__(BLOCK) void k_apply_vbcs(DevState s)
{
    BIND_NODE(s, n);
    unsigned flag = s.bcflag[n];
    if (!(flag & BOUND_ANY)) return;
    double v[ND];
#pragma unroll
    for (int d = 0; d < ND; ++d) v[d] = s.vel[d * s.N + n];
    if (!vbc_node<ND>(s, flag, v)) set_error(s.sc, DES_ERR_ASSERT);
#pragma unroll
    for (int d = 0; d < ND; ++d) s.vel[d * s.N + n] = v[d];
}

template <int ND> __global__ __launch_bounds__(BLOCK) void k_update_coordinate(DevState s)
{
    BIND_NODE(s, n);
    const double dt = s.sc->dt;
#pragma unroll
    for (int d = 0; d < ND; ++d) s.coord[d * s.N + n] += s.vel[d * s.N + n] * dt;
}

// rotate_stress (fields.cxx:411-483) with stored shape functions
template <int ND> __global__ __launch_bounds__(BLOCK) void k_rotate(DevState s, int use_dt_prev)
{
    BIND_ELEM(s, e);
    // batched models close their steps in one launch: only the elastic ones
    // rotate (dynearthsol.cxx:346-347)
    if (s.eblk_model && !(s.P->rheol_type & DES_RH_ELASTIC)) return;
    constexpr int NPE = Dim<ND>::NPE, NSTR = Dim<ND>::NSTR;
    int c[NPE];
    load_conn<ND>(s, e, c);
    double v[NPE][ND], sh[ND][NPE], w[3], st[NSTR], es[NSTR];
    gather_vec<ND>(s.vel, s.N, c, v);
    load_shp<ND>(s, e, sh);
    spin<ND>(v, sh, w);
    const double dt = use_dt_prev ? s.sc->dt_prev : s.sc->dt;
    load_tensor<ND>(s.stress, s.E, e, st);
    load_tensor<ND>(s.strain, s.E, e, es);
    jaumann<ND>(st, dt, w);
    jaumann<ND>(es, dt, w);
    store_tensor<ND>(s.stress, s.E, e, st);
    store_tensor<ND>(s.strain, s.E, e, es);
}

// ---- compute_dt (geometry.cxx:239-310): block minima, then one block
template <int ND> __device__ __forceinline__ void dt_terms(const DevState &s, int e, double &dtm, double &dtd,
                                                         double &minl)
{
    constexpr int NPE = Dim<ND>::NPE;
    const des_params *P = s.P;
    int c[NPE];
    load_conn<ND>(s, e, c);
    double x[NPE][ND];
    gather_vec<ND>(s.coord, s.N, c, x);
    double vol = elem_volume<ND>(x);
    double minh;
    if constexpr (ND == 3) {
        double maxa = dmax(dmax(triangle_area<3>(x[0], x[1], x[2]), triangle_area<3>(x[0], x[1], x[3])),
                           dmax(triangle_area<3>(x[2], x[3], x[0]), triangle_area<3>(x[2], x[3], x[1])));
        minh = 3 * vol / maxa;
    } else {
        double maxl = sqrt(dmax(dmax(dist2<2>(x[0], x[1]), dist2<2>(x[1], x[2])), dist2<2>(x[0], x[2])));
        minh = 2 * vol / maxl;
    }
    dtm = dmin(dtm, 0.5 * P->visc_min / (1e-40 + harmonic_mean(s, P->shear_modulus, e)));
    if (P->has_thermal_diffusion) dtd = dmin(dtd, 0.5 * minh * minh / P->therm_diff_max);
    minl = dmin(minl, minh);
}

// minimum of v[0..2] over the block, left in sm[j][0]
__device__ __forceinline__ void block_min3(double (*sm)[BLOCK], const double v[3])
{
    for (int j = 0; j < 3; ++j) sm[j][threadIdx.x] = v[j];
    __syncthreads();
    for (int w = BLOCK / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w)
            for (int j = 0; j < 3; ++j) sm[j][threadIdx.x] = dmin(sm[j][threadIdx.x], sm[j][threadIdx.x + w]);
        __syncthreads();
    }
}

__device__ __forceinline__ void store_dt(const des_params *P, DevScalars *sc, double dt_maxwell,
                                         double dt_diffusion, double minl, double bulkm0, double rho0)
{
    double dt;
    if (P->fixed_dt != 0) {
        dt = P->fixed_dt;
    } else {
        double dt_advection = 0.5 * minl / P->max_vbc_val;
        double dt_elastic = P->is_quasi_static ?
            0.5 * minl / (P->max_vbc_val * P->inertial_scaling) :
            0.5 * minl / sqrt(bulkm0 / rho0);
        dt = dmin(dmin(dt_elastic, dt_maxwell), dmin(dt_advection, dt_diffusion)) * P->dt_fraction;
        if (!(dt > 0)) set_error(sc, DES_ERR_RUNTIME);
    }
    sc->dt = dt;
}

template <int ND> __global__ __launch_bounds__(BLOCK) void k_dt_partial(DevState s)
{
    double v[3] = {DBL_MAX, DBL_MAX, DBL_MAX};
    for (int e = blockIdx.x * BLOCK + threadIdx.x; e < s.E; e += gridDim.x * BLOCK) dt_terms<ND>(s, e, v[0], v[1], v[2]);
    __shared__ double sm[3][BLOCK];
    block_min3(sm, v);
    if (threadIdx.x < 3) s.red[blockIdx.x * 3 + threadIdx.x] = sm[threadIdx.x][0];
}

// bulkm(0) and rho(0) of the dynamic elastic dt (geometry.cxx:297): element 0
// is the reference's (host) element 0, wherever the device numbering put it
template <int ND> __device__ __forceinline__ void elem0_props(const DevState &s, int e0, double &bulkm0,
                                                            double &rho0)
{
    int c[Dim<ND>::NPE];
    load_conn<ND>(s, e0, c);
    bulkm0 = harmonic_mean(s, s.P->bulk_modulus, e0);
    rho0 = mat_rho(s, e0, elem_mean<ND>(s.T, c));
}

template <int ND> __device__ __forceinline__ void dt_final_body(const DevState &s, int nb, double (*sm)[BLOCK])
{
    double v[3] = {DBL_MAX, DBL_MAX, DBL_MAX};
    for (int b = threadIdx.x; b < nb; b += BLOCK)
        for (int j = 0; j < 3; ++j) v[j] = dmin(v[j], s.red[b * 3 + j]);
    block_min3(sm, v);
    if (threadIdx.x == 0) {
        double bulkm0 = 0, rho0 = 1;
        if (!s.P->is_quasi_static && s.P->fixed_dt == 0) elem0_props<ND>(s, s.e0dev[0], bulkm0, rho0);
        store_dt(s.P, s.sc, sm[0][0], sm[1][0], sm[2][0], bulkm0, rho0);
    }
}

template <int ND> __global__ __launch_bounds__(BLOCK) void k_dt_final(DevState s, int nb)
{
    __shared__ double sm[3][BLOCK];
    dt_final_body<ND>(s, nb, sm);
}

// compute_dt of batched models: element block b holds elements of one model
// only, so its minima go to slot b; one block per model then reduces that
// model's slots and evaluates bulkm and rho of the model's first element.
// the minima of element block blk in red[3 blk ..]: every lane of the
// workgroup takes part (the block reduction has barriers)
template <int ND> __device__ __forceinline__ void dt_partial_blk(DevState s, int blk, double (*sm)[BLOCK])
{
    const int e = blk * BLOCK + threadIdx.x;
    bool valid = e < s.E;
    if (s.eblk_model) {
        bind_model<ND>(s, s.eblk_model[blk]);
        valid = e < s.eblk_lim[blk];
    }
    double v[3] = {DBL_MAX, DBL_MAX, DBL_MAX};
    if (valid) dt_terms<ND>(s, e, v[0], v[1], v[2]);
    block_min3(sm, v);
    if (threadIdx.x < 3) s.red[blk * 3 + threadIdx.x] = sm[threadIdx.x][0];
}

template <int ND> __global__ __launch_bounds__(BLOCK) void k_dt_partial_b(DevState s)
{
    __shared__ double sm[3][BLOCK];
    dt_partial_blk<ND>(s, xcd_block(), sm);
}

template <int ND> __device__ __forceinline__ void dt_final_b_body(DevState s, int m, double (*sm)[BLOCK])
{
    bind_model<ND>(s, m);
    const int b0 = s.mblk_off[m], b1 = s.mblk_off[m + 1];
    double v[3] = {DBL_MAX, DBL_MAX, DBL_MAX};
    for (int b = b0 + threadIdx.x; b < b1; b += BLOCK)
        for (int j = 0; j < 3; ++j) v[j] = dmin(v[j], s.red[b * 3 + j]);
    block_min3(sm, v);
    if (threadIdx.x == 0) {
        double bulkm0 = 0, rho0 = 1;
        if (!s.P->is_quasi_static && s.P->fixed_dt == 0) elem0_props<ND>(s, s.e0dev[m], bulkm0, rho0);
        store_dt(s.P, s.sc, sm[0][0], sm[1][0], sm[2][0], bulkm0, rho0);
    }
}

template <int ND> __global__ __launch_bounds__(BLOCK) void k_dt_final_b(DevState s)
{
    __shared__ double sm[3][BLOCK];
    dt_final_b_body<ND>(s, blockIdx.x, sm);
}

// the bottom nodes' coord0 after a remesh (remeshing.cxx:2020-2026)
template <int ND> __global__ __launch_bounds__(BLOCK) void k_reset_bottom_coord0(DevState s)
{
    BIND_NODE(s, n);
    if (s.bcflag[n] & BOUNDZ0) s.coord0[(ND - 1) * s.N + n] = -s.P->zlength;
}

// ---- worst_elem_quality (geometry.cxx:458-512)
template <int ND> __device__ __forceinline__ double elem_quality(const DevState &s, int e)
{
    constexpr int NPE = Dim<ND>::NPE;
    int c[NPE];
    load_conn<ND>(s, e, c);
    double x[NPE][ND];
    gather_vec<ND>(s.coord, s.N, c, x);
    const double vol = s.volume[e];
    if constexpr (ND == 3) {
        double normalization_factor = 216 * sqrt(3.0);
        double area_sum = (triangle_area<3>(x[0], x[1], x[2]) + triangle_area<3>(x[0], x[1], x[3]) +
                           triangle_area<3>(x[2], x[3], x[0]) + triangle_area<3>(x[2], x[3], x[1]));
        return normalization_factor * vol * vol / (area_sum * area_sum * area_sum);
    } else {
        double normalization_factor = 4 * sqrt(3.0);
        double dist2_sum = dist2<2>(x[0], x[1]) + dist2<2>(x[1], x[2]) + dist2<2>(x[0], x[2]);
        return normalization_factor * vol / dist2_sum;
    }
}

// (quality, host element index) minimum over the block, smaller index on ties; slot b
__device__ __forceinline__ void block_argmin_store(DevState &s, int slot, double q, int qi)
{
    __shared__ double sq[BLOCK];
    __shared__ int si[BLOCK];
    sq[threadIdx.x] = q;
    si[threadIdx.x] = qi;
    __syncthreads();
    for (int w = BLOCK / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w) {
            double q2 = sq[threadIdx.x + w];
            int i2 = si[threadIdx.x + w];
            if (q2 < sq[threadIdx.x] || (q2 == sq[threadIdx.x] && i2 < si[threadIdx.x])) {
                sq[threadIdx.x] = q2;
                si[threadIdx.x] = i2;
            }
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        s.red[slot] = sq[0];
        s.redi[slot] = si[0];
    }
}

template <int ND> __global__ __launch_bounds__(BLOCK) void k_quality_partial(DevState s)
{
    double q = DBL_MAX;
    int qi = 0x7fffffff;
    for (int e = blockIdx.x * BLOCK + threadIdx.x; e < s.E; e += gridDim.x * BLOCK) {
        const double quality = elem_quality<ND>(s, e);
        const int h = s.ehost[e];
        if (quality < q || (quality == q && h < qi)) { q = quality; qi = h; }
    }
    block_argmin_store(s, blockIdx.x, q, qi);
}

// batched models: one slot per element block (blocks never mix models)
template <int ND> __global__ __launch_bounds__(BLOCK) void k_quality_partial_b(DevState s)
{
    const int blk = xcd_block();
    const int e = blk * BLOCK + threadIdx.x;
    double q = DBL_MAX;
    int qi = 0x7fffffff;
    if (e < s.eblk_lim[blk]) {
        q = elem_quality<ND>(s, e);
        qi = s.ehost[e];
    }
    block_argmin_store(s, blk, q, qi);
}

// per-element fields of an output frame (output.cxx:147-166): MatProps::rho,
// elem_quality and MatProps::visc of the last step's strain rate
template <int ND> __global__ __launch_bounds__(BLOCK) void k_elem_diagnostics(DevState s, double *density,
                                                                            double *quality, double *viscosity)
{
    BIND_ELEM(s, e);
    constexpr int NPE = Dim<ND>::NPE, NSTR = Dim<ND>::NSTR;
    int c[NPE];
    load_conn<ND>(s, e, c);
    const double Tm = elem_mean<ND>(s.T, c);
    density[e] = mat_rho(s, e, Tm);
    double sr[NSTR];
    load_tensor<ND>(s.strain_rate, s.E, e, sr);
    viscosity[e] = mat_visc<ND>(s, e, Tm, sr);
    quality[e] = elem_quality<ND>(s, e);
}

// ---- bad_mesh_quality's first two checks (remeshing.cxx:1893-1918): the
// first element smaller than smallest_vol and the first bottom node farther
// than dist from the initial bottom depth, as block minima of their host indices
template <int ND> __global__ __launch_bounds__(BLOCK) void k_badmesh_partial(DevState s, double smallest_vol,
                                                                           int check_bottom, double bottom,
                                                                           double dist)
{
    int first_e = 0x7fffffff, first_n = 0x7fffffff;
    for (int e = blockIdx.x * BLOCK + threadIdx.x; e < s.E; e += gridDim.x * BLOCK)
        if (s.volume[e] < smallest_vol) first_e = min(first_e, s.ehost[e]);
    if (check_bottom) {
        for (int n = blockIdx.x * BLOCK + threadIdx.x; n < s.N; n += gridDim.x * BLOCK)
            if ((s.bcflag[n] & BOUNDZ0) && fabs(s.coord[(ND - 1) * s.N + n] - bottom) > dist)
                first_n = min(first_n, s.nhost[n]);
    }
    __shared__ int se[BLOCK], sn[BLOCK];
    se[threadIdx.x] = first_e;
    sn[threadIdx.x] = first_n;
    __syncthreads();
    for (int w = BLOCK / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w) {
            se[threadIdx.x] = min(se[threadIdx.x], se[threadIdx.x + w]);
            sn[threadIdx.x] = min(sn[threadIdx.x], sn[threadIdx.x + w]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        s.redi[2 * blockIdx.x] = se[0];
        s.redi[2 * blockIdx.x + 1] = sn[0];
    }
}

// ---- Output::average_fields (output.cxx:211-251) after a step.  snapshot
// (the interval's first step, var.steps % average_interval == 1): keep
// strain, stress and delta_plstrain (coord and the time are kept by the
// caller and lane 0); otherwise add stress and delta_plstrain to the sums.
// A rotation the fused driver still owes (rot_pending) is applied to the
// copies here exactly as the next pass C applies it, so the fields summed are
// the reference's after rotate_stress.  One model per context.
template <int ND> __global__ __launch_bounds__(BLOCK) void k_average_fields(DevState s, int snapshot)
{
    const int e = xcd_block() * BLOCK + threadIdx.x;
    if (e >= s.E) return;
    constexpr int NPE = Dim<ND>::NPE, NSTR = Dim<ND>::NSTR;
    const long E = s.E;
    if (snapshot && e == 0) s.sc->time0 = s.sc->time;
    double st[NSTR], es[NSTR];
    load_tensor<ND>(s.stress, E, e, st);
    if (snapshot) load_tensor<ND>(s.strain, E, e, es);
    if (s.sc->rot_pending) {
        int c[NPE];
        load_conn<ND>(s, e, c);
        double x[NPE][ND], v[NPE][ND], Tn[NPE], aux[NPE], sh[ND][NPE], edot[NSTR], w[3];
        load_nrec<ND>(s.nrec, c, x, v, Tn, aux);
        shape_fn<ND>(x, elem_volume<ND>(x), sh);
        strain_rate_spin<ND>(v, sh, edot, w, true);
        const double dtp = s.sc->dt_prev;
        jaumann<ND>(st, dtp, w);
        if (snapshot) jaumann<ND>(es, dtp, w);
    }
    if (snapshot) {
        store_tensor<ND>(s.avg_stress, E, e, st);
        store_tensor<ND>(s.avg_strain0, E, e, es);
        s.avg_dpl[e] = s.delta_plstrain[e];
    } else {
#pragma unroll
        for (int k = 0; k < NSTR; ++k) s.avg_stress[k * E + e] += st[k];
        s.avg_dpl[e] += s.delta_plstrain[e];
    }
}

// ============================================================ fused step
//
// Pass A (elements): volume, shape functions, thermal contributions,
//   volume-weighted dilatation, element mass and thermal mass.
// Pass B (nodes):    temperature update, volume_n, mass, dvoldt.
// Pass C (elements): deferred Jaumann rotation, strain rate, anti-locking,
//   rheology.
// Pass D (nodes):    volume-weighted nodal pressure change.
// Pass E (elements): nodal-mixed pressure correction, force contributions.
// Pass F (nodes):    force gather and BCs, damping, velocity, velocity BCs,
//   coordinates.

struct FusedFlags {
    int use_stored_mass = 0;   // first step after a sync: mass/tmass/volume_n from compute_mass
    int store_aux = 0;         // last step before a sync: keep strain_rate, edvoldt, force
    int step_end = 0;          // pass F also ends the step: time, steps, dt_prev, rotation flag
    // the differentiated run's step records, written by the pass that has
    // the value at hand (null in plain runs): pass C (FP64) the stress,
    // plastic strain and volume_old it starts from, pass B the node records
    // after it (x and v of the step's start, T, dvoldt), mass and volume_n,
    // pass D dp_nd, pass F the force
    double *rec_stress = nullptr, *rec_plstrain = nullptr, *rec_volume_old = nullptr;
    double *rec_nrec = nullptr, *rec_mass = nullptr, *rec_volume_n = nullptr;
    double *rec_dpn = nullptr, *rec_force = nullptr;
};

// the scalar bookkeeping at the end of a step (var.time, var.steps), and the
// Jaumann rotation owed by the next pass C when the rheology is elastic
__device__ __forceinline__ void step_end_scalars(DevScalars *sc, int rheol_type)
{
    sc->time += sc->dt;
    sc->steps += 1;
    sc->dt_prev = sc->dt;
    sc->rot_pending = (rheol_type & DES_RH_ELASTIC) ? 1 : 0;
}

// Packed per-element record written by pass A and gathered by pass B: the
// thermal contributions to the element's nodes, tr(edot) V, the element's
// mass and thermal mass and its volume, one 64-byte line per element.
constexpr int REC = 8;
constexpr int REC_DJ = 4, REC_M = 5, REC_TM = 6, REC_V = 7;

// xbuf: this block's LDS for the records' transposed store (BLOCK * REC / 2 chunks)
template <int ND> __device__ __forceinline__ void pass_a_body(DevState s, const FusedFlags &fl, int blk,
                                                            bool staged, const double2 *lds, double2 *xbuf)
{
    BIND_ELEM_AT(s, e, blk);
    constexpr int NPE = Dim<ND>::NPE, NSTR = Dim<ND>::NSTR;
    const des_params *P = s.P;
    double x[NPE][ND], sh[ND][NPE], v[NPE][ND], Tn[NPE], aux[NPE];
    load_elem_nodes<ND>(s, e, staged, lds, x, v, Tn, aux);
    double vol = elem_volume<ND>(x);
    shape_fn<ND>(x, vol, sh);
    s.volume[e] = vol;
    double rec[REC];
#pragma unroll
    for (int k = 0; k < REC; ++k) rec[k] = 0;
    if (P->has_thermal_diffusion) temperature_contrib<ND>(s, e, vol, sh, Tn, rec);
    double sr[NSTR];
    strain_rate<ND>(v, sh, sr);
    rec[REC_DJ] = trace<ND>(sr) * vol;
    if (!fl.use_stored_mass) {
        double Tm = 0;
#pragma unroll
        for (int i = 0; i < NPE; ++i) Tm += Tn[i];
        Tm /= NPE;
        mass_of<ND>(s, e, vol, Tm, rec[REC_M], rec[REC_TM]);
        rec[REC_V] = vol;
    }
    double2 r2[REC / 2];
#pragma unroll
    for (int k = 0; k < REC / 2; ++k) r2[k] = make_double2(rec[2 * k], rec[2 * k + 1]);
    double2 *erec2 = reinterpret_cast<double2 *>(s.erec);
    if (!wave_store_records<REC / 2, REC / 2>(erec2, e - (int)__lane_id(), r2, xbuf + (threadIdx.x & ~63) * (REC / 2))) {
#pragma unroll
        for (int k = 0; k < REC / 2; ++k) erec2[(long)e * (REC / 2) + k] = r2[k];
    }
}

template <int ND, bool STAGE, int WPE = 1> __global__ __launch_bounds__(BLOCK, WPE) void k_pass_a(DevState s, FusedFlags fl)
{
    __shared__ double2 lds[STAGE ? STAGE_MAX * (NREC / 2) : 1];
    __shared__ double2 xbuf[BLOCK * (REC / 2)];
    const int blk = xcd_block(s.rev);
    const bool staged = STAGE && stage_block(s, blk, lds);
    pass_a_body<ND>(s, fl, blk, staged, lds, xbuf);
}

// value of lane k (0..3) of this lane's quad
template <int K> __device__ __forceinline__ int quad_bcast(int v)
{
    return __builtin_amdgcn_mov_dpp(v, K * 0x55, 0xf, 0xf, false);
}
template <int K> __device__ __forceinline__ double quad_bcast(double v)
{
    const long long b = __double_as_longlong(v);
    const int lo = quad_bcast<K>((int)(b & 0xffffffff)), hi = quad_bcast<K>((int)(b >> 32));
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

template <int ND, int DEPTH> __device__ __forceinline__ void pass_b_body(DevState s, const FusedFlags &fl, int blk)
{
    // four lanes (a quad) per node: for every entry of the node's list each
    // lane loads the 16-byte quarter of the element's 64-byte record that
    // holds its own sums' terms and adds them in the list's (ascending
    // element) order; lane 0 collects the sums and finishes the node
    const int blk_ = blk;
    const int t = blk_ * BLOCK + threadIdx.x, n = t >> 2, q = t & 3;
    if (s.nblk_model) {
        const int nb = blk_ >> 2;
        if (n >= s.nblk_lim[nb]) return;
        bind_model<ND>(s, s.nblk_model[nb]);
    } else if (n >= s.N) {
        return;
    }
    constexpr int NPE = Dim<ND>::NPE;
    const int thermal = s.P->has_thermal_diffusion;
    double T_old = 0;
    unsigned bcf = 0;
    if (thermal && q == 0) {      // the node's own values, requested before the gather
        T_old = __builtin_nontemporal_load(s.T + n);
        bcf = s.bcflag[n];
    }
    // Record quarters: (c0, c1), (c2, c3), (dj, m), (tm, V).  Lane 0 sums the
    // node's thermal term c_i (quarter i/2, .x or .y by i's parity) in accx;
    // lane 1 sums dj and m, lane 2 tm and V (accx, accy); lane 3 repeats
    // lane 2's loads and its sums are not used.
    double accx = 0, accy = 0;
    const int qrt = q == 1 ? 2 : 3;      // the quarter of lanes 1..3
    const int p0 = s.csr_off[n], p1 = s.csr_off[n + 1];
    // DEPTH entries per iteration: their records are all requested before
    // the first is summed, and the next iteration's entries are requested
    // before these records are summed
    int mine[DEPTH / 4];
#pragma unroll
    for (int g = 0; g < DEPTH / 4; ++g) {
        const int pp = p0 + 4 * g + q;
        mine[g] = pp < p1 ? s.csr_ent[pp] : -1;
    }
    for (int p = p0; p < p1; p += DEPTH) {
        int en[DEPTH];
#pragma unroll
        for (int g = 0; g < DEPTH / 4; ++g) {
            en[4 * g + 0] = quad_bcast<0>(mine[g]);
            en[4 * g + 1] = quad_bcast<1>(mine[g]);
            en[4 * g + 2] = quad_bcast<2>(mine[g]);
            en[4 * g + 3] = quad_bcast<3>(mine[g]);
        }
        double2 rq[DEPTH];
#pragma unroll
        for (int j = 0; j < DEPTH; ++j)
            if (en[j] >= 0)
                rq[j] = reinterpret_cast<const double2 *>(s.erec + (long)(en[j] / NPE) * REC)
                    [q == 0 ? (en[j] % NPE) >> 1 : qrt];
#pragma unroll
        for (int g = 0; g < DEPTH / 4; ++g) {
            const int pp = p + DEPTH + 4 * g + q;
            mine[g] = pp < p1 ? s.csr_ent[pp] : -1;
        }
#pragma unroll
        for (int j = 0; j < DEPTH; ++j) {
            if (en[j] < 0) break;
            accx += (q == 0 && (en[j] % NPE) & 1) ? rq[j].y : rq[j].x;
            accy += rq[j].y;
        }
    }
    const double tdot = accx, dj = quad_bcast<1>(accx);
    double m = quad_bcast<1>(accy), tm = quad_bcast<2>(accx), vn = quad_bcast<2>(accy);
    if (q != 0) return;
    if (!thermal) tm = 0;      // the thermal mass is summed for thermal models only
    if (fl.use_stored_mass) {
        vn = s.volume_n[n];
        m = s.mass[n];
        tm = s.tmass[n];
    } else {
        s.volume_n[n] = vn;
        s.mass[n] = m;
        __builtin_nontemporal_store(tm, s.tmass + n);      // read again only by a run's first step
    }
    double *rn = s.nrec + (long)NREC * n;
    double Tnew = 0;
    if (thermal) {      // the SoA temperature: read only by this pass (the others read the records)
        Tnew = temperature_new_f(s, bcf, T_old, tdot, s.sc->dt, tm);
        __builtin_nontemporal_store(Tnew, s.T + n);
    }
    // dvoldt goes to the records only: ntmp is left to pass D (NMD_stress's
    // dp_nd, its last writer in the reference's step).  T and aux share the
    // record's last 16-byte chunk: one store
    const double dvn = dj / vn;
    static_assert(NrecSlot<ND>::T % 2 == 0 && NrecSlot<ND>::AUX == NrecSlot<ND>::T + 1, "T, aux: one chunk");
    if (thermal)
        reinterpret_cast<double2 *>(rn)[NrecSlot<ND>::T / 2] = make_double2(Tnew, dvn);
    else
        rn[NrecSlot<ND>::AUX] = dvn;
    if (fl.rec_nrec) {
        double *r = fl.rec_nrec + (long)NREC * n;
#pragma unroll
        for (int d = 0; d < ND; ++d) {
            r[NrecSlot<ND>::X + d] = s.coord[d * s.N + n];
            r[NrecSlot<ND>::V + d] = s.vel[d * s.N + n];
        }
        r[NrecSlot<ND>::T] = thermal ? Tnew : rn[NrecSlot<ND>::T];
        r[NrecSlot<ND>::AUX] = dvn;
        fl.rec_mass[n] = m;
        fl.rec_volume_n[n] = vn;
    }
}

template <int ND, int DEPTH> __global__ __launch_bounds__(BLOCK) void k_pass_b(DevState s, FusedFlags fl)
{
    if (blockIdx.x == 0 && threadIdx.x == 0 && s.ret_count) *s.ret_count = 0;   // pass C appends
    pass_b_body<ND, DEPTH>(s, fl, xcd_block(s.rev));
}

template <int ND, bool DEFER, bool PLASTIC_ONLY = false>
__device__ __forceinline__ void pass_c_body(const DevState &s, const FusedFlags &fl,
                                                                          int e, bool staged, const double2 *lds,
                                            const unsigned *pli = nullptr, const double *pes = nullptr,
                                            const double *pst = nullptr)
{
    // pli, pes, pst: the element's stage_lidx entry, strain and stress when
    // the caller loaded them
    constexpr int NPE = Dim<ND>::NPE, NSTR = Dim<ND>::NSTR;
    const long E = s.E;
    double sh[ND][NPE], v[NPE][ND], Tn[NPE], aux[NPE], vol;
    {
        // the volume pass A stored, recomputed from the same coordinates
        // (the same operations on the same values: the same bits) instead of
        // read back
        double x[NPE][ND];
        load_elem_nodes<ND>(s, e, staged, lds, x, v, Tn, aux, staged ? pli : nullptr);
        vol = elem_volume<ND>(x);
        shape_fn<ND>(x, vol, sh);
    }
    double w[3];
    const bool rot = s.sc->rot_pending;
    const double dtp = s.sc->dt_prev;
    double edot[NSTR];
    strain_rate_spin<ND>(v, sh, edot, w, rot);
    double edv = 0;
#pragma unroll
    for (int i = 0; i < NPE; ++i) edv += aux[i];
    edv = edv / NPE;
    const double dt = s.sc->dt;
    // anti-locking and the strain update of update_stress (rheology.cxx:528-540)
    // first, so the strain is stored before the rheology needs the registers
    double total_dv;
    {
        double div = trace<ND>(edot);
#pragma unroll
        for (int i = 0; i < ND; ++i) edot[i] += (edv - div) / ND;
        double es[NSTR];
        if (pes) {
#pragma unroll
            for (int i = 0; i < NSTR; ++i) es[i] = pes[i];
        } else {
            load_tensor_nt<ND>(s.strain, E, e, es);
        }
        {
            // rotated unconditionally and selected: a branch here keeps ~30
            // more registers live through the rheology below
            double er[NSTR];
#pragma unroll
            for (int i = 0; i < NSTR; ++i) er[i] = es[i];
            jaumann<ND>(er, dtp, w);
#pragma unroll
            for (int i = 0; i < NSTR; ++i) es[i] = rot ? er[i] : es[i];
        }
#pragma unroll
        for (int i = 0; i < NSTR; ++i) es[i] += edot[i] * dt;
        store_tensor_nt<ND>(s.strain, E, e, es);
        total_dv = trace<ND>(es);
    }
    if (fl.store_aux) {
        store_tensor<ND>(s.strain_rate, E, e, edot);
        s.edvoldt[e] = edv;
    }
    double st[NSTR];
    if (pst) {
#pragma unroll
        for (int i = 0; i < NSTR; ++i) st[i] = pst[i];
    } else {
        load_tensor<ND>(s.stress, E, e, st);
    }
    if (fl.rec_stress) store_tensor<ND>(fl.rec_stress, E, e, st);
    if (rot) jaumann<ND>(st, dtp, w);
    // stressyy exists only under plane strain (2D); nothing else reads or writes it
    double syy = (ND == 2 && s.P->is_plane_strain) ? s.stressyy[e] : 0.0, dp;
    double pl = __builtin_nontemporal_load(s.plstrain + e), dpl = 0;
    if (fl.rec_plstrain) fl.rec_plstrain[e] = pl;
    if (fl.rec_volume_old) fl.rec_volume_old[e] = s.volume_old[e];
    bool ok = true, defer = false;
    // which of plastic strain and its increment this pass changes (and stores)
    bool store_pl = true, store_dpl = true;
    double trial[6], old_s = 0;
    const double Tmean = node_mean<ND>(Tn);
    if (DEFER && ND == 3 && (s.P->rheol_type == DES_RH_EP || s.P->rheol_type == DES_RH_EVP)) {
        // the plastic strain stays; its increment is zeroed or left as it is,
        // so neither is read back to be stored unchanged
        defer = plastic_deferred3(s, e, dt, Tmean, vol, s.volume_old[e], st, edot, pl, store_dpl, dp, trial,
                                  old_s);
        store_pl = false;
    } else if constexpr (!PLASTIC_ONLY) {
        // 3D plastic rheologies are deferred above: their inline return is not compiled in here
        dpl = __builtin_nontemporal_load(s.delta_plstrain + e);
        ok = rheology_after_strain<ND, !(DEFER && ND == 3)>(s, e, dt, total_dv, Tmean, vol, s.volume_old[e], st,
                                                           syy, edot, pl, dpl, dp);
    } else {
        ok = false;        // launched for elasto-(visco-)plastic 3D models only
        dp = 0;
        store_pl = store_dpl = false;
    }
    if (!ok) set_error(s.sc, DES_ERR_INPUT);
    if constexpr (DEFER) {
        // append the deferred elements, one atomic per wavefront
        const unsigned long long mask = __ballot(defer);
        if (mask) {
            const int lane = __lane_id();
            const int leader = __ffsll((long long)mask) - 1;
            int base = 0;
            if (lane == leader) base = atomicAdd(s.ret_count, __popcll(mask));
            base = __shfl(base, leader);
            if (defer) {
                const int k = base + __popcll(mask & ((1ull << lane) - 1));
                s.ret_elem[k] = e;
                double *rv = s.ret_val + 8 * (long)k;
#pragma unroll
                for (int i = 0; i < 6; ++i) rv[i] = trial[i];
                rv[6] = old_s;
            }
        }
    }
    store_tensor<ND>(s.stress, E, e, st);
    if (s.P->is_plane_strain) s.stressyy[e] = syy;
    if (store_pl) __builtin_nontemporal_store(pl, s.plstrain + e);
    if (store_dpl) __builtin_nontemporal_store(dpl, s.delta_plstrain + e);
    s.dpressure[e] = dp;
    s.escal[e] = dp * vol;          // NMD_stress's dp * volume (geometry.cxx:206)
}

template <int ND, bool PLASTIC_ONLY = false>
__device__ __forceinline__ void pass_c_blk(DevState s, const FusedFlags &fl, int blk, bool staged, const double2 *lds,
                                           const unsigned *pli = nullptr, const double *pes = nullptr,
                                           const double *pst = nullptr)
{
    BIND_ELEM_AT(s, e, blk);
    pass_c_body<ND, true, PLASTIC_ONLY>(s, fl, e, staged, lds, pli, pes, pst);
}

// PLASTIC_ONLY: the instantiation for 3D contexts whose models are all
// elasto-(visco-)plastic, without the other rheologies' code (whose registers
// would otherwise set the kernel's allocation: 119 VGPRs direct and 107 staged
// instead of 131 and 133, four waves per SIMD instead of three)
template <int ND, int WPE, bool STAGE, bool PLASTIC_ONLY = false>
__global__ __launch_bounds__(BLOCK, WPE) void k_pass_c(DevState s, FusedFlags fl)
{
    __shared__ double2 lds[STAGE ? STAGE_MAX * (NREC / 2) : 1];
    const int blk = xcd_block(s.rev);
    // the staged variants request the element's staged-node positions,
    // strain and stress before the staging barrier (as in k_pass_e)
    constexpr int NSTR = Dim<ND>::NSTR;
    unsigned pli = 0;
    double pes[NSTR], pst[NSTR];
    if constexpr (STAGE) {
        const int e0 = blk * BLOCK + threadIdx.x;
        if (s.eblk_model ? e0 < s.eblk_lim[blk] : e0 < s.E) {
            if (s.stage) pli = s.stage_lidx[e0];
            load_tensor_nt<ND>(s.strain, s.E, e0, pes);
            load_tensor<ND>(s.stress, s.E, e0, pst);
        }
    }
    const bool staged = STAGE && stage_block(s, blk, lds);
    pass_c_blk<ND, PLASTIC_ONLY>(s, fl, blk, staged, lds, staged ? &pli : nullptr, STAGE ? pes : nullptr,
                                 STAGE ? pst : nullptr);
}

// Pass C's Mohr-Coulomb returns (rheology.cxx:151-299, 594-635) of the 3D
// elements that pass C deferred, from their trial stresses; the Maxwell
// stress that pass C stored stays when its J2 is the smaller one.
__device__ __forceinline__ void pass_c_ret_item(const DevState &s, int k)
{
    {
        const int e = s.ret_elem[k];
        DevState st = s;
        if (s.eblk_model) bind_model<3>(st, s.eblk_model[e / BLOCK]);
        const des_params *P = st.P;
        const double *rv = s.ret_val + 8 * (long)k;
        double sp[6];
#pragma unroll
        for (int i = 0; i < 6; ++i) sp[i] = rv[i];
        const double old_s = rv[6];
        const double bulkm = harmonic_mean(st, P->bulk_modulus, e);
        const double shearm = harmonic_mean(st, P->shear_modulus, e);
        const double pl = s.plstrain[e];
        double amc, anphi, anpsi, hardn, ten_max, depls;
        mat_plastic_props(st, e, pl, amc, anphi, anpsi, hardn, ten_max);
        mc_return<3>(bulkm, shearm, amc, anphi, anpsi, hardn, ten_max, depls, sp);
        bool take = true;
        if (P->rheol_type == DES_RH_EVP) {
            double sv[6];
            load_tensor<3>(s.stress, s.E, e, sv);
            take = !(second_invariant2<3>(sv) < second_invariant2<3>(sp));
        }
        if (take) {
            store_tensor<3>(s.stress, s.E, e, sp);
            s.plstrain[e] = pl + depls;
            s.delta_plstrain[e] = depls;
            const double dp = trace<3>(sp) - old_s;
            s.dpressure[e] = dp;
            s.escal[e] = dp * s.volume[e];
        }
    }
}

__global__ __launch_bounds__(BLOCK) void k_pass_c_ret(DevState s)
{
    const int cnt = *s.ret_count;
    for (int k = blockIdx.x * BLOCK + threadIdx.x; k < cnt; k += gridDim.x * BLOCK) pass_c_ret_item(s, k);
}

// pass D: node sum of dp * volume (NMD_stress, geometry.cxx:203-216)
template <int ND, int NB, int DEPTH> __device__ __forceinline__ void pass_d_body(DevState s, const FusedFlags &fl, int blk)
{
    BIND_NODE_NB_AT(s, n, NB, blk);
    constexpr int NPE = Dim<ND>::NPE;
    const double vn = s.volume_n[n];       // requested before the gather
    double sum = 0;
    const int p1 = s.csr_off[n + 1];
    int p = s.csr_off[n];
    for (; p + DEPTH <= p1; p += DEPTH) {       // DEPTH entries in flight, summed in order
        int en[DEPTH];
        double v[DEPTH];
#pragma unroll
        for (int j = 0; j < DEPTH; ++j) en[j] = s.csr_ent[p + j];
#pragma unroll
        for (int j = 0; j < DEPTH; ++j) v[j] = s.escal[en[j] / NPE];
#pragma unroll
        for (int j = 0; j < DEPTH; ++j) sum += v[j];
    }
    for (; p < p1; ++p) sum += s.escal[s.csr_ent[p] / NPE];
    const double dpn = sum / vn;
    __builtin_nontemporal_store(dpn, s.ntmp + n);          // read only at a sync
    s.nrec[(long)NREC * n + NrecSlot<ND>::AUX] = dpn;
    if (fl.rec_dpn) fl.rec_dpn[n] = dpn;
}

template <int ND, int NB, int DEPTH = 4> __global__ __launch_bounds__(NB) void k_pass_d(DevState s, FusedFlags fl)
{
    pass_d_body<ND, NB, DEPTH>(s, fl, xcd_block(s.rev));
}

// wbuf: with EREC, this wave's LDS for the records' transposed store
// loff: with staged, the staged nodes' list starts (stage_block)
template <int ND, bool EREC = false>
__device__ __forceinline__ void pass_e_body(DevState s, const FusedFlags &fl, int blk, bool staged, const double2 *lds,
                                            const int *loff, double2 *wbuf = nullptr, const double *pre = nullptr,
                                            const unsigned *pli = nullptr)
{
    // pre: the element's stress and dpressure, pli its stage_lidx and srank
    // entries, when the caller loaded them
    BIND_ELEM_AT(s, e, blk);
    constexpr int NPE = Dim<ND>::NPE, NSTR = Dim<ND>::NSTR;
    const long E = s.E;
    double x[NPE][ND], v[NPE][ND], Tn[NPE], dpn[NPE];
    load_elem_nodes<ND>(s, e, staged, lds, x, v, Tn, dpn, staged ? pli : nullptr);
    double st[NSTR];
    if (pre) {
#pragma unroll
        for (int i = 0; i < NSTR; ++i) st[i] = pre[i];
    } else {
        load_tensor<ND>(s.stress, E, e, st);
    }
    {   // NMD_stress's element side (geometry.cxx:221-235), dp_nd from the records
        const double dp_el = node_mean<ND>(dpn);
        const double dp_orig = pre ? pre[NSTR] : s.dpressure[e];
#pragma unroll
        for (int i = 0; i < ND; ++i) st[i] += (-dp_orig + dp_el) / ND;
    }
#pragma unroll
    for (int i = 0; i < ND; ++i) s.stress[i * E + e] = st[i];
    double sh[ND][NPE], out[NPE][ND];
    const double vol = elem_volume<ND>(x);       // pass A's volume, recomputed (same bits)
    shape_fn<ND>(x, vol, sh);
    force_contrib<ND>(s, e, vol, sh, st, node_mean<ND>(Tn), out);
    if constexpr (EREC) {
        store_force_rec<ND>(s, e, out, wbuf);
    } else {
        int pos[NPE];
        force_slots<ND>(s, e, staged, loff, pos, staged ? pli : nullptr);
        store_force_contrib<ND>(s, pos, out);
    }
}

template <int ND, bool STAGE, int WPE = 1> __global__ __launch_bounds__(BLOCK, WPE) void k_pass_e(DevState s, FusedFlags fl)
{
    __shared__ double2 lds[STAGE ? STAGE_MAX * (NREC / 2) : 1];
    __shared__ int loff[STAGE ? STAGE_MAX : 1];
    const int blk = xcd_block(s.rev);
    // the element's stress and pressure change, its staged-node positions
    // and its slot positions do not depend on the staged records: requested
    // before the block's staging barrier, so that their HBM latency overlaps
    // the staging's (the compiler keeps loads written after a barrier after it)
    constexpr int NSTR = Dim<ND>::NSTR;
    double pre[NSTR + 1];
    unsigned pli[2] = {0u, 0u};
    if constexpr (STAGE) {
        const int e0 = blk * BLOCK + threadIdx.x;
        if (s.eblk_model ? e0 < s.eblk_lim[blk] : e0 < s.E) {
            load_tensor<ND>(s.stress, s.E, e0, pre);
            pre[NSTR] = s.dpressure[e0];
            if (s.stage) pli[0] = s.stage_lidx[e0];
            if (s.srank) pli[1] = s.srank[e0];
        }
    }
    const bool staged = STAGE && stage_block(s, blk, lds, loff);
    pass_e_body<ND>(s, fl, blk, staged, lds, loff, nullptr, STAGE ? pre : nullptr, STAGE ? pli : nullptr);
}

// pass E writing element-major force records
template <int ND, bool STAGE> __global__ __launch_bounds__(BLOCK) void k_pass_e_er(DevState s, FusedFlags fl)
{
    __shared__ double2 lds[STAGE ? STAGE_MAX * (NREC / 2) : 1];
    __shared__ double2 wbuf[BLOCK * 4];
    const int blk = xcd_block(s.rev);
    const bool staged = STAGE && stage_block(s, blk, lds);
    pass_e_body<ND, true>(s, fl, blk, staged, lds, nullptr, wbuf + (threadIdx.x & ~63) * 4);
}

// wbuf: this wave's LDS for the node records' transposed store (null: the
// lanes store their own records)
template <int ND, bool F32, int DEPTH, bool EREC = false>
__device__ __forceinline__ void pass_f_body(DevState &s, const FusedFlags &fl, int n, int tb_, double2 *wbuf)
{
    const double dt = s.sc->dt;
    // the lane of each model's first node ends the step; nothing in this pass
    // reads what it updates
    if (fl.step_end && n % BLOCK == 0 &&
        (s.nblk_model ? (tb_ == 0 || s.nblk_model[tb_ - 1] != s.nblk_model[tb_]) : n == 0))
        step_end_scalars(s.sc, s.P->rheol_type);
    double f[ND], v[ND], x0[ND];
#pragma unroll
    for (int d = 0; d < ND; ++d) v[d] = __builtin_nontemporal_load(s.vel + d * s.N + n);
    // the node's own values, requested before the gather
    const double m = s.mass[n];
    const unsigned flag = s.bcflag[n];
#pragma unroll
    for (int d = 0; d < ND; ++d) x0[d] = __builtin_nontemporal_load(s.coord + d * s.N + n);
    node_force<ND, F32, DEPTH, EREC>(s, n, f, v);
    if (fl.store_aux) {
#pragma unroll
        for (int d = 0; d < ND; ++d) s.force[d * s.N + n] = f[d];
    }
    if (fl.rec_force) {
#pragma unroll
        for (int d = 0; d < ND; ++d) fl.rec_force[d * s.N + n] = f[d];
    }
#pragma unroll
    for (int d = 0; d < ND; ++d) v[d] += dt * f[d] / m;
    if (flag & BOUND_ANY)
        if (!vbc_node<ND>(s, flag, v)) set_error(s.sc, DES_ERR_ASSERT);
    double xn[ND];
#pragma unroll
    for (int d = 0; d < ND; ++d) {
        __builtin_nontemporal_store(v[d], s.vel + d * s.N + n);
        xn[d] = x0[d] + v[d] * dt;
        __builtin_nontemporal_store(xn[d], s.coord + d * s.N + n);
    }
    // the record's x and v: its first ND 16-byte chunks (T and aux stay)
    double2 c[ND];
    if constexpr (ND == 3) {
        c[0] = make_double2(xn[0], xn[1]);
        c[1] = make_double2(xn[2], v[0]);
        c[2] = make_double2(v[1], v[2]);
    } else {
        c[0] = make_double2(xn[0], xn[1]);
        c[1] = make_double2(v[0], v[1]);
    }
    double2 *nrec2 = reinterpret_cast<double2 *>(s.nrec);
    if (!wbuf || !wave_store_records<ND, NREC / 2>(nrec2, n - (int)__lane_id(), c, wbuf)) {
#pragma unroll
        for (int k = 0; k < ND; ++k) nrec2[(long)n * (NREC / 2) + k] = c[k];
    }
}

template <int ND, bool F32, int NB, int DEPTH, bool EREC = false>
__device__ __forceinline__ void pass_f_blk(DevState s, const FusedFlags &fl, int blk, double2 *wbuf)
{
    BIND_NODE_NB_AT(s, n, NB, blk);
    pass_f_body<ND, F32, DEPTH, EREC>(s, fl, n, tb_, wbuf);
}

template <int ND, int NB, int DEPTH = 4> __global__ __launch_bounds__(NB) void k_pass_f(DevState s, FusedFlags fl)
{
    __shared__ double2 wbuf[NB * ND];
    pass_f_blk<ND, false, NB, DEPTH>(s, fl, xcd_block(s.rev), wbuf + (threadIdx.x & ~63) * ND);
}

// pass F of the mixed-precision schedule: FP32 force contributions
// pass F gathering element-major force records
template <int ND, int NB, int DEPTH = 4> __global__ __launch_bounds__(NB) void k_pass_f_er(DevState s, FusedFlags fl)
{
    __shared__ double2 wbuf[NB * ND];
    pass_f_blk<ND, false, NB, DEPTH, true>(s, fl, xcd_block(s.rev), wbuf + (threadIdx.x & ~63) * ND);
}

template <int ND, int NB, int DEPTH = 4> __global__ __launch_bounds__(NB) void k_pass_f32(DevState s, FusedFlags fl)
{
    __shared__ double2 wbuf[NB * ND];
    pass_f_blk<ND, true, NB, DEPTH>(s, fl, xcd_block(s.rev), wbuf + (threadIdx.x & ~63) * ND);
}

// ---- mixed precision (DES_MIXED): passes C and E on (pres, sdev) -------

template <int ND> __global__ __launch_bounds__(BLOCK) void k_to_mixed(DevState s)
{
    BIND_ELEM(s, e);
    constexpr int NSTR = Dim<ND>::NSTR;
    const long E = s.E;
    double st[NSTR], p;
    float sd[NSTR];
    load_tensor<ND>(s.stress, E, e, st);
    split_stress<ND>(st, p, sd);
    s.pres[e] = p;
#pragma unroll
    for (int k = 0; k < NSTR; ++k) {
        s.sdev[k * E + e] = sd[k];
        s.strain32[k * E + e] = (float)s.strain[k * E + e];
    }
    s.dpl32[e] = (float)s.delta_plstrain[e];
    if (s.visc32) const_cast<float *>(s.visc32)[e] = (float)s.elem_visc[e];
    if (s.coh32) const_cast<float *>(s.coh32)[e] = (float)s.elem_cohesion[e];
}

template <int ND> __global__ __launch_bounds__(BLOCK) void k_from_mixed(DevState s)
{
    BIND_ELEM(s, e);
    constexpr int NSTR = Dim<ND>::NSTR;
    const long E = s.E;
    float sd[NSTR];
#pragma unroll
    for (int k = 0; k < NSTR; ++k) sd[k] = s.sdev[k * E + e];
    double st[NSTR];
    join_stress<ND>(s.pres[e], sd, st);
    store_tensor<ND>(s.stress, E, e, st);
#pragma unroll
    for (int k = 0; k < NSTR; ++k) s.strain[k * E + e] = (double)s.strain32[k * E + e];
    s.delta_plstrain[e] = (double)s.dpl32[e];
}

// staged, lds: the block's staged node records (k_pass_c_mixed's staged
// variants); pli, psd, pes, pp: the element's stage_lidx entry, deviator,
// strain and pressure when the caller loaded them
template <int ND>
__device__ __forceinline__ void pass_c_mixed_body(DevState s, const FusedFlags &fl, int blk, bool staged = false,
                                                  const double2 *lds = nullptr, const unsigned *pli = nullptr,
                                                  const float *psd = nullptr, const float *pes = nullptr,
                                                  const double *pp = nullptr)
{
    BIND_ELEM_AT(s, e, blk);
    constexpr int NPE = Dim<ND>::NPE, NSTR = Dim<ND>::NSTR;
    const long E = s.E;
    double x[NPE][ND], sh[ND][NPE], v[NPE][ND], Tn[NPE], aux[NPE];
    load_elem_nodes<ND>(s, e, staged, lds, x, v, Tn, aux, staged ? pli : nullptr);
    const double vol = elem_volume<ND>(x);       // pass A's volume, recomputed (same bits)
    shape_fn<ND>(x, vol, sh);
    // velocities relative to node 0 (sum_i dN_i = 0): no cancellation of the
    // rigid-body part when the rates are rounded to FP32
#pragma unroll
    for (int i = NPE - 1; i >= 0; --i)
#pragma unroll
        for (int d = 0; d < ND; ++d) v[i][d] -= v[0][d];
    double p = pp ? *pp : s.pres[e];
    float sd[NSTR], es[NSTR], edot[NSTR];
#pragma unroll
    for (int k = 0; k < NSTR; ++k) {       // strain and plastic strain: non-temporal streams
        sd[k] = psd ? psd[k] : s.sdev[k * E + e];
        es[k] = pes ? pes[k] : __builtin_nontemporal_load(s.strain32 + k * E + e);
    }
    // strain rate and spin from the same products, before the rotation, so
    // that the velocities and shape functions are no longer held across it
    const bool rot = s.sc->rot_pending;
    double w[3];
    {
        double sr[NSTR];
        strain_rate_spin<ND>(v, sh, sr, w, rot);
#pragma unroll
        for (int k = 0; k < NSTR; ++k) edot[k] = (float)sr[k];
    }
    const float wf[3] = {(float)w[0], (float)w[1], (float)w[2]};
    const float dtp = (float)s.sc->dt_prev;
    double edv = 0;
#pragma unroll
    for (int i = 0; i < NPE; ++i) edv += aux[i];
    edv = edv / NPE;
    // in 3D (returns deferred) the update never changes the plastic strain
    // and only zeroes its increment (reset_dpl): neither is read back to be
    // stored unchanged
    double pl = __builtin_nontemporal_load(s.plstrain + e);
    float dpl = ND == 3 ? 0.f : __builtin_nontemporal_load(s.dpl32 + e);
    bool reset_dpl;
    double dp;
    const float Tm = (float)node_mean<ND>(Tn);
    bool deferred;
    MixedRet rec;
    bool ok = stress_update_mixed<ND, ND == 3>(s, e, (float)s.sc->dt, (float)edv, Tm, vol, s.volume_old[e], p,
                                               sd, es, wf, dtp, rot, edot, pl, dpl, reset_dpl, dp, deferred, rec);
    if (!ok) set_error(s.sc, DES_ERR_INPUT);
    if constexpr (ND == 3) {
        // append the deferred elements, one atomic per wavefront
        const unsigned long long mask = __ballot(deferred);
        if (mask) {
            const int lane = __lane_id();
            const int leader = __ffsll((long long)mask) - 1;
            int base = 0;
            if (lane == leader) base = atomicAdd(s.ret_count, __popcll(mask));
            base = __shfl(base, leader);
            if (deferred) {
                const int k = base + __popcll(mask & ((1ull << lane) - 1));
                s.ret_elem[k] = e;
                s.ret_mix[k] = rec;
            }
        }
    }

    if (!deferred) {
        s.pres[e] = p;
#pragma unroll
        for (int k = 0; k < NSTR; ++k) s.sdev[k * E + e] = sd[k];
        if (ND == 2) __builtin_nontemporal_store(pl, s.plstrain + e);
        if (ND == 2 || reset_dpl) __builtin_nontemporal_store(dpl, s.dpl32 + e);
        s.dpressure[e] = dp;
        s.escal[e] = dp * vol;
    }
    if (fl.store_aux) {
#pragma unroll
        for (int k = 0; k < NSTR; ++k) s.strain_rate[k * E + e] = (double)edot[k];
        s.edvoldt[e] = edv;
    }
}

// STAGE: the block's node records staged in LDS, with the element's own data
// requested before the staging barrier (as in k_pass_e)
template <int ND, int WPE, bool STAGE = false>
__global__ __launch_bounds__(BLOCK, WPE) void k_pass_c_mixed(DevState s, FusedFlags fl)
{
    if constexpr (STAGE) {
        __shared__ double2 lds[STAGE_MAX * (NREC / 2)];
        constexpr int NSTR = Dim<ND>::NSTR;
        const int blk = xcd_block(s.rev);
        unsigned pli = 0;
        float psd[NSTR], pes[NSTR];
        double pp = 0;
        const int e0 = blk * BLOCK + threadIdx.x;
        if (s.eblk_model ? e0 < s.eblk_lim[blk] : e0 < s.E) {
            if (s.stage) pli = s.stage_lidx[e0];
#pragma unroll
            for (int k = 0; k < NSTR; ++k) {
                psd[k] = s.sdev[k * (long)s.E + e0];
                pes[k] = __builtin_nontemporal_load(s.strain32 + k * (long)s.E + e0);
            }
            pp = s.pres[e0];
        }
        const bool staged = stage_block(s, blk, lds);
        pass_c_mixed_body<ND>(s, fl, blk, staged, lds, staged ? &pli : nullptr, psd, pes, &pp);
    } else {
        pass_c_mixed_body<ND>(s, fl, xcd_block(s.rev));
    }
}

// The Mohr-Coulomb returns that the mixed pass C deferred (3D), completing
// each element as stress_update_mixed's inline return would.
__device__ __forceinline__ void pass_c_ret_mixed_item(const DevState &s, int k)
{
    const long E = s.E;
    {
        const int e = s.ret_elem[k];
        DevState st = s;
        if (s.eblk_model) bind_model<3>(st, s.eblk_model[e / BLOCK]);
        const des_params *P = st.P;
        const MixedRet &r = s.ret_mix[k];
        const float K = (float)harmonic_mean(st, P->bulk_modulus, e);
        const float G = (float)harmonic_mean(st, P->shear_modulus, e);
        double pl = s.plstrain[e];
        float dpl = s.dpl32[e];
        double amc, anphi, anpsi, hardn, ten_max;
        mat_plastic_props_f(st, e, (float)pl, amc, anphi, anpsi, hardn, ten_max);
        double s64[6];
#pragma unroll
        for (int i = 0; i < 6; ++i) s64[i] = r.s64[i];
        double depls;
        mc_return<3>((double)K, (double)G, amc, anphi, anpsi, hardn, ten_max, depls, s64);
        float sd[6];
#pragma unroll
        for (int i = 0; i < 6; ++i) sd[i] = s.sdev[i * E + e];
        double p;
        if (P->rheol_type == DES_RH_EVP && (double)r.svII < second_invariant2<3>(s64)) {
            p = r.pv;
#pragma unroll
            for (int i = 0; i < 6; ++i) sd[i] += r.incv[i];
        } else if (depls == 0) {
            p = r.pp;
#pragma unroll
            for (int i = 0; i < 6; ++i) sd[i] += r.inc[i];
            dpl = 0;
        } else {
            split_stress<3>(s64, p, sd);
            pl += depls;
            dpl = (float)depls;
        }
        s.pres[e] = p;
#pragma unroll
        for (int i = 0; i < 6; ++i) s.sdev[i * E + e] = sd[i];
        s.plstrain[e] = pl;
        s.dpl32[e] = dpl;
        const double dp = 3 * (p - r.p_old);
        s.dpressure[e] = dp;
        s.escal[e] = dp * s.volume[e];
    }
}

__global__ __launch_bounds__(BLOCK) void k_pass_c_ret_mixed(DevState s)
{
    const int cnt = *s.ret_count;
    for (int k = blockIdx.x * BLOCK + threadIdx.x; k < cnt; k += gridDim.x * BLOCK) pass_c_ret_mixed_item(s, k);
}

template <int ND> __device__ __forceinline__ void pass_e_mixed_body(DevState s, const FusedFlags &fl, int blk,
                                                                  bool staged, const double2 *lds,
                                                                  const double *pre = nullptr, const float *psd = nullptr,
                                                                  const unsigned *pli = nullptr, const int *loff = nullptr)
{
    BIND_ELEM_AT(s, e, blk);
    constexpr int NPE = Dim<ND>::NPE, NSTR = Dim<ND>::NSTR;
    const long E = s.E;
    double x[NPE][ND], v[NPE][ND], Tn[NPE], dpn[NPE];
    load_elem_nodes<ND>(s, e, staged, lds, x, v, Tn, dpn, staged ? pli : nullptr);
    const double dp_el = node_mean<ND>(dpn);
    double p = (pre ? pre[0] : s.pres[e]) + (-(pre ? pre[1] : s.dpressure[e]) + dp_el) / ND;
    s.pres[e] = p;
    float sd[NSTR];
#pragma unroll
    for (int k = 0; k < NSTR; ++k) sd[k] = psd ? psd[k] : s.sdev[k * E + e];
    double st[NSTR];
    join_stress<ND>(p, sd, st);
    double sh[ND][NPE], out[NPE][ND];
    const double vol = elem_volume<ND>(x);       // pass A's volume, recomputed (same bits)
    shape_fn<ND>(x, vol, sh);
    force_contrib<ND>(s, e, vol, sh, st, node_mean<ND>(Tn), out);
    int pos[NPE];
    force_slots<ND>(s, e, staged, loff, pos, staged ? pli : nullptr);
    store_force_contrib_f32<ND>(s, pos, out);
}

template <int ND, bool STAGE> __global__ __launch_bounds__(BLOCK) void k_pass_e_mixed(DevState s, FusedFlags fl)
{
    __shared__ double2 lds[STAGE ? STAGE_MAX * (NREC / 2) : 1];
    __shared__ int loff[STAGE ? STAGE_MAX : 1];
    const int blk = xcd_block(s.rev);
    // pressure, pressure change, deviator, staged-node and slot positions
    // requested before the staging barrier (as in k_pass_e)
    constexpr int NSTR = Dim<ND>::NSTR;
    double pre[2];
    float psd[NSTR];
    unsigned pli[2] = {0u, 0u};
    if constexpr (STAGE) {
        const int e0 = blk * BLOCK + threadIdx.x;
        if (s.eblk_model ? e0 < s.eblk_lim[blk] : e0 < s.E) {
            pre[0] = s.pres[e0];
            pre[1] = s.dpressure[e0];
#pragma unroll
            for (int k = 0; k < NSTR; ++k) psd[k] = s.sdev[k * (long)s.E + e0];
            if (s.stage) pli[0] = s.stage_lidx[e0];
            if (s.srank) pli[1] = s.srank[e0];
        }
    }
    const bool staged = STAGE && stage_block(s, blk, lds, loff);
    pass_e_mixed_body<ND>(s, fl, blk, staged, lds, STAGE ? pre : nullptr, STAGE ? psd : nullptr,
                          STAGE ? pli : nullptr, loff);
}

// node records from the SoA state, at the start of a fused run
template <int ND> __global__ __launch_bounds__(BLOCK) void k_pack_nrec(DevState s)
{
    const int n = blockIdx.x * BLOCK + threadIdx.x;
    if (n >= s.N) return;
    double *rn = s.nrec + (long)NREC * n;
#pragma unroll
    for (int d = 0; d < ND; ++d) {
        rn[NrecSlot<ND>::X + d] = s.coord[d * s.N + n];
        rn[NrecSlot<ND>::V + d] = s.vel[d * s.N + n];
    }
    rn[NrecSlot<ND>::T] = s.T[n];
    rn[NrecSlot<ND>::AUX] = s.ntmp[n];
#pragma unroll
    for (int k = 2 * ND + 2; k < NREC; ++k) rn[k] = 0;
}

// step bookkeeping after pass F: var.time, var.steps, the rotation to defer

__global__ void k_clear_rotation_b(DevState s)
{
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m < s.nmodels) s.scb[m].rot_pending = 0;
}

__global__ void k_clear_rotation(DevState s)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    s.sc->rot_pending = 0;
}

__global__ void k_step_begin(DevState s)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    s.sc->time += s.sc->dt;
    s.sc->steps += 1;
}

// ---- the persistent launch: a whole run of fused steps in one kernel
//
// One cooperative launch holds as many workgroups as can be resident at once
// (a multiple of 8) and runs every step of the run: the passes in the fused
// order, each a walk over its logical blocks, separated by grid barriers
// (every pass reads what the one before it wrote, on all XCDs).  Workgroup w
// sits on XCD w mod 8 and walks that XCD's contiguous share of each pass's
// logical blocks, in the directions of the per-pass launches, so each XCD
// works on the same part of the mesh in every pass.  Up to two groups run in
// the one launch (the 3D and the 2D models of a batch): each pass walks the
// 3D group's blocks, then the 2D group's.  The per-block code is the
// per-pass kernels' own, with the variants fixed: direct node-record
// gathers, 256-thread node blocks, 8 entries in flight.

struct PersistGroup {
    int neb, nqb, nnb;          // element blocks, pass B blocks (a quad per node), node blocks
    int plastic;                // 3D and some model is plastic: the return list runs
    int nactive;                // traction-loaded facets
    const int *active_facets;
    int nmodels;                // batched group: per-model compute_dt
};

struct PersistArgs {
    int nsteps;                 // fused steps in this launch
    int steps0;                 // var.steps before its first step
    int use_stored_mass;        // its first step is a run's first (mass from compute_mass)
    int swap0;                  // its first step swaps volume and volume_old too (a continued run)
    int store_aux;              // its last step ends the run
    int rev;                    // node passes walk downwards
    PersistGroup g3, g2;
};

// logical blocks of this workgroup's walk over n3 + n2 blocks (3D group's
// first): body3(b) for b < n3, body2(b - n3) after
template <class F3, class F2>
__device__ __forceinline__ void pwalk(int n3, int n2, int rev, F3 &&body3, F2 &&body2)
{
    const int nvb = n3 + n2, G8 = gridDim.x >> 3, x = blockIdx.x & 7;
    const int q = nvb >> 3, r = nvb & 7, len = q + (x < r ? 1 : 0), base = x * q + min(x, r);
#pragma nounroll
    for (int j = blockIdx.x >> 3; j < len; j += G8) {
        const int b = base + (rev ? len - 1 - j : j);
        if (b < n3) body3(b);
        else body2(b - n3);
    }
}

// a grid-stride walk over n items
template <class F> __device__ __forceinline__ void pitems(int n, F &&body)
{
#pragma nounroll
    for (int t = blockIdx.x * BLOCK + threadIdx.x; t < n; t += gridDim.x * BLOCK) body(t);
}

// the return-list items as calls: the eigen solve's registers are not
// reserved across the other passes' code
__device__ __attribute__((noinline)) void pass_c_ret_call(const DevState &s, int t) { pass_c_ret_item(s, t); }
__device__ __attribute__((noinline)) void pass_c_ret_mixed_call(const DevState &s, int t) { pass_c_ret_mixed_item(s, t); }

// group state at step k of the launch: the state in device memory (read
// after each barrier, so no pass keeps it in registers across the others),
// with volume and volume_old swapped on odd steps (update_mesh's swap)
__device__ __forceinline__ DevState pstate(const DevState *ps, int k)
{
    DevState s = *ps;
    if (k & 1) {         // k counts the swaps made since the launch
        double *t = s.volume;
        s.volume = s.volume_old;
        s.volume_old = t;
    }
    return s;
}

// a pass of both groups: CALL3 / CALL2 on the 3D / 2D group's block b, with
// the group's state in s
#define PWALK(n3, n2, rev, CALL3, CALL2)                                                    \
    pwalk((n3), (n2), (rev),                                                                 \
          [&](int b) { if constexpr (HAS3) { DevState s = pstate(ps3, k + pa.swap0); CALL3; } },        \
          [&](int b) { if constexpr (HAS2) { DevState s = pstate(ps2, k + pa.swap0); CALL2; } })

template <bool MIXED, bool HAS3, bool HAS2>
__global__ __launch_bounds__(BLOCK, 2) void k_persistent(const DevState *__restrict__ ps3, const DevState *__restrict__ ps2, PersistArgs pa)
{
    cooperative_groups::grid_group grid = cooperative_groups::this_grid();
    __shared__ double2 xbuf[BLOCK * (REC / 2)];       // pass A's records, pass F's node records
    __shared__ double sm[3][BLOCK];                   // compute_dt's block minima
    double2 *wbuf = xbuf + (threadIdx.x & ~63) * (REC / 2);
    const PersistGroup &g3 = pa.g3, &g2 = pa.g2;
    const int ne3 = HAS3 ? g3.neb : 0, ne2 = HAS2 ? g2.neb : 0;
    const int nq3 = HAS3 ? g3.nqb : 0, nq2 = HAS2 ? g2.nqb : 0;
    const int nn3 = HAS3 ? g3.nnb : 0, nn2 = HAS2 ? g2.nnb : 0;
    const int nf3 = HAS3 ? g3.nactive : 0, nf2 = HAS2 ? g2.nactive : 0;
#pragma nounroll
    for (int k = 0; k < pa.nsteps; ++k) {
        FusedFlags fl;
        fl.use_stored_mass = pa.use_stored_mass && k == 0;
        fl.store_aux = pa.store_aux && k == pa.nsteps - 1;
        PWALK(ne3, ne2, 0, pass_a_body<3>(s, fl, b, false, nullptr, xbuf),
              pass_a_body<2>(s, fl, b, false, nullptr, xbuf));
        if (blockIdx.x == 0 && threadIdx.x == 0) {      // pass C appends to the return lists
            if constexpr (HAS3) if (ps3->ret_count) *ps3->ret_count = 0;
            if constexpr (HAS2) if (ps2->ret_count) *ps2->ret_count = 0;
        }
        grid.sync();
        PWALK(nq3, nq2, pa.rev, (pass_b_body<3, 8>(s, fl, b)), (pass_b_body<2, 8>(s, fl, b)));
        grid.sync();
        if constexpr (MIXED)
            PWALK(ne3, ne2, 0, pass_c_mixed_body<3>(s, fl, b), pass_c_mixed_body<2>(s, fl, b));
        else
            PWALK(ne3, ne2, 0, pass_c_blk<3>(s, fl, b, false, nullptr), pass_c_blk<2>(s, fl, b, false, nullptr));
        grid.sync();
        if constexpr (HAS3) {
            if (g3.plastic) {       // the deferred Mohr-Coulomb returns (3D only)
                const DevState s = pstate(ps3, k + pa.swap0);
                const int cnt = __hip_atomic_load(s.ret_count, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                if constexpr (MIXED) pitems(cnt, [&](int t) { pass_c_ret_mixed_call(s, t); });
                else pitems(cnt, [&](int t) { pass_c_ret_call(s, t); });
                grid.sync();
            }
        }
        PWALK(nn3, nn2, pa.rev, (pass_d_body<3, BLOCK, 8>(s, fl, b)), (pass_d_body<2, BLOCK, 8>(s, fl, b)));
        grid.sync();
        if constexpr (MIXED)
            PWALK(ne3, ne2, 0, pass_e_mixed_body<3>(s, fl, b, false, nullptr),
                  pass_e_mixed_body<2>(s, fl, b, false, nullptr));
        else
            PWALK(ne3, ne2, 0, pass_e_body<3>(s, fl, b, false, nullptr, nullptr),
                  pass_e_body<2>(s, fl, b, false, nullptr, nullptr));
        grid.sync();
        if (nf3 + nf2) {
            pitems(nf3 + nf2, [&](int t) {
                if constexpr (HAS3) if (t < nf3) facet_traction_item<3>(pstate(ps3, k + pa.swap0), g3.active_facets, t);
                if constexpr (HAS2) if (t >= nf3) facet_traction_item<2>(pstate(ps2, k + pa.swap0), g2.active_facets, t - nf3);
            });
            grid.sync();
        }
        FusedFlags ff = fl;
        ff.step_end = 1;
        PWALK(nn3, nn2, pa.rev, (pass_f_blk<3, MIXED, BLOCK, 8>(s, ff, b, wbuf)),
              (pass_f_blk<2, MIXED, BLOCK, 8>(s, ff, b, wbuf)));
        grid.sync();
        if ((pa.steps0 + k + 1) % 10 == 0) {          // compute_dt every tenth step
            PWALK(ne3, ne2, 0, dt_partial_blk<3>(s, b, sm), dt_partial_blk<2>(s, b, sm));
            grid.sync();
            const int nm3 = HAS3 ? max(g3.nmodels, 1) : 0, nm2 = HAS2 ? max(g2.nmodels, 1) : 0;
            for (int m = blockIdx.x; m < nm3 + nm2; m += gridDim.x) {
                if constexpr (HAS3) {
                    if (m < nm3) {
                        const DevState s = pstate(ps3, k + pa.swap0);
                        if (g3.nmodels) dt_final_b_body<3>(s, m, sm);
                        else dt_final_body<3>(s, ne3, sm);
                    }
                }
                if constexpr (HAS2) {
                    if (m >= nm3) {
                        const DevState s = pstate(ps2, k + pa.swap0);
                        if (g2.nmodels) dt_final_b_body<2>(s, m - nm3, sm);
                        else dt_final_body<2>(s, ne2, sm);
                    }
                }
            }
            grid.sync();
        }
    }
}
#undef PWALK

}  // namespace

// ================================================================ host side

// a checkpoint of the differentiated run: the state a fused step reads at its start
struct AdjCkpt {
    std::vector<double *> a;      // device copies
    DevScalars sc;
    int vol_cur, steps;
    bool closing_pending;
};

struct des_ctx {
    des_params P;
    int nd, npe, nstr, N, E, nbf;
    int device = 0;                // the HIP device current at des_create
    hipStream_t stream = nullptr;
    // start/stop events handed to the next launches (timed runs only): the
    // dispatch itself timestamps them, so they time the kernel's execution
    hipEvent_t ev_a = nullptr, ev_b = nullptr;
    DevState s{};
    std::vector<void *> allocs;
    std::string err;
    double *vol_buf[2] = {nullptr, nullptr};
    int vol_cur = 0;               // s.volume == vol_buf[vol_cur]
    int *active_facets = nullptr;  // facets whose traction enters the force
    int nactive = 0;
    int steps = 0;                 // host mirror of var.steps
    int avg_interval = 0;          // Output::average_interval (0: averaged fields off)
    int launch_mode = 0;           // des_run: 0 one launch per fused pass, 1 one persistent launch
    DevState *pstate_dev = nullptr;  // the persistent launch's copy of s in device memory
    DevState pstate_host{};          // its host source (kept until the copy has run)
    bool closing_pending = false;  // des_run left volume/mass/shape functions/rotation deferred
    bool continuing = false;       // this run carries on from the last one's deferred tail
    int run_precision = -1;        // the precision of the last run
    const unsigned *srank = nullptr;   // the entries' positions in their node lists (DevState::srank)
    std::vector<int> eh, nh;       // device element / node -> host index (model-local in a group)
    double *xfer = nullptr;        // device staging of host-order transfers, grown on demand
    size_t xfer_cap = 0;           // its size in doubles
    std::vector<int> h_conn;       // host copies for output frames
    std::vector<unsigned> h_bcflag;
    // the autotuner's choice of variant per tunable fused pass (DES_TUNE_*
    // order: A, C, D, E, F, B), per precision (0 FP64, 1 mixed); -1 untuned
    int tune_var[2][DES_NTUNE] = {{-1, -1, -1, -1, -1, -1}, {-1, -1, -1, -1, -1, -1}};
    double tune_ms[2][DES_NTUNE][DES_TUNE_MAXV] = {};   // mean ms per variant it measured
    struct TuneLaunch { int pass, variant; hipEvent_t a, b; };
    std::vector<TuneLaunch> tune;          // timed tuning launches in flight
    // batched group (nmodels > 0): models of one dimension concatenated, each
    // starting on a block boundary of the element and of the node arrays
    int nmodels = 0;
    std::vector<des_params> hP;
    std::vector<int> e_off, n_off, e_cnt, n_cnt, nbf_cnt;
    std::vector<char> has_visc, has_coh;  // per model: per-element viscosity / cohesion uploaded
    // differentiated run (des_adj_forward / des_adj_backward)
    std::vector<void *> adj_allocs;
    std::vector<AdjCkpt> adj_ckpt;
    int adj_nsteps = 0, adj_window = 0;
    // taped differentiated run (window 0): every step's records kept from
    // the forward, 8 buffers per step (NR, S, P, VO, F, m, vn, dpn), and each
    // step's deferred-rotation flag
    std::vector<double *> adj_tape;
    std::vector<int> adj_tape_rot;
    // checkpointed run: one window's records (8 buffers per step), and
    // whether they still hold the run's last window, recorded by the forward
    std::vector<double *> adj_win;
    std::vector<int> adj_win_rot;
    bool adj_win_last = false;
};

// host description of a batched group's concatenated mesh
struct GroupLayout {
    std::vector<des_params> P;
    std::vector<int> e_off, n_off, e_cnt, n_cnt, nbf_cnt;
    std::vector<int> facet_model;          // model of each boundary facet
    std::vector<double> bn, ev;            // per model 10 x nd, 10 x 10 x nd
    std::vector<int> evv;                  // per model 10 x 10
};

namespace {

int fail(des_ctx *c, int code, const std::string &msg)
{
    if (c) c->err = msg;
    return code;
}

#define HIPCHK(call)                                                                     \
    do {                                                                                 \
        hipError_t e_ = (call);                                                          \
        if (e_ != hipSuccess)                                                            \
            return fail(ctx, DES_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

int dalloc(des_ctx *ctx, void **p, size_t bytes)
{
    if (bytes == 0) bytes = 8;
    HIPCHK(hipMalloc(p, bytes));
    ctx->allocs.push_back(*p);
    HIPCHK(hipMemsetAsync(*p, 0, bytes, ctx->stream));
    return DES_OK;
}

template <typename T> int dalloc_n(des_ctx *ctx, T **p, size_t n)
{
    return dalloc(ctx, reinterpret_cast<void **>(p), n * sizeof(T));
}

int check_launch(des_ctx *ctx)
{
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(ctx, DES_ERR_HIP, std::string("kernel launch: ") + hipGetErrorString(e));
    return DES_OK;
}

#define LAUNCH_ND(kern, grid, ...)                                                        \
    do {                                                                                  \
        if (ctx->nd == 3)                                                                 \
            hipExtLaunchKernelGGL(kern<3>, dim3(grid), dim3(BLOCK), 0, ctx->stream, ctx->ev_a, ctx->ev_b, 0, __VA_ARGS__); \
        else                                                                              \
            hipExtLaunchKernelGGL(kern<2>, dim3(grid), dim3(BLOCK), 0, ctx->stream, ctx->ev_a, ctx->ev_b, 0, __VA_ARGS__); \
        int rc_ = check_launch(ctx);                                                      \
        if (rc_) return rc_;                                                              \
    } while (0)

bool facet_loaded(const des_params &P, int i)
{
    if (P.gravity == 0) return false;
    int t = P.vbc_types[i];
    if (t != 0 && t != 2 && t != 4) return false;
    if (i == 4 && !P.has_winkler_foundation) return false;
    if (i == 5 && !P.has_water_loading) return false;
    return true;
}

// host AoS [n][k] <-> device SoA [k][stride], rows off .. off+n-1; device
// row off + i holds host row perm[i] (i without perm)
// Host-order AoS (row perm[i], k values per row) <-> device SoA (component j
// of device row i at soa[j * stride + i]), one thread per device row
__global__ __launch_bounds__(BLOCK) void k_soa_to_aos(const double *soa, long stride, long n, int k,
                                                      const int *perm, double *aos)
{
    const long i = blockIdx.x * (long)BLOCK + threadIdx.x;
    if (i >= n) return;
    const long h = perm ? perm[i] : i;
    for (int j = 0; j < k; ++j) aos[h * k + j] = soa[j * stride + i];
}

__global__ __launch_bounds__(BLOCK) void k_aos_to_soa(const double *aos, long n, int k, const int *perm,
                                                      double *soa, long stride)
{
    const long i = blockIdx.x * (long)BLOCK + threadIdx.x;
    if (i >= n) return;
    const long h = perm ? perm[i] : i;
    for (int j = 0; j < k; ++j) soa[j * stride + i] = aos[h * k + j];
}

// the device copy of a host permutation (a range of eh or nh)
const int *device_perm(const des_ctx *ctx, const int *perm)
{
    const auto in = [&](const std::vector<int> &v) {
        return !v.empty() && perm >= v.data() && perm < v.data() + v.size();
    };
    if (in(ctx->eh)) return ctx->s.ehost + (perm - ctx->eh.data());
    if (in(ctx->nh)) return ctx->s.nhost + (perm - ctx->nh.data());
    return nullptr;
}

int ensure_xfer(des_ctx *ctx, size_t n)
{
    if (ctx->xfer_cap >= n) return DES_OK;
    if (ctx->xfer) {
        HIPCHK(hipStreamSynchronize(ctx->stream));
        HIPCHK(hipFree(ctx->xfer));
        ctx->xfer = nullptr;
        ctx->xfer_cap = 0;
    }
    HIPCHK(hipMalloc(&ctx->xfer, std::max<size_t>(n, 1) * sizeof(double)));
    ctx->xfer_cap = n;
    return DES_OK;
}

// Transfers in the host's numbering: the rows are permuted and interleaved
// on the device, and one copy crosses PCIe.
// src_dev: src is device memory (host row order, AoS) and is read in place
int upload_soa(des_ctx *ctx, double *dst, const double *src, long n, int k, long stride = -1, long off = 0,
               const int *perm = nullptr, bool src_dev = false)
{
    if (!src || n <= 0) return DES_OK;
    if (stride < 0) stride = n;
    const int *dp = device_perm(ctx, perm);
    if (perm && !dp) return fail(ctx, DES_ERR_INPUT, "upload: unknown row order");
    int rc;
    const double *aos = src;
    if (!src_dev) {
        if ((rc = ensure_xfer(ctx, (size_t)n * k))) return rc;
        HIPCHK(hipMemcpyAsync(ctx->xfer, src, sizeof(double) * n * k, hipMemcpyHostToDevice, ctx->stream));
        aos = ctx->xfer;
    }
    hipLaunchKernelGGL(k_aos_to_soa, dim3(nblocks(n)), dim3(BLOCK), 0, ctx->stream, aos, n, k, dp,
                       dst + off, stride);
    if ((rc = check_launch(ctx))) return rc;
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return DES_OK;
}

// dst_dev: dst is device memory (host row order, AoS) written in place
int download_soa(des_ctx *ctx, double *dst, const double *src, long n, int k, long stride = -1, long off = 0,
                 const int *perm = nullptr, bool dst_dev = false)
{
    if (!dst || n <= 0) return DES_OK;
    if (stride < 0) stride = n;
    const int *dp = device_perm(ctx, perm);
    if (perm && !dp) return fail(ctx, DES_ERR_INPUT, "download: unknown row order");
    int rc;
    if (!dst_dev && (rc = ensure_xfer(ctx, (size_t)n * k))) return rc;
    hipLaunchKernelGGL(k_soa_to_aos, dim3(nblocks(n)), dim3(BLOCK), 0, ctx->stream, src + off, stride, n, k, dp,
                       dst_dev ? dst : ctx->xfer);
    if ((rc = check_launch(ctx))) return rc;
    if (!dst_dev)
        HIPCHK(hipMemcpyAsync(dst, ctx->xfer, sizeof(double) * n * k, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return DES_OK;
}

int read_scalars(des_ctx *ctx, DevScalars &h)
{
    HIPCHK(hipMemcpyAsync(&h, ctx->s.sc, sizeof(DevScalars), hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return DES_OK;
}

void swap_volume(des_ctx *ctx)
{
    ctx->vol_cur ^= 1;
    ctx->s.volume = ctx->vol_buf[ctx->vol_cur];
    ctx->s.volume_old = ctx->vol_buf[ctx->vol_cur ^ 1];
}

int launch_compute_dt(des_ctx *ctx)
{
    if (ctx->nmodels) {          // per model, entirely on the device
        LAUNCH_ND(k_dt_partial_b, nblocks(ctx->E), ctx->s);
        LAUNCH_ND(k_dt_final_b, ctx->nmodels, ctx->s);
        return DES_OK;
    }
    if (ctx->P.fixed_dt == 0) LAUNCH_ND(k_dt_partial, RED_BLOCKS, ctx->s);
    LAUNCH_ND(k_dt_final, 1, ctx->s, RED_BLOCKS);
    return DES_OK;
}

int launch_mass(des_ctx *ctx)
{
    LAUNCH_ND(k_mass_elem, nblocks(ctx->E), ctx->s);
    LAUNCH_ND(k_mass_node, nblocks(ctx->N), ctx->s);
    return DES_OK;
}

int launch_update_mesh(des_ctx *ctx)
{
    LAUNCH_ND(k_update_coordinate, nblocks(ctx->N), ctx->s);
    swap_volume(ctx);
    LAUNCH_ND(k_volume, nblocks(ctx->E), ctx->s);
    int rc = launch_mass(ctx);
    if (rc) return rc;
    LAUNCH_ND(k_shape_fn, nblocks(ctx->E), ctx->s);
    return DES_OK;
}

int launch_force(des_ctx *ctx)
{
    LAUNCH_ND(k_force_elem, nblocks(ctx->E), ctx->s);
    if (ctx->nactive) LAUNCH_ND(k_facet_traction, nblocks(ctx->nactive), ctx->s, ctx->active_facets, ctx->nactive);
    LAUNCH_ND(k_force_node, nblocks(ctx->N), ctx->s);
    return DES_OK;
}

// bring the device state back to the reference's Variables after des_run
int flush_deferred(des_ctx *ctx)
{
    if (!ctx->closing_pending) return DES_OK;
    // the tail of the last step: update_mesh after update_coordinate, then
    // rotate_stress with that step's dt (dynearthsol.cxx:226-231, 346-347)
    swap_volume(ctx);
    LAUNCH_ND(k_volume, nblocks(ctx->E), ctx->s);
    int rc = launch_mass(ctx);
    if (rc) return rc;
    LAUNCH_ND(k_shape_fn, nblocks(ctx->E), ctx->s);
    if (ctx->nmodels) {          // k_rotate skips the models without elasticity
        LAUNCH_ND(k_rotate, nblocks(ctx->E), ctx->s, 1);
        hipLaunchKernelGGL(k_clear_rotation_b, dim3(nblocks(ctx->nmodels)), dim3(BLOCK), 0, ctx->stream, ctx->s);
    } else {
        if (ctx->P.rheol_type & DES_RH_ELASTIC) LAUNCH_ND(k_rotate, nblocks(ctx->E), ctx->s, 1);
        hipLaunchKernelGGL(k_clear_rotation, dim3(1), dim3(1), 0, ctx->stream, ctx->s);
    }
    ctx->closing_pending = false;
    return check_launch(ctx);
}

}  // namespace

// ------------------------------------------------------------------ C-ABI

namespace {

// ---- device numbering of a mesh
//
// Elements and nodes live on the device in Hilbert-curve order of their
// centroids and positions: the 64 elements of a wavefront share nodes and the
// 64 nodes of a wavefront share elements, so the gathers of every pass touch
// few cache lines, and the contiguous range each XCD runs is a compact region
// whose data its L2 holds.  (The Hilbert curve has fewer long jumps between
// neighbouring cells than the Morton curve: 1.7 % faster steps.)  The host
// numbering stays the interface: every transfer permutes rows, every index
// handed back is a host index, and each node's CSR lists its elements in
// ascending host order, so every sum runs in the reference's order and the
// results are bitwise those of the host numbering.
int g_mesh_order = 1;    // 1: Hilbert-curve order, 0: host numbering (des_set_mesh_order)
int g_sweep = 1;         // 1: node passes walk their blocks downwards (des_set_sweep_order)
int g_force_layout = 0;  // fused FP64 force contributions: 0 node-major slots, 1 element-major records,
                         // 2 node-major slots addressed by their full 4-byte positions

// Position along the Hilbert curve of the cell X[0..nd) (bits per axis):
// J. Skilling, "Programming the Hilbert curve", AIP Conf. Proc. 707, 381
// (2004) -- the axes are transformed in place into the curve's transposed
// index, whose bits, interleaved from the top, are the position.
uint64_t hilbert_index(uint32_t X[3], int nd, int bits)
{
    const uint32_t M = 1u << (bits - 1);
    for (uint32_t Q = M; Q > 1; Q >>= 1) {        // inverse undo of the excess work
        const uint32_t P = Q - 1;
        for (int d = 0; d < nd; ++d) {
            if (X[d] & Q) {
                X[0] ^= P;
            } else {
                const uint32_t t = (X[0] ^ X[d]) & P;
                X[0] ^= t;
                X[d] ^= t;
            }
        }
    }
    for (int d = 1; d < nd; ++d) X[d] ^= X[d - 1];  // Gray encode
    uint32_t t = 0;
    for (uint32_t Q = M; Q > 1; Q >>= 1)
        if (X[nd - 1] & Q) t ^= Q - 1;
    for (int d = 0; d < nd; ++d) X[d] ^= t;
    uint64_t k = 0;
    for (int b = bits - 1; b >= 0; --b)
        for (int d = 0; d < nd; ++d) k = (k << 1) | ((X[d] >> b) & 1u);
    return k;
}

// f(t) for t = 0 .. T-1 on T host threads
template <class F> void run_threads(int T, F f)
{
    if (T <= 1) {
        f(0);
        return;
    }
    std::vector<std::thread> th;
    for (int t = 0; t < T; ++t) th.emplace_back(f, t);
    for (auto &x : th) x.join();
}

// host threads for n items (at most 16, at least 64 k items each)
int host_threads(long n)
{
    const long hw = std::max(1u, std::thread::hardware_concurrency());
    return (int)std::max(1L, std::min({hw, 16L, n / 65536}));
}

// perm[device index] = host index, by Hilbert index of the n points x[n][nd]
// (one scale for every axis, so the curve's cells are cubes).  Keys and the
// sort run on host threads: chunks sorted in parallel, then merged pairwise.
void curve_order(const double *x, long n, int nd, std::vector<int> &perm)
{
    double lo[3] = {DBL_MAX, DBL_MAX, DBL_MAX}, hi[3] = {-DBL_MAX, -DBL_MAX, -DBL_MAX};
    for (long i = 0; i < n; ++i)
        for (int d = 0; d < nd; ++d) {
            lo[d] = std::min(lo[d], x[i * nd + d]);
            hi[d] = std::max(hi[d], x[i * nd + d]);
        }
    double ext = 0;
    for (int d = 0; d < nd; ++d) ext = std::max(ext, hi[d] - lo[d]);
    const double cells = (double)((1u << (nd == 3 ? 21 : 31)) - 1);
    const double scale = ext > 0 ? cells / ext : 0;
    std::vector<std::pair<uint64_t, int>> key(n);
    const int T = host_threads(n);
    std::vector<long> b(T + 1);
    for (int t = 0; t <= T; ++t) b[t] = n * t / T;
    run_threads(T, [&](int t) {
        for (long i = b[t]; i < b[t + 1]; ++i) {
            uint32_t X[3] = {0, 0, 0};
            for (int d = 0; d < nd; ++d) {
                double q = std::floor((x[i * nd + d] - lo[d]) * scale);
                q = std::min(std::max(q, 0.0), cells);
                X[d] = (uint32_t)q;
            }
            key[i] = {hilbert_index(X, nd, nd == 3 ? 21 : 31), (int)i};
        }
        std::sort(key.begin() + b[t], key.begin() + b[t + 1]);
    });
    for (int w = 1; w < T; w *= 2) {
        const int pairs = (T + 2 * w - 1) / (2 * w);
        run_threads(pairs, [&](int p) {
            const int lo_ = 2 * w * p, mid = lo_ + w, hi_ = std::min(lo_ + 2 * w, T);
            if (mid < hi_)
                std::inplace_merge(key.begin() + b[lo_], key.begin() + b[mid], key.begin() + b[hi_]);
        });
    }
    perm.resize(n);
    for (long i = 0; i < n; ++i) perm[i] = key[i].second;
}

// A mesh renumbered into device order, with the maps back to host indices.
struct OrderedMesh {
    std::vector<int> eh, nh;            // device element / node -> host index
    std::vector<int> conn, felem, markers;
    std::vector<double> coord;
    std::vector<unsigned> bcflag;
    des_mesh m;                         // in device order (other arrays: the input's)
};

void order_mesh(const des_mesh &in, int nd, int nmat, int mode, OrderedMesh &o)
{
    const long E = in.nelem, N = in.nnode;
    const int npe = nd + 1;
    o.nh.resize(N);
    o.eh.resize(E);
    if (mode == 0) {
        for (long i = 0; i < N; ++i) o.nh[i] = (int)i;
        for (long i = 0; i < E; ++i) o.eh[i] = (int)i;
    } else {
        std::vector<double> cen((size_t)E * nd, 0.0);
        for (long e = 0; e < E; ++e)
            for (int i = 0; i < npe; ++i)
                for (int d = 0; d < nd; ++d) cen[e * nd + d] += in.coord[(size_t)in.connectivity[e * npe + i] * nd + d];
        curve_order(cen.data(), E, nd, o.eh);
        if (mode == 1) {
            curve_order(in.coord, N, nd, o.nh);
        } else {
            // nodes follow the elements: numbered in the order the curve-ordered
            // elements first touch them (mode 2), or by the mean device index
            // of their elements (mode 3), so that a block of nodes gathers from
            // one compact window of element records
            std::vector<long> key(N, -1), sum(N, 0);
            std::vector<int> cnt(N, 0);
            long next = 0;
            for (long e = 0; e < E; ++e)
                for (int i = 0; i < npe; ++i) {
                    const int n = in.connectivity[(size_t)o.eh[e] * npe + i];
                    if (key[n] < 0) key[n] = next++;
                    sum[n] += e;
                    cnt[n]++;
                }
            std::vector<std::pair<double, long>> k2(N);
            for (long n = 0; n < N; ++n) {
                const long first = key[n] < 0 ? next + n : key[n];      // unused nodes last
                k2[n] = {mode == 2 || !cnt[n] ? (double)first : (double)sum[n] / cnt[n], first};
            }
            for (long n = 0; n < N; ++n) o.nh[n] = (int)n;
            std::sort(o.nh.begin(), o.nh.end(), [&](int a, int b) { return k2[a] < k2[b]; });
        }
    }
    std::vector<int> ninv(N), einv(E);
    for (long i = 0; i < N; ++i) ninv[o.nh[i]] = (int)i;
    for (long i = 0; i < E; ++i) einv[o.eh[i]] = (int)i;
    o.conn.resize((size_t)E * npe);
    for (long e = 0; e < E; ++e)
        for (int i = 0; i < npe; ++i) o.conn[e * npe + i] = ninv[in.connectivity[(size_t)o.eh[e] * npe + i]];
    o.coord.resize((size_t)N * nd);
    o.bcflag.resize(N);
    for (long n = 0; n < N; ++n) {
        for (int d = 0; d < nd; ++d) o.coord[n * nd + d] = in.coord[(size_t)o.nh[n] * nd + d];
        o.bcflag[n] = in.bcflag[o.nh[n]];
    }
    o.felem.resize(std::max(in.nbfacet, 0));
    for (int k = 0; k < in.nbfacet; ++k) o.felem[k] = einv[in.bfacet_elem[k]];
    if (in.elemmarkers) {
        o.markers.resize((size_t)E * nmat);
        for (long e = 0; e < E; ++e)
            for (int j = 0; j < nmat; ++j) o.markers[e * nmat + j] = in.elemmarkers[(size_t)o.eh[e] * nmat + j];
    }
    o.m = in;
    o.m.connectivity = o.conn.data();
    o.m.coord = o.coord.data();
    o.m.bcflag = o.bcflag.data();
    o.m.bfacet_elem = in.nbfacet > 0 ? o.felem.data() : in.bfacet_elem;
    o.m.elemmarkers = in.elemmarkers ? o.markers.data() : nullptr;
}

// One context: a single model (g == nullptr) or a batched group whose models
// are concatenated into the mesh m, laid out as g describes.  m is in device
// order; eh / nh map its elements / nodes to (model-local) host indices, -1
// on a group's padding.
int create_impl(des_ctx *ctx, const des_params *p, const des_mesh *m, const GroupLayout *g,
                const std::vector<int> &eh, const std::vector<int> &nh)
{
    ctx->P = *p;
    const int nd = ctx->nd = p->ndims;
    const int npe = ctx->npe = nd + 1;
    ctx->nstr = nd == 3 ? 6 : 3;
    const int N = ctx->N = m->nnode, E = ctx->E = m->nelem;
    ctx->nbf = m->nbfacet;
    const int nmodels = g ? (int)g->P.size() : 0;
    ctx->nmodels = nmodels;
    HIPCHK(hipGetDevice(&ctx->device));
    HIPCHK(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking));
    DevState &s = ctx->s;
    s.N = N;
    s.E = E;
    s.nbf = m->nbfacet;
    s.mpe = m->markers_per_element;
    s.nmodels = nmodels;
    int rc;
#define ALLOC(ptr, n) do { if ((rc = dalloc_n(ctx, &(ptr), (size_t)(n)))) return rc; } while (0)
#define H2D(dst, src, n) HIPCHK(hipMemcpyAsync((void *)(dst), (src), sizeof(*(src)) * (size_t)(n), hipMemcpyHostToDevice, ctx->stream))
    des_params *dP;
    ALLOC(dP, nmodels ? nmodels : 1);
    if (nmodels) H2D(dP, g->P.data(), nmodels);
    else H2D(dP, p, 1);
    s.P = dP;
    s.Pb = dP;
    int *conn;
    ALLOC(conn, (size_t)E * npe);
    H2D(conn, m->connectivity, (size_t)E * npe);
    s.conn = conn;
    unsigned *bcflag;
    ALLOC(bcflag, N);
    H2D(bcflag, m->bcflag, N);
    s.bcflag = bcflag;
    if (m->elemmarkers) {
        int *em;
        ALLOC(em, (size_t)E * p->nmat);
        H2D(em, m->elemmarkers, (size_t)E * p->nmat);
        s.elemmarkers = em;
    }
    // live elements: a batched group's padding between models is not
    std::vector<char> live(E, 1);
    if (nmodels)
        for (int k = 0; k < nmodels; ++k)
            for (int e = g->e_off[k] + g->e_cnt[k]; e < (k + 1 < nmodels ? g->e_off[k + 1] : E); ++e) live[e] = 0;
    // create_support (mesh.cxx:1707-1721) as a CSR, ascending elements per node
    std::vector<int> off(N + 1, 0);
    for (int e = 0; e < E; ++e)
        if (live[e])
            for (int i = 0; i < npe; ++i) off[m->connectivity[(size_t)e * npe + i] + 1]++;
    for (int n = 0; n < N; ++n) off[n + 1] += off[n];
    std::vector<int> ent(std::max(off[N], 1));
    {
        // elements visited in ascending host index, so that every node lists
        // its elements in ascending host order, the reference's summation
        // order (the elements of a node all belong to one model of a group)
        int hmax = 0;
        for (int e = 0; e < E; ++e)
            if (live[e]) hmax = std::max(hmax, eh[e] + 1);
        std::vector<int> hoff(hmax + 1, 0);
        for (int e = 0; e < E; ++e)
            if (live[e]) hoff[eh[e] + 1]++;
        for (int h = 0; h < hmax; ++h) hoff[h + 1] += hoff[h];
        const int nlive = hoff[hmax];
        std::vector<int> byh(std::max(nlive, 1));
        for (int e = 0; e < E; ++e)
            if (live[e]) byh[hoff[eh[e]]++] = e;
        std::vector<int> fill(off.begin(), off.end() - 1);
        for (int k = 0; k < nlive; ++k) {
            const int e = byh[k];
            for (int i = 0; i < npe; ++i) ent[fill[m->connectivity[(size_t)e * npe + i]]++] = e * npe + i;
        }
    }
    ctx->eh = eh;
    ctx->nh = nh;
    {
        int *d_eh, *d_nh;
        ALLOC(d_eh, E);
        ALLOC(d_nh, N);
        H2D(d_eh, eh.data(), E);
        H2D(d_nh, nh.data(), N);
        s.ehost = d_eh;
        s.nhost = d_nh;
    }
    int *d_off, *d_ent;
    ALLOC(d_off, N + 1);
    ALLOC(d_ent, ent.size());
    H2D(d_off, off.data(), N + 1);
    H2D(d_ent, ent.data(), ent.size());
    s.csr_off = d_off;
    s.csr_ent = d_ent;
    {
        // per element block: its distinct nodes, ascending, and each element's
        // nodes as positions in that list (blocks above STAGE_MAX: count -1)
        const int neb = nblocks(E);
        std::vector<int2> stage(neb);
        std::vector<int> snode, u, mark(N, -1), pos(N, 0);
        std::vector<unsigned> lidx(E, 0);
        snode.reserve((size_t)E);
        for (int b = 0; b < neb; ++b) {
            const int e0 = b * BLOCK, e1 = std::min(E, e0 + BLOCK);
            u.clear();
            for (int e = e0; e < e1; ++e)
                if (live[e])
                    for (int i = 0; i < npe; ++i) {
                        const int nd_ = m->connectivity[(size_t)e * npe + i];
                        if (mark[nd_] != b) {
                            mark[nd_] = b;
                            u.push_back(nd_);
                        }
                    }
            if ((int)u.size() > STAGE_MAX) {
                stage[b] = make_int2((int)snode.size(), -1);
                continue;
            }
            std::sort(u.begin(), u.end());
            for (int k = 0; k < (int)u.size(); ++k) pos[u[k]] = k;
            stage[b] = make_int2((int)snode.size(), (int)u.size());
            for (int e = e0; e < e1; ++e)
                if (live[e])
                    for (int i = 0; i < npe; ++i)
                        lidx[e] |= (unsigned)pos[m->connectivity[(size_t)e * npe + i]] << (8 * i);
            snode.insert(snode.end(), u.begin(), u.end());
        }
        if (snode.empty()) snode.push_back(0);
        int2 *d_stage;
        int *d_snode;
        unsigned *d_lidx;
        ALLOC(d_stage, neb);
        ALLOC(d_snode, snode.size());
        ALLOC(d_lidx, E);
        H2D(d_stage, stage.data(), neb);
        H2D(d_snode, snode.data(), snode.size());
        H2D(d_lidx, lidx.data(), E);
        HIPCHK(hipStreamSynchronize(ctx->stream));
        s.stage = d_stage;
        s.stage_node = d_snode;
        s.stage_lidx = d_lidx;
    }
    {   // CSR position of every entry, for the node-major force contributions
        std::vector<int> slot((size_t)E * npe, 0);
        for (int q = 0; q < off[N]; ++q) slot[ent[q]] = q;
        int *d_slot;
        ALLOC(d_slot, slot.size());
        H2D(d_slot, slot.data(), slot.size());
        s.slot = d_slot;
        // and each entry's position in its node's list, one byte per entry,
        // when no list is longer than 256 (4 bytes per element instead of 16)
        int maxdeg = 0;
        for (int n = 0; n < N; ++n) maxdeg = std::max(maxdeg, off[n + 1] - off[n]);
        if (maxdeg <= 256 && E > 0) {
            std::vector<unsigned> rk((size_t)E, 0u);
            for (int n = 0; n < N; ++n)
                for (int q = off[n]; q < off[n + 1]; ++q)
                    rk[ent[q] / npe] |= (unsigned)(q - off[n]) << (8 * (ent[q] % npe));
            unsigned *d_rk;
            ALLOC(d_rk, rk.size());
            H2D(d_rk, rk.data(), rk.size());
            s.srank = d_rk;
            ctx->srank = d_rk;
        }
        ALLOC(s.fwork, (size_t)std::max(off[N], 1) * nd);
        ALLOC(s.frec, (size_t)std::max(E, 1) * 2 * (nd == 3 ? 8 : 4));
    }
    // boundary facets; the loaded ones get a per-node list in reference order
    // (boundary by boundary, facets ascending)
    if (m->nbfacet > 0) {
        int *b, *e, *f;
        ALLOC(b, m->nbfacet);
        ALLOC(e, m->nbfacet);
        ALLOC(f, m->nbfacet);
        H2D(b, m->bfacet_bdry, m->nbfacet);
        H2D(e, m->bfacet_elem, m->nbfacet);
        H2D(f, m->bfacet_face, m->nbfacet);
        s.bf_bdry = b;
        s.bf_elem = e;
        s.bf_face = f;
        static const int NOF3[4][3] = {{1, 2, 3}, {0, 3, 2}, {0, 1, 3}, {0, 2, 1}};
        static const int NOF2[3][2] = {{1, 2}, {2, 0}, {0, 1}};
        std::vector<int> act, act_model;
        for (int i = 0; i < DES_NBDRY; ++i)
            for (int k = 0; k < m->nbfacet; ++k) {
                if (m->bfacet_bdry[k] != i) continue;
                const int mk = nmodels ? g->facet_model[k] : 0;
                if (!facet_loaded(nmodels ? g->P[mk] : *p, i)) continue;
                act.push_back(k);
                act_model.push_back(mk);
            }
        if (!act.empty()) {
            std::vector<int> foff(N + 1, 0);
            auto facet_node = [&](int k, int j) {
                int el = m->bfacet_elem[k], fc = m->bfacet_face[k];
                int loc = nd == 3 ? NOF3[fc][j] : NOF2[fc][j];
                return m->connectivity[(size_t)el * npe + loc];
            };
            for (int k : act)
                for (int j = 0; j < nd; ++j) foff[facet_node(k, j) + 1]++;
            for (int n = 0; n < N; ++n) foff[n + 1] += foff[n];
            std::vector<int> fent(foff[N]), fill(foff.begin(), foff.end() - 1);
            for (int k : act)
                for (int j = 0; j < nd; ++j) fent[fill[facet_node(k, j)]++] = k;
            int *d_foff, *d_fent, *d_act;
            ALLOC(d_foff, N + 1);
            ALLOC(d_fent, fent.size());
            ALLOC(d_act, act.size());
            H2D(d_foff, foff.data(), N + 1);
            H2D(d_fent, fent.data(), fent.size());
            H2D(d_act, act.data(), act.size());
            s.ff_off = d_foff;
            s.ff_ent = d_fent;
            ctx->active_facets = d_act;
            ctx->nactive = (int)act.size();
            ALLOC(s.ftrac, (size_t)m->nbfacet * nd);
            if (nmodels) {
                int *fm;
                ALLOC(fm, act_model.size());
                H2D(fm, act_model.data(), act_model.size());
                s.fmodel = fm;
            }
        }
    }
    double *bn, *ev;
    int *evv;
    const int ngeo = nmodels ? nmodels : 1;
    ALLOC(bn, (size_t)ngeo * DES_NBDRY * nd);
    ALLOC(ev, (size_t)ngeo * DES_NBDRY * DES_NBDRY * nd);
    ALLOC(evv, (size_t)ngeo * DES_NBDRY * DES_NBDRY);
    if (nmodels) {
        H2D(bn, g->bn.data(), g->bn.size());
        H2D(ev, g->ev.data(), g->ev.size());
        H2D(evv, g->evv.data(), g->evv.size());
    } else {
        if (m->bnormals) H2D(bn, m->bnormals, DES_NBDRY * nd);
        if (m->edge_vectors) H2D(ev, m->edge_vectors, DES_NBDRY * DES_NBDRY * nd);
        if (m->edge_valid) H2D(evv, m->edge_valid, DES_NBDRY * DES_NBDRY);
    }
    s.bnormals = s.bnb = bn;
    s.edge_vectors = s.evb = ev;
    s.edge_valid = s.evvb = evv;
    // fields (zero, as allocate_variables leaves them)
    ALLOC(s.coord, (size_t)N * nd);
    ALLOC(s.coord0, (size_t)N * nd);
    ALLOC(s.vel, (size_t)N * nd);
    ALLOC(s.force, (size_t)N * nd);
    ALLOC(s.T, N);
    ALLOC(s.mass, N);
    ALLOC(s.tmass, N);
    ALLOC(s.volume_n, N);
    ALLOC(s.ntmp, N);
    ALLOC(ctx->vol_buf[0], E);
    ALLOC(ctx->vol_buf[1], E);
    s.volume = ctx->vol_buf[0];
    s.volume_old = ctx->vol_buf[1];
    ALLOC(s.edvoldt, E);
    ALLOC(s.plstrain, E);
    ALLOC(s.delta_plstrain, E);
    ALLOC(s.dpressure, E);
    ALLOC(s.stressyy, E);
    ALLOC(s.strain_rate, (size_t)E * ctx->nstr);
    ALLOC(s.strain, (size_t)E * ctx->nstr);
    ALLOC(s.stress, (size_t)E * ctx->nstr);
    ALLOC(s.shp, (size_t)E * npe * nd);
    ALLOC(s.ework, (size_t)E * npe * nd);
    ALLOC(s.escal, (size_t)E * 4);
    ALLOC(s.erec, (size_t)E * 8);
    ALLOC(s.sc, ngeo);
    s.scb = s.sc;
    const int nred = std::max(RED_BLOCKS, nblocks(E));
    ALLOC(s.red, (size_t)nred * 3);
    ALLOC(s.redi, (size_t)nred * 2);
    {
        // the device index of each model's host element 0 (compute_dt's
        // bulkm(0) and rho(0), geometry.cxx:297)
        std::vector<int> e0(std::max(nmodels, 1), 0);
        for (int k = 0; k < (int)e0.size(); ++k) {
            const int lo = nmodels ? g->e_off[k] : 0, hi = nmodels ? lo + g->e_cnt[k] : E;
            for (int e = lo; e < hi; ++e)
                if (eh[e] == 0) { e0[k] = e; break; }
        }
        int *d_e0;
        ALLOC(d_e0, e0.size());
        H2D(d_e0, e0.data(), e0.size());
        s.e0dev = d_e0;
    }
    if (nmodels) {
        // block -> model tables; a block's real elements (nodes) end at lim
        const int neb = nblocks(E), nnb = nblocks(N);
        std::vector<int> ebm(neb, 0), ebl(neb, 0), nbm(nnb, 0), nbl(nnb, 0), mblk(nmodels + 1, neb);
        for (int k = 0; k < nmodels; ++k) {
            const int e0 = g->e_off[k], e1 = e0 + g->e_cnt[k], n0 = g->n_off[k], n1 = n0 + g->n_cnt[k];
            mblk[k] = e0 / BLOCK;
            for (int b = e0 / BLOCK; b < nblocks(e1); ++b) { ebm[b] = k; ebl[b] = e1; }
            for (int b = n0 / BLOCK; b < nblocks(n1); ++b) { nbm[b] = k; nbl[b] = n1; }
        }
        int *d_ebm, *d_ebl, *d_nbm, *d_nbl, *d_mblk;
        ALLOC(d_ebm, neb);
        ALLOC(d_ebl, neb);
        ALLOC(d_nbm, nnb);
        ALLOC(d_nbl, nnb);
        ALLOC(d_mblk, nmodels + 1);
        H2D(d_ebm, ebm.data(), neb);
        H2D(d_ebl, ebl.data(), neb);
        H2D(d_nbm, nbm.data(), nnb);
        H2D(d_nbl, nbl.data(), nnb);
        H2D(d_mblk, mblk.data(), nmodels + 1);
        s.eblk_model = d_ebm;
        s.eblk_lim = d_ebl;
        s.nblk_model = d_nbm;
        s.nblk_lim = d_nbl;
        s.mblk_off = d_mblk;
        ctx->hP = g->P;
        ctx->e_off = g->e_off;
        ctx->n_off = g->n_off;
        ctx->e_cnt = g->e_cnt;
        ctx->n_cnt = g->n_cnt;
        ctx->nbf_cnt = g->nbf_cnt;
        ctx->has_visc.assign(nmodels, 0);
        ctx->has_coh.assign(nmodels, 0);
    }
#undef ALLOC
#undef H2D
    // initial coordinates, also coord0 (dynearthsol.cxx:82-84)
    if ((rc = upload_soa(ctx, s.coord, m->coord, N, nd))) return rc;
    if ((rc = upload_soa(ctx, s.coord0, m->coord, N, nd))) return rc;
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return DES_OK;
}

// every index the kernels will follow, checked on the host
int validate_mesh(const des_params *p, const des_mesh *m)
{
    if (p->ndims != 2 && p->ndims != 3) return DES_ERR_INPUT;
    if (p->nmat < 1 || p->nmat > DES_MAX_MAT) return DES_ERR_INPUT;
    if (m->nnode <= 0 || m->nelem <= 0 || !m->connectivity || !m->coord || !m->bcflag) return DES_ERR_MESH;
    const int nd = p->ndims, npe = nd + 1;
    if ((long)m->nelem * npe >= (1L << 31)) return DES_ERR_MESH;      // CSR entries e * npe + i are 32-bit
    for (long k = 0; k < (long)m->nelem * npe; ++k)
        if (m->connectivity[k] < 0 || m->connectivity[k] >= m->nnode) return DES_ERR_MESH;
    if (m->nbfacet > 0 && (!m->bfacet_bdry || !m->bfacet_elem || !m->bfacet_face)) return DES_ERR_MESH;
    for (int k = 0; k < m->nbfacet; ++k)
        if (m->bfacet_bdry[k] < 0 || m->bfacet_bdry[k] >= DES_NBDRY || m->bfacet_elem[k] < 0 ||
            m->bfacet_elem[k] >= m->nelem || m->bfacet_face[k] < 0 || m->bfacet_face[k] > nd)
            return DES_ERR_MESH;
    if (!m->elemmarkers && p->nmat != 1) return DES_ERR_INPUT;
    return DES_OK;
}

}  // namespace

extern "C" {

const char *des_last_error(const des_ctx *ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int des_create(const des_params *p, const des_mesh *m, des_ctx **out)
{
    if (!p || !m || !out) return DES_ERR_INPUT;
    *out = nullptr;
    int rc = validate_mesh(p, m);
    if (rc) return rc;
    des_ctx *ctx = new des_ctx();
    *out = ctx;
    OrderedMesh om;
    order_mesh(*m, p->ndims, p->nmat, g_mesh_order, om);
    if ((rc = create_impl(ctx, p, &om.m, nullptr, om.eh, om.nh))) return rc;
    ctx->h_conn.assign(m->connectivity, m->connectivity + (size_t)m->nelem * (p->ndims + 1));
    ctx->h_bcflag.assign(m->bcflag, m->bcflag + m->nnode);
    return DES_OK;
}

void des_set_mesh_order(int mode) { g_mesh_order = (mode >= 0 && mode <= 3) ? mode : 1; }
void des_set_sweep_order(int mode) { g_sweep = mode ? 1 : 0; }
void des_set_force_layout(int mode) { g_force_layout = mode == 1 || mode == 2 ? mode : 0; }

void des_destroy(des_ctx *ctx)
{
    if (!ctx) return;
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    for (void *p : ctx->allocs) (void)hipFree(p);
    for (void *p : ctx->adj_allocs) (void)hipFree(p);
    if (ctx->xfer) (void)hipFree(ctx->xfer);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
}

}  // extern "C"

namespace {

// per-element optional field, allocated on first upload
int field_opt(des_ctx *ctx, double **slot, const double *src, long n, long off, bool dev = false)
{
    if (!src) return DES_OK;
    if (!*slot) {
        int rc = dalloc_n(ctx, slot, (size_t)ctx->E);
        if (rc) return rc;
    }
    return upload_soa(ctx, *slot, src, n, 1, ctx->E, off, ctx->eh.data() + off, dev);
}

// The fields of des_fields between AoS arrays in host row order and the
// device SoA state: the whole context (model < 0) or batched model `model` of
// a group.  dev: the des_fields pointers are device memory.
int xfer_fields(des_ctx *ctx, des_fields *f, bool up, int model, bool dev = false)
{
    int rc = flush_deferred(ctx);
    if (rc) return rc;
    DevState &s = ctx->s;
    const long NS = ctx->N, ES = ctx->E;
    const long N = model < 0 ? NS : ctx->n_cnt[model], E = model < 0 ? ES : ctx->e_cnt[model];
    const long no = model < 0 ? 0 : ctx->n_off[model], eo = model < 0 ? 0 : ctx->e_off[model];
    const int nd = ctx->nd, npe = ctx->npe, nstr = ctx->nstr;
    struct { double *dev; double *host; long n; int k; bool elem; } tab[] = {
        {s.coord, f->coord, N, nd, false}, {s.coord0, f->coord0, N, nd, false},
        {s.vel, f->vel, N, nd, false}, {s.force, f->force, N, nd, false},
        {s.T, f->temperature, N, 1, false}, {s.mass, f->mass, N, 1, false},
        {s.tmass, f->tmass, N, 1, false}, {s.volume_n, f->volume_n, N, 1, false},
        {s.ntmp, f->ntmp, N, 1, false},
        {s.volume, f->volume, E, 1, true}, {s.volume_old, f->volume_old, E, 1, true},
        {s.edvoldt, f->edvoldt, E, 1, true}, {s.plstrain, f->plstrain, E, 1, true},
        {s.delta_plstrain, f->delta_plstrain, E, 1, true}, {s.dpressure, f->dpressure, E, 1, true},
        {s.stressyy, f->stressyy, E, 1, true}, {s.strain_rate, f->strain_rate, E, nstr, true},
        {s.strain, f->strain, E, nstr, true}, {s.stress, f->stress, E, nstr, true},
    };
    for (auto &t : tab) {
        const long stride = t.elem ? ES : NS, off = t.elem ? eo : no;
        const int *perm = (t.elem ? ctx->eh.data() : ctx->nh.data()) + off;
        rc = up ? upload_soa(ctx, t.dev, t.host, t.n, t.k, stride, off, perm, dev)
                : download_soa(ctx, t.host, t.dev, t.n, t.k, stride, off, perm, dev);
        if (rc) return rc;
    }
    double *shp_host[3] = {f->shpdx, nd == 3 ? f->shpdy : f->shpdz, nd == 3 ? f->shpdz : nullptr};
    for (int d = 0; d < nd; ++d) {
        double *shpd = s.shp + (size_t)d * npe * ES;
        const int *perm = ctx->eh.data() + eo;
        rc = up ? upload_soa(ctx, shpd, shp_host[d], E, npe, ES, eo, perm, dev)
                : download_soa(ctx, shp_host[d], shpd, E, npe, ES, eo, perm, dev);
        if (rc) return rc;
    }
    if (up) {
        double *ev = const_cast<double *>(s.elem_visc), *ec = const_cast<double *>(s.elem_cohesion);
        if ((rc = field_opt(ctx, &ev, f->elem_visc, E, eo, dev))) return rc;
        if ((rc = field_opt(ctx, &ec, f->elem_cohesion, E, eo, dev))) return rc;
        s.elem_visc = ev;
        s.elem_cohesion = ec;
        if (model >= 0) {
            if (f->elem_visc) ctx->has_visc[model] = 1;
            if (f->elem_cohesion) ctx->has_coh[model] = 1;
        }
        HIPCHK(hipStreamSynchronize(ctx->stream));
    }
    return DES_OK;
}

}  // namespace

extern "C" {

int des_upload_fields(des_ctx *ctx, const des_fields *f)
{
    if (!ctx || !f) return DES_ERR_INPUT;
    return xfer_fields(ctx, const_cast<des_fields *>(f), true, -1);
}

int des_upload_fields_device(des_ctx *ctx, const des_fields *f)
{
    if (!ctx || !f) return DES_ERR_INPUT;
    if (ctx->nmodels) return fail(ctx, DES_ERR_INPUT, "device-pointer transfers are for single models");
    des_fields c = *f;
    return xfer_fields(ctx, &c, true, -1, true);
}

int des_download_fields_device(des_ctx *ctx, des_fields *f)
{
    if (!ctx || !f) return DES_ERR_INPUT;
    if (ctx->nmodels) return fail(ctx, DES_ERR_INPUT, "device-pointer transfers are for single models");
    return xfer_fields(ctx, f, false, -1, true);
}

int des_download_fields(des_ctx *ctx, des_fields *f)
{
    if (!ctx || !f) return DES_ERR_INPUT;
    return xfer_fields(ctx, f, false, -1);
}

// One output frame in the reference's binary format (binaryio.cxx:16-36: a
// 4096-byte ASCII header "# DynEarthSol ndims=N revision=3" followed by
// "name<TAB>offset" lines, then the raw arrays) with the field set of
// Output::write (output.cxx:93-179, instantaneous fields), and one line of
// the .info table (output.cxx:43-76).
}  // extern "C"

namespace {

// Output::write (output.cxx:79-208); averaged: with the averaged fields of
// the current interval (average_interval > 0), as a regular frame writes them
int write_frame_impl(des_ctx *ctx, const char *modelname, int frame, double run_time, bool averaged)
{
    if (!ctx || !modelname || frame < 0) return DES_ERR_INPUT;
    if (averaged && ctx->avg_interval <= 0) return fail(ctx, DES_ERR_STATE, "averaged fields are not enabled");
    int rc = flush_deferred(ctx);
    if (rc) return rc;
    const long N = ctx->N, E = ctx->E;
    const int nd = ctx->nd, npe = ctx->npe, nstr = ctx->nstr;
    std::vector<double> coord(N * nd), vel(N * nd), force(N * nd), coord0(N * nd), T(N);
    std::vector<double> pl(E), dpl(E), sr(E * nstr), strain(E * nstr), stress(E * nstr);
    std::vector<double> density(E), quality(E), viscosity(E), material(E, 0.0);
    des_fields f{};
    f.coord = coord.data(); f.vel = vel.data(); f.force = force.data(); f.coord0 = coord0.data();
    f.temperature = T.data(); f.plstrain = pl.data(); f.delta_plstrain = dpl.data();
    f.strain_rate = sr.data(); f.strain = strain.data(); f.stress = stress.data();
    if ((rc = des_download_fields(ctx, &f))) return rc;
    double *d_diag;
    HIPCHK(hipMallocAsync((void **)&d_diag, sizeof(double) * 3 * E, ctx->stream));
    LAUNCH_ND(k_elem_diagnostics, nblocks(E), ctx->s, d_diag, d_diag + E, d_diag + 2 * E);
    {
        const int *perm = ctx->eh.data();
        if ((rc = download_soa(ctx, density.data(), d_diag, E, 1, E, 0, perm))) return rc;
        if ((rc = download_soa(ctx, quality.data(), d_diag + E, E, 1, E, 0, perm))) return rc;
        if ((rc = download_soa(ctx, viscosity.data(), d_diag + 2 * E, E, 1, E, 0, perm))) return rc;
    }
    HIPCHK(hipFreeAsync(d_diag, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    if (ctx->s.elemmarkers) {       // most abundant material (output.cxx:168-173)
        std::vector<int> em((size_t)E * ctx->P.nmat);
        HIPCHK(hipMemcpy(em.data(), ctx->s.elemmarkers, sizeof(int) * em.size(), hipMemcpyDeviceToHost));
        for (long e = 0; e < E; ++e) {
            const int *a = em.data() + e * ctx->P.nmat;
            material[ctx->eh[e]] = (double)(std::max_element(a, a + ctx->P.nmat) - a);
        }
    }
    DevScalars sc{};
    if ((rc = read_scalars(ctx, sc))) return rc;
    // the averaged variants (output.cxx:83-86, 97-105, 113-119, 121-131, 136-143)
    double info_dt = sc.dt;
    std::vector<double> vel_avg, stress_avg;
    if (averaged) {
        const double span = sc.time - sc.time0;
        const double inv_dt = 1.0 / span;
        info_dt = span / ctx->avg_interval;
        std::vector<double> c0(N * nd), s0(E * nstr), davg(E);
        stress_avg.resize(E * nstr);
        const int *np = ctx->nh.data(), *ep = ctx->eh.data();
        if ((rc = download_soa(ctx, c0.data(), ctx->s.avg_coord0, N, nd, N, 0, np))) return rc;
        if ((rc = download_soa(ctx, s0.data(), ctx->s.avg_strain0, E, nstr, E, 0, ep))) return rc;
        if ((rc = download_soa(ctx, stress_avg.data(), ctx->s.avg_stress, E, nstr, E, 0, ep))) return rc;
        if ((rc = download_soa(ctx, davg.data(), ctx->s.avg_dpl, E, 1, E, 0, ep))) return rc;
        vel_avg.resize(N * nd);
        for (long i = 0; i < N * nd; ++i) vel_avg[i] = (coord[i] - c0[i]) * inv_dt;
        for (long i = 0; i < E; ++i) dpl[i] = davg[i] * inv_dt;
        for (long i = 0; i < E * nstr; ++i) sr[i] = (strain[i] - s0[i]) * inv_dt;
        const double tmp = 1.0 / (ctx->avg_interval + 1);
        for (double &x : stress_avg) x *= tmp;
    }

    char fname[1024];
    std::snprintf(fname, sizeof fname, "%s.save.%06d", modelname, frame);
    std::FILE *fp = std::fopen(fname, "wb");
    if (!fp) return fail(ctx, DES_ERR_IO, std::string("cannot open ") + fname);
    const size_t headerlen = 4096;
    std::string header = "# DynEarthSol ndims=" + std::to_string(nd) + " revision=3\n";
    long eof = (long)headerlen;
    std::fseek(fp, eof, SEEK_SET);
    bool ok = true;
    auto put = [&](const char *name, const void *data, size_t bytes) {
        char line[256];
        std::snprintf(line, sizeof line, "%s\t%ld\n", name, eof);
        header += line;
        if (std::fwrite(data, 1, bytes, fp) != bytes) ok = false;
        eof += (long)bytes;
    };
    put("coordinate", coord.data(), sizeof(double) * coord.size());
    put("connectivity", ctx->h_conn.data(), sizeof(int) * ctx->h_conn.size());
    put("velocity", vel.data(), sizeof(double) * vel.size());
    if (averaged) put("velocity averaged", vel_avg.data(), sizeof(double) * vel_avg.size());
    put("temperature", T.data(), sizeof(double) * T.size());
    put("plastic strain", pl.data(), sizeof(double) * pl.size());
    put("plastic strain-rate", dpl.data(), sizeof(double) * dpl.size());
    put("strain-rate", sr.data(), sizeof(double) * sr.size());
    put("strain", strain.data(), sizeof(double) * strain.size());
    put("stress", stress.data(), sizeof(double) * stress.size());
    if (averaged) put("stress averaged", stress_avg.data(), sizeof(double) * stress_avg.size());
    put("density", density.data(), sizeof(double) * E);
    put("mesh quality", quality.data(), sizeof(double) * E);
    put("viscosity", viscosity.data(), sizeof(double) * E);
    put("material", material.data(), sizeof(double) * E);
    put("force", force.data(), sizeof(double) * force.size());
    put("coord0", coord0.data(), sizeof(double) * coord0.size());
    put("bcflag", ctx->h_bcflag.data(), sizeof(unsigned) * ctx->h_bcflag.size());
    if (header.size() >= headerlen) ok = false;
    header.resize(headerlen, '\0');
    std::fseek(fp, 0, SEEK_SET);
    if (std::fwrite(header.data(), 1, headerlen, fp) != headerlen) ok = false;
    std::fclose(fp);
    if (!ok) return fail(ctx, DES_ERR_IO, std::string("failed writing ") + fname);

    std::string info = std::string(modelname) + ".info";
    std::FILE *fi = std::fopen(info.c_str(), frame == 0 ? "w" : "a");
    if (!fi) return fail(ctx, DES_ERR_IO, "cannot open " + info);
    std::fprintf(fi, "%6d\t%10d\t%12.6e\t%12.4e\t%12.6e\t%8d\t%8d\t%8d\n", frame, sc.steps, sc.time, info_dt,
                 run_time, ctx->N, ctx->E, ctx->nbf);
    std::fclose(fi);
    for (double x : coord)      // output.cxx:194-208
        if (!std::isfinite(x)) return fail(ctx, DES_ERR_RUNTIME, "coordinate becomes NaN or Infinity");
    return DES_OK;
}

// Output::write_checkpoint (output.cxx:254-276) in the format of binaryio.cxx
int write_checkpoint_impl(des_ctx *ctx, const char *modelname, int frame, int nseg, const int *segment,
                          const int *segflag, int nextra, const des_named_array *extra)
{
    if (!ctx || !modelname || frame < 0 || nseg < 0 || (nseg && (!segment || !segflag)) || nextra < 0 ||
        (nextra && !extra) || ctx->nmodels)
        return DES_ERR_INPUT;
    int rc = flush_deferred(ctx);
    if (rc) return rc;
    const long E = ctx->E;
    std::vector<double> vold(E), syy;
    des_fields f{};
    f.volume_old = vold.data();
    if (ctx->P.is_plane_strain) {
        syy.resize(E);
        f.stressyy = syy.data();
    }
    if ((rc = des_download_fields(ctx, &f))) return rc;
    DevScalars sc{};
    if ((rc = read_scalars(ctx, sc))) return rc;
    const double tcp[2] = {sc.time, sc.compensation_pressure};

    char fname[1024];
    std::snprintf(fname, sizeof fname, "%s.chkpt.%06d", modelname, frame);
    std::FILE *fp = std::fopen(fname, "wb");
    if (!fp) return fail(ctx, DES_ERR_IO, std::string("cannot open ") + fname);
    const size_t headerlen = 4096;
    std::string header = "# DynEarthSol ndims=" + std::to_string(ctx->nd) + " revision=3\n";
    long eof = (long)headerlen;
    std::fseek(fp, eof, SEEK_SET);
    bool ok = true;
    auto put = [&](const char *name, const void *data, size_t bytes) {
        char line[256];
        std::snprintf(line, sizeof line, "%s\t%ld\n", name, eof);
        header += line;
        if (bytes && std::fwrite(data, 1, bytes, fp) != bytes) ok = false;
        eof += (long)bytes;
    };
    put("time compensation_pressure", tcp, sizeof tcp);
    put("segment", segment, sizeof(int) * (size_t)nseg * ctx->nd);
    put("segflag", segflag, sizeof(int) * (size_t)nseg);
    put("volume_old", vold.data(), sizeof(double) * vold.size());
    if (ctx->P.is_plane_strain) put("stressyy", syy.data(), sizeof(double) * syy.size());
    for (int k = 0; k < nextra; ++k) {
        if (!extra[k].name || (extra[k].bytes && !extra[k].data)) ok = false;
        else put(extra[k].name, extra[k].data, extra[k].bytes);
    }
    if (header.size() >= headerlen) ok = false;
    header.resize(headerlen, '\0');
    std::fseek(fp, 0, SEEK_SET);
    if (std::fwrite(header.data(), 1, headerlen, fp) != headerlen) ok = false;
    std::fclose(fp);
    if (!ok) return fail(ctx, DES_ERR_IO, std::string("failed writing ") + fname);
    return DES_OK;
}

// Output::average_fields for the step just completed (ctx->steps is var.steps)
int average_fields(des_ctx *ctx)
{
    const int snapshot = (ctx->steps % ctx->avg_interval == 1) ? 1 : 0;
    if (snapshot)
        HIPCHK(hipMemcpyAsync(ctx->s.avg_coord0, ctx->s.coord, sizeof(double) * ctx->N * ctx->nd,
                              hipMemcpyDeviceToDevice, ctx->stream));
    LAUNCH_ND(k_average_fields, nblocks(ctx->E), ctx->s, snapshot);
    return DES_OK;
}

}  // namespace

extern "C" {

int des_write_frame(des_ctx *ctx, const char *modelname, int frame, double run_time)
{
    return write_frame_impl(ctx, modelname, frame, run_time, false);
}

int des_write_checkpoint(des_ctx *ctx, const char *modelname, int frame, int nseg, const int *segment,
                         const int *segflag, int nextra, const des_named_array *extra)
{
    return write_checkpoint_impl(ctx, modelname, frame, nseg, segment, segflag, nextra, extra);
}

int des_write_frame_averaged(des_ctx *ctx, const char *modelname, int frame, double run_time)
{
    return write_frame_impl(ctx, modelname, frame, run_time, true);
}

int des_set_averaging(des_ctx *ctx, int average_interval)
{
    if (!ctx || average_interval < 0) return DES_ERR_INPUT;
    if (ctx->nmodels) return fail(ctx, DES_ERR_INPUT, "averaged fields are per model, not for batched groups");
    if (average_interval > 0 && !ctx->s.avg_stress) {
        DevState &s = ctx->s;
        const size_t E = ctx->E, N = ctx->N;
        int rc;
        if ((rc = dalloc_n(ctx, &s.avg_coord0, N * ctx->nd))) return rc;
        if ((rc = dalloc_n(ctx, &s.avg_strain0, E * ctx->nstr))) return rc;
        if ((rc = dalloc_n(ctx, &s.avg_stress, E * ctx->nstr))) return rc;
        if ((rc = dalloc_n(ctx, &s.avg_dpl, E))) return rc;
    }
    ctx->avg_interval = average_interval;
    return DES_OK;
}

int des_average_fields(des_ctx *ctx)
{
    if (!ctx) return DES_ERR_INPUT;
    if (ctx->avg_interval <= 0) return fail(ctx, DES_ERR_STATE, "averaged fields are not enabled");
    return average_fields(ctx);
}

int des_upload_elemmarkers(des_ctx *ctx, const int *counts)
{
    if (!ctx || !counts) return DES_ERR_INPUT;
    int rc = flush_deferred(ctx);
    if (rc) return rc;
    const size_t n = (size_t)ctx->E * ctx->P.nmat;
    if (!ctx->s.elemmarkers) {
        int *em;
        if ((rc = dalloc_n(ctx, &em, n))) return rc;
        ctx->s.elemmarkers = em;
    }
    const int nmat = ctx->P.nmat;
    std::vector<int> dev(n);
    for (long e = 0; e < ctx->E; ++e)
        for (int j = 0; j < nmat; ++j) dev[(size_t)e * nmat + j] = counts[(size_t)ctx->eh[e] * nmat + j];
    HIPCHK(hipMemcpyAsync(const_cast<int *>(ctx->s.elemmarkers), dev.data(), sizeof(int) * n,
                          hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return DES_OK;
}

int des_set_scalars(des_ctx *ctx, double dt, double time, int steps, double compensation_pressure)
{
    if (!ctx) return DES_ERR_INPUT;
    int rc = flush_deferred(ctx);
    if (rc) return rc;
    DevScalars h{};
    if ((rc = read_scalars(ctx, h))) return rc;
    h.dt = dt;
    h.time = time;
    h.steps = steps;
    h.compensation_pressure = compensation_pressure;
    h.rot_pending = 0;
    HIPCHK(hipMemcpyAsync(ctx->s.sc, &h, sizeof h, hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    ctx->steps = steps;
    return DES_OK;
}

int des_get_scalars(des_ctx *ctx, double *dt, double *time, int *steps)
{
    if (!ctx) return DES_ERR_INPUT;
    DevScalars h{};
    int rc = read_scalars(ctx, h);
    if (rc) return rc;
    if (dt) *dt = h.dt;
    if (time) *time = h.time;
    if (steps) *steps = h.steps;
    if (h.err) return fail(ctx, h.err, "device reported an error");
    return DES_OK;
}

#define ENTRY(name, body)                        \
    int name(des_ctx *ctx)                       \
    {                                            \
        if (!ctx) return DES_ERR_INPUT;          \
        int rc = flush_deferred(ctx);            \
        if (rc) return rc;                       \
        body;                                    \
        return DES_OK;                           \
    }

ENTRY(des_compute_volume, LAUNCH_ND(k_volume, nblocks(ctx->E), ctx->s))
ENTRY(des_compute_mass, if ((rc = launch_mass(ctx))) return rc)
ENTRY(des_compute_shape_fn, LAUNCH_ND(k_shape_fn, nblocks(ctx->E), ctx->s))
ENTRY(des_update_temperature, LAUNCH_ND(k_temp_elem, nblocks(ctx->E), ctx->s);
      LAUNCH_ND(k_temp_node, nblocks(ctx->N), ctx->s))
ENTRY(des_update_strain_rate, LAUNCH_ND(k_strain_rate, nblocks(ctx->E), ctx->s))
ENTRY(des_compute_dvoldt, LAUNCH_ND(k_dvoldt_node, nblocks(ctx->N), ctx->s))
ENTRY(des_compute_edvoldt, LAUNCH_ND(k_edvoldt, nblocks(ctx->E), ctx->s))
ENTRY(des_update_stress, LAUNCH_ND(k_update_stress, nblocks(ctx->E), ctx->s))
ENTRY(des_nmd_stress, LAUNCH_ND(k_nmd_node, nblocks(ctx->N), ctx->s);
      LAUNCH_ND(k_nmd_elem, nblocks(ctx->E), ctx->s))
ENTRY(des_update_force, if ((rc = launch_force(ctx))) return rc)
ENTRY(des_update_velocity, LAUNCH_ND(k_update_velocity, nblocks(ctx->N), ctx->s))
ENTRY(des_apply_vbcs, LAUNCH_ND(k_apply_vbcs, nblocks(ctx->N), ctx->s))
ENTRY(des_update_mesh, if ((rc = launch_update_mesh(ctx))) return rc)
ENTRY(des_rotate_stress, LAUNCH_ND(k_rotate, nblocks(ctx->E), ctx->s, 0))

int des_init_derived(des_ctx *ctx)
{
    if (!ctx) return DES_ERR_INPUT;
    int rc = flush_deferred(ctx);
    if (rc) return rc;
    LAUNCH_ND(k_volume, nblocks(ctx->E), ctx->s);
    HIPCHK(hipMemcpyAsync(ctx->s.volume_old, ctx->s.volume, sizeof(double) * ctx->E, hipMemcpyDeviceToDevice, ctx->stream));
    if ((rc = launch_mass(ctx))) return rc;
    LAUNCH_ND(k_shape_fn, nblocks(ctx->E), ctx->s);
    return DES_OK;
}

int des_remesh_finish(des_ctx *ctx, int remeshing_option)
{
    int rc = des_init_derived(ctx);
    if (rc) return rc;
    if (remeshing_option == 1 || remeshing_option == 2 || remeshing_option == 11)
        LAUNCH_ND(k_reset_bottom_coord0, nblocks(ctx->N), ctx->s);
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return DES_OK;
}

int des_compute_dt(des_ctx *ctx, double *dt_out)
{
    if (!ctx) return DES_ERR_INPUT;
    int rc = flush_deferred(ctx);
    if (rc) return rc;
    if ((rc = launch_compute_dt(ctx))) return rc;
    DevScalars h{};
    if ((rc = read_scalars(ctx, h))) return rc;
    if (dt_out) *dt_out = h.dt;
    if (h.err) return fail(ctx, h.err, "compute_dt: dt <= 0");
    return DES_OK;
}

int des_worst_quality(des_ctx *ctx, double *q, int *worst_elem)
{
    if (!ctx) return DES_ERR_INPUT;
    int rc = flush_deferred(ctx);
    if (rc) return rc;
    LAUNCH_ND(k_quality_partial, RED_BLOCKS, ctx->s);
    std::vector<double> hq(RED_BLOCKS);
    std::vector<int> hi(RED_BLOCKS);
    HIPCHK(hipMemcpyAsync(hq.data(), ctx->s.red, sizeof(double) * RED_BLOCKS, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipMemcpyAsync(hi.data(), ctx->s.redi, sizeof(int) * RED_BLOCKS, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    // the reference starts from q = 1, element 0, and keeps the first strict minimum
    double best = 1;
    int bi = 0;
    for (int b = 0; b < RED_BLOCKS; ++b)
        if (hq[b] < best || (hq[b] == best && best < 1 && hi[b] < bi)) {
            best = hq[b];
            bi = hi[b];
        }
    if (q) *q = best;
    if (worst_elem) *worst_elem = bi;
    return DES_OK;
}

}  // extern "C"

// ---- worst_elem_quality without allocating
//
// The reduction's partial results -- one (quality, host element index) pair
// per block of k_quality_partial's fixed grid of RED_BLOCKS blocks -- in a
// file-scope device array instead of the context's buffers, for callers that
// cannot allocate device memory.  Each GPU holds its own copy of the array
// until the process ends; every call overwrites every slot, so nothing needs
// clearing, and calls that overlap on one GPU race for it.
constexpr int NOALLOC_SLOTS = RED_BLOCKS;
constexpr size_t NOALLOC_BYTES = NOALLOC_SLOTS * (sizeof(double) + sizeof(int));
static __device__ __align__(256) char scratch[NOALLOC_BYTES];

namespace {

constexpr int MAX_DEVICES = 64;
std::mutex g_scratch_mu;
void *g_scratch_dev[MAX_DEVICES];      // scratch's address on each device

// scratch's device address on the context's device, looked up once per
// device with that device current (the array's name is only a host-side
// stand-in for it)
int scratch_on(des_ctx *ctx, void **ptr)
{
    const int dev = ctx->device;
    if (dev < 0 || dev >= MAX_DEVICES) return fail(ctx, DES_ERR_HIP, "device index out of range");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_dev[dev]) {
        int cur = 0;
        HIPCHK(hipGetDevice(&cur));
        if (cur != dev) HIPCHK(hipSetDevice(dev));
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (cur != dev) HIPCHK(hipSetDevice(cur));
        if (e != hipSuccess || !p)
            return fail(ctx, DES_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_scratch_dev[dev] = p;
    }
    *ptr = g_scratch_dev[dev];
    return DES_OK;
}

}  // namespace

extern "C" {

int des_worst_quality_noalloc(des_ctx *ctx, double *q, int *worst_elem)
{
    if (!ctx) return DES_ERR_INPUT;
    if (ctx->nmodels)
        return fail(ctx, DES_ERR_INPUT, "des_worst_quality_noalloc: batched groups use des_batch_worst_quality");
    void *p = nullptr;
    int rc = scratch_on(ctx, &p);
    if (rc) return rc;
    if ((rc = flush_deferred(ctx))) return rc;      // kernel launches only
    DevState s = ctx->s;              // the context's state with its partials in scratch
    s.red = static_cast<double *>(p);
    s.redi = reinterpret_cast<int *>(static_cast<char *>(p) + NOALLOC_SLOTS * sizeof(double));
    LAUNCH_ND(k_quality_partial, NOALLOC_SLOTS, s);
    double hq[NOALLOC_SLOTS];
    int hi[NOALLOC_SLOTS];
    HIPCHK(hipMemcpyAsync(hq, s.red, sizeof hq, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipMemcpyAsync(hi, s.redi, sizeof hi, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    // the reference starts from q = 1, element 0, and keeps the first strict minimum
    double best = 1;
    int bi = 0;
    for (int b = 0; b < NOALLOC_SLOTS; ++b)
        if (hq[b] < best || (hq[b] == best && best < 1 && hi[b] < bi)) {
            best = hq[b];
            bi = hi[b];
        }
    if (q) *q = best;
    if (worst_elem) *worst_elem = bi;
    return DES_OK;
}

int des_bad_mesh_quality(des_ctx *ctx, int remeshing_option, double max_boundary_distortion,
                         int *reason, int *index)
{
    if (!ctx || !reason || !index) return DES_ERR_INPUT;
    int rc = flush_deferred(ctx);
    if (rc) return rc;
    const des_params &P = ctx->P;
    const double sizefactor = ctx->nd == 3 ? 0.118 : 0.433;   // remeshing.cxx:76-82
    const double smallest_vol = P.smallest_size * sizefactor * std::pow(P.resolution, ctx->nd);
    const int check_bottom = remeshing_option == 1 || remeshing_option == 2 || remeshing_option == 11;
    LAUNCH_ND(k_badmesh_partial, RED_BLOCKS, ctx->s, smallest_vol, check_bottom, -P.zlength,
              max_boundary_distortion * P.resolution);
    std::vector<int> h(2 * RED_BLOCKS);
    HIPCHK(hipMemcpyAsync(h.data(), ctx->s.redi, sizeof(int) * 2 * RED_BLOCKS, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    int fe = 0x7fffffff, fn = 0x7fffffff;
    for (int b = 0; b < RED_BLOCKS; ++b) {
        fe = std::min(fe, h[2 * b]);
        fn = std::min(fn, h[2 * b + 1]);
    }
    *reason = 0;
    *index = 0;
    if (fe != 0x7fffffff) {
        *reason = 3;
        *index = fe;
        return DES_OK;
    }
    if (fn != 0x7fffffff) {
        *reason = 2;
        *index = fn;
        return DES_OK;
    }
    double q;
    int worst;
    if ((rc = des_worst_quality(ctx, &q, &worst))) return rc;
    if (ctx->nd == 3) q = std::pow(q, 1.0 / 3);      // remeshing.cxx:1924-1927
    if (q < P.min_quality) {
        *reason = 1;
        *index = worst;
    }
    return DES_OK;
}

int des_step(des_ctx *ctx)
{
    if (!ctx) return DES_ERR_INPUT;
    int rc = flush_deferred(ctx);
    if (rc) return rc;
    hipLaunchKernelGGL(k_step_begin, dim3(1), dim3(1), 0, ctx->stream, ctx->s);
    if ((rc = check_launch(ctx))) return rc;
    ctx->steps++;
    if (ctx->P.has_thermal_diffusion) {
        LAUNCH_ND(k_temp_elem, nblocks(ctx->E), ctx->s);
        LAUNCH_ND(k_temp_node, nblocks(ctx->N), ctx->s);
    }
    LAUNCH_ND(k_strain_rate, nblocks(ctx->E), ctx->s);
    LAUNCH_ND(k_dvoldt_node, nblocks(ctx->N), ctx->s);
    LAUNCH_ND(k_edvoldt, nblocks(ctx->E), ctx->s);
    LAUNCH_ND(k_update_stress, nblocks(ctx->E), ctx->s);
    LAUNCH_ND(k_nmd_node, nblocks(ctx->N), ctx->s);
    LAUNCH_ND(k_nmd_elem, nblocks(ctx->E), ctx->s);
    if ((rc = launch_force(ctx))) return rc;
    LAUNCH_ND(k_update_velocity, nblocks(ctx->N), ctx->s);
    LAUNCH_ND(k_apply_vbcs, nblocks(ctx->N), ctx->s);
    if ((rc = launch_update_mesh(ctx))) return rc;
    if (ctx->P.rheol_type & DES_RH_ELASTIC) LAUNCH_ND(k_rotate, nblocks(ctx->E), ctx->s, 0);
    if (ctx->steps % 10 == 0)
        if ((rc = launch_compute_dt(ctx))) return rc;
    return DES_OK;
}

}  // extern "C"

namespace {

// Timing of the fused passes.  arm() before a pass hands its launch a
// start/stop event pair that the dispatch itself timestamps: the kernel's
// execution time, as a kernel trace reports it.  arm(true) brackets several
// launches with events recorded on the stream instead (launch gaps included).
// mark(k) books the armed pair under pass kind k.
struct PassTimer {
    struct Span {
        hipEvent_t a, b;
        int kind;
    };
    des_ctx *ctx;
    std::vector<Span> sp;
    hipEvent_t a = nullptr, b = nullptr;
    bool bracket = false;
    bool on;
    explicit PassTimer(des_ctx *c, bool enable) : ctx(c), on(enable) {}
    ~PassTimer()
    {
        if (!on) return;
        ctx->ev_a = ctx->ev_b = nullptr;
        for (auto &x : sp) {
            (void)hipEventDestroy(x.a);
            (void)hipEventDestroy(x.b);
        }
        if (a) (void)hipEventDestroy(a);
        if (b) (void)hipEventDestroy(b);
    }
    int arm(bool several = false)
    {
        if (!on) return DES_OK;
        HIPCHK(hipEventCreate(&a));
        HIPCHK(hipEventCreate(&b));
        bracket = several;
        if (bracket) {
            HIPCHK(hipEventRecord(a, ctx->stream));
        } else {
            ctx->ev_a = a;
            ctx->ev_b = b;
        }
        return DES_OK;
    }
    int mark(int k)
    {
        if (!on) return DES_OK;
        if (bracket) HIPCHK(hipEventRecord(b, ctx->stream));
        ctx->ev_a = ctx->ev_b = nullptr;
        sp.push_back({a, b, k});
        a = b = nullptr;
        return DES_OK;
    }
    int collect(double *ms, int *counts)
    {
        if (!on || sp.empty()) return DES_OK;
        HIPCHK(hipStreamSynchronize(ctx->stream));
        for (auto &x : sp) {
            float t = 0;
            HIPCHK(hipEventElapsedTime(&t, x.a, x.b));
            ms[x.kind] += t;
            counts[x.kind] += 1;
        }
        return DES_OK;
    }
};

#define ARM(several) do { if ((rc = tm.arm(several))) return rc; } while (0)
#define MARK(k) do { if ((rc = tm.mark(k))) return rc; } while (0)

// the deferred-return list of pass C (3D), allocated on first use
int ensure_retlist(des_ctx *ctx)
{
    DevState &s = ctx->s;
    if (!s.nrec) {
        int rc = dalloc_n(ctx, &s.nrec, (size_t)ctx->N * NREC);
        if (rc) return rc;
    }
    if (s.ret_count || ctx->nd != 3) return DES_OK;
    int rc;
    if ((rc = dalloc_n(ctx, &s.ret_elem, (size_t)ctx->E))) return rc;
    if ((rc = dalloc_n(ctx, &s.ret_val, (size_t)ctx->E * 8))) return rc;
    if ((rc = dalloc_n(ctx, &s.ret_count, 1))) return rc;
    return DES_OK;
}

// device buffers of the mixed-precision state, allocated on first use
int ensure_mixed(des_ctx *ctx)
{
    DevState &s = ctx->s;
    if (s.pres) return DES_OK;
    const size_t E = ctx->E, nstr = ctx->nstr;
    int rc;
    if ((rc = dalloc_n(ctx, &s.pres, E))) return rc;
    if ((rc = dalloc_n(ctx, &s.sdev, E * nstr))) return rc;
    if ((rc = dalloc_n(ctx, &s.strain32, E * nstr))) return rc;
    if ((rc = dalloc_n(ctx, &s.dpl32, E))) return rc;
    if (ctx->nd == 3 && (rc = dalloc_n(ctx, &s.ret_mix, E))) return rc;
    if (s.elem_visc) {
        float *v;
        if ((rc = dalloc_n(ctx, &v, E))) return rc;
        s.visc32 = v;
    }
    if (s.elem_cohesion) {
        float *v;
        if ((rc = dalloc_n(ctx, &v, E))) return rc;
        s.coh32 = v;
    }
    return DES_OK;
}

// ---- the autotuner's variants of the fused passes.  All variants of a pass
// perform the same arithmetic in the same order (bitwise equal results);
// they differ in how the work maps on the machine:
//   A, E: the block's node records staged in LDS (0) or gathered directly (1);
//   C:    a minimum of 1, 2, 3 or 4 waves per SIMD (__launch_bounds__): fewer
//         registers per lane buy occupancy at the price of spills; and 1 or 2
//         with the block's node records staged in LDS (FP64 only; the mixed
//         pass C has one variant);
//   D, F: node blocks of 256, 128 or 64 threads; D also 256 threads with 8
//         instead of 4 gathered entries in flight;
//   B:    4 or 8 element records in flight per node.
// The first TUNE_STEPS steps of a run of at least TUNE_MIN_RUN steps time
// every variant of every untuned pass with HIP events (variant k mod the
// variant count at tuning step k) and keep the fastest mean.
constexpr int TUNE_A = DES_TUNE_A, TUNE_C = DES_TUNE_C, TUNE_D = DES_TUNE_D, TUNE_E = DES_TUNE_E,
              TUNE_F = DES_TUNE_F, TUNE_B = DES_TUNE_B;
constexpr int TUNE_NV[2][DES_NTUNE] = {{3, 6, 5, 3, 5, 3}, {3, 6, 5, 2, 5, 3}};
constexpr int TUNE_DEFAULT[DES_NTUNE] = {0, 1, 3, 0, 3, 1};
constexpr int TUNE_STEPS = 2 * DES_TUNE_MAXV - 2, TUNE_MIN_RUN = 16;
constexpr int PASSC_VARIANTS = 6;

void clear_tuning_events(des_ctx *ctx)
{
    for (auto &t : ctx->tune) {
        (void)hipEventDestroy(t.a);
        (void)hipEventDestroy(t.b);
    }
    ctx->tune.clear();
}

int tuned_variant(const des_ctx *ctx, bool mixed, int pass)
{
    const int v = ctx->tune_var[mixed ? 1 : 0][pass];
    return v < 0 ? std::min(TUNE_DEFAULT[pass], TUNE_NV[mixed ? 1 : 0][pass] - 1) : v;
}

// kernel<ND, P2> in blocks of nthreads
#define LAUNCH_T2(kern, P2, nthreads, grid, ...)                                                    \
    do {                                                                                          \
        if (ctx->nd == 3)                                                                         \
            hipExtLaunchKernelGGL((kern<3, P2>), dim3(grid), dim3(nthreads), 0, ctx->stream,      \
                                  ctx->ev_a, ctx->ev_b, 0, __VA_ARGS__);                          \
        else                                                                                      \
            hipExtLaunchKernelGGL((kern<2, P2>), dim3(grid), dim3(nthreads), 0, ctx->stream,      \
                                  ctx->ev_a, ctx->ev_b, 0, __VA_ARGS__);                          \
        int rc_ = check_launch(ctx);                                                              \
        if (rc_) return rc_;                                                                      \
    } while (0)

// kernel<ND, P2, P3> in blocks of BLOCK threads
#define LAUNCH_E3(kern, P2, P3, grid, ...)                                                          \
    do {                                                                                          \
        if (ctx->nd == 3)                                                                         \
            hipExtLaunchKernelGGL((kern<3, P2, P3>), dim3(grid), dim3(BLOCK), 0, ctx->stream,      \
                                  ctx->ev_a, ctx->ev_b, 0, __VA_ARGS__);                          \
        else                                                                                      \
            hipExtLaunchKernelGGL((kern<2, P2, P3>), dim3(grid), dim3(BLOCK), 0, ctx->stream,      \
                                  ctx->ev_a, ctx->ev_b, 0, __VA_ARGS__);                          \
        int rc_ = check_launch(ctx);                                                              \
        if (rc_) return rc_;                                                                      \
    } while (0)

// pass A: node records staged in LDS (0), gathered directly (1), gathered
// directly with at least 6 waves per SIMD (2)
int launch_pass_a(des_ctx *ctx, const FusedFlags &fl, int variant)
{
    if (variant == 2) LAUNCH_E3(k_pass_a, false, 6, nblocks(ctx->E), ctx->s, fl);
    else if (variant == 1) LAUNCH_T2(k_pass_a, false, BLOCK, nblocks(ctx->E), ctx->s, fl);
    else LAUNCH_T2(k_pass_a, true, BLOCK, nblocks(ctx->E), ctx->s, fl);
    return DES_OK;
}

// kernel<ND, NB, DEPTH> in blocks of NB threads over nthreads threads
#define LAUNCH_T3(kern, NB, DEPTH, nthreads, ...)                                                   \
    do {                                                                                          \
        const dim3 g_((unsigned)(((nthreads) + (NB) - 1) / (NB))), b_(NB);                        \
        if (ctx->nd == 3)                                                                         \
            hipExtLaunchKernelGGL((kern<3, NB, DEPTH>), g_, b_, 0, ctx->stream, ctx->ev_a,        \
                                  ctx->ev_b, 0, __VA_ARGS__);                                     \
        else                                                                                      \
            hipExtLaunchKernelGGL((kern<2, NB, DEPTH>), g_, b_, 0, ctx->stream, ctx->ev_a,        \
                                  ctx->ev_b, 0, __VA_ARGS__);                                     \
        int rc_ = check_launch(ctx);                                                              \
        if (rc_) return rc_;                                                                      \
    } while (0)

int launch_pass_b(des_ctx *ctx, const FusedFlags &fl, int variant)
{
    const long nthreads = 4L * ctx->N;
    if (variant == 2) LAUNCH_T2(k_pass_b, 16, BLOCK, nblocks(nthreads), ctx->s, fl);
    else if (variant == 1) LAUNCH_T2(k_pass_b, 8, BLOCK, nblocks(nthreads), ctx->s, fl);
    else LAUNCH_T2(k_pass_b, 4, BLOCK, nblocks(nthreads), ctx->s, fl);
    return DES_OK;
}

// the node passes D and F: blocks of 256/128/64 threads with 4 entries in
// flight (variants 0..2), 256 threads with 8 or 16 in flight (3, 4)
#define NODE_PASS_VARIANTS(kern, ...)                                                             \
    switch (variant) {                                                                            \
    case 1: LAUNCH_T3(kern, 128, 4, ctx->N, __VA_ARGS__); break;                                  \
    case 2: LAUNCH_T3(kern, 64, 4, ctx->N, __VA_ARGS__); break;                                   \
    case 3: LAUNCH_T3(kern, 256, 8, ctx->N, __VA_ARGS__); break;                                  \
    case 4: LAUNCH_T3(kern, 256, 16, ctx->N, __VA_ARGS__); break;                                 \
    default: LAUNCH_T3(kern, 256, 4, ctx->N, __VA_ARGS__); break;                                 \
    }

int launch_pass_d(des_ctx *ctx, const FusedFlags &fl, int variant)
{
    NODE_PASS_VARIANTS(k_pass_d, ctx->s, fl);
    return DES_OK;
}

int launch_pass_e(des_ctx *ctx, const FusedFlags &fl, bool mixed, int variant)
{
    if (mixed) {
        if (variant == 1) LAUNCH_T2(k_pass_e_mixed, false, BLOCK, nblocks(ctx->E), ctx->s, fl);
        else LAUNCH_T2(k_pass_e_mixed, true, BLOCK, nblocks(ctx->E), ctx->s, fl);
    } else if (g_force_layout == 1) {
        if (variant == 1) LAUNCH_T2(k_pass_e_er, false, BLOCK, nblocks(ctx->E), ctx->s, fl);
        else LAUNCH_T2(k_pass_e_er, true, BLOCK, nblocks(ctx->E), ctx->s, fl);
    } else {     // staged (0), direct (1), staged with at least 8 waves per SIMD (2)
        if (variant == 2) LAUNCH_E3(k_pass_e, true, 8, nblocks(ctx->E), ctx->s, fl);
        else if (variant == 1) LAUNCH_T2(k_pass_e, false, BLOCK, nblocks(ctx->E), ctx->s, fl);
        else LAUNCH_T2(k_pass_e, true, BLOCK, nblocks(ctx->E), ctx->s, fl);
    }
    return DES_OK;
}

int launch_pass_f(des_ctx *ctx, const FusedFlags &fl, bool mixed, int variant)
{
    if (mixed) {
        NODE_PASS_VARIANTS(k_pass_f32, ctx->s, fl);
    } else if (g_force_layout == 1) {
        NODE_PASS_VARIANTS(k_pass_f_er, ctx->s, fl);
    } else {
        NODE_PASS_VARIANTS(k_pass_f, ctx->s, fl);
    }
    return DES_OK;
}

// whether every model of a 3D context is elasto-plastic or elasto-visco-plastic
bool all_plastic3(const des_ctx *ctx)
{
    if (ctx->nd != 3) return false;
    auto ep = [](const des_params &P) { return P.rheol_type == DES_RH_EP || P.rheol_type == DES_RH_EVP; };
    if (!ctx->nmodels) return ep(ctx->P);
    for (const des_params &P : ctx->hP)
        if (!ep(P)) return false;
    return true;
}

int launch_pass_c(des_ctx *ctx, const FusedFlags &fl, int variant)
{
    dim3 g(nblocks(ctx->E)), b(BLOCK);
#define PC(NDV, W, ST) \
    hipExtLaunchKernelGGL((k_pass_c<NDV, W, ST>), g, b, 0, ctx->stream, ctx->ev_a, ctx->ev_b, 0, ctx->s, fl)
#define PCP(W, ST) \
    hipExtLaunchKernelGGL((k_pass_c<3, W, ST, true>), g, b, 0, ctx->stream, ctx->ev_a, ctx->ev_b, 0, ctx->s, fl)
    if (ctx->nd == 3 && all_plastic3(ctx)) {
        switch (variant) {
        case 0: PCP(1, false); break;
        case 1: PCP(2, false); break;
        case 2: PCP(3, false); break;
        case 3: PCP(4, false); break;
        case 4: PCP(1, true); break;
        default: PCP(2, true); break;
        }
    } else if (ctx->nd == 3) {
        switch (variant) {
        case 0: PC(3, 1, false); break;
        case 1: PC(3, 2, false); break;
        case 2: PC(3, 3, false); break;
        case 3: PC(3, 4, false); break;
        case 4: PC(3, 1, true); break;
        default: PC(3, 2, true); break;
        }
    } else {
        switch (variant) {
        case 0: PC(2, 1, false); break;
        case 1: PC(2, 2, false); break;
        case 2: PC(2, 3, false); break;
        case 3: PC(2, 4, false); break;
        case 4: PC(2, 1, true); break;
        default: PC(2, 2, true); break;
        }
    }
#undef PC
#undef PCP
    return check_launch(ctx);
}

// whether any model of the context has a plastic rheology (only those defer
// returns to the list)
bool any_plastic(const des_ctx *ctx)
{
    if (!ctx->nmodels) return (ctx->P.rheol_type & DES_RH_PLASTIC) != 0;
    for (const des_params &P : ctx->hP)
        if (P.rheol_type & DES_RH_PLASTIC) return true;
    return false;
}

// the Mohr-Coulomb returns that pass C deferred (3D); the list is short, so a
// small grid walks it
int launch_pass_c_ret(des_ctx *ctx, bool mixed = false)
{
    if (ctx->nd != 3 || !any_plastic(ctx)) return DES_OK;
    const dim3 g(std::min(nblocks(ctx->E), 512)), b(BLOCK);
    if (mixed) hipExtLaunchKernelGGL(k_pass_c_ret_mixed, g, b, 0, ctx->stream, ctx->ev_a, ctx->ev_b, 0, ctx->s);
    else hipExtLaunchKernelGGL(k_pass_c_ret, g, b, 0, ctx->stream, ctx->ev_a, ctx->ev_b, 0, ctx->s);
    return check_launch(ctx);
}

int fused_begin(des_ctx *ctx, int precision)
{
    if (precision != DES_FP64 && precision != DES_MIXED)
        return fail(ctx, DES_ERR_INPUT, "unknown precision");
    const int nm = std::max(ctx->nmodels, 1);
    for (int k = 0; k < nm; ++k) {
        const des_params &P = ctx->nmodels ? ctx->hP[k] : ctx->P;
        if (precision == DES_MIXED && P.is_plane_strain)
            return fail(ctx, DES_ERR_INPUT, "mixed precision does not cover plane strain");
        if (precision == DES_MIXED && ctx->avg_interval > 0)
            return fail(ctx, DES_ERR_INPUT, "averaged fields are accumulated by FP64 runs");
    }
    if (ctx->nmodels) {
        // per-element viscosity / cohesion are read for every model or none
        for (int k = 1; k < ctx->nmodels; ++k)
            if (ctx->has_visc[k] != ctx->has_visc[0] || ctx->has_coh[k] != ctx->has_coh[0])
                return fail(ctx, DES_ERR_INPUT, "per-element viscosity / cohesion uploaded for some models only");
    }
    // A run right after a run of the same precision (nothing in between
    // touched the state, or it would have completed the deferred tail) carries
    // on as the next step of the last one: its first step swaps the volumes,
    // rebuilds the mass in pass B and applies the owed rotation in pass C, as
    // every later step does, and the node records and the FP32 stress,
    // strain and plastic-strain increment of a mixed run are current.  The
    // results are those of one run of all the steps, without the tail's extra
    // passes between the runs.
    int rc;
    if (ctx->closing_pending && ctx->run_precision != precision && (rc = flush_deferred(ctx))) return rc;
    ctx->continuing = ctx->closing_pending;
    ctx->run_precision = precision;
    if ((rc = ensure_retlist(ctx))) return rc;
    ctx->s.srank = g_force_layout == 2 ? nullptr : ctx->srank;
    clear_tuning_events(ctx);          // left by a run that stopped while tuning
    if (!ctx->continuing) LAUNCH_ND(k_pack_nrec, nblocks(ctx->N), ctx->s);
    if (precision == DES_MIXED) {
        if ((rc = ensure_mixed(ctx))) return rc;
        if (!ctx->continuing) LAUNCH_ND(k_to_mixed, nblocks(ctx->E), ctx->s);
    }
    return DES_OK;
}

// Step k of an nsteps run: the six passes, the step's bookkeeping and, every
// tenth step, compute_dt.
int fused_step(des_ctx *ctx, int k, int nsteps, bool mixed, PassTimer &tm)
{
    int rc;
    FusedFlags fl;
    fl.use_stored_mass = (k == 0 && !ctx->continuing);
    fl.store_aux = (k == nsteps - 1);
    if (k > 0 || ctx->continuing) swap_volume(ctx);
    // a tuning step: the untuned passes run variant k mod their variant count
    // between HIP events
    const int pr = mixed ? 1 : 0;
    bool tuning = false;
    if (k < TUNE_STEPS && nsteps >= TUNE_MIN_RUN)
        for (int p = 0; p < DES_NTUNE; ++p) tuning |= ctx->tune_var[pr][p] < 0;
    auto tunable = [&](int p, auto &&launch) -> int {
        const bool timed = tuning && ctx->tune_var[pr][p] < 0;
        const int v = timed ? k % TUNE_NV[pr][p] : tuned_variant(ctx, mixed, p);
        hipEvent_t a = nullptr, b = nullptr;
        if (timed) {
            HIPCHK(hipEventCreate(&a));
            HIPCHK(hipEventCreate(&b));
            HIPCHK(hipEventRecord(a, ctx->stream));
        }
        const int rc2 = launch(v);
        if (rc2) return rc2;
        if (timed) {
            HIPCHK(hipEventRecord(b, ctx->stream));
            ctx->tune.push_back({p, v, a, b});
        }
        return DES_OK;
    };
    // the element passes walk their blocks upwards, the node passes (B, D, F)
    // downwards: each pass starts on what the one before it wrote last
    const int down = g_sweep ? 1 : 0;
    ARM(false);
    ctx->s.rev = 0;
    if ((rc = tunable(TUNE_A, [&](int v) { return launch_pass_a(ctx, fl, v); }))) return rc;
    MARK(0);
    ARM(false);
    ctx->s.rev = down;
    rc = tunable(TUNE_B, [&](int v) { return launch_pass_b(ctx, fl, v); });
    ctx->s.rev = 0;
    if (rc) return rc;
    MARK(1);
    ARM(false);
    rc = tunable(TUNE_C, [&](int v) -> int {
        if (mixed) {       // a minimum of 1..4 waves per SIMD (variants 0..3), LDS-staged 1..2 (4, 5)
            const dim3 g(nblocks(ctx->E)), b(BLOCK);
#define PCM(W, ST)                                                                                  \
    (ctx->nd == 3 ? hipExtLaunchKernelGGL((k_pass_c_mixed<3, W, ST>), g, b, 0, ctx->stream, ctx->ev_a,  \
                                          ctx->ev_b, 0, ctx->s, fl)                                 \
                  : hipExtLaunchKernelGGL((k_pass_c_mixed<2, W, ST>), g, b, 0, ctx->stream, ctx->ev_a,  \
                                          ctx->ev_b, 0, ctx->s, fl))
            switch (v) {
            case 1: PCM(2, false); break;
            case 2: PCM(3, false); break;
            case 3: PCM(4, false); break;
            case 4: PCM(1, true); break;
            case 5: PCM(2, true); break;
            default: PCM(1, false); break;
            }
#undef PCM
            return check_launch(ctx);
        }
        return launch_pass_c(ctx, fl, v);
    });
    if (rc) return rc;
    MARK(2);
    if (ctx->nd == 3 && any_plastic(ctx)) {
        ARM(false);
        if ((rc = launch_pass_c_ret(ctx, mixed))) return rc;
        MARK(8);
    }
    ARM(false);
    ctx->s.rev = down;
    rc = tunable(TUNE_D, [&](int v) { return launch_pass_d(ctx, fl, v); });
    ctx->s.rev = 0;
    if (rc) return rc;
    MARK(3);
    ARM(false);
    if ((rc = tunable(TUNE_E, [&](int v) { return launch_pass_e(ctx, fl, mixed, v); }))) return rc;
    MARK(4);
    if (ctx->nactive) {
        ARM(false);
        LAUNCH_ND(k_facet_traction, nblocks(ctx->nactive), ctx->s, ctx->active_facets, ctx->nactive);
        MARK(5);
    }
    ARM(false);
    FusedFlags ff = fl;
    ff.step_end = 1;
    ctx->s.rev = down;
    rc = tunable(TUNE_F, [&](int v) { return launch_pass_f(ctx, ff, mixed, v); });
    ctx->s.rev = 0;
    if (rc) return rc;
    MARK(6);
    if (tuning && k == TUNE_STEPS - 1 && !ctx->tune.empty()) {
        // the fastest mean per untuned pass
        HIPCHK(hipEventSynchronize(ctx->tune.back().b));
        double sum[DES_NTUNE][DES_TUNE_MAXV] = {};
        int cnt[DES_NTUNE][DES_TUNE_MAXV] = {};
        for (auto &t : ctx->tune) {
            float ms = 0;
            HIPCHK(hipEventElapsedTime(&ms, t.a, t.b));
            sum[t.pass][t.variant] += ms;
            cnt[t.pass][t.variant]++;
        }
        clear_tuning_events(ctx);
        for (int p = 0; p < DES_NTUNE; ++p) {
            if (ctx->tune_var[pr][p] >= 0) continue;
            int best = -1;
            for (int v = 0; v < TUNE_NV[pr][p]; ++v) {
                ctx->tune_ms[pr][p][v] = cnt[p][v] ? sum[p][v] / cnt[p][v] : 0;
                if (cnt[p][v] && (best < 0 || ctx->tune_ms[pr][p][v] < ctx->tune_ms[pr][p][best])) best = v;
            }
            ctx->tune_var[pr][p] = best < 0 ? TUNE_DEFAULT[p] : best;
        }
    }
    ctx->steps++;
    ctx->closing_pending = true;
    if (ctx->steps % 10 == 0) {
        ARM(true);
        if ((rc = launch_compute_dt(ctx))) return rc;
        MARK(7);
    }
    if (ctx->avg_interval > 0 && (rc = average_fields(ctx))) return rc;   // dynearthsol.cxx:363-364
    return DES_OK;
}

int fused_end(des_ctx *ctx, bool mixed)
{
    if (mixed) LAUNCH_ND(k_from_mixed, nblocks(ctx->E), ctx->s);
    return DES_OK;
}

// ---- the persistent launch (k_persistent)

template <bool M, bool H3, bool H2> const void *persistent_kernel()
{
    return reinterpret_cast<const void *>(&k_persistent<M, H3, H2>);
}

const void *persistent_kernel(bool mixed, bool h3, bool h2)
{
    if (mixed)
        return h3 ? (h2 ? persistent_kernel<true, true, true>() : persistent_kernel<true, true, false>())
                  : persistent_kernel<true, false, true>();
    return h3 ? (h2 ? persistent_kernel<false, true, true>() : persistent_kernel<false, true, false>())
              : persistent_kernel<false, false, true>();
}

PersistGroup persist_group(const des_ctx *c)
{
    PersistGroup g{};
    g.neb = nblocks(c->E);
    g.nqb = nblocks(4L * c->N);
    g.nnb = nblocks(c->N);
    g.plastic = c->nd == 3 && any_plastic(c);
    g.nactive = c->nactive;
    g.active_facets = c->active_facets;
    g.nmodels = c->nmodels;
    return g;
}

// A whole run of nsteps fused steps of c3 (3D) and c2 (2D) -- either may be
// null -- in one cooperative launch on ctx's stream (ctx is one of them).
// Both are past fused_begin and at the same step count.  The grid is as many
// workgroups as can be resident together (hipLaunchCooperativeKernel refuses
// more), a multiple of 8, and no more than the largest pass has blocks.
int run_persistent(des_ctx *ctx, des_ctx *c3, des_ctx *c2, int nsteps, bool mixed)
{
    const void *kern = persistent_kernel(mixed, c3 != nullptr, c2 != nullptr);
    int per_cu = 0, cus = 0;
    HIPCHK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, BLOCK, 0));
    HIPCHK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, ctx->device));
    PersistArgs pa{};
    if (c3) pa.g3 = persist_group(c3);
    if (c2) pa.g2 = persist_group(c2);
    const long most = std::max({(long)pa.g3.neb + pa.g2.neb, (long)pa.g3.nqb + pa.g2.nqb, (long)pa.g3.nnb + pa.g2.nnb});
    const long G = std::min((long)per_cu * cus, (most + 7) / 8 * 8) / 8 * 8;
    if (G < 8) return fail(ctx, DES_ERR_HIP, "persistent launch: fewer than 8 workgroups fit the device");
    pa.nsteps = nsteps;
    pa.steps0 = ctx->steps;
    pa.use_stored_mass = ctx->continuing ? 0 : 1;
    pa.swap0 = ctx->continuing ? 1 : 0;
    pa.store_aux = 1;
    pa.rev = g_sweep ? 1 : 0;
    // the groups' states go to device memory, where each pass of the kernel
    // reads what it needs (scalar loads) instead of holding both in registers
    const DevState *ps[2] = {nullptr, nullptr};
    int k = 0;
    for (des_ctx *c : {c3, c2}) {
        if (c) {
            if (!c->pstate_dev) {      // no clearing: the copy below fills it, on this stream
                void *p = nullptr;
                HIPCHK(hipMalloc(&p, sizeof(DevState)));
                c->allocs.push_back(p);
                c->pstate_dev = static_cast<DevState *>(p);
            }
            c->pstate_host = c->s;
            c->pstate_host.rev = 0;
            HIPCHK(hipMemcpyAsync(c->pstate_dev, &c->pstate_host, sizeof(DevState), hipMemcpyHostToDevice,
                                  ctx->stream));
            ps[k] = c->pstate_dev;
        }
        ++k;
    }
    void *args[] = {&ps[0], &ps[1], &pa};
    HIPCHK(hipLaunchCooperativeKernel(kern, dim3((unsigned)G), dim3(BLOCK), args, 0, ctx->stream));
    for (des_ctx *c : {c3, c2}) {
        if (!c) continue;
        if ((nsteps - 1 + pa.swap0) & 1) swap_volume(c);   // the swaps the kernel made
        c->steps += nsteps;
        c->closing_pending = true;
    }
    return DES_OK;
}

int run_fused(des_ctx *ctx, int nsteps, int precision, PassTimer &tm)
{
    if (nsteps <= 0) return precision == DES_FP64 || precision == DES_MIXED ? DES_OK : DES_ERR_INPUT;
    int rc = fused_begin(ctx, precision);
    if (rc) return rc;
    const bool mixed = precision == DES_MIXED;
    if (ctx->launch_mode == 1 && !tm.on && ctx->avg_interval == 0) {
        if ((rc = run_persistent(ctx, ctx->nd == 3 ? ctx : nullptr, ctx->nd == 2 ? ctx : nullptr, nsteps, mixed)))
            return rc;
    } else {
        for (int k = 0; k < nsteps; ++k)
            if ((rc = fused_step(ctx, k, nsteps, mixed, tm))) return rc;
    }
    return fused_end(ctx, mixed);
}

}  // namespace

extern "C" {

int des_run(des_ctx *ctx, int nsteps, int precision)
{
    if (!ctx) return DES_ERR_INPUT;
    PassTimer tm(ctx, false);
    return run_fused(ctx, nsteps, precision, tm);
}

int des_set_launch_mode(des_ctx *ctx, int mode)
{
    if (!ctx || mode < 0 || mode > 1) return DES_ERR_INPUT;
    ctx->launch_mode = mode;
    return DES_OK;
}

int des_set_tuning(des_ctx *ctx, int variant)
{
    if (!ctx || variant < -1 || variant >= PASSC_VARIANTS) return DES_ERR_INPUT;
    ctx->tune_var[0][TUNE_C] = variant;
    return DES_OK;
}

int des_get_tuning_all(des_ctx *ctx, int precision, int variant[DES_NTUNE],
                       double ms[DES_TUNE_MAXV * DES_NTUNE])
{
    if (!ctx || !variant || (precision != DES_FP64 && precision != DES_MIXED)) return DES_ERR_INPUT;
    const int pr = precision == DES_MIXED ? 1 : 0;
    for (int p = 0; p < DES_NTUNE; ++p) {
        variant[p] = ctx->tune_var[pr][p];
        if (ms)
            for (int v = 0; v < DES_TUNE_MAXV; ++v) ms[DES_TUNE_MAXV * p + v] = ctx->tune_ms[pr][p][v];
    }
    return DES_OK;
}

int des_set_tuning_all(des_ctx *ctx, int precision, const int variant[DES_NTUNE])
{
    if (!ctx || !variant || (precision != DES_FP64 && precision != DES_MIXED)) return DES_ERR_INPUT;
    const int pr = precision == DES_MIXED ? 1 : 0;
    for (int p = 0; p < DES_NTUNE; ++p)
        if (variant[p] < -1 || variant[p] >= TUNE_NV[pr][p])
            return fail(ctx, DES_ERR_INPUT, "tuning variant out of range for pass " + std::to_string(p));
    for (int p = 0; p < DES_NTUNE; ++p) ctx->tune_var[pr][p] = variant[p];
    return DES_OK;
}

int des_run_timed(des_ctx *ctx, int nsteps, int precision, double *ms, int *counts)
{
    if (!ctx || !ms || !counts) return DES_ERR_INPUT;
    PassTimer tm(ctx, true);
    int rc = run_fused(ctx, nsteps, precision, tm);
    if (rc) return rc;
    return tm.collect(ms, counts);
}

int des_sync(des_ctx *ctx)
{
    if (!ctx) return DES_ERR_INPUT;
    HIPCHK(hipStreamSynchronize(ctx->stream));
    DevScalars h{};
    int rc = read_scalars(ctx, h);
    if (rc) return rc;
    if (h.err) return fail(ctx, h.err, "device reported an error");
    return DES_OK;
}

}  // extern "C"

// ================================================================ batches
//
// Many independent models advance together.  Models of one dimension form a
// group: their meshes are concatenated into one context, each model's
// elements and nodes starting on a block boundary, so that every launch of
// the fused schedule covers all models of the group and each block works for
// exactly one model (its parameters and scalars are bound from the block ->
// model tables).  The 2D and the 3D group run on their own streams, one step
// of each enqueued in turn.

struct des_batch {
    std::vector<des_ctx *> groups;
    std::vector<int> group_of, index_of;   // per model
    std::vector<int> msteps;               // per model: var.steps as last set
    std::string err;
    int launch_mode = 0;                   // des_batch_run: 0 per-pass launches, 1 one persistent launch
};

namespace {

int bfail(des_batch *b, int code, const std::string &msg)
{
    b->err = msg;
    return code;
}

// a group call's status, with the group's message on failure
int bcheck(des_batch *b, des_ctx *ctx, int rc)
{
    if (rc) b->err = ctx->err.empty() ? "status " + std::to_string(rc) : ctx->err;
    return rc;
}

int build_group(des_batch *b, const std::vector<int> &ids, const des_params *params, const des_mesh *meshes)
{
    const des_params &P0 = params[ids[0]];
    const des_mesh &M0 = meshes[ids[0]];
    const int nd = P0.ndims, npe = nd + 1, nmat = P0.nmat;
    GroupLayout g;
    long E = 0, N = 0, F = 0;
    for (int id : ids) {
        const des_params &p = params[id];
        const des_mesh &m = meshes[id];
        if (p.nmat != nmat || m.markers_per_element != M0.markers_per_element)
            return bfail(b, DES_ERR_INPUT, "models of one dimension must share nmat and markers_per_element");
        g.P.push_back(p);
        g.e_off.push_back((int)E);
        g.n_off.push_back((int)N);
        g.e_cnt.push_back(m.nelem);
        g.n_cnt.push_back(m.nnode);
        g.nbf_cnt.push_back(m.nbfacet);
        E += (long)nblocks(m.nelem) * BLOCK;
        N += (long)nblocks(m.nnode) * BLOCK;
        F += m.nbfacet;
        if (E * npe >= (1L << 31) || N >= (1L << 31))
            return bfail(b, DES_ERR_INPUT, "batched group too large for 32-bit indices");
    }
    std::vector<int> conn((size_t)E * npe, 0), bdry, felem, fface;
    std::vector<double> coord((size_t)N * nd, 0.0);
    std::vector<unsigned> bcflag(N, 0u);
    std::vector<int> markers(nmat > 1 ? (size_t)E * nmat : 0, 0);
    std::vector<int> eh(E, -1), nh(N, -1);
    bdry.reserve(F);
    felem.reserve(F);
    fface.reserve(F);
    for (size_t k = 0; k < ids.size(); ++k) {
        OrderedMesh om;
        order_mesh(meshes[ids[k]], nd, nmat, g_mesh_order, om);
        const des_mesh &m = om.m;
        const int eo = g.e_off[k], no = g.n_off[k];
        std::copy(om.eh.begin(), om.eh.end(), eh.begin() + eo);
        std::copy(om.nh.begin(), om.nh.end(), nh.begin() + no);
        for (long i = 0; i < (long)m.nelem * npe; ++i) conn[(size_t)eo * npe + i] = m.connectivity[i] + no;
        std::memcpy(coord.data() + (size_t)no * nd, m.coord, sizeof(double) * m.nnode * nd);
        std::memcpy(bcflag.data() + no, m.bcflag, sizeof(unsigned) * m.nnode);
        if (nmat > 1) std::memcpy(markers.data() + (size_t)eo * nmat, m.elemmarkers, sizeof(int) * m.nelem * nmat);
        for (int f = 0; f < m.nbfacet; ++f) {
            bdry.push_back(m.bfacet_bdry[f]);
            felem.push_back(m.bfacet_elem[f] + eo);
            fface.push_back(m.bfacet_face[f]);
            g.facet_model.push_back((int)k);
        }
        const size_t nbn = DES_NBDRY * nd, nev = DES_NBDRY * DES_NBDRY * nd, nevv = DES_NBDRY * DES_NBDRY;
        const size_t bn0 = g.bn.size(), ev0 = g.ev.size(), evv0 = g.evv.size();
        g.bn.resize(bn0 + nbn, 0.0);
        g.ev.resize(ev0 + nev, 0.0);
        g.evv.resize(evv0 + nevv, 0);
        if (m.bnormals) std::memcpy(g.bn.data() + bn0, m.bnormals, sizeof(double) * nbn);
        if (m.edge_vectors) std::memcpy(g.ev.data() + ev0, m.edge_vectors, sizeof(double) * nev);
        if (m.edge_valid) std::memcpy(g.evv.data() + evv0, m.edge_valid, sizeof(int) * nevv);
    }
    des_mesh sm{};
    sm.nnode = (int)N;
    sm.nelem = (int)E;
    sm.connectivity = conn.data();
    sm.coord = coord.data();
    sm.bcflag = bcflag.data();
    sm.nbfacet = (int)F;
    sm.bfacet_bdry = bdry.data();
    sm.bfacet_elem = felem.data();
    sm.bfacet_face = fface.data();
    sm.elemmarkers = nmat > 1 ? markers.data() : nullptr;
    sm.markers_per_element = M0.markers_per_element;
    des_ctx *ctx = new des_ctx();
    b->groups.push_back(ctx);
    int rc = create_impl(ctx, &P0, &sm, &g, eh, nh);
    if (rc) return bcheck(b, ctx, rc);
    for (size_t k = 0; k < ids.size(); ++k) {
        b->group_of[ids[k]] = (int)b->groups.size() - 1;
        b->index_of[ids[k]] = (int)k;
    }
    return DES_OK;
}

int model_ctx(des_batch *b, int model, des_ctx **ctx, int *idx)
{
    if (model < 0 || model >= (int)b->group_of.size()) return bfail(b, DES_ERR_INPUT, "model index out of range");
    *ctx = b->groups[b->group_of[model]];
    *idx = b->index_of[model];
    return DES_OK;
}

int read_model_scalars(des_ctx *ctx, int idx, DevScalars &h)
{
    HIPCHK(hipMemcpyAsync(&h, ctx->s.scb + idx, sizeof(DevScalars), hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return DES_OK;
}

}  // namespace

extern "C" {

const char *des_batch_last_error(const des_batch *b) { return b ? b->err.c_str() : "null batch"; }

int des_batch_create(int nmodels, const des_params *params, const des_mesh *meshes, des_batch **out)
{
    if (!out) return DES_ERR_INPUT;
    *out = nullptr;
    if (nmodels <= 0 || !params || !meshes) return DES_ERR_INPUT;
    for (int k = 0; k < nmodels; ++k) {
        int rc = validate_mesh(&params[k], &meshes[k]);
        if (rc) return rc;
    }
    des_batch *b = new des_batch();
    *out = b;
    b->group_of.assign(nmodels, -1);
    b->index_of.assign(nmodels, -1);
    b->msteps.assign(nmodels, 0);
    for (int nd = 2; nd <= 3; ++nd) {
        std::vector<int> ids;
        for (int k = 0; k < nmodels; ++k)
            if (params[k].ndims == nd) ids.push_back(k);
        if (ids.empty()) continue;
        int rc = build_group(b, ids, params, meshes);
        if (rc) return rc;
    }
    return DES_OK;
}

void des_batch_destroy(des_batch *b)
{
    if (!b) return;
    for (des_ctx *c : b->groups) des_destroy(c);
    delete b;
}

int des_batch_upload_fields(des_batch *b, int model, const des_fields *f)
{
    if (!b || !f) return DES_ERR_INPUT;
    des_ctx *ctx;
    int idx, rc = model_ctx(b, model, &ctx, &idx);
    if (rc) return rc;
    return bcheck(b, ctx, xfer_fields(ctx, const_cast<des_fields *>(f), true, idx));
}

int des_batch_download_fields(des_batch *b, int model, des_fields *f)
{
    if (!b || !f) return DES_ERR_INPUT;
    des_ctx *ctx;
    int idx, rc = model_ctx(b, model, &ctx, &idx);
    if (rc) return rc;
    return bcheck(b, ctx, xfer_fields(ctx, f, false, idx));
}

int des_batch_set_scalars(des_batch *b, int model, double dt, double time, int steps,
                          double compensation_pressure)
{
    if (!b) return DES_ERR_INPUT;
    des_ctx *ctx;
    int idx, rc = model_ctx(b, model, &ctx, &idx);
    if (rc) return rc;
    if ((rc = bcheck(b, ctx, flush_deferred(ctx)))) return rc;
    DevScalars h{};
    if ((rc = bcheck(b, ctx, read_model_scalars(ctx, idx, h)))) return rc;
    h.dt = dt;
    h.time = time;
    h.steps = steps;
    h.compensation_pressure = compensation_pressure;
    h.rot_pending = 0;
    hipError_t e = hipMemcpyAsync(ctx->s.scb + idx, &h, sizeof h, hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) return bfail(b, DES_ERR_HIP, hipGetErrorString(e));
    b->msteps[model] = steps;
    ctx->steps = steps;
    return DES_OK;
}

int des_batch_get_scalars(des_batch *b, int model, double *dt, double *time, int *steps)
{
    if (!b) return DES_ERR_INPUT;
    des_ctx *ctx;
    int idx, rc = model_ctx(b, model, &ctx, &idx);
    if (rc) return rc;
    DevScalars h{};
    if ((rc = bcheck(b, ctx, read_model_scalars(ctx, idx, h)))) return rc;
    if (dt) *dt = h.dt;
    if (time) *time = h.time;
    if (steps) *steps = h.steps;
    if (h.err) return bfail(b, h.err, "device reported an error in model " + std::to_string(model));
    return DES_OK;
}

int des_batch_init_derived(des_batch *b)
{
    if (!b) return DES_ERR_INPUT;
    for (des_ctx *c : b->groups) {
        int rc = des_init_derived(c);
        if (rc) return bcheck(b, c, rc);
    }
    return DES_OK;
}

int des_batch_apply_vbcs(des_batch *b)
{
    if (!b) return DES_ERR_INPUT;
    for (des_ctx *c : b->groups) {
        int rc = des_apply_vbcs(c);
        if (rc) return bcheck(b, c, rc);
    }
    return DES_OK;
}

int des_batch_compute_dt(des_batch *b, double *dt)
{
    if (!b) return DES_ERR_INPUT;
    for (des_ctx *ctx : b->groups) {
        int rc = flush_deferred(ctx);
        if (!rc) rc = launch_compute_dt(ctx);
        if (rc) return bcheck(b, ctx, rc);
    }
    for (int m = 0; m < (int)b->group_of.size(); ++m) {
        des_ctx *ctx = b->groups[b->group_of[m]];
        DevScalars h{};
        int rc = read_model_scalars(ctx, b->index_of[m], h);
        if (rc) return bcheck(b, ctx, rc);
        if (dt) dt[m] = h.dt;
        if (h.err) return bfail(b, h.err, "compute_dt: dt <= 0 in model " + std::to_string(m));
    }
    return DES_OK;
}

int des_batch_run(des_batch *b, int nsteps, int precision)
{
    if (!b) return DES_ERR_INPUT;
    if (precision != DES_FP64 && precision != DES_MIXED) return bfail(b, DES_ERR_INPUT, "unknown precision");
    if (nsteps <= 0) return DES_OK;
    for (size_t m = 1; m < b->msteps.size(); ++m)
        if (b->msteps[m] != b->msteps[0])
            return bfail(b, DES_ERR_STATE, "batched models must start a run at the same step count");
    const bool mixed = precision == DES_MIXED;
    std::vector<PassTimer> tms;
    tms.reserve(b->groups.size());       // timers disabled: no events to own
    for (des_ctx *c : b->groups) tms.emplace_back(c, false);
    for (des_ctx *c : b->groups) {
        int rc = fused_begin(c, precision);
        if (rc) return bcheck(b, c, rc);
    }
    bool persistent = b->launch_mode == 1;
    des_ctx *c3 = nullptr, *c2 = nullptr;
    for (des_ctx *c : b->groups) {
        (c->nd == 3 ? c3 : c2) = c;
        persistent = persistent && c->avg_interval == 0 && c->steps == b->groups[0]->steps &&
                     c->continuing == b->groups[0]->continuing;
    }
    if (persistent) {
        // both groups in one launch on the 3D group's stream, ordered after
        // the 2D group's earlier work and before its later work
        des_ctx *owner = c3 ? c3 : c2;
        hipEvent_t before = nullptr, after = nullptr;
        if (c3 && c2) {
            if (hipEventCreateWithFlags(&before, hipEventDisableTiming) != hipSuccess ||
                hipEventCreateWithFlags(&after, hipEventDisableTiming) != hipSuccess ||
                hipEventRecord(before, c2->stream) != hipSuccess ||
                hipStreamWaitEvent(c3->stream, before, 0) != hipSuccess)
                return bfail(b, DES_ERR_HIP, "persistent launch: stream ordering");
        }
        int rc = run_persistent(owner, c3, c2, nsteps, mixed);
        if (c3 && c2) {
            if (!rc && (hipEventRecord(after, c3->stream) != hipSuccess ||
                        hipStreamWaitEvent(c2->stream, after, 0) != hipSuccess))
                rc = fail(owner, DES_ERR_HIP, "persistent launch: stream ordering");
            (void)hipEventDestroy(before);
            (void)hipEventDestroy(after);
        }
        if (rc) return bcheck(b, owner, rc);
    } else {
        for (int k = 0; k < nsteps; ++k)
            for (size_t g = 0; g < b->groups.size(); ++g) {
                int rc = fused_step(b->groups[g], k, nsteps, mixed, tms[g]);
                if (rc) return bcheck(b, b->groups[g], rc);
            }
    }
    for (des_ctx *c : b->groups) {
        int rc = fused_end(c, mixed);
        if (rc) return bcheck(b, c, rc);
    }
    for (int &s : b->msteps) s += nsteps;
    return DES_OK;
}

int des_batch_set_launch_mode(des_batch *b, int mode)
{
    if (!b || mode < 0 || mode > 1) return DES_ERR_INPUT;
    b->launch_mode = mode;
    return DES_OK;
}

int des_batch_sync(des_batch *b)
{
    if (!b) return DES_ERR_INPUT;
    for (des_ctx *ctx : b->groups) {
        hipError_t e = hipStreamSynchronize(ctx->stream);
        if (e != hipSuccess) return bfail(b, DES_ERR_HIP, hipGetErrorString(e));
    }
    for (des_ctx *ctx : b->groups) {
        std::vector<DevScalars> h(ctx->nmodels);
        hipError_t e = hipMemcpy(h.data(), ctx->s.scb, sizeof(DevScalars) * h.size(), hipMemcpyDeviceToHost);
        if (e != hipSuccess) return bfail(b, DES_ERR_HIP, hipGetErrorString(e));
        for (int k = 0; k < ctx->nmodels; ++k)
            if (h[k].err) {
                for (size_t m = 0; m < b->group_of.size(); ++m)
                    if (b->groups[b->group_of[m]] == ctx && b->index_of[m] == k)
                        return bfail(b, h[k].err, "device reported an error in model " + std::to_string(m));
            }
    }
    return DES_OK;
}

int des_batch_worst_quality(des_batch *b, double *q, int *worst_elem)
{
    if (!b || !q || !worst_elem) return DES_ERR_INPUT;
    for (des_ctx *ctx : b->groups) {
        int rc = flush_deferred(ctx);
        if (rc) return bcheck(b, ctx, rc);
        const int neb = nblocks(ctx->E);
        if (ctx->nd == 3) hipLaunchKernelGGL(k_quality_partial_b<3>, dim3(neb), dim3(BLOCK), 0, ctx->stream, ctx->s);
        else hipLaunchKernelGGL(k_quality_partial_b<2>, dim3(neb), dim3(BLOCK), 0, ctx->stream, ctx->s);
        if ((rc = check_launch(ctx))) return bcheck(b, ctx, rc);
        std::vector<double> hq(neb);
        std::vector<int> hi(neb);
        hipError_t e = hipMemcpyAsync(hq.data(), ctx->s.red, sizeof(double) * neb, hipMemcpyDeviceToHost, ctx->stream);
        if (e == hipSuccess)
            e = hipMemcpyAsync(hi.data(), ctx->s.redi, sizeof(int) * neb, hipMemcpyDeviceToHost, ctx->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
        if (e != hipSuccess) return bfail(b, DES_ERR_HIP, hipGetErrorString(e));
        for (size_t m = 0; m < b->group_of.size(); ++m) {
            if (b->groups[b->group_of[m]] != ctx) continue;
            const int k = b->index_of[m];
            // from q = 1 at element 0, the first strict minimum (geometry.cxx:458-512)
            double best = 1;
            int bi = 0;
            const int b0 = ctx->e_off[k] / BLOCK, b1 = nblocks(ctx->e_off[k] + ctx->e_cnt[k]);
            for (int blk = b0; blk < b1; ++blk)
                if (hq[blk] < best || (hq[blk] == best && best < 1 && hi[blk] < bi)) {
                    best = hq[blk];
                    bi = hi[blk];
                }
            q[m] = best;
            worst_elem[m] = bi;
        }
    }
    return DES_OK;
}

int des_batch_tuning(des_batch *b, int ndims, int *variant, double ms[DES_TUNE_MAXV])
{
    if (!b || !variant) return DES_ERR_INPUT;
    for (des_ctx *ctx : b->groups)
        if (ctx->nd == ndims) return des_get_tuning(ctx, variant, ms);
    return bfail(b, DES_ERR_INPUT, "no group of that dimension");
}

int des_batch_tuning_all(des_batch *b, int ndims, int precision, int variant[DES_NTUNE],
                         double ms[DES_TUNE_MAXV * DES_NTUNE])
{
    if (!b || !variant) return DES_ERR_INPUT;
    for (des_ctx *ctx : b->groups)
        if (ctx->nd == ndims) return des_get_tuning_all(ctx, precision, variant, ms);
    return bfail(b, DES_ERR_INPUT, "no group of that dimension");
}

int des_get_tuning(des_ctx *ctx, int *variant, double ms[DES_TUNE_MAXV])
{
    if (!ctx || !variant) return DES_ERR_INPUT;
    *variant = ctx->tune_var[0][TUNE_C];
    if (ms)
        for (int v = 0; v < DES_TUNE_MAXV; ++v) ms[v] = ctx->tune_ms[0][TUNE_C][v];
    return DES_OK;
}

}  // extern "C"

#include "des_adjoint.inc"
