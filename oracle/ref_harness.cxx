/*
 * ref_harness.cxx -- TEST INFRASTRUCTURE ONLY.
 *
 * Drives the reference's own step functions (fields.cxx, geometry.cxx, bc.cxx,
 * matprops.cxx, rheology.cxx, 3x3-C), compiled from the reference tree by
 * oracle/Makefile into oracle/_ref/.  It fills the reference's Param and
 * Variables from the oracle's state struct, runs the reference functions in
 * the order of the main loop (dynearthsol.cxx:325-361, update_mesh
 * dynearthsol.cxx:221-232), and copies the results back.  It is the checker of
 * the oracle; the product never loads it.
 *
 * Built once per dimension (-DTHREED for 3D) exactly like the reference.
 */
#include <chrono>
#include <cstring>
#include <map>
#include <vector>

#include "constants.hpp"
#include "parameters.hpp"
#include "matprops.hpp"
#include "fields.hpp"
#include "geometry.hpp"
#include "rheology.hpp"
#include "bc.hpp"
#include "utils.hpp"

#include "des_amd.h"

#ifdef USE_OMP
#include <omp.h>
#endif

#ifdef THREED
#define RFN(x) ref3_##x
#else
#define RFN(x) ref2_##x
#endif

/* same layout as the oracle's state struct (oracle/des_oracle.c) */
struct ostate {
    int nnode, nelem, nmat, markers_per_element;
    const int *conn;
    const unsigned *bcflag;
    const int *elemmarkers;
    const int *csr_off;
    const int *csr_ent;
    int nbfacet;
    const int *bf_bdry, *bf_elem, *bf_face;
    const double *bnormals;
    const double *edge_vectors;
    const int *edge_valid;
    double *coord, *coord0, *vel, *force, *temperature, *mass, *tmass, *volume_n, *ntmp;
    double *volume, *volume_old, *edvoldt, *plstrain, *delta_plstrain, *dpressure, *stressyy;
    double *strain_rate, *strain, *stress, *shpdx, *shpdy, *shpdz;
    const double *elem_visc, *elem_cohesion;
    double *ework;
    double dt, time, compensation_pressure;
    int steps;
    int status;
};

namespace {

void vec_param(double_vec &dst, const double *src, int n) { dst.assign(src, src + n); }

void fill_param(const des_params *p, Param &P)
{
    const int nm = p->nmat;
    P.mat.rheol_type = p->rheol_type;
    P.mat.phase_change_option = 0;
    P.mat.nmat = nm;
    P.mat.is_plane_strain = p->is_plane_strain != 0;
    P.mat.visc_min = p->visc_min;
    P.mat.visc_max = p->visc_max;
    P.mat.tension_max = p->tension_max;
    P.mat.therm_diff_max = p->therm_diff_max;
    vec_param(P.mat.rho0, p->rho0, nm);
    vec_param(P.mat.alpha, p->alpha, nm);
    vec_param(P.mat.bulk_modulus, p->bulk_modulus, nm);
    vec_param(P.mat.shear_modulus, p->shear_modulus, nm);
    vec_param(P.mat.visc_exponent, p->visc_exponent, nm);
    vec_param(P.mat.visc_coefficient, p->visc_coefficient, nm);
    vec_param(P.mat.visc_activation_energy, p->visc_activation_energy, nm);
    vec_param(P.mat.heat_capacity, p->heat_capacity, nm);
    vec_param(P.mat.therm_cond, p->therm_cond, nm);
    vec_param(P.mat.radiogenic_heating, p->radiogenic_heating, nm);
    vec_param(P.mat.pls0, p->pls0, nm);
    vec_param(P.mat.pls1, p->pls1, nm);
    vec_param(P.mat.cohesion0, p->cohesion0, nm);
    vec_param(P.mat.cohesion1, p->cohesion1, nm);
    vec_param(P.mat.friction_angle0, p->friction_angle0, nm);
    vec_param(P.mat.friction_angle1, p->friction_angle1, nm);
    vec_param(P.mat.dilation_angle0, p->dilation_angle0, nm);
    vec_param(P.mat.dilation_angle1, p->dilation_angle1, nm);

    P.control.gravity = p->gravity;
    P.control.characteristic_speed = p->characteristic_speed;
    P.control.inertial_scaling = p->inertial_scaling;
    P.control.dt_fraction = p->dt_fraction;
    P.control.fixed_dt = p->fixed_dt;
    P.control.damping_factor = p->damping_factor;
    P.control.damping_option = p->damping_option;
    P.control.ref_pressure_option = p->ref_pressure_option;
    P.control.surface_process_option = p->surface_process_option;
    P.control.surface_diffusivity = 0;
    P.control.is_quasi_static = p->is_quasi_static != 0;
    P.control.has_thermal_diffusion = p->has_thermal_diffusion != 0;
    P.control.has_hydration_processes = false;

    P.bc.surface_temperature = p->surface_temperature;
    P.bc.mantle_temperature = p->mantle_temperature;
    P.bc.has_fixed_bottom_temperature = p->has_fixed_bottom_temperature != 0;
    P.bc.winkler_delta_rho = p->winkler_delta_rho;
    P.bc.has_winkler_foundation = p->has_winkler_foundation != 0;
    P.bc.elastic_foundation_constant = p->elastic_foundation_constant;
    P.bc.has_elastic_foundation = p->has_elastic_foundation != 0;
    P.bc.has_water_loading = p->has_water_loading != 0;
    int *types[10] = {&P.bc.vbc_x0, &P.bc.vbc_x1, &P.bc.vbc_y0, &P.bc.vbc_y1, &P.bc.vbc_z0,
                      &P.bc.vbc_z1, &P.bc.vbc_n0, &P.bc.vbc_n1, &P.bc.vbc_n2, &P.bc.vbc_n3};
    double *vals[10] = {&P.bc.vbc_val_x0, &P.bc.vbc_val_x1, &P.bc.vbc_val_y0, &P.bc.vbc_val_y1,
                        &P.bc.vbc_val_z0, &P.bc.vbc_val_z1, &P.bc.vbc_val_n0, &P.bc.vbc_val_n1,
                        &P.bc.vbc_val_n2, &P.bc.vbc_val_n3};
    for (int i = 0; i < 10; ++i) {
        *types[i] = p->vbc_types[i];
        *vals[i] = p->vbc_values[i];
    }
    P.mesh.zlength = p->zlength;
    P.mesh.resolution = p->resolution;
    P.mesh.smallest_size = p->smallest_size;
    P.mesh.min_quality = p->min_quality;
    P.debug.dt = false;
}

/* The reference model built around copies of the oracle's arrays. */
struct RefModel {
    Param param;
    Variables var;
    std::vector<int> conn_store;

    RefModel(const des_params *p, const ostate *S)
    {
        fill_param(p, param);
        const int N = S->nnode, E = S->nelem;
        var.time = S->time;
        var.dt = S->dt;
        var.steps = S->steps;
        var.nnode = N;
        var.nelem = E;
        var.nseg = 0;
        var.max_vbc_val = p->max_vbc_val;
        var.compensation_pressure = S->compensation_pressure;

        var.coord = new array_t(N);
        std::memcpy(var.coord->data(), S->coord, sizeof(double) * N * NDIMS);
        var.connectivity = new conn_t(E);
        std::memcpy(var.connectivity->data(), S->conn, sizeof(int) * E * NODES_PER_ELEM);
        var.segment = NULL;
        var.segflag = NULL;
        var.regattr = NULL;
        var.bcflag = new uint_vec(S->bcflag, S->bcflag + N);
        for (int n = 0; n < N; ++n)
            for (int j = 0; j < nbdrytypes; ++j)
                if (S->bcflag[n] & (1u << j)) var.bnodes[j].push_back(n);
        for (int k = 0; k < S->nbfacet; ++k)
            var.bfacets[S->bf_bdry[k]].push_back(std::make_pair(S->bf_elem[k], S->bf_face[k]));
        for (int i = 0; i < nbdrytypes; ++i) {
            for (int d = 0; d < NDIMS; ++d) var.bnormals[i][d] = S->bnormals[i * NDIMS + d];
            var.vbc_types[i] = p->vbc_types[i];
            var.vbc_values[i] = p->vbc_values[i];
        }
        for (int i = 0; i < nbdrytypes; ++i)
            for (int j = 0; j < nbdrytypes; ++j)
                if (S->edge_valid && S->edge_valid[i * nbdrytypes + j]) {
                    double *s = new double[NDIMS];
                    for (int d = 0; d < NDIMS; ++d)
                        s[d] = S->edge_vectors[(i * nbdrytypes + j) * NDIMS + d];
                    var.edge_vectors[std::make_pair(i, j)] = s;
                }
        var.support = NULL;

        /* create_elem_groups (mesh.cxx:1724-1817) without the overlap warning */
        var.egroups.clear();
#ifdef USE_OMP
        {
            int ngroups = 2 * omp_get_max_threads();
            int el_per_group = E / ngroups;
            for (int i = 0; i < ngroups; i++) var.egroups.push_back(i * el_per_group);
            var.egroups.push_back(E);
        }
#else
        var.egroups.push_back(0);
        var.egroups.push_back(E);
#endif

        var.elemmarkers = new int_vec2D(E, std::vector<int>(S->nmat, 0));
        for (int e = 0; e < E; ++e)
            for (int m = 0; m < S->nmat; ++m)
                (*var.elemmarkers)[e][m] = S->elemmarkers ? S->elemmarkers[e * S->nmat + m]
                                                          : (m == 0 ? S->markers_per_element : 0);
        var.hydrous_elemmarkers = NULL;
        var.phch = NULL;

        /* allocate_variables (fields.cxx:10-48) */
        var.volume = new double_vec(S->volume, S->volume + E);
        var.volume_old = new double_vec(S->volume_old, S->volume_old + E);
        var.volume_n = new double_vec(S->volume_n, S->volume_n + N);
        var.mass = new double_vec(S->mass, S->mass + N);
        var.tmass = new double_vec(S->tmass, S->tmass + N);
        var.edvoldt = new double_vec(S->edvoldt, S->edvoldt + E);
        var.temperature = new double_vec(S->temperature, S->temperature + N);
        var.coord0 = new array_t(N);
        std::memcpy(var.coord0->data(), S->coord0, sizeof(double) * N * NDIMS);
        var.plstrain = new double_vec(S->plstrain, S->plstrain + E);
        var.delta_plstrain = new double_vec(S->delta_plstrain, S->delta_plstrain + E);
        var.vel = new array_t(N);
        std::memcpy(var.vel->data(), S->vel, sizeof(double) * N * NDIMS);
        var.strain = new tensor_t(E);
        std::memcpy(var.strain->data(), S->strain, sizeof(double) * E * NSTR);
        var.stress = new tensor_t(E);
        std::memcpy(var.stress->data(), S->stress, sizeof(double) * E * NSTR);
        var.stressyy = new double_vec(S->stressyy, S->stressyy + E);
        var.ntmp = new double_vec(S->ntmp, S->ntmp + N);
        var.dpressure = new double_vec(S->dpressure, S->dpressure + E);
        var.force = new array_t(N);
        std::memcpy(var.force->data(), S->force, sizeof(double) * N * NDIMS);
        var.strain_rate = new tensor_t(E);
        std::memcpy(var.strain_rate->data(), S->strain_rate, sizeof(double) * E * NSTR);
        var.shpdx = new shapefn(E);
        std::memcpy(var.shpdx->data(), S->shpdx, sizeof(double) * E * NODES_PER_ELEM);
#ifdef THREED
        var.shpdy = new shapefn(E);
        std::memcpy(var.shpdy->data(), S->shpdy, sizeof(double) * E * NODES_PER_ELEM);
#else
        var.shpdy = NULL;
#endif
        var.shpdz = new shapefn(E);
        std::memcpy(var.shpdz->data(), S->shpdz, sizeof(double) * E * NODES_PER_ELEM);
        var.mat = new MatProps(param, var);
    }

    void copy_back(ostate *S) const
    {
        const int N = S->nnode, E = S->nelem;
        std::memcpy(S->coord, var.coord->data(), sizeof(double) * N * NDIMS);
        std::memcpy(S->vel, var.vel->data(), sizeof(double) * N * NDIMS);
        std::memcpy(S->force, var.force->data(), sizeof(double) * N * NDIMS);
        std::memcpy(S->temperature, var.temperature->data(), sizeof(double) * N);
        std::memcpy(S->mass, var.mass->data(), sizeof(double) * N);
        std::memcpy(S->tmass, var.tmass->data(), sizeof(double) * N);
        std::memcpy(S->volume_n, var.volume_n->data(), sizeof(double) * N);
        std::memcpy(S->ntmp, var.ntmp->data(), sizeof(double) * N);
        std::memcpy(S->volume, var.volume->data(), sizeof(double) * E);
        std::memcpy(S->volume_old, var.volume_old->data(), sizeof(double) * E);
        std::memcpy(S->edvoldt, var.edvoldt->data(), sizeof(double) * E);
        std::memcpy(S->plstrain, var.plstrain->data(), sizeof(double) * E);
        std::memcpy(S->delta_plstrain, var.delta_plstrain->data(), sizeof(double) * E);
        std::memcpy(S->dpressure, var.dpressure->data(), sizeof(double) * E);
        std::memcpy(S->stressyy, var.stressyy->data(), sizeof(double) * E);
        std::memcpy(S->strain_rate, var.strain_rate->data(), sizeof(double) * E * NSTR);
        std::memcpy(S->strain, var.strain->data(), sizeof(double) * E * NSTR);
        std::memcpy(S->stress, var.stress->data(), sizeof(double) * E * NSTR);
        std::memcpy(S->shpdx, var.shpdx->data(), sizeof(double) * E * NODES_PER_ELEM);
#ifdef THREED
        std::memcpy(S->shpdy, var.shpdy->data(), sizeof(double) * E * NODES_PER_ELEM);
#endif
        std::memcpy(S->shpdz, var.shpdz->data(), sizeof(double) * E * NODES_PER_ELEM);
        S->dt = var.dt;
        S->time = var.time;
        S->steps = var.steps;
    }

    /* update_mesh (dynearthsol.cxx:221-232) */
    void update_mesh()
    {
        update_coordinate(var, *var.coord);
        surface_processes(param, var, *var.coord);
        var.volume->swap(*var.volume_old);
        compute_volume(*var.coord, *var.connectivity, *var.volume);
        compute_mass(param, var.egroups, *var.connectivity, *var.volume, *var.mat,
                     var.max_vbc_val, *var.volume_n, *var.mass, *var.tmass);
        compute_shape_fn(*var.coord, *var.connectivity, *var.volume, var.egroups,
                         *var.shpdx, *var.shpdy, *var.shpdz);
    }

    /* main loop body (dynearthsol.cxx:325-361) without phase changes and markers */
    void step()
    {
        var.steps++;
        var.time += var.dt;
        if (param.control.has_thermal_diffusion)
            update_temperature(param, var, *var.temperature, *var.ntmp);
        update_strain_rate(var, *var.strain_rate);
        compute_dvoldt(var, *var.ntmp);
        compute_edvoldt(var, *var.ntmp, *var.edvoldt);
        update_stress(var, *var.stress, *var.stressyy, *var.dpressure, *var.strain,
                      *var.plstrain, *var.delta_plstrain, *var.strain_rate);
        NMD_stress(var, *var.ntmp, *var.stress);
        update_force(param, var, *var.force);
        update_velocity(var, *var.vel);
        apply_vbcs(param, var, *var.vel);
        update_mesh();
        if (var.mat->rheol_type & MatProps::rh_elastic)
            rotate_stress(var, *var.stress, *var.strain);
        if (var.steps % 10 == 0) var.dt = compute_dt(param, var);
    }
};

} // namespace

extern "C" {

/* Run nsteps reference steps on a copy of S; results copied back.
   Returns the wall time of the step loop in seconds (setup excluded). */
double RFN(run)(const des_params *p, ostate *S, int nsteps)
{
    RefModel m(p, S);
    auto t0 = std::chrono::steady_clock::now();
    for (int k = 0; k < nsteps; ++k) m.step();
    auto t1 = std::chrono::steady_clock::now();
    m.copy_back(S);
    return std::chrono::duration<double>(t1 - t0).count();
}

/* Call one reference function by id on a copy of S; results copied back. */
int RFN(call)(const des_params *p, ostate *S, int which)
{
    RefModel m(p, S);
    Variables &var = m.var;
    Param &param = m.param;
    switch (which) {
    case 0: compute_volume(*var.coord, *var.connectivity, *var.volume); break;
    case 1:
        compute_mass(param, var.egroups, *var.connectivity, *var.volume, *var.mat,
                     var.max_vbc_val, *var.volume_n, *var.mass, *var.tmass);
        break;
    case 2:
        compute_shape_fn(*var.coord, *var.connectivity, *var.volume, var.egroups,
                         *var.shpdx, *var.shpdy, *var.shpdz);
        break;
    case 3: update_temperature(param, var, *var.temperature, *var.ntmp); break;
    case 4: update_strain_rate(var, *var.strain_rate); break;
    case 5: compute_dvoldt(var, *var.ntmp); break;
    case 6: compute_edvoldt(var, *var.ntmp, *var.edvoldt); break;
    case 7:
        update_stress(var, *var.stress, *var.stressyy, *var.dpressure, *var.strain,
                      *var.plstrain, *var.delta_plstrain, *var.strain_rate);
        break;
    case 8: NMD_stress(var, *var.ntmp, *var.stress); break;
    case 9: update_force(param, var, *var.force); break;
    case 10: update_velocity(var, *var.vel); break;
    case 11: apply_vbcs(param, var, *var.vel); break;
    case 12: m.update_mesh(); break;
    case 13: rotate_stress(var, *var.stress, *var.strain); break;
    case 14: var.dt = compute_dt(param, var); break;
    default: return 1;
    }
    m.copy_back(S);
    return 0;
}

double RFN(worst_quality)(const des_params *p, ostate *S, int *worst)
{
    RefModel m(p, S);
    return worst_elem_quality(*m.var.coord, *m.var.connectivity, *m.var.volume, *worst);
}

int RFN(nthreads)(void)
{
#ifdef USE_OMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

} // extern "C"

#ifdef THREED
#include "3x3-C/dsyevh3.h"
extern "C" int ref3_dsyevh3(const double *a6, double *w, double *q9)
{
    double A[3][3] = {{a6[0], a6[3], a6[4]}, {0, a6[1], a6[5]}, {0, 0, a6[2]}};
    double Q[3][3];
    int r = dsyevh3(A, Q, w);
    std::memcpy(q9, Q, sizeof Q);
    return r;
}
#endif

#ifdef DES_REF_IO
/* Output::write (output.cxx:79-208) and Output::write_checkpoint
   (output.cxx:254-276) of the reference, on the reference model built from S:
   "<modelname>.save.<frame>", "<modelname>.chkpt.<frame>" and a line of
   "<modelname>.info" (frame 0 creates it).  The boundary segments are the
   caller's (nseg x NDIMS node indices and their flags); no marker sets. */
#include "output.hpp"
extern "C" int RFN(write_output)(const des_params *p, ostate *S, const char *modelname, int frame, int nseg,
                                 const int *segment, const int *segflag)
{
    RefModel m(p, S);
    m.param.sim.modelname = modelname;
    m.param.sim.is_outputting_averaged_fields = false;
    m.param.sim.has_marker_output = false;
    m.param.mesh.quality_check_step_interval = 100;
    m.var.nseg = nseg;
    m.var.segment = new segment_t(nseg);
    m.var.segflag = new segflag_t(nseg);
    if (nseg) {
        std::memcpy(m.var.segment->data(), segment, sizeof(int) * nseg * NDIMS);
        std::memcpy(m.var.segflag->data(), segflag, sizeof(int) * nseg);
    }
    Output out(m.param, 0.0, frame);
    out.write_checkpoint(m.param, m.var);      // before the frame, as the main loop does (dynearthsol.cxx:381-384)
    out.write(m.var, false);
    return 0;
}
#endif
