/*
 * des_amd.h -- C-ABI of the MI355X-native DynEarthSol explicit time step.
 *
 * The reference (luh1202/DynEarthSol) has no FFI: its per-step physics is a set
 * of free C++ functions `f(const Param&, const Variables&, <out arrays>)`
 * called in sequence by the main loop (dynearthsol.cxx:324-406).  This header
 * replaces that surface with plain C: one opaque context per model, host
 * arrays in the reference's own AoS layout (array2d.hpp:9-104) at the upload /
 * download boundary, and one entry point per reference function plus fused
 * multi-step drivers.  Every entry point returns a status code whose numbering
 * follows the reference's exit codes (README:139-144) instead of calling
 * std::exit.
 *
 * Dimensions are a run-time choice (p->ndims = 2 or 3) where the reference
 * fixes them at compile time (constants.hpp:11-23, Makefile:233-238).
 */
#ifndef DES_AMD_H
#define DES_AMD_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define DES_MAX_MAT 8   /* max material types per model (Mat::nmat) */
#define DES_NBDRY 10    /* boundary types x0,x1,y0,y1,z0,z1,n0..n3 (constants.hpp:26-36) */

/* Status codes (README:139-144; rheology.cxx:636-639; geometry.cxx:304-308). */
enum des_status {
    DES_OK = 0,
    DES_ERR_INPUT = 1,     /* invalid parameter, unknown rheology or damping option */
    DES_ERR_IO = 2,
    DES_ERR_MESH = 10,     /* inconsistent mesh (bad connectivity, node out of range), or nelem * (ndims + 1) >= 2^31 */
    DES_ERR_RUNTIME = 11,  /* dt <= 0, non-finite coordinates */
    DES_ERR_ASSERT = 12,   /* missing edge vector of a boundary intersection */
    DES_ERR_HIP = 100,     /* HIP runtime failure (message in des_last_error) */
    DES_ERR_STATE = 101    /* call out of order (e.g. run before upload) */
};

/* Rheology bit flags (matprops.hpp:47-53). */
enum des_rheology {
    DES_RH_ELASTIC = 1,
    DES_RH_VISCOUS = 2,
    DES_RH_PLASTIC = 4,
    DES_RH_MAXWELL = 3,
    DES_RH_EP = 5,
    DES_RH_EVP = 7
};

/* Precision of the fused step (des_run). */
enum des_precision {
    DES_FP64 = 0,   /* every field and operation in double */
    DES_MIXED = 1   /* stress deviator / strain / strain rate / rheology arithmetic in
                       float; coordinates, volumes, mass, pressure, force accumulation
                       and the eigen-return in double */
};

/*
 * Model parameters: the subset of the reference's Param (parameters.hpp:39-221)
 * that the explicit step reads, with the reference's names.  Per-material
 * vectors hold nmat entries (the reference's one-entry broadcast,
 * matprops.cxx:257-268, is expanded by the caller).
 */
typedef struct des_params {
    int ndims;                    /* 2 or 3 */
    /* Mat (parameters.hpp:158-186) */
    int rheol_type;               /* enum des_rheology */
    int nmat;
    int is_plane_strain;
    double visc_min, visc_max, tension_max, therm_diff_max;
    double rho0[DES_MAX_MAT], alpha[DES_MAX_MAT];
    double bulk_modulus[DES_MAX_MAT], shear_modulus[DES_MAX_MAT];
    double visc_exponent[DES_MAX_MAT], visc_coefficient[DES_MAX_MAT];
    double visc_activation_energy[DES_MAX_MAT];
    double heat_capacity[DES_MAX_MAT], therm_cond[DES_MAX_MAT];
    double radiogenic_heating[DES_MAX_MAT];
    double pls0[DES_MAX_MAT], pls1[DES_MAX_MAT];
    double cohesion0[DES_MAX_MAT], cohesion1[DES_MAX_MAT];
    double friction_angle0[DES_MAX_MAT], friction_angle1[DES_MAX_MAT];
    double dilation_angle0[DES_MAX_MAT], dilation_angle1[DES_MAX_MAT];
    /* Control (parameters.hpp:79-96) */
    double gravity, characteristic_speed, inertial_scaling, dt_fraction, fixed_dt;
    double damping_factor;
    int damping_option, ref_pressure_option, is_quasi_static, has_thermal_diffusion;
    int surface_process_option;   /* only 0 (no surface process) is on the device path */
    /* BC (parameters.hpp:98-135) */
    double surface_temperature, mantle_temperature;
    int has_fixed_bottom_temperature;
    int has_winkler_foundation;
    double winkler_delta_rho;
    int has_elastic_foundation;
    double elastic_foundation_constant;
    int has_water_loading;
    int vbc_types[DES_NBDRY];     /* Variables::vbc_types (dynearthsol.cxx:43-52) */
    double vbc_values[DES_NBDRY]; /* Variables::vbc_values (dynearthsol.cxx:54-63) */
    /* Mesh (parameters.hpp:54-77): only what the step and its checks read */
    double zlength;               /* Winkler foundation depth (bc.cxx:532) */
    double resolution, smallest_size, min_quality;
    /* Derived by the caller as the reference's init_var does (dynearthsol.cxx:36-39) */
    double max_vbc_val;
} des_params;

/*
 * Mesh and setup artefacts (mesh.cxx:1546-1721, bc.cxx:87-176), host AoS.
 * Boundary facets are var.bfacets[0..9] concatenated in boundary order.
 */
typedef struct des_mesh {
    int nnode, nelem;
    const int *connectivity;      /* nelem x (ndims+1) */
    const double *coord;          /* nnode x ndims: initial coordinates */
    const unsigned *bcflag;       /* nnode */
    int nbfacet;
    const int *bfacet_bdry;       /* nbfacet: boundary index 0..9 */
    const int *bfacet_elem;       /* nbfacet: element owning the facet */
    const int *bfacet_face;       /* nbfacet: local facet number (NODE_OF_FACET) */
    const double *bnormals;       /* DES_NBDRY x ndims */
    const double *edge_vectors;   /* DES_NBDRY x DES_NBDRY x ndims, entry [i][j], i<j */
    const int *edge_valid;        /* DES_NBDRY x DES_NBDRY: 1 where the reference map has the pair */
    const int *elemmarkers;       /* nelem x nmat marker counts, or NULL: every element
                                     holds markers_per_element markers of material 0 */
    int markers_per_element;
} des_mesh;

/*
 * Field arrays, host AoS in the reference's layout (parameters.hpp:229-279).
 * Any NULL pointer is skipped by upload and download.
 */
typedef struct des_fields {
    /* node fields */
    double *coord, *coord0, *vel, *force;           /* nnode x ndims */
    double *temperature, *mass, *tmass, *volume_n, *ntmp;   /* nnode */
    /* element fields */
    double *volume, *volume_old, *edvoldt, *plstrain, *delta_plstrain;
    double *dpressure, *stressyy;                   /* nelem */
    double *strain_rate, *strain, *stress;          /* nelem x nstr */
    double *shpdx, *shpdy, *shpdz;                  /* nelem x (ndims+1) */
    /* per-element material fields (extension, no counterpart in the reference):
       viscosity replacing the power law of MatProps::visc, and cohesion replacing
       cohesion0 in MatProps::plastic_weakening */
    double *elem_visc, *elem_cohesion;              /* nelem */
} des_fields;

typedef struct des_ctx des_ctx;

/* ---- lifetime ---------------------------------------------------------- */

/* Create a model on the current HIP device: uploads the mesh, builds the
   node->element CSR (create_support, mesh.cxx:1707-1721) and the per-node
   boundary-facet lists.  Fields start at zero, as allocate_variables leaves
   them (fields.cxx:10-48).  On failure *out may hold a partly built context
   whose des_last_error explains it: release it with des_destroy. */
int des_create(const des_params *p, const des_mesh *m, des_ctx **out);
void des_destroy(des_ctx *ctx);
/* Device numbering of the meshes given to later des_create / des_batch_create
   calls (process-wide; no counterpart in the reference): 1 (default) stores
   elements and nodes in Hilbert-curve order of their positions, 0 in the host's
   order; 2 and 3 keep the elements' curve order and number the nodes by the
   order in which those elements first touch them (2) or by the mean device
   index of their elements (3).  Every array crossing this interface stays in
   the host's numbering and every index returned is a host index; each node
   sums its elements in ascending host order, so results do not depend on the
   setting. */
void des_set_mesh_order(int mode);
/* Block order of the fused schedule's node passes (process-wide; no
   counterpart in the reference): 1 (default) walks passes B, D and F from the
   top of each XCD's block range down, so that each pass starts on the data the
   pass before it wrote last; 0 walks every pass upwards.  Each element and
   node is computed the same way in either order, so results are the same. */
void des_set_sweep_order(int mode);
/* Layout of the fused FP64 schedule's force contributions (process-wide; no
   counterpart in the reference): 0 (default) node-major CSR slots, written
   scattered by pass E and read contiguously by pass F, pass E finding each
   slot from the start of the node's list and the entry's position in it (one
   byte per entry; full 4-byte positions when a node list is longer than 256);
   1 element-major records (an element's contributions to its nodes in one
   128-byte line in 3D, 64 bytes in 2D), written coalesced by pass E and
   gathered by pass F; 2 the node-major slots of 0 always found by their full
   positions (this also for the mixed schedule's FP32 slots, which otherwise
   are found as in 0).  Pass F sums a node's contributions in its CSR order
   every way: the same results. */
void des_set_force_layout(int mode);
const char *des_last_error(const des_ctx *ctx);

int des_upload_fields(des_ctx *ctx, const des_fields *f);
/* Completes any deferred work of the fused driver first, so downloaded
   arrays equal the reference's Variables after the same number of steps. */
int des_download_fields(des_ctx *ctx, des_fields *f);
/* The same with every des_fields pointer in device memory of the context's
   GPU (still the host's row order and AoS layout): no host copy, for callers
   whose arrays live on the device (a torch.autograd binding's tensors).
   Single-model contexts; the context's stream is synchronized on return, so
   the caller's own stream must have finished writing the inputs. */
int des_upload_fields_device(des_ctx *ctx, const des_fields *f);
int des_download_fields_device(des_ctx *ctx, des_fields *f);

/* Output::write (output.cxx:79-208) of the instantaneous fields: writes
   "<modelname>.save.<frame %06d>" in the reference's binary format
   (binaryio.cxx:16-36) and appends the frame's line to "<modelname>.info"
   (created at frame 0).  Returns 11 when a coordinate is NaN or infinite. */
int des_write_frame(des_ctx *ctx, const char *modelname, int frame, double run_time);

/* Averaged fields (Output::average_fields, output.cxx:211-251), on the
   device.  des_set_averaging: average_interval = the reference's
   mesh.quality_check_step_interval when sim.is_outputting_averaged_fields is
   on, 0 to turn it off.  des_run then accumulates after every step (FP64
   runs); a per-function driver calls des_average_fields after each step, as
   the main loop does (dynearthsol.cxx:363-364).  des_write_frame_averaged is
   Output::write (output.cxx:79-208) of a regular frame with them: "velocity
   averaged", "stress averaged", strain-rate and plastic strain-rate averaged
   over the interval, and the interval's mean dt in the .info line. */
int des_set_averaging(des_ctx *ctx, int average_interval);
int des_average_fields(des_ctx *ctx);
int des_write_frame_averaged(des_ctx *ctx, const char *modelname, int frame, double run_time);

/* A named block of bytes written into a checkpoint as it is. */
typedef struct {
    const char *name;
    const void *data;
    size_t bytes;
} des_named_array;

/* Output::write_checkpoint (output.cxx:254-276): writes
   "<modelname>.chkpt.<frame %06d>" in the binary format with "time
   compensation_pressure", the host's boundary segments ("segment", nseg x
   ndims, and "segflag", nseg) as given, "volume_old" from the device and, in
   plane strain, "stressyy" -- then the nextra host-owned arrays in order (the
   marker sets' "<name> size", ".eta", ".elem", ".mattype", ".id" of
   MarkerSet::write_chkpt_file, markerset.cxx:386-398, which the host keeps).
   With the frame of the same number from des_write_frame this is what
   restart() (dynearthsol.cxx:104-218) reads. */
int des_write_checkpoint(des_ctx *ctx, const char *modelname, int frame, int nseg, const int *segment,
                         const int *segflag, int nextra, const des_named_array *extra);

/* Marker counts per element and material (var.elemmarkers, nelem x nmat),
   e.g. after the host's phase changes (phasechanges.cxx:232-258). */
int des_upload_elemmarkers(des_ctx *ctx, const int *counts);

/* Scalar state: var.dt, var.time, var.steps, var.compensation_pressure. */
int des_set_scalars(des_ctx *ctx, double dt, double time, int steps,
                    double compensation_pressure);
/* des_get_scalars waits for the queued work and, like des_sync, returns the
   first device error the steps detected (the scalars are filled either way). */
int des_get_scalars(des_ctx *ctx, double *dt, double *time, int *steps);

/* ---- per-function entry points (same semantics as the named function) -- */

int des_compute_volume(des_ctx *ctx);       /* geometry.cxx:86-108 */
int des_compute_mass(des_ctx *ctx);         /* geometry.cxx:313-365 */
int des_compute_shape_fn(des_ctx *ctx);     /* geometry.cxx:368-455 */
/* init()'s derived-state block: compute_volume, volume_old = volume,
   compute_mass, compute_shape_fn (dynearthsol.cxx:86-91). */
int des_init_derived(des_ctx *ctx);
int des_update_temperature(des_ctx *ctx);   /* fields.cxx:96-163 */
int des_update_strain_rate(des_ctx *ctx);   /* fields.cxx:166-232 */
int des_compute_dvoldt(des_ctx *ctx);       /* geometry.cxx:111-155 (into ntmp) */
int des_compute_edvoldt(des_ctx *ctx);      /* geometry.cxx:158-179 (from ntmp) */
int des_update_stress(des_ctx *ctx);        /* rheology.cxx:510-646 */
int des_nmd_stress(des_ctx *ctx);           /* geometry.cxx:182-236 */
int des_update_force(des_ctx *ctx);         /* fields.cxx:291-340 with bc.cxx:493-566 */
int des_update_velocity(des_ctx *ctx);      /* fields.cxx:343-355 */
int des_apply_vbcs(des_ctx *ctx);           /* bc.cxx:179-490 */
int des_update_mesh(des_ctx *ctx);          /* dynearthsol.cxx:221-232 */
int des_rotate_stress(des_ctx *ctx);        /* fields.cxx:411-483 */
int des_compute_dt(des_ctx *ctx, double *dt_out);   /* geometry.cxx:239-310; sets var.dt */
/* The tail of remesh() (remeshing.cxx:2006-2029), after the host has built
   the new mesh, created a context on it and uploaded the interpolated fields:
   volume, volume_old = volume, mass / thermal mass / nodal volume, shape
   functions and, for remeshing options 1, 2 and 11, the bottom nodes' coord0
   reset to -zlength.  dt and the other scalars carry over (des_set_scalars). */
int des_remesh_finish(des_ctx *ctx, int remeshing_option);
int des_worst_quality(des_ctx *ctx, double *q, int *worst_elem); /* geometry.cxx:499-512 */
/* des_worst_quality without allocating device memory: no counterpart in the
   reference; for callers that cannot allocate (HIP graph capture, tight
   serving loops).  The reduction's partial results (1024 quality / element
   pairs, 12 KiB) live in a static device array that each GPU holds until the
   process ends and that every call overwrites; calls that overlap on one GPU
   race for it.  Same results as des_worst_quality; single-model contexts only
   (a batched group's context is refused with DES_ERR_INPUT). */
int des_worst_quality_noalloc(des_ctx *ctx, double *q, int *worst_elem);
/* bad_mesh_quality (remeshing.cxx:1882-1934): *reason = 3 (an element smaller
   than mesh.smallest_size of the resolution), 2 (a bottom node drifted more than
   max_boundary_distortion * resolution, remeshing options 1, 2, 11), 1 (worst
   quality, cube-rooted in 3D, below mesh.min_quality) or 0; *index names the
   element or node. */
int des_bad_mesh_quality(des_ctx *ctx, int remeshing_option, double max_boundary_distortion,
                         int *reason, int *index);

/* ---- step drivers --------------------------------------------------------- */

/* One step of the main loop, dynearthsol.cxx:325-361 (no phase changes, markers,
   output or remeshing), through the per-function kernels above. */
int des_step(des_ctx *ctx);

/* nsteps steps of the same loop through the fused six-pass schedule, with
   compute_dt every 10 steps as the reference (dynearthsol.cxx:349-361).
   The last step's mesh update and stress rotation stay deferred until a call
   reads or writes the state, which completes them with the per-function
   kernels.  A des_run right after a des_run of the same precision carries on
   as the next step of the last one and completes them inside its first step:
   runs of 7 + 11 + 7 steps end bitwise where one run of 25 ends (in FP64 also
   where the same runs end with the state read between them). */
int des_run(des_ctx *ctx, int nsteps, int precision);

/* des_run with the duration of each fused pass measured by HIP events on the
   context's stream: ms[k] accumulates the milliseconds and counts[k] the
   launches of pass k (0..4 = passes A..E, 5 = facet tractions, 6 = pass F,
   7 = compute_dt, 8 = pass C's deferred Mohr-Coulomb returns; 9 unused).
   Measurement aid with no counterpart in the reference. */
int des_run_timed(des_ctx *ctx, int nsteps, int precision, double ms[10], int counts[10]);

/* How des_run launches its steps (no counterpart in the reference): 0
   (default) one kernel launch per fused pass, with the autotuner's variants;
   1 one persistent cooperative launch for the whole run, its workgroups as
   many as fit the device at once, grid barriers between the passes and the
   variants fixed (direct node-record gathers, 256-thread node blocks).  Both
   compute the same values in the same order: results are bitwise equal.
   Runs with averaged fields on, and des_run_timed, launch per pass. */
int des_set_launch_mode(des_ctx *ctx, int mode);

/* Wait for all work queued on the context's stream; returns the first device
   error the step detected (dt <= 0, missing edge vector, unknown rheology).
   des_run only queues work: its device-side failures surface here (or at the
   next compute_dt, every 10 steps). */
int des_sync(des_ctx *ctx);

/* The autotuner of des_run (no counterpart in the reference).  The tunable
   fused passes and their variants -- every variant performs the same
   arithmetic in the same order, so results do not depend on the choice:
     DES_TUNE_A, DES_TUNE_E: 0 the element block's node records staged in LDS,
                             1 gathered directly, 2 pass A direct with at
                             least 6 waves per SIMD, FP64 pass E staged with
                             at least 8 (the mixed pass E has 0 and 1);
     DES_TUNE_C:             pass C (FP64 and mixed) with a minimum of 1..4
                             waves per SIMD (0..3), or of 1 or 2 with the
                             element block's node records staged in LDS
                             (4, 5);
     DES_TUNE_D, DES_TUNE_F: node blocks of 256, 128 or 64 threads with 4
                             gathered entries in flight per node (0..2), or
                             256 threads with 8 or 16 in flight (3, 4);
     DES_TUNE_B:             4, 8 or 16 element records in flight per node
                             (0..2).
   Until tuned a pass runs variant 0, except D and F (3) and B (1).  The first
   10 steps of a run of 16 or more steps time every variant of each pass not
   tuned yet (variant k mod the variant count at step k) and keep the fastest
   mean; FP64 and mixed-precision runs tune separately, and each dimension's
   group of a batch tunes on its own. */
#define DES_NTUNE 6
#define DES_TUNE_A 0
#define DES_TUNE_C 1
#define DES_TUNE_D 2
#define DES_TUNE_E 3
#define DES_TUNE_F 4
#define DES_TUNE_B 5
#define DES_TUNE_MAXV 6     /* variants per pass at most */
/* variant[p]: the choice for pass p at this precision (DES_FP64 or DES_MIXED),
   -1 while untuned; ms[DES_TUNE_MAXV * p + v]: the mean milliseconds measured
   for variant v of pass p (0 where not measured). */
int des_get_tuning_all(des_ctx *ctx, int precision, int variant[DES_NTUNE],
                       double ms[DES_TUNE_MAXV * DES_NTUNE]);
/* Fix the choices (each -1 to let the next long run tune that pass). */
int des_set_tuning_all(des_ctx *ctx, int precision, const int variant[DES_NTUNE]);
/* The FP64 pass-C entry of the above: variant 0..5 or -1, ms per variant. */
int des_get_tuning(des_ctx *ctx, int *variant, double ms[DES_TUNE_MAXV]);
int des_set_tuning(des_ctx *ctx, int variant);

/* ---- differentiated run (no counterpart in the reference) ------------------

   Reverse-mode gradient of a loss of the final coordinates with respect to
   the per-element viscosity (elem_visc) and cohesion (elem_cohesion), through
   the fused step, for 3D models with one material, alpha = 0, fixed dt,
   damping option 0 or 1, no boundary tractions and axis-aligned velocity
   boundaries; rheology elastic, maxwell, elasto-plastic or
   elasto-visco-plastic (maxwell and elasto-visco-plastic need elem_visc).
   Other configurations return 1 (DES_ERR_INPUT). */

/* nsteps fused FP64 steps from the current state, keeping a device checkpoint
   every `window` steps and the records of one window ((14 nnode + 8 nelem)
   doubles per step, window + 1 steps; the last window's are recorded here);
   leaves the context at the final state.  window = 0 tapes the run instead:
   every step's records stay on the device (nsteps + 1 steps), and the
   backward recomputes nothing.  Records beyond 80% of free device memory are
   refused with DES_ERR_INPUT. */
int des_adj_forward(des_ctx *ctx, int nsteps, int window);

/* coord_bar (nnode x ndims, AoS) = d loss / d final coordinates.  Recomputes
   each window from its checkpoint (the first call takes the last window's
   records from the forward; taped run: reads the tape) and runs the reverse
   sweep; may be called again for another coord_bar; writes
   visc_bar[e] = d loss / d elem_visc[e] (0 where the viscosity is clamped to
   the min / max) and cohesion_bar[e] = d loss / d elem_cohesion[e] (either
   may be null).  The context is left at the final state of the run. */
int des_adj_backward(des_ctx *ctx, const double *coord_bar, double *visc_bar, double *cohesion_bar);
/* des_adj_backward with coord_bar, visc_bar and cohesion_bar in device memory
   (host row order). */
int des_adj_backward_device(des_ctx *ctx, const double *coord_bar, double *visc_bar, double *cohesion_bar);

/* Frees the checkpoints and the reverse sweep's buffers. */
void des_adj_release(des_ctx *ctx);

/* ---- batched models ----------------------------------------------------------

   Many independent models stepping together, as a caller would run many
   instances of the reference's main loop (dynearthsol.cxx:316-361) one after
   another; the reference itself has no batch interface.  Models are numbered
   0..nmodels-1 in the order given; each keeps its own parameters, mesh, fields
   and scalars.  Models of one dimension must share nmat and
   markers_per_element.  All models advance together: des_batch_run steps every
   model nsteps times, with compute_dt every 10 steps per model. */
typedef struct des_batch des_batch;

/* On failure *out may hold a partly built batch (des_batch_last_error):
   release it with des_batch_destroy. */
int des_batch_create(int nmodels, const des_params *params, const des_mesh *meshes, des_batch **out);
void des_batch_destroy(des_batch *b);
const char *des_batch_last_error(const des_batch *b);
int des_batch_upload_fields(des_batch *b, int model, const des_fields *f);
int des_batch_download_fields(des_batch *b, int model, des_fields *f);
int des_batch_set_scalars(des_batch *b, int model, double dt, double time, int steps,
                          double compensation_pressure);
/* As des_get_scalars: also returns the model's first pending device error. */
int des_batch_get_scalars(des_batch *b, int model, double *dt, double *time, int *steps);
int des_batch_init_derived(des_batch *b);            /* des_init_derived for every model */
int des_batch_apply_vbcs(des_batch *b);              /* des_apply_vbcs for every model */
int des_batch_compute_dt(des_batch *b, double *dt);  /* dt[model] */
int des_batch_run(des_batch *b, int nsteps, int precision);
/* des_set_launch_mode for a batch: with mode 1 every model of both dimension
   groups runs in one persistent launch per des_batch_run (each pass walks the
   3D models' blocks, then the 2D models'). */
int des_batch_set_launch_mode(des_batch *b, int mode);
int des_batch_sync(des_batch *b);
int des_batch_worst_quality(des_batch *b, double *q, int *worst_elem);   /* per model */
int des_batch_tuning(des_batch *b, int ndims, int *variant, double ms[DES_TUNE_MAXV]);
int des_batch_tuning_all(des_batch *b, int ndims, int precision, int variant[DES_NTUNE],
                         double ms[DES_TUNE_MAXV * DES_NTUNE]);

#ifdef __cplusplus
}
#endif

#endif /* DES_AMD_H */
