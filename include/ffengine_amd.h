/*
 * ffengine_amd.h -- C ABI of the MI355X FLIP engine (libffengine.so).
 *
 * The entry points carry the names, argument meanings and error convention of
 * the reference engine's C bindings (rlguy/Blender-FLIP-Fluids 1.8.5,
 * src/engine/c_bindings/fluidsimulation_c.cpp and cbindings.cpp), so the
 * reference's ctypes layer (src/engine/ffengine/) can load this library in
 * place of libffengine.so for the per-substep path.  Citations are
 * file:line into that tree.
 *
 * Error convention (cbindings.h:48-155, cbindings.cpp:35-43): every call
 * except FluidSimulation_destroy and CBindings_get_error_message ends with
 * int *err; *err = 1 on success and 0 on failure, in which case
 * CBindings_get_error_message() returns the message of the exception (one
 * process-wide buffer, like the reference).  No C++ exception crosses the ABI.
 *
 * Threading: one caller thread per FluidSimulation; calls are not re-entrant.
 */
#ifndef FFENGINE_AMD_H
#define FFENGINE_AMD_H

#ifdef __cplusplus
extern "C" {
#endif

/* Opaque handle (the reference returns FluidSimulation*). */
typedef struct FluidSimulation FluidSimulation;

/* Vector3_t (c_bindings/vector3_c.h) */
typedef struct FFVector3 {
    float x, y, z;
} FFVector3;

/* FluidSimulationMarkerParticleData (fluidsimulation.h:153-157): packed float
 * xyz positions and velocities, passed by value. */
typedef struct FluidSimulationMarkerParticleData {
    int size;
    char *positions;
    char *velocities;
} FluidSimulationMarkerParticleData;

/* FluidSimulationMarkerParticleAffineData (fluidsimulation.h:159-164): packed
 * float xyz AFFINEX, AFFINEY and AFFINEZ vectors of one load batch. */
typedef struct FluidSimulationMarkerParticleAffineData {
    int size;
    char *affineX;
    char *affineY;
    char *affineZ;
} FluidSimulationMarkerParticleAffineData;

/* FluidSimulationMeshStats / TimingStats / FrameStats (fluidsimulation.h:62-151),
 * same field order and types, returned by value. */
typedef struct FFMeshStats {
    int enabled;
    int vertices;
    int triangles;
    unsigned int bytes;
} FFMeshStats;
typedef struct FFTimingStats {
    double total, mesh, advection, particles, pressure, diffuse, viscosity, objects;
} FFTimingStats;
typedef struct FFFrameStats {
    int frame;
    int substeps;
    double deltaTime;
    int fluidParticles;
    int diffuseParticles;
    int performanceScore;
    int pressureSolverEnabled;
    int pressureSolverSuccess;
    double pressureSolverError;
    int pressureSolverIterations;
    int pressureSolverMaxIterations;
    int viscositySolverEnabled;
    int viscositySolverSuccess;
    double viscositySolverError;
    int viscositySolverIterations;
    int viscositySolverMaxIterations;
    FFMeshStats mesh[50]; /* surface ... forcefield, in the reference's order */
    FFTimingStats timing;
} FFFrameStats;
typedef FFFrameStats FluidSimulationFrameStats;

/* ---- lifecycle ---- */
FluidSimulation *FluidSimulation_new_from_dimensions(int isize, int jsize, int ksize, double dx,
                                                     int *err); /* fluidsimulation_c.cpp:56-69 */
void FluidSimulation_destroy(FluidSimulation *obj);             /* fluidsimulation_c.cpp:71-73 */
void FluidSimulation_initialize(FluidSimulation *obj, int *err); /* fluidsimulation_c.cpp:97-99 */
int FluidSimulation_is_initialized(FluidSimulation *obj, int *err); /* fluidsimulation_c.cpp:101-105 */
void FluidSimulation_update(FluidSimulation *obj, double dt, int *err); /* fluidsimulation_c.cpp:115-119 */
int FluidSimulation_get_current_frame(FluidSimulation *obj, int *err); /* fluidsimulation_c.cpp:121-125 */
void FluidSimulation_set_current_frame(FluidSimulation *obj, int frameno, int *err); /* :127-131 */
int FluidSimulation_is_current_frame_finished(FluidSimulation *obj, int *err); /* :148-153 */
double FluidSimulation_get_cell_size(FluidSimulation *obj, int *err);          /* :155-159 */
void FluidSimulation_get_grid_dimensions(FluidSimulation *obj, int *i, int *j, int *k,
                                         int *err);                            /* :161-166 */

/* ---- settings on the substep path ---- */
double FluidSimulation_get_density(FluidSimulation *obj, int *err);            /* :253-257 */
void FluidSimulation_set_density(FluidSimulation *obj, double density, int *err); /* :259-265 */
int FluidSimulation_get_pressure_solver_max_iterations(FluidSimulation *obj, int *err); /* :318-323 */
void FluidSimulation_set_pressure_solver_max_iterations(FluidSimulation *obj, int n, int *err); /* :325-331 */
void FluidSimulation_enable_surface_reconstruction(FluidSimulation *obj, int *err);  /* :607-612 */
void FluidSimulation_disable_surface_reconstruction(FluidSimulation *obj, int *err); /* :614-619 */
int FluidSimulation_is_surface_reconstruction_enabled(FluidSimulation *obj, int *err); /* :621-626 */
int FluidSimulation_get_max_thread_count(FluidSimulation *obj, int *err);       /* :2583-2589 */
void FluidSimulation_set_max_thread_count(FluidSimulation *obj, int n, int *err); /* :2591-2597 */
void FluidSimulation_add_body_force(FluidSimulation *obj, double fx, double fy, double fz,
                                    int *err);                                 /* :2599-2605 */
FFVector3 FluidSimulation_get_constant_body_force(FluidSimulation *obj, int *err); /* :2607-2613 */
void FluidSimulation_reset_body_force(FluidSimulation *obj, int *err);         /* :2615-2619 */
int FluidSimulation_get_CFL_condition_number(FluidSimulation *obj, int *err);  /* :2825-2831 */
void FluidSimulation_set_CFL_condition_number(FluidSimulation *obj, int n, int *err); /* :2833-2839 */
int FluidSimulation_get_min_time_steps_per_frame(FluidSimulation *obj, int *err); /* :2878-2884 */
void FluidSimulation_set_min_time_steps_per_frame(FluidSimulation *obj, int n, int *err); /* :2886-2892 */
int FluidSimulation_get_max_time_steps_per_frame(FluidSimulation *obj, int *err); /* :2894-2900 */
void FluidSimulation_set_max_time_steps_per_frame(FluidSimulation *obj, int n, int *err); /* :2902-2908 */
void FluidSimulation_enable_extreme_velocity_removal(FluidSimulation *obj, int *err); /* :2952-2957 */
void FluidSimulation_disable_extreme_velocity_removal(FluidSimulation *obj, int *err); /* :2959-2964 */
int FluidSimulation_is_extreme_velocity_removal_enabled(FluidSimulation *obj, int *err); /* :2966-2971 */
void FluidSimulation_set_velocity_transfer_method_FLIP(FluidSimulation *obj, int *err); /* :2973-2978 */
void FluidSimulation_set_velocity_transfer_method_APIC(FluidSimulation *obj, int *err); /* :2980-2985 */
int FluidSimulation_is_velocity_transfer_method_FLIP(FluidSimulation *obj, int *err); /* :2987-2992 */
int FluidSimulation_is_velocity_transfer_method_APIC(FluidSimulation *obj, int *err); /* :2994-2999 */
double FluidSimulation_get_PICFLIP_ratio(FluidSimulation *obj, int *err);      /* :3001-3005 */
void FluidSimulation_set_PICFLIP_ratio(FluidSimulation *obj, double ratio, int *err); /* :3007-3013 */
double FluidSimulation_get_PICAPIC_ratio(FluidSimulation *obj, int *err);      /* :3014-3018 */
void FluidSimulation_set_PICAPIC_ratio(FluidSimulation *obj, double ratio, int *err); /* :3020-3026 */
/* surface tension (fluidsimulation_c.cpp:2757-2770, 2841-2876;
 * fluidsimulation.cpp:2604-2744): enabled above 1e-6; the condition number of
 * the substep restriction (> 0); the smooth kernel doubles the particle radius
 * of the liquid SDF and the velocity transfer. */
double FluidSimulation_get_surface_tension(FluidSimulation *obj, int *err);               /* :2757-2761 */
void FluidSimulation_set_surface_tension(FluidSimulation *obj, double k, int *err);       /* :2763-2769 */
double FluidSimulation_get_surface_tension_condition_number(FluidSimulation *obj, int *err); /* :2841-2847 */
void FluidSimulation_set_surface_tension_condition_number(FluidSimulation *obj, double n,
                                                          int *err);                      /* :2849-2855 */
void FluidSimulation_enable_smooth_surface_tension_kernel(FluidSimulation *obj, int *err);  /* :2857-2862 */
void FluidSimulation_disable_smooth_surface_tension_kernel(FluidSimulation *obj, int *err); /* :2864-2869 */
int FluidSimulation_is_smooth_surface_tension_kernel_enabled(FluidSimulation *obj, int *err); /* :2871-2876 */
/* constant viscosity and its solver (fluidsimulation_c.cpp:333-346, 2728-2754;
 * fluidsimulation.cpp:303-312, 2565-2601): the first set enables the
 * viscosity solve after the body forces of every substep; negative values
 * and tolerances are rejected. */
int FluidSimulation_get_viscosity_solver_max_iterations(FluidSimulation *obj, int *err);     /* :333-338 */
void FluidSimulation_set_viscosity_solver_max_iterations(FluidSimulation *obj, int n, int *err); /* :340-346 */
double FluidSimulation_get_viscosity(FluidSimulation *obj, int *err);                        /* :2728-2732 */
void FluidSimulation_set_viscosity(FluidSimulation *obj, double v, int *err);                /* :2734-2740 */
double FluidSimulation_get_viscosity_solver_error_tolerance(FluidSimulation *obj, int *err); /* :2742-2746 */
void FluidSimulation_set_viscosity_solver_error_tolerance(FluidSimulation *obj, double tol,
                                                          int *err);                          /* :2748-2754 */

/* ---- particle I/O ---- */
void FluidSimulation_load_marker_particle_data(FluidSimulation *obj,
                                               FluidSimulationMarkerParticleData data,
                                               int *err);                      /* :4872-4878 */
void FluidSimulation_load_marker_particle_affine_data(FluidSimulation *obj,
                                                      FluidSimulationMarkerParticleAffineData data,
                                                      int *err);               /* :4880-4886 */
void FluidSimulation_get_marker_particle_affinex_data_range(FluidSimulation *obj, int start_idx,
                                                            int end_idx, char *data,
                                                            int *err);         /* :4647-4656 */
void FluidSimulation_get_marker_particle_affiney_data_range(FluidSimulation *obj, int start_idx,
                                                            int end_idx, char *data,
                                                            int *err);         /* :4658-4667 */
void FluidSimulation_get_marker_particle_affinez_data_range(FluidSimulation *obj, int start_idx,
                                                            int end_idx, char *data,
                                                            int *err);         /* :4669-4678 */
int FluidSimulation_get_num_marker_particles(FluidSimulation *obj, int *err);  /* :3151-3156 */
unsigned int FluidSimulation_get_marker_particle_position_data_size(FluidSimulation *obj,
                                                                   int *err); /* :3952-3956 */
unsigned int FluidSimulation_get_marker_particle_velocity_data_size(FluidSimulation *obj,
                                                                   int *err); /* :3958-3962 */
void FluidSimulation_get_marker_particle_position_data(FluidSimulation *obj, char *data,
                                                       int *err);              /* :4823-4828 */
/* Returns positions, like the reference (fluidsimulation.cpp:3727-3735). */
void FluidSimulation_get_marker_particle_velocity_data(FluidSimulation *obj, char *data,
                                                       int *err);              /* :4830-4835 */
void FluidSimulation_get_marker_particle_position_data_range(FluidSimulation *obj, int start_idx,
                                                             int end_idx, char *data,
                                                             int *err);        /* :4625-4634 */
void FluidSimulation_get_marker_particle_velocity_data_range(FluidSimulation *obj, int start_idx,
                                                             int end_idx, char *data,
                                                             int *err);        /* :4636-4645 */

/* ---- stats ---- */
FluidSimulationFrameStats FluidSimulation_get_frame_stats_data(FluidSimulation *obj,
                                                               int *err);      /* :4618-4623 */
char *CBindings_get_error_message(void);                                       /* cbindings.cpp:101-104 */

/* ---- additive entry points (no counterpart in the reference) ---- */

/* Relative pressure tolerance of the PCG stop test |r|_inf <= tol*|b|_inf.
 * The reference hard-codes 1e-9 (fluidsimulation.h:2378). */
void FluidSimulation_set_pressure_solver_tolerance(FluidSimulation *obj, double tol, int *err);
double FluidSimulation_get_pressure_solver_tolerance(FluidSimulation *obj, int *err);

/* Pressure precision (no counterpart in the reference, which always solves
 * in FP64): 0 = FP64 PCG with the reference's MIC(0) preconditioner and
 * float-rounded step lengths; 1 = FP32 PCG restarted inside FP64 iterative
 * refinement, inner relative tolerance inner_tol; 2 = FP32 PCG with FP64
 * reliable updates, the true residual recomputed whenever the FP32 residual
 * has dropped by inner_tol; 3 = FP64 PCG whose MIC(0) application runs on
 * the FP32 copy of the factor (inner_tol unused).  In modes 1 and 2 the FP64
 * true residual decides convergence with the reference's rule. */
void FluidSimulation_set_pressure_solver_precision(FluidSimulation *obj, int mode,
                                                   double inner_tol, int *err);
int FluidSimulation_get_pressure_solver_precision(FluidSimulation *obj, int *err);

/* Dot products of the FP64 solve in the reference's summation order (no
 * counterpart in the reference, which always sums that way): BLAS::dot
 * (pcgsolver/blaswrapper.h:27-63) splits the rows into
 * min(max thread count, ceil(M / 500000)) intervals, sums each from zero in
 * index order and adds the interval sums in order.  With it on (and the same
 * thread count as the reference run) the FP64 solve takes the reference's
 * iterates bit for bit; off (the default), the sums use a fixed parallel tree. */
void FluidSimulation_set_pressure_solver_reference_summation(FluidSimulation *obj, int enabled,
                                                            int *err);
int FluidSimulation_get_pressure_solver_reference_summation(FluidSimulation *obj, int *err);
/* FP64 PCG (precision modes 0 and 3) started from the previous substep's
 * pressure at the rows instead of 0 (no counterpart in the reference, whose
 * PCG starts from 0); the stopping threshold stays tolerance x |b|_inf. */
void FluidSimulation_set_pressure_solver_warm_start(FluidSimulation *obj, int enabled, int *err);
int FluidSimulation_get_pressure_solver_warm_start(FluidSimulation *obj, int *err);
/* Schedule of the MIC(0) preconditioner application (no counterpart in the
 * reference, whose triangular solves are serial loops, pcgsolver.h:211-241):
 * 2 = bands of i+j+k hyperplanes, one launch per band; 4 = the line
 * pipeline, one launch per triangular sweep; 5 = the bands of schedule 2
 * recorded once per solve as a HIP graph and replayed per application.  All
 * give the same bits. */
void FluidSimulation_set_pressure_solver_preconditioner_schedule(FluidSimulation *obj, int schedule,
                                                                 int *err);
int FluidSimulation_get_pressure_solver_preconditioner_schedule(FluidSimulation *obj, int *err);

/* ---- mesh obstacles --------------------------------------------------------
 * MeshObject (meshobject.h) and its C ABI (c_bindings/meshobject_c.cpp, lines
 * cited per entry); the mesh is MeshUtils::TriangleMesh_t (meshutils.h:34-39).
 * Obstacles enter the solid SDF through FluidSimulation_add_mesh_obstacle
 * (fluidsimulation_c.cpp:3113-3135); the solid is rebuilt at the next
 * substep when the list or an object's enabled state changed, and at every
 * substep while an enabled animated object's mesh moves (its previous and
 * current frame differ), at the substep's frame progress
 * (fluidsimulation.cpp:5219-5426).  Static and animated meshes, normal and
 * inversed, with friction (the velocity constraint's closest-object
 * friction).  An animated object's solid velocity (from its vertices'
 * frame-to-frame translations times the velocity scale) enters the pressure
 * solve and the velocity constraint.  The three meshes of an animated update
 * must have the same vertex and triangle counts (the reference also accepts
 * a changing topology; this engine rejects it with an error). */
typedef struct MeshObject MeshObject;
typedef struct TriangleMesh_t {
    float *vertices;
    int *triangles;
    int numVertices;
    int numTriangles;
} TriangleMesh_t;
MeshObject *MeshObject_new(int i, int j, int k, double dx, int *err);                 /* :38-50 */
void MeshObject_destroy(MeshObject *obj);                                               /* :52-54 */
void MeshObject_update_mesh_static(MeshObject *obj, TriangleMesh_t mesh, int *err);     /* :56-69 */
void MeshObject_update_mesh_animated(MeshObject *obj, TriangleMesh_t previous, TriangleMesh_t current,
                                     TriangleMesh_t next, int *err);                   /* :71-88 */
void MeshObject_enable(MeshObject *obj, int *err);                                      /* :90-94 */
void MeshObject_disable(MeshObject *obj, int *err);                                     /* :96-100 */
int MeshObject_is_enabled(MeshObject *obj, int *err);                                   /* :102-106 */
void MeshObject_inverse(MeshObject *obj, int *err);                                     /* :108-112 */
int MeshObject_is_inversed(MeshObject *obj, int *err);                                  /* :114-118 */
float MeshObject_get_friction(MeshObject *obj, int *err);                               /* :120-124 */
void MeshObject_set_friction(MeshObject *obj, float f, int *err);                       /* :126-130 */
float MeshObject_get_velocity_scale(MeshObject *obj, int *err);                         /* :132-136 */
void MeshObject_set_velocity_scale(MeshObject *obj, float scale, int *err);             /* :138-142 */
float MeshObject_get_whitewater_influence(MeshObject *obj, int *err);                   /* :144-148 */
void MeshObject_set_whitewater_influence(MeshObject *obj, float value, int *err);       /* :150-154 */
float MeshObject_get_dust_emission_strength(MeshObject *obj, int *err);                 /* :156-160 */
void MeshObject_set_dust_emission_strength(MeshObject *obj, float value, int *err);     /* :162-166 */
float MeshObject_get_sheeting_strength(MeshObject *obj, int *err);                      /* :168-172 */
void MeshObject_set_sheeting_strength(MeshObject *obj, float value, int *err);          /* :174-178 */
float MeshObject_get_mesh_expansion(MeshObject *obj, int *err);                         /* :180-184 */
void MeshObject_set_mesh_expansion(MeshObject *obj, float ex, int *err);                /* :186-192 */
void MeshObject_enable_append_object_velocity(MeshObject *obj, int *err);               /* :194-198 */
void MeshObject_disable_append_object_velocity(MeshObject *obj, int *err);              /* :200-204 */
int MeshObject_is_append_object_velocity_enabled(MeshObject *obj, int *err);            /* :206-210 */
float MeshObject_get_object_velocity_influence(MeshObject *obj, int *err);              /* :212-216 */
void MeshObject_set_object_velocity_influence(MeshObject *obj, float value, int *err);  /* :218-224 */
int MeshObject_get_priority(MeshObject *obj, int *err);                                 /* :226-230 */
void MeshObject_set_priority(MeshObject *obj, int n, int *err);                         /* :232-236 */
int MeshObject_get_source_id(MeshObject *obj, int *err);                                /* :238-242 */
void MeshObject_set_source_id(MeshObject *obj, int id, int *err);                       /* :244-248 */
int MeshObject_get_viscosity(MeshObject *obj, int *err);                                /* :250-254 */
void MeshObject_set_viscosity(MeshObject *obj, float v, int *err);                      /* :256-260 */
int MeshObject_get_density(MeshObject *obj, int *err);                                  /* :262-266 */
void MeshObject_set_density(MeshObject *obj, float d, int *err);                        /* :268-272 */
int MeshObject_get_lifetime(MeshObject *obj, int *err);                                 /* :274-278 */
void MeshObject_set_lifetime(MeshObject *obj, float v, int *err);                       /* :280-284 */
int MeshObject_get_lifetime_variance(MeshObject *obj, int *err);                        /* :286-290 */
void MeshObject_set_lifetime_variance(MeshObject *obj, float v, int *err);              /* :292-296 */
FFVector3 MeshObject_get_source_color(MeshObject *obj, int *err);                       /* :298-303 */
void MeshObject_set_source_color(MeshObject *obj, double r, double g, double b, int *err); /* :305-312 */
void FluidSimulation_add_mesh_obstacle(FluidSimulation *obj, MeshObject *obstacle, int *err);    /* fluidsimulation_c.cpp:3113-3119 */
void FluidSimulation_remove_mesh_obstacle(FluidSimulation *obj, MeshObject *obstacle, int *err); /* :3121-3127 */
void FluidSimulation_remove_mesh_obstacles(FluidSimulation *obj, int *err);                      /* :3129-3134 */
void FluidSimulation_set_boundary_friction(FluidSimulation *obj, double f, int *err);            /* :2813-2817 */
double FluidSimulation_get_boundary_friction(FluidSimulation *obj, int *err);                   /* :2819-2823 */
/* rand() values the last solid rebuild drew for its sign tests' jitter
 * (three per test; no counterpart in the reference, for tests). */
int FFAMD_get_last_solid_rand_draws(FluidSimulation *obj, int *err);

/* Nodal solid SDF ((isize+1)*(jsize+1)*(ksize+1) floats, i fastest, negative
 * inside solids) replacing the domain-boundary SDF the engine builds itself.
 * Solids are static; their face velocity is zero.  Call before initialize. */
void FluidSimulation_set_solid_phi(FluidSimulation *obj, const float *phi_nodes, int *err);

/* Grid getters (host copies of the last substep's grids).  Sizes: liquid SDF
 * and pressure isize*jsize*ksize; solid SDF nodal; u (isize+1)*jsize*ksize,
 * v isize*(jsize+1)*ksize, w isize*jsize*(ksize+1). */
void FluidSimulation_get_liquid_sdf_data(FluidSimulation *obj, float *data, int *err);
void FluidSimulation_get_solid_sdf_data(FluidSimulation *obj, float *data, int *err);
/* the solid's face velocities of the last solid rebuild (MeshLevelSet velocity
 * grids, normalized and extrapolated; zero unless an animated obstacle took
 * part), u/v/w sized as the velocity field.  No counterpart in the reference's
 * C ABI: a getter for tests and tools. */
void FluidSimulation_get_solid_velocity_field(FluidSimulation *obj, float *u, float *v, float *w, int *err);
void FluidSimulation_get_velocity_field_data(FluidSimulation *obj, float *u, float *v, float *w,
                                             int *err);
void FluidSimulation_get_pressure_grid_data(FluidSimulation *obj, float *data, int *err);
/* iterations, |r|_inf, success flag and unknown count of the last pressure solve */
void FluidSimulation_get_last_pressure_solve_stats(FluidSimulation *obj, int *iterations,
                                                   double *error, int *success, int *unknowns,
                                                   int *err);
/* |b|_inf of the last pressure solve's right-hand side (no counterpart in the
 * reference): with get_last_pressure_solve_stats' |r|_inf it gives the
 * relative residual the stop test uses. */
double FluidSimulation_get_last_pressure_solve_rhs_norm(FluidSimulation *obj, int *err);
/* iterations, |r|_inf, success flag and matrix size of the last viscosity
 * solve (no counterpart in the reference, which logs them) */
void FluidSimulation_get_last_viscosity_solve_stats(FluidSimulation *obj, int *iterations,
                                                    double *error, int *success, int *unknowns,
                                                    int *err);
/* GPU milliseconds of the last substep's phases: 0 level sets, 1 P2G,
 * extrapolation and body forces, 2 pressure system, 3 PCG, 4 projection and
 * constraints, 5 G2P and advection, 6 removal, 7 whole substep, 8 the
 * viscosity solve. */
double FluidSimulation_get_last_substep_phase_ms(FluidSimulation *obj, int phase, int *err);
/* HIP-event timing of kernel groups on the engine's stream (off by default).
 * Enabling resets the counters; ids: 0 cell sort, 1 liquid SDF, 2 P2G,
 * 3 extrapolation, 4 pressure assembly, 5 MIC(0) factor, 6 SpMV, 7 MIC(0)
 * apply (both sweeps), 8 (unused), 9 G2P + advection, 10 removal,
 * 11 pressure gradient, 12 viscosity solve.  ms = summed kernel time,
 * launches = kernel count. */
void FluidSimulation_enable_kernel_timing(FluidSimulation *obj, int on, int *err);
void FluidSimulation_get_kernel_timing(FluidSimulation *obj, int id, double *ms, long long *launches,
                                       int *err);
/* Run one substep of length dt with the current state (no frame split). */
void FluidSimulation_step(FluidSimulation *obj, double dt, int *err);

/* ---- kernel-level entry points on host arrays (parity tests) ---- */
/* Stable cell-key counting sort: keys[np], counts[N], offsets[N+1], perm[np]. */
void FFAMD_cell_sort(int isize, int jsize, int ksize, double dx, int np, const float *pos,
                     int *keys, int *counts, int *offsets, int *perm, int *err);
void FFAMD_liquid_sdf(int isize, int jsize, int ksize, double dx, int np, const float *pos,
                      double radius, float *phi, int *err);
void FFAMD_p2g(int isize, int jsize, int ksize, double dx, int np, const float *pos,
               const float *vel, double radius, float *u, float *v, float *w,
               unsigned char *vu, unsigned char *vv, unsigned char *vw, int *err);
void FFAMD_extrapolate(int wi, int wj, int wk, float *field, const unsigned char *valid,
                       int layers, int *err);
/* FFAMD_extrapolate without any allocation (no counterpart in the
 * reference): for callers that cannot allocate device memory, such as code
 * under HIP graph capture, tight serving loops and harnesses that forbid
 * hipMalloc.  field (in place) and valid are DEVICE pointers of wi*wj*wk
 * entries; the kernels are enqueued on `stream` (a hipStream_t, NULL for the
 * default stream) and the call returns without waiting.  Its only scratch,
 * one status byte per face, is a static device array of 64 MiB shared by
 * every call on a GPU: at most 67108864 faces per call (larger sizes fail
 * with an error), and calls that overlap on one GPU race. */
void FFAMD_extrapolate_noalloc(int wi, int wj, int wk, float *field, const unsigned char *valid,
                               int layers, void *stream, int *err);
/* Adjoint of FFAMD_extrapolate (no counterpart in the reference): for the
 * same flags and layer count the extrapolation is linear in the field, and
 * this replaces bar (wk x wj x wi) by its transpose applied to bar. */
void FFAMD_extrapolate_vjp(int wi, int wj, int wk, float *bar, const unsigned char *valid,
                           int layers, int *err);
void FFAMD_solid_box_phi(int isize, int jsize, int ksize, double dx, float *phi_nodes, int *err);
/* MeshLevelSet::fastCalculateSignedDistanceField (meshlevelset.cpp:572-601)
 * of one mesh on a level set of isize x jsize x ksize cells at node offset
 * (oi, oj, ok): exact band of 3 nodes, crossing-count signs with rand()
 * jitter.  phi_nodes: (isize+1)(jsize+1)(ksize+1).  Returns the rand() draws. */
int FFAMD_mesh_levelset_fast(int isize, int jsize, int ksize, double dx, int oi, int oj, int ok, int nv,
                             const float *vertices, int nt, const int *triangles, float *phi_nodes, int *err);
/* FluidSimulation::_addStaticObjectsToSDF (fluidsimulation.cpp:5277-5339):
 * the domain box SDF with nobj static mesh obstacles unioned in (obstacle q:
 * nv[q] vertices, nt[q] triangles, concatenated, indices local to it;
 * inversed[q] != 0 for an inversed one; expansion (optional) the objects'
 * mesh expansion).  Returns the rand() draws. */
int FFAMD_solid_phi_meshes(int isize, int jsize, int ksize, double dx, int nobj, const int *nv,
                           const float *vertices, const int *nt, const int *triangles, const int *inversed,
                           const float *expansion, float *phi_nodes, int *err);
/* FluidSimulation::_updateSolidLevelSet with moving obstacles
 * (fluidsimulation.cpp:5219-5426; MeshObject::getMesh(f) and
 * getVertexVelocities(dt, f), meshobject.cpp:158-220; the MeshLevelSet
 * velocity data, meshlevelset.cpp:168-205, 577-718, 1320-1425, 1568-1845):
 * obstacle q has nv[q] vertices in three frames (prev, cur, next; one
 * concatenated array each) and nt[q] triangles (indices local to it);
 * flags[q] bit 0 animated, bit 1 inversed; expansion and velocity_scale per
 * object (optional).  At frame_progress of a frame of frame_dt seconds: the
 * nodal solid phi and the normalized, extrapolated solid face velocities
 * u, v, w.  Returns the rand() draws. */
int FFAMD_solid_update_moving(int isize, int jsize, int ksize, double dx, int nobj, const int *nv,
                              const int *nt, const float *prev, const float *cur, const float *next,
                              const int *triangles, const int *flags, const float *expansion,
                              const float *velocity_scale, double frame_dt, float frame_progress,
                              float *phi_nodes, float *u, float *v, float *w, int *err);
/* The viscosity stage (ViscositySolver, viscositysolver.cpp:50-79) on host
 * arrays: liquid_phi cell-centred, solid_phi_nodes and viscosity on the
 * (n+1)^3 nodes, u/v/w the MAC field.
 * FFAMD_viscosity_system builds the system and returns the matrix size m,
 * writing the nonzero count to *nnz; outputs (NULL skips): the face states
 * of the u, v, w faces (0 air, 1 fluid, 2 solid), the volume grids center,
 * U, V, W, edgeU, edgeV, edgeW concatenated, the face-to-row table (-1 off
 * the matrix), the CSR rows (rowstart[m+1], ascending cols, vals) and the
 * right-hand side.  FFAMD_viscosity_solve runs the stage in place on u, v,
 * w (cleared and set on the matrix faces when the solve succeeds) and
 * returns m; solution (NULL skips) receives the m-row solution. */
int FFAMD_viscosity_system(int isize, int jsize, int ksize, double dx, double dt, const float *liquid_phi,
                           const float *solid_phi_nodes, const float *viscosity, const float *u, const float *v,
                           const float *w, int *nnz, unsigned char *states, float *volumes, int *table,
                           int *rowstart, int *cols, float *vals, float *rhs, int *err);
int FFAMD_viscosity_solve(int isize, int jsize, int ksize, double dx, double dt, const float *liquid_phi,
                          const float *solid_phi_nodes, const float *viscosity, double tol, int max_iter, float *u,
                          float *v, float *w, int *iterations, float *error, int *success, float *solution,
                          int *err);
/* Surface tension stages on host arrays (liquid phi cell-centred, ni nj nk):
 * FFAMD_curvature_grid: ParticleLevelSet::calculateCurvatureGrid
 * (particlelevelset.cpp:196-292), the reinitialized surface phi and the
 * extrapolated curvature.  FFAMD_tension_cluster_status: 16 border, 1 / 2
 * cells of 4^3 blocks wholly inside / outside, 4 / 8 BAD clusters
 * (pressuresolver.cpp:314-606), 0 elsewhere.  FFAMD_pressure_solve_tension:
 * the FP64 pressure solve with the tension terms of constant k (none at
 * k <= 1e-6) and, on success, the gradient applied to u, v, w in place;
 * reference_threads > 0 takes the dot products in the reference's order for
 * that thread count.  Returns 1 on success. */
void FFAMD_curvature_grid(int isize, int jsize, int ksize, double dx, const float *liquid_phi, float *surface_phi,
                          float *curvature, int *err);
void FFAMD_tension_cluster_status(int isize, int jsize, int ksize, const float *liquid_phi, signed char *status,
                                  int *err);
int FFAMD_pressure_solve_tension(int isize, int jsize, int ksize, double dx, double dt, float *u, float *v, float *w,
                                 const float *wu, const float *wv, const float *ww, const float *liquid_phi, double k,
                                 double tol, int max_iter, int reference_threads, float *pressure, int *iterations,
                                 double *error, unsigned char *vu, unsigned char *vv, unsigned char *vw, int *err);
/* the cells' solid weights 1 - volume fraction of the solid
 * (fluidsimulation.cpp:6140-6147, meshlevelset.cpp:1491-1515) */
void FFAMD_cell_weights(int isize, int jsize, int ksize, const float *phi_nodes, float *wc, int *err);
/* PressureSolver::_conditionSolidVelocityField (pressuresolver.cpp:161-312):
 * the solid face velocities us, vs, ws (in place) zeroed around isolated
 * pockets of the liquid (phi cell-centred) */
void FFAMD_condition_solid_velocity(int isize, int jsize, int ksize, const float *liquid_phi,
                                    const float *wu, const float *wv, const float *ww, float *us,
                                    float *vs, float *ws, int *err);
void FFAMD_face_weights(int isize, int jsize, int ksize, const float *phi_nodes, float *wu,
                        float *wv, float *ww, int *err);
/* Pressure solve + projection on host grids; mode as in
 * FluidSimulation_set_pressure_solver_precision.  u, v, w updated in place. */
int FFAMD_pressure_solve(int isize, int jsize, int ksize, double dx, double dt, float *u,
                         float *v, float *w, const float *wu, const float *wv, const float *ww,
                         const float *liquid_phi, double tol, int max_iter, int mode,
                         float *pressure, int *iterations, double *error, unsigned char *vu,
                         unsigned char *vv, unsigned char *vw, int *err);
/* MIC(0) system of the pressure solve: rows[M], rhs, diag, ox, oy, oz,
 * nbr[6M], invdiag; returns M. Arrays must hold the interior cell count. */
int FFAMD_pressure_system(int isize, int jsize, int ksize, double dx, double dt,
                          const float *u, const float *v, const float *w, const float *wu,
                          const float *wv, const float *ww, const float *liquid_phi, int *rows,
                          double *rhs, double *diag, double *ox, double *oy, double *oz,
                          int *nbr, double *invdiag, int *err);
/* One MIC(0) preconditioner application z = (L L^T)^-1 r on the system and
 * factor FFAMD_pressure_system builds from the same inputs
 * (solveLower + solveLowerTransposeInPlace, pcgsolver.h:211-241).  r and z
 * hold n values in row order and n must equal M (else an error); precision
 * 0 = FP64, 1 = the FP32 copy of the factor used by the mixed-precision solve
 * (r rounded to float, z widened).  schedule 0 = one launch per hyperplane,
 * 1 = the tile-wavefront kernel, which is switched off and fails with an
 * error, 2 = bands of hyperplanes (the solver's: the launches name the tiles
 * that hold rows of their band), 3 = the same bands with each workgroup
 * reading its tile's hyperplane range on the device (used when the tiles
 * outnumber what a launch can list), 4 = the line pipeline (pencils of
 * 16 x 16 lines along i, one launch per triangular sweep).  The inputs must give a non-zero
 * divergence.  Returns M. */
int FFAMD_mic0_apply(int isize, int jsize, int ksize, double dx, double dt, const float *u,
                     const float *v, const float *w, const float *wu, const float *wv,
                     const float *ww, const float *liquid_phi, int precision, int schedule,
                     int n, const double *r, double *z, int *err);
/* Adjoints of the particle/grid transfers for fixed positions (no
 * counterpart in the reference).  FFAMD_p2g_vjp: P2G (as FFAMD_p2g, radius r)
 * is linear in the particle velocities; returns vbar (np x 3) = J^T of the
 * three face-grid cotangents grad_u/v/w.  FFAMD_sample_velocity_vjp:
 * trilinear sampling (as FFAMD_sample_velocity) is linear in the face grids;
 * returns the face-grid cotangents u, v, w of the per-point cotangents vbar
 * (np x 3).  Both gather in a fixed order (no float atomics).  With the
 * particle velocities vel (np x 3), FFAMD_p2g_vjp also returns xbar (np x 3),
 * the cotangent of the positions through the transfer weights (vel and xbar
 * may be null). */
void FFAMD_p2g_vjp(int isize, int jsize, int ksize, double dx, int np, const float *pos,
                   const float *vel, double radius, const float *grad_u, const float *grad_v,
                   const float *grad_w, float *vbar, float *xbar, int *err);
void FFAMD_sample_velocity_vjp(int isize, int jsize, int ksize, double dx, int np, const float *pos,
                               const float *vbar, float *u, float *v, float *w, int *err);

/* RK3 advection of positions through a MAC field (_RK3,
 * fluidsimulation.cpp:7616-7623) without the collision step, and its adjoint
 * (no counterpart in the reference): for the cotangent xbar (np x 3) of the
 * new positions, pbar (np x 3) is the cotangent of the old positions (the
 * trilinear field's gradient inside the cell the forward samples) and
 * grad_u/v/w that of the face grids. */
void FFAMD_advect_rk3(int isize, int jsize, int ksize, double dx, const float *u, const float *v,
                      const float *w, int np, const float *pos, double dt, float *out, int *err);
void FFAMD_advect_rk3_vjp(int isize, int jsize, int ksize, double dx, const float *u, const float *v,
                          const float *w, int np, const float *pos, double dt, const float *xbar,
                          float *pbar, float *grad_u, float *grad_v, float *grad_w, int *err);

/* Adjoint of the pressure projection (no counterpart in the reference):
 * for the liquid SDF and face weights of a FFAMD_pressure_solve call, the
 * projection u' = P(u) is linear, and this returns ubar = P^T u'bar on the
 * three face grids.  The pressure system's matrix is symmetric, so the
 * adjoint solve is the same FP64 MIC(0)-PCG (tolerance tol relative to its
 * right-hand side).  The forward's rounding of the pressure to float is not
 * differentiated.  Returns 1 when the solve converged. */
int FFAMD_projection_vjp(int isize, int jsize, int ksize, double dx, double dt, const float *wu,
                         const float *wv, const float *ww, const float *liquid_phi,
                         const float *grad_u, const float *grad_v, const float *grad_w, double tol,
                         int max_iter, float *out_u, float *out_v, float *out_w, int *iterations,
                         int *err);

/* Shape of the band schedule of the MIC(0) application (no counterpart in
 * the reference): hyperplanes per launch and tile side in lines, for the FP64
 * (fp64 = 1) or the FP32 factor. */
void FFAMD_mic0_band_shape(int fp64, int *levels, int *tile, int *err);
void FFAMD_sample_velocity(int isize, int jsize, int ksize, double dx, const float *u,
                           const float *v, const float *w, int np, const float *pos, float *out,
                           int *err);

/* Kernel-level APIC transfers for parity checks (the engine runs them inside
 * update()): FFAMD_p2g_apic replaces VelocityAdvector::advect with the APIC
 * producer (velocityadvector.cpp:38-465, 543-623) and FFAMD_g2p_apic
 * _updatePICAPICMarkerParticleVelocitiesThread (fluidsimulation.cpp:6709-6769,
 * 6791-6843).  aff: np x 9 (AFFINEX, AFFINEY, AFFINEZ per particle). */
void FFAMD_p2g_apic(int isize, int jsize, int ksize, double dx, int np, const float *pos,
                    const float *vel, const float *aff, double radius, float *u, float *v,
                    float *w, unsigned char *valid_u, unsigned char *valid_v,
                    unsigned char *valid_w, int *err);
void FFAMD_g2p_apic(int isize, int jsize, int ksize, double dx, const float *u, const float *v,
                    const float *w, int np, const float *pos, float *vel, float *aff, int *err);

/* Position cotangent of trilinear sampling (no counterpart in the reference):
 * xbar (np x 3) = sum_d vbar_d grad V_d(pos) for the face grids u/v/w, inside
 * the cell each sample reads; 0 outside the domain. */
void FFAMD_sample_velocity_position_vjp(int isize, int jsize, int ksize, double dx, const float *u,
                                        const float *v, const float *w, int np, const float *pos,
                                        const float *vbar, float *xbar, int *err);

/* Differentiable run of substeps (no counterpart in the reference: the
 * reference's grid is always P2G of its particles and it has no adjoint).
 * A rollout holds np particles at fixed positions pos (np x 3) whose initial
 * velocities are the trilinear sample of an initial MAC field, and runs
 * `substeps` substeps of length dt in the engine's order
 * (fluidsimulation.cpp:10078-10121; no particle removal), then returns the
 * particle density on the cells: each cell centre sums the transfer kernel's
 * weights w(|x_p - c|^2) of the particles within its radius.  The backward
 * gives the gradient of <density_bar, density> with respect to the initial
 * field, recomputing substeps from particle checkpoints kept every
 * checkpoint_every substeps.  The liquid SDF and P2G flags are those the
 * last forward with freeze = 0 recorded, held fixed (freeze = 1 makes a
 * forward reuse them, for finite differences of the same map), and the
 * collision response is taken as the identity.  solid_nodes: nodal solid SDF
 * ((isize+1)(jsize+1)(ksize+1)) or null for the domain walls; picflip: the
 * PIC share of the G2P blend; tol/max_iter: every pressure solve, forward
 * and adjoint (FP64 MIC(0)-PCG). */
typedef struct FFAMD_Rollout FFAMD_Rollout;
FFAMD_Rollout *FFAMD_rollout_new(int isize, int jsize, int ksize, double dx, int np,
                                 const float *pos, const float *solid_nodes, double dt,
                                 int substeps, int checkpoint_every, double picflip, double gx,
                                 double gy, double gz, double cfl, double tol, int max_iter,
                                 int *err);
void FFAMD_rollout_destroy(FFAMD_Rollout *r);
void FFAMD_rollout_forward(FFAMD_Rollout *r, const float *u0, const float *v0, const float *w0,
                           int freeze, float *density, int *err);
void FFAMD_rollout_backward(FFAMD_Rollout *r, const float *density_bar, float *grad_u,
                            float *grad_v, float *grad_w, int *err);
void FFAMD_rollout_positions(FFAMD_Rollout *r, float *pos, int *err);
/* Dot products of every pressure solve (forward and adjoint) in the
 * reference's summation order for `threads` threads (blaswrapper.h:27-57),
 * as FluidSimulation_set_pressure_solver_reference_summation does for the
 * engine; 0 restores the parallel tree.  With it a forward takes the CPU
 * oracle's bits. */
void FFAMD_rollout_set_reference_summation(FFAMD_Rollout *r, int threads, int *err);
/* particles moved by the collision response (summed over the last forward's
 * substeps) and the pressure iterations of the last forward and backward */
void FFAMD_rollout_stats(FFAMD_Rollout *r, long long *collided, long long *forward_iterations,
                         long long *backward_iterations, int *err);

#ifdef __cplusplus
}
#endif
#endif
