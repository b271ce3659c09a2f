/*
 * fforacle.h -- C ABI of the CPU oracle for the FLIP substep path.
 *
 * TEST INFRASTRUCTURE ONLY.  The oracle is a from-scratch, single-threaded
 * restatement of the reference engine's per-substep algorithms
 * (rlguy/Blender-FLIP-Fluids 1.8.5, src/engine/).  It exists so that tests,
 * the smoke check and the benchmark's CPU-baseline leg can check the HIP
 * engine; the shipped library never links, loads or calls it.
 *
 * Grids are i-fastest (flat = i + ni*(j + nj*k)), matching Array3d
 * (array3d.h:774-777).  Face arrays: u is (ni+1)*nj*nk, v is ni*(nj+1)*nk,
 * w is ni*nj*(nk+1).  Nodal arrays are (ni+1)*(nj+1)*(nk+1).  Particles are
 * packed float xyz triples, like vmath::vec3 (vmath.h:37-59).
 */
#ifndef FFORACLE_H
#define FFORACLE_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- stage functions (each one reference function; see fforacle.cpp) ---- */

/* Domain-boundary solid SDF at the nodes: the negated mesh SDF of the domain
 * box inset by 3*dx+1e-4 (fluidsimulation.cpp:5175-5186, 5277-5287). */
void oracle_solid_box_phi(int ni, int nj, int nk, double dx, float *phi_nodes);

/* MeshLevelSet::fastCalculateSignedDistanceField on a level set of
 * isize x jsize x ksize cells at node offset (oi, oj, ok): exact-band
 * distances (band nodes around each triangle), then crossing-count signs.
 * rnd[nrnd] are the rand() values the sign test draws (three per jitter
 * vector; rand_max = RAND_MAX).  phi_nodes: (isize+1)(jsize+1)(ksize+1).
 * Returns the number of rand() values used, -1 on error. */
int oracle_mesh_levelset_fast(int isize, int jsize, int ksize, double dx, int oi, int oj, int ok, int nv,
                              const float *verts, int nt, const int *tris, int band, const int *rnd, int nrnd,
                              int rand_max, float *phi_nodes);

/* The static solid SDF with nobj mesh obstacles: the domain box, then the
 * obstacles' level sets (fluidsimulation.cpp:5277-5339).  Obstacle q has
 * nv[q] vertices and nt[q] triangles, concatenated in verts/tris (triangle
 * indices local to the obstacle); inversed[q] != 0 for an inversed one;
 * expansion (optional) the objects' mesh expansion.
 * Returns the number of rand() values used, -1 on error. */
int oracle_solid_phi_meshes(int ni, int nj, int nk, double dx, int nobj, const int *nv, const float *verts,
                            const int *nt, const int *tris, const int *inversed, const float *expansion,
                            const int *rnd, int nrnd, int rand_max, float *phi_nodes);

/* Counting sort of particles by flat cell id of floor(p/dx).  keys[P],
 * counts[N], offsets[N+1] (exclusive scan), perm[P] (stable). */
void oracle_cell_sort(int ni, int nj, int nk, double dx, int np, const float *pos,
                      int *keys, int *counts, int *offsets, int *perm);

/* Liquid SDF from particles (particlelevelset.cpp:335-666). */
void oracle_liquid_sdf(int ni, int nj, int nk, double dx, int np, const float *pos,
                       double radius, float *phi);

/* Post-process of the liquid SDF against the solid SDF (particlelevelset.cpp:170-194). */
void oracle_liquid_sdf_postprocess(int ni, int nj, int nk, double dx,
                                   const float *solid_phi_nodes, float *phi);

/* P2G of one velocity component (dir 0/1/2 = U/V/W) (velocityadvector.cpp:38-536).
 * field and valid have that component's face-array size and are overwritten. */
void oracle_p2g(int ni, int nj, int nk, double dx, int np, const float *pos,
                const float *vel, double radius, int dir, float *field,
                unsigned char *valid);

/* APIC P2G of one velocity component (velocityadvector.cpp:38-465, 543-623):
 * trilinear deposits of v + affine . (node - p) to the 8 nodes of the
 * particle's cell in each 10^3 block it belongs to.  aff is np x 9: the
 * AFFINEX, AFFINEY and AFFINEZ vectors of each particle. */
void oracle_p2g_apic(int ni, int nj, int nk, double dx, int np, const float *pos,
                     const float *vel, const float *aff, double radius, int dir,
                     float *field, unsigned char *valid);

/* APIC G2P (fluidsimulation.cpp:6709-6769, 6791-6843): vel = the trilinear
 * sample, aff (np x 9) = the gradients of the three components. */
void oracle_g2p_apic(int ni, int nj, int nk, double dx, const float *u, const float *v,
                     const float *w, int np, const float *pos, float *vel, float *aff);

/* Layered extrapolation of one face array (gridutils.h:43-160). */
void oracle_extrapolate(int wi, int wj, int wk, float *field,
                        const unsigned char *valid, int layers);

/* Face weights from the nodal solid SDF (fluidsimulation.cpp:6066-6150). */
void oracle_face_weights(int ni, int nj, int nk, const float *solid_phi_nodes,
                         float *wu, float *wv, float *ww);

/* Pressure system: cells (k,j,i order), rhs, matrix, MIC(0), PCG
 * (pressuresolver.cpp:52-916, pcgsolver.h).  Returns the number of unknowns.
 * pressure[N] receives the float pressure grid.  status[0] = iterations,
 * status[1] = success flag (after the acceptable-tolerance rule),
 * err[0] = |r|_inf at exit.  nthreads mimics the reference's BLAS chunking
 * for the FP64 dot products (blaswrapper.h:27-57). */
/* The weight grid's cell weights: 1 - the solid's volume fraction per cell,
 * clamped to [0, 1] (fluidsimulation.cpp:6140-6147, meshlevelset.cpp:1491-1515,
 * levelsetutils.cpp:213-258). */
void oracle_cell_weights(int ni, int nj, int nk, const float *solid_phi_nodes, float *wc);
/* PressureSolver::_conditionSolidVelocityField (pressuresolver.cpp:161-312):
 * the solid face velocities us, vs, ws zeroed in place around isolated
 * pockets of liquid (phi cell-centred, face weights wu, wv, ww). */
void oracle_condition_solid_velocity(int ni, int nj, int nk, const float *phi, const float *wu, const float *wv,
                                     const float *ww, float *us, float *vs, float *ws);
/* Surface tension.  curvature_grid: ParticleLevelSet::calculateCurvatureGrid
 * (particlelevelset.cpp:196-292) of the cell-centred liquid phi: the
 * reinitialized surface phi and the extrapolated curvature (ni nj nk each).
 * tension_cluster_status: PressureSolver's cluster status per cell
 * (pressuresolver.cpp:314-606; 1 OK inside, 2 OK outside, 4 BAD inside,
 * 8 BAD outside, 16 border, 0 unset).  pressure_solve_tension: the pressure
 * solve with the surface tension constant k (no term when k <= 1e-6) and, on
 * success, the pressure applied to u, v, w in place (pressuresolver.cpp:
 * 52-112, 624-749, 1047-1205); returns the row count. */
void oracle_curvature_grid(int ni, int nj, int nk, double dx, const float *liquid_phi, float *surface_phi,
                           float *curvature);
void oracle_tension_cluster_status(int ni, int nj, int nk, const float *liquid_phi, char *status);
int oracle_pressure_solve_tension(int ni, int nj, int nk, double dx, double dt, float *u, float *v, float *w,
                                  const float *wu, const float *wv, const float *ww, const float *liquid_phi,
                                  double k, double tol, int max_iter, float *pressure, int *status, double *err,
                                  unsigned char *vu, unsigned char *vv, unsigned char *vw);

/* The viscosity solve (ViscositySolver, viscositysolver.cpp:50-79).  Grids:
 * liquid_phi cell-centred (ni nj nk), solid_phi_nodes and viscosity on
 * (ni+1)(nj+1)(nk+1) nodes, u/v/w the MAC field.
 * system_new builds the face states, the seven control-volume grids, the
 * matrix index and the linear system, returning the matrix size m and its
 * nonzero count; system_get copies out (NULL skips): states of the u, v, w
 * faces concatenated (0 air, 1 fluid, 2 solid), volumes center, U, V, W,
 * edgeU, edgeV, edgeW concatenated (sizes ni nj nk, (ni+1) nj nk,
 * ni (nj+1) nk, ni nj (nk+1), ni (nj+1)(nk+1), (ni+1) nj (nk+1),
 * (ni+1)(nj+1) nk), the face-to-row table (u, v, w faces, -1 when not in the
 * matrix), the rows as CSR with sorted columns, and the right-hand side.
 * pcg solves a CSR system with PCGSolver<float> (pcgsolver.h:246-300) and
 * returns 1 on success.  solve runs the whole stage in place on u, v, w
 * (viscositysolver.cpp:832-905) and returns m. */
void *oracle_viscosity_system_new(int ni, int nj, int nk, double dx, double dt, const float *liquid_phi,
                                  const float *solid_phi_nodes, const float *viscosity, const float *u,
                                  const float *v, const float *w, int *m, int *nnz);
void oracle_viscosity_system_get(void *handle, unsigned char *states, float *volumes, int *table, int *rowstart,
                                 int *cols, float *vals, float *rhs);
void oracle_viscosity_system_free(void *handle);
int oracle_viscosity_pcg(int m, const int *rowstart, const int *cols, const float *vals, const float *rhs,
                         double tol, int max_iter, int nthreads, float *x, int *iterations, float *error);
int oracle_viscosity_solve(int ni, int nj, int nk, double dx, double dt, const float *liquid_phi,
                           const float *solid_phi_nodes, const float *viscosity, double tol, int max_iter,
                           int nthreads, float *u, float *v, float *w, int *iterations, float *error,
                           int *success);

int oracle_pressure_solve(int ni, int nj, int nk, double dx, double dt,
                          const float *u, const float *v, const float *w,
                          const float *wu, const float *wv, const float *ww,
                          const float *liquid_phi, double tol, int max_iter,
                          int nthreads, float *pressure, int *status, double *err);

/* Pressure system pieces for kernel-level checks (same reference lines).
 * rows[M] = flat cell ids; rhs[M]; A: diag[M], ox[M], oy[M], oz[M]
 * (coefficient to the +x/+y/+z neighbour row, 0 when absent), nbr[6M]
 * (row ids of -x,+x,-y,+y,-z,+z or -1); invdiag[M]. */
int oracle_pressure_system(int ni, int nj, int nk, double dx, double dt,
                           const float *u, const float *v, const float *w,
                           const float *wu, const float *wv, const float *ww,
                           const float *liquid_phi, int *rows, double *rhs,
                           double *diag, double *ox, double *oy, double *oz,
                           int *nbr, double *invdiag);

/* One application of the MIC(0) preconditioner z = (L L^T)^-1 r (pcgsolver.h:211-241). */
void oracle_apply_mic0(int m, const double *diag, const double *ox, const double *oy,
                       const double *oz, const int *nbr, const double *invdiag,
                       const double *r, double *z);

/* Apply the pressure gradient (pressuresolver.cpp:94-112, 1021-1205). */
void oracle_apply_pressure(int ni, int nj, int nk, double dx, double dt,
                           const float *pressure, const float *liquid_phi,
                           const float *wu, const float *wv, const float *ww,
                           float *u, float *v, float *w,
                           unsigned char *vu, unsigned char *vv, unsigned char *vw);

/* Velocity at a point, FP64-internal trilinear (macvelocityfield.cpp:519-645). */
void oracle_sample_velocity(int ni, int nj, int nk, double dx, const float *u,
                            const float *v, const float *w, int np,
                            const float *pos, float *out);

/* Index of the element the reference's load queue returns for position i of
 * one load_marker_particle_data batch (see fforacle.cpp). */
size_t oracle_load_queue_index(size_t i);

/* ---- whole-engine oracle with the FluidSimulation stepping semantics ---- */
/* Threads for the loops whose outputs are disjoint and keep the serial
 * per-output order (liquid SDF and P2G per block, SpMV and vector updates per
 * row, G2P and advection per particle): the same bits at any count.
 * Default 1. */
void oracle_set_threads(int n);
int oracle_get_threads(void);

/* The solid update with moving mesh obstacles: the nodal solid phi and the
 * normalized solid face velocities (u: (ni+1) nj nk, v: ni (nj+1) nk, w:
 * ni nj (nk+1)).  Per object: nv vertices in each of prev/cur/next (3 floats
 * each), nt triangles (one list), flags bit 0 animated, bit 1 inversed.
 * Returns the rand() values used, or -1. */
int oracle_solid_update(int ni, int nj, int nk, double dx, int nobj, const int *nv, const int *nt,
                        const float *prev, const float *cur, const float *next, const int *tris, const int *flags,
                        const float *expansion, const float *vscale, double frame_dt, float frame_progress,
                        const int *rnd, int nrnd, int rand_max, float *phi_nodes, float *us, float *vs, float *ws);

typedef struct OracleSim OracleSim;
OracleSim *oracle_sim_new(int ni, int nj, int nk, double dx);
void oracle_sim_destroy(OracleSim *s);
void oracle_sim_add_body_force(OracleSim *s, double fx, double fy, double fz);
void oracle_sim_set_picflip_ratio(OracleSim *s, double r);
/* velocity transfer method: 1 = APIC, 0 = FLIP (the default); set before
 * initialize.  oracle_sim_get_affine writes 9 floats per particle. */
void oracle_sim_set_apic(OracleSim *s, int on);
void oracle_sim_get_affine(OracleSim *s, float *aff);
void oracle_sim_set_cfl(OracleSim *s, int cfl);
void oracle_sim_set_pressure_tolerance(OracleSim *s, double tol);
void oracle_sim_set_max_pressure_iterations(OracleSim *s, int n);
void oracle_sim_set_thread_chunks(OracleSim *s, int n);
/* setViscosity with a constant value, the viscosity solver's tolerance and
 * iteration limit (fluidsimulation.cpp:2569-2601, 303-312); the last frame's
 * viscosity solve statistics: returns success, iterations and error of the
 * first failing or else the most iterated substep. */
void oracle_sim_set_viscosity(OracleSim *s, double visc, double tol, int max_iter);
int oracle_sim_last_viscosity_stats(OracleSim *s, int *iterations, float *error);
/* setSurfaceTension (enabled above 1e-6), setSurfaceTensionConditionNumber
 * and the smooth surface tension kernel (fluidsimulation.cpp:2608-2744) */
void oracle_sim_set_surface_tension(OracleSim *s, double k, double condition, int smooth);
void oracle_sim_set_solid_phi(OracleSim *s, const float *phi_nodes);
void oracle_sim_load_particles(OracleSim *s, int np, const float *pos, const float *vel);
/* Mesh obstacles with velocities (static or animated, normal or inversed:
 * flags bit 0 animated, bit 1 inversed), one triangle list and three vertex
 * sets each; replaces the box/set solid by the solid update of every substep
 * whose meshes move.  The sign tests draw from rand() seeded with
 * frame + substep at each substep, as _stepFluid seeds it
 * (fluidsimulation.cpp:10079), through a restatement of the C library's
 * generator (oracle_glibc_rand). */
int oracle_sim_add_moving_obstacle(OracleSim *s, int nv, const float *prev, const float *cur, const float *next,
                                   int nt, const int *tris, int flags, float expansion, float vscale);
int oracle_sim_update_moving_obstacle(OracleSim *s, int index, const float *prev, const float *cur,
                                      const float *next);
int oracle_sim_rand_used(OracleSim *s);
/* The first n values of the C library's rand() after srand(seed), from the
 * oracle's restatement of glibc's generator. */
int oracle_glibc_rand(unsigned seed, int n, int *out);
void oracle_sim_get_solid_velocity(OracleSim *s, float *u, float *v, float *w);
void oracle_sim_get_solid_phi(OracleSim *s, float *phi);
void oracle_sim_initialize(OracleSim *s);
void oracle_sim_update(OracleSim *s, double frame_dt);
void oracle_sim_step(OracleSim *s, double dt);
int oracle_sim_num_particles(OracleSim *s);
void oracle_sim_get_particles(OracleSim *s, float *pos, float *vel);
int oracle_sim_last_substeps(OracleSim *s);
int oracle_sim_last_pressure_iterations(OracleSim *s);
double oracle_sim_last_pressure_error(OracleSim *s);
/* copies of the last substep's grids (sizes as above); pass NULL to skip */
void oracle_sim_get_grids(OracleSim *s, float *liquid_phi, float *u, float *v,
                          float *w, float *pressure);

/* ---- the differentiable rollout and its FP64 adjoint (no counterpart in the
 * reference; the HIP version is FFAMD_rollout_*) ----
 * substeps substeps of length dt of the engine's substep without particle
 * removal, from np particles at pos whose velocities are the trilinear sample
 * of the initial MAC field (u0, v0, w0), then the particle density on the
 * cells.  forward: freeze = 0 records the liquid SDF and P2G flags of every
 * substep, freeze = 1 reuses the recorded ones; returns 0, or -1 when a
 * pressure solve failed.  backward: the gradient of <density_bar, density>
 * with respect to the initial field of the last recording forward, in
 * double (masks held fixed, the collision response as the identity);
 * returns 0 or -1. */
typedef struct OracleRollout OracleRollout;
OracleRollout *oracle_rollout_new(int ni, int nj, int nk, double dx, int np, const float *pos,
                                  const float *solid_nodes, double dt, int substeps, double picflip,
                                  double gx, double gy, double gz, double cfl, double tol, int max_iter);
void oracle_rollout_destroy(OracleRollout *r);
int oracle_rollout_forward(OracleRollout *r, const float *u0, const float *v0, const float *w0,
                           int freeze, float *density);
int oracle_rollout_backward(OracleRollout *r, const double *density_bar, double *gu, double *gv,
                            double *gw);
void oracle_rollout_positions(OracleRollout *r, float *pos);
long long oracle_rollout_iterations(OracleRollout *r);

#ifdef __cplusplus
}
#endif
#endif
