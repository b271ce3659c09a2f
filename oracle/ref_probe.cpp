/*
 * ref_probe.cpp -- drives the reference engine's own classes, stage by stage.
 *
 * TEST INFRASTRUCTURE ONLY.  Compiled by oracle/Makefile against the headers
 * of the reference sources under $(REF)/src/engine and linked with the
 * reference engine built from them (oracle/_ref/libffengine_ref.so).  Tests
 * call these entry points to pin the oracle restatement to the reference bit
 * for bit.  Layouts and argument meanings match fforacle.h.
 *
 * Private members of FluidSimulation and ParticleLevelSet are reached by
 * compiling their headers with class privacy lifted; the probe only reads
 * state or calls the private step functions the engine itself calls.
 */
#include <algorithm>
#include <array>
#include <atomic>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstdint>
#include <cstring>
#include <deque>
#include <fstream>
#include <functional>
#include <iomanip>
#include <iostream>
#include <limits>
#include <map>
#include <memory>
#include <mutex>
#include <queue>
#include <random>
#include <set>
#include <sstream>
#include <stdexcept>
#include <string>
#include <thread>
#include <unordered_map>
#include <unordered_set>
#include <utility>
#include <vector>

// standard headers first: privacy is lifted only for the engine's own headers
#define private public
#define protected public
#include "fluidsimulation.h"
#include "gridutils.h"
#include "macvelocityfield.h"
#include "meshlevelset.h"
#include "meshobject.h"
#include "trianglemesh.h"
#include "particlelevelset.h"
#include "particlesystem.h"
#include "pressuresolver.h"
#include "velocityadvector.h"
#include "viscositysolver.h"
#undef private
#undef protected

namespace {
void fillParticles(ParticleSystem &ps, int np, const float *pos, const float *vel) {
    ps.addAttributeVector3("POSITION");
    ps.addAttributeVector3("VELOCITY");
    std::vector<vmath::vec3> *P, *V;
    ps.getAttributeValues("POSITION", P);
    ps.getAttributeValues("VELOCITY", V);
    for (int p = 0; p < np; p++) {
        P->push_back(vmath::vec3(pos[3 * p], pos[3 * p + 1], pos[3 * p + 2]));
        V->push_back(vel ? vmath::vec3(vel[3 * p], vel[3 * p + 1], vel[3 * p + 2]) : vmath::vec3());
    }
    ps.update();
}
template <class T>
void copyOut(Array3d<T> &a, T *out) {
    std::memcpy(out, a.getRawArray(), sizeof(T) * (size_t)a.width * a.height * a.depth);
}
template <class T>
void copyIn(Array3d<T> &a, const T *in) {
    std::memcpy(a.getRawArray(), in, sizeof(T) * (size_t)a.width * a.height * a.depth);
}
TriangleMesh makeMesh(int nv, const float *verts, int nt, const int *tris) {
    TriangleMesh m;
    for (int q = 0; q < nv; q++) m.vertices.push_back(vmath::vec3(verts[3 * q], verts[3 * q + 1], verts[3 * q + 2]));
    for (int q = 0; q < nt; q++) {
        Triangle t;
        t.tri[0] = tris[3 * q];
        t.tri[1] = tris[3 * q + 1];
        t.tri[2] = tris[3 * q + 2];
        m.triangles.push_back(t);
    }
    return m;
}
void boolOut(Array3d<bool> &a, unsigned char *out) {
    size_t n = (size_t)a.width * a.height * a.depth;
    bool *b = a.getRawArray();
    for (size_t i = 0; i < n; i++) out[i] = b[i] ? 1 : 0;
}
}  // namespace

extern "C" {

void refprobe_set_threads(int n) { ThreadUtils::setMaxThreadCount(n); }

/* FluidSimulation::_addStaticObjectsToSDF on a fresh level set (fluidsimulation.cpp:5277-5339). */
void refprobe_solid_box_phi(int ni, int nj, int nk, double dx, float *out) {
    FluidSimulation sim(ni, nj, nk, dx);
    sim.initialize();
    MeshLevelSet sdf(ni, nj, nk, dx);
    sim._addStaticObjectsToSDF(0.0, sdf);
    copyOut(*sdf.getPhiArray3d(), out);
}

/* MeshLevelSet::fastCalculateSignedDistanceField on a level set of
 * isize x jsize x ksize cells at node offset (oi, oj, ok), after srand(seed)
 * (meshlevelset.cpp:572-601). */
void refprobe_mesh_levelset_fast(int isize, int jsize, int ksize, double dx, int oi, int oj, int ok, int nv,
                                 const float *verts, int nt, const int *tris, int band, unsigned seed,
                                 float *out) {
    TriangleMesh m = makeMesh(nv, verts, nt, tris);
    MeshLevelSet ls(isize, jsize, ksize, dx);
    ls.setGridOffset(GridIndex(oi, oj, ok));
    ls.disableVelocityData();
    srand(seed);
    ls.fastCalculateSignedDistanceField(m, band);
    copyOut(*ls.getPhiArray3d(), out);
}

/* FluidSimulation::_addStaticObjectsToSDF with static mesh obstacles, after
 * srand(seed) (fluidsimulation.cpp:5277-5339).  The temporary level set is
 * sized to the domain as _updateSolidLevelSet does (5396-5405).  Returns the
 * wall-clock seconds of _addStaticObjectsToSDF alone. */
double refprobe_solid_phi_meshes(int ni, int nj, int nk, double dx, int nobj, const int *nv, const float *verts,
                               const int *nt, const int *tris, const int *inversed, const float *expansion,
                               unsigned seed, float *out) {
    FluidSimulation sim(ni, nj, nk, dx);
    sim.initialize();
    std::vector<std::unique_ptr<MeshObject>> objs;
    size_t vo = 0, to = 0;
    for (int q = 0; q < nobj; q++) {
        objs.emplace_back(new MeshObject(ni, nj, nk, dx));
        objs.back()->updateMeshStatic(makeMesh(nv[q], verts + 3 * vo, nt[q], tris + 3 * to));
        if (inversed[q]) objs.back()->inverse();
        if (expansion) objs.back()->setMeshExpansion(expansion[q]);
        sim.addMeshObstacle(objs.back().get());
        vo += nv[q];
        to += nt[q];
    }
    sim._tempSolidSDF = MeshLevelSet(ni, nj, nk, dx);
    MeshLevelSet sdf(ni, nj, nk, dx);
    srand(seed);
    auto t0 = std::chrono::steady_clock::now();
    sim._addStaticObjectsToSDF(1.0 / 30.0, sdf);
    double seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    copyOut(*sdf.getPhiArray3d(), out);
    sim.removeMeshObstacles();
    return seconds;
}

void refprobe_liquid_sdf(int ni, int nj, int nk, double dx, int np, const float *pos,
                         double radius, float *out) {
    ParticleSystem ps;
    fillParticles(ps, np, pos, nullptr);
    ParticleLevelSet pls(ni, nj, nk, dx);
    pls.calculateSignedDistanceField(ps, radius);
    copyOut(pls._phi, out);
}

void refprobe_liquid_sdf_postprocess(int ni, int nj, int nk, double dx, const float *solid,
                                     float *phi) {
    MeshLevelSet m(ni, nj, nk, dx);
    copyIn(*m.getPhiArray3d(), solid);
    ParticleLevelSet pls(ni, nj, nk, dx);
    copyIn(pls._phi, (const float *)phi);
    pls.postProcessSignedDistanceField(m);
    copyOut(pls._phi, phi);
}

void refprobe_p2g(int ni, int nj, int nk, double dx, int np, const float *pos, const float *vel,
                  double radius, float *u, float *v, float *w, unsigned char *vu,
                  unsigned char *vv, unsigned char *vw) {
    ParticleSystem ps;
    fillParticles(ps, np, pos, vel);
    MACVelocityField mac(ni, nj, nk, dx);
    ValidVelocityComponentGrid valid(ni, nj, nk);
    VelocityAdvectorParameters params;
    params.particles = &ps;
    params.vfield = &mac;
    params.validVelocities = &valid;
    params.particleRadius = radius;
    VelocityAdvector adv;
    adv.advect(params);
    copyOut(*mac.getArray3dU(), u);
    copyOut(*mac.getArray3dV(), v);
    copyOut(*mac.getArray3dW(), w);
    boolOut(valid.validU, vu);
    boolOut(valid.validV, vv);
    boolOut(valid.validW, vw);
}

void refprobe_p2g_apic(int ni, int nj, int nk, double dx, int np, const float *pos, const float *vel,
                       const float *aff, double radius, float *u, float *v, float *w, unsigned char *vu,
                       unsigned char *vv, unsigned char *vw) {
    ParticleSystem ps;
    ps.addAttributeVector3("POSITION");
    ps.addAttributeVector3("VELOCITY");
    const char *names[3] = {"AFFINEX", "AFFINEY", "AFFINEZ"};
    for (int d = 0; d < 3; d++) ps.addAttributeVector3(names[d]);
    std::vector<vmath::vec3> *Pv, *Vv, *A[3];
    ps.getAttributeValues("POSITION", Pv);
    ps.getAttributeValues("VELOCITY", Vv);
    for (int d = 0; d < 3; d++) ps.getAttributeValues(names[d], A[d]);
    for (int p = 0; p < np; p++) {
        Pv->push_back(vmath::vec3(pos[3 * p], pos[3 * p + 1], pos[3 * p + 2]));
        Vv->push_back(vmath::vec3(vel[3 * p], vel[3 * p + 1], vel[3 * p + 2]));
        for (int d = 0; d < 3; d++)
            A[d]->push_back(vmath::vec3(aff[9 * p + 3 * d], aff[9 * p + 3 * d + 1], aff[9 * p + 3 * d + 2]));
    }
    ps.update();
    MACVelocityField mac(ni, nj, nk, dx);
    ValidVelocityComponentGrid valid(ni, nj, nk);
    VelocityAdvectorParameters params;
    params.particles = &ps;
    params.vfield = &mac;
    params.validVelocities = &valid;
    params.particleRadius = radius;
    params.velocityTransferMethod = VelocityAdvectorTransferMethod::APIC;
    VelocityAdvector adv;
    adv.advect(params);
    copyOut(*mac.getArray3dU(), u);
    copyOut(*mac.getArray3dV(), v);
    copyOut(*mac.getArray3dW(), w);
    boolOut(valid.validU, vu);
    boolOut(valid.validV, vv);
    boolOut(valid.validW, vw);
}

/* FluidSimulation::_updatePICAPICMarkerParticleVelocitiesThread on a field
 * copied into a fresh simulation object: velocities and affine vectors. */
void refprobe_g2p_apic(int ni, int nj, int nk, double dx, const float *u, const float *v, const float *w, int np,
                       const float *pos, float *vel, float *aff) {
    FluidSimulation sim(ni, nj, nk, dx);
    sim.setVelocityTransferMethodAPIC();
    sim._initializeParticleSystems();
    std::vector<vmath::vec3> *P, *V, *A[3];
    sim._markerParticles.getAttributeValues("POSITION", P);
    sim._markerParticles.getAttributeValues("VELOCITY", V);
    for (int p = 0; p < np; p++) {
        P->push_back(vmath::vec3(pos[3 * p], pos[3 * p + 1], pos[3 * p + 2]));
        V->push_back(vmath::vec3());
    }
    sim._markerParticles.update();
    sim._markerParticles.getAttributeValues("AFFINEX", A[0]);
    sim._markerParticles.getAttributeValues("AFFINEY", A[1]);
    sim._markerParticles.getAttributeValues("AFFINEZ", A[2]);
    sim._MACVelocity = MACVelocityField(ni, nj, nk, dx);
    copyIn(*sim._MACVelocity.getArray3dU(), u);
    copyIn(*sim._MACVelocity.getArray3dV(), v);
    copyIn(*sim._MACVelocity.getArray3dW(), w);
    sim._updatePICAPICMarkerParticleVelocitiesThread(0, np);
    for (int p = 0; p < np; p++) {
        vel[3 * p] = V->at(p).x;
        vel[3 * p + 1] = V->at(p).y;
        vel[3 * p + 2] = V->at(p).z;
        for (int d = 0; d < 3; d++) {
            aff[9 * p + 3 * d] = A[d]->at(p).x;
            aff[9 * p + 3 * d + 1] = A[d]->at(p).y;
            aff[9 * p + 3 * d + 2] = A[d]->at(p).z;
        }
    }
}

void refprobe_extrapolate(int wi, int wj, int wk, float *field, const unsigned char *valid,
                          int layers) {
    Array3d<float> g(wi, wj, wk, 0.0f);
    copyIn(g, (const float *)field);
    Array3d<bool> vg(wi, wj, wk, false);
    size_t n = (size_t)wi * wj * wk;
    for (size_t i = 0; i < n; i++) vg.getRawArray()[i] = valid[i] != 0;
    GridUtils::extrapolateGrid(&g, &vg, layers);
    copyOut(g, field);
}

/* FluidSimulation::_updateWeightGridThread (fluidsimulation.cpp:6110-6150). */
void refprobe_face_weights(int ni, int nj, int nk, double dx, const float *solid, float *wu,
                           float *wv, float *ww, float *wc) {
    MeshLevelSet m(ni, nj, nk, dx);
    copyIn(*m.getPhiArray3d(), solid);
    auto cl = [](float x) { return std::max(0.0f, std::min(x, 1.0f)); };
    for (int k = 0; k < nk; k++)
        for (int j = 0; j < nj; j++)
            for (int i = 0; i < ni + 1; i++)
                wu[i + (ni + 1) * (j + nj * k)] = cl(1.0f - m.getFaceWeightU(i, j, k));
    for (int k = 0; k < nk; k++)
        for (int j = 0; j < nj + 1; j++)
            for (int i = 0; i < ni; i++)
                wv[i + ni * (j + (nj + 1) * k)] = cl(1.0f - m.getFaceWeightV(i, j, k));
    for (int k = 0; k < nk + 1; k++)
        for (int j = 0; j < nj; j++)
            for (int i = 0; i < ni; i++)
                ww[i + ni * (j + nj * k)] = cl(1.0f - m.getFaceWeightW(i, j, k));
    if (wc)
        for (int k = 0; k < nk; k++)
            for (int j = 0; j < nj; j++)
                for (int i = 0; i < ni; i++)
                    wc[i + ni * (j + nj * k)] = cl(1.0f - m.getCellWeight(i, j, k));
}

/* PressureSolver::solve + applySolutionToVelocityField, surface tension off,
 * density 1, zero solid velocity (fluidsimulation.cpp:6206-6249).  u, v, w are
 * updated in place when the solve succeeds; returns 1 on success. */
int refprobe_pressure_solve(int ni, int nj, int nk, double dx, double dt, float *u, float *v,
                            float *w, const float *wu, const float *wv, const float *ww,
                            const float *wc, const float *phi, double tol, int maxIter,
                            float *pressure, int *iterations, float *error,
                            unsigned char *vu, unsigned char *vv, unsigned char *vw) {
    MACVelocityField fluid(ni, nj, nk, dx), solid(ni, nj, nk, dx);
    copyIn(*fluid.getArray3dU(), (const float *)u);
    copyIn(*fluid.getArray3dV(), (const float *)v);
    copyIn(*fluid.getArray3dW(), (const float *)w);
    ValidVelocityComponentGrid valid(ni, nj, nk);
    Array3d<float> liquid(ni, nj, nk, 0.0f);
    copyIn(liquid, phi);
    WeightGrid wg(ni, nj, nk);
    copyIn(wg.U, wu);
    copyIn(wg.V, wv);
    copyIn(wg.W, ww);
    copyIn(wg.center, wc);
    Array3d<float> pgrid(ni, nj, nk, 0.0f), density(ni, nj, nk, 1.0f);
    PressureSolverParameters params;
    params.cellwidth = dx;
    params.deltaTime = dt;
    params.tolerance = tol;
    params.acceptableTolerance = 1.0;
    params.maxIterations = maxIter;
    params.velocityFieldFluid = &fluid;
    params.velocityFieldSolid = &solid;
    params.validVelocities = &valid;
    params.liquidSDF = &liquid;
    params.weightGrid = &wg;
    params.pressureGrid = &pgrid;
    params.densityGrid = &density;
    params.isSurfaceTensionEnabled = false;
    PressureSolver ps;
    bool ok = ps.solve(params);
    if (ok) ps.applySolutionToVelocityField();
    copyOut(pgrid, pressure);
    *iterations = ps.getIterations();
    *error = ps.getError();
    copyOut(*fluid.getArray3dU(), u);
    copyOut(*fluid.getArray3dV(), v);
    copyOut(*fluid.getArray3dW(), w);
    boolOut(valid.validU, vu);
    boolOut(valid.validV, vv);
    boolOut(valid.validW, vw);
    return ok ? 1 : 0;
}

void refprobe_sample_velocity(int ni, int nj, int nk, double dx, const float *u, const float *v,
                              const float *w, int np, const float *pos, float *out) {
    MACVelocityField mac(ni, nj, nk, dx);
    copyIn(*mac.getArray3dU(), u);
    copyIn(*mac.getArray3dV(), v);
    copyIn(*mac.getArray3dW(), w);
    for (int p = 0; p < np; p++) {
        vmath::vec3 r = mac.evaluateVelocityAtPositionLinear(vmath::vec3(pos[3 * p], pos[3 * p + 1], pos[3 * p + 2]));
        out[3 * p] = r.x;
        out[3 * p + 1] = r.y;
        out[3 * p + 2] = r.z;
    }
}

/* Whole engine through the reference's own object, with the phi grids of the
 * last substep exposed for comparison. */
void *refprobe_sim_new(int ni, int nj, int nk, double dx) { return new FluidSimulation(ni, nj, nk, dx); }
void refprobe_sim_get_liquid_phi(void *s, float *out) {
    copyOut(*((FluidSimulation *)s)->_liquidSDF.getPhiGrid(), out);
}
void refprobe_sim_get_solid_phi(void *s, float *out) {
    copyOut(*((FluidSimulation *)s)->_solidSDF.getPhiArray3d(), out);
}
void refprobe_sim_get_mac(void *s, float *u, float *v, float *w) {
    FluidSimulation *f = (FluidSimulation *)s;
    copyOut(*f->_MACVelocity.getArray3dU(), u);
    copyOut(*f->_MACVelocity.getArray3dV(), v);
    copyOut(*f->_MACVelocity.getArray3dW(), w);
}
/* The solid update of one substep (_updateSolidLevelSet, fluidsimulation.cpp:
 * 5387-5426) forced with the given frame timing (frame progress = 1 -
 * (remaining + step) / frame_dt), then its phi and normalized velocity grids. */
void refprobe_sim_update_solid(void *s, double frame_dt, double remaining, double step) {
    FluidSimulation *f = (FluidSimulation *)s;
    f->_currentFrameDeltaTime = frame_dt;
    f->_currentFrameDeltaTimeRemaining = remaining;
    f->_currentFrameTimeStep = step;
    f->_isSolidLevelSetUpToDate = false;
    f->_updateSolidLevelSet(frame_dt);
}
void refprobe_sim_get_solid_velocity(void *s, float *u, float *v, float *w) {
    MACVelocityField &m = ((FluidSimulation *)s)->_solidSDF.getVelocityDataGrid()->field;
    copyOut(*m.getArray3dU(), u);
    copyOut(*m.getArray3dV(), v);
    copyOut(*m.getArray3dW(), w);
}
/* ParticleLevelSet::calculateCurvatureGrid (particlelevelset.cpp:196-292).
 * The level set solver's second buffer is an uninitialized allocation
 * (levelsetsolver.cpp:97, array3d.h:706) whose cells outside the solver
 * region are read after the first swap: what the reference computes depends
 * on what that memory held.  Callers that compare with the oracle run the
 * process with every allocation of a page or more on fresh zeroed pages
 * (MALLOC_MMAP_THRESHOLD_=4096 from the start), so the buffer reads 0. */
namespace {
void curvatureOf(int ni, int nj, int nk, double dx, const float *phi, Array3d<float> &surface,
                 Array3d<float> &kgrid) {
    ParticleLevelSet pls(ni, nj, nk, dx);
    copyIn(pls._phi, phi);
    pls.calculateCurvatureGrid(surface, kgrid);
}
}  // namespace

void refprobe_curvature_grid(int ni, int nj, int nk, double dx, const float *phi, float *surface_phi,
                             float *curvature) {
    Array3d<float> surface(ni, nj, nk, 0.0f), kgrid(ni, nj, nk, 0.0f);
    curvatureOf(ni, nj, nk, dx, phi, surface, kgrid);
    copyOut(surface, surface_phi);
    copyOut(kgrid, curvature);
}

/* PressureSolver::solve + applySolutionToVelocityField with surface tension
 * constant k and the curvature grid of the liquid phi computed as the engine
 * does (fluidsimulation.cpp:5704-5730, 6206-6249); static solid, density 1.
 * Writes the cluster status of the solve (pressuresolver.cpp:314-606). */
int refprobe_pressure_solve_tension(int ni, int nj, int nk, double dx, double dt, float *u, float *v, float *w,
                                    const float *wu, const float *wv, const float *ww, const float *phi,
                                    double k, double tol, int maxIter, float *pressure, int *iterations,
                                    float *error, unsigned char *vu, unsigned char *vv, unsigned char *vw,
                                    char *cluster) {
    MACVelocityField fluid(ni, nj, nk, dx), solid(ni, nj, nk, dx);
    copyIn(*fluid.getArray3dU(), (const float *)u);
    copyIn(*fluid.getArray3dV(), (const float *)v);
    copyIn(*fluid.getArray3dW(), (const float *)w);
    ValidVelocityComponentGrid valid(ni, nj, nk);
    Array3d<float> liquid(ni, nj, nk, 0.0f);
    copyIn(liquid, phi);
    WeightGrid wg(ni, nj, nk);
    copyIn(wg.U, wu);
    copyIn(wg.V, wv);
    copyIn(wg.W, ww);
    wg.center.fill(1.0f);
    Array3d<float> surface(ni, nj, nk, 0.0f), kgrid(ni, nj, nk, 0.0f);
    curvatureOf(ni, nj, nk, dx, phi, surface, kgrid);
    Array3d<float> pgrid(ni, nj, nk, 0.0f), density(ni, nj, nk, 1.0f);
    PressureSolverParameters params;
    params.cellwidth = dx;
    params.deltaTime = dt;
    params.tolerance = tol;
    params.acceptableTolerance = 1.0;
    params.maxIterations = maxIter;
    params.velocityFieldFluid = &fluid;
    params.velocityFieldSolid = &solid;
    params.validVelocities = &valid;
    params.liquidSDF = &liquid;
    params.weightGrid = &wg;
    params.pressureGrid = &pgrid;
    params.densityGrid = &density;
    params.isSurfaceTensionEnabled = k > 1e-6;
    params.surfaceTensionConstant = k;
    params.curvatureGrid = &kgrid;
    PressureSolver ps;
    bool ok = ps.solve(params);
    if (ok) ps.applySolutionToVelocityField();
    copyOut(pgrid, pressure);
    *iterations = ps.getIterations();
    *error = ps.getError();
    copyOut(*fluid.getArray3dU(), u);
    copyOut(*fluid.getArray3dV(), v);
    copyOut(*fluid.getArray3dW(), w);
    boolOut(valid.validU, vu);
    boolOut(valid.validV, vv);
    boolOut(valid.validW, vw);
    if (cluster && params.isSurfaceTensionEnabled) copyOut(ps._surfaceTensionClusterStatus, cluster);
    return ok ? 1 : 0;
}

/* ViscositySolver stage by stage (viscositysolver.cpp:50-79): the solver's
 * own face states, volume grids, matrix index and linear system for the
 * given fields; outputs as oracle_viscosity_system_get (NULL skips).  Returns
 * the matrix size and writes the nonzero count. */
namespace {
struct ViscInputs {
    MACVelocityField vf;
    ParticleLevelSet pls;
    MeshLevelSet solid;
    Array3d<float> visc;
    ViscInputs(int ni, int nj, int nk, double dx, const float *liquid, const float *sphi, const float *vg,
               const float *u, const float *v, const float *w)
        : vf(ni, nj, nk, dx), pls(ni, nj, nk, dx), solid(ni, nj, nk, dx), visc(ni + 1, nj + 1, nk + 1, 0.0f) {
        copyIn(*vf.getArray3dU(), u);
        copyIn(*vf.getArray3dV(), v);
        copyIn(*vf.getArray3dW(), w);
        copyIn(pls._phi, liquid);
        copyIn(*solid.getPhiArray3d(), sphi);
        copyIn(visc, vg);
    }
    ViscositySolverParameters params(double dx, double dt, double tol, int maxIter) {
        ViscositySolverParameters p;
        p.cellwidth = dx;
        p.deltaTime = dt;
        p.velocityField = &vf;
        p.liquidSDF = &pls;
        p.solidSDF = &solid;
        p.viscosity = &visc;
        p.errorTolerance = tol;
        p.maxIterations = maxIter;
        return p;
    }
};
}  // namespace

int refprobe_viscosity_system(int ni, int nj, int nk, double dx, double dt, const float *liquid,
                              const float *solid_nodes, const float *viscosity, const float *u, const float *v,
                              const float *w, int *nnz, unsigned char *states, float *volumes, int *table,
                              int *rowstart, int *cols, float *vals, float *rhs) {
    ViscInputs in(ni, nj, nk, dx, liquid, solid_nodes, viscosity, u, v, w);
    ViscositySolver vs;
    vs._initialize(in.params(dx, dt, 1e-4, 900));
    vs._computeFaceStateGrid();
    vs._computeVolumeGrid();
    vs._computeMatrixIndexTable();
    const int m = vs._matrixIndex.matrixSize;
    SparseMatrixf matrix(m, 15);
    std::vector<float> b(m, 0.0f);
    if (m > 0) vs._initializeLinearSystem(matrix, b);
    int q = 0;
    for (int r = 0; r < m; r++) q += (int)matrix.index[r].size();
    *nnz = q;
    if (states) {
        for (auto *g : {&vs._state.U, &vs._state.V, &vs._state.W}) {
            size_t n = (size_t)g->width * g->height * g->depth;
            for (size_t i = 0; i < n; i++) *states++ = (unsigned char)g->getRawArray()[i];
        }
    }
    if (volumes) {
        auto &V = vs._volumes;
        for (Array3d<float> *g : {&V.center, &V.U, &V.V, &V.W, &V.edgeU, &V.edgeV, &V.edgeW}) {
            copyOut(*g, volumes);
            volumes += (size_t)g->width * g->height * g->depth;
        }
    }
    if (table) std::copy(vs._matrixIndex.indexTable.begin(), vs._matrixIndex.indexTable.end(), table);
    if (rowstart) {
        q = 0;
        rowstart[0] = 0;
        for (int r = 0; r < m; r++) {
            for (size_t e = 0; e < matrix.index[r].size(); e++) {
                if (cols) cols[q] = (int)matrix.index[r][e];
                if (vals) vals[q] = matrix.value[r][e];
                q++;
            }
            rowstart[r + 1] = q;
        }
    }
    if (rhs) std::copy(b.begin(), b.end(), rhs);
    return m;
}

/* ViscositySolver::applyViscosityToVelocityField on the given fields; u, v, w
 * updated in place.  Returns 1 on success. */
int refprobe_viscosity_solve(int ni, int nj, int nk, double dx, double dt, const float *liquid,
                             const float *solid_nodes, const float *viscosity, double tol, int maxIter, float *u,
                             float *v, float *w, int *iterations, float *error) {
    ViscInputs in(ni, nj, nk, dx, liquid, solid_nodes, viscosity, u, v, w);
    ViscositySolver vs;
    bool ok = vs.applyViscosityToVelocityField(in.params(dx, dt, tol, maxIter));
    *iterations = vs.getIterations();
    *error = vs.getError();
    copyOut(*in.vf.getArray3dU(), u);
    copyOut(*in.vf.getArray3dV(), v);
    copyOut(*in.vf.getArray3dW(), w);
    return ok ? 1 : 0;
}

int refprobe_sim_last_substeps(void *s) { return ((FluidSimulation *)s)->_outputData.frameData.substeps; }
int refprobe_sim_pressure_iterations(void *s) { return ((FluidSimulation *)s)->_pressureSolverIterations; }
int refprobe_sim_viscosity_stats(void *s, int *iterations, float *error) {
    FluidSimulation *f = (FluidSimulation *)s;
    *iterations = f->_viscositySolverIterations;
    *error = f->_viscositySolverError;
    return f->_viscositySolverSuccess ? 1 : 0;
}

}  // extern "C"
