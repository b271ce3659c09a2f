// capi.hip -- the extern "C" boundary of libffengine.so.
//
// Every entry point converts exceptions into *err = 0 plus one process-wide
// message buffer, like the reference's safe_execute_* wrappers
// (c_bindings/cbindings.h:48-155, cbindings.cpp:35-43).
#include <cstring>
#include <exception>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "engine.h"
#include "rollout.h"

namespace {
const int SUCCESS = 1, FAIL = 0;
char g_error[4096];

void setError(const std::exception &ex) {
    std::string msg = ex.what();
    size_t n = std::min(msg.size(), sizeof(g_error) - 1);
    std::memcpy(g_error, msg.data(), n);
    g_error[n] = '\0';
}

template <class F>
auto guarded(int *err, F &&f) -> decltype(f()) {
    *err = SUCCESS;
    try {
        return f();
    } catch (const std::exception &ex) {
        setError(ex);
        *err = FAIL;
    }
    return decltype(f())();
}

ff::MeshObstacle *M(MeshObject *o) {
    if (!o) throw std::runtime_error("Error: null MeshObject handle.\n");
    return reinterpret_cast<ff::MeshObstacle *>(o);
}

ff::FluidSimulation *S(FluidSimulation *o) {
    if (!o) throw std::runtime_error("Error: null FluidSimulation handle.\n");
    return reinterpret_cast<ff::FluidSimulation *>(o);
}

// A scratch engine state on device for the kernel-level entry points.
struct Scratch {
    hipStream_t s = nullptr;
    Scratch() { ff::hipCheck(hipStreamCreateWithFlags(&s, hipStreamNonBlocking), "hipStreamCreate"); }
    ~Scratch() {
        if (s) (void)hipStreamDestroy(s);
    }
};
template <class T>
struct DBuf {
    T *p = nullptr;
    explicit DBuf(size_t n) { ff::hipCheck(hipMalloc((void **)&p, sizeof(T) * (n ? n : 1)), "hipMalloc"); }
    ~DBuf() {
        if (p) (void)hipFree(p);
    }
    operator T *() const { return p; }
};
template <class T>
void up(T *d, const T *h, size_t n, hipStream_t s) {
    if (n) ff::hipCheck(hipMemcpyAsync(d, h, sizeof(T) * n, hipMemcpyHostToDevice, s), "H2D");
}
template <class T>
void down(T *h, const T *d, size_t n, hipStream_t s) {
    if (n) ff::hipCheck(hipMemcpyAsync(h, d, sizeof(T) * n, hipMemcpyDeviceToHost, s), "D2H");
}
void sync(hipStream_t s) { ff::hipCheck(hipStreamSynchronize(s), "sync"); }
void checkGrid(int i, int j, int k, double dx) {
    if (i <= 0 || j <= 0 || k <= 0 || !(dx > 0.0))
        throw std::domain_error("Error: dimensions and cell size must be greater than 0.\n");
}
}  // namespace

extern "C" {

char *CBindings_get_error_message(void) { return g_error; }

FluidSimulation *FluidSimulation_new_from_dimensions(int isize, int jsize, int ksize, double dx, int *err) {
    return guarded(err, [&] {
        return reinterpret_cast<FluidSimulation *>(new ff::FluidSimulation(isize, jsize, ksize, dx));
    });
}
void FluidSimulation_destroy(FluidSimulation *obj) {
    try {
        delete reinterpret_cast<ff::FluidSimulation *>(obj);
    } catch (...) {
    }
}
void FluidSimulation_initialize(FluidSimulation *o, int *err) {
    guarded(err, [&] { S(o)->initialize(); });
}
int FluidSimulation_is_initialized(FluidSimulation *o, int *err) {
    return guarded(err, [&] { return (int)S(o)->isInitialized(); });
}
void FluidSimulation_update(FluidSimulation *o, double dt, int *err) {
    guarded(err, [&] { S(o)->update(dt); });
}
int FluidSimulation_get_current_frame(FluidSimulation *o, int *err) {
    return guarded(err, [&] { return S(o)->getCurrentFrame(); });
}
void FluidSimulation_set_current_frame(FluidSimulation *o, int f, int *err) {
    guarded(err, [&] { S(o)->setCurrentFrame(f); });
}
int FluidSimulation_is_current_frame_finished(FluidSimulation *o, int *err) {
    return guarded(err, [&] { return (int)S(o)->isCurrentFrameFinished(); });
}
double FluidSimulation_get_cell_size(FluidSimulation *o, int *err) {
    return guarded(err, [&] { return S(o)->getCellSize(); });
}
void FluidSimulation_get_grid_dimensions(FluidSimulation *o, int *i, int *j, int *k, int *err) {
    guarded(err, [&] { S(o)->getGridDimensions(i, j, k); });
}
double FluidSimulation_get_density(FluidSimulation *o, int *err) {
    return guarded(err, [&] { return S(o)->getDensity(); });
}
void FluidSimulation_set_density(FluidSimulation *o, double d, int *err) {
    guarded(err, [&] { S(o)->setDensity(d); });
}
int FluidSimulation_get_pressure_solver_max_iterations(FluidSimulation *o, int *err) {
    return guarded(err, [&] { return S(o)->getPressureSolverMaxIterations(); });
}
void FluidSimulation_set_pressure_solver_max_iterations(FluidSimulation *o, int n, int *err) {
    guarded(err, [&] { S(o)->setPressureSolverMaxIterations(n); });
}
void FluidSimulation_enable_surface_reconstruction(FluidSimulation *o, int *err) {
    guarded(err, [&] { S(o)->enableSurfaceReconstruction(); });
}
void FluidSimulation_disable_surface_reconstruction(FluidSimulation *o, int *err) {
    guarded(err, [&] { S(o)->disableSurfaceReconstruction(); });
}
int FluidSimulation_is_surface_reconstruction_enabled(FluidSimulation *o, int *err) {
    return guarded(err, [&] { return (int)S(o)->isSurfaceReconstructionEnabled(); });
}
int FluidSimulation_get_max_thread_count(FluidSimulation *o, int *err) {
    return guarded(err, [&] { return S(o)->getMaxThreadCount(); });
}
void FluidSimulation_set_max_thread_count(FluidSimulation *o, int n, int *err) {
    guarded(err, [&] { S(o)->setMaxThreadCount(n); });
}
void FluidSimulation_add_body_force(FluidSimulation *o, double fx, double fy, double fz, int *err) {
    guarded(err, [&] { S(o)->addBodyForce(fx, fy, fz); });
}
FFVector3 FluidSimulation_get_constant_body_force(FluidSimulation *o, int *err) {
    return guarded(err, [&] { return S(o)->getConstantBodyForce(); });
}
void FluidSimulation_reset_body_force(FluidSimulation *o, int *err) {
    guarded(err, [&] { S(o)->resetBodyForce(); });
}
int FluidSimulation_get_CFL_condition_number(FluidSimulation *o, int *err) {
    return guarded(err, [&] { return S(o)->getCFLConditionNumber(); });
}
void FluidSimulation_set_CFL_condition_number(FluidSimulation *o, int n, int *err) {
    guarded(err, [&] { S(o)->setCFLConditionNumber(n); });
}
int FluidSimulation_get_min_time_steps_per_frame(FluidSimulation *o, int *err) {
    return guarded(err, [&] { return S(o)->getMinTimeStepsPerFrame(); });
}
void FluidSimulation_set_min_time_steps_per_frame(FluidSimulation *o, int n, int *err) {
    guarded(err, [&] { S(o)->setMinTimeStepsPerFrame(n); });
}
int FluidSimulation_get_max_time_steps_per_frame(FluidSimulation *o, int *err) {
    return guarded(err, [&] { return S(o)->getMaxTimeStepsPerFrame(); });
}
void FluidSimulation_set_max_time_steps_per_frame(FluidSimulation *o, int n, int *err) {
    guarded(err, [&] { S(o)->setMaxTimeStepsPerFrame(n); });
}
void FluidSimulation_enable_extreme_velocity_removal(FluidSimulation *o, int *err) {
    guarded(err, [&] { S(o)->enableExtremeVelocityRemoval(); });
}
void FluidSimulation_disable_extreme_velocity_removal(FluidSimulation *o, int *err) {
    guarded(err, [&] { S(o)->disableExtremeVelocityRemoval(); });
}
int FluidSimulation_is_extreme_velocity_removal_enabled(FluidSimulation *o, int *err) {
    return guarded(err, [&] { return (int)S(o)->isExtremeVelocityRemovalEnabled(); });
}
void FluidSimulation_set_velocity_transfer_method_FLIP(FluidSimulation *o, int *err) {
    guarded(err, [&] { S(o)->setVelocityTransferMethodFLIP(); });
}
void FluidSimulation_set_velocity_transfer_method_APIC(FluidSimulation *o, int *err) {
    guarded(err, [&] { S(o)->setVelocityTransferMethodAPIC(); });
}
int FluidSimulation_is_velocity_transfer_method_FLIP(FluidSimulation *o, int *err) {
    return guarded(err, [&] { return (int)S(o)->isVelocityTransferMethodFLIP(); });
}
int FluidSimulation_is_velocity_transfer_method_APIC(FluidSimulation *o, int *err) {
    return guarded(err, [&] { return (int)S(o)->isVelocityTransferMethodAPIC(); });
}
double FluidSimulation_get_PICAPIC_ratio(FluidSimulation *o, int *err) {
    return guarded(err, [&] { return S(o)->getPICAPICRatio(); });
}
void FluidSimulation_set_PICAPIC_ratio(FluidSimulation *o, double r, int *err) {
    guarded(err, [&] { S(o)->setPICAPICRatio(r); });
}
double FluidSimulation_get_PICFLIP_ratio(FluidSimulation *o, int *err) {
    return guarded(err, [&] { return S(o)->getPICFLIPRatio(); });
}
void FluidSimulation_set_PICFLIP_ratio(FluidSimulation *o, double r, int *err) {
    guarded(err, [&] { S(o)->setPICFLIPRatio(r); });
}

int FluidSimulation_get_viscosity_solver_max_iterations(FluidSimulation *o, int *err) {
    return guarded(err, [&] { return S(o)->getViscositySolverMaxIterations(); });
}
void FluidSimulation_set_viscosity_solver_max_iterations(FluidSimulation *o, int n, int *err) {
    guarded(err, [&] { S(o)->setViscositySolverMaxIterations(n); });
}
double FluidSimulation_get_viscosity(FluidSimulation *o, int *err) {
    return guarded(err, [&] { return S(o)->getViscosity(); });
}
void FluidSimulation_set_viscosity(FluidSimulation *o, double v, int *err) {
    guarded(err, [&] { S(o)->setViscosity(v); });
}
double FluidSimulation_get_viscosity_solver_error_tolerance(FluidSimulation *o, int *err) {
    return guarded(err, [&] { return S(o)->getViscositySolverErrorTolerance(); });
}
void FluidSimulation_set_viscosity_solver_error_tolerance(FluidSimulation *o, double tol, int *err) {
    guarded(err, [&] { S(o)->setViscositySolverErrorTolerance(tol); });
}

double FluidSimulation_get_surface_tension(FluidSimulation *o, int *err) {
    return guarded(err, [&] { return S(o)->getSurfaceTension(); });
}
void FluidSimulation_set_surface_tension(FluidSimulation *o, double k, int *err) {
    guarded(err, [&] { S(o)->setSurfaceTension(k); });
}
double FluidSimulation_get_surface_tension_condition_number(FluidSimulation *o, int *err) {
    return guarded(err, [&] { return S(o)->getSurfaceTensionConditionNumber(); });
}
void FluidSimulation_set_surface_tension_condition_number(FluidSimulation *o, double n, int *err) {
    guarded(err, [&] { S(o)->setSurfaceTensionConditionNumber(n); });
}
void FluidSimulation_enable_smooth_surface_tension_kernel(FluidSimulation *o, int *err) {
    guarded(err, [&] { S(o)->setSmoothSurfaceTensionKernel(true); });
}
void FluidSimulation_disable_smooth_surface_tension_kernel(FluidSimulation *o, int *err) {
    guarded(err, [&] { S(o)->setSmoothSurfaceTensionKernel(false); });
}
int FluidSimulation_is_smooth_surface_tension_kernel_enabled(FluidSimulation *o, int *err) {
    return guarded(err, [&] { return S(o)->isSmoothSurfaceTensionKernelEnabled() ? 1 : 0; });
}

void FluidSimulation_load_marker_particle_data(FluidSimulation *o, FluidSimulationMarkerParticleData d,
                                               int *err) {
    guarded(err, [&] { S(o)->loadMarkerParticleData(d.size, d.positions, d.velocities); });
}
void FluidSimulation_load_marker_particle_affine_data(FluidSimulation *o, FluidSimulationMarkerParticleAffineData d,
                                                      int *err) {
    guarded(err, [&] { S(o)->loadMarkerParticleAffineData(d.size, d.affineX, d.affineY, d.affineZ); });
}
void FluidSimulation_get_marker_particle_affinex_data_range(FluidSimulation *o, int a, int b, char *data, int *err) {
    guarded(err, [&] { S(o)->getMarkerParticleAffineDataRange(0, a, b, data); });
}
void FluidSimulation_get_marker_particle_affiney_data_range(FluidSimulation *o, int a, int b, char *data, int *err) {
    guarded(err, [&] { S(o)->getMarkerParticleAffineDataRange(1, a, b, data); });
}
void FluidSimulation_get_marker_particle_affinez_data_range(FluidSimulation *o, int a, int b, char *data, int *err) {
    guarded(err, [&] { S(o)->getMarkerParticleAffineDataRange(2, a, b, data); });
}
int FluidSimulation_get_num_marker_particles(FluidSimulation *o, int *err) {
    return guarded(err, [&] { return S(o)->getNumMarkerParticles(); });
}
unsigned int FluidSimulation_get_marker_particle_position_data_size(FluidSimulation *o, int *err) {
    return guarded(err, [&] { return (unsigned int)(12u * (unsigned)S(o)->getNumMarkerParticles()); });
}
unsigned int FluidSimulation_get_marker_particle_velocity_data_size(FluidSimulation *o, int *err) {
    return guarded(err, [&] { return (unsigned int)(12u * (unsigned)S(o)->getNumMarkerParticles()); });
}
void FluidSimulation_get_marker_particle_position_data(FluidSimulation *o, char *data, int *err) {
    guarded(err, [&] { S(o)->getMarkerParticlePositionData(data); });
}
void FluidSimulation_get_marker_particle_velocity_data(FluidSimulation *o, char *data, int *err) {
    guarded(err, [&] { S(o)->getMarkerParticleVelocityData(data); });
}
void FluidSimulation_get_marker_particle_position_data_range(FluidSimulation *o, int a, int b, char *data,
                                                             int *err) {
    guarded(err, [&] { S(o)->getMarkerParticlePositionDataRange(a, b, data); });
}
void FluidSimulation_get_marker_particle_velocity_data_range(FluidSimulation *o, int a, int b, char *data,
                                                             int *err) {
    guarded(err, [&] { S(o)->getMarkerParticleVelocityDataRange(a, b, data); });
}
FluidSimulationFrameStats FluidSimulation_get_frame_stats_data(FluidSimulation *o, int *err) {
    return guarded(err, [&] { return S(o)->getFrameStatsData(); });
}

void FluidSimulation_set_pressure_solver_tolerance(FluidSimulation *o, double tol, int *err) {
    guarded(err, [&] { S(o)->setPressureSolverTolerance(tol); });
}
double FluidSimulation_get_pressure_solver_tolerance(FluidSimulation *o, int *err) {
    return guarded(err, [&] { return S(o)->getPressureSolverTolerance(); });
}
void FluidSimulation_set_pressure_solver_precision(FluidSimulation *o, int mode, double inner, int *err) {
    guarded(err, [&] { S(o)->setPressureSolverPrecision(mode, inner); });
}
int FluidSimulation_get_pressure_solver_precision(FluidSimulation *o, int *err) {
    return guarded(err, [&] { return S(o)->getPressureSolverPrecision(); });
}
void FluidSimulation_set_pressure_solver_reference_summation(FluidSimulation *o, int on, int *err) {
    guarded(err, [&] { S(o)->setPressureSolverReferenceSummation(on != 0); });
}
int FluidSimulation_get_pressure_solver_reference_summation(FluidSimulation *o, int *err) {
    return guarded(err, [&] { return S(o)->getPressureSolverReferenceSummation() ? 1 : 0; });
}
void FluidSimulation_set_pressure_solver_warm_start(FluidSimulation *o, int on, int *err) {
    guarded(err, [&] { S(o)->setPressureSolverWarmStart(on != 0); });
}
int FluidSimulation_get_pressure_solver_warm_start(FluidSimulation *o, int *err) {
    return guarded(err, [&] { return S(o)->getPressureSolverWarmStart() ? 1 : 0; });
}
void FluidSimulation_set_pressure_solver_preconditioner_schedule(FluidSimulation *o, int schedule, int *err) {
    guarded(err, [&] { S(o)->setPressureSolverPreconditionerSchedule(schedule); });
}
int FluidSimulation_get_pressure_solver_preconditioner_schedule(FluidSimulation *o, int *err) {
    return guarded(err, [&] { return S(o)->getPressureSolverPreconditionerSchedule(); });
}
// ---- static mesh obstacles ----
MeshObject *MeshObject_new(int i, int j, int k, double dx, int *err) {
    return guarded(err, [&] {
        checkGrid(i, j, k, dx);
        return reinterpret_cast<MeshObject *>(new ff::MeshObstacle(i, j, k, dx));
    });
}
void MeshObject_destroy(MeshObject *obj) { delete reinterpret_cast<ff::MeshObstacle *>(obj); }
void MeshObject_update_mesh_static(MeshObject *obj, TriangleMesh_t mesh, int *err) {
    guarded(err, [&] { M(obj)->updateMeshStatic(mesh.numVertices, mesh.vertices, mesh.numTriangles, mesh.triangles); });
}
void MeshObject_update_mesh_animated(MeshObject *obj, TriangleMesh_t previous, TriangleMesh_t current,
                                     TriangleMesh_t next, int *err) {
    guarded(err, [&] {
        ff::MeshObstacle *o = M(obj);
        if (previous.numVertices != current.numVertices || next.numVertices != current.numVertices ||
            previous.numTriangles != current.numTriangles || next.numTriangles != current.numTriangles)
            throw std::runtime_error("Error: animated mesh obstacles must keep their vertex and triangle counts "
                                     "between frames.\n");
        o->updateMeshAnimated(current.numVertices, previous.vertices, current.vertices, next.vertices,
                              current.numTriangles, current.triangles);
    });
}
void MeshObject_enable(MeshObject *obj, int *err) {
    guarded(err, [&] { M(obj)->enable(); });
}
void MeshObject_disable(MeshObject *obj, int *err) {
    guarded(err, [&] { M(obj)->disable(); });
}
int MeshObject_is_enabled(MeshObject *obj, int *err) {
    return guarded(err, [&] { return (int)M(obj)->enabled; });
}
void MeshObject_inverse(MeshObject *obj, int *err) {
    guarded(err, [&] { M(obj)->inverse(); });
}
int MeshObject_is_inversed(MeshObject *obj, int *err) {
    return guarded(err, [&] { return (int)M(obj)->inversed; });
}
float MeshObject_get_friction(MeshObject *obj, int *err) {
    return guarded(err, [&] { return M(obj)->friction; });
}
void MeshObject_set_friction(MeshObject *obj, float f, int *err) {
    guarded(err, [&] { M(obj)->setFriction(f); });
}
float MeshObject_get_velocity_scale(MeshObject *obj, int *err) {
    return guarded(err, [&] { return M(obj)->velocityScale; });
}
void MeshObject_set_velocity_scale(MeshObject *obj, float scale, int *err) {
    guarded(err, [&] { M(obj)->velocityScale = scale; });
}
float MeshObject_get_whitewater_influence(MeshObject *obj, int *err) {
    return guarded(err, [&] { return M(obj)->whitewaterInfluence; });
}
void MeshObject_set_whitewater_influence(MeshObject *obj, float value, int *err) {
    guarded(err, [&] { M(obj)->setNonNegative(M(obj)->whitewaterInfluence, value); });
}
float MeshObject_get_dust_emission_strength(MeshObject *obj, int *err) {
    return guarded(err, [&] { return M(obj)->dustEmissionStrength; });
}
void MeshObject_set_dust_emission_strength(MeshObject *obj, float value, int *err) {
    guarded(err, [&] { M(obj)->setNonNegative(M(obj)->dustEmissionStrength, value); });
}
float MeshObject_get_sheeting_strength(MeshObject *obj, int *err) {
    return guarded(err, [&] { return M(obj)->sheetingStrength; });
}
void MeshObject_set_sheeting_strength(MeshObject *obj, float value, int *err) {
    guarded(err, [&] { M(obj)->setNonNegative(M(obj)->sheetingStrength, value); });
}
float MeshObject_get_mesh_expansion(MeshObject *obj, int *err) {
    return guarded(err, [&] { return M(obj)->meshExpansion; });
}
void MeshObject_set_mesh_expansion(MeshObject *obj, float ex, int *err) {
    guarded(err, [&] { M(obj)->meshExpansion = ex; });
}
void MeshObject_enable_append_object_velocity(MeshObject *obj, int *err) {
    guarded(err, [&] { M(obj)->appendObjectVelocity = true; });
}
void MeshObject_disable_append_object_velocity(MeshObject *obj, int *err) {
    guarded(err, [&] { M(obj)->appendObjectVelocity = false; });
}
int MeshObject_is_append_object_velocity_enabled(MeshObject *obj, int *err) {
    return guarded(err, [&] { return (int)M(obj)->appendObjectVelocity; });
}
float MeshObject_get_object_velocity_influence(MeshObject *obj, int *err) {
    return guarded(err, [&] { return M(obj)->objectVelocityInfluence; });
}
void MeshObject_set_object_velocity_influence(MeshObject *obj, float value, int *err) {
    guarded(err, [&] { M(obj)->objectVelocityInfluence = value; });
}
int MeshObject_get_priority(MeshObject *obj, int *err) {
    return guarded(err, [&] { return M(obj)->priority; });
}
void MeshObject_set_priority(MeshObject *obj, int n, int *err) {
    guarded(err, [&] { M(obj)->priority = n; });
}
int MeshObject_get_source_id(MeshObject *obj, int *err) {
    return guarded(err, [&] { return M(obj)->sourceID; });
}
void MeshObject_set_source_id(MeshObject *obj, int id, int *err) {
    guarded(err, [&] { M(obj)->sourceID = id; });
}
// the reference declares these four float getters as returning int
int MeshObject_get_viscosity(MeshObject *obj, int *err) {
    return guarded(err, [&] { return (int)M(obj)->viscosity; });
}
void MeshObject_set_viscosity(MeshObject *obj, float v, int *err) {
    guarded(err, [&] { M(obj)->viscosity = v; });
}
int MeshObject_get_density(MeshObject *obj, int *err) {
    return guarded(err, [&] { return (int)M(obj)->density; });
}
void MeshObject_set_density(MeshObject *obj, float d, int *err) {
    guarded(err, [&] { M(obj)->density = d; });
}
int MeshObject_get_lifetime(MeshObject *obj, int *err) {
    return guarded(err, [&] { return (int)M(obj)->lifetime; });
}
void MeshObject_set_lifetime(MeshObject *obj, float v, int *err) {
    guarded(err, [&] { M(obj)->lifetime = v; });
}
int MeshObject_get_lifetime_variance(MeshObject *obj, int *err) {
    return guarded(err, [&] { return (int)M(obj)->lifetimeVariance; });
}
void MeshObject_set_lifetime_variance(MeshObject *obj, float v, int *err) {
    guarded(err, [&] { M(obj)->lifetimeVariance = v; });
}
FFVector3 MeshObject_get_source_color(MeshObject *obj, int *err) {
    return guarded(err, [&] { return M(obj)->sourceColor; });
}
void MeshObject_set_source_color(MeshObject *obj, double r, double g, double b, int *err) {
    guarded(err, [&] { M(obj)->sourceColor = FFVector3{(float)r, (float)g, (float)b}; });
}
void FluidSimulation_add_mesh_obstacle(FluidSimulation *o, MeshObject *obstacle, int *err) {
    guarded(err, [&] { S(o)->addMeshObstacle(M(obstacle)); });
}
void FluidSimulation_remove_mesh_obstacle(FluidSimulation *o, MeshObject *obstacle, int *err) {
    guarded(err, [&] { S(o)->removeMeshObstacle(M(obstacle)); });
}
void FluidSimulation_remove_mesh_obstacles(FluidSimulation *o, int *err) {
    guarded(err, [&] { S(o)->removeMeshObstacles(); });
}
void FluidSimulation_set_boundary_friction(FluidSimulation *o, double f, int *err) {
    guarded(err, [&] { S(o)->setBoundaryFriction(f); });
}
double FluidSimulation_get_boundary_friction(FluidSimulation *o, int *err) {
    return guarded(err, [&] { return S(o)->getBoundaryFriction(); });
}
int FFAMD_get_last_solid_rand_draws(FluidSimulation *o, int *err) {
    return guarded(err, [&] { return S(o)->getLastSolidRandDraws(); });
}

void FluidSimulation_set_solid_phi(FluidSimulation *o, const float *phi, int *err) {
    guarded(err, [&] { S(o)->setSolidPhi(phi); });
}
void FluidSimulation_get_liquid_sdf_data(FluidSimulation *o, float *d, int *err) {
    guarded(err, [&] { S(o)->getLiquidSDF(d); });
}
void FluidSimulation_get_solid_velocity_field(FluidSimulation *o, float *u, float *v, float *w, int *err) {
    guarded(err, [&] { S(o)->getSolidVelocityField(u, v, w); });
}
void FluidSimulation_get_solid_sdf_data(FluidSimulation *o, float *d, int *err) {
    guarded(err, [&] { S(o)->getSolidSDF(d); });
}
void FluidSimulation_get_velocity_field_data(FluidSimulation *o, float *u, float *v, float *w, int *err) {
    guarded(err, [&] { S(o)->getVelocityField(u, v, w); });
}
void FluidSimulation_get_pressure_grid_data(FluidSimulation *o, float *d, int *err) {
    guarded(err, [&] { S(o)->getPressureGrid(d); });
}
void FluidSimulation_get_last_pressure_solve_stats(FluidSimulation *o, int *it, double *e, int *ok, int *m,
                                                   int *err) {
    guarded(err, [&] { S(o)->getLastPressureSolveStats(it, e, ok, m); });
}
double FluidSimulation_get_last_pressure_solve_rhs_norm(FluidSimulation *o, int *err) {
    return guarded(err, [&] { return S(o)->getLastPressureRhsNorm(); });
}
void FluidSimulation_get_last_viscosity_solve_stats(FluidSimulation *o, int *it, double *e, int *ok, int *m,
                                                    int *err) {
    guarded(err, [&] { S(o)->getLastViscositySolveStats(it, e, ok, m); });
}
double FluidSimulation_get_last_substep_phase_ms(FluidSimulation *o, int phase, int *err) {
    return guarded(err, [&] { return S(o)->getLastPhaseMilliseconds(phase); });
}
void FluidSimulation_enable_kernel_timing(FluidSimulation *o, int on, int *err) {
    guarded(err, [&] { S(o)->enableKernelTiming(on != 0); });
}
void FluidSimulation_get_kernel_timing(FluidSimulation *o, int id, double *ms, long long *launches, int *err) {
    guarded(err, [&] { S(o)->getKernelTiming(id, ms, launches); });
}
void FluidSimulation_step(FluidSimulation *o, double dt, int *err) {
    guarded(err, [&] { S(o)->step(dt); });
}

// ---------------------------------------------------------------------------
// kernel-level entry points on host arrays

void FFAMD_cell_sort(int ni, int nj, int nk, double dx, int np, const float *pos, int *keys, int *counts,
                     int *offsets, int *perm, int *err) {
    guarded(err, [&] {
        checkGrid(ni, nj, nk, dx);
        ff::Grid g{ni, nj, nk, dx};
        size_t N = (size_t)ni * nj * nk, P = (size_t)np;
        Scratch sc;
        std::vector<float> c(P);
        DBuf<float> px(P), py(P), pz(P), vx(P), vy(P), vz(P);
        for (int a = 0; a < 3; a++) {
            for (size_t q = 0; q < P; q++) c[q] = pos[3 * q + a];
            up(a == 0 ? px.p : (a == 1 ? py.p : pz.p), c.data(), P, sc.s);
            sync(sc.s);
        }
        DBuf<int> dk(P), ds(P), dc(N + 1), doff(N + 1), dp(P), scan(ff::scanScratchInts(std::max(N + 1, P + 1)));
        DBuf<float> s0(P), s1(P), s2(P), s3(P), s4(P), s5(P);
        ff::Particles pp{px, py, pz, vx, vy, vz, np};
        ff::CellSort cs{dk, ds, dc, doff, dp, s0, s1, s2, s3, s4, s5};
        ff::cellSort(g, pp, cs, scan, sc.s);
        down(keys, dk.p, P, sc.s);
        down(counts, dc.p, N, sc.s);
        down(offsets, doff.p, N + 1, sc.s);
        down(perm, dp.p, P, sc.s);
        sync(sc.s);
    });
}

namespace {
struct SortedParticles {
    ff::CellSort cs;
    std::vector<std::unique_ptr<DBuf<float>>> f;
    std::vector<std::unique_ptr<DBuf<int>>> i;
    SortedParticles(ff::Grid g, int np, const float *pos, const float *vel, hipStream_t s) {
        size_t N = (size_t)g.ni * g.nj * g.nk, P = (size_t)np;
        for (int q = 0; q < 12; q++) f.push_back(std::make_unique<DBuf<float>>(P));
        for (int q = 0; q < 6; q++)
            i.push_back(std::make_unique<DBuf<int>>(
                q == 2 || q == 3 ? N + 1 : (q == 5 ? ff::scanScratchInts(std::max(N + 1, P + 1)) : P)));
        std::vector<float> c(P);
        for (int a = 0; a < 6; a++) {
            const float *src = a < 3 ? pos : vel;
            for (size_t q = 0; q < P; q++) c[q] = src ? src[3 * q + a % 3] : 0.0f;
            up(f[a]->p, c.data(), P, s);
            sync(s);
        }
        ff::Particles pp{*f[0], *f[1], *f[2], *f[3], *f[4], *f[5], np};
        cs = ff::CellSort{*i[0], *i[1], *i[2], *i[3], *i[4], *f[6], *f[7], *f[8], *f[9], *f[10], *f[11]};
        ff::cellSort(g, pp, cs, *i[5], s);
    }
};
}  // namespace

void FFAMD_liquid_sdf(int ni, int nj, int nk, double dx, int np, const float *pos, double radius, float *phi,
                      int *err) {
    guarded(err, [&] {
        checkGrid(ni, nj, nk, dx);
        ff::Grid g{ni, nj, nk, dx};
        Scratch sc;
        size_t N = (size_t)ni * nj * nk;
        SortedParticles sp(g, np, pos, nullptr, sc.s);
        DBuf<float> d(N);
        ff::liquidSDF(g, sp.cs, np, radius, d, sc.s);
        down(phi, d.p, N, sc.s);
        sync(sc.s);
    });
}

void FFAMD_p2g(int ni, int nj, int nk, double dx, int np, const float *pos, const float *vel, double radius,
               float *u, float *v, float *w, unsigned char *vu, unsigned char *vv, unsigned char *vw, int *err) {
    guarded(err, [&] {
        checkGrid(ni, nj, nk, dx);
        ff::Grid g{ni, nj, nk, dx};
        Scratch sc;
        size_t nu = (size_t)(ni + 1) * nj * nk, nv = (size_t)ni * (nj + 1) * nk, nw = (size_t)ni * nj * (nk + 1);
        SortedParticles sp(g, np, pos, vel, sc.s);
        DBuf<float> du(nu), dv(nv), dw(nw);
        DBuf<uint8_t> bu(nu), bv(nv), bw(nw);
        ff::p2g(g, sp.cs, radius, du, dv, dw, bu, bv, bw, sc.s);
        down(u, du.p, nu, sc.s);
        down(v, dv.p, nv, sc.s);
        down(w, dw.p, nw, sc.s);
        down(vu, bu.p, nu, sc.s);
        down(vv, bv.p, nv, sc.s);
        down(vw, bw.p, nw, sc.s);
        sync(sc.s);
    });
}

void FFAMD_extrapolate(int wi, int wj, int wk, float *field, const unsigned char *valid, int layers, int *err) {
    guarded(err, [&] {
        if (wi < 3 || wj < 3 || wk < 3 || layers < 0 || layers > 250)
            throw std::domain_error("Error: face array dims must be >= 3 and layers in [0, 250].\n");
        Scratch sc;
        size_t n = (size_t)wi * wj * wk;
        DBuf<float> a(n);
        DBuf<uint8_t> vb(n), st(n);
        up(a.p, field, n, sc.s);
        up(vb.p, valid, n, sc.s);
        ff::extrapolateInit(wi, wj, wk, vb, st, sc.s);
        for (int l = 0; l < layers; l++) ff::extrapolateLayer(wi, wj, wk, a, st, l, sc.s);
        down(field, a.p, n, sc.s);
        sync(sc.s);
    });
}

// ---- the layered extrapolation without allocation ----
// Its one scratch buffer, the status byte of every face, lives in this
// file-scope device array: each GPU the library runs on holds its own copy
// until the process ends, at the address hipGetSymbolAddress gives on that
// device (the array's name is only a host-side stand-in).
namespace {
constexpr size_t kNoAllocScratchBytes = (size_t)64 << 20;  // faces of one call, at most
__device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];
constexpr int kMaxDevices = 64;
std::mutex g_scratchLock;
unsigned char *g_scratchAddr[kMaxDevices] = {};

// the scratch array's address on the current device, looked up on the
// device's first call
unsigned char *deviceScratch() {
    int dev = -1;
    ff::hipCheck(hipGetDevice(&dev), "hipGetDevice");
    if (dev < 0 || dev >= kMaxDevices) throw std::runtime_error("device ordinal out of range");
    std::lock_guard<std::mutex> lock(g_scratchLock);
    if (!g_scratchAddr[dev]) {
        void *p = nullptr;
        ff::hipCheck(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
        if (!p) throw std::runtime_error("hipGetSymbolAddress returned no address for the scratch array");
        g_scratchAddr[dev] = (unsigned char *)p;
    }
    return g_scratchAddr[dev];
}
}  // namespace

void FFAMD_extrapolate_noalloc(int wi, int wj, int wk, float *field, const unsigned char *valid, int layers,
                               void *stream, int *err) {
    guarded(err, [&] {
        if (wi < 3 || wj < 3 || wk < 3 || layers < 0 || layers > 250)
            throw std::domain_error("Error: face array dims must be >= 3 and layers in [0, 250].\n");
        const size_t n = (size_t)wi * wj * wk;
        if (n > kNoAllocScratchBytes)
            throw std::domain_error("Error: FFAMD_extrapolate_noalloc takes at most " +
                                    std::to_string(kNoAllocScratchBytes) + " faces per call; got " +
                                    std::to_string(n) + ".\n");
        if (!field || !valid) throw std::domain_error("Error: null field or flag array.\n");
        unsigned char *st = deviceScratch();
        hipStream_t s = (hipStream_t)stream;
        // every status byte is written by the first kernel: nothing to clear
        ff::extrapolateInit(wi, wj, wk, valid, st, s);
        for (int l = 0; l < layers; l++) ff::extrapolateLayer(wi, wj, wk, field, st, l, s);
        ff::hipCheck(hipGetLastError(), "extrapolation launch");
    });
}

void FFAMD_extrapolate_vjp(int wi, int wj, int wk, float *bar, const unsigned char *valid, int layers, int *err) {
    guarded(err, [&] {
        if (wi < 3 || wj < 3 || wk < 3 || layers < 0 || layers > 250)
            throw std::domain_error("Error: face array dims must be >= 3 and layers in [0, 250].\n");
        Scratch sc;
        size_t n = (size_t)wi * wj * wk;
        DBuf<float> a(n), dummy(n);
        DBuf<uint8_t> vb(n), st(n);
        up(a.p, bar, n, sc.s);
        up(vb.p, valid, n, sc.s);
        // the status bytes depend on the flags only: run the forward's layers
        // on a scratch field to get them
        ff::hipCheck(hipMemsetAsync(dummy, 0, sizeof(float) * n, sc.s), "memset");
        ff::extrapolateInit(wi, wj, wk, vb, st, sc.s);
        for (int l = 0; l < layers; l++) ff::extrapolateLayer(wi, wj, wk, dummy, st, l, sc.s);
        ff::extrapolateAdjoint(wi, wj, wk, a, st, layers, sc.s);
        down(bar, a.p, n, sc.s);
        sync(sc.s);
    });
}

void FFAMD_solid_box_phi(int ni, int nj, int nk, double dx, float *phi, int *err) {
    guarded(err, [&] {
        checkGrid(ni, nj, nk, dx);
        Scratch sc;
        size_t nn = (size_t)(ni + 1) * (nj + 1) * (nk + 1);
        DBuf<float> d(nn);
        ff::solidBoxPhi(ff::Grid{ni, nj, nk, dx}, d, sc.s);
        down(phi, d.p, nn, sc.s);
        sync(sc.s);
    });
}

int FFAMD_mesh_levelset_fast(int isize, int jsize, int ksize, double dx, int oi, int oj, int ok, int nv,
                             const float *vertices, int nt, const int *triangles, float *phi, int *err) {
    return guarded(err, [&] {
        checkGrid(isize, jsize, ksize, dx);
        Scratch sc;
        size_t nn = (size_t)(isize + 1) * (jsize + 1) * (ksize + 1);
        DBuf<float> d(nn);
        ff::ObstacleMesh m{vertices, nv, triangles, nt, false, 0.0f};
        int draws = ff::meshLevelSetFastDevice(ff::Grid{isize, jsize, ksize, dx}, ff::I3{oi, oj, ok}, m, d, sc.s);
        down(phi, d.p, nn, sc.s);
        sync(sc.s);
        return draws;
    });
}

int FFAMD_solid_phi_meshes(int ni, int nj, int nk, double dx, int nobj, const int *nv, const float *vertices,
                           const int *nt, const int *triangles, const int *inversed, const float *expansion,
                           float *phi, int *err) {
    return guarded(err, [&] {
        checkGrid(ni, nj, nk, dx);
        if (nobj < 0) throw std::invalid_argument("Error: negative obstacle count.\n");
        Scratch sc;
        size_t nn = (size_t)(ni + 1) * (nj + 1) * (nk + 1);
        DBuf<float> d(nn);
        std::vector<ff::ObstacleMesh> meshes;
        size_t vo = 0, to = 0;
        for (int q = 0; q < nobj; q++) {
            meshes.push_back(ff::ObstacleMesh{vertices + 3 * vo, nv[q], triangles + 3 * to, nt[q], inversed[q] != 0,
                                              expansion ? expansion[q] : 0.0f});
            vo += (size_t)nv[q];
            to += (size_t)nt[q];
        }
        ff::solidBoxPhi(ff::Grid{ni, nj, nk, dx}, d, sc.s);
        int draws = ff::meshObstacleSolid(ff::Grid{ni, nj, nk, dx}, meshes.data(), nobj, d, nullptr, false, sc.s);
        down(phi, d.p, nn, sc.s);
        sync(sc.s);
        return draws;
    });
}

int FFAMD_solid_update_moving(int ni, int nj, int nk, double dx, int nobj, const int *nv, const int *nt,
                              const float *prev, const float *cur, const float *next, const int *triangles,
                              const int *flags, const float *expansion, const float *velocity_scale,
                              double frame_dt, float frame_progress, float *phi, float *u, float *v, float *w,
                              int *err) {
    return guarded(err, [&] {
        checkGrid(ni, nj, nk, dx);
        if (nobj < 0) throw std::invalid_argument("Error: negative obstacle count.\n");
        const ff::Grid g{ni, nj, nk, dx};
        Scratch sc;
        const size_t nn = (size_t)(ni + 1) * (nj + 1) * (nk + 1);
        const size_t nf[3] = {(size_t)(ni + 1) * nj * nk, (size_t)ni * (nj + 1) * nk, (size_t)ni * nj * (nk + 1)};
        // the objects as the engine holds them, at the frame progress
        std::vector<ff::MeshObstacle> objs;
        objs.reserve(nobj);
        std::vector<std::vector<float>> pos(nobj), vel(nobj);
        std::vector<ff::ObstacleMesh> meshes;
        size_t vo = 0, to = 0;
        for (int q = 0; q < nobj; q++) {
            objs.emplace_back(ni, nj, nk, dx);
            ff::MeshObstacle &o = objs.back();
            if (flags[q] & 1)
                o.updateMeshAnimated(nv[q], prev + 3 * vo, cur + 3 * vo, next + 3 * vo, nt[q], triangles + 3 * to);
            else
                o.updateMeshStatic(nv[q], cur + 3 * vo, nt[q], triangles + 3 * to);
            o.inversed = (flags[q] & 2) != 0;
            o.meshExpansion = expansion ? expansion[q] : 0.0f;
            o.velocityScale = velocity_scale ? velocity_scale[q] : 1.0f;
            vo += (size_t)nv[q];
            to += (size_t)nt[q];
        }
        for (int q = 0; q < nobj; q++) {
            objs[q].meshAtProgress(frame_progress, frame_dt, pos[q], vel[q]);
            ff::ObstacleMesh m{pos[q].data(), nv[q], objs[q].triangles.data(), nt[q], objs[q].inversed,
                               objs[q].meshExpansion};
            m.velocities = vel[q].data();
            m.animated = objs[q].animated;
            meshes.push_back(m);
        }
        DBuf<float> d(nn), us(nf[0]), vs(nf[1]), ws(nf[2]);
        DBuf<char> block(ff::solidVelocityBytes(g));
        const size_t big = std::max(nf[0], std::max(nf[1], nf[2]));
        DBuf<uint8_t> valid(big), status(big);
        ff::SolidVelocityData vd = ff::solidVelocityCarve(g, block.p);
        ff::solidBoxPhi(g, d, sc.s);
        int draws = ff::meshObstacleSolid(g, meshes.data(), nobj, d, nullptr, false, sc.s, &vd);
        ff::normalizeSolidVelocity(g, vd, us, vs, ws, valid, status, sc.s);
        down(phi, d.p, nn, sc.s);
        down(u, us.p, nf[0], sc.s);
        down(v, vs.p, nf[1], sc.s);
        down(w, ws.p, nf[2], sc.s);
        sync(sc.s);
        return draws;
    });
}

void FFAMD_cell_weights(int ni, int nj, int nk, const float *phi, float *wc, int *err) {
    guarded(err, [&] {
        checkGrid(ni, nj, nk, 1.0);
        Scratch sc;
        const size_t nn = (size_t)(ni + 1) * (nj + 1) * (nk + 1), N = (size_t)ni * nj * nk;
        DBuf<float> d(nn), c(N);
        up(d.p, phi, nn, sc.s);
        ff::cellWeights(ff::Grid{ni, nj, nk, 1.0}, d, c, sc.s);
        down(wc, c.p, N, sc.s);
        sync(sc.s);
    });
}

namespace {
// the viscosity stage's inputs on the device
struct ViscInputs {
    DBuf<float> liquid, solid, visc, u, v, w;
    size_t nf[3];
    ViscInputs(int ni, int nj, int nk, const float *lp, const float *sp, const float *vp, const float *hu,
               const float *hv, const float *hw, hipStream_t s)
        : liquid((size_t)ni * nj * nk), solid((size_t)(ni + 1) * (nj + 1) * (nk + 1)),
          visc((size_t)(ni + 1) * (nj + 1) * (nk + 1)), u((size_t)(ni + 1) * nj * nk), v((size_t)ni * (nj + 1) * nk),
          w((size_t)ni * nj * (nk + 1)), nf{(size_t)(ni + 1) * nj * nk, (size_t)ni * (nj + 1) * nk,
                                           (size_t)ni * nj * (nk + 1)} {
        const size_t nn = (size_t)(ni + 1) * (nj + 1) * (nk + 1);
        up(liquid.p, lp, (size_t)ni * nj * nk, s);
        up(solid.p, sp, nn, s);
        up(visc.p, vp, nn, s);
        up(u.p, hu, nf[0], s);
        up(v.p, hv, nf[1], s);
        up(w.p, hw, nf[2], s);
    }
};
}  // namespace

int FFAMD_viscosity_system(int ni, int nj, int nk, double dx, double dt, const float *liquid_phi,
                           const float *solid_phi_nodes, const float *viscosity, const float *u, const float *v,
                           const float *w, int *nnz, unsigned char *states, float *volumes, int *table,
                           int *rowstart, int *cols, float *vals, float *rhs, int *err) {
    return guarded(err, [&] {
        checkGrid(ni, nj, nk, dx);
        Scratch sc;
        ViscInputs in(ni, nj, nk, liquid_phi, solid_phi_nodes, viscosity, u, v, w, sc.s);
        ff::ViscositySolver vs;
        const int m = vs.build(ff::Grid{ni, nj, nk, dx}, dt, in.liquid, in.solid, in.visc, in.u, in.v, in.w, sc.s);
        ff::ViscosityHostSystem H;
        vs.copySystem(H, sc.s);
        *nnz = (int)H.cols.size();
        if (states) std::copy(H.states.begin(), H.states.end(), states);
        if (volumes) std::copy(H.volumes.begin(), H.volumes.end(), volumes);
        if (table) std::copy(H.table.begin(), H.table.end(), table);
        if (rowstart) std::copy(H.rowstart.begin(), H.rowstart.end(), rowstart);
        if (cols) std::copy(H.cols.begin(), H.cols.end(), cols);
        if (vals) std::copy(H.vals.begin(), H.vals.end(), vals);
        if (rhs) std::copy(H.rhs.begin(), H.rhs.end(), rhs);
        return m;
    });
}

int FFAMD_viscosity_solve(int ni, int nj, int nk, double dx, double dt, const float *liquid_phi,
                          const float *solid_phi_nodes, const float *viscosity, double tol, int max_iter, float *u,
                          float *v, float *w, int *iterations, float *error, int *success, float *solution,
                          int *err) {
    return guarded(err, [&] {
        checkGrid(ni, nj, nk, dx);
        if (max_iter < 0) throw std::domain_error("Error: iteration limit must be >= 0.\n");
        Scratch sc;
        ViscInputs in(ni, nj, nk, liquid_phi, solid_phi_nodes, viscosity, u, v, w, sc.s);
        ff::ViscositySolver vs;
        const ff::ViscosityResult r = vs.solve(ff::Grid{ni, nj, nk, dx}, dt, in.liquid, in.solid, in.visc, tol,
                                               max_iter, in.u, in.v, in.w, sc.s);
        down(u, in.u.p, in.nf[0], sc.s);
        down(v, in.v.p, in.nf[1], sc.s);
        down(w, in.w.p, in.nf[2], sc.s);
        sync(sc.s);
        if (solution) {
            std::vector<float> x;
            vs.copySolution(x, sc.s);
            std::copy(x.begin(), x.end(), solution);
        }
        *iterations = r.iterations;
        *error = r.error;
        *success = r.success ? 1 : 0;
        return r.unknowns;
    });
}

void FFAMD_condition_solid_velocity(int ni, int nj, int nk, const float *phi, const float *wu, const float *wv,
                                    const float *ww, float *us, float *vs, float *ws, int *err) {
    guarded(err, [&] {
        checkGrid(ni, nj, nk, 1.0);
        Scratch sc;
        const size_t N = (size_t)ni * nj * nk;
        const size_t nf[3] = {(size_t)(ni + 1) * nj * nk, (size_t)ni * (nj + 1) * nk, (size_t)ni * nj * (nk + 1)};
        DBuf<float> p(N), a(nf[0]), b(nf[1]), c(nf[2]), x(nf[0]), y(nf[1]), z(nf[2]);
        DBuf<int> scratch(2 * N);
        up(p.p, phi, N, sc.s);
        up(a.p, wu, nf[0], sc.s);
        up(b.p, wv, nf[1], sc.s);
        up(c.p, ww, nf[2], sc.s);
        up(x.p, us, nf[0], sc.s);
        up(y.p, vs, nf[1], sc.s);
        up(z.p, ws, nf[2], sc.s);
        ff::conditionSolidVelocity(ff::Grid{ni, nj, nk, 1.0}, p, a, b, c, x, y, z, scratch, sc.s);
        down(us, x.p, nf[0], sc.s);
        down(vs, y.p, nf[1], sc.s);
        down(ws, z.p, nf[2], sc.s);
        sync(sc.s);
    });
}

void FFAMD_face_weights(int ni, int nj, int nk, const float *phi, float *wu, float *wv, float *ww, int *err) {
    guarded(err, [&] {
        checkGrid(ni, nj, nk, 1.0);
        Scratch sc;
        size_t nn = (size_t)(ni + 1) * (nj + 1) * (nk + 1);
        size_t nu = (size_t)(ni + 1) * nj * nk, nv = (size_t)ni * (nj + 1) * nk, nw = (size_t)ni * nj * (nk + 1);
        DBuf<float> d(nn), a(nu), b(nv), c(nw);
        up(d.p, phi, nn, sc.s);
        ff::faceWeights(ff::Grid{ni, nj, nk, 1.0}, d, a, b, c, sc.s);
        down(wu, a.p, nu, sc.s);
        down(wv, b.p, nv, sc.s);
        down(ww, c.p, nw, sc.s);
        sync(sc.s);
    });
}

void FFAMD_sample_velocity(int ni, int nj, int nk, double dx, const float *u, const float *v, const float *w,
                           int np, const float *pos, float *out, int *err) {
    guarded(err, [&] {
        checkGrid(ni, nj, nk, dx);
        // sampling is G2P with PIC ratio 1, zero saved field and no motion:
        // the new velocity is the sampled field
        ff::Grid g{ni, nj, nk, dx};
        Scratch sc;
        size_t nu = (size_t)(ni + 1) * nj * nk, nv = (size_t)ni * (nj + 1) * nk, nw = (size_t)ni * nj * (nk + 1);
        size_t P = (size_t)np;
        DBuf<float> du(nu), dv(nv), dw(nw), zu(nu), zv(nv), zw(nw);
        up(du.p, u, nu, sc.s);
        up(dv.p, v, nv, sc.s);
        up(dw.p, w, nw, sc.s);
        ff::hipCheck(hipMemsetAsync(zu, 0, sizeof(float) * nu, sc.s), "memset");
        ff::hipCheck(hipMemsetAsync(zv, 0, sizeof(float) * nv, sc.s), "memset");
        ff::hipCheck(hipMemsetAsync(zw, 0, sizeof(float) * nw, sc.s), "memset");
        std::vector<std::unique_ptr<DBuf<float>>> pa;
        std::vector<float> c(P);
        for (int a = 0; a < 6; a++) {
            pa.push_back(std::make_unique<DBuf<float>>(P));
            for (size_t q = 0; q < P; q++) c[q] = a < 3 ? pos[3 * q + a] : 0.0f;
            up(pa.back()->p, c.data(), P, sc.s);
            sync(sc.s);
        }
        size_t NN = (size_t)(ni + 1) * (nj + 1) * (nk + 1);
        DBuf<float> solid(NN);
        // positive solid SDF everywhere and an empty near-solid grid: no collision
        std::vector<float> one(NN, 1.0f);
        up(solid.p, one.data(), NN, sc.s);
        DBuf<uint8_t> near(1);
        ff::hipCheck(hipMemsetAsync(near, 0, 1, sc.s), "memset");
        ff::Particles pp{*pa[0], *pa[1], *pa[2], *pa[3], *pa[4], *pa[5], np};
        ff::g2pAdvect(g, pp, ff::MacView{du, dv, dw}, ff::MacView{zu, zv, zw}, solid, near, 1, 1, 1, 1.0f, 0.0f,
                      1e-30, 5.0, sc.s);
        std::vector<float> x(P), y(P), z(P);
        down(x.data(), pa[3]->p, P, sc.s);
        down(y.data(), pa[4]->p, P, sc.s);
        down(z.data(), pa[5]->p, P, sc.s);
        sync(sc.s);
        for (size_t q = 0; q < P; q++) {
            out[3 * q] = x[q];
            out[3 * q + 1] = y[q];
            out[3 * q + 2] = z[q];
        }
    });
}

int FFAMD_pressure_solve(int ni, int nj, int nk, double dx, double dt, float *u, float *v, float *w,
                         const float *wu, const float *wv, const float *ww, const float *phi, double tol,
                         int maxIter, int mode, float *pressure, int *iterations, double *error,
                         unsigned char *vu, unsigned char *vv, unsigned char *vw, int *err) {
    return guarded(err, [&] {
        checkGrid(ni, nj, nk, dx);
        ff::Grid g{ni, nj, nk, dx};
        Scratch sc;
        size_t N = (size_t)ni * nj * nk;
        size_t nu = (size_t)(ni + 1) * nj * nk, nv = (size_t)ni * (nj + 1) * nk, nw = (size_t)ni * nj * (nk + 1);
        DBuf<float> du(nu), dv(nv), dw(nw), a(nu), b(nv), c(nw), dphi(N), dp(N);
        DBuf<uint8_t> bu(nu), bv(nv), bw(nw);
        DBuf<int> scan(ff::scanScratchInts(N + 1));
        up(du.p, (const float *)u, nu, sc.s); up(dv.p, (const float *)v, nv, sc.s); up(dw.p, (const float *)w, nw, sc.s);
        up(a.p, wu, nu, sc.s); up(b.p, wv, nv, sc.s); up(c.p, ww, nw, sc.s);
        up(dphi.p, phi, N, sc.s);
        ff::PressureSolver ps;
        ps.allocate(g);
        if (mode < 0 || mode > 3) throw std::domain_error("mode must be 0, 1, 2 or 3");
        ff::PressureResult r = ps.solve(dt, du, dv, dw, a, b, c, dphi, tol, maxIter, mode, mode == 2 ? 0.1 : 1e-4,
                                        dp, scan, sc.s, nullptr);
        ff::hipCheck(hipMemsetAsync(bu, 0, nu, sc.s), "memset");
        ff::hipCheck(hipMemsetAsync(bv, 0, nv, sc.s), "memset");
        ff::hipCheck(hipMemsetAsync(bw, 0, nw, sc.s), "memset");
        if (r.success) ff::applyPressure(g, dt, dp, dphi, a, b, c, du, dv, dw, bu, bv, bw, sc.s);
        down(pressure, dp.p, N, sc.s);
        down(u, du.p, nu, sc.s); down(v, dv.p, nv, sc.s); down(w, dw.p, nw, sc.s);
        down(vu, bu.p, nu, sc.s); down(vv, bv.p, nv, sc.s); down(vw, bw.p, nw, sc.s);
        sync(sc.s);
        *iterations = r.iterations;
        *error = r.error;
        return r.success ? 1 : 0;
    });
}

namespace {
struct TensionScratch {
    DBuf<char> block;
    ff::TensionBuffers B;
    TensionScratch(ff::Grid g) : block(ff::tensionBytes(g)) { ff::tensionCarve(g, block.p, B); }
};
}  // namespace

void FFAMD_curvature_grid(int ni, int nj, int nk, double dx, const float *phi, float *surface_phi, float *curvature,
                          int *err) {
    guarded(err, [&] {
        checkGrid(ni, nj, nk, dx);
        ff::Grid g{ni, nj, nk, dx};
        Scratch sc;
        const size_t N = (size_t)ni * nj * nk;
        DBuf<float> dphi(N);
        up(dphi.p, phi, N, sc.s);
        TensionScratch T(g);
        ff::curvatureGrid(g, dphi, T.B, sc.s);
        down(surface_phi, T.B.surface, N, sc.s);
        down(curvature, T.B.curvature, N, sc.s);
        sync(sc.s);
    });
}

void FFAMD_tension_cluster_status(int ni, int nj, int nk, const float *phi, signed char *status, int *err) {
    guarded(err, [&] {
        checkGrid(ni, nj, nk, 1.0);
        ff::Grid g{ni, nj, nk, 1.0};
        Scratch sc;
        const size_t N = (size_t)ni * nj * nk;
        DBuf<float> dphi(N);
        up(dphi.p, phi, N, sc.s);
        TensionScratch T(g);
        ff::tensionClusterStatus(g, dphi, T.B, sc.s);
        down((int8_t *)status, T.B.status, N, sc.s);
        sync(sc.s);
    });
}

int FFAMD_pressure_solve_tension(int ni, int nj, int nk, double dx, double dt, float *u, float *v, float *w,
                                 const float *wu, const float *wv, const float *ww, const float *phi, double k,
                                 double tol, int maxIter, int reference_threads, float *pressure, int *iterations,
                                 double *error, unsigned char *vu, unsigned char *vv, unsigned char *vw, int *err) {
    return guarded(err, [&] {
        checkGrid(ni, nj, nk, dx);
        if (reference_threads < 0) throw std::domain_error("thread count must be >= 0");
        ff::Grid g{ni, nj, nk, dx};
        Scratch sc;
        size_t N = (size_t)ni * nj * nk;
        size_t nu = (size_t)(ni + 1) * nj * nk, nv = (size_t)ni * (nj + 1) * nk, nw = (size_t)ni * nj * (nk + 1);
        DBuf<float> du(nu), dv(nv), dw(nw), a(nu), b(nv), c(nw), dphi(N), dp(N);
        DBuf<uint8_t> bu(nu), bv(nv), bw(nw);
        DBuf<int> scan(ff::scanScratchInts(N + 1));
        up(du.p, (const float *)u, nu, sc.s); up(dv.p, (const float *)v, nv, sc.s); up(dw.p, (const float *)w, nw, sc.s);
        up(a.p, wu, nu, sc.s); up(b.p, wv, nv, sc.s); up(c.p, ww, nw, sc.s);
        up(dphi.p, phi, N, sc.s);
        TensionScratch T(g);
        ff::TensionView tv{};
        if (k > 1e-6) {
            ff::curvatureGrid(g, dphi, T.B, sc.s);
            ff::tensionClusterStatus(g, dphi, T.B, sc.s);
            tv = ff::TensionView{T.B.status, T.B.curvature, dphi.p, k};
        }
        ff::PressureSolver ps;
        ps.allocate(g);
        ps.setTension(tv);
        ff::PressureResult r = ps.solve(dt, du, dv, dw, a, b, c, dphi, tol, maxIter, 0, 1e-4, dp, scan, sc.s, nullptr,
                                        nullptr, reference_threads);
        ff::hipCheck(hipMemsetAsync(bu, 0, nu, sc.s), "memset");
        ff::hipCheck(hipMemsetAsync(bv, 0, nv, sc.s), "memset");
        ff::hipCheck(hipMemsetAsync(bw, 0, nw, sc.s), "memset");
        if (r.success) ff::applyPressure(g, dt, dp, dphi, a, b, c, du, dv, dw, bu, bv, bw, sc.s, tv);
        down(pressure, dp.p, N, sc.s);
        down(u, du.p, nu, sc.s); down(v, dv.p, nv, sc.s); down(w, dw.p, nw, sc.s);
        down(vu, bu.p, nu, sc.s); down(vv, bv.p, nv, sc.s); down(vw, bw.p, nw, sc.s);
        sync(sc.s);
        *iterations = r.iterations;
        *error = r.error;
        return r.success ? 1 : 0;
    });
}

void FFAMD_p2g_vjp(int ni, int nj, int nk, double dx, int np, const float *pos, const float *vel, double radius,
                   const float *gu, const float *gv, const float *gw, float *vbar, float *xbar, int *err) {
    guarded(err, [&] {
        checkGrid(ni, nj, nk, dx);
        if (np < 0) throw std::domain_error("negative particle count");
        ff::Grid g{ni, nj, nk, dx};
        Scratch sc;
        size_t nu = (size_t)(ni + 1) * nj * nk, nv = (size_t)ni * (nj + 1) * nk, nw = (size_t)ni * nj * (nk + 1);
        // the forward: face values and weight sums
        if (xbar && !vel) throw std::domain_error("the position cotangent needs the particle velocities");
        SortedParticles sp(g, np, pos, vel, sc.s);
        DBuf<float> fu(nu), fv(nv), fw(nw), wsu(nu), wsv(nv), wsw(nw), gu_(nu), gv_(nv), gw_(nw);
        DBuf<uint8_t> bu(nu), bv(nv), bw(nw);
        ff::p2g(g, sp.cs, radius, fu, fv, fw, bu, bv, bw, sc.s, wsu, wsv, wsw);
        up(gu_.p, gu, nu, sc.s); up(gv_.p, gv, nv, sc.s); up(gw_.p, gw, nw, sc.s);
        DBuf<float> vx(np), vy(np), vz(np), xb(3 * (size_t)np);
        ff::p2gAdjoint(g, np, *sp.f[0], *sp.f[1], *sp.f[2], vel ? (const float *)*sp.f[3] : nullptr, *sp.f[4],
                       *sp.f[5], radius, gu_, gv_, gw_, wsu, wsv, wsw, fu, fv, fw, vx, vy, vz, xbar ? xb.p : nullptr,
                       sc.s);
        std::vector<float> hx(np), hy(np), hz(np);
        down(hx.data(), vx.p, np, sc.s); down(hy.data(), vy.p, np, sc.s); down(hz.data(), vz.p, np, sc.s);
        if (xbar) down(xbar, xb.p, 3 * (size_t)np, sc.s);
        sync(sc.s);
        for (int q = 0; q < np; q++) {
            vbar[3 * (size_t)q] = hx[q];
            vbar[3 * (size_t)q + 1] = hy[q];
            vbar[3 * (size_t)q + 2] = hz[q];
        }
    });
}

void FFAMD_sample_velocity_vjp(int ni, int nj, int nk, double dx, int np, const float *pos, const float *vbar,
                               float *u, float *v, float *w, int *err) {
    guarded(err, [&] {
        checkGrid(ni, nj, nk, dx);
        if (np < 0) throw std::domain_error("negative particle count");
        ff::Grid g{ni, nj, nk, dx};
        Scratch sc;
        const size_t P = (size_t)np;
        std::vector<float> c(P);
        DBuf<float> px(P), py(P), pz(P), vb(3 * P);
        float *dst[3] = {px.p, py.p, pz.p};
        for (int a = 0; a < 3; a++) {
            for (size_t q = 0; q < P; q++) c[q] = pos[3 * q + a];
            up(dst[a], c.data(), P, sc.s);
            sync(sc.s);
        }
        up(vb.p, vbar, 3 * P, sc.s);
        size_t nbMax = 0;
        for (int d = 0; d < 3; d++) nbMax = std::max(nbMax, ff::sampleBuckets(g, d));
        DBuf<int> key(P), slot(P), perm(P), counts(nbMax + 1), offsets(nbMax + 1),
            scan(ff::scanScratchInts(nbMax + 1));
        float *outs[3] = {u, v, w};
        for (int d = 0; d < 3; d++) {
            const size_t nf = (size_t)(ni + (d == 0)) * (nj + (d == 1)) * (nk + (d == 2));
            DBuf<float> o(nf);
            ff::sampleAdjoint(g, d, np, px, py, pz, vb, key, slot, counts, offsets, perm, scan, o, sc.s);
            down(outs[d], o.p, nf, sc.s);
            sync(sc.s);
        }
    });
}

namespace {
struct HostSoA {  // device SoA copies of a host (P x 3) array
    DBuf<float> x, y, z;
    HostSoA(const float *a, size_t P, hipStream_t s) : x(P), y(P), z(P) {
        std::vector<float> c(P);
        float *d[3] = {x.p, y.p, z.p};
        for (int e = 0; e < 3; e++) {
            for (size_t q = 0; q < P; q++) c[q] = a[3 * q + e];
            up(d[e], c.data(), P, s);
            sync(s);
        }
    }
};
}  // namespace

void FFAMD_advect_rk3(int ni, int nj, int nk, double dx, const float *u, const float *v, const float *w, int np,
                      const float *pos, double dt, float *out, int *err) {
    guarded(err, [&] {
        checkGrid(ni, nj, nk, dx);
        if (np < 0) throw std::domain_error("negative particle count");
        ff::Grid g{ni, nj, nk, dx};
        Scratch sc;
        size_t nu = (size_t)(ni + 1) * nj * nk, nv = (size_t)ni * (nj + 1) * nk, nw = (size_t)ni * nj * (nk + 1);
        DBuf<float> du(nu), dv(nv), dw(nw);
        up(du.p, u, nu, sc.s); up(dv.p, v, nv, sc.s); up(dw.p, w, nw, sc.s);
        HostSoA p(pos, np, sc.s);
        DBuf<float> ox(np), oy(np), oz(np);
        ff::advectRK3(g, np, p.x, p.y, p.z, ff::MacView{du, dv, dw}, dt, ox, oy, oz, sc.s);
        std::vector<float> hx(np), hy(np), hz(np);
        down(hx.data(), ox.p, np, sc.s); down(hy.data(), oy.p, np, sc.s); down(hz.data(), oz.p, np, sc.s);
        sync(sc.s);
        for (int q = 0; q < np; q++) {
            out[3 * (size_t)q] = hx[q];
            out[3 * (size_t)q + 1] = hy[q];
            out[3 * (size_t)q + 2] = hz[q];
        }
    });
}

void FFAMD_advect_rk3_vjp(int ni, int nj, int nk, double dx, const float *u, const float *v, const float *w, int np,
                          const float *pos, double dt, const float *xbar, float *pbar, float *gu, float *gv,
                          float *gw, int *err) {
    guarded(err, [&] {
        checkGrid(ni, nj, nk, dx);
        if (np < 0) throw std::domain_error("negative particle count");
        ff::Grid g{ni, nj, nk, dx};
        Scratch sc;
        size_t nu = (size_t)(ni + 1) * nj * nk, nv = (size_t)ni * (nj + 1) * nk, nw = (size_t)ni * nj * (nk + 1);
        DBuf<float> du(nu), dv(nv), dw(nw);
        up(du.p, u, nu, sc.s); up(dv.p, v, nv, sc.s); up(dw.p, w, nw, sc.s);
        HostSoA p(pos, np, sc.s);
        const size_t P = (size_t)np, P3 = 3 * P;
        DBuf<float> xb(P3), pb(P3), pts(3 * P3), cot(3 * P3);
        up(xb.p, xbar, P3, sc.s);
        ff::advectRK3Adjoint(g, np, p.x, p.y, p.z, ff::MacView{du, dv, dw}, dt, xb, pb, pts, cot, sc.s);
        down(pbar, pb.p, P3, sc.s);
        // the face grids' cotangent: sampling adjoints of the three stages
        size_t nbMax = 0;
        for (int d = 0; d < 3; d++) nbMax = std::max(nbMax, ff::sampleBuckets(g, d));
        DBuf<int> key(P3), slot(P3), perm(P3), counts(nbMax + 1), offsets(nbMax + 1),
            scan(ff::scanScratchInts(nbMax + 1));
        float *outs[3] = {gu, gv, gw};
        for (int d = 0; d < 3; d++) {
            const size_t nf = (size_t)(ni + (d == 0)) * (nj + (d == 1)) * (nk + (d == 2));
            DBuf<float> o(nf);
            ff::sampleAdjoint(g, d, (int)P3, pts.p, pts.p + P3, pts.p + 2 * P3, cot, key, slot, counts, offsets, perm,
                              scan, o, sc.s);
            down(outs[d], o.p, nf, sc.s);
            sync(sc.s);
        }
    });
}

int FFAMD_projection_vjp(int ni, int nj, int nk, double dx, double dt, const float *wu, const float *wv,
                         const float *ww, const float *phi, const float *gu, const float *gv, const float *gw,
                         double tol, int maxIter, float *ou, float *ov, float *ow, int *iterations, int *err) {
    return guarded(err, [&] {
        checkGrid(ni, nj, nk, dx);
        ff::Grid g{ni, nj, nk, dx};
        Scratch sc;
        size_t N = (size_t)ni * nj * nk;
        size_t nu = (size_t)(ni + 1) * nj * nk, nv = (size_t)ni * (nj + 1) * nk, nw = (size_t)ni * nj * (nk + 1);
        DBuf<float> zu(nu), zv(nv), zw(nw), a(nu), b(nv), c(nw), dphi(N), dp(N), gu_(nu), gv_(nv), gw_(nw);
        DBuf<float> du(nu), dv(nv), dw(nw);
        DBuf<double> pbar(N), lambda(N);
        DBuf<int> scan(ff::scanScratchInts(N + 1));
        up(a.p, wu, nu, sc.s); up(b.p, wv, nv, sc.s); up(c.p, ww, nw, sc.s);
        up(dphi.p, phi, N, sc.s);
        up(gu_.p, gu, nu, sc.s); up(gv_.p, gv, nv, sc.s); up(gw_.p, gw, nw, sc.s);
        ff::hipCheck(hipMemsetAsync(zu, 0, sizeof(float) * nu, sc.s), "memset");
        ff::hipCheck(hipMemsetAsync(zv, 0, sizeof(float) * nv, sc.s), "memset");
        ff::hipCheck(hipMemsetAsync(zw, 0, sizeof(float) * nw, sc.s), "memset");
        // C^T u'bar on the cells, then A lambda = that on the rows (FP64 PCG)
        ff::projectionAdjointCells(g, dt, dphi, a, b, c, gu_, gv_, gw_, pbar, sc.s);
        ff::PressureSolver ps;
        ps.allocate(g);
        ff::PressureResult r = ps.solve(dt, zu, zv, zw, a, b, c, dphi, tol, maxIter, 0, 1e-4, dp, scan, sc.s, nullptr,
                                        nullptr, 0, pbar, lambda);
        ff::projectionAdjointFaces(g, dt, dphi, a, b, c, gu_, gv_, gw_, lambda, du, dv, dw, sc.s);
        down(ou, du.p, nu, sc.s); down(ov, dv.p, nv, sc.s); down(ow, dw.p, nw, sc.s);
        sync(sc.s);
        *iterations = r.iterations;
        return r.success ? 1 : 0;
    });
}

void FFAMD_mic0_band_shape(int fp64, int *levels, int *tile, int *err) {
    guarded(err, [&] {
        int k = 0, bt = 0;
        ff::mic0BandShape(fp64 != 0, k, bt);
        *levels = k;
        *tile = bt;
    });
}

int FFAMD_mic0_apply(int ni, int nj, int nk, double dx, double dt, const float *u, const float *v,
                     const float *w, const float *wu, const float *wv, const float *ww, const float *phi,
                     int precision, int schedule, int n, const double *r, double *z, int *err) {
    return guarded(err, [&] {
        checkGrid(ni, nj, nk, dx);
        if (precision != 0 && precision != 1) throw std::domain_error("precision must be 0 (FP64) or 1 (FP32)");
        if (schedule < 0 || schedule > 4)
            throw std::domain_error("schedule must be 0 (hyperplanes), 1 (tiles), 2 (bands), 3 (bands, device "
                                    "tile check) or 4 (line pipeline)");
        ff::Grid g{ni, nj, nk, dx};
        Scratch sc;
        size_t N = (size_t)ni * nj * nk;
        size_t nu = (size_t)(ni + 1) * nj * nk, nv = (size_t)ni * (nj + 1) * nk, nw = (size_t)ni * nj * (nk + 1);
        DBuf<float> du(nu), dv(nv), dw(nw), a(nu), b(nv), c(nw), dphi(N), dp(N);
        DBuf<int> scan(ff::scanScratchInts(N + 1));
        up(du.p, u, nu, sc.s); up(dv.p, v, nv, sc.s); up(dw.p, w, nw, sc.s);
        up(a.p, wu, nu, sc.s); up(b.p, wv, nv, sc.s); up(c.p, ww, nw, sc.s);
        up(dphi.p, phi, N, sc.s);
        ff::PressureSolver ps;
        ps.allocate(g);
        // one PCG iteration builds the system and the factor
        ps.solve(dt, du, dv, dw, a, b, c, dphi, 1e-300, 1, precision, 1e-4, dp, scan, sc.s, nullptr);
        size_t m = (size_t)ps.system().m;
        if ((size_t)n != m)
            throw std::domain_error("r and z hold " + std::to_string(n) + " values; the system has " +
                                    std::to_string(m) + " rows");
        DBuf<double> dr(m), dz(m);
        up(dr.p, r, m, sc.s);
        if (precision == 0) {
            ps.applyPreconditioner(dr, dz, schedule, sc.s);
        } else {
            DBuf<float> fr(m), fz(m);
            ff::castToFloat(dr, fr, m, sc.s);
            ps.applyPreconditionerF(fr, fz, schedule, sc.s);
            ff::hipCheck(hipMemsetAsync(dz, 0, sizeof(double) * m, sc.s), "memset");
            ff::addFloatToDouble(dz, fz, m, sc.s);
            sync(sc.s);
        }
        down(z, dz.p, m, sc.s);
        sync(sc.s);
        ps.checkError(sc.s);
        return (int)m;
    });
}

int FFAMD_pressure_system(int ni, int nj, int nk, double dx, double dt, const float *u, const float *v,
                          const float *w, const float *wu, const float *wv, const float *ww, const float *phi,
                          int *rows, double *rhs, double *diag, double *ox, double *oy, double *oz, int *nbr,
                          double *invdiag, int *err) {
    return guarded(err, [&] {
        checkGrid(ni, nj, nk, dx);
        ff::Grid g{ni, nj, nk, dx};
        Scratch sc;
        size_t N = (size_t)ni * nj * nk;
        size_t nu = (size_t)(ni + 1) * nj * nk, nv = (size_t)ni * (nj + 1) * nk, nw = (size_t)ni * nj * (nk + 1);
        DBuf<float> du(nu), dv(nv), dw(nw), a(nu), b(nv), c(nw), dphi(N), dp(N);
        DBuf<int> scan(ff::scanScratchInts(N + 1));
        up(du.p, u, nu, sc.s); up(dv.p, v, nv, sc.s); up(dw.p, w, nw, sc.s);
        up(a.p, wu, nu, sc.s); up(b.p, wv, nv, sc.s); up(c.p, ww, nw, sc.s);
        up(dphi.p, phi, N, sc.s);
        ff::PressureSolver ps;
        ps.allocate(g);
        // one PCG iteration builds the system and the factor; the tolerance is
        // irrelevant for the system itself
        ps.solve(dt, du, dv, dw, a, b, c, dphi, 1e-300, 1, 0, 1e-4, dp, scan, sc.s, nullptr);
        const ff::PressureSys &S = ps.system();
        size_t m = (size_t)S.m;
        down(rows, S.rows, m, sc.s);
        down(rhs, S.rhs, m, sc.s);
        down(diag, S.diag, m, sc.s);
        down(ox, S.ox, m, sc.s);
        down(oy, S.oy, m, sc.s);
        down(oz, S.oz, m, sc.s);
        down(nbr, S.nbr, 6 * m, sc.s);
        down(invdiag, S.invdiag, m, sc.s);
        sync(sc.s);
        return S.m;
    });
}

void FFAMD_sample_velocity_position_vjp(int ni, int nj, int nk, double dx, const float *u, const float *v,
                                        const float *w, int np, const float *pos, const float *vbar, float *xbar,
                                        int *err) {
    guarded(err, [&] {
        checkGrid(ni, nj, nk, dx);
        if (np < 0) throw std::domain_error("negative particle count");
        ff::Grid g{ni, nj, nk, dx};
        Scratch sc;
        size_t nu = (size_t)(ni + 1) * nj * nk, nv = (size_t)ni * (nj + 1) * nk, nw = (size_t)ni * nj * (nk + 1);
        DBuf<float> du(nu), dv(nv), dw(nw), vb(3 * (size_t)np), xb(3 * (size_t)np), pts(9 * (size_t)np),
            cot(9 * (size_t)np), zu(nu), zv(nv), zw(nw);
        up(du.p, u, nu, sc.s); up(dv.p, v, nv, sc.s); up(dw.p, w, nw, sc.s);
        HostSoA p(pos, np, sc.s);
        up(vb.p, vbar, 3 * (size_t)np, sc.s);
        ff::hipCheck(hipMemsetAsync(xb, 0, sizeof(float) * 3 * np, sc.s), "memset");
        // the G2P blend's adjoint with PIC share 1, a zero saved field, no
        // motion (dt = 0) and a zero position cotangent leaves exactly
        // grad V(pos)^T vbar
        ff::hipCheck(hipMemsetAsync(zu, 0, sizeof(float) * nu, sc.s), "memset");
        ff::hipCheck(hipMemsetAsync(zv, 0, sizeof(float) * nv, sc.s), "memset");
        ff::hipCheck(hipMemsetAsync(zw, 0, sizeof(float) * nw, sc.s), "memset");
        DBuf<float> vdummy(3 * (size_t)np), sdummy(3 * (size_t)np), out(3 * (size_t)np);
        ff::BlendAdj B;
        B.sav = ff::MacView{zu, zv, zw};
        B.r1 = 1.0f;
        B.r2 = 0.0f;
        B.vbarOut = vb;
        B.vbar = vdummy;
        B.cotSav = sdummy;
        ff::advectRK3Adjoint(g, np, p.x, p.y, p.z, ff::MacView{du, dv, dw}, 0.0, xb, out, pts, cot, sc.s, &B);
        down(xbar, out.p, 3 * (size_t)np, sc.s);
        sync(sc.s);
    });
}

FFAMD_Rollout *FFAMD_rollout_new(int ni, int nj, int nk, double dx, int np, const float *pos, const float *solidNodes,
                                 double dt, int substeps, int checkpointEvery, double picflip, double gx, double gy,
                                 double gz, double cfl, double tol, int maxIter, int *err) {
    return guarded(err, [&] {
        checkGrid(ni, nj, nk, dx);
        ff::RolloutSettings st;
        st.dt = dt;
        st.substeps = substeps;
        st.checkpointEvery = checkpointEvery;
        st.picflip = picflip;
        st.gx = gx;
        st.gy = gy;
        st.gz = gz;
        st.cfl = cfl;
        st.tol = tol;
        st.maxIter = maxIter;
        return reinterpret_cast<FFAMD_Rollout *>(new ff::Rollout(ff::Grid{ni, nj, nk, dx}, np, pos, solidNodes, st));
    });
}

void FFAMD_rollout_destroy(FFAMD_Rollout *r) {
    try {
        delete reinterpret_cast<ff::Rollout *>(r);
    } catch (...) {
    }
}

namespace {
ff::Rollout *R(FFAMD_Rollout *r) {
    if (!r) throw std::runtime_error("null rollout handle");
    return reinterpret_cast<ff::Rollout *>(r);
}
}  // namespace

void FFAMD_rollout_forward(FFAMD_Rollout *r, const float *u0, const float *v0, const float *w0, int freeze,
                           float *density, int *err) {
    guarded(err, [&] { R(r)->forward(u0, v0, w0, density, freeze != 0); });
}

void FFAMD_rollout_backward(FFAMD_Rollout *r, const float *densityBar, float *gu, float *gv, float *gw, int *err) {
    guarded(err, [&] { R(r)->backward(densityBar, gu, gv, gw); });
}

void FFAMD_rollout_positions(FFAMD_Rollout *r, float *pos, int *err) {
    guarded(err, [&] { R(r)->positions(pos); });
}

void FFAMD_rollout_set_reference_summation(FFAMD_Rollout *r, int threads, int *err) {
    guarded(err, [&] {
        if (threads < 0 || threads > 4096) throw std::domain_error("Error: thread count must be in [0, 4096].\n");
        R(r)->setReferenceSummation(threads);
    });
}

void FFAMD_rollout_stats(FFAMD_Rollout *r, long long *collided, long long *fwd, long long *bwd, int *err) {
    guarded(err, [&] {
        *collided = R(r)->collided();
        *fwd = R(r)->forwardIterations();
        *bwd = R(r)->backwardIterations();
    });
}

void FFAMD_p2g_apic(int ni, int nj, int nk, double dx, int np, const float *pos, const float *vel, const float *aff,
                    double radius, float *u, float *v, float *w, unsigned char *vu, unsigned char *vv,
                    unsigned char *vw, int *err) {
    guarded(err, [&] {
        checkGrid(ni, nj, nk, dx);
        if (np < 0) throw std::domain_error("negative particle count");
        ff::Grid g{ni, nj, nk, dx};
        Scratch sc;
        const size_t P = (size_t)np;
        HostSoA p(pos, P, sc.s), vl(vel, P, sc.s);
        std::vector<std::unique_ptr<DBuf<float>>> a;
        std::vector<float> c(P);
        for (int e = 0; e < 9; e++) {
            a.push_back(std::make_unique<DBuf<float>>(P));
            for (size_t q = 0; q < P; q++) c[q] = aff[9 * q + e];
            up(a.back()->p, c.data(), P, sc.s);
            sync(sc.s);
        }
        const size_t nf[3] = {(size_t)(ni + 1) * nj * nk, (size_t)ni * (nj + 1) * nk, (size_t)ni * nj * (nk + 1)};
        DBuf<float> f0(nf[0]), f1(nf[1]), f2(nf[2]);
        DBuf<uint8_t> b0(nf[0]), b1(nf[1]), b2(nf[2]);
        const size_t nb = ff::apicBuckets(g), R = ff::apicRecords(np);
        DBuf<uint8_t> mark(ff::apicBlocks(g)), active(ff::apicBlocks(g));
        DBuf<int> keys(R), slot(R), perm(R), counts(nb + 1), offsets(nb + 1), scan(ff::scanScratchInts(nb + 1));
        ff::ApicScratch A{mark, active, keys, slot, perm, counts, offsets, scan};
        const float *velp[3] = {vl.x, vl.y, vl.z};
        const float *affp[9];
        for (int e = 0; e < 9; e++) affp[e] = *a[e];
        float *fields[3] = {f0, f1, f2};
        uint8_t *valid[3] = {b0, b1, b2};
        ff::p2gApic(g, np, p.x, p.y, p.z, velp, affp, radius, fields, valid, A, sc.s);
        down(u, f0.p, nf[0], sc.s); down(v, f1.p, nf[1], sc.s); down(w, f2.p, nf[2], sc.s);
        down(vu, b0.p, nf[0], sc.s); down(vv, b1.p, nf[1], sc.s); down(vw, b2.p, nf[2], sc.s);
        sync(sc.s);
    });
}

void FFAMD_g2p_apic(int ni, int nj, int nk, double dx, const float *u, const float *v, const float *w, int np,
                    const float *pos, float *vel, float *aff, int *err) {
    guarded(err, [&] {
        checkGrid(ni, nj, nk, dx);
        if (np < 0) throw std::domain_error("negative particle count");
        ff::Grid g{ni, nj, nk, dx};
        Scratch sc;
        const size_t P = (size_t)np;
        size_t nu = (size_t)(ni + 1) * nj * nk, nv = (size_t)ni * (nj + 1) * nk, nw = (size_t)ni * nj * (nk + 1);
        DBuf<float> du(nu), dv(nv), dw(nw);
        up(du.p, u, nu, sc.s); up(dv.p, v, nv, sc.s); up(dw.p, w, nw, sc.s);
        std::vector<std::unique_ptr<DBuf<float>>> pa, aa;
        std::vector<float> c(P);
        for (int a = 0; a < 6; a++) {
            pa.push_back(std::make_unique<DBuf<float>>(P));
            for (size_t q = 0; q < P; q++) c[q] = a < 3 ? pos[3 * q + a] : 0.0f;
            up(pa.back()->p, c.data(), P, sc.s);
            sync(sc.s);
        }
        ff::AffineArrays A;
        for (int e = 0; e < 9; e++) {
            aa.push_back(std::make_unique<DBuf<float>>(P));
            A.a[e] = *aa.back();
        }
        size_t NN = (size_t)(ni + 1) * (nj + 1) * (nk + 1);
        DBuf<float> solid(NN);
        // positive solid SDF everywhere, an empty near-solid grid and a zero
        // substep: G2P only, no motion and no collision
        std::vector<float> one(NN, 1.0f);
        up(solid.p, one.data(), NN, sc.s);
        DBuf<uint8_t> near(1);
        ff::hipCheck(hipMemsetAsync(near, 0, 1, sc.s), "memset");
        ff::Particles pp{*pa[0], *pa[1], *pa[2], *pa[3], *pa[4], *pa[5], np};
        ff::g2pAdvect(g, pp, ff::MacView{du, dv, dw}, ff::MacView{du, dv, dw}, solid, near, 1, 1, 1, 1.0f, 0.0f, 0.0,
                      5.0, sc.s, nullptr, &A);
        std::vector<float> x(P);
        for (int e = 0; e < 3; e++) {
            down(x.data(), pa[3 + e]->p, P, sc.s);
            sync(sc.s);
            for (size_t q = 0; q < P; q++) vel[3 * q + e] = x[q];
        }
        for (int e = 0; e < 9; e++) {
            down(x.data(), aa[e]->p, P, sc.s);
            sync(sc.s);
            for (size_t q = 0; q < P; q++) aff[9 * q + e] = x[q];
        }
    });
}

}  // extern "C"
