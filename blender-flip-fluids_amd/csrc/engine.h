// engine.h -- the MI355X FLIP engine behind the FluidSimulation C ABI.
//
// FluidSimulation keeps the reference engine's stepping semantics
// (rlguy/Blender-FLIP-Fluids 1.8.5, fluidsimulation.cpp:10078-10543): a frame
// is split into CFL substeps on the host, and every substep runs on one HIP
// stream of the device that was current at initialize().  Particle and grid
// state stays resident in HBM between calls; particle I/O copies at the
// C ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/ffengine_amd.h"
#include "kernels.h"
#include "pressure.h"
#include "viscosity.h"
#include "timers.h"

namespace ff {

// RAII device allocation
template <class T>
struct DevBuf {
    T *p = nullptr;
    size_t n = 0;
    DevBuf() = default;
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    ~DevBuf() { release(); }
    void alloc(size_t count);
    void release();
    operator T *() const { return p; }
};

void hipCheck(hipError_t e, const char *what);

// A static triangle-mesh obstacle (MeshObject, meshobject.h/.cpp): its mesh
// in host memory and the reference's per-object settings.  Only enable and
// disable mark the object's state as changed (meshobject.cpp:306-318), which
// makes the simulation rebuild its solid SDF.
class MeshObstacle {
public:
    MeshObstacle(int i, int j, int k, double dx) : isize(i), jsize(j), ksize(k), dx(dx) {}
    // MeshObject::updateMeshStatic (meshobject.cpp:50-59)
    void updateMeshStatic(int nv, const float *vertices, int nt, const int *triangles);
    // MeshObject::updateMeshAnimated (meshobject.cpp:61-95): the previous,
    // current and next frame's meshes; the triangles are the current mesh's.
    // The three must have the same vertex and triangle counts (a mesh whose
    // topology changes between frames is rejected).
    void updateMeshAnimated(int nv, const float *previous, const float *current, const float *next, int nt,
                            const int *triangles);
    // MeshObject::_isMeshChanged (meshobject.cpp:893-912): animated and some
    // vertex moved more than 1e-5 from the previous frame's
    bool isMeshChanged() const;
    // the mesh at frame progress f and its vertex velocities for a frame of
    // length frameDt (MeshObject::getMesh(f), getVertexVelocities(dt, f),
    // meshobject.cpp:158-220): positions and velocities, 3 floats per vertex
    void meshAtProgress(float f, double frameDt, std::vector<float> &positions, std::vector<float> &velocities) const;
    void enable() {
        if (!enabled) stateChanged = true;
        enabled = true;
    }
    void disable() {
        if (enabled) stateChanged = true;
        enabled = false;
    }
    void inverse() { inversed = !inversed; }
    // setters with the reference's clamping (meshobject.cpp:340-387)
    void setFriction(float f) { friction = std::fmax(std::fmin(f, 1.0f), 0.0f); }
    void setNonNegative(float &field, float v) { field = std::fmax(v, 0.0f); }

    int isize, jsize, ksize;
    double dx;
    std::vector<float> vertices;  // the current frame's
    std::vector<int> triangles;
    std::vector<float> previousVertices, nextVertices;  // animated objects only
    bool animated = false;
    bool enabled = true, inversed = false, stateChanged = false, appendObjectVelocity = false;
    float friction = 0.0f, velocityScale = 1.0f, whitewaterInfluence = 1.0f, dustEmissionStrength = 1.0f,
          sheetingStrength = 1.0f, meshExpansion = 0.0f, objectVelocityInfluence = 1.0f;
    int priority = 0, sourceID = 0;
    float viscosity = 0.0f, density = 1.0f, lifetime = 0.0f, lifetimeVariance = 0.0f;
    FFVector3 sourceColor{0.0f, 0.0f, 0.0f};
};

enum class PressurePrecision { FP64 = 0, MixedRefined = 1, MixedReliable = 2, FP32Preconditioner = 3 };

class FluidSimulation {
public:
    FluidSimulation(int isize, int jsize, int ksize, double dx);
    ~FluidSimulation();

    // lifecycle (fluidsimulation.cpp:84-94, 10410-10543)
    void initialize();
    bool isInitialized() const { return _initialized; }
    void update(double dt);
    int getCurrentFrame() const { return _currentFrame; }
    void setCurrentFrame(int f) { _currentFrame = f; }
    bool isCurrentFrameFinished() const { return _frameFinished; }

    // grid
    double getCellSize() const { return _g.dx; }
    void getGridDimensions(int *i, int *j, int *k) const;

    // settings used on the substep path, with the reference's validation
    void addBodyForce(double fx, double fy, double fz);
    void resetBodyForce() { _bodyForces.clear(); }
    FFVector3 getConstantBodyForce() const;
    void setPICFLIPRatio(double r);
    double getPICFLIPRatio() const { return _picflip; }
    void setCFLConditionNumber(int n);
    int getCFLConditionNumber() const { return (int)_cfl; }
    void setMinTimeStepsPerFrame(int n);
    int getMinTimeStepsPerFrame() const { return _minSteps; }
    void setMaxTimeStepsPerFrame(int n);
    int getMaxTimeStepsPerFrame() const { return _maxSteps; }
    void setPressureSolverMaxIterations(int n) { _maxPressureIterations = n; }
    int getPressureSolverMaxIterations() const { return _maxPressureIterations; }
    void setMaxThreadCount(int n);
    int getMaxThreadCount() const { return _maxThreads; }
    void setDensity(double d);
    double getDensity() const { return _density; }
    void enableExtremeVelocityRemoval() { _extremeRemoval = true; }
    void disableExtremeVelocityRemoval() { _extremeRemoval = false; }
    bool isExtremeVelocityRemovalEnabled() const { return _extremeRemoval; }
    void enableSurfaceReconstruction() { _surfaceReconstruction = true; }
    void disableSurfaceReconstruction() { _surfaceReconstruction = false; }
    bool isSurfaceReconstructionEnabled() const { return _surfaceReconstruction; }
    // velocity transfer method (fluidsimulation.cpp:2837-2857) and the PIC/APIC
    // ratio (2877-2893; stored and validated, unused by the transfer itself)
    void setVelocityTransferMethodFLIP() { _apic = false; }
    void setVelocityTransferMethodAPIC();
    bool isVelocityTransferMethodFLIP() const { return !_apic; }
    bool isVelocityTransferMethodAPIC() const { return _apic; }
    void setPICAPICRatio(double r);
    double getPICAPICRatio() const { return _picapic; }

    // additive settings
    void setPressureSolverTolerance(double tol);
    double getPressureSolverTolerance() const { return _pressureTolerance; }
    void setPressureSolverPrecision(int mode, double innerTol);
    int getPressureSolverPrecision() const { return (int)_precision; }
    // FP64 solve: dot products in the reference's order for the thread count
    void setPressureSolverReferenceSummation(bool on) { _referenceSummation = on; }
    bool getPressureSolverReferenceSummation() const { return _referenceSummation; }
    // FP64 PCG modes: start each solve from the previous substep's pressure
    void setPressureSolverWarmStart(bool on) {
        _warmStart = on;
        _ps.setWarmStart(on);
    }
    bool getPressureSolverWarmStart() const { return _warmStart; }
    // the MIC(0) application's schedule: 2 = bands of hyperplanes, 4 = the line pipeline
    void setPressureSolverPreconditionerSchedule(int schedule) { _ps.setSchedule(schedule); }
    int getPressureSolverPreconditionerSchedule() const { return _ps.schedule(); }
    // constant viscosity (fluidsimulation.cpp:2565-2601, 303-312): the first
    // set enables the viscosity solve; a value of 0 keeps it enabled and
    // skips it.  Negative values and tolerances raise.
    void setViscosity(double v);
    double getViscosity() const { return _viscosityValue; }
    bool isViscosityEnabled() const { return _viscEnabled; }
    void setViscositySolverErrorTolerance(double tol);
    double getViscositySolverErrorTolerance() const { return _viscTolerance; }
    void setViscositySolverMaxIterations(int n) { _viscMaxIterations = n; }
    int getViscositySolverMaxIterations() const { return _viscMaxIterations; }
    // surface tension (fluidsimulation.cpp:2604-2744): enabled above 1e-6;
    // the substep restriction's condition number; the smooth kernel, which
    // doubles the particle radius of the liquid SDF and the transfer, is
    // stored but a substep with it and surface tension on raises
    void setSurfaceTension(double k);
    double getSurfaceTension() const { return _tensionK; }
    void setSurfaceTensionConditionNumber(double n);
    double getSurfaceTensionConditionNumber() const { return _tensionCondition; }
    void setSmoothSurfaceTensionKernel(bool on) { _smoothTension = on; }
    bool isSmoothSurfaceTensionKernelEnabled() const { return _smoothTension; }
    // the last viscosity solve: iterations, |r|_inf, success, matrix size
    void getLastViscositySolveStats(int *iterations, double *error, int *success, int *unknowns) const;
    void setSolidPhi(const float *phiNodes);

    // static mesh obstacles (fluidsimulation.cpp:2990-3032); the solid SDF is
    // rebuilt at the next update (fluidsimulation.cpp:5387-5426)
    void addMeshObstacle(MeshObstacle *obstacle);
    void removeMeshObstacle(MeshObstacle *obstacle);
    void removeMeshObstacles();
    // FluidSimulation::setBoundaryFriction (fluidsimulation.cpp:2675-2691):
    // the domain object's friction
    void setBoundaryFriction(double f);
    double getBoundaryFriction() const { return _boundaryFriction; }
    // rand() draws of the last solid SDF rebuild (the sign tests' jitter)
    int getLastSolidRandDraws() const { return _solidDraws; }

    // particle I/O (fluidsimulation.cpp:3426-3839)
    void loadMarkerParticleData(int n, const char *positions, const char *velocities);
    // APIC affine vectors of one load batch (fluidsimulation.cpp:3841-3863)
    void loadMarkerParticleAffineData(int n, const char *ax, const char *ay, const char *az);
    // affine vectors of particles [start, end) (fluidsimulation.cpp:3455-3501)
    void getMarkerParticleAffineDataRange(int component, int start, int end, char *out);
    int getNumMarkerParticles() const;
    void getMarkerParticlePositionDataRange(int start, int end, char *out);
    void getMarkerParticleVelocityDataRange(int start, int end, char *out);
    void getMarkerParticlePositionData(char *out);
    void getMarkerParticleVelocityData(char *out);

    FFFrameStats getFrameStatsData() const { return _stats; }

    // grid getters for tests and tools (host copies)
    void getLiquidSDF(float *out);
    void getSolidSDF(float *out);
    // the solid's face velocities of the last rebuild (zero for a static solid)
    void getSolidVelocityField(float *u, float *v, float *w);
    void getVelocityField(float *u, float *v, float *w);
    void getPressureGrid(float *out);
    void getLastPressureSolveStats(int *iterations, double *error, int *success, int *unknowns);
    double getLastPressureRhsNorm() const { return _lastRhsInf; }
    double getLastPhaseMilliseconds(int phase);

    // one substep of dt (for tests and the benchmark; no frame bookkeeping)
    void step(double dt);

    // HIP-event timing of kernel groups (see timers.h)
    void enableKernelTiming(bool on) { _timers.enable(on); }
    void getKernelTiming(int id, double *ms, long long *launches) const;

private:
    Grid _g;
    hipStream_t _stream = nullptr;
    bool _initialized = false;
    int _currentFrame = 0;
    int _stepNumber = 0;
    bool _frameFinished = true;
    bool _updatedOnce = false;
    double _frameDt = 0, _remaining = 0;

    double _cfl = 5.0;
    double _picflip = 0.05;
    bool _apic = false;
    double _picapic = 0.0;
    int _minSteps = 1, _maxSteps = 6;
    int _maxPressureIterations = 900;
    double _pressureTolerance = 1e-9;
    PressurePrecision _precision = PressurePrecision::FP64;
    double _innerTolerance = 1e-4;
    bool _referenceSummation = false;
    bool _warmStart = false;
    int _maxThreads = 1;
    double _density = 20.0;
    bool _extremeRemoval = true;
    bool _surfaceReconstruction = true;
    std::vector<FFVector3> _bodyForces;
    bool _solidGiven = false;
    std::vector<float> _solidHost;
    std::vector<MeshObstacle *> _obstacles;
    bool _solidUpToDate = false;
    // closest mesh object of each solid node (-1 none, 0 domain, q + 1 the
    // q-th enabled obstacle of the last rebuild) and the objects' friction,
    // read live each substep as the reference reads getFriction()
    DevBuf<int> _nodeObj;
    DevBuf<float> _objFric;
    std::vector<MeshObstacle *> _solidObjects;
    std::vector<float> _objFricHost;
    bool _frictionActive = false;
    double _boundaryFriction = 0.0;
    void refreshFriction();
    int _solidDraws = 0;
    void updateSolidLevelSet(double dt);
    // moving obstacles: the solid's velocity data and face velocities, the
    // cells' solid weights, and the scratch of their rebuild and conditioning
    bool _movingSolid = false;  // the last rebuild had an animated obstacle
    DevBuf<char> _solidVelBlock;
    DevBuf<float> _us, _vs, _ws, _wc;
    DevBuf<uint8_t> _svValid, _svStatus;
    DevBuf<int> _condScratch;

    std::vector<std::vector<float>> _loadPos, _loadVel, _loadAff;  // _loadAff: 9 floats per particle

    // particles (double-buffered for compaction)
    int _capacity = 0;
    DevBuf<float> _pa[6], _pb[6];
    Particles _P{};
    Particles _Q{};
    // APIC: AFFINEX.xyz, AFFINEY.xyz, AFFINEZ.xyz (SoA), double-buffered like the particles
    DevBuf<float> _aa[9], _ab[9];
    AffineArrays _A{}, _B{};
    DevBuf<uint8_t> _apMark, _apActive;
    DevBuf<int> _apKeys, _apSlot, _apPerm, _apCounts, _apOffsets, _apScan;
    ApicScratch _apic_scratch{};

    CellSort _cs{};
    DevBuf<int> _keys, _slot, _perm, _counts, _offsets, _keep, _rank, _scan, _maxCount;
    DevBuf<float> _sorted[6];

    DevBuf<float> _phi, _solid, _wu, _wv, _ww, _pressure;
    DevBuf<uint8_t> _near, _nearTmp;
    int _ci = 0, _cj = 0, _ck = 0;
    DevBuf<float> _u, _v, _w, _su, _sv, _sw;
    DevBuf<uint8_t> _vu, _vv, _vw, _stu, _stv, _stw;

    PressureSolver _ps;
    // viscosity: the constant value on the (n+1)^3 nodes and its solver
    bool _viscEnabled = false;
    double _viscosityValue = 0.0;
    float _viscosityGridValue = -1.0f;  // the value the node grid holds
    double _viscTolerance = 1e-4;
    int _viscMaxIterations = 900;
    DevBuf<float> _viscNodes;
    ViscositySolver _visc;
    bool _vSuccess = true;
    int _vIters = 0;
    float _vErr = 0.0f;
    ViscosityResult _lastVisc{};
    void applyViscosity(double dt);
    // the reference's _constrainVelocityField on one MAC field
    void constrainField(float *u, float *v, float *w);
    // surface tension: settings and the curvature / cluster buffers
    bool _tensionEnabled = false, _smoothTension = false;
    double _tensionK = 0.0, _tensionCondition = 10.0;
    DevBuf<char> _tensionBlock;
    TensionBuffers _tb{};
    double liquidRadius() const;
    DevBuf<double> _partial, _scal;
    DevBuf<int> _bins, _outl;

    // status of the last pressure solve / frame
    bool _pSuccess = true;
    int _pIters = 0;
    float _pErr = 0.0f;
    int _lastIters = 0, _lastUnknowns = 0, _lastSuccess = 1;
    double _lastRhsInf = 0.0;
    double _lastErr = 0.0;
    FFFrameStats _stats{};
    double _phaseMs[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    double _viscMs = 0;  // the last substep's viscosity stage
    hipEvent_t _ev[9] = {};
    KernelTimers _timers;

    void allocate();
    // affine arrays and the APIC transfer's scratch for the current capacity
    // (zero: the particles take zero affine vectors)
    void allocateAffine(bool zero);
    void uploadParticles();
    double nextTimeStep(double dt);
    void stepFluid(double dt);
    void extrapolate(int layers);
    bool pressureSolve(double dt);
    void removeParticles();
    double hostScalar(const double *d);
    FFVector3 bodyForce() const;
    void requireInitialized(const char *what) const;
};

}  // namespace ff
