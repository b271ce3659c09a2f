// engine.hip -- host orchestration of the FLIP substep on one HIP stream.
#include <cstdlib>

#include "engine.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <sstream>

namespace ff {

void hipCheck(hipError_t e, const char *what) {
    if (e != hipSuccess) {
        std::string msg = std::string("HIP error in ") + what + ": " + hipGetErrorString(e);
        throw std::runtime_error(msg);
    }
}

template <class T>
void DevBuf<T>::alloc(size_t count) {
    release();
    if (count == 0) count = 1;
    hipCheck(hipMalloc((void **)&p, sizeof(T) * count), "hipMalloc");
    n = count;
}
template <class T>
void DevBuf<T>::release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
}
template struct DevBuf<float>;
template struct DevBuf<double>;
template struct DevBuf<int>;
template struct DevBuf<uint8_t>;

namespace {
std::string toString(double v) {
    std::ostringstream ss;
    ss << v;
    return ss.str();
}
}  // namespace

FluidSimulation::FluidSimulation(int isize, int jsize, int ksize, double dx) {
    if (isize <= 0 || jsize <= 0 || ksize <= 0 || dx <= 0.0)
        throw std::domain_error("Error: dimensions and cell size must be greater than 0.\n");
    _g = Grid{isize, jsize, ksize, dx};
}

FluidSimulation::~FluidSimulation() {
    for (auto &e : _ev)
        if (e) (void)hipEventDestroy(e);
    if (_stream) (void)hipStreamDestroy(_stream);
}

void FluidSimulation::getGridDimensions(int *i, int *j, int *k) const {
    *i = _g.ni;
    *j = _g.nj;
    *k = _g.nk;
}

// setters (fluidsimulation.cpp:227-238, 2427-2462, 2697-2781, 2863-2876)
void FluidSimulation::addBodyForce(double fx, double fy, double fz) {
    _bodyForces.push_back(FFVector3{(float)fx, (float)fy, (float)fz});
}
FFVector3 FluidSimulation::bodyForce() const {
    FFVector3 b{0.0f, 0.0f, 0.0f};
    for (const FFVector3 &f : _bodyForces) {
        b.x += f.x;
        b.y += f.y;
        b.z += f.z;
    }
    return b;
}
FFVector3 FluidSimulation::getConstantBodyForce() const { return bodyForce(); }
void FluidSimulation::setPICFLIPRatio(double r) {
    if (r < 0.0 || r > 1.0)
        throw std::domain_error("Error: PICFLIP ratio must be in range [0.0, 1.0].\nratio: " + toString(r) + "\n");
    _picflip = r;
}
void FluidSimulation::setViscosity(double v) {
    if (v < 0.0)
        throw std::domain_error("Error: viscosity must be greater than or equal to 0.\nviscosity: " + toString(v) +
                                "\n");
    _viscEnabled = true;
    _viscosityValue = v;
}
void FluidSimulation::setSurfaceTension(double k) {
    if (k < 0.0)
        throw std::domain_error("Error: surface tension must be greater than or equal to 0.\nsurface tension: " +
                                toString(k) + "\n");
    const double eps = 1e-6;
    _tensionEnabled = k > eps;
    _tensionK = k;
}
void FluidSimulation::setSurfaceTensionConditionNumber(double n) {
    if (n <= 0.0)
        throw std::domain_error("Error: condition number must be greater than 0.0.\nnumber: " + toString(n) + "\n");
    _tensionCondition = n;
}
// fluidsimulation.cpp:4351, doubled for the smooth surface tension kernel
// (5593-5597, 5635-5638; fluidsimulation.h:2239)
double FluidSimulation::liquidRadius() const {
    const double r = 0.5 * 1.0 * _g.dx * std::sqrt(3.0);
    // the liquid SDF's block halo and the transfer's 18 neighbour lists
    // reach particles within the default radius only
    if (_tensionEnabled && _smoothTension)
        throw std::runtime_error(
            "Error: the smooth surface tension kernel (twice the particle radius) is not supported on the GPU.\n");
    return r;
}
void FluidSimulation::setViscositySolverErrorTolerance(double tol) {
    if (tol < 0.0)
        throw std::domain_error(
            "Error: viscosity solver error tolerance must be greater than or equal to 0.\nerror tolerance: " +
            toString(tol) + "\n");
    _viscTolerance = tol;
}
void FluidSimulation::getLastViscositySolveStats(int *iterations, double *error, int *success, int *unknowns) const {
    *iterations = _lastVisc.iterations;
    *error = _lastVisc.error;
    *success = _lastVisc.success ? 1 : 0;
    *unknowns = _lastVisc.unknowns;
}
void FluidSimulation::setCFLConditionNumber(int n) {
    if (n < 1)
        throw std::domain_error("Error: CFL must be greater than or equal to 1.\nCFL: " + toString(n) + "\n");
    _cfl = n;
}
void FluidSimulation::setMinTimeStepsPerFrame(int n) {
    if (n < 1)
        throw std::domain_error("Error: min step count must be greater than or equal to 1.\nStep count: " +
                                toString(n) + "\n");
    _minSteps = n;
}
void FluidSimulation::setMaxTimeStepsPerFrame(int n) {
    if (n < 1)
        throw std::domain_error("Error: max step count must be greater than or equal to 1.\nStep count: " +
                                toString(n) + "\n");
    _maxSteps = n;
}
void FluidSimulation::setMaxThreadCount(int n) {
    if (n < 1)
        throw std::domain_error("Error: thread count must be greater than or equal to 1.\nthread count: " +
                                toString(n) + "\n");
    _maxThreads = n;
}
void FluidSimulation::setDensity(double d) {
    if (d <= 0.0)
        throw std::domain_error("Error: density must be greater than 0.\ndensity: " + toString(d) + "\n");
    _density = d;
}
void FluidSimulation::setVelocityTransferMethodAPIC() {
    _apic = true;
    if (_initialized && !_A.a[0]) allocateAffine(true);
}
void FluidSimulation::setPICAPICRatio(double r) {
    if (r < 0.0 || r > 1.0)
        throw std::domain_error("Error: PICAPIC ratio must be in range [0.0, 1.0].\nratio: " + toString(r) + "\n");
    _picapic = r;
}
void FluidSimulation::setPressureSolverTolerance(double tol) {
    if (!(tol > 0.0))
        throw std::domain_error("Error: pressure solver tolerance must be greater than 0.\ntolerance: " +
                                toString(tol) + "\n");
    _pressureTolerance = tol;
}
void FluidSimulation::setPressureSolverPrecision(int mode, double innerTol) {
    if (mode < 0 || mode > 3)
        throw std::domain_error("Error: pressure precision mode must be 0 (FP64), 1 (mixed, restarted "
                                "refinement), 2 (mixed, reliable updates) or 3 (FP64 with an FP32 "
                                "preconditioner).\n");
    if ((mode == 1 || mode == 2) && !(innerTol > 0.0 && innerTol < 1.0))
        throw std::domain_error("Error: inner tolerance must be in (0, 1).\n");
    _precision = (PressurePrecision)mode;
    if (mode == 1 || mode == 2) _innerTolerance = innerTol;
}
void FluidSimulation::setSolidPhi(const float *phi) {
    if (_initialized)
        throw std::runtime_error("Error: the solid SDF must be set before initialize.\n");
    size_t nn = (size_t)(_g.ni + 1) * (_g.nj + 1) * (_g.nk + 1);
    _solidHost.assign(phi, phi + nn);
    _solidGiven = true;
}

void MeshObstacle::updateMeshStatic(int nv, const float *v, int nt, const int *t) {
    if (nv < 0 || nt < 0 || (nv > 0 && !v) || (nt > 0 && !t))
        throw std::invalid_argument("Error: invalid triangle mesh data.\n");
    for (int q = 0; q < 3 * nt; q++)
        if (t[q] < 0 || t[q] >= nv) throw std::out_of_range("Error: triangle vertex index out of range.\n");
    vertices.assign(v, v + 3 * (size_t)nv);
    triangles.assign(t, t + 3 * (size_t)nt);
}

void MeshObstacle::updateMeshAnimated(int nv, const float *prev, const float *cur, const float *next, int nt,
                                      const int *t) {
    if (nv < 0 || nt < 0 || (nv > 0 && (!prev || !cur || !next)) || (nt > 0 && !t))
        throw std::invalid_argument("Error: invalid triangle mesh data.\n");
    for (int q = 0; q < 3 * nt; q++)
        if (t[q] < 0 || t[q] >= nv) throw std::out_of_range("Error: triangle vertex index out of range.\n");
    previousVertices.assign(prev, prev + 3 * (size_t)nv);
    vertices.assign(cur, cur + 3 * (size_t)nv);
    nextVertices.assign(next, next + 3 * (size_t)nv);
    triangles.assign(t, t + 3 * (size_t)nt);
    animated = true;
}

bool MeshObstacle::isMeshChanged() const {
    if (!animated) return false;
    if (previousVertices.size() != vertices.size()) return true;
    const float eps = 1e-5f;
    for (size_t q = 0; q < vertices.size(); q += 3) {
        // vmath::length of the float difference (vmath.h): sqrt of the float dot
        const float dx = previousVertices[q] - vertices[q], dy = previousVertices[q + 1] - vertices[q + 1],
                    dz = previousVertices[q + 2] - vertices[q + 2];
        if (std::sqrt(dx * dx + dy * dy + dz * dz) > eps) return true;
    }
    return false;
}

void MeshObstacle::meshAtProgress(float f, double frameDt, std::vector<float> &pos, std::vector<float> &vel) const {
    f = std::fmax(0.0f, f);
    f = std::fmin(1.0f, f);
    const size_t n = vertices.size();
    pos.resize(n);
    vel.assign(n, 0.0f);
    if (!animated) {  // v + f (v - v): -0 becomes +0, as in getMesh(f)
        for (size_t q = 0; q < n; q++) pos[q] = vertices[q] + f * (vertices[q] - vertices[q]);
        return;
    }
    for (size_t q = 0; q < n; q++) pos[q] = vertices[q] + f * (nextVertices[q] - vertices[q]);
    if (frameDt < 1e-10) return;
    const float invdt = (float)(1.0 / frameDt);
    for (size_t q = 0; q < n; q++) {
        const float p1 = vertices[q] - previousVertices[q], p2 = nextVertices[q] - vertices[q];
        const float t = p1 + f * (p2 - p1);
        vel[q] = velocityScale * (invdt * t);
    }
}

void FluidSimulation::addMeshObstacle(MeshObstacle *obstacle) {
    if (!obstacle) throw std::invalid_argument("Error: null mesh obstacle.\n");
    for (MeshObstacle *o : _obstacles)
        if (o == obstacle) throw std::runtime_error("Error: mesh obstacle has already been added.\n");
    _obstacles.push_back(obstacle);
    _solidUpToDate = false;
}

void FluidSimulation::removeMeshObstacle(MeshObstacle *obstacle) {
    for (size_t q = 0; q < _obstacles.size(); q++)
        if (_obstacles[q] == obstacle) {
            _obstacles.erase(_obstacles.begin() + q);
            _solidUpToDate = false;
            return;
        }
    throw std::invalid_argument("Error: could not find mesh obstacle to remove.\n");
}

void FluidSimulation::removeMeshObstacles() {
    _obstacles.clear();
    _solidUpToDate = false;
}

// _updateSolidLevelSet + _addStaticObjectsToSDF + _addAnimatedObjectsToSolidSDF
// + _updateNearSolidGrid (fluidsimulation.cpp:5219-5426, 5448-5480), run at the
// start of a substep of length dt while particles exist, as the reference's
// obstacle update is (5550-5570): the objects' state flags are read and
// cleared, and a stale solid -- or one with an enabled animated obstacle whose
// mesh changed -- is rebuilt from the domain SDF (or the one set) and the
// enabled obstacles at the substep's frame progress, with its face weights and
// near-solid grid.  With an animated obstacle the rebuild also makes the
// solid's face velocities and the cells' solid weights (6066-6148).
void FluidSimulation::updateSolidLevelSet(double dt) {
    bool changed = false;
    for (MeshObstacle *o : _obstacles) {
        if (o->stateChanged || (o->enabled && o->isMeshChanged())) changed = true;
        o->stateChanged = false;
    }
    if (changed) _solidUpToDate = false;
    if (_solidUpToDate) return;
    _solidObjects.clear();
    bool anyAnimated = false;
    for (MeshObstacle *o : _obstacles) {
        if (!o->enabled) continue;
        _solidObjects.push_back(o);
        anyAnimated |= o->animated;
    }
    // the substep's progress through the frame (fluidsimulation.cpp:5242-5243)
    const float progress = 1.0f - (float)(_remaining + dt) / (float)_frameDt;
    std::vector<std::vector<float>> pos(_solidObjects.size()), vel(_solidObjects.size());
    std::vector<ObstacleMesh> meshes;
    for (size_t q = 0; q < _solidObjects.size(); q++) {
        const MeshObstacle *o = _solidObjects[q];
        o->meshAtProgress(progress, _frameDt, pos[q], vel[q]);
        ObstacleMesh m{pos[q].data(), (int)(pos[q].size() / 3), o->triangles.data(), (int)(o->triangles.size() / 3),
                       o->inversed, o->meshExpansion};
        m.velocities = vel[q].data();
        m.animated = o->animated;
        meshes.push_back(m);
    }
    if (_nodeObj.n == 0) _nodeObj.alloc((size_t)(_g.ni + 1) * (_g.nj + 1) * (_g.nk + 1));
    if (_solidGiven)
        hipCheck(hipMemcpyAsync(_solid, _solidHost.data(), sizeof(float) * _solidHost.size(), hipMemcpyHostToDevice,
                                _stream),
                 "upload solid SDF");
    else
        solidBoxPhi(_g, _solid, _stream);
    SolidVelocityData vd{};
    if (anyAnimated) {
        const size_t N = (size_t)_g.ni * _g.nj * _g.nk;
        if (_solidVelBlock.n == 0) {
            _solidVelBlock.alloc(solidVelocityBytes(_g));
            _us.alloc(_u.n);
            _vs.alloc(_v.n);
            _ws.alloc(_w.n);
            _wc.alloc(N);
            const size_t nf = std::max(_u.n, std::max(_v.n, _w.n));
            _svValid.alloc(nf);
            _svStatus.alloc(nf);
            _condScratch.alloc(2 * N);
        }
        vd = solidVelocityCarve(_g, _solidVelBlock.p);
    }
    // the domain object holds the domain SDF's band unless a solid SDF was set
    _solidDraws = meshObstacleSolid(_g, meshes.data(), (int)meshes.size(), _solid, _nodeObj.p, !_solidGiven, _stream,
                                    anyAnimated ? &vd : nullptr);
    _movingSolid = anyAnimated;
    if (anyAnimated) {
        normalizeSolidVelocity(_g, vd, _us, _vs, _ws, _svValid, _svStatus, _stream);
        cellWeights(_g, _solid, _wc, _stream);
        _ps.setSolidTerms(SolidTerms{_us.p, _vs.p, _ws.p, _wc.p});
    } else {
        _ps.setSolidTerms(SolidTerms{});
    }
    _objFricHost.clear();  // re-upload the table for the new object numbering
    faceWeights(_g, _solid, _wu, _wv, _ww, _stream);
    int layers = (int)std::ceil((float)_cfl / (float)3);
    nearSolidGrid(_g, _solid, _near, _nearTmp, _ci, _cj, _ck, layers, _stream);
    hipCheck(hipStreamSynchronize(_stream), "solid level set");
    _solidUpToDate = true;
}

// the objects' friction table: [0] the domain object, [q + 1] the q-th
// enabled obstacle of the last rebuild; uploaded when a value changed
void FluidSimulation::refreshFriction() {
    std::vector<float> table(1 + _solidObjects.size());
    table[0] = _solidGiven ? 0.0f : (float)_boundaryFriction;
    for (size_t q = 0; q < _solidObjects.size(); q++) table[q + 1] = _solidObjects[q]->friction;
    _frictionActive = false;
    for (float f : table) _frictionActive |= f != 0.0f;
    if (table == _objFricHost) return;
    if (_objFric.n < table.size()) _objFric.alloc(std::max<size_t>(table.size(), 16));
    _objFricHost = table;
    hipCheck(hipMemcpyAsync(_objFric.p, _objFricHost.data(), sizeof(float) * table.size(), hipMemcpyHostToDevice,
                            _stream),
             "upload friction table");
}

void FluidSimulation::setBoundaryFriction(double f) {
    if (f < 0.0 || f > 1.0)
        throw std::domain_error("Error: boundary friction must be in range [0.0, 1.0].\nfriction: " + toString(f) +
                                "\n");
    _boundaryFriction = f;
}

void FluidSimulation::requireInitialized(const char *what) const {
    if (!_initialized)
        throw std::runtime_error(std::string("Error: FluidSimulation must be initialized before ") + what + ".\n");
}

// FluidSimulation::loadMarkerParticleData (fluidsimulation.cpp:3818-3839):
// a synchronous copy into a load queue consumed by initialize()
void FluidSimulation::loadMarkerParticleData(int n, const char *positions, const char *velocities) {
    if (n <= 0) return;
    const float *p = (const float *)positions, *v = (const float *)velocities;
    _loadPos.emplace_back(p, p + 3 * (size_t)n);
    _loadVel.emplace_back(v, v + 3 * (size_t)n);
}

// FluidSimulation::loadMarkerParticleAffineData (fluidsimulation.cpp:3841-3863)
void FluidSimulation::loadMarkerParticleAffineData(int n, const char *ax, const char *ay, const char *az) {
    if (n <= 0) return;
    const float *a[3] = {(const float *)ax, (const float *)ay, (const float *)az};
    std::vector<float> b(9 * (size_t)n);
    for (size_t q = 0; q < (size_t)n; q++)
        for (int d = 0; d < 3; d++)
            for (int e = 0; e < 3; e++) b[9 * q + 3 * d + e] = a[d][3 * q + e];
    _loadAff.push_back(std::move(b));
}

void FluidSimulation::allocateAffine(bool zero) {
    const size_t P = (size_t)std::max(_capacity, 1);
    for (int e = 0; e < 9; e++) {
        _aa[e].alloc(P);
        _ab[e].alloc(P);
        _A.a[e] = _aa[e];
        _B.a[e] = _ab[e];
        if (zero) hipCheck(hipMemsetAsync(_aa[e], 0, sizeof(float) * P, _stream), "memset");
    }
    const size_t nb = apicBuckets(_g), R = apicRecords((int)P);
    _apMark.alloc(apicBlocks(_g));
    _apActive.alloc(apicBlocks(_g));
    _apKeys.alloc(R);
    _apSlot.alloc(R);
    _apPerm.alloc(R);
    _apCounts.alloc(nb + 1);
    _apOffsets.alloc(nb + 1);
    _apScan.alloc(scanScratchInts(nb + 1));
    _apic_scratch = ApicScratch{_apMark, _apActive, _apKeys, _apSlot, _apPerm, _apCounts, _apOffsets, _apScan};
    if (zero) hipCheck(hipStreamSynchronize(_stream), "sync");
}

void FluidSimulation::allocate() {
    const size_t N = (size_t)_g.ni * _g.nj * _g.nk;
    const size_t NN = (size_t)(_g.ni + 1) * (_g.nj + 1) * (_g.nk + 1);
    const size_t nu = (size_t)(_g.ni + 1) * _g.nj * _g.nk, nv = (size_t)_g.ni * (_g.nj + 1) * _g.nk,
                 nw = (size_t)_g.ni * _g.nj * (_g.nk + 1);
    const size_t P = (size_t)std::max(_capacity, 1);
    for (int q = 0; q < 6; q++) {
        _pa[q].alloc(P);
        _pb[q].alloc(P);
        _sorted[q].alloc(P);
    }
    _keys.alloc(P);
    _slot.alloc(P);
    _perm.alloc(P);
    _keep.alloc(P + 1);
    _rank.alloc(P + 1);
    _counts.alloc(N + 1);
    _offsets.alloc(N + 1);
    _scan.alloc(scanScratchInts(std::max(N + 1, P + 1)));
    _maxCount.alloc(1);
    _phi.alloc(N);
    _pressure.alloc(N);
    _solid.alloc(NN);
    _wu.alloc(nu);
    _wv.alloc(nv);
    _ww.alloc(nw);
    _u.alloc(nu); _v.alloc(nv); _w.alloc(nw);
    _su.alloc(nu); _sv.alloc(nv); _sw.alloc(nw);
    _vu.alloc(nu); _vv.alloc(nv); _vw.alloc(nw);
    _stu.alloc(nu); _stv.alloc(nv); _stw.alloc(nw);
    _ci = (int)std::ceil(_g.ni * _g.dx / (3 * _g.dx));
    _cj = (int)std::ceil(_g.nj * _g.dx / (3 * _g.dx));
    _ck = (int)std::ceil(_g.nk * _g.dx / (3 * _g.dx));
    _near.alloc((size_t)_ci * _cj * _ck);
    _nearTmp.alloc((size_t)_ci * _cj * _ck);
    _partial.alloc(1024);
    _scal.alloc(16);
    _bins.alloc(std::max(_maxSteps, 1) + 16);
    _outl.alloc(2);

    _cs = CellSort{_keys, _slot, _counts, _offsets, _perm, _sorted[0], _sorted[1], _sorted[2],
                   _sorted[3], _sorted[4], _sorted[5]};
    _ps.allocate(_g);
}

// FluidSimulation::initialize / _initializeSimulation (fluidsimulation.cpp:84-90, 4362-4398)
void FluidSimulation::initialize() {
    if (_initialized) return;
    int dev = 0;
    hipCheck(hipGetDevice(&dev), "hipGetDevice");
    hipCheck(hipStreamCreateWithFlags(&_stream, hipStreamNonBlocking), "hipStreamCreate");
    for (auto &e : _ev) hipCheck(hipEventCreate(&e), "hipEventCreate");

    // _loadMarkerParticles (fluidsimulation.cpp:4903-5016): the load queue is a
    // FragmentedVector of 5e6-byte fragments (E = 208333 particles); its
    // operator[] picks the fragment as (int)(i * (1.0 / E)) in double
    // (fragmentedvector.h:140-151, 294-301), which reads element (c-1)*E at
    // some exact multiples i = c*E.  Particles outside [0, n*dx) are dropped.
    // APIC affine vectors load with a batch when every batch has them and the
    // sizes match (_loadParticles, 5078-5110, and _loadMarkerParticles,
    // 4918-4919); their queue is a FragmentedVector<MarkerParticleAffine> of
    // 36-byte elements, read with the same fragment rule.
    std::vector<float> hp, hv, ha;
    const int E = (int)(5000000u / 24u), EA = (int)(5000000u / 36u);
    const double inv = 1.0 / (double)E, invA = 1.0 / (double)EA;
    const bool affQueue = _loadAff.size() == _loadPos.size();
    Box bounds{V3{0, 0, 0}, _g.ni * _g.dx, _g.nj * _g.dx, _g.nk * _g.dx};
    for (size_t b = 0; b < _loadPos.size(); b++) {
        const std::vector<float> &P = _loadPos[b], &Vl = _loadVel[b];
        size_t n = P.size() / 3;
        const bool useAff = _apic && affQueue && _loadAff[b].size() == 9 * n;
        for (size_t i = 0; i < n; i++) {
            size_t q = (size_t)((int)((int)i * inv)) * E + (int)i % E;
            V3 x{P[3 * q], P[3 * q + 1], P[3 * q + 2]};
            if (!bounds.inside(x)) continue;
            hp.insert(hp.end(), {x.x, x.y, x.z});
            hv.insert(hv.end(), {Vl[3 * q], Vl[3 * q + 1], Vl[3 * q + 2]});
            if (_apic) {
                const size_t qa = (size_t)((int)((int)i * invA)) * EA + (int)i % EA;
                for (int e = 0; e < 9; e++) ha.push_back(useAff ? _loadAff[b][9 * qa + e] : 0.0f);
            }
        }
    }
    _loadPos.clear();
    _loadVel.clear();
    _loadAff.clear();
    int np = (int)(hp.size() / 3);
    _capacity = np;
    allocate();
    // SoA split on the host, one copy per attribute
    std::vector<float> col(std::max(np, 1));
    for (int a = 0; a < 6; a++) {
        const std::vector<float> &src = a < 3 ? hp : hv;
        int c = a % 3;
        for (int q = 0; q < np; q++) col[q] = src[3 * q + c];
        hipCheck(hipMemcpy(_pa[a], col.data(), sizeof(float) * np, hipMemcpyHostToDevice), "upload particles");
    }
    _P = Particles{_pa[0], _pa[1], _pa[2], _pa[3], _pa[4], _pa[5], np};
    _Q = Particles{_pb[0], _pb[1], _pb[2], _pb[3], _pb[4], _pb[5], 0};
    if (_apic) {
        allocateAffine(false);
        for (int e = 0; e < 9; e++) {
            for (int q = 0; q < np; q++) col[q] = ha[9 * (size_t)q + e];
            hipCheck(hipMemcpy(_aa[e], col.data(), sizeof(float) * np, hipMemcpyHostToDevice), "upload affine");
        }
    }

    // static solids: boundary SDF (or the given one), weights, near-solid grid
    if (_solidGiven)
        hipCheck(hipMemcpy(_solid, _solidHost.data(), sizeof(float) * _solidHost.size(), hipMemcpyHostToDevice),
                 "upload solid SDF");
    else
        solidBoxPhi(_g, _solid, _stream);
    faceWeights(_g, _solid, _wu, _wv, _ww, _stream);
    int layers = (int)std::ceil((float)_cfl / (float)3);
    nearSolidGrid(_g, _solid, _near, _nearTmp, _ci, _cj, _ck, layers, _stream);
    hipCheck(hipMemsetAsync(_phi, 0, sizeof(float) * _phi.n, _stream), "memset");
    hipCheck(hipMemsetAsync(_pressure, 0, sizeof(float) * _pressure.n, _stream), "memset");
    hipCheck(hipMemsetAsync(_u, 0, sizeof(float) * _u.n, _stream), "memset");
    hipCheck(hipMemsetAsync(_v, 0, sizeof(float) * _v.n, _stream), "memset");
    hipCheck(hipMemsetAsync(_w, 0, sizeof(float) * _w.n, _stream), "memset");
    hipCheck(hipStreamSynchronize(_stream), "initialize");
    _initialized = true;
    _solidUpToDate = false;  // rebuilt, with the obstacles, on the first substep
}

double FluidSimulation::hostScalar(const double *d) {
    double h = 0.0;
    hipCheck(hipMemcpyAsync(&h, d, sizeof(double), hipMemcpyDeviceToHost, _stream), "D2H scalar");
    hipCheck(hipStreamSynchronize(_stream), "sync");
    return h;
}

// _calculateNextTimeStep (fluidsimulation.cpp:10165-10263)
double FluidSimulation::nextTimeStep(double dt) {
    double maxu = 0.0;
    if (_currentFrame == 0 && _stepNumber == 0) {
        FFVector3 b = bodyForce();
        float len = std::sqrt(b.x * b.x + b.y * b.y + b.z * b.z);
        maxu += len * dt;
    } else if (_P.n > 0) {
        maxSpeedSq(_P, _partial, _scal + 0, _stream);
        maxu = std::sqrt(hostScalar(_scal + 0));
    }
    if (_P.n > 0) maxu = std::max(0.0, maxu);
    double eps = 1e-6;
    double timeStep = _cfl * _g.dx / (maxu + eps);
    if (_tensionEnabled && _P.n > 0) {  // fluidsimulation.cpp:10248-10251
        double restriction = std::sqrt(_g.dx * _g.dx * _g.dx) * std::sqrt(1.0 / (_tensionK + eps));
        timeStep = std::min(timeStep, _tensionCondition * restriction);
    }
    int nsub = std::max((int)std::ceil(dt / timeStep), 1);
    return dt / nsub;
}

// FluidSimulation::update (fluidsimulation.cpp:10410-10543)
void FluidSimulation::update(double dt) {
    if (!_initialized) throw std::runtime_error("Error: FluidSimulation must be initialized before update.\n");
    if (dt < 0.0)
        throw std::domain_error("Error: delta time must be greater than or equal to 0.\ndelta time: " +
                                toString(dt) + "\n");
    // the frame's two timing events, destroyed also when a substep throws
    struct FrameEvents {
        hipEvent_t a = nullptr, b = nullptr;
        ~FrameEvents() {
            if (a) (void)hipEventDestroy(a);
            if (b) (void)hipEventDestroy(b);
        }
    } ev;
    hipCheck(hipEventCreate(&ev.a), "event");
    hipCheck(hipEventCreate(&ev.b), "event");
    hipEvent_t t0 = ev.a, t1 = ev.b;
    hipCheck(hipEventRecord(t0, _stream), "event");
    double epsdt = 1e-6;
    bool zeroLen = dt < epsdt;
    dt = std::max(dt, epsdt);
    _frameFinished = false;
    _frameDt = dt;
    _remaining = dt;
    _stepNumber = 0;
    bool skipped = zeroLen && _updatedOnce;
    double substepTime = _frameDt / (double)_minSteps;
    _pSuccess = true;
    _pIters = 0;
    _pErr = 0.0f;
    double timing[4] = {0, 0, 0, 0};
    double viscTotal = 0;
    size_t processed = 0;
    double eps = 1e-9;
    do {
        double step = std::fmin(nextTimeStep(dt), _remaining);
        double done = _frameDt - _remaining;
        double limit = (_stepNumber + 1) * substepTime;
        if (done + step > limit) step = std::fmin(substepTime, _remaining);
        if (_stepNumber == _maxSteps - 1) step = _remaining;
        _remaining -= step;
        // _stepFluid seeds the C library's rand() with frame + substep
        // (fluidsimulation.cpp:10079): the obstacles' sign tests draw from it
        srand((unsigned)(_currentFrame + _stepNumber));
        if (!skipped) {
            stepFluid(step);
            timing[0] += _phaseMs[0] + _phaseMs[1];
            timing[1] += _phaseMs[2] + _phaseMs[3] + _phaseMs[4];
            timing[2] += _phaseMs[5] + _phaseMs[6];
            timing[3] += _phaseMs[7];
            viscTotal += _viscMs;
        }
        processed += _P.n;
        _stepNumber++;
    } while (_remaining > eps);
    hipCheck(hipEventRecord(t1, _stream), "event");
    hipCheck(hipEventSynchronize(t1), "event");
    float total = 0;
    hipCheck(hipEventElapsedTime(&total, t0, t1), "event");

    FFFrameStats st{};
    st.frame = _currentFrame;
    st.substeps = _stepNumber;
    st.deltaTime = dt;
    st.fluidParticles = _P.n;
    st.diffuseParticles = 0;
    double secs = timing[3] * 1e-3;
    st.performanceScore = (processed == 0 || secs < 1e-9) ? -1 : (int)(((double)processed / secs) / 1000.0);
    st.pressureSolverEnabled = 1;
    st.pressureSolverSuccess = _pSuccess ? 1 : 0;
    st.pressureSolverError = (double)_pErr;
    st.pressureSolverIterations = _pIters;
    st.pressureSolverMaxIterations = _maxPressureIterations;
    st.viscositySolverEnabled = _viscEnabled ? 1 : 0;
    st.viscositySolverSuccess = _vSuccess ? 1 : 0;
    st.viscositySolverError = (double)_vErr;
    st.viscositySolverIterations = _vIters;
    st.viscositySolverMaxIterations = _viscMaxIterations;
    for (auto &m : st.mesh) m = FFMeshStats{0, -1, -1, 0};
    st.timing.total = total * 1e-3;
    st.timing.advection = timing[0] * 1e-3;
    st.timing.pressure = timing[1] * 1e-3;
    st.timing.particles = timing[2] * 1e-3;
    st.timing.viscosity = viscTotal * 1e-3;
    _stats = st;
    _updatedOnce = true;
    _currentFrame++;
    _frameFinished = true;
}

void FluidSimulation::step(double dt) {
    requireInitialized("step");
    if (!(dt > 0.0)) throw std::domain_error("Error: substep length must be greater than 0.\n");
    _frameDt = dt;
    _remaining = 0.0;  // one substep is the whole frame: progress 0
    srand((unsigned)(_currentFrame + _stepNumber));  // as _stepFluid does (fluidsimulation.cpp:10079)
    stepFluid(dt);
    _stepNumber++;
}

void FluidSimulation::extrapolate(int layers) {
    const int ni = _g.ni, nj = _g.nj, nk = _g.nk;
    _timers.begin(T_EXTRAPOLATE, _stream, 3 * layers + 3);
    extrapolateInit(ni + 1, nj, nk, _vu, _stu, _stream);
    extrapolateInit(ni, nj + 1, nk, _vv, _stv, _stream);
    extrapolateInit(ni, nj, nk + 1, _vw, _stw, _stream);
    for (int l = 0; l < layers; l++) {
        extrapolateLayer(ni + 1, nj, nk, _u, _stu, l, _stream);
        extrapolateLayer(ni, nj + 1, nk, _v, _stv, l, _stream);
        extrapolateLayer(ni, nj, nk + 1, _w, _stw, l, _stream);
    }
    _timers.end(T_EXTRAPOLATE, _stream);
}

// FluidSimulation::_stepFluid (fluidsimulation.cpp:10078-10121), plain configuration
void FluidSimulation::stepFluid(double dt) {
    const size_t nu = _u.n, nv = _v.n, nw = _w.n;
    const double radius = liquidRadius();
    const int layers = (int)std::ceil(std::sqrt(3.0) * _cfl) + 3;  // :6282-6286
    if (_P.n == 0) return;
    updateSolidLevelSet(dt);
    refreshFriction();
    hipCheck(hipEventRecord(_ev[0], _stream), "event");
    // liquid level set (5586-5617)
    _timers.begin(T_CELLSORT, _stream, 5);
    cellSort(_g, _P, _cs, _scan, _stream);
    _timers.end(T_CELLSORT, _stream);
    _timers.begin(T_LIQUID_SDF, _stream);
    liquidSDF(_g, _cs, _P.n, radius, _phi, _stream);
    _timers.end(T_LIQUID_SDF, _stream);
    liquidSDFPostprocess(_g, _solid, _phi, _stream);
    hipCheck(hipEventRecord(_ev[1], _stream), "event");
    // advect velocity field: P2G + extrapolation (5623-5661)
    _timers.begin(T_P2G, _stream, 3);
    if (_apic) {
        const float *vel[3] = {_P.vx, _P.vy, _P.vz};
        const float *aff[9];
        for (int e = 0; e < 9; e++) aff[e] = _A.a[e];
        float *fields[3] = {_u, _v, _w};
        uint8_t *valid[3] = {_vu, _vv, _vw};
        p2gApic(_g, _P.n, _P.px, _P.py, _P.pz, vel, aff, radius, fields, valid, _apic_scratch, _stream);
    } else {
        p2g(_g, _cs, radius, _u, _v, _w, _vu, _vv, _vw, _stream);
    }
    _timers.end(T_P2G, _stream);
    extrapolate(layers);
    // save (5671-5685) and body force (5837-5876)
    hipCheck(hipMemcpyAsync(_su, _u, sizeof(float) * nu, hipMemcpyDeviceToDevice, _stream), "save");
    hipCheck(hipMemcpyAsync(_sv, _v, sizeof(float) * nv, hipMemcpyDeviceToDevice, _stream), "save");
    hipCheck(hipMemcpyAsync(_sw, _w, sizeof(float) * nw, hipMemcpyDeviceToDevice, _stream), "save");
    FFVector3 bf = bodyForce();
    const float beps = 1e-6f;
    if (std::fabs(bf.x) > beps) addToFaces(_u, nu, (float)(bf.x * dt), _stream);
    if (std::fabs(bf.y) > beps) addToFaces(_v, nv, (float)(bf.y * dt), _stream);
    if (std::fabs(bf.z) > beps) addToFaces(_w, nw, (float)(bf.z * dt), _stream);
    hipCheck(hipEventRecord(_ev[8], _stream), "event");
    applyViscosity(dt);
    hipCheck(hipEventRecord(_ev[2], _stream), "event");
    // pressure (6152-6276)
    // the liquid surface's curvature and the tension terms' clusters
    // (fluidsimulation.cpp:5704-5730, pressuresolver.cpp:314-606)
    TensionView tension{};
    if (_tensionEnabled) {
        const size_t bytes = tensionBytes(_g);
        if (_tensionBlock.n < bytes) {
            _tensionBlock.alloc(bytes);
            tensionCarve(_g, _tensionBlock.p, _tb);
        }
        curvatureGrid(_g, _phi, _tb, _stream);
        tensionClusterStatus(_g, _phi, _tb, _stream);
        tension = TensionView{_tb.status, _tb.curvature, _phi, _tensionK};
    }
    _ps.setTension(tension);
    bool success = pressureSolve(dt);
    if (success) {
        _timers.begin(T_APPLY_PRESSURE, _stream, 3);
        applyPressure(_g, dt, _pressure, _phi, _wu, _wv, _ww, _u, _v, _w, _vu, _vv, _vw, _stream, tension);
        _timers.end(T_APPLY_PRESSURE, _stream);
    }
    if (_stepNumber == 0) {
        _pSuccess = success;
        _pIters = _lastIters;
        _pErr = (float)_lastErr;
    } else if (_pSuccess && (!success || _lastIters > _pIters)) {
        _pSuccess = success;
        _pIters = _lastIters;
        _pErr = (float)_lastErr;
    }
    extrapolate(layers);
    // constrain saved then current (6444-6459)
    constrainField(_su, _sv, _sw);
    constrainField(_u, _v, _w);
    hipCheck(hipEventRecord(_ev[5], _stream), "event");
    // G2P (6932-6945) + RK3 advection and collision (7853-7890)
    _timers.begin(T_G2P_ADVECT, _stream);
    g2pAdvect(_g, _P, MacView{_u, _v, _w}, MacView{_su, _sv, _sw}, _solid, _near, _ci, _cj, _ck,
              (float)_picflip, (float)(1 - _picflip), dt, _cfl, _stream, nullptr, _apic ? &_A : nullptr);
    _timers.end(T_G2P_ADVECT, _stream);
    hipCheck(hipEventRecord(_ev[6], _stream), "event");
    // removal (7773-7851)
    _timers.begin(T_REMOVAL, _stream);
    removeParticles();
    _timers.end(T_REMOVAL, _stream);
    hipCheck(hipEventRecord(_ev[7], _stream), "event");
    hipCheck(hipEventSynchronize(_ev[7]), "event");
    float ms[7];
    hipCheck(hipEventElapsedTime(&ms[0], _ev[0], _ev[1]), "hipEventElapsedTime");
    hipCheck(hipEventElapsedTime(&ms[1], _ev[1], _ev[8]), "hipEventElapsedTime");
    float viscMs = 0;
    hipCheck(hipEventElapsedTime(&viscMs, _ev[8], _ev[2]), "hipEventElapsedTime");
    _viscMs = viscMs;
    hipCheck(hipEventElapsedTime(&ms[2], _ev[2], _ev[3]), "hipEventElapsedTime");
    hipCheck(hipEventElapsedTime(&ms[3], _ev[3], _ev[4]), "hipEventElapsedTime");
    hipCheck(hipEventElapsedTime(&ms[4], _ev[4], _ev[5]), "hipEventElapsedTime");
    hipCheck(hipEventElapsedTime(&ms[5], _ev[5], _ev[6]), "hipEventElapsedTime");
    hipCheck(hipEventElapsedTime(&ms[6], _ev[6], _ev[7]), "hipEventElapsedTime");
    float whole = 0;
    hipCheck(hipEventElapsedTime(&whole, _ev[0], _ev[7]), "hipEventElapsedTime");
    for (int q = 0; q < 7; q++) _phaseMs[q] = ms[q];
    _phaseMs[7] = whole;
    _timers.collect();
}

// _constrainVelocityField (fluidsimulation.cpp:6355-6442): closed faces take
// the solid's face velocity (zero for a static solid); with obstacle friction
// the partly open faces blend toward it
void FluidSimulation::constrainField(float *u, float *v, float *w) {
    const float *svel[3] = {_movingSolid ? _us.p : nullptr, _movingSolid ? _vs.p : nullptr,
                            _movingSolid ? _ws.p : nullptr};
    float *faces[3] = {u, v, w};
    const float *wts[3] = {_wu, _wv, _ww};
    const size_t counts[3] = {_u.n, _v.n, _w.n};
    for (int q = 0; q < 3; q++) {
        if (_frictionActive)
            constrainFacesFriction(_g, q, faces[q], wts[q], _nodeObj, _objFric, _stream, svel[q]);
        else
            constrainFaces(faces[q], wts[q], counts[q], _stream, svel[q]);
    }
}

// _applyViscosityToVelocityField (fluidsimulation.cpp:5983-6056) with a
// constant viscosity: skipped while the value is 0, else the field is
// constrained and the viscosity solve runs on it; the frame keeps the first
// failing, else the most iterated, solve's statistics
void FluidSimulation::applyViscosity(double dt) {
    if (!_viscEnabled || _P.n == 0) return;
    const float value = (float)_viscosityValue;
    if (value == 0.0f) return;
    if (_viscosityGridValue != value || _viscNodes.n == 0) {
        const size_t n = (size_t)(_g.ni + 1) * (_g.nj + 1) * (_g.nk + 1);
        if (_viscNodes.n != n) _viscNodes.alloc(n);
        std::vector<float> h(n, value);
        hipCheck(hipMemcpyAsync(_viscNodes.p, h.data(), sizeof(float) * n, hipMemcpyHostToDevice, _stream),
                 "viscosity H2D");
        hipCheck(hipStreamSynchronize(_stream), "sync");
        _viscosityGridValue = value;
    }
    constrainField(_u, _v, _w);
    _timers.begin(T_VISCOSITY, _stream);
    ViscosityResult r = _visc.solve(_g, dt, _phi, _solid, _viscNodes, _viscTolerance, _viscMaxIterations, _u, _v,
                                    _w, _stream);
    _timers.end(T_VISCOSITY, _stream);
    _lastVisc = r;
    if (_stepNumber == 0) {
        _vSuccess = r.success;
        _vIters = r.iterations;
        _vErr = r.error;
    } else if (_vSuccess && (!r.success || r.iterations > _vIters)) {
        _vSuccess = r.success;
        _vIters = r.iterations;
        _vErr = r.error;
    }
}

// PressureSolver::solve + applySolutionToVelocityField gate (fluidsimulation.cpp:6206-6249)
bool FluidSimulation::pressureSolve(double dt) {
    // PressureSolver::solve conditions the moving solid's face velocities
    // first (pressuresolver.cpp:52-56); with a static solid they are all 0
    if (_movingSolid)
        conditionSolidVelocity(_g, _phi, _wu, _wv, _ww, _us, _vs, _ws, _condScratch, _stream);
    PressureResult r = _ps.solve(dt, _u, _v, _w, _wu, _wv, _ww, _phi, _pressureTolerance,
                                 _maxPressureIterations, (int)_precision, _innerTolerance, _pressure,
                                 _scan, _stream, _ev[3], &_timers, _referenceSummation ? _maxThreads : 0);
    _lastIters = r.iterations;
    _lastErr = r.error;
    _lastRhsInf = r.rhsInf;
    _lastSuccess = r.success ? 1 : 0;
    _lastUnknowns = r.unknowns;
    hipCheck(hipEventRecord(_ev[4], _stream), "event");
    return r.success;
}

// _removeMarkerParticles (fluidsimulation.cpp:7773-7851) with the speed limit
// of _getMarkerParticleSpeedLimit (7723-7771), evaluated with the frame delta
void FluidSimulation::removeParticles() {
    const double dt = _frameDt;
    const int nbins = _maxSteps;
    if (_bins.n < (size_t)nbins) _bins.alloc(nbins);
    double stepv = _cfl * _g.dx / dt;
    speedHistogram(_P, stepv, nbins, _bins, _partial, _scal + 6, _stream);
    std::vector<int> counts(nbins);
    hipCheck(hipMemcpyAsync(counts.data(), _bins, sizeof(int) * nbins, hipMemcpyDeviceToHost, _stream), "D2H");
    double maxParticleSpeed = hostScalar(_scal + 6);
    int maxRemoval = (int)std::fmin((int)((double)_P.n * 0.0005), 35);
    double maxspeed = _maxSteps * stepv;
    int cur = 0;
    for (int i = nbins - 1; i > 0; i--) {
        if (cur + counts[i] > maxRemoval) break;
        cur += counts[i];
        maxspeed = std::max(i + 4, _maxSteps) * stepv;
    }
    double lower = 0.90 * maxParticleSpeed, thr = 0.99999 * maxParticleSpeed;
    speedOutliers(_P, lower, thr, _outl, _stream);
    int outl[2];
    hipCheck(hipMemcpyAsync(outl, _outl, sizeof(int) * 2, hipMemcpyDeviceToHost, _stream), "D2H");
    hipCheck(hipStreamSynchronize(_stream), "sync");
    if (outl[1] <= 6 && outl[0] <= 6) maxspeed = std::min(thr, maxspeed);
    float fmaxspeed = (float)maxspeed;
    double maxspeedsq = fmaxspeed * fmaxspeed;
    removalFlags(_g, _P, _solid, _keys, _counts, _slot, _offsets, _perm, _scan, 250, _extremeRemoval,
                 maxspeedsq, _keep, _maxCount, _stream);
    const int before = _P.n;
    compactParticles(_P, _Q, _keep, _rank, _scan, _stream);
    if (_apic) {
        compactAffine(_A, _B, before, _keep, _rank, _stream);
        std::swap(_A, _B);
    }
    std::swap(_P, _Q);
    _Q.n = 0;
}

int FluidSimulation::getNumMarkerParticles() const { return _initialized ? _P.n : 0; }

namespace {
void interleave(const float *x, const float *y, const float *z, int start, int end, float *out,
                hipStream_t s) {
    int n = end - start;
    if (n <= 0) return;
    std::vector<float> a(n), b(n), c(n);
    hipCheck(hipMemcpyAsync(a.data(), x + start, sizeof(float) * n, hipMemcpyDeviceToHost, s), "D2H");
    hipCheck(hipMemcpyAsync(b.data(), y + start, sizeof(float) * n, hipMemcpyDeviceToHost, s), "D2H");
    hipCheck(hipMemcpyAsync(c.data(), z + start, sizeof(float) * n, hipMemcpyDeviceToHost, s), "D2H");
    hipCheck(hipStreamSynchronize(s), "sync");
    for (int q = 0; q < n; q++) {
        out[3 * q] = a[q];
        out[3 * q + 1] = b[q];
        out[3 * q + 2] = c[q];
    }
}
}  // namespace

// getMarkerParticleAffine{X,Y,Z}DataRange (fluidsimulation.cpp:3455-3501)
void FluidSimulation::getMarkerParticleAffineDataRange(int component, int start, int end, char *out) {
    if (start < 0 || end > getNumMarkerParticles() || start > end)
        throw std::domain_error("Error: invalid range.\nrange: [" + toString(start) + ", " + toString(end) + "]\n");
    if (!_apic || !_A.a[0])
        throw std::runtime_error("Error: affine data exists with the APIC velocity transfer only.\n");
    if (end > start)
        interleave(_A.a[3 * component], _A.a[3 * component + 1], _A.a[3 * component + 2], start, end, (float *)out,
                   _stream);
}

// getMarkerParticle{Position,Velocity}DataRange (fluidsimulation.cpp:3426-3460)
void FluidSimulation::getMarkerParticlePositionDataRange(int start, int end, char *out) {
    if (start < 0 || end > getNumMarkerParticles() || start > end)
        throw std::domain_error("Error: invalid range.\nrange: [" + toString(start) + ", " + toString(end) + "]\n");
    if (end > start) interleave(_P.px, _P.py, _P.pz, start, end, (float *)out, _stream);
}
void FluidSimulation::getMarkerParticleVelocityDataRange(int start, int end, char *out) {
    if (start < 0 || end > getNumMarkerParticles() || start > end)
        throw std::domain_error("Error: invalid range.\nrange: [" + toString(start) + ", " + toString(end) + "]\n");
    if (end > start) interleave(_P.vx, _P.vy, _P.vz, start, end, (float *)out, _stream);
}
void FluidSimulation::getMarkerParticlePositionData(char *out) {
    getMarkerParticlePositionDataRange(0, getNumMarkerParticles(), out);
}
// the reference's non-range velocity getter copies positions (fluidsimulation.cpp:3727-3735)
void FluidSimulation::getMarkerParticleVelocityData(char *out) {
    getMarkerParticlePositionDataRange(0, getNumMarkerParticles(), out);
}

void FluidSimulation::getLiquidSDF(float *out) {
    requireInitialized("reading grids");
    hipCheck(hipMemcpyAsync(out, _phi, sizeof(float) * _phi.n, hipMemcpyDeviceToHost, _stream), "D2H");
    hipCheck(hipStreamSynchronize(_stream), "sync");
}
void FluidSimulation::getSolidVelocityField(float *u, float *v, float *w) {
    requireInitialized("getSolidVelocityField");
    float *out[3] = {u, v, w};
    const float *src[3] = {_us.p, _vs.p, _ws.p};
    const size_t n[3] = {_u.n, _v.n, _w.n};
    for (int d = 0; d < 3; d++) {
        if (_movingSolid)
            hipCheck(hipMemcpyAsync(out[d], src[d], sizeof(float) * n[d], hipMemcpyDeviceToHost, _stream), "D2H");
        else
            std::fill(out[d], out[d] + n[d], 0.0f);
    }
    hipCheck(hipStreamSynchronize(_stream), "solid velocity");
}

void FluidSimulation::getSolidSDF(float *out) {
    requireInitialized("reading grids");
    hipCheck(hipMemcpyAsync(out, _solid, sizeof(float) * _solid.n, hipMemcpyDeviceToHost, _stream), "D2H");
    hipCheck(hipStreamSynchronize(_stream), "sync");
}
void FluidSimulation::getVelocityField(float *u, float *v, float *w) {
    requireInitialized("reading grids");
    hipCheck(hipMemcpyAsync(u, _u, sizeof(float) * _u.n, hipMemcpyDeviceToHost, _stream), "D2H");
    hipCheck(hipMemcpyAsync(v, _v, sizeof(float) * _v.n, hipMemcpyDeviceToHost, _stream), "D2H");
    hipCheck(hipMemcpyAsync(w, _w, sizeof(float) * _w.n, hipMemcpyDeviceToHost, _stream), "D2H");
    hipCheck(hipStreamSynchronize(_stream), "sync");
}
void FluidSimulation::getPressureGrid(float *out) {
    requireInitialized("reading grids");
    hipCheck(hipMemcpyAsync(out, _pressure, sizeof(float) * _pressure.n, hipMemcpyDeviceToHost, _stream), "D2H");
    hipCheck(hipStreamSynchronize(_stream), "sync");
}
void FluidSimulation::getLastPressureSolveStats(int *iterations, double *error, int *success, int *unknowns) {
    *iterations = _lastIters;
    *error = _lastErr;
    *success = _lastSuccess;
    *unknowns = _lastUnknowns;
}
void FluidSimulation::getKernelTiming(int id, double *ms, long long *launches) const {
    if (id < 0 || id >= T_COUNT) throw std::domain_error("Error: kernel timer id out of range.\n");
    *ms = _timers.ms(id);
    *launches = _timers.launches(id);
}
double FluidSimulation::getLastPhaseMilliseconds(int phase) {
    if (phase < 0 || phase > 8) throw std::domain_error("Error: phase must be in [0, 8].\n");
    return phase == 8 ? _viscMs : _phaseMs[phase];
}

}  // namespace ff
