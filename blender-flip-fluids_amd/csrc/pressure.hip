// pressure.hip -- host side of the GPU pressure solve.
#include "pressure.h"

#include <algorithm>
#include <cmath>
#include <stdexcept>
#include <string>

namespace ff {

void hipCheck(hipError_t e, const char *what);

PressureSolver::~PressureSolver() {
    dropGraphs();
    if (_block) (void)hipFree(_block);
    if (_hostScal) (void)hipHostFree(_hostScal);
}

void PressureSolver::allocate(Grid g) {
    _g = g;
    if (!_hostScal) {
        hipCheck(hipHostMalloc((void **)&_hostScal, 64, hipHostMallocMapped), "hipHostMalloc");
        hipCheck(hipHostGetDevicePointer((void **)&_hostScalDev, _hostScal, 0), "hipHostGetDevicePointer");
    }
    if (_block) (void)hipFree(_block);
    const size_t N = (size_t)g.ni * g.nj * g.nk;
    // pressure rows live on interior cells only
    const size_t M = (size_t)std::max(g.ni - 2, 1) * std::max(g.nj - 2, 1) * std::max(g.nk - 2, 1);
    const size_t L = (size_t)g.ni + g.nj + g.nk;
    size_t bytes = 0;
    auto take = [&](size_t nbytes) {
        size_t off = bytes;
        bytes += (nbytes + 255) / 256 * 256;
        return off;
    };
    size_t oMask = take(4 * (N + 1)), oRowOf = take(4 * (N + 1)), oRows = take(4 * M), oNbr = take(24 * M);
    size_t oCellRow = take(4 * N);
    {
        int ti, tj, tk;
        micTileCounts(g, std::min(micTileDepth(true), micTileDepth(false)), ti, tj, tk);
        _maxTiles = ti * tj * tk;
    }
    size_t oOrder = take(4 * (size_t)_maxTiles), oTileRows = take(4 * (size_t)_maxTiles);
    size_t oFlags = take(4 * (size_t)_maxTiles), oCounter = take(4), oError = take(4);
    _bandTilesMax = std::max(mic0BandTiles(g, true), mic0BandTiles(g, false));
    size_t oTileLo = take(4 * (size_t)_bandTilesMax), oTileHi = take(4 * (size_t)_bandTilesMax);
    size_t oLc = take(4 * (L + 1)), oLs = take(4 * (L + 1)), oLf = take(4 * L), oLr = take(4 * M);
    size_t oD[15], oF[14];
    for (auto &o : oD) o = take(8 * M);
    for (auto &o : oF) o = take(4 * M);
    size_t oV[5], oFV[6];
    for (auto &o : oV) o = take(8 * M);
    for (auto &o : oFV) o = take(4 * M);
    size_t oRec64[2], oRec32[2];
    for (auto &o : oRec64) o = take(sizeof(MicRec<double>) * M);
    for (auto &o : oRec32) o = take(sizeof(MicRec<float>) * M);
    pencilShape(g, _pp.npj, _pp.npk, _pp.nw);
    _pp.np = _pp.npj * _pp.npk;
    size_t oPOrder = take(4 * (size_t)_pp.np), oPWords = take(sizeof(LineWord) * (size_t)g.nj * g.nk * _pp.nw);
    const size_t exWords = pencilExchangeWords(g);
    size_t oPEx = take(8 * exWords), oPCounter = take(4);
    size_t oPart = take(8 * 1024), oScal = take(8 * 16);
    hipCheck(hipMalloc(&_block, bytes), "hipMalloc pressure");
    char *b = (char *)_block;
    _S.mask = (int *)(b + oMask);
    _S.rowOf = (int *)(b + oRowOf);
    _S.rows = (int *)(b + oRows);
    _S.nbr = (int *)(b + oNbr);
    _S.levelCount = (int *)(b + oLc);
    _S.levelStart = (int *)(b + oLs);
    _S.levelFill = (int *)(b + oLf);
    _S.levelRows = (int *)(b + oLr);
    _S.cellRow = (int *)(b + oCellRow);
    _ms.order = (int *)(b + oOrder);
    _ms.tileRows = (int *)(b + oTileRows);
    _ms.flags = (int *)(b + oFlags);
    _ms.counter = (int *)(b + oCounter);
    _ms.error = (int *)(b + oError);
    _tileLo = (int *)(b + oTileLo);
    _tileHi = (int *)(b + oTileHi);
    _hTileLo.assign(_bandTilesMax, 0);
    _hTileHi.assign(_bandTilesMax, 0);
    double *d[15];
    for (int q = 0; q < 15; q++) d[q] = (double *)(b + oD[q]);
    _S.rhs = d[0]; _S.diag = d[1]; _S.ox = d[2]; _S.oy = d[3]; _S.oz = d[4];
    _S.mx = d[5]; _S.my = d[6]; _S.mz = d[7]; _S.invdiag = d[8];
    _S.lx = d[9]; _S.ly = d[10]; _S.lz = d[11];
    _S.lmx = d[12]; _S.lmy = d[13]; _S.lmz = d[14];
    float *f[14];
    for (int q = 0; q < 14; q++) f[q] = (float *)(b + oF[q]);
    _S.finvdiag = f[0]; _S.flx = f[1]; _S.fly = f[2]; _S.flz = f[3];
    _S.fdiag = f[4]; _S.fox = f[5]; _S.foy = f[6]; _S.foz = f[7];
    _S.fmx = f[8]; _S.fmy = f[9]; _S.fmz = f[10];
    _S.flmx = f[11]; _S.flmy = f[12]; _S.flmz = f[13];
    _S.recFwd = (MicRec<double> *)(b + oRec64[0]);
    _S.recBwd = (MicRec<double> *)(b + oRec64[1]);
    _S.frecFwd = (MicRec<float> *)(b + oRec32[0]);
    _S.frecBwd = (MicRec<float> *)(b + oRec32[1]);
    _x = (double *)(b + oV[0]); _r = (double *)(b + oV[1]); _z = (double *)(b + oV[2]); _s = (double *)(b + oV[3]);
    _y = (double *)(b + oV[4]);
    _fr = (float *)(b + oFV[0]); _fz = (float *)(b + oFV[1]); _fs = (float *)(b + oFV[2]);
    _fd = (float *)(b + oFV[3]); _fq = (float *)(b + oFV[4]); _fy = (float *)(b + oFV[5]);
    _pp.order = (int *)(b + oPOrder);
    _pp.words = (LineWord *)(b + oPWords);
    _S.words = _pp.words;
    _S.nw = _pp.nw;
    _S.ni = g.ni;
    _S.nj = g.nj;
    _S.nk = g.nk;
    _pp.ex = (unsigned long long *)(b + oPEx);
    _pp.counter = (unsigned *)(b + oPCounter);
    _pp.error = _ms.error;
    _pp.base = 0;
    _pp.epoch = 0;
    {
        std::vector<int> order;
        pencilOrder(g, order);
        hipCheck(hipMemcpy(_pp.order, order.data(), sizeof(int) * order.size(), hipMemcpyHostToDevice),
                 "H2D pencil order");
        // epoch 0 marks a word no launch has published; the counter starts at 0
        hipCheck(hipMemset(_pp.ex, 0, 8 * exWords), "clear exchange words");
        hipCheck(hipMemset(_pp.counter, 0, 4), "clear ticket counter");
    }
    _partial = (double *)(b + oPart);
    _scal = (double *)(b + oScal);
    _S.m = 0;
    _S.nLevels = g.ni + g.nj + g.nk - 2;
    int dev = 0, cus = 0;
    hipCheck(hipGetDevice(&dev), "hipGetDevice");
    hipCheck(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev), "device attribute");
    _ms.grid = std::max(cus, 1);
    int wallKHz = 0;
    hipCheck(hipDeviceGetAttribute(&wallKHz, hipDeviceAttributeWallClockRate, dev), "device attribute");
    _ms.waitTicks = 2000LL * std::max(wallKHz, 1);  // 2 s
    _pp.waitTicks = _ms.waitTicks;
    int lds = 0;
    hipCheck(hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev), "device attribute");
    if ((size_t)lds < pencilLdsBytes(g) + 8448)
        throw std::runtime_error("the MIC(0) line pipeline needs " + std::to_string(pencilLdsBytes(g) + 8448) +
                                 " bytes of LDS per workgroup at this grid size; the device offers " +
                                 std::to_string(lds));
    if ((size_t)lds < micTileLds(true))
        throw std::runtime_error("the MIC(0) tile kernel needs " + std::to_string(micTileLds(true)) +
                                 " bytes of LDS per workgroup; the device offers " + std::to_string(lds));
    _ms.g = g;
}

// Non-empty tiles of the preconditioner's tile wavefront, listed by tile
// level (any order inside a level), with the flags, ticket counter and error
// word cleared for a new system.
void PressureSolver::buildSchedule(bool fp64, hipStream_t s) {
    _ms.depth = micTileDepth(fp64);
    micTileCounts(_g, _ms.depth, _ms.ti, _ms.tj, _ms.tk);
    _ms.nTiles = _ms.ti * _ms.tj * _ms.tk;
    micMarkTiles(_g, _S, _ms.depth, _ms.tileRows, s);
    _hTileRows.resize(_ms.nTiles);
    hipCheck(hipMemcpyAsync(_hTileRows.data(), _ms.tileRows, sizeof(int) * _ms.nTiles, hipMemcpyDeviceToHost, s),
             "D2H tiles");
    hipCheck(hipStreamSynchronize(s), "sync");
    const int nl = _ms.ti + _ms.tj + _ms.tk;
    std::vector<int> count(nl + 1, 0);
    for (int t = 0; t < _ms.nTiles; t++)
        if (_hTileRows[t]) count[t % _ms.ti + (t / _ms.ti) % _ms.tj + t / (_ms.ti * _ms.tj) + 1]++;
    for (int l = 0; l < nl; l++) count[l + 1] += count[l];
    _hOrder.assign(count[nl], 0);
    for (int t = 0; t < _ms.nTiles; t++)
        if (_hTileRows[t]) _hOrder[count[t % _ms.ti + (t / _ms.ti) % _ms.tj + t / (_ms.ti * _ms.tj)]++] = t;
    _ms.nActive = (int)_hOrder.size();
    if (_ms.nActive > 0)
        hipCheck(hipMemcpyAsync(_ms.order, _hOrder.data(), sizeof(int) * _ms.nActive, hipMemcpyHostToDevice, s),
                 "H2D tile order");
    hipCheck(hipMemsetAsync(_ms.flags, 0, sizeof(int) * _ms.nTiles, s), "memset flags");
    hipCheck(hipMemsetAsync(_ms.counter, 0, sizeof(int), s), "memset counter");
    _ms.base = 0;
    _ms.epoch = 0;
}

// hyperplane range of each band tile's rows, on the device and (after the
// stream's next synchronisation) in _hTileLo/_hTileHi
void PressureSolver::enqueueTileRanges(bool fp64, hipStream_t s) {
    const int nt = mic0BandTiles(_g, fp64);
    mic0BandTileRanges(_g, _S, fp64, _tileLo, _tileHi, s);
    hipCheck(hipMemcpyAsync(_hTileLo.data(), _tileLo, sizeof(int) * nt, hipMemcpyDeviceToHost, s), "D2H tiles");
    hipCheck(hipMemcpyAsync(_hTileHi.data(), _tileHi, sizeof(int) * nt, hipMemcpyDeviceToHost, s), "D2H tiles");
}

void PressureSolver::checkError(hipStream_t s) {
    int e = 0;
    hipCheck(hipMemcpyAsync(&e, _ms.error, sizeof(int), hipMemcpyDeviceToHost, s), "D2H error word");
    hipCheck(hipStreamSynchronize(s), "sync");
    if (e) throw std::runtime_error("MIC(0) preconditioner: a wait for a neighbour's results exceeded its time limit");
}

void PressureSolver::setSchedule(int schedule) {
    if (schedule != 2 && schedule != 4 && schedule != 5)
        throw std::domain_error(
            "the solver's preconditioner schedule must be 2 (bands), 4 (line pipeline) or 5 (bands from a graph)");
    _sched = schedule;
}

void PressureSolver::dropGraphs() {
    for (BandGraph &b : _graphs) {
        if (b.exec) (void)hipGraphExecDestroy(b.exec);
        b = BandGraph{};
    }
    _graphNext = 0;
}

// The band launches of one application are the same for every application
// of a solve (same plan, same buffers), so they are recorded once per (r, z)
// pair by stream capture and replayed with one hipGraphLaunch.  The capture
// holds kernel launches only.
template <class T>
void PressureSolver::graphApply(const T *r, T *y, T *z, hipStream_t s) {
    if (_S.m == 0) return;
    BandGraph *slot = nullptr;
    for (BandGraph &b : _graphs)
        if (b.exec && b.r == r && b.z == z) slot = &b;
    if (!slot) {
        slot = &_graphs[_graphNext];
        _graphNext = (_graphNext + 1) % 2;
        if (slot->exec) (void)hipGraphExecDestroy(slot->exec);
        *slot = BandGraph{};
        hipGraph_t graph = nullptr;
        hipCheck(hipStreamBeginCapture(s, hipStreamCaptureModeRelaxed), "begin capture");
        try {
            if constexpr (sizeof(T) == 8)
                mic0ApplyBands(_g, _S, _plan, _tileLo, _tileHi, r, y, z, s);
            else
                mic0ApplyBandsF(_g, _S, _plan, _tileLo, _tileHi, r, y, z, s);
        } catch (...) {
            (void)hipStreamEndCapture(s, &graph);
            if (graph) (void)hipGraphDestroy(graph);
            throw;
        }
        hipCheck(hipStreamEndCapture(s, &graph), "end capture");
        const hipError_t e = hipGraphInstantiate(&slot->exec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        hipCheck(e, "instantiate the band graph");
        slot->r = r;
        slot->z = z;
    }
    hipCheck(hipGraphLaunch(slot->exec, s), "launch the band graph");
}

void PressureSolver::precond(const double *r, double *y, double *z, hipStream_t s) {
    if (_sched == 4)
        pencilApply(_g, _S, _pp, r, y, z, s);
    else if (_sched == 5)
        graphApply(r, y, z, s);
    else
        mic0ApplyBands(_g, _S, _plan, _tileLo, _tileHi, r, y, z, s);
}
void PressureSolver::precondF(const float *r, float *y, float *z, hipStream_t s) {
    if (_sched == 4)
        pencilApplyF(_g, _S, _pp, r, y, z, s);
    else if (_sched == 5)
        graphApply(r, y, z, s);
    else
        mic0ApplyBandsF(_g, _S, _plan, _tileLo, _tileHi, r, y, z, s);
}

void PressureSolver::applyPreconditioner(const double *r, double *z, int schedule, hipStream_t s) {
    if (schedule == 4) {
        pencilLineWords(_g, _S, _pp, s);
        pencilApply(_g, _S, _pp, r, _y, z, s);
    } else if (schedule == 1) {
        buildSchedule(true, s);
        micApplyTiles(_S, _ms, r, z, s);
    } else if (schedule == 2 || schedule == 3) {
        enqueueTileRanges(true, s);
        hipCheck(hipStreamSynchronize(s), "sync");
        const bool host = schedule == 2;
        mic0BandPlan(_g, _S, _hLevelStart.data(), true, host ? _hTileLo.data() : nullptr,
                     host ? _hTileHi.data() : nullptr, _plan);
        mic0ApplyBands(_g, _S, _plan, _tileLo, _tileHi, r, _y, z, s);
    } else {
        mic0Apply(_S, _hLevelStart.data(), r, z, s);
    }
}
void PressureSolver::applyPreconditionerF(const float *r, float *z, int schedule, hipStream_t s) {
    if (schedule == 4) {
        pencilLineWords(_g, _S, _pp, s);
        pencilApplyF(_g, _S, _pp, r, _fy, z, s);
    } else if (schedule == 1) {
        buildSchedule(false, s);
        micApplyTilesF(_S, _ms, r, z, s);
    } else if (schedule == 2 || schedule == 3) {
        enqueueTileRanges(false, s);
        hipCheck(hipStreamSynchronize(s), "sync");
        const bool host = schedule == 2;
        mic0BandPlan(_g, _S, _hLevelStart.data(), false, host ? _hTileLo.data() : nullptr,
                     host ? _hTileHi.data() : nullptr, _plan);
        mic0ApplyBandsF(_g, _S, _plan, _tileLo, _tileHi, r, _fy, z, s);
    } else {
        mic0ApplyF(_S, _hLevelStart.data(), r, z, s);
    }
}

// Reads one device scalar.  (The solver's band schedule never raises the tile
// wavefront's error word; applyPreconditioner checks it for that schedule.)
double PressureSolver::hostScalar(const double *d, hipStream_t s) {
    double h = 0.0;
    hipCheck(hipMemcpyAsync(&h, d, sizeof(double), hipMemcpyDeviceToHost, s), "D2H scalar");
    hipCheck(hipStreamSynchronize(s), "sync");
    return h;
}

// The |r|_inf the last PCG step's reduction also stored in mapped pinned host
// memory: a stream synchronize instead of a copy per iteration.
double PressureSolver::mappedResidual(hipStream_t s) {
    hipCheck(hipStreamSynchronize(s), "sync");
    return *(volatile const double *)_hostScal;
}

// PressureSolver::solve (pressuresolver.cpp:52-92) and _solveLinearSystem (869-916)
PressureResult PressureSolver::solve(double dt, const float *u, const float *v, const float *w,
                                     const float *wu, const float *wv, const float *ww, const float *phi,
                                     double tol, int maxIter, int mode, double innerTol, float *pressure,
                                     int *scanScratch, hipStream_t s, hipEvent_t afterAssembly,
                                     KernelTimers *timers, int referenceDots, const double *cellRhs,
                                     double *cellSolution) {
    PressureResult res;
    _refDots = referenceDots;
    if (cellSolution)
        hipCheck(hipMemsetAsync(cellSolution, 0, sizeof(double) * (size_t)_g.ni * _g.nj * _g.nk, s), "memset");
    _t = timers ? timers : &_noTimers;
    hipCheck(hipMemsetAsync(_ms.error, 0, sizeof(int), s), "memset error word");
    _t->begin(T_ASSEMBLE, s);
    const size_t N = (size_t)_g.ni * _g.nj * _g.nk;
    pressureMask(_g, phi, _S.mask, s);
    exclusiveScan(_S.mask, _S.rowOf, N, scanScratch, s);
    int m = 0;
    hipCheck(hipMemcpyAsync(&m, _S.rowOf + N, sizeof(int), hipMemcpyDeviceToHost, s), "D2H");
    hipCheck(hipStreamSynchronize(s), "sync");
    _S.m = m;
    res.unknowns = m;
    pressureRows(_g, _S.mask, _S.rowOf, _S.rows, _S.cellRow, s);
    pressureAssemble(_g, dt, u, v, w, wu, wv, ww, phi, _S, s, _solid, _tension);
    if (cellRhs) gatherRows(_S, cellRhs, _S.rhs, s);
    pressureLevels(_g, _S, scanScratch, s);
    if (m > 0) pencilLineWords(_g, _S, _pp, s);  // SpMV and the line pipeline find neighbour rows in them
    _t->end(T_ASSEMBLE, s);
    _hLevelStart.resize(_S.nLevels + 1);
    hipCheck(hipMemcpyAsync(_hLevelStart.data(), _S.levelStart, sizeof(int) * (_S.nLevels + 1),
                            hipMemcpyDeviceToHost, s), "D2H levels");
    enqueueTileRanges(mode == 0, s);
    double maxAbs = 0.0;
    if (m > 0) {
        absMax(_S.rhs, m, _partial, _scal + 1, s);
        maxAbs = hostScalar(_scal + 1, s);  // also waits for the level table and the tile ranges
    } else {
        hipCheck(hipStreamSynchronize(s), "sync");
    }
    res.rhsInf = maxAbs;
    if (afterAssembly) hipCheck(hipEventRecord(afterAssembly, s), "event");
    _levelLaunches = mic0FactorBandLaunches(_S, _hLevelStart.data());
    mic0BandPlan(_g, _S, _hLevelStart.data(), mode == 0, _hTileLo.data(), _hTileHi.data(), _plan);
    _bandLaunches = (int)_plan.l0.size();
    dropGraphs();  // recorded with the previous system's plan
    // The forward solve skips a right-hand side below the tolerance as an
    // absolute threshold, as the reference does (pressuresolver.cpp:69-75).  A
    // given right-hand side (the adjoint's) is solved unless it is exactly 0,
    // so the solution stays linear in it at every scale.
    if (cellRhs ? maxAbs == 0.0 : maxAbs < tol) {
        hipCheck(hipMemsetAsync(pressure, 0, sizeof(float) * N, s), "memset");
        res.success = true;
        res.iterations = 0;
        res.error = 0.0;
        return res;
    }
    int iters = 0;
    double resid = -1.0;
    _prevPressure = pressure;
    bool conv = mode == 0   ? pcgFP64(tol, maxIter, maxAbs, iters, resid, s)
                : mode == 1 ? pcgMixed(tol, maxIter, maxAbs, innerTol, iters, resid, s)
                : mode == 2 ? pcgReliable(tol, maxIter, maxAbs, innerTol, iters, resid, s)
                            : pcgPrecond32(tol, maxIter, maxAbs, iters, resid, s);
    checkError(s);
    pressureToGrid(_g, _S, _x, pressure, s);
    if (cellSolution) scatterRows(_S, _x, cellSolution, s);
    float e = (float)resid;
    res.success = conv || (iters == maxIter && e < 1.0);
    res.iterations = iters;
    res.error = e;
    return res;
}

// PCGSolver<double>::solve (pcgsolver.h:268-322): x0 = 0, MIC(0), float step
// lengths, |r|_inf stop test against min(tol * |b|_inf, 1.0)
void PressureSolver::initialGuess(hipStream_t st) {
    const int m = _S.m;
    if (_warm && _prevPressure) {
        gatherRowsF(_S, _prevPressure, _x, st);
        residual(_S, _S.rhs, _x, _r, st);
    } else {
        hipCheck(hipMemcpyAsync(_r, _S.rhs, sizeof(double) * m, hipMemcpyDeviceToDevice, st), "copy");
        hipCheck(hipMemsetAsync(_x, 0, sizeof(double) * m, st), "memset");
    }
}

bool PressureSolver::pcgFP64(double tolFactor, int maxIter, double resid0, int &iters, double &resid,
                             hipStream_t st) {
    const int m = _S.m;
    const int *hs = _hLevelStart.data();
    if (tolFactor < 1e-30) tolFactor = 1e-30;
    initialGuess(st);
    resid = resid0;
    double tol = tolFactor * resid;
    _t->begin(T_MIC0_FACTOR, st, _levelLaunches + 2);
    mic0FactorBands(_g, _S, hs, _hTileLo.data(), _hTileHi.data(), _tileLo, _tileHi, st);
    mic0LFactor(_S, st);
    mic0LowerEntries(_S, true, st);
    _t->end(T_MIC0_FACTOR, st);
    _t->begin(T_MIC0_APPLY, st, applyLaunches());
    precond(_r, _y, _z, st);
    _t->end(T_MIC0_APPLY, st);
    double *rho = _scal + 2, *rhoNew = _scal + 3, *sz = _scal + 4, *rn = _scal + 5;
    // dot products: the reference's order, or the deterministic tree
    auto dotp = [&](const double *a, const double *b, double *out) {
        if (_refDots > 0)
            dotReference(a, b, m, _refDots, _partial, out, st);
        else
            dot(a, b, m, _partial, out, st);
    };
    dotp(_z, _r, rho);
    double hrho = hostScalar(rho, st);
    if (hrho == 0 || hrho != hrho) {
        iters = 0;
        return false;
    }
    hipCheck(hipMemcpyAsync(_s, _z, sizeof(double) * m, hipMemcpyDeviceToDevice, st), "copy");
    double *s = _s, *z = _z;
    int it;
    for (it = 0; it < maxIter; it++) {
        _t->begin(T_SPMV, st);
        if (_refDots > 0) {
            spmv(_S, s, z, st);
            dotReference(s, z, m, _refDots, _partial, sz, st);
        } else {
            spmvDot(_S, s, z, _partial, sz, st);
        }
        _t->end(T_SPMV, st);
        pcgStep(_x, s, _r, z, m, rho, sz, _partial, rn, st, nullptr, _hostScalDev);
        resid = mappedResidual(st);
        if (resid <= std::min(tol, 1.0)) {
            iters = it + 1;
            return true;
        }
        _t->begin(T_MIC0_APPLY, st, applyLaunches());
        precond(_r, _y, z, st);
        _t->end(T_MIC0_APPLY, st);
        dotp(z, _r, rhoNew);
        // beta = rhoNew / rho; z += (float)beta * s; then s <-> z
        axpyFloatStep(z, s, m, rhoNew, rho, 1.0, st);
        std::swap(s, z);
        std::swap(rho, rhoNew);
    }
    iters = it;
    return false;
}

// FP64 PCG (pcgsolver.h:268-322) whose MIC(0) application runs on the FP32
// copy of the factor: the step that updates r also rounds it to float for the
// sweeps, and their result is widened into the search-direction update.  The
// SpMV, the dot products, the step lengths and every vector stay FP64, so the
// iteration follows the FP64 solve except for the preconditioner's rounding.
bool PressureSolver::pcgPrecond32(double tolFactor, int maxIter, double resid0, int &iters, double &resid,
                                  hipStream_t st) {
    const int m = _S.m;
    const int *hs = _hLevelStart.data();
    if (tolFactor < 1e-30) tolFactor = 1e-30;
    initialGuess(st);
    resid = resid0;
    double tol = tolFactor * resid;
    _t->begin(T_MIC0_FACTOR, st, _levelLaunches + 2);
    mic0FactorBands(_g, _S, hs, _hTileLo.data(), _hTileHi.data(), _tileLo, _tileHi, st);
    mic0LFactor(_S, st);
    _t->end(T_MIC0_FACTOR, st);
    mic0ToFloat(_S, st);
    mic0LowerEntries(_S, false, st);
    castToFloat(_r, _fr, m, st);
    _t->begin(T_MIC0_APPLY, st, applyLaunches());
    precondF(_fr, _fy, _fz, st);
    _t->end(T_MIC0_APPLY, st);
    double *rho = _scal + 2, *rhoNew = _scal + 3, *sz = _scal + 4, *rn = _scal + 5;
    dotfd(_fz, _r, m, _partial, rho, st);
    double hrho = hostScalar(rho, st);
    if (hrho == 0 || hrho != hrho) {
        iters = 0;
        return false;
    }
    hipCheck(hipMemsetAsync(_s, 0, sizeof(double) * m, st), "memset");
    addFloatToDouble(_s, _fz, m, st);
    double *s = _s, *z = _z;
    int it;
    for (it = 0; it < maxIter; it++) {
        _t->begin(T_SPMV, st);
        spmvDot(_S, s, z, _partial, sz, st);
        _t->end(T_SPMV, st);
        pcgStep(_x, s, _r, z, m, rho, sz, _partial, rn, st, _fr, _hostScalDev);
        resid = mappedResidual(st);
        if (resid <= std::min(tol, 1.0)) {
            iters = it + 1;
            return true;
        }
        _t->begin(T_MIC0_APPLY, st, applyLaunches());
        precondF(_fr, _fy, _fz, st);
        _t->end(T_MIC0_APPLY, st);
        dotfd(_fz, _r, m, _partial, rhoNew, st);
        // beta = rhoNew / rho; z = (double)fz + (float)beta * s; then s <-> z
        widenAxpyStep(z, _fz, s, m, rhoNew, rho, st);
        std::swap(s, z);
        std::swap(rho, rhoNew);
    }
    iters = it;
    return false;
}

// FP32 PCG inside FP64 iterative refinement.  The outer loop keeps the FP64
// solution and its true residual r = b - A x; each inner solve runs PCG with
// the same MIC(0) factor rounded to FP32 on FP32 vectors (FP64 dot products)
// to a relative tolerance on its own right-hand side.  The stop test is the
// reference's |r|_inf <= min(tol * |b|_inf, 1.0), on the true residual.
bool PressureSolver::pcgMixed(double tolFactor, int maxIter, double resid0, double innerRel, int &iters,
                              double &resid, hipStream_t st) {
    const int m = _S.m;
    const int *hs = _hLevelStart.data();
    if (tolFactor < 1e-30) tolFactor = 1e-30;
    double tol = std::min(tolFactor * resid0, 1.0);
    hipCheck(hipMemcpyAsync(_r, _S.rhs, sizeof(double) * m, hipMemcpyDeviceToDevice, st), "copy");
    hipCheck(hipMemsetAsync(_x, 0, sizeof(double) * m, st), "memset");
    _t->begin(T_MIC0_FACTOR, st, _levelLaunches + 2);
    mic0FactorBands(_g, _S, hs, _hTileLo.data(), _hTileHi.data(), _tileLo, _tileHi, st);
    mic0LFactor(_S, st);
    _t->end(T_MIC0_FACTOR, st);
    mic0ToFloat(_S, st);
    mic0LowerEntries(_S, false, st);
    double *rho = _scal + 2, *rhoNew = _scal + 3, *sz = _scal + 4, *rn = _scal + 5;
    resid = resid0;
    iters = 0;
    int outer = 0;
    while (resid > tol && iters < maxIter && outer < 100) {
        outer++;
        castToFloat(_r, _fr, m, st);
        absMaxf(_fr, m, _partial, rn, st);
        double innerTol = innerRel * hostScalar(rn, st);
        hipCheck(hipMemsetAsync(_fd, 0, sizeof(float) * m, st), "memset");
        _t->begin(T_MIC0_APPLY, st, applyLaunches());
        precondF(_fr, _fy, _fz, st);
        _t->end(T_MIC0_APPLY, st);
        dotf(_fz, _fr, m, _partial, rho, st);
        double hrho = hostScalar(rho, st);
        if (hrho == 0 || hrho != hrho) return false;
        hipCheck(hipMemcpyAsync(_fs, _fz, sizeof(float) * m, hipMemcpyDeviceToDevice, st), "copy");
        float *s = _fs, *z = _fz;
        while (iters < maxIter) {
            _t->begin(T_SPMV, st);
            spmvDotF(_S, s, _fq, _partial, sz, st);
            _t->end(T_SPMV, st);
            pcgStepF(_fd, s, _fr, _fq, m, rho, sz, _partial, rn, st, _hostScalDev);
            iters++;
            if (mappedResidual(st) <= innerTol) break;
            _t->begin(T_MIC0_APPLY, st, applyLaunches());
            precondF(_fr, _fy, z, st);
            _t->end(T_MIC0_APPLY, st);
            dotf(z, _fr, m, _partial, rhoNew, st);
            axpyFloatStepF(z, s, m, rhoNew, rho, 1.0, st);
            std::swap(s, z);
            std::swap(rho, rhoNew);
        }
        addFloatToDouble(_x, _fd, m, st);
        residual(_S, _S.rhs, _x, _r, st);
        absMax(_r, m, _partial, rn, st);
        resid = hostScalar(rn, st);
    }
    return resid <= tol;
}

// FP32 PCG whose FP64 solution is kept exact by reliable updates (Sleijpen &
// van der Vorst, 1996).  The search directions, the FP32 residual, SpMV and the
// preconditioner run in FP32; x accumulates in FP64 with the float-rounded
// step length; dot products accumulate in FP64.  Whenever the FP32 residual
// has dropped by `delta` since the last update (or meets the tolerance), the
// true residual r = b - A x is recomputed in FP64, tested against the
// reference's stop rule |r|_inf <= min(tol * |b|_inf, 1.0), and replaces the
// FP32 residual; the search direction is kept, so the Krylov space is not
// restarted.  Iterations counts every FP32 PCG step.
bool PressureSolver::pcgReliable(double tolFactor, int maxIter, double resid0, double delta, int &iters,
                                 double &resid, hipStream_t st) {
    const int m = _S.m;
    const int *hs = _hLevelStart.data();
    if (tolFactor < 1e-30) tolFactor = 1e-30;
    const double tol = std::min(tolFactor * resid0, 1.0);
    hipCheck(hipMemsetAsync(_x, 0, sizeof(double) * m, st), "memset");
    _t->begin(T_MIC0_FACTOR, st, _levelLaunches + 2);
    mic0FactorBands(_g, _S, hs, _hTileLo.data(), _hTileHi.data(), _tileLo, _tileHi, st);
    mic0LFactor(_S, st);
    _t->end(T_MIC0_FACTOR, st);
    mic0ToFloat(_S, st);
    mic0LowerEntries(_S, false, st);
    double *rho = _scal + 2, *rhoNew = _scal + 3, *sz = _scal + 4, *rn = _scal + 5;
    castToFloat(_S.rhs, _fr, m, st);
    double sinceUpdate = resid0;
    resid = resid0;
    iters = 0;
    _t->begin(T_MIC0_APPLY, st, applyLaunches());
    precondF(_fr, _fy, _fz, st);
    _t->end(T_MIC0_APPLY, st);
    dotf(_fz, _fr, m, _partial, rho, st);
    double hrho = hostScalar(rho, st);
    if (hrho == 0 || hrho != hrho) return false;
    hipCheck(hipMemcpyAsync(_fs, _fz, sizeof(float) * m, hipMemcpyDeviceToDevice, st), "copy");
    float *s = _fs, *z = _fz;
    bool trueKnown = true;  // resid holds the true residual of the current x
    while (iters < maxIter) {
        _t->begin(T_SPMV, st);
        spmvDotF(_S, s, _fq, _partial, sz, st);
        _t->end(T_SPMV, st);
        pcgStepMixed(_x, s, _fr, _fq, m, rho, sz, _partial, rn, st, _hostScalDev);
        iters++;
        trueKnown = false;
        const double rf = mappedResidual(st);
        if (rf <= delta * sinceUpdate || rf <= tol) {
            residual(_S, _S.rhs, _x, _r, st);
            absMax(_r, m, _partial, rn, st);
            resid = hostScalar(rn, st);
            trueKnown = true;
            if (resid <= tol) return true;
            castToFloat(_r, _fr, m, st);
            sinceUpdate = resid;
        }
        _t->begin(T_MIC0_APPLY, st, applyLaunches());
        precondF(_fr, _fy, z, st);
        _t->end(T_MIC0_APPLY, st);
        dotf(z, _fr, m, _partial, rhoNew, st);
        axpyFloatStepF(z, s, m, rhoNew, rho, 1.0, st);
        std::swap(s, z);
        std::swap(rho, rhoNew);
    }
    if (!trueKnown) {
        residual(_S, _S.rhs, _x, _r, st);
        absMax(_r, m, _partial, rn, st);
        resid = hostScalar(rn, st);
    }
    return resid <= tol;
}

}  // namespace ff
