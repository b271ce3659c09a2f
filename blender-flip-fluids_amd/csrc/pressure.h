// pressure.h -- GPU pressure solve of the FLIP substep (host side).
//
// PressureSolver mirrors the reference's PressureSolver::solve +
// _solveLinearSystem (pressuresolver.cpp:52-92, 869-916) and
// PCGSolver<double>::solve (pcgsolver.h:268-322), with an additional
// mixed-precision mode.  It owns the system buffers for one grid size.
#pragma once

#include <vector>

#include "kernels.h"
#include "timers.h"

namespace ff {

struct PressureResult {
    bool success = true;  // after the acceptable-tolerance rule
    int iterations = 0;
    double error = 0.0;   // |r|_inf at exit (the reference rounds it to float)
    double rhsInf = 0.0;  // |b|_inf of the right-hand side
    int unknowns = 0;
};

class PressureSolver {
public:
    PressureSolver() = default;
    ~PressureSolver();
    PressureSolver(const PressureSolver &) = delete;
    PressureSolver &operator=(const PressureSolver &) = delete;

    void allocate(Grid g);
    // Builds the system from the current fields, solves it and writes the
    // float pressure grid.  cellRhs (N doubles, optional) replaces the
    // assembled right-hand side by its values at the rows' cells (the matrix
    // still comes from phi and the weights); cellSolution (N doubles,
    // optional) receives the FP64 solution on the rows' cells and 0 elsewhere.
    // referenceDots > 0 makes the FP64 solve (mode 0)
    // take its dot products in the reference's order for that thread count.
    // mode 0 = FP64 PCG; 1 = FP32 PCG restarted inside
    // FP64 iterative refinement (innerTol = relative tolerance of each inner
    // solve); 2 = FP32 PCG with FP64 reliable updates (innerTol = drop of the
    // FP32 residual between true-residual updates); 3 = FP64 PCG with the
    // MIC(0) application in FP32 (innerTol unused).
    PressureResult solve(double dt, const float *u, const float *v, const float *w, const float *wu,
                         const float *wv, const float *ww, const float *phi, double tol, int maxIter,
                         int mode, double innerTol, float *pressure, int *scanScratch, hipStream_t s,
                         hipEvent_t afterAssembly, KernelTimers *timers = nullptr, int referenceDots = 0,
                         const double *cellRhs = nullptr, double *cellSolution = nullptr);
    const PressureSys &system() const { return _S; }
    // FP64 PCG (modes 0 and 3): start from the pressure grid's values at
    // the rows (the previous solve's result) instead of 0.  The stopping
    // test keeps the zero-start threshold, tolerance x |b|_inf.
    void setWarmStart(bool on) { _warm = on; }
    // the moving solid's face velocities and cell weights for the divergence
    // of the following solves (all null: a static solid)
    void setSolidTerms(SolidTerms t) { _solid = t; }
    // the surface tension terms of the following solves' divergence (status
    // null: none)
    void setTension(TensionView t) { _tension = t; }
    // the solver's MIC(0) application: 2 = bands of hyperplanes, 4 = the
    // line pipeline (two launches per application), 5 = the bands of one
    // application replayed from a HIP graph captured once per solve
    void setSchedule(int schedule);
    int schedule() const { return _sched; }
    const std::vector<int> &levelTable() const { return _hLevelStart; }
    // one preconditioner application z = (L L^T)^-1 r on the current factor
    // (after solve()); FP64 or the FP32 copy.  schedule 0 = one launch per
    // hyperplane, 1 = tile wavefront (switched off: raises an error),
    // 2 = bands of hyperplanes, 3 = the same bands with the tile ranges
    // checked on the device instead of passed in the launches, 4 = the line
    // pipeline.
    void applyPreconditioner(const double *r, double *z, int schedule, hipStream_t s);
    void applyPreconditionerF(const float *r, float *z, int schedule, hipStream_t s);
    // throws when a preconditioner launch gave up waiting (waits for the stream)
    void checkError(hipStream_t s);

private:
    Grid _g{};
    PressureSys _S{};
    void *_block = nullptr;  // one allocation for everything
    double *_x = nullptr, *_r = nullptr, *_z = nullptr, *_s = nullptr, *_y = nullptr;
    float *_fr = nullptr, *_fz = nullptr, *_fs = nullptr, *_fd = nullptr, *_fq = nullptr, *_fy = nullptr;
    double *_partial = nullptr, *_scal = nullptr;
    double *_hostScal = nullptr, *_hostScalDev = nullptr;  // mapped pinned host slot and its device alias
    std::vector<int> _hLevelStart;
    MicSchedule _ms;
    int _maxTiles = 0;
    int *_tileLo = nullptr, *_tileHi = nullptr;  // band schedule: hyperplane range of each tile's rows
    int _bandTilesMax = 0;
    std::vector<int> _hTileLo, _hTileHi;  // host copies of the tile ranges
    BandPlan _plan;                       // band launches of the current system
    std::vector<int> _hTileRows, _hOrder;
    KernelTimers *_t = nullptr;
    KernelTimers _noTimers;
    int _levelLaunches = 0, _bandLaunches = 0;
    PencilPlan _pp;  // line-pipeline buffers (inside _block)
    int _sched = 2;
    // z = M^-1 r with the solver's schedule (y: forward result)
    void precond(const double *r, double *y, double *z, hipStream_t s);
    void precondF(const float *r, float *y, float *z, hipStream_t s);
    int applyLaunches() const { return _sched == 4 ? 2 : _bandLaunches; }
    // schedule 5: executable graphs of the band launches for the (r, z)
    // pairs of the current solve's applications (the PCG's z alternates
    // between two buffers), dropped when a new system is assembled
    struct BandGraph {
        const void *r = nullptr, *z = nullptr;
        hipGraphExec_t exec = nullptr;
    };
    BandGraph _graphs[2];
    int _graphNext = 0;
    void dropGraphs();
    template <class T>
    void graphApply(const T *r, T *y, T *z, hipStream_t s);
    int _refDots = 0;  // thread count of the reference's dot order, or 0 for the tree reductions
    bool _warm = false;
    SolidTerms _solid{};
    TensionView _tension{};
    const float *_prevPressure = nullptr;  // the pressure grid at solve() entry
    // x = 0, r = b; or, warm, x from the pressure grid and r = b - A x
    void initialGuess(hipStream_t st);

    double hostScalar(const double *d, hipStream_t s);
    double mappedResidual(hipStream_t s);
    void enqueueTileRanges(bool fp64, hipStream_t s);
    // tile schedule of the preconditioner application for this system
    void buildSchedule(bool fp64, hipStream_t s);
    bool pcgFP64(double tolFactor, int maxIter, double resid0, int &iters, double &resid, hipStream_t s);
    bool pcgMixed(double tolFactor, int maxIter, double resid0, double innerTol, int &iters,
                  double &resid, hipStream_t s);
    bool pcgReliable(double tolFactor, int maxIter, double resid0, double delta, int &iters, double &resid,
                     hipStream_t s);
    bool pcgPrecond32(double tolFactor, int maxIter, double resid0, int &iters, double &resid, hipStream_t s);
};

}  // namespace ff
