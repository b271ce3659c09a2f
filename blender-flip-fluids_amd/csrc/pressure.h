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
    // float pressure grid.  mode 0 = FP64 PCG, 1 = FP32 PCG in FP64 refinement.
    PressureResult solve(double dt, const float *u, const float *v, const float *w, const float *wu,
                         const float *wv, const float *ww, const float *phi, double tol, int maxIter,
                         int mode, double innerTol, float *pressure, int *scanScratch, hipStream_t s,
                         hipEvent_t afterAssembly, KernelTimers *timers = nullptr);
    const PressureSys &system() const { return _S; }
    const std::vector<int> &levelTable() const { return _hLevelStart; }
    // one preconditioner application z = (L L^T)^-1 r on the current factor
    // (after solve()); FP64 or the FP32 copy.  tiles selects the tile-wavefront
    // kernel, which is switched off and raises an error.
    void applyPreconditioner(const double *r, double *z, bool tiles, hipStream_t s);
    void applyPreconditionerF(const float *r, float *z, bool tiles, hipStream_t s);
    // throws when a preconditioner launch gave up waiting (waits for the stream)
    void checkError(hipStream_t s);

private:
    Grid _g{};
    PressureSys _S{};
    void *_block = nullptr;  // one allocation for everything
    double *_x = nullptr, *_r = nullptr, *_z = nullptr, *_s = nullptr;
    float *_fr = nullptr, *_fz = nullptr, *_fs = nullptr, *_fd = nullptr, *_fq = nullptr;
    double *_partial = nullptr, *_scal = nullptr;
    std::vector<int> _hLevelStart;
    MicSchedule _ms;
    int _maxTiles = 0;
    std::vector<int> _hTileRows, _hOrder;
    KernelTimers *_t = nullptr;
    KernelTimers _noTimers;
    int _levelLaunches = 0;

    double hostScalar(const double *d, hipStream_t s);
    // tile schedule of the preconditioner application for this system
    void buildSchedule(bool fp64, hipStream_t s);
    bool pcgFP64(double tolFactor, int maxIter, double resid0, int &iters, double &resid, hipStream_t s);
    bool pcgMixed(double tolFactor, int maxIter, double resid0, double innerTol, int &iters,
                  double &resid, hipStream_t s);
};

}  // namespace ff
