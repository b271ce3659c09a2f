// viscosity.h -- GPU viscosity solve of the FLIP substep (host side).
//
// ViscositySolver mirrors the reference's ViscositySolver::
// applyViscosityToVelocityField (viscositysolver.cpp:50-79): one symmetric
// positive definite system over the u, v and w faces together, built from
// the solid SDF (face states), the liquid SDF (control volumes on a twice
// finer grid) and the viscosity on the grid's nodes.  The system is
// assembled bit for bit as the reference assembles it; the solve is a
// Jacobi-preconditioned CG in float with deterministic FP64 reductions,
// stopped on the device by the reference's rule (|r|_inf <=
// min(tol |b|_inf, 1), success also at the iteration limit below an error of
// 10).
#pragma once

#include <vector>

#include "kernels.h"

namespace ff {

struct ViscosityResult {
    bool success = true;
    int iterations = 0;
    float error = 0.0f;  // |r|_inf at exit
    int unknowns = 0;    // matrix size
};

// Host copies of the assembled system for checks: face states of the u, v, w
// faces concatenated (0 air, 1 fluid, 2 solid), the seven volume grids
// (center, U, V, W, edgeU, edgeV, edgeW) concatenated, the face-to-row table
// (-1 off the matrix), the rows as CSR with ascending columns, and the
// right-hand side.
struct ViscosityHostSystem {
    std::vector<unsigned char> states;
    std::vector<float> volumes;
    std::vector<int> table, rowstart, cols;
    std::vector<float> vals, rhs;
};

class ViscositySolver {
public:
    ViscositySolver() = default;
    ~ViscositySolver();
    ViscositySolver(const ViscositySolver &) = delete;
    ViscositySolver &operator=(const ViscositySolver &) = delete;

    // Builds the system for the fields (liquid phi cell-centred, solid phi
    // and viscosity on the (n+1)^3 nodes, u/v/w the MAC field before the
    // solve) and returns its size.
    int build(Grid g, double dt, const float *liquidPhi, const float *solidNodes, const float *viscNodes,
              const float *u, const float *v, const float *w, hipStream_t s);
    // build, solve, and on success clear the field and write the solution
    // on the matrix faces; on failure, or with no matrix faces, the field is
    // left unchanged.
    ViscosityResult solve(Grid g, double dt, const float *liquidPhi, const float *solidNodes,
                          const float *viscNodes, double tol, int maxIter, float *u, float *v, float *w,
                          hipStream_t s);
    // the last built system (waits for the stream)
    void copySystem(ViscosityHostSystem &out, hipStream_t s);
    // the last solve's solution on the rows (waits for the stream)
    void copySolution(std::vector<float> &x, hipStream_t s);

private:
    Grid _g{};
    size_t _faces = 0, _nodes = 0, _subcells = 0;
    int _m = 0, _rowCap = 0;
    // grid-sized buffers (one block)
    void *_gridBlock = nullptr;
    unsigned char *_state = nullptr, *_valid = nullptr, *_valid2 = nullptr;
    float *_sub = nullptr, *_vol[7] = {};
    int *_flag = nullptr, *_rowOf = nullptr, *_table = nullptr, *_scanScratch = nullptr;
    // row-sized buffers (one block, grown on demand)
    void *_rowBlock = nullptr;
    int *_ellCol = nullptr, *_rowFace = nullptr, *_nnz = nullptr;
    float *_ellVal = nullptr, *_rhs = nullptr, *_dinv = nullptr;
    float *_x = nullptr, *_r = nullptr, *_z = nullptr, *_s = nullptr, *_q = nullptr;
    double *_partial = nullptr, *_scal = nullptr;
    double *_hostScal = nullptr;  // pinned
    void allocateGrid(Grid g);
    void allocateRows(int m);
};

}  // namespace ff
