// kernels_adjoint.hip -- adjoint (vector-Jacobian product) of the pressure
// projection u' = P(u) of PressureSolver::solve + applySolutionToVelocityField
// (pressuresolver.cpp:52-104, 624-713, 1047-1205) for a fixed liquid SDF and
// fixed face weights, where P is linear:
//
//   u' = K u + C p,   p = A^-1 B u
//
//   B  the negative divergence of the rhs (assembleKernel: per row,
//      (1/dx) * (w_left u_left - w_right u_right + ...)),
//   A  the symmetric pressure matrix,
//   C  the pressure-gradient update of applyPressureKernel (ghost-fluid
//      weights theta at the free surface, density 1),
//   K  1 on faces the projection keeps or updates, 0 on faces it zeroes
//      (no fluid on either side, or weight 0).
//
// The adjoint is  ubar = K u'bar + B^T A^-1 (C^T u'bar):  C^T is a gather over
// each cell's six faces, the solve reuses the FP64 MIC(0)-PCG (A = A^T), B^T
// is a gather over each face's two cells.  The forward rounds the pressure to
// float before applying it; that rounding is not differentiated.
#include "kernels.h"

namespace ff {

namespace {

unsigned blocksFor(size_t n, unsigned t) { return (unsigned)((n + t - 1) / t); }
constexpr int TPB = 256;

// What applyPressureKernel does to face (i, j, k) of direction DIR:
// kind 0 = kept unchanged (index 0 or n-1 along its axis), 1 = updated by the
// pressure gradient, 2 = set to zero.  For kind 1, dLo / dHi are the
// derivatives of the new face value with respect to the pressure of the
// low-side cell la and the high-side cell hb.
struct FaceRule {
    int kind;
    double dLo, dHi;
    size_t la, hb;
    bool hasLo, hasHi;  // la / hb are cells of the grid
};

template <int DIR>
__device__ FaceRule faceRule(Grid g, double dt, const float *phi, const float *wgt, int i, int j, int k) {
    const int wi = g.ni + (DIR == 0), wj = g.nj + (DIR == 1);
    const size_t f = flat(i, j, k, wi, wj);
    const int c = DIR == 0 ? i : (DIR == 1 ? j : k);
    const int nc = DIR == 0 ? g.ni : (DIR == 1 ? g.nj : g.nk);
    FaceRule r{0, 0.0, 0.0, 0, 0, c > 0, c < nc};
    const int ai = i - (DIR == 0), aj = j - (DIR == 1), ak = k - (DIR == 2);
    r.la = c > 0 ? flat(ai, aj, ak, g.ni, g.nj) : 0;
    r.hb = c < nc ? flat(i, j, k, g.ni, g.nj) : r.la;
    if (c == 0 || c == nc - 1) return r;  // kept
    const float factor = (float)(dt / g.dx);
    const float eps = 1e-6f;
    const double maxtheta = 25;
    const bool lo = c > 0 && phi[r.la] < 0.0f;
    const bool hi = c < nc && phi[r.hb] < 0.0f;
    if (!(wgt[f] > 0 && (lo || hi))) {
        r.kind = 2;
        return r;
    }
    r.kind = 1;
    if (lo && hi) {
        r.dLo = (double)factor;
        r.dHi = -(double)factor;
    } else if (lo) {
        // p2 = thetaP * p1: u' = u - factor * (thetaP - 1) * p1
        const float phi1 = phi[r.la], phi2 = phi[r.hb];
        const float thetaP = (float)fmax(-maxtheta, fmin((double)(phi2 / (phi1 + eps)), maxtheta));
        r.dLo = -(double)factor * ((double)thetaP - 1.0);
    } else {
        // p1 = thetaP * p2: u' = u - factor * (1 - thetaP) * p2
        const float phi1 = phi[r.la], phi2 = phi[r.hb];
        const float thetaP = (float)fmax(-maxtheta, fmin((double)(phi1 / (phi2 + eps)), maxtheta));
        r.dHi = -(double)factor * (1.0 - (double)thetaP);
    }
    return r;
}

// pbar[c] = (C^T u'bar)[c]: the six faces of cell c, in the order -x, +x,
// -y, +y, -z, +z
__global__ void projAdjCellKernel(Grid g, double dt, const float *phi, const float *wu, const float *wv,
                                  const float *ww, const float *gu, const float *gv, const float *gw,
                                  double *pbar) {
    const size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t n = (size_t)g.ni * g.nj * g.nk;
    if (c >= n) return;
    const int i = (int)(c % g.ni), j = (int)((c / g.ni) % g.nj), k = (int)(c / ((size_t)g.ni * g.nj));
    double acc = 0.0;
    FaceRule r;
    // -x face (this cell is its high side), +x face (low side)
    r = faceRule<0>(g, dt, phi, wu, i, j, k);
    if (r.kind == 1) acc += r.dHi * (double)gu[flat(i, j, k, g.ni + 1, g.nj)];
    r = faceRule<0>(g, dt, phi, wu, i + 1, j, k);
    if (r.kind == 1) acc += r.dLo * (double)gu[flat(i + 1, j, k, g.ni + 1, g.nj)];
    r = faceRule<1>(g, dt, phi, wv, i, j, k);
    if (r.kind == 1) acc += r.dHi * (double)gv[flat(i, j, k, g.ni, g.nj + 1)];
    r = faceRule<1>(g, dt, phi, wv, i, j + 1, k);
    if (r.kind == 1) acc += r.dLo * (double)gv[flat(i, j + 1, k, g.ni, g.nj + 1)];
    r = faceRule<2>(g, dt, phi, ww, i, j, k);
    if (r.kind == 1) acc += r.dHi * (double)gw[flat(i, j, k, g.ni, g.nj)];
    r = faceRule<2>(g, dt, phi, ww, i, j, k + 1);
    if (r.kind == 1) acc += r.dLo * (double)gw[flat(i, j, k + 1, g.ni, g.nj)];
    pbar[c] = acc;
}

// ubar[f] = K u'bar[f] + (B^T lambda)[f], lambda on cells (0 off the rows):
// B^T at face f is (w_f / dx) (lambda[hb] - lambda[la])
template <int DIR>
__global__ void projAdjFaceKernel(Grid g, double dt, const float *phi, const float *wgt, const float *gin,
                                  const double *lambda, float *gout) {
    const int wi = g.ni + (DIR == 0), wj = g.nj + (DIR == 1), wk = g.nk + (DIR == 2);
    const size_t f = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= (size_t)wi * wj * wk) return;
    const int i = (int)(f % wi), j = (int)((f / wi) % wj), k = (int)(f / ((size_t)wi * wj));
    const FaceRule r = faceRule<DIR>(g, dt, phi, wgt, i, j, k);
    double v = r.kind == 2 ? 0.0 : (double)gin[f];
    const double lHi = r.hasHi ? lambda[r.hb] : 0.0, lLo = r.hasLo ? lambda[r.la] : 0.0;
    v += (1.0 / g.dx) * (double)wgt[f] * (lHi - lLo);
    gout[f] = (float)v;
}

}  // namespace

void projectionAdjointCells(Grid g, double dt, const float *phi, const float *wu, const float *wv, const float *ww,
                            const float *gu, const float *gv, const float *gw, double *pbar, hipStream_t s) {
    const size_t n = (size_t)g.ni * g.nj * g.nk;
    hipLaunchKernelGGL(projAdjCellKernel, dim3(blocksFor(n, TPB)), dim3(TPB), 0, s, g, dt, phi, wu, wv, ww, gu, gv,
                       gw, pbar);
}

void projectionAdjointFaces(Grid g, double dt, const float *phi, const float *wu, const float *wv, const float *ww,
                            const float *gu, const float *gv, const float *gw, const double *lambda, float *ou,
                            float *ov, float *ow, hipStream_t s) {
    const size_t nu = (size_t)(g.ni + 1) * g.nj * g.nk, nv = (size_t)g.ni * (g.nj + 1) * g.nk,
                 nw = (size_t)g.ni * g.nj * (g.nk + 1);
    hipLaunchKernelGGL(projAdjFaceKernel<0>, dim3(blocksFor(nu, TPB)), dim3(TPB), 0, s, g, dt, phi, wu, gu, lambda, ou);
    hipLaunchKernelGGL(projAdjFaceKernel<1>, dim3(blocksFor(nv, TPB)), dim3(TPB), 0, s, g, dt, phi, wv, gv, lambda, ov);
    hipLaunchKernelGGL(projAdjFaceKernel<2>, dim3(blocksFor(nw, TPB)), dim3(TPB), 0, s, g, dt, phi, ww, gw, lambda, ow);
}

}  // namespace ff
