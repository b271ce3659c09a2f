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
//
// Adjoints of the transfers, for fixed particle positions:
// - P2G (velocityadvector.cpp:467-536): u_f = S_f / W_f (S_f when W_f <= 1e-6)
//   with S_f = sum w_pf v_p and W_f = sum w_pf, so vbar_p = sum over the faces
//   within r of p of (w_pf / W_f) ubar_f: one thread per particle gathers over
//   the 18 faces per direction whose candidate cells hold the particle.
// - Trilinear sampling of the MAC field (macvelocityfield.cpp:519-645):
//   v_p = sum of the 8 corner weights times the faces, so ubar_f = sum over
//   the points whose stencil holds f of the corner weight times vbar_p.  The
//   points are bucketed by stencil base per component (a counting sort,
//   each bucket in ascending point index), and one thread per face gathers
//   over its 8 candidate buckets: deterministic, no atomics on floats.
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

// ---- transfer adjoints ----
namespace {

// P2G adjoint of particle q: vbar (the velocity cotangent) and, when vel is
// given, xbar (the position cotangent): with u_f = S_f / W_f,
// du_f/dx_p = (dw_pf/dx_p) (v_p - u_f) / W_f  (S_f alone when W_f <= 1e-6),
// dw/dx_e = w'(d2) * (-2) (gpos_e - p'_e) in the face's frame.  The kernel
// weight and its slope vanish at d2 = r^2, so the derivative is continuous
// as particles enter or leave a face's support.
__global__ void p2gAdjKernel(Grid g, int np, const float *px, const float *py, const float *pz, const float *vx,
                             const float *vy, const float *vz, double radius, const float *gu, const float *gv,
                             const float *gw, const float *wsu, const float *wsv, const float *wsw, const float *fu,
                             const float *fv, const float *fw, float *vbx, float *vby, float *vbz, float *xbar) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= np) return;
    const float x = px[q], y = py[q], z = pz[q];
    const I3 c = posToIndex(V3{x, y, z}, g.dx);
    double acc[3] = {0.0, 0.0, 0.0}, xb[3] = {0.0, 0.0, 0.0};
    if (inRange(c.i, c.j, c.k, g.ni, g.nj, g.nk)) {
        const P2GWeight W = p2gWeight(radius);
        const float eps = 1e-6f;
#pragma unroll
        for (int dir = 0; dir < 3; dir++) {
            const int wi = g.ni + (dir == 0), wj = g.nj + (dir == 1), wk = g.nk + (dir == 2);
            const float *gb = dir == 0 ? gu : (dir == 1 ? gv : gw);
            const float *ws = dir == 0 ? wsu : (dir == 1 ? wsv : wsw);
            const float *fo = dir == 0 ? fu : (dir == 1 ? fv : fw);
            const double vd = vx ? (double)(dir == 0 ? vx[q] : (dir == 1 ? vy[q] : vz[q])) : 0.0;
            // faces whose candidate cells (2 along the normal, 3 across) hold cell c
            const int i0 = dir == 0 ? c.i : c.i - 1, i1 = c.i + 1;
            const int j0 = dir == 1 ? c.j : c.j - 1, j1 = c.j + 1;
            const int k0 = dir == 2 ? c.k : c.k - 1, k1 = c.k + 1;
            for (int fk = k0; fk <= k1; fk++)
                for (int fj = j0; fj <= j1; fj++)
                    for (int fi = i0; fi <= i1; fi++) {
                        if (!inRange(fi, fj, fk, wi, wj, wk)) continue;
                        const FaceFrame F = p2gFrame(g, dir, fi, fj, fk);
                        const float d2 = p2gDist2(F, x, y, z);
                        if (!(d2 < W.rsq)) continue;
                        const size_t f = flat(fi, fj, fk, wi, wj);
                        const double w = (double)p2gWeightAt(W, d2);
                        const double wsum = (double)ws[f];
                        const bool norm = ws[f] > eps;
                        acc[dir] += (norm ? w / wsum : w) * (double)gb[f];
                        if (xbar) {
                            const double dd = (double)d2;
                            const double dw = -3.0 * (double)W.coef1 * dd * dd + 2.0 * (double)W.coef2 * dd -
                                              (double)W.coef3;
                            const double s = (norm ? (vd - (double)fo[f]) / wsum : vd) * (double)gb[f] * dw * -2.0;
                            const V3 pl = sub(sub(V3{x, y, z}, F.offd), F.offb);
                            xb[0] += s * ((double)F.gpos.x - (double)pl.x);
                            xb[1] += s * ((double)F.gpos.y - (double)pl.y);
                            xb[2] += s * ((double)F.gpos.z - (double)pl.z);
                        }
                    }
        }
    }
    vbx[q] = (float)acc[0];
    vby[q] = (float)acc[1];
    vbz[q] = (float)acc[2];
    if (xbar)
        for (int e = 0; e < 3; e++) xbar[3 * (size_t)q + e] = (float)xb[e];
}

// stencil of point q for component dir of the MAC field (interpFace after the
// in-domain test): base corner (i, j, k) and fractions; false outside
__device__ bool sampleStencil(Grid g, int dir, float px, float py, float pz, int &i, int &j, int &k, double &ix,
                              double &iy, double &iz) {
    double x = px, y = py, z = pz;
    if (!(x >= 0 && y >= 0 && z >= 0 && x < g.dx * g.ni && y < g.dx * g.nj && z < g.dx * g.nk)) return false;
    const double h = 0.5 * g.dx;
    if (dir != 0) x -= h;
    if (dir != 1) y -= h;
    if (dir != 2) z -= h;
    const double invdx = 1.0 / g.dx;
    i = (int)floor(x * invdx);
    j = (int)floor(y * invdx);
    k = (int)floor(z * invdx);
    const double gx = (double)i * g.dx, gy = (double)j * g.dx, gz = (double)k * g.dx;
    const double inv_dx = 1 / g.dx;
    ix = (x - gx) * inv_dx;
    iy = (y - gy) * inv_dx;
    iz = (z - gz) * inv_dx;
    return true;
}

// bucket of point q: its stencil base shifted by one (the base can be -1) in
// a (wi+1) x (wj+1) x (wk+1) grid, or -1 outside the domain
__global__ void sampleKeyKernel(Grid g, int dir, int np, const float *px, const float *py, const float *pz, int *key,
                                int *slot, int *counts) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= np) return;
    const int wi = g.ni + (dir == 0), wj = g.nj + (dir == 1);
    int i, j, k;
    double ix, iy, iz;
    int b = -1;
    if (sampleStencil(g, dir, px[q], py[q], pz[q], i, j, k, ix, iy, iz))
        b = (int)flat(i + 1, j + 1, k + 1, wi + 1, wj + 1);
    key[q] = b;
    if (b >= 0) slot[q] = atomicAdd(&counts[b], 1);
}

__global__ void bucketScatterKernel(int np, const int *key, const int *slot, const int *offsets, int *perm) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= np) return;
    if (key[q] >= 0) perm[offsets[key[q]] + slot[q]] = q;
}

// each bucket's points in ascending index
__global__ void bucketSortKernel(const int *offsets, size_t nb, int *perm) {
    const size_t b = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nb) return;
    const int lo = offsets[b], hi = offsets[b + 1];
    for (int a = lo + 1; a < hi; a++) {
        const int v = perm[a];
        int q = a - 1;
        while (q >= lo && perm[q] > v) {
            perm[q + 1] = perm[q];
            q--;
        }
        perm[q + 1] = v;
    }
}

// sampling adjoint: ubar[f] for component dir, over the 8 buckets whose
// stencil holds f (corner (fi - bi, fj - bj, fk - bk) of the base)
__global__ void sampleAdjFaceKernel(Grid g, int dir, const float *px, const float *py, const float *pz,
                                    const float *vbar, const int *offsets, const int *perm, float *out) {
    const int wi = g.ni + (dir == 0), wj = g.nj + (dir == 1), wk = g.nk + (dir == 2);
    const size_t f = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= (size_t)wi * wj * wk) return;
    const int fi = (int)(f % wi), fj = (int)((f / wi) % wj), fk = (int)(f / ((size_t)wi * wj));
    double acc = 0.0;
    for (int oz = 0; oz < 2; oz++)
        for (int oy = 0; oy < 2; oy++)
            for (int ox = 0; ox < 2; ox++) {
                const int bi = fi - ox, bj = fj - oy, bk = fk - oz;  // >= -1
                const size_t b = flat(bi + 1, bj + 1, bk + 1, wi + 1, wj + 1);
                for (int s = offsets[b]; s < offsets[b + 1]; s++) {
                    const int q = perm[s];
                    int i, j, k;
                    double ix, iy, iz;
                    sampleStencil(g, dir, px[q], py[q], pz[q], i, j, k, ix, iy, iz);
                    const double w = (ox ? ix : 1 - ix) * (oy ? iy : 1 - iy) * (oz ? iz : 1 - iz);
                    acc += w * (double)vbar[3 * (size_t)q + dir];
                }
            }
    out[f] = (float)acc;
}

}  // namespace

void p2gAdjoint(Grid g, int np, const float *px, const float *py, const float *pz, const float *vx, const float *vy,
                const float *vz, double radius, const float *gu, const float *gv, const float *gw, const float *wsu,
                const float *wsv, const float *wsw, const float *fu, const float *fv, const float *fw, float *vbx,
                float *vby, float *vbz, float *xbar, hipStream_t s) {
    if (np == 0) return;
    hipLaunchKernelGGL(p2gAdjKernel, dim3(blocksFor(np, TPB)), dim3(TPB), 0, s, g, np, px, py, pz, vx, vy, vz, radius,
                       gu, gv, gw, wsu, wsv, wsw, fu, fv, fw, vbx, vby, vbz, xbar);
}

size_t sampleBuckets(Grid g, int dir) {
    return (size_t)(g.ni + 1 + (dir == 0)) * (g.nj + 1 + (dir == 1)) * (g.nk + 1 + (dir == 2));
}

void sampleAdjoint(Grid g, int dir, int np, const float *px, const float *py, const float *pz, const float *vbar,
                   int *key, int *slot, int *counts, int *offsets, int *perm, int *scanScratch, float *out,
                   hipStream_t s) {
    const size_t nb = sampleBuckets(g, dir);
    hipCheck(hipMemsetAsync(counts, 0, sizeof(int) * (nb + 1), s), "memset");
    if (np > 0) {
        hipLaunchKernelGGL(sampleKeyKernel, dim3(blocksFor(np, TPB)), dim3(TPB), 0, s, g, dir, np, px, py, pz, key,
                           slot, counts);
    }
    exclusiveScan(counts, offsets, nb, scanScratch, s);
    if (np > 0)
        hipLaunchKernelGGL(bucketScatterKernel, dim3(blocksFor(np, TPB)), dim3(TPB), 0, s, np, key, slot, offsets,
                           perm);
    hipLaunchKernelGGL(bucketSortKernel, dim3(blocksFor(nb, TPB)), dim3(TPB), 0, s, offsets, nb, perm);
    const size_t nf = (size_t)(g.ni + (dir == 0)) * (g.nj + (dir == 1)) * (g.nk + (dir == 2));
    hipLaunchKernelGGL(sampleAdjFaceKernel, dim3(blocksFor(nf, TPB)), dim3(TPB), 0, s, g, dir, px, py, pz, vbar,
                       offsets, perm, out);
}

}  // namespace ff
