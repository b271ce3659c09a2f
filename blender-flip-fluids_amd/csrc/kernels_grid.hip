// kernels_grid.hip -- face-array stencils of the FLIP substep: layered
// velocity extrapolation, body force, solid constraint and the pressure
// gradient update.
#include "kernels.h"

namespace ff {

namespace {

unsigned blocksFor(size_t n, unsigned t) { return (unsigned)((n + t - 1) / t); }
constexpr int TPB = 256;
constexpr uint8_t ST_BORDER = 254, ST_UNKNOWN = 255;

// GridUtils::extrapolateGrid (gridutils.h:95-160).  The status byte holds the
// layer at which a face became known: 0 for valid faces, L+1 for faces filled
// in layer L, ST_BORDER for the array border (always DONE, never a source of
// new work) and ST_UNKNOWN.  In layer L the reference turns the KNOWN front
// (status L) into DONE and fills every UNKNOWN face next to it with the mean
// of its DONE neighbours in the order +i,-i,+j,-j,+k,-k (gridutils.h:43-92,
// gridutils.cpp:120-175).  Faces written in this layer get L+1, which neither
// the front test (== L) nor the DONE test (<= L) accepts, so one in-place pass
// per layer is race free.
__global__ void extrapInitKernel(int wi, int wj, int wk, const uint8_t *valid, uint8_t *st) {
    size_t f = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t n = (size_t)wi * wj * wk;
    if (f >= n) return;
    int i = (int)(f % wi), j = (int)((f / wi) % wj), k = (int)(f / ((size_t)wi * wj));
    uint8_t s;
    if (i == 0 || j == 0 || k == 0 || i == wi - 1 || j == wj - 1 || k == wk - 1)
        s = ST_BORDER;
    else
        s = valid[f] ? 0 : ST_UNKNOWN;
    st[f] = s;
}

__global__ void extrapLayerKernel(int wi, int wj, int wk, float *a, uint8_t *st, int L) {
    size_t f = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t n = (size_t)wi * wj * wk;
    if (f >= n || st[f] != ST_UNKNOWN) return;
    const long long nb[6] = {1, -1, wi, -(long long)wi, (long long)wi * wj, -(long long)wi * wj};
    uint8_t s[6];
    bool front = false;
#pragma unroll
    for (int q = 0; q < 6; q++) {
        s[q] = st[(long long)f + nb[q]];
        front |= s[q] == (uint8_t)L;
    }
    if (!front) return;
    float sum = 0.0f;
    int count = 0;
#pragma unroll
    for (int q = 0; q < 6; q++)
        if (s[q] == ST_BORDER || s[q] <= (uint8_t)L) {
            sum += a[(long long)f + nb[q]];
            count++;
        }
    a[f] = sum / (float)count;
    st[f] = (uint8_t)(L + 1);
}

// Adjoint of layer L of extrapLayerKernel, given the status bytes the
// forward left.  Face f filled in layer L (status L+1) holds the mean of
// its neighbours that were DONE then (status <= L or the border), so each
// such neighbour g gathers bar[f] / count_f from its filled neighbours.
// Writes go to faces with status <= L or the border and reads to faces with
// status L+1, so one pass per layer is race free.
__global__ void extrapAdjLayerKernel(int wi, int wj, int wk, float *bar, const uint8_t *st, int L) {
    size_t g = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t n = (size_t)wi * wj * wk;
    if (g >= n) return;
    const uint8_t sg = st[g];
    if (!(sg == ST_BORDER || sg <= (uint8_t)L)) return;
    const int i = (int)(g % wi), j = (int)((g / wi) % wj), k = (int)(g / ((size_t)wi * wj));
    const long long nb[6] = {1, -1, wi, -(long long)wi, (long long)wi * wj, -(long long)wi * wj};
    const bool in[6] = {i + 1 < wi, i > 0, j + 1 < wj, j > 0, k + 1 < wk, k > 0};
    double acc = 0.0;
    for (int q = 0; q < 6; q++) {
        if (!in[q]) continue;
        const long long f = (long long)g + nb[q];
        if (st[f] != (uint8_t)(L + 1)) continue;
        // a filled face is never on the border, so its six neighbours exist
        int count = 0;
        for (int r = 0; r < 6; r++) {
            const uint8_t s = st[f + nb[r]];
            count += (s == ST_BORDER || s <= (uint8_t)L) ? 1 : 0;
        }
        acc += (double)(bar[f] / (float)count);
    }
    if (acc != 0.0) bar[g] = (float)((double)bar[g] + acc);
}

// after the layers: the filled faces' own input values were overwritten
__global__ void extrapAdjClearKernel(size_t n, float *bar, const uint8_t *st) {
    size_t f = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= n) return;
    const uint8_t s = st[f];
    if (s != 0 && s != ST_BORDER && s != ST_UNKNOWN) bar[f] = 0.0f;
}

// _applyConstantBodyForces (fluidsimulation.cpp:5837-5876): no inflow
// constraints on this path, so every face gets the increment
__global__ void addKernel(float *a, size_t n, float inc) {
    size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q < n) a[q] += inc;
}

// _constrainVelocityFieldThread (fluidsimulation.cpp:6388-6442): domain
// friction 0 and zero face velocity of static solids
__global__ void constrainKernel(float *a, const float *w, size_t n, const float *sv) {
    size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    const float f = 0.0f, sface = sv ? sv[q] : 0.0f;
    float wt = w[q];
    if (wt == 0)
        a[q] = sface;
    else if (wt < 1.0f)
        a[q] = f * sface + (1.0f - f) * a[q];
}

// _constrainVelocityFieldThread with friction (fluidsimulation.cpp:6292-6442):
// on a partly open face f is 0.25 x the sum, in the reference's order, of the
// friction of the closest mesh object at the face's four nodes (read live
// from the objects' table, none counting 0); the solid's face velocity sv, or
// 0 for static solids (sv null).
template <int DIR>
__global__ void constrainFrictionKernel(Grid g, float *a, const float *w, const int *obj, const float *fr,
                                        const float *sv) {
    const int wi = g.ni + (DIR == 0), wj = g.nj + (DIR == 1), wk = g.nk + (DIR == 2);
    const size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= (size_t)wi * wj * wk) return;
    const float sface = sv ? sv[q] : 0.0f;
    const float wt = w[q];
    if (wt == 0) {
        a[q] = sface;
    } else if (wt < 1.0f) {
        const int i = (int)(q % wi), j = (int)((q / wi) % wj), k = (int)(q / ((size_t)wi * wj));
        const int pi = g.ni + 1, pj = g.nj + 1;
        // the two node offsets spanning the face: u (j, k), v (i, k), w (i, j)
        const int ai = DIR == 0 ? 0 : 1, aj = DIR == 0 ? 1 : 0, ak = 0;
        const int bi = 0, bj = DIR == 2 ? 1 : 0, bk = DIR == 2 ? 0 : 1;
        auto of = [&](int x, int y, int z) {
            const int m = obj[flat(x, y, z, pi, pj)];
            return m < 0 ? 0.0f : fr[m];
        };
        float friction = 0.0f;
        friction += of(i, j, k);
        friction += of(i + ai, j + aj, k + ak);
        friction += of(i + bi, j + bj, k + bk);
        friction += of(i + ai + bi, j + aj + bj, k + ak + bk);
        const float f = 0.25f * friction;
        a[q] = f * sface + (1.0f - f) * a[q];
    }
}

// PressureSolver::_applyPressureToVelocityFieldThread (pressuresolver.cpp:1047-1205),
// surface tension off and density 1; faces with index 0 or n-1 along their
// axis are skipped (1058, 1109, 1159) and keep their value.
template <int DIR>
__global__ void applyPressureKernel(Grid g, double dt, const float *p, const float *phi,
                                    const float *wgt, float *a, uint8_t *valid, TensionView tv) {
    const int wi = g.ni + (DIR == 0), wj = g.nj + (DIR == 1), wk = g.nk + (DIR == 2);
    size_t f = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t n = (size_t)wi * wj * wk;
    if (f >= n) return;
    int i = (int)(f % wi), j = (int)((f / wi) % wj), k = (int)(f / ((size_t)wi * wj));
    const int c = DIR == 0 ? i : (DIR == 1 ? j : k);
    const int nc = DIR == 0 ? g.ni : (DIR == 1 ? g.nj : g.nk);
    valid[f] = 0;
    if (c == 0 || c == nc - 1) return;
    const float factor = (float)(dt / g.dx);
    const float eps = 1e-6f;
    const double maxtheta = 25;
    int ai = i - (DIR == 0), aj = j - (DIR == 1), ak = k - (DIR == 2);  // low-side cell
    // isFaceBorderingMaterial (fluidmaterialgrid.cpp:126-146)
    // The reference reads the cell past the last face (c == nc) when that face
    // is open; with the domain walls it never is.  Reads stay in bounds here by
    // using the low-side cell in its place.
    const size_t la = flat(ai, aj, ak, g.ni, g.nj);
    const size_t hb = c < nc ? flat(i, j, k, g.ni, g.nj) : la;
    bool lo = c > 0 && phi[la] < 0.0f;
    bool hi = c < nc && phi[hb] < 0.0f;
    if (wgt[f] > 0 && (lo || hi)) {
        float p1 = 0.0f, p2 = 0.0f;
        if (lo && hi) {
            p1 = p[la];
            p2 = p[hb];
        } else {
            float phi1 = phi[la], phi2 = phi[hb];
            // the surface tension term of the pair (pressuresolver.cpp:1076)
            const double tension = tv.status && c < nc ? tensionTerm(g, tv, ai, aj, ak, i, j, k) : 0.0;
            if (lo) {
                float thetaP = (float)fmax(-maxtheta, fmin((double)(phi2 / (phi1 + eps)), maxtheta));
                float thetaT = (float)fmax(-maxtheta, fmin((double)((phi1 - phi2) / (phi1 + eps)), maxtheta));
                p1 = p[la];
                p2 = (float)((double)thetaT * tension + (double)(thetaP * p1));
            } else {
                float thetaP = (float)fmax(-maxtheta, fmin((double)(phi1 / (phi2 + eps)), maxtheta));
                float thetaT = (float)fmax(-maxtheta, fmin((double)((phi2 - phi1) / (phi2 + eps)), maxtheta));
                p2 = p[hb];
                p1 = (float)((double)thetaT * tension + (double)(thetaP * p2));
            }
        }
        float rho = (1.0f + 1.0f) / 2.0f;
        a[f] += -factor * (p2 - p1) / rho;
        valid[f] = 1;
    } else {
        a[f] = 0.0f;
    }
}

}  // namespace

void extrapolateInit(int wi, int wj, int wk, const uint8_t *valid, uint8_t *st, hipStream_t s) {
    size_t n = (size_t)wi * wj * wk;
    hipLaunchKernelGGL(extrapInitKernel, dim3(blocksFor(n, TPB)), dim3(TPB), 0, s, wi, wj, wk, valid, st);
}

void extrapolateLayer(int wi, int wj, int wk, float *a, uint8_t *st, int layer, hipStream_t s) {
    size_t n = (size_t)wi * wj * wk;
    hipLaunchKernelGGL(extrapLayerKernel, dim3(blocksFor(n, TPB)), dim3(TPB), 0, s, wi, wj, wk, a, st, layer);
}

void extrapolateAdjoint(int wi, int wj, int wk, float *bar, const uint8_t *st, int layers, hipStream_t s) {
    size_t n = (size_t)wi * wj * wk;
    for (int L = layers - 1; L >= 0; L--)
        hipLaunchKernelGGL(extrapAdjLayerKernel, dim3(blocksFor(n, TPB)), dim3(TPB), 0, s, wi, wj, wk, bar, st, L);
    hipLaunchKernelGGL(extrapAdjClearKernel, dim3(blocksFor(n, TPB)), dim3(TPB), 0, s, n, bar, st);
}

void addToFaces(float *a, size_t n, float inc, hipStream_t s) {
    hipLaunchKernelGGL(addKernel, dim3(blocksFor(n, TPB)), dim3(TPB), 0, s, a, n, inc);
}

void constrainFaces(float *a, const float *w, size_t n, hipStream_t s, const float *sv) {
    hipLaunchKernelGGL(constrainKernel, dim3(blocksFor(n, TPB)), dim3(TPB), 0, s, a, w, n, sv);
}

void constrainFacesFriction(Grid g, int dir, float *a, const float *w, const int *obj, const float *fr,
                            hipStream_t s, const float *sv) {
    const size_t n = (size_t)(g.ni + (dir == 0)) * (g.nj + (dir == 1)) * (g.nk + (dir == 2));
    if (dir == 0)
        hipLaunchKernelGGL(constrainFrictionKernel<0>, dim3(blocksFor(n, TPB)), dim3(TPB), 0, s, g, a, w, obj, fr,
                           sv);
    else if (dir == 1)
        hipLaunchKernelGGL(constrainFrictionKernel<1>, dim3(blocksFor(n, TPB)), dim3(TPB), 0, s, g, a, w, obj, fr,
                           sv);
    else
        hipLaunchKernelGGL(constrainFrictionKernel<2>, dim3(blocksFor(n, TPB)), dim3(TPB), 0, s, g, a, w, obj, fr,
                           sv);
}

void applyPressure(Grid g, double dt, const float *p, const float *phi, const float *wu,
                   const float *wv, const float *ww, float *u, float *v, float *w, uint8_t *vu,
                   uint8_t *vv, uint8_t *vw, hipStream_t s, TensionView tension) {
    size_t nu = (size_t)(g.ni + 1) * g.nj * g.nk, nv = (size_t)g.ni * (g.nj + 1) * g.nk,
           nw = (size_t)g.ni * g.nj * (g.nk + 1);
    hipLaunchKernelGGL(applyPressureKernel<0>, dim3(blocksFor(nu, TPB)), dim3(TPB), 0, s, g, dt, p, phi, wu, u, vu, tension);
    hipLaunchKernelGGL(applyPressureKernel<1>, dim3(blocksFor(nv, TPB)), dim3(TPB), 0, s, g, dt, p, phi, wv, v, vv, tension);
    hipLaunchKernelGGL(applyPressureKernel<2>, dim3(blocksFor(nw, TPB)), dim3(TPB), 0, s, g, dt, p, phi, ww, w, vw, tension);
}

}  // namespace ff
