// kernels_setup.hip -- static solid data, built once per solid configuration:
// the domain-boundary solid SDF at the nodes, the face-open weights and the
// coarse near-solid grid used by particle collision.
#include "kernels.h"

namespace ff {

namespace {

struct BoxMesh {
    V3 v[8];
    int t[12][3];
    int gmin[12][3], gmax[12][3];
    Box box;
    float ub;
};

// One thread per node: min over the box triangles whose node band holds the
// node, in 10^3-node block-local coordinates, then the inside/outside sign and
// the final negation (meshlevelset.cpp:781-1163, 1298-1318, 702-708;
// fluidsimulation.cpp:5277-5287).
__global__ void solidBoxPhiKernel(Grid g, BoxMesh M, float *phi) {
    const int pi = g.ni + 1, pj = g.nj + 1, pk = g.nk + 1;
    size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t nn = (size_t)pi * pj * pk;
    if (idx >= nn) return;
    int i = (int)(idx % pi), j = (int)((idx / pi) % pj), k = (int)(idx / ((size_t)pi * pj));
    const int bw = 10;
    int bi = i / bw, bj = j / bw, bk = k / bw;
    V3 off = indexToPos(bi, bj, bk, bw * g.dx);
    V3 gpos = indexToPos(i - bi * bw, j - bj * bw, k - bk * bw, g.dx);
    float best = M.ub;
    for (int t = 0; t < 12; t++) {
        if (i < M.gmin[t][0] || i > M.gmax[t][0] || j < M.gmin[t][1] || j > M.gmax[t][1] ||
            k < M.gmin[t][2] || k > M.gmax[t][2])
            continue;
        V3 a = sub(M.v[M.t[t][0]], off), b = sub(M.v[M.t[t][1]], off), c = sub(M.v[M.t[t][2]], off);
        V3 cp = closestPointOnTriangle(gpos, a, b, c);
        float dist = len3(sub(cp, gpos));
        if (dist < best) best = dist;
    }
    bool in = M.box.inside(indexToPos(i, j, k, g.dx));
    float v = in ? -best : best;
    phi[idx] = -v;
}

__device__ float clampw(float v) { return fmaxf(0.0f, fminf(v, 1.0f)); }

// FluidSimulation::_updateWeightGridThread (fluidsimulation.cpp:6110-6150) with
// MeshLevelSet::getFaceWeightU/V/W (meshlevelset.cpp:357-391).
__global__ void faceWeightsKernel(Grid g, const float *sp, float *wu, float *wv, float *ww) {
    const int ni = g.ni, nj = g.nj, nk = g.nk, pi = ni + 1, pj = nj + 1;
    size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t nu = (size_t)(ni + 1) * nj * nk, nv = (size_t)ni * (nj + 1) * nk, nw = (size_t)ni * nj * (nk + 1);
#define P(a, b, c) sp[flat(a, b, c, pi, pj)]
    if (idx < nu) {
        int i = (int)(idx % (ni + 1)), j = (int)((idx / (ni + 1)) % nj), k = (int)(idx / ((size_t)(ni + 1) * nj));
        wu[idx] = clampw(1.0f - fraction4(P(i, j, k), P(i, j + 1, k), P(i, j, k + 1), P(i, j + 1, k + 1)));
    } else if (idx < nu + nv) {
        size_t q = idx - nu;
        int i = (int)(q % ni), j = (int)((q / ni) % (nj + 1)), k = (int)(q / ((size_t)ni * (nj + 1)));
        wv[q] = clampw(1.0f - fraction4(P(i, j, k), P(i, j, k + 1), P(i + 1, j, k), P(i + 1, j, k + 1)));
    } else if (idx < nu + nv + nw) {
        size_t q = idx - nu - nv;
        int i = (int)(q % ni), j = (int)((q / ni) % nj), k = (int)(q / ((size_t)ni * nj));
        ww[q] = clampw(1.0f - fraction4(P(i, j, k), P(i, j + 1, k), P(i + 1, j, k), P(i + 1, j + 1, k)));
    }
#undef P
}

// FluidSimulation::_initializeNearSolidGridThread (fluidsimulation.cpp:5436-5446).
// Writes 1 into a zeroed coarse grid; racing writers all store the same value.
__global__ void nearSolidMark(Grid g, const float *sp, uint8_t *near, int ci, int cj) {
    size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t n = (size_t)g.ni * g.nj * g.nk;
    if (idx >= n) return;
    int i = (int)(idx % g.ni), j = (int)((idx / g.ni) % g.nj), k = (int)(idx / ((size_t)g.ni * g.nj));
    float maxd = (float)(3 * g.dx);
    if (fabsf(sp[flat(i, j, k, g.ni + 1, g.nj + 1)]) < maxd) near[flat(i / 3, j / 3, k / 3, ci, cj)] = 1;
}

// GridUtils::featherGrid6 (gridutils.cpp:229-262): out = in dilated by the 6-neighbourhood
__global__ void feather6Kernel(const uint8_t *in, uint8_t *out, int ci, int cj, int ck) {
    size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t n = (size_t)ci * cj * ck;
    if (idx >= n) return;
    int i = (int)(idx % ci), j = (int)((idx / ci) % cj), k = (int)(idx / ((size_t)ci * cj));
    uint8_t v = in[idx];
    if (i > 0) v |= in[idx - 1];
    if (i < ci - 1) v |= in[idx + 1];
    if (j > 0) v |= in[idx - ci];
    if (j < cj - 1) v |= in[idx + ci];
    if (k > 0) v |= in[idx - (size_t)ci * cj];
    if (k < ck - 1) v |= in[idx + (size_t)ci * cj];
    out[idx] = v;
}

unsigned blocksFor(size_t n, unsigned t) { return (unsigned)((n + t - 1) / t); }

}  // namespace

// Host part of the box mesh: vertices, triangles and node bands
// (fluidsimulation.cpp:5148-5186; meshlevelset.cpp:878-918; aabb.cpp:84-115).
void solidBoxPhi(Grid g, float *phi, hipStream_t s) {
    BoxMesh M;
    M.box = boundaryBox(g.ni, g.nj, g.nk, g.dx);
    Box &b = M.box;
    V3 p = b.pos;
    V3 verts[8] = {
        mk(p.x, p.y, p.z),
        mk((double)p.x + b.w, p.y, p.z),
        mk((double)p.x + b.w, p.y, (double)p.z + b.d),
        mk(p.x, p.y, (double)p.z + b.d),
        mk(p.x, (double)p.y + b.h, p.z),
        mk((double)p.x + b.w, (double)p.y + b.h, p.z),
        mk((double)p.x + b.w, (double)p.y + b.h, (double)p.z + b.d),
        mk(p.x, (double)p.y + b.h, (double)p.z + b.d),
    };
    const int tris[12][3] = {{0, 1, 2}, {0, 2, 3}, {4, 7, 6}, {4, 6, 5}, {0, 3, 7}, {0, 7, 4},
                             {1, 5, 6}, {1, 6, 2}, {0, 4, 5}, {0, 5, 1}, {3, 2, 6}, {3, 6, 7}};
    for (int q = 0; q < 8; q++) M.v[q] = verts[q];
    const int band = 3;
    for (int t = 0; t < 12; t++) {
        for (int c = 0; c < 3; c++) M.t[t][c] = tris[t][c];
        V3 a = verts[tris[t][0]];
        double mnx = a.x, mny = a.y, mnz = a.z, mxx = a.x, mxy = a.y, mxz = a.z;
        for (int q = 0; q < 3; q++) {
            V3 v = verts[tris[t][q]];
            mnx = fmin((double)v.x, mnx); mny = fmin((double)v.y, mny); mnz = fmin((double)v.z, mnz);
            mxx = fmax((double)v.x, mxx); mxy = fmax((double)v.y, mxy); mxz = fmax((double)v.z, mxz);
        }
        Box tb{mk(mnx, mny, mnz), mxx - mnx, mxy - mny, mxz - mnz};
        I3 g0 = posToIndex(tb.pos, g.dx), g1 = posToIndex(tb.maxPoint(), g.dx);
        g0.i -= band; g0.j -= band; g0.k -= band;
        g1.i += band + 1; g1.j += band + 1; g1.k += band + 1;
        bool skip = g1.i < 0 || g0.i >= g.ni + 1 || g1.j < 0 || g0.j >= g.nj + 1 || g1.k < 0 ||
                    g0.k >= g.nk + 1;
        if (skip) {  // empty band
            g0 = I3{1, 1, 1};
            g1 = I3{0, 0, 0};
        } else {
            g0.i = g0.i < 0 ? 0 : g0.i; g0.j = g0.j < 0 ? 0 : g0.j; g0.k = g0.k < 0 ? 0 : g0.k;
            g1.i = g1.i > g.ni ? g.ni : g1.i; g1.j = g1.j > g.nj ? g.nj : g1.j; g1.k = g1.k > g.nk ? g.nk : g1.k;
        }
        M.gmin[t][0] = g0.i; M.gmin[t][1] = g0.j; M.gmin[t][2] = g0.k;
        M.gmax[t][0] = g1.i; M.gmax[t][1] = g1.j; M.gmax[t][2] = g1.k;
    }
    M.ub = (float)((double)((g.ni + 1) + (g.nj + 1) + (g.nk + 1)) * g.dx);
    size_t nn = (size_t)(g.ni + 1) * (g.nj + 1) * (g.nk + 1);
    hipLaunchKernelGGL(solidBoxPhiKernel, dim3(blocksFor(nn, 256)), dim3(256), 0, s, g, M, phi);
}

void faceWeights(Grid g, const float *sp, float *wu, float *wv, float *ww, hipStream_t s) {
    size_t n = (size_t)(g.ni + 1) * g.nj * g.nk + (size_t)g.ni * (g.nj + 1) * g.nk +
               (size_t)g.ni * g.nj * (g.nk + 1);
    hipLaunchKernelGGL(faceWeightsKernel, dim3(blocksFor(n, 256)), dim3(256), 0, s, g, sp, wu, wv, ww);
}

void nearSolidGrid(Grid g, const float *sp, uint8_t *near, uint8_t *tmp, int ci, int cj, int ck,
                   int layers, hipStream_t s) {
    size_t nc = (size_t)ci * cj * ck;
    hipCheck(hipMemsetAsync(near, 0, nc, s), "hipMemsetAsync");
    size_t n = (size_t)g.ni * g.nj * g.nk;
    hipLaunchKernelGGL(nearSolidMark, dim3(blocksFor(n, 256)), dim3(256), 0, s, g, sp, near, ci, cj);
    for (int l = 0; l < layers; l++) {
        hipLaunchKernelGGL(feather6Kernel, dim3(blocksFor(nc, 256)), dim3(256), 0, s, near, tmp, ci, cj, ck);
        hipCheck(hipMemcpyAsync(near, tmp, nc, hipMemcpyDeviceToDevice, s), "hipMemcpyAsync");
    }
}

}  // namespace ff
