// kernels_solid.hip -- what a moving solid adds to the pressure solve: the
// cells' solid volume weights and the conditioning of the solid's face
// velocities around isolated pockets of liquid.
//
// Cell weights (fluidsimulation.cpp:6140-6147, MeshLevelSet::_getCellWeight
// meshlevelset.cpp:1491-1515): one thread per cell evaluates the volume
// fraction of the solid in the cell from its eight nodal SDF values with the
// reference's ten-tetrahedron split (levelsetutils.cpp:213-258) in float.
//
// Conditioning (PressureSolver::_conditionSolidVelocityField,
// pressuresolver.cpp:161-312): the reference grows the groups of connected
// liquid cells one by one with a stack.  The groups are the connected
// components of a graph (interior liquid cells; an edge where the shared face
// has weight >= 1e-6), so the GPU finds them with a lock-free union-find in
// one pass over the cells' +x/+y/+z edges: every root hooks onto the smaller
// of two roots with a compare-and-swap, so parent indices only decrease and
// every find and retry ends.  A second pass flags each component that has an
// edge (more than one cell) and each that borders air; a third zeroes the
// faces of the cells of isolated components with an edge.
#include <stdexcept>

#include "kernels.h"
#include "unionfind.h"

namespace ff {

namespace {

constexpr int TPB = 256;
unsigned blocksFor(size_t n, unsigned t) { return (unsigned)((n + t - 1) / t); }

__global__ void cellWeightKernel(Grid g, const float *solid, float *wc) {
    const size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= (size_t)g.ni * g.nj * g.nk) return;
    const int i = (int)(c % g.ni), j = (int)((c / g.ni) % g.nj), k = (int)(c / ((size_t)g.ni * g.nj));
    const int pi = g.ni + 1, pj = g.nj + 1;
    auto P = [&](int a, int b, int d) { return solid[flat(i + a, j + b, k + d, pi, pj)]; };
    const float p000 = P(0, 0, 0), p001 = P(0, 0, 1), p010 = P(0, 1, 0), p011 = P(0, 1, 1);
    const float p100 = P(1, 0, 0), p101 = P(1, 0, 1), p110 = P(1, 1, 0), p111 = P(1, 1, 1);
    float weight;
    if (p000 < 0 && p001 < 0 && p010 < 0 && p011 < 0 && p100 < 0 && p101 < 0 && p110 < 0 && p111 < 0)
        weight = 1.0f;
    else if (p000 >= 0 && p001 >= 0 && p010 >= 0 && p011 >= 0 && p100 >= 0 && p101 >= 0 && p110 >= 0 &&
             p111 >= 0)
        weight = 0.0f;
    else
        weight = volumeFraction8(p000, p100, p010, p110, p001, p101, p011, p111);
    // clamp(1 - weight, 0, 1) as max(0, min(v, 1))
    float v = 1.0f - weight;
    v = 1.0f < v ? 1.0f : v;
    wc[c] = 0.0f < v ? v : 0.0f;
}

// ---- conditioning ----

struct CondView {
    int ni, nj, nk;
    const float *phi, *wu, *wv, *ww;
    __device__ bool border(int i, int j, int k) const {
        return i == 0 || j == 0 || k == 0 || i == ni - 1 || j == nj - 1 || k == nk - 1;
    }
    // a member of the graph: an interior liquid cell
    __device__ bool member(int i, int j, int k) const { return !border(i, j, k) && phi[flat(i, j, k, ni, nj)] < 0.0f; }
    __device__ float U(int i, int j, int k) const { return wu[flat(i, j, k, ni + 1, nj)]; }
    __device__ float V(int i, int j, int k) const { return wv[flat(i, j, k, ni, nj + 1)]; }
    __device__ float W(int i, int j, int k) const { return ww[flat(i, j, k, ni, nj)]; }
    __device__ float L(int i, int j, int k) const { return phi[flat(i, j, k, ni, nj)]; }
    // _computeBordersAirGridThread (pressuresolver.cpp:314-360): an open face
    // (weight >= 1e-6) to an air cell (phi >= 0)
    __device__ bool bordersAir(int i, int j, int k) const {
        const float eps = 1e-6f;
        return (U(i, j, k) >= 0.0f + eps && L(i - 1, j, k) >= 0.0f) ||
               (U(i + 1, j, k) >= 0.0f + eps && L(i + 1, j, k) >= 0.0f) ||
               (V(i, j, k) >= 0.0f + eps && L(i, j - 1, k) >= 0.0f) ||
               (V(i, j + 1, k) >= 0.0f + eps && L(i, j + 1, k) >= 0.0f) ||
               (W(i, j, k) >= 0.0f + eps && L(i, j, k - 1) >= 0.0f) ||
               (W(i, j, k + 1) >= 0.0f + eps && L(i, j, k + 1) >= 0.0f);
    }
};

__global__ void condInitKernel(CondView v, int *parent, int *info) {
    const size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= (size_t)v.ni * v.nj * v.nk) return;
    const int i = (int)(c % v.ni), j = (int)((c / v.ni) % v.nj), k = (int)(c / ((size_t)v.ni * v.nj));
    parent[c] = v.member(i, j, k) ? (int)c : -1;
    info[c] = 0;
}

__global__ void condUniteKernel(CondView v, int *parent) {
    const size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= (size_t)v.ni * v.nj * v.nk) return;
    const int i = (int)(c % v.ni), j = (int)((c / v.ni) % v.nj), k = (int)(c / ((size_t)v.ni * v.nj));
    if (!v.member(i, j, k)) return;
    const float eps = 1e-6f;
    if (v.member(i + 1, j, k) && v.U(i + 1, j, k) >= eps) unite(parent, (int)c, (int)flat(i + 1, j, k, v.ni, v.nj));
    if (v.member(i, j + 1, k) && v.V(i, j + 1, k) >= eps) unite(parent, (int)c, (int)flat(i, j + 1, k, v.ni, v.nj));
    if (v.member(i, j, k + 1) && v.W(i, j, k + 1) >= eps) unite(parent, (int)c, (int)flat(i, j, k + 1, v.ni, v.nj));
}

// info[root]: bit 1 = the component has an edge (more than one cell), bit 0 =
// a cell borders air.  Nearly every cell of the liquid body shares one root,
// so each flag is set with an atomic only while a plain read still finds it
// clear: once set, the other cells only read it.
__global__ void condFlagKernel(CondView v, int *parent, int *info) {
    const size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= (size_t)v.ni * v.nj * v.nk) return;
    const int i = (int)(c % v.ni), j = (int)((c / v.ni) % v.nj), k = (int)(c / ((size_t)v.ni * v.nj));
    if (!v.member(i, j, k)) return;
    const float eps = 1e-6f;
    const bool edge = (v.member(i + 1, j, k) && v.U(i + 1, j, k) >= eps) ||
                      (v.member(i, j + 1, k) && v.V(i, j + 1, k) >= eps) ||
                      (v.member(i, j, k + 1) && v.W(i, j, k + 1) >= eps);
    const bool air = v.bordersAir(i, j, k);
    if (!edge && !air) return;
    const int r = findRoot(parent, (int)c);
    const int have = __hip_atomic_load(info + r, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const int want = (edge ? 2 : 0) | (air ? 1 : 0);
    if ((have & want) != want) atomicOr(info + r, want);
}

__global__ void condZeroKernel(CondView v, int *parent, const int *info, float *us, float *vs, float *ws) {
    const size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= (size_t)v.ni * v.nj * v.nk) return;
    const int i = (int)(c % v.ni), j = (int)((c / v.ni) % v.nj), k = (int)(c / ((size_t)v.ni * v.nj));
    if (!v.member(i, j, k)) return;
    const int r = findRoot(parent, (int)c);
    const int in = info[r];
    if (in != 2) return;  // isolated (no air) and more than one cell
    const int ni = v.ni, nj = v.nj;
    us[flat(i, j, k, ni + 1, nj)] = 0.0f;
    us[flat(i + 1, j, k, ni + 1, nj)] = 0.0f;
    vs[flat(i, j, k, ni, nj + 1)] = 0.0f;
    vs[flat(i, j + 1, k, ni, nj + 1)] = 0.0f;
    ws[flat(i, j, k, ni, nj)] = 0.0f;
    ws[flat(i, j, k + 1, ni, nj)] = 0.0f;
}

}  // namespace

void cellWeights(Grid g, const float *solid, float *wc, hipStream_t s) {
    const size_t n = (size_t)g.ni * g.nj * g.nk;
    hipLaunchKernelGGL(cellWeightKernel, dim3(blocksFor(n, TPB)), dim3(TPB), 0, s, g, solid, wc);
}

void conditionSolidVelocity(Grid g, const float *phi, const float *wu, const float *wv, const float *ww, float *us,
                            float *vs, float *ws, int *scratch, hipStream_t s) {
    const size_t n = (size_t)g.ni * g.nj * g.nk;
    if (n > (size_t)0x7fffffff) throw std::runtime_error("conditionSolidVelocity: grid too large for int cell indices");
    int *parent = scratch, *info = scratch + n;
    const CondView v{g.ni, g.nj, g.nk, phi, wu, wv, ww};
    const dim3 grid(blocksFor(n, TPB));
    hipLaunchKernelGGL(condInitKernel, grid, dim3(TPB), 0, s, v, parent, info);
    hipLaunchKernelGGL(condUniteKernel, grid, dim3(TPB), 0, s, v, parent);
    hipLaunchKernelGGL(condFlagKernel, grid, dim3(TPB), 0, s, v, parent, info);
    hipLaunchKernelGGL(condZeroKernel, grid, dim3(TPB), 0, s, v, parent, info, us, vs, ws);
}

}  // namespace ff
