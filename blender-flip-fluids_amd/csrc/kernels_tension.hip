// kernels_tension.hip -- what surface tension adds to the substep: the
// curvature of the liquid surface and the cluster status of the pressure
// solve's tension terms.
//
// Curvature (ParticleLevelSet::calculateCurvatureGrid,
// particlelevelset.cpp:196-292): the cells within 2 dx of the surface, grown
// to whole 6^3 blocks and one block further through the blocks' faces, are
// the solver cells of LevelSetSolver::reinitializeUpwind
// (levelsetsolver.cpp:82-205, 317-390): Jacobi steps of the upwind
// reinitialization between two buffers, each step one launch over the cells
// with a per-block maximum of the change, the host comparing successive
// maxima for the reference's convergence test.  The reference's cell list
// starts with one copy of cell (0, 0, 0) per solver cell, so that cell is a
// solver cell too; its second buffer is never initialized and is read after
// the first swap outside the solver cells: here it starts at 0.  The
// curvature is the mean curvature of the result at the cells whose six
// neighbours lie within 2 dx of the surface, extrapolated 3 layers.
//
// Cluster status (PressureSolver::_initializeSurfaceTensionClusterData,
// pressuresolver.cpp:314-606): the reference grows a breadth-first search
// from each cell beside the surface and marks the cells it visits BAD when
// the search ends with fewer than 36 cells without meeting a cell of a 4^3
// block wholly in one phase.  A search that meets such a cell, or queues 36
// cells, marks OK instead, and later searches that meet those marks inherit
// OK, so the cells that end BAD are exactly those of the connected same-phase
// regions (through cells outside such blocks and off the border) with fewer
// than 36 cells, no same-phase neighbour in such a block, and a cell beside
// the surface.  The GPU finds the regions with the lock-free union-find and
// counts each one's cells up to 36.  Only the BAD marks enter the tension
// terms; cells the reference's searches mark OK stay 0 here.
#include <algorithm>
#include <cmath>
#include <limits>
#include <stdexcept>
#include <vector>

#include "kernels.h"
#include "unionfind.h"

namespace ff {

namespace {

constexpr int TPB = 256;
constexpr int BAND = 3;        // _curvatureGridExactBand (particlelevelset.h:143)
constexpr int BW = 2 * BAND;   // solver block width
constexpr int LAYERS = 3;      // _curvatureGridExtrapolationLayers
constexpr int CBW = 4;         // cluster block width (pressuresolver.h:211)
constexpr int THRESHOLD = 36;  // _surfaceTensionClusterThreshold (pressuresolver.h:210)
enum : int8_t { ST_UNSET = 0, ST_OK_IN = 1, ST_OK_OUT = 2, ST_BAD_IN = 4, ST_BAD_OUT = 8, ST_BORDER = 16 };

unsigned blocksFor(size_t n, unsigned t) { return (unsigned)((n + t - 1) / t); }
size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

struct Dims {
    int ni, nj, nk;
    __device__ size_t n() const { return (size_t)ni * nj * nk; }
    __device__ void ijk(size_t c, int &i, int &j, int &k) const {
        i = (int)(c % ni);
        j = (int)((c / ni) % nj);
        k = (int)(c / ((size_t)ni * nj));
    }
};

// ------------------------------------------------------------- curvature ---
__global__ void validKernel(Dims D, const float *phi, float maxDist, uint8_t *valid) {
    const size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= D.n()) return;
    valid[c] = fabsf(phi[c]) < maxDist ? 1 : 0;
}
// a 6^3 block holds a valid cell
__global__ void blockAnyKernel(Dims D, int bi, int bj, int bk, const uint8_t *valid, uint8_t *blocks) {
    const size_t b = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= (size_t)bi * bj * bk) return;
    const int x = (int)(b % bi), y = (int)((b / bi) % bj), z = (int)(b / ((size_t)bi * bj));
    uint8_t any = 0;
    for (int k = z * BW; k < min((z + 1) * BW, D.nk) && !any; k++)
        for (int j = y * BW; j < min((y + 1) * BW, D.nj) && !any; j++)
            for (int i = x * BW; i < min((x + 1) * BW, D.ni); i++)
                if (valid[flat(i, j, k, D.ni, D.nj)]) {
                    any = 1;
                    break;
                }
    blocks[b] = any;
}
// GridUtils::featherGrid6 (gridutils.cpp:229-262); *any: some block is set
__global__ void featherKernel(int bi, int bj, int bk, const uint8_t *in, uint8_t *out, int *any) {
    const size_t b = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= (size_t)bi * bj * bk) return;
    const int x = (int)(b % bi), y = (int)((b / bi) % bj), z = (int)(b / ((size_t)bi * bj));
    uint8_t v = in[b];
    const int nb[6][3] = {{-1, 0, 0}, {1, 0, 0}, {0, -1, 0}, {0, 1, 0}, {0, 0, -1}, {0, 0, 1}};
    for (int q = 0; q < 6 && !v; q++) {
        const int a = x + nb[q][0], c = y + nb[q][1], e = z + nb[q][2];
        if (inRange(a, c, e, bi, bj, bk) && in[flat(a, c, e, bi, bj)]) v = 1;
    }
    out[b] = v;
    if (v) atomicOr(any, 1);
}
// the cells of the grown blocks: valid nodes and solver cells; cell (0, 0, 0)
// is a solver cell whenever there is any
__global__ void solverKernel(Dims D, int bi, int bj, const uint8_t *blocks, const int *any, uint8_t *valid,
                             uint8_t *solver) {
    const size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= D.n()) return;
    int i, j, k;
    D.ijk(c, i, j, k);
    const uint8_t v = blocks[flat(i / BW, j / BW, k / BW, bi, bj)];
    valid[c] = v;
    solver[c] = (v || (c == 0 && *any)) ? 1 : 0;
}
__device__ inline float signOf(float phi, float dx) {
    const double d = phi;
    return (float)(d / sqrt(d * d + (double)(dx * dx)));
}
// per-block maximum of the sign function over all cells
__global__ __launch_bounds__(TPB) void signMaxKernel(Dims D, const float *phi, float dx, float *partial) {
    __shared__ float sh[TPB];
    const size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    sh[threadIdx.x] = c < D.n() ? signOf(phi[c], dx) : -3.4028235e38f;
    __syncthreads();
    for (int o = TPB / 2; o > 0; o >>= 1) {
        if (threadIdx.x < (unsigned)o) sh[threadIdx.x] = fmaxf(sh[threadIdx.x], sh[threadIdx.x + o]);
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = sh[0];
}
__global__ void fillKernel(float *a, size_t n, float v) {
    const size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c < n) a[c] = v;
}
__device__ inline float sq(float x) { return x * x; }
// one reinitialization step on the solver cells (levelsetsolver.cpp:173-205
// with _getDerivativesUpwind 351-390) and the per-block maximum change
__global__ __launch_bounds__(TPB) void upwindStepKernel(Dims D, const float *out, float *tmp, const uint8_t *solver,
                                                        float dx, float dtau, float *partial) {
    __shared__ float sh[TPB];
    const size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    float diff = 0.0f;
    if (c < D.n() && solver[c]) {
        int i, j, k;
        D.ijk(c, i, j, k);
        const float invdx = 1.0f / dx;
        const float C0 = out[c];
        const float s = signOf(C0, dx);
        const int im = i < 1 ? 0 : i - 1, ip = min(i + 1, D.ni - 1);
        const int jm = j < 1 ? 0 : j - 1, jp = min(j + 1, D.nj - 1);
        const int km = k < 1 ? 0 : k - 1, kp = min(k + 1, D.nk - 1);
        const float dx0 = invdx * (C0 - out[flat(im, j, k, D.ni, D.nj)]);
        const float dx1 = invdx * (out[flat(ip, j, k, D.ni, D.nj)] - C0);
        const float dy0 = invdx * (C0 - out[flat(i, jm, k, D.ni, D.nj)]);
        const float dy1 = invdx * (out[flat(i, jp, k, D.ni, D.nj)] - C0);
        const float dz0 = invdx * (C0 - out[flat(i, j, km, D.ni, D.nj)]);
        const float dz1 = invdx * (out[flat(i, j, kp, D.ni, D.nj)] - C0);
        const float val =
            C0 -
            dtau * fmaxf(s, 0.0f) *
                (sqrtf(sq(fmaxf(dx0, 0.0f)) + sq(fminf(dx1, 0.0f)) + sq(fmaxf(dy0, 0.0f)) + sq(fminf(dy1, 0.0f)) +
                       sq(fmaxf(dz0, 0.0f)) + sq(fminf(dz1, 0.0f))) -
                 1.0f) -
            dtau * fminf(s, 0.0f) *
                (sqrtf(sq(fminf(dx0, 0.0f)) + sq(fmaxf(dx1, 0.0f)) + sq(fminf(dy0, 0.0f)) + sq(fmaxf(dy1, 0.0f)) +
                       sq(fminf(dz0, 0.0f)) + sq(fmaxf(dz1, 0.0f))) -
                 1.0f);
        tmp[c] = val;
        diff = fabsf(val - C0);
    }
    sh[threadIdx.x] = diff;
    __syncthreads();
    for (int o = TPB / 2; o > 0; o >>= 1) {
        if (threadIdx.x < (unsigned)o) sh[threadIdx.x] = fmaxf(sh[threadIdx.x], sh[threadIdx.x + o]);
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = sh[0];
}
__global__ void outOfRangeKernel(Dims D, float *surface, const uint8_t *valid, float dist) {
    const size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c < D.n() && !valid[c]) surface[c] = dist;
}
// _getValidCurvatureNodes (particlelevelset.cpp:692-726) and _getCurvature
// (728-823): the mean curvature at interior cells whose six neighbours and
// themselves lie within (band - 1) dx of the surface, clamped to +-1/dx
__global__ void curvatureKernel(Dims D, const float *surface, float upper, double dx, uint8_t *cv,
                                float *kgrid) {
    const size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= D.n()) return;
    int i, j, k;
    D.ijk(c, i, j, k);
    auto P = [&](int a, int b, int e) { return surface[flat(a, b, e, D.ni, D.nj)]; };
    auto near = [&](int a, int b, int e) { return fabsf(P(a, b, e)) < upper; };
    bool ok = i > 0 && j > 0 && k > 0 && i < D.ni - 1 && j < D.nj - 1 && k < D.nk - 1;
    ok = ok && near(i, j, k) && near(i + 1, j, k) && near(i - 1, j, k) && near(i, j + 1, k) && near(i, j - 1, k) &&
         near(i, j, k + 1) && near(i, j, k - 1);
    cv[c] = ok ? 1 : 0;
    float kv = 0.0f;
    if (ok) {
        const float x = 0.5f * (P(i + 1, j, k) - P(i - 1, j, k));
        const float y = 0.5f * (P(i, j + 1, k) - P(i, j - 1, k));
        const float z = 0.5f * (P(i, j, k + 1) - P(i, j, k - 1));
        const float xx = P(i + 1, j, k) - 2.0f * P(i, j, k) + P(i - 1, j, k);
        const float yy = P(i, j + 1, k) - 2.0f * P(i, j, k) + P(i, j - 1, k);
        const float zz = P(i, j, k + 1) - 2.0f * P(i, j, k) + P(i, j, k - 1);
        const float xy = 0.25f * (P(i + 1, j + 1, k) - P(i - 1, j + 1, k) - P(i + 1, j - 1, k) + P(i - 1, j - 1, k));
        const float xz = 0.25f * (P(i + 1, j, k + 1) - P(i - 1, j, k + 1) - P(i + 1, j, k - 1) + P(i - 1, j, k - 1));
        const float yz = 0.25f * (P(i, j + 1, k + 1) - P(i, j - 1, k + 1) - P(i, j + 1, k - 1) + P(i, j - 1, k - 1));
        float den = x * x + y * y + z * z;
        den = (float)sqrt((double)(den * den * den));
        if (!(den < 1e-9f)) {
            const float kmax = 1.0f / (float)dx;
            kv = (float)((double)((xx * (y * y + z * z) + yy * (x * x + z * z) + zz * (x * x + y * y) -
                                   2 * xy * x * y - 2 * xz * x * z - 2 * yz * y * z) /
                                  den) /
                         dx);
            kv = fminf(kv, kmax);
            kv = fmaxf(kv, -kmax);
        }
    }
    kgrid[c] = kv;
}

// -------------------------------------------------------- cluster status ---
__global__ void block4Kernel(Dims D, int bi, int bj, int bk, const float *phi, uint8_t *blocks) {
    const size_t b = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= (size_t)bi * bj * bk) return;
    const int x = (int)(b % bi), y = (int)((b / bi) % bj), z = (int)(b / ((size_t)bi * bj));
    uint8_t f = 0;
    for (int k = z * CBW; k < (z + 1) * CBW; k++)
        for (int j = y * CBW; j < (y + 1) * CBW; j++)
            for (int i = x * CBW; i < (x + 1) * CBW; i++) f |= phi[flat(i, j, k, D.ni, D.nj)] < 0 ? 1 : 2;
    blocks[b] = f;
}
__global__ void cellStatusKernel(Dims D, int bi, int bj, int bk, const uint8_t *blocks, int8_t *status,
                                 int *parent, int *count, int *flags) {
    const size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= D.n()) return;
    int i, j, k;
    D.ijk(c, i, j, k);
    int8_t s = ST_UNSET;
    if (i == 0 || j == 0 || k == 0 || i == D.ni - 1 || j == D.nj - 1 || k == D.nk - 1) {
        s = ST_BORDER;
    } else if (inRange(i / CBW, j / CBW, k / CBW, bi, bj, bk)) {
        const uint8_t b = blocks[flat(i / CBW, j / CBW, k / CBW, bi, bj)];
        if ((b & 1) && !(b & 2))
            s = ST_OK_IN;
        else if (!(b & 1) && (b & 2))
            s = ST_OK_OUT;
    }
    status[c] = s;
    parent[c] = s == ST_UNSET ? (int)c : -1;
    count[c] = 0;
    flags[c] = 0;
}
__global__ void clusterUniteKernel(Dims D, const float *phi, const int8_t *status, int *parent) {
    const size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= D.n() || status[c] != ST_UNSET) return;
    int i, j, k;
    D.ijk(c, i, j, k);
    const bool in = phi[c] < 0.0f;
    const size_t nb[3] = {flat(i + 1, j, k, D.ni, D.nj), flat(i, j + 1, k, D.ni, D.nj), flat(i, j, k + 1, D.ni, D.nj)};
    for (int q = 0; q < 3; q++)
        if (status[nb[q]] == ST_UNSET && (phi[nb[q]] < 0.0f) == in) unite(parent, (int)c, (int)nb[q]);
}
// count[root] counts the region's cells up to 36; flags[root]: bit 0 a cell
// beside the surface, bit 1 a same-phase neighbour in a one-phase block.
// Atomics run only while a plain read still finds them needed.
__global__ void clusterCountKernel(Dims D, const float *phi, const int8_t *status, int *parent, int *count,
                                   int *flags) {
    const size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= D.n() || status[c] != ST_UNSET) return;
    int i, j, k;
    D.ijk(c, i, j, k);
    const int r = findRoot(parent, (int)c);
    if (__hip_atomic_load(count + r, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < THRESHOLD) atomicAdd(count + r, 1);
    const bool in = phi[c] < 0.0f;
    int f = 0;
    const size_t nb[6] = {flat(i + 1, j, k, D.ni, D.nj), flat(i - 1, j, k, D.ni, D.nj), flat(i, j + 1, k, D.ni, D.nj),
                          flat(i, j - 1, k, D.ni, D.nj), flat(i, j, k + 1, D.ni, D.nj), flat(i, j, k - 1, D.ni, D.nj)};
    for (int q = 0; q < 6; q++) {
        const bool nin = phi[nb[q]] < 0.0f;
        if (nin != in) f |= 1;
        else if (status[nb[q]] == ST_OK_IN || status[nb[q]] == ST_OK_OUT) f |= 2;
    }
    if (f && (__hip_atomic_load(flags + r, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) & f) != f) atomicOr(flags + r, f);
}
__global__ void clusterMarkKernel(Dims D, const float *phi, int8_t *status, int *parent, const int *count,
                                  const int *flags) {
    const size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= D.n() || status[c] != ST_UNSET) return;
    const int r = findRoot(parent, (int)c);
    if (count[r] < THRESHOLD && flags[r] == 1) status[c] = phi[c] < 0.0f ? ST_BAD_IN : ST_BAD_OUT;
}

struct Layout {
    size_t N, nb6, nb4, nPartial;
    int bi, bj, bk, ci, cj, ck;
};
Layout layoutOf(Grid g) {
    Layout L;
    L.N = (size_t)g.ni * g.nj * g.nk;
    L.bi = (g.ni + BW - 1) / BW;
    L.bj = (g.nj + BW - 1) / BW;
    L.bk = (g.nk + BW - 1) / BW;
    L.nb6 = (size_t)L.bi * L.bj * L.bk;
    L.ci = g.ni / CBW;
    L.cj = g.nj / CBW;
    L.ck = g.nk / CBW;
    L.nb4 = (size_t)L.ci * L.cj * L.ck;
    L.nPartial = blocksFor(L.N, TPB);
    return L;
}

float hostMax(const float *dev, size_t n, hipStream_t s) {
    std::vector<float> h(n);
    hipCheck(hipMemcpyAsync(h.data(), dev, sizeof(float) * n, hipMemcpyDeviceToHost, s), "D2H");
    hipCheck(hipStreamSynchronize(s), "sync");
    float m = -std::numeric_limits<float>::max();
    for (float x : h) m = std::max(x, m);
    return m;
}

}  // namespace

size_t tensionBytes(Grid g) {
    const Layout L = layoutOf(g);
    return 3 * align256(4 * L.N) + 5 * align256(L.N) + 3 * align256(4 * L.N) +
           3 * align256(std::max<size_t>(L.nb6, L.nb4) + 1) + align256(4 * (L.nPartial + 1)) + align256(4);
}

void tensionCarve(Grid g, void *block, TensionBuffers &B) {
    const Layout L = layoutOf(g);
    char *p = (char *)block;
    auto take = [&](size_t b) {
        char *q = p;
        p += align256(b);
        return q;
    };
    B.a = (float *)take(4 * L.N);
    B.b = (float *)take(4 * L.N);
    B.curvature = (float *)take(4 * L.N);
    B.status = (int8_t *)take(L.N);
    B.valid = (uint8_t *)take(L.N);
    B.solver = (uint8_t *)take(L.N);
    B.cv = (uint8_t *)take(L.N);
    B.ext = (uint8_t *)take(L.N);
    B.parent = (int *)take(4 * L.N);
    B.count = (int *)take(4 * L.N);
    B.rootFlags = (int *)take(4 * L.N);
    const size_t nbmax = std::max<size_t>(L.nb6, L.nb4) + 1;
    B.blocks = (uint8_t *)take(nbmax);
    B.blocks2 = (uint8_t *)take(nbmax);
    B.blocks4 = (uint8_t *)take(nbmax);
    B.partial = (float *)take(4 * (L.nPartial + 1));
    B.any = (int *)take(4);
    B.surface = nullptr;
}

void curvatureGrid(Grid g, const float *phi, TensionBuffers &B, hipStream_t s) {
    const Layout L = layoutOf(g);
    const Dims D{g.ni, g.nj, g.nk};
    const unsigned nbN = blocksFor(L.N, TPB);
    int *any = B.any;
    hipCheck(hipMemsetAsync(any, 0, sizeof(int), s), "hipMemsetAsync");
    const float maxSurfaceDist = 2.0f * g.dx;
    validKernel<<<nbN, TPB, 0, s>>>(D, phi, maxSurfaceDist, B.valid);
    blockAnyKernel<<<blocksFor(L.nb6, TPB), TPB, 0, s>>>(D, L.bi, L.bj, L.bk, B.valid, B.blocks);
    featherKernel<<<blocksFor(L.nb6, TPB), TPB, 0, s>>>(L.bi, L.bj, L.bk, B.blocks, B.blocks2, any);
    solverKernel<<<nbN, TPB, 0, s>>>(D, L.bi, L.bj, B.blocks2, any, B.valid, B.solver);
    hipCheck(hipGetLastError(), "curvature launch");
    // reinitializeUpwind: pseudo time step and step count on the host in the
    // reference's float arithmetic
    const float fdx = (float)g.dx;
    const float width = BAND * g.dx;
    signMaxKernel<<<nbN, TPB, 0, s>>>(D, phi, fdx, B.partial);
    const float maxS = std::max(hostMax(B.partial, nbN, s), -std::numeric_limits<float>::max());
    const float maxCFL = 0.5f;
    float dtau = maxCFL * fdx;
    while (dtau * maxS / fdx > maxCFL) dtau *= 0.5;
    const int iterations = (int)std::ceil(width / dtau);
    hipCheck(hipMemcpyAsync(B.a, phi, sizeof(float) * L.N, hipMemcpyDeviceToDevice, s), "D2D");
    fillKernel<<<nbN, TPB, 0, s>>>(B.b, L.N, 0.0f);
    float *out = B.a, *tmp = B.b;
    float lastMaxDiff = -1.0f;
    for (int n = 0; n < iterations; n++) {
        upwindStepKernel<<<nbN, TPB, 0, s>>>(D, out, tmp, B.solver, fdx, dtau, B.partial);
        hipCheck(hipGetLastError(), "curvature step launch");
        const float maxDiff = std::max(hostMax(B.partial, nbN, s), 0.0f);
        std::swap(tmp, out);
        if (std::abs(maxDiff - lastMaxDiff) < 0.01f * fdx) break;
        lastMaxDiff = maxDiff;
    }
    const float outDist = 5.0f * g.dx;
    outOfRangeKernel<<<nbN, TPB, 0, s>>>(D, out, B.valid, outDist);
    const float upper = (BAND - 1) * g.dx;
    curvatureKernel<<<nbN, TPB, 0, s>>>(D, out, upper, g.dx, B.cv, B.curvature);
    hipCheck(hipGetLastError(), "curvature launch");
    extrapolateInit(g.ni, g.nj, g.nk, B.cv, B.ext, s);
    for (int l = 0; l < LAYERS; l++) extrapolateLayer(g.ni, g.nj, g.nk, B.curvature, B.ext, l, s);
    B.surface = out;
}

void tensionClusterStatus(Grid g, const float *phi, TensionBuffers &B, hipStream_t s) {
    const Layout L = layoutOf(g);
    const Dims D{g.ni, g.nj, g.nk};
    if (L.N > (size_t)0x7fffffff) throw std::runtime_error("cluster status: grid exceeds the 32-bit cell range");
    const unsigned nbN = blocksFor(L.N, TPB);
    int *flags = B.rootFlags;
    if (L.nb4 > 0) block4Kernel<<<blocksFor(L.nb4, TPB), TPB, 0, s>>>(D, L.ci, L.cj, L.ck, phi, B.blocks4);
    cellStatusKernel<<<nbN, TPB, 0, s>>>(D, L.ci, L.cj, L.ck, B.blocks4, B.status, B.parent, B.count, flags);
    clusterUniteKernel<<<nbN, TPB, 0, s>>>(D, phi, B.status, B.parent);
    clusterCountKernel<<<nbN, TPB, 0, s>>>(D, phi, B.status, B.parent, B.count, flags);
    clusterMarkKernel<<<nbN, TPB, 0, s>>>(D, phi, B.status, B.parent, B.count, flags);
    hipCheck(hipGetLastError(), "cluster launch");
}

}  // namespace ff
