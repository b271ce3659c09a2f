// kernels_mesh.hip -- the solid SDF of static triangle-mesh obstacles: the
// domain SDF unioned with each obstacle's signed distance level set
// (FluidSimulation::_addStaticObjectsToSDF, fluidsimulation.cpp:5277-5339).
//
// An obstacle's level set covers the obstacle's bounding box +-3 nodes
// (MeshObject::_getMeshIslandLevelSet, meshobject.cpp:667-701) and is the
// reference's "fast" signed distance field (meshlevelset.cpp:572-601):
//   - exact band: every node within a triangle's node box (its bounding box
//     +-3 nodes) takes the min over those triangles of the point-triangle
//     distance.  GPU: one workgroup per triangle walks its node box; the min
//     is an integer atomicMin on the float bits, exact and order-free since
//     every distance is >= +0;
//   - signs: a node is inside when the +z line through it crosses the mesh an
//     even number of times in all and an odd number of times below it
//     (meshutils.cpp:57-437).  GPU: one thread per triangle tests its
//     columns, per-column crossing lists are built by count, scan and fill,
//     then one thread per cell counts the crossings below it.
// Host code prepares what is per triangle or per mesh (node boxes, island
// grids, the half-cell shifted frame of the sign test) with the reference's
// float/double steps, so every device input is bit-identical to the
// reference's.  The lines through the cells are jittered by +-0.001 dx from
// the C library's rand(), three draws per test, as in the reference
// (meshutils.cpp:93-117): a program that seeds rand() identically gets the
// same jitter from both engines.
#include <algorithm>
#include <array>
#include <cmath>
#include <cstdlib>
#include <memory>
#include <stdexcept>
#include <vector>

#include "kernels.h"

namespace ff {

namespace {

constexpr int EXACT_BAND = 3;  // _solidLevelSetExactBand (fluidsimulation.h)
constexpr int LS_BLOCK = 10;   // MeshLevelSet::_blockwidth (meshlevelset.h:369)

// ------------------------------------------------------------ device --------

// one triangle of the exact band: vertices in the level set's frame and the
// inclusive node box it writes
struct BandTri {
    V3 a, b, c;
    int lo[3], hi[3];
};

// BLOCK: distances in 10^3-node block-local coordinates, the multi-threaded
// band (meshlevelset.cpp:1115-1163); otherwise in the level set's own frame,
// the single-threaded band (1165-1234).
template <bool BLOCK>
__global__ void meshBandKernel(const BandTri *tris, int pi, int pj, double dx, int *phiBits) {
    const BandTri t = tris[blockIdx.x];
    const int ni = t.hi[0] - t.lo[0] + 1, nj = t.hi[1] - t.lo[1] + 1, nk = t.hi[2] - t.lo[2] + 1;
    const long long count = (long long)ni * nj * nk;
    for (long long q = threadIdx.x; q < count; q += blockDim.x) {
        const int i = t.lo[0] + (int)(q % ni);
        const int j = t.lo[1] + (int)((q / ni) % nj);
        const int k = t.lo[2] + (int)(q / ((long long)ni * nj));
        V3 gpos, a = t.a, b = t.b, c = t.c;
        if (BLOCK) {
            const int bi = i / LS_BLOCK, bj = j / LS_BLOCK, bk = k / LS_BLOCK;
            const V3 off = indexToPos(bi, bj, bk, LS_BLOCK * dx);
            gpos = indexToPos(i - bi * LS_BLOCK, j - bj * LS_BLOCK, k - bk * LS_BLOCK, dx);
            a = sub(a, off);
            b = sub(b, off);
            c = sub(c, off);
        } else {
            gpos = indexToPos(i, j, k, dx);
        }
        const float dist = len3(sub(closestPointOnTriangle(gpos, a, b, c), gpos));
        atomicMin(phiBits + flat(i, j, k, pi, pj), __float_as_int(dist));
    }
}

// one triangle of a sign test: vertices in the shifted frame, inclusive column range
struct ColTri {
    V3 a, b, c;
    int i0, i1, j0, j1;
};

FF_HD V3 cross3(V3 a, V3 b) {  // vmath::cross (vmath.h:79-83)
    return V3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}

// Collision::lineIntersectsTriangle (collision.cpp:79-118), float
// Moller-Trumbore, hit point p + (float)t * dir
__device__ bool lineHitZ(V3 p, V3 v0, V3 v1, V3 v2, float *z) {
    const V3 dir{0.0f, 0.0f, 1.0f};
    const V3 e1 = sub(v1, v0), e2 = sub(v2, v0);
    const V3 h = cross3(dir, e2);
    const float a = dot3(e1, h);
    const double eps = 10e-9;
    if (fabsf(a) < eps) return false;
    const float f = 1 / a;
    const V3 s = sub(p, v0);
    const float u = f * dot3(s, h);
    if (u < 0.0 || u > 1.0) return false;
    const V3 q = cross3(s, e1);
    const float v = f * dot3(dir, q);
    if (v < 0.0 || u + v > 1.0) return false;
    const double t = f * dot3(e2, q);
    *z = add(p, scale((float)t, dir)).z;
    return true;
}

// the line of column (i, j): through the jittered cell centre (i, j, -1)
__device__ V3 columnOrigin(int i, int j, double dx, V3 jitter) {
    return add(indexToCellCenter(i, j, -1, dx), jitter);
}

// FILL = false: count the crossings of each column; true: write their z
template <bool FILL>
__global__ void crossingKernel(const ColTri *tris, int nt, int isz, double dx, V3 jitter, int *count,
                               const int *offs, int *cursor, float *zs) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nt) return;
    const ColTri T = tris[t];
    for (int j = T.j0; j <= T.j1; j++)
        for (int i = T.i0; i <= T.i1; i++) {
            float z;
            if (!lineHitZ(columnOrigin(i, j, dx, jitter), T.a, T.b, T.c, &z)) continue;
            const size_t col = (size_t)i + (size_t)isz * j;
            if (FILL) {
                // both passes make the same float operations on the same
                // inputs, so the fill finds exactly the counted crossings;
                // the bound keeps every write inside the column's segment
                const int slot = atomicAdd(cursor + col, 1);
                if (slot < offs[col + 1] - offs[col]) zs[offs[col] + slot] = z;
            } else {
                atomicAdd(count + col, 1);
            }
        }
}

// a cell of the shifted grid is inside when its column holds an even number
// of crossings and an odd number below its centre; it marks node cell + go
__global__ void insideKernel(int isz, int jsz, int ksz, double dx, const int *offs, const float *zs, I3 go, int pi,
                             int pj, int pk, uint8_t *inside) {
    const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t cells = (size_t)isz * jsz * ksz;
    if (idx >= cells) return;
    const int i = (int)(idx % isz), j = (int)((idx / isz) % jsz), k = (int)(idx / ((size_t)isz * jsz));
    const size_t col = (size_t)i + (size_t)isz * j;
    const int b = offs[col], n = offs[col + 1] - b;
    if (n % 2 != 0) return;
    const float zc = indexToCellCenter(i, j, k, dx).z;
    int less = 0;
    for (int q = 0; q < n; q++) less += zs[b + q] < zc;
    if (less % 2 == 1) {
        const int a = i + go.i, c = j + go.j, d = k + go.k;
        if (inRange(a, c, d, pi, pj, pk)) inside[flat(a, c, d, pi, pj)] = 1;
    }
}

__global__ void fillFloatKernel(float *a, size_t n, float v) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = v;
}

__global__ void negateInsideKernel(float *phi, const uint8_t *inside, size_t n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && inside[i]) phi[i] = -phi[i];
}

// MeshLevelSet::calculateUnion (meshlevelset.cpp:1759-1794): dst = min(dst,
// [-]src) over src's nodes at offset o.  dstObj (optional) follows the
// closest mesh object: a node takes the source's object when the source wins
// with a smaller magnitude and has a closest triangle there.  The source's
// object is given by srcObjArr (-1 = none) or, for a mesh level set, is srcObj
// where the node lies in the exact band (|phi| below the set's upper bound srcUb).
__global__ void unionKernel(const float *src, int qi, int qj, int qk, I3 o, bool negate, float *dst, int pi, int pj,
                            int pk, int *dstObj, const int *srcObjArr, int srcObj, float srcUb) {
    const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (size_t)qi * qj * qk) return;
    const int i = (int)(idx % qi), j = (int)((idx / qi) % qj), k = (int)(idx / ((size_t)qi * qj));
    const int a = i + o.i, b = j + o.j, c = k + o.k;
    if (!inRange(a, b, c, pi, pj, pk)) return;
    const float raw = src[idx];
    const float v = negate ? -raw : raw;
    const size_t t = flat(a, b, c, pi, pj);
    float &d = dst[t];
    if (v < d) {
        if (dstObj) {
            const int so = srcObjArr ? srcObjArr[idx] : (fabsf(raw) < srcUb ? srcObj : -1);
            if (so >= 0 && fabsf(v) < fabsf(d)) dstObj[t] = so;
        }
        d = v;
    }
}

unsigned blocksFor(size_t n, unsigned t) { return (unsigned)((n + t - 1) / t); }

// ------------------------------------------------------------ host ----------

template <class T>
struct DevArray {
    T *p = nullptr;
    size_t n = 0;
    explicit DevArray(size_t count) : n(count) {
        if (count) hipCheck(hipMalloc(&p, sizeof(T) * count), "hipMalloc (mesh obstacle)");
    }
    DevArray(const DevArray &) = delete;
    DevArray &operator=(const DevArray &) = delete;
    ~DevArray() {
        if (p) (void)hipFree(p);
    }
};

template <class T>
void upload(DevArray<T> &d, const std::vector<T> &h, hipStream_t s) {
    if (!h.empty())
        hipCheck(hipMemcpyAsync(d.p, h.data(), sizeof(T) * h.size(), hipMemcpyHostToDevice, s), "upload mesh data");
}

struct HMesh {
    std::vector<V3> v;
    std::vector<std::array<int, 3>> t;
};

// AABB(points) / AABB(Triangle, vertices): double min/max, extents + 1e-9 (aabb.cpp:54-112)
Box aabbOf(const V3 *p, size_t n) {
    if (n == 0) return Box{V3{0.0f, 0.0f, 0.0f}, 0.0, 0.0, 0.0};  // aabb.h:60-63
    double x0 = p[0].x, y0 = p[0].y, z0 = p[0].z, x1 = p[0].x, y1 = p[0].y, z1 = p[0].z;
    for (size_t q = 0; q < n; q++) {
        x0 = fmin((double)p[q].x, x0);
        y0 = fmin((double)p[q].y, y0);
        z0 = fmin((double)p[q].z, z0);
        x1 = fmax((double)p[q].x, x1);
        y1 = fmax((double)p[q].y, y1);
        z1 = fmax((double)p[q].z, z1);
    }
    const double eps = 1e-9;
    return Box{mk(x0, y0, z0), x1 - x0 + eps, y1 - y0 + eps, z1 - z0 + eps};
}
Box aabbTri(const HMesh &m, size_t t) {
    V3 p[3] = {m.v[m.t[t][0]], m.v[m.t[t][1]], m.v[m.t[t][2]]};
    return aabbOf(p, 3);
}
bool boxContains(const Box &g, const Box &b) { return g.inside(b.pos) && g.inside(b.maxPoint()); }
// AABB::getIntersection has positive extent (aabb.cpp:170-191)
bool boxIntersects(const Box &a, const Box &b) {
    V3 n1 = a.pos, x1 = a.maxPoint(), n2 = b.pos, x2 = b.maxPoint();
    if (n1.x > x2.x || n1.y > x2.y || n1.z > x2.z || x1.x < n2.x || x1.y < n2.y || x1.z < n2.z) return false;
    return fmin((double)x1.x, (double)x2.x) - fmax((double)n1.x, (double)n2.x) > 0.0 ||
           fmin((double)x1.y, (double)x2.y) - fmax((double)n1.y, (double)n2.y) > 0.0 ||
           fmin((double)x1.z, (double)x2.z) - fmax((double)n1.z, (double)n2.z) > 0.0;
}
void translate(HMesh &m, V3 d) {
    for (V3 &p : m.v) p = add(p, d);
}

// connected components over shared vertices, island vertices in first-use
// order (MeshUtils::_splitIntoMeshIslands, meshutils.cpp:464-573)
std::vector<HMesh> splitIslands(const HMesh &m) {
    const int nv = (int)m.v.size();
    std::vector<std::vector<int>> nb(nv);
    for (const auto &t : m.t)
        for (int a = 0; a < 3; a++)
            for (int b = 1; b < 3; b++) nb[t[a]].push_back(t[(a + b) % 3]);
    std::vector<int> group(nv, -1), stack;
    int ng = 0;
    for (int s = 0; s < nv; s++) {
        if (group[s] >= 0) continue;
        stack.assign(1, s);
        group[s] = ng;
        while (!stack.empty()) {
            int v = stack.back();
            stack.pop_back();
            for (int w : nb[v])
                if (group[w] < 0) {
                    group[w] = ng;
                    stack.push_back(w);
                }
        }
        ng++;
    }
    std::vector<HMesh> isl(ng);
    std::vector<int> tr(nv, -1);
    for (auto t : m.t) {
        HMesh &I = isl[group[t[0]]];
        for (int c = 0; c < 3; c++) {
            if (tr[t[c]] < 0) {
                tr[t[c]] = (int)I.v.size();
                I.v.push_back(m.v[t[c]]);
            }
            t[c] = tr[t[c]];
        }
        I.t.push_back(t);
    }
    std::vector<HMesh> out;
    for (HMesh &I : isl)
        if (!I.t.empty()) out.push_back(std::move(I));
    return out;
}

struct RandJitter {
    int draws = 0;
    // vec3(_randomDouble(jit, -jit) x3): the reference's compiler evaluates the
    // three arguments right to left, so z takes the first draw
    V3 next(double dx) {
        const double jit = 0.001 * dx;
        auto draw = [&]() {
            draws++;
            return jit + ((double)rand() / (double)RAND_MAX) * (-jit - jit);
        };
        const double z = draw(), y = draw(), x = draw();
        return mk(x, y, z);
    }
};

// One sign test (MeshUtils::getGridNodesInsideTriangleMesh, list form,
// meshutils.cpp:356-380) of mesh m in the level set's frame: marks the
// inside nodes of the (pi, pj, pk)-node level set.
void markInside(HMesh m, double dx, int pi, int pj, int pk, RandJitter &rj, uint8_t *inside, hipStream_t s) {
    if (m.v.empty()) {  // a 1x1x1 grid with no triangles: draws, marks nothing
        rj.next(dx);
        return;
    }
    translate(m, mk(0.5 * dx, 0.5 * dx, 0.5 * dx));
    Box b = aabbOf(m.v.data(), m.v.size());
    const I3 go = posToIndex(b.pos, dx);
    const V3 off = indexToPos(go.i, go.j, go.k, dx);
    translate(m, V3{-off.x, -off.y, -off.z});
    b.pos = sub(b.pos, off);
    const I3 g2 = posToIndex(add(b.pos, mk(b.w, b.h, b.d)), dx);
    const int isz = g2.i + 1, jsz = g2.j + 1, ksz = g2.k + 1;
    if (isz <= 0 || jsz <= 0 || ksz <= 0) throw std::runtime_error("mesh obstacle: empty sign-test grid");
    // columns of each triangle (Grid3d::getGridIndexBounds, grid3d.h:423-436)
    std::vector<ColTri> ct;
    ct.reserve(m.t.size());
    for (size_t t = 0; t < m.t.size(); t++) {
        const Box tb = aabbTri(m, t);
        const I3 a = posToIndex(tb.pos, dx), c = posToIndex(add(tb.pos, mk(tb.w, tb.h, tb.d)), dx);
        ColTri T{m.v[m.t[t][0]], m.v[m.t[t][1]], m.v[m.t[t][2]], std::max(a.i, 0), std::min(c.i, isz - 1),
                 std::max(a.j, 0), std::min(c.j, jsz - 1)};
        if (T.i0 <= T.i1 && T.j0 <= T.j1) ct.push_back(T);
    }
    const V3 jitter = rj.next(dx);
    const size_t ncol = (size_t)isz * jsz;
    DevArray<ColTri> dct(ct.size());
    DevArray<int> count(ncol + 1), offs(ncol + 1), cursor(ncol), scan(scanScratchInts(ncol + 1));
    upload(dct, ct, s);
    hipCheck(hipMemsetAsync(count.p, 0, sizeof(int) * (ncol + 1), s), "hipMemsetAsync");
    hipCheck(hipMemsetAsync(cursor.p, 0, sizeof(int) * ncol, s), "hipMemsetAsync");
    const int nt = (int)ct.size();
    if (nt > 0)
        hipLaunchKernelGGL(crossingKernel<false>, dim3(blocksFor(nt, 128)), dim3(128), 0, s, dct.p, nt, isz, dx,
                           jitter, count.p, nullptr, nullptr, nullptr);
    exclusiveScan(count.p, offs.p, ncol, scan.p, s);
    int total = 0;
    hipCheck(hipMemcpyAsync(&total, offs.p + ncol, sizeof(int), hipMemcpyDeviceToHost, s), "D2H crossings");
    hipCheck(hipStreamSynchronize(s), "crossing count");
    DevArray<float> zs(total > 0 ? total : 1);
    if (nt > 0)
        hipLaunchKernelGGL(crossingKernel<true>, dim3(blocksFor(nt, 128)), dim3(128), 0, s, dct.p, nt, isz, dx,
                           jitter, nullptr, offs.p, cursor.p, zs.p);
    const size_t cells = ncol * ksz;
    hipLaunchKernelGGL(insideKernel, dim3(blocksFor(cells, 256)), dim3(256), 0, s, isz, jsz, ksz, dx, offs.p, zs.p,
                       go, pi, pj, pk, inside);
    hipCheck(hipStreamSynchronize(s), "sign test");
}

// MeshLevelSet::fastCalculateSignedDistanceField of mesh m on a level set of
// (isize, jsize, ksize) cells at node offset o (meshlevelset.cpp:572-601);
// block = the multi-threaded band.  phi: device, (isize+1)(jsize+1)(ksize+1).
void levelSetFast(const HMesh &m, int isize, int jsize, int ksize, I3 o, double dx, bool block, RandJitter &rj,
                  float *phi, hipStream_t s) {
    const int pi = isize + 1, pj = jsize + 1, pk = ksize + 1, band = EXACT_BAND;
    const size_t nn = (size_t)pi * pj * pk;
    const float ub = (float)((pi + pj + pk) * dx);  // getDistanceUpperBound (meshlevelset.cpp:769-771)
    hipLaunchKernelGGL(fillFloatKernel, dim3(blocksFor(nn, 256)), dim3(256), 0, s, phi, nn, ub);
    const V3 po = indexToPos(o.i, o.j, o.k, dx);
    std::vector<BandTri> bt;
    bt.reserve(m.t.size());
    for (size_t t = 0; t < m.t.size(); t++) {
        const V3 a = sub(m.v[m.t[t][0]], po), b = sub(m.v[m.t[t][1]], po), c = sub(m.v[m.t[t][2]], po);
        BandTri T{a, b, c, {0, 0, 0}, {0, 0, 0}};
        if (block) {  // _initializeTriangleData (meshlevelset.cpp:878-918)
            Box tb = aabbTri(m, t);
            tb.pos = sub(tb.pos, po);
            I3 g0 = posToIndex(tb.pos, dx), g1 = posToIndex(tb.maxPoint(), dx);
            g0 = I3{g0.i - band, g0.j - band, g0.k - band};
            g1 = I3{g1.i + band + 1, g1.j + band + 1, g1.k + band + 1};
            if (g1.i < 0 || g0.i >= pi || g1.j < 0 || g0.j >= pj || g1.k < 0 || g0.k >= pk) continue;
            T.lo[0] = std::max(g0.i, 0);
            T.lo[1] = std::max(g0.j, 0);
            T.lo[2] = std::max(g0.k, 0);
            T.hi[0] = std::min(g1.i, isize);
            T.hi[1] = std::min(g1.j, jsize);
            T.hi[2] = std::min(g1.k, ksize);
        } else {  // _computeExactBandDistanceFieldSingleThreaded (1165-1234)
            const double inv = 1.0 / dx;
            const double f[3][3] = {{(double)a.x * inv, (double)a.y * inv, (double)a.z * inv},
                                    {(double)b.x * inv, (double)b.y * inv, (double)b.z * inv},
                                    {(double)c.x * inv, (double)c.y * inv, (double)c.z * inv}};
            const int dim[3] = {pi, pj, pk};
            for (int q = 0; q < 3; q++) {
                const int lo = int(fmin(f[0][q], fmin(f[1][q], f[2][q]))) - band;
                const int hi = int(fmax(f[0][q], fmax(f[1][q], f[2][q]))) + band + 1;
                T.lo[q] = std::min(std::max(lo, 0), dim[q] - 1);
                T.hi[q] = std::min(std::max(hi, 0), dim[q] - 1);
            }
        }
        bt.push_back(T);
    }
    if (!bt.empty()) {
        DevArray<BandTri> dbt(bt.size());
        upload(dbt, bt, s);
        if (block)
            hipLaunchKernelGGL(meshBandKernel<true>, dim3((unsigned)bt.size()), dim3(256), 0, s, dbt.p, pi, pj, dx,
                               (int *)phi);
        else
            hipLaunchKernelGGL(meshBandKernel<false>, dim3((unsigned)bt.size()), dim3(256), 0, s, dbt.p, pi, pj, dx,
                               (int *)phi);
        hipCheck(hipStreamSynchronize(s), "exact band");
    }
    // signs (_computeDistanceFieldSigns, meshlevelset.cpp:1298-1318, and the
    // grid form of the inside test, meshutils.cpp:382-437)
    HMesh local = m;
    translate(local, V3{-po.x, -po.y, -po.z});
    DevArray<uint8_t> inside(nn);
    hipCheck(hipMemsetAsync(inside.p, 0, nn, s), "hipMemsetAsync");
    const Box grid{V3{0.0f, 0.0f, 0.0f}, isize * dx, jsize * dx, ksize * dx};
    bool contained = true;
    for (const V3 &p : local.v)
        if (!grid.inside(p)) {
            contained = false;
            break;
        }
    if (contained) {
        markInside(local, dx, pi, pj, pk, rj, inside.p, s);
    } else {
        HMesh in;
        std::vector<HMesh> out;
        for (HMesh &I : splitIslands(local)) {
            const Box b = aabbOf(I.v.data(), I.v.size());
            if (boxContains(grid, b)) {
                const int base = (int)in.v.size();
                in.v.insert(in.v.end(), I.v.begin(), I.v.end());
                for (auto t : I.t) in.t.push_back({t[0] + base, t[1] + base, t[2] + base});
            } else if (boxIntersects(grid, b)) {
                out.push_back(std::move(I));
            }
        }
        markInside(in, dx, pi, pj, pk, rj, inside.p, s);
        for (HMesh &I : out) markInside(I, dx, pi, pj, pk, rj, inside.p, s);
    }
    hipLaunchKernelGGL(negateInsideKernel, dim3(blocksFor(nn, 256)), dim3(256), 0, s, phi, inside.p, nn);
    hipCheck(hipStreamSynchronize(s), "mesh level set");
}

HMesh hostMesh(const ObstacleMesh &o) {
    HMesh m;
    if (o.numVertices < 0 || o.numTriangles < 0) throw std::runtime_error("mesh obstacle: negative size");
    m.v.resize(o.numVertices);
    for (int q = 0; q < o.numVertices; q++)
        m.v[q] = V3{o.vertices[3 * q], o.vertices[3 * q + 1], o.vertices[3 * q + 2]};
    m.t.resize(o.numTriangles);
    for (int q = 0; q < o.numTriangles; q++)
        for (int c = 0; c < 3; c++) {
            const int v = o.triangles[3 * q + c];
            if (v < 0 || v >= o.numVertices) throw std::runtime_error("mesh obstacle: triangle index out of range");
            m.t[q][c] = v;
        }
    return m;
}

// MeshObject::getMeshLevelSet for a static mesh, unioned into levelset
// (meshobject.cpp:239-264, 633-701, 738-839); obj (optional) tracks the
// closest object beside it, this obstacle being objId
void obstacleLevelSet(const ObstacleMesh &o, int objId, Grid g, RandJitter &rj, float *levelset, int *obj,
                      hipStream_t s) {
    HMesh m = hostMesh(o);
    // getMesh(frameInterpolation): v + f (v - v), which turns -0 into +0
    for (V3 &p : m.v) p = add(p, V3{0.0f, 0.0f, 0.0f});
    {  // TriangleMesh::removeExtraneousVertices (trianglemesh.cpp:775-822)
        std::vector<char> isUsed(m.v.size(), 0);
        for (const auto &t : m.t) isUsed[t[0]] = isUsed[t[1]] = isUsed[t[2]] = 1;
        std::vector<int> tr(m.v.size(), -1);
        std::vector<V3> kept;
        for (size_t q = 0; q < m.v.size(); q++)
            if (isUsed[q]) {
                tr[q] = (int)kept.size();
                kept.push_back(m.v[q]);
            }
        for (auto &t : m.t)
            for (int c = 0; c < 3; c++) t[c] = tr[t[c]];
        m.v = kept;
    }
    if (m.t.empty()) return;
    const Box domain{V3{0.0f, 0.0f, 0.0f}, g.ni * g.dx, g.nj * g.dx, g.nk * g.dx};
    std::vector<HMesh> islands;
    for (HMesh &I : splitIslands(m)) {
        const Box b = aabbOf(I.v.data(), I.v.size());
        if (boxContains(domain, b) || boxIntersects(domain, b)) islands.push_back(std::move(I));
    }
    const bool fracture = islands.size() >= 25;  // _numIslandsForFractureOptimizationTrigger
    // that path's islands are first expanded (_expandMeshIslands,
    // meshobject.cpp:703-736): every vertex moves 0.5 x expansion along its
    // normalized offset from the island's float centroid
    if (fracture && std::fabs(o.meshExpansion) >= 1e-9f) {
        for (HMesh &I : islands) {
            V3 sum{0.0f, 0.0f, 0.0f};
            for (const V3 &p : I.v) sum = add(sum, p);
            const V3 centroid = divide3(sum, (float)I.v.size());
            const float expval = 0.5f * o.meshExpansion, eps = 1e-9f;
            for (V3 &p : I.v) {
                V3 d = sub(p, centroid);
                if (std::fabs(d.x) < eps && std::fabs(d.y) < eps && std::fabs(d.z) < eps) continue;
                p = add(p, scale(expval, normalize3(d)));
            }
        }
    }
    std::vector<HMesh> parts;
    if (fracture)  // the island work queue pops from its back (boundedbuffer.h:119-135)
        parts.assign(std::make_move_iterator(islands.rbegin()), std::make_move_iterator(islands.rend()));
    else
        parts.push_back(std::move(m));
    const int band = EXACT_BAND;
    for (const HMesh &part : parts) {
        const Box b = aabbOf(part.v.data(), part.v.size());
        I3 g0 = posToIndex(b.pos, g.dx), g1 = posToIndex(b.maxPoint(), g.dx);
        g0 = I3{(int)fmax(g0.i - band, 0), (int)fmax(g0.j - band, 0), (int)fmax(g0.k - band, 0)};
        g1 = I3{(int)fmin(g1.i + band + 1, g.ni - 1), (int)fmin(g1.j + band + 1, g.nj - 1),
                (int)fmin(g1.k + band + 1, g.nk - 1)};
        const int w = g1.i - g0.i, h = g1.j - g0.j, d = g1.k - g0.k;
        if (w <= 0 || h <= 0 || d <= 0) continue;
        const size_t nn = (size_t)(w + 1) * (h + 1) * (d + 1);
        DevArray<float> phi(nn);
        // fracture path: per-island level sets with multi-threading off (meshobject.cpp:798-839)
        levelSetFast(part, w, h, d, g0, g.dx, !fracture, rj, phi.p, s);
        const float ub = (float)(((w + 1) + (h + 1) + (d + 1)) * g.dx);  // the island set's upper bound
        hipLaunchKernelGGL(unionKernel, dim3(blocksFor(nn, 256)), dim3(256), 0, s, phi.p, w + 1, h + 1, d + 1, g0,
                           false, levelset, g.ni + 1, g.nj + 1, g.nk + 1, obj, (const int *)nullptr, objId, ub);
        hipCheck(hipStreamSynchronize(s), "obstacle union");
    }
}

// the domain object (0) on the domain SDF's exact band, none (-1) elsewhere
__global__ void domainObjectKernel(const float *solid, size_t n, float ub, int domain, int *obj) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) obj[i] = (domain && fabsf(solid[i]) < ub) ? 0 : -1;
}

}  // namespace

int meshLevelSetFastDevice(Grid cells, I3 offset, const ObstacleMesh &mesh, float *phi, hipStream_t s) {
    RandJitter rj;
    HMesh m = hostMesh(mesh);
    if (m.v.empty()) throw std::runtime_error("mesh level set: empty mesh");
    levelSetFast(m, cells.ni, cells.nj, cells.nk, offset, cells.dx, true, rj, phi, s);
    return rj.draws;
}

int meshObstacleSolid(Grid g, const ObstacleMesh *obs, int n, float *solid, int *nodeObject, bool domainObject,
                      hipStream_t s) {
    RandJitter rj;
    rj.next(g.dx);  // the domain box's own sign test draws first (fluidsimulation.cpp:5277-5287)
    const size_t nn = (size_t)(g.ni + 1) * (g.nj + 1) * (g.nk + 1);
    const float ubDomain = (float)(((g.ni + 1) + (g.nj + 1) + (g.nk + 1)) * g.dx);
    if (nodeObject)
        hipLaunchKernelGGL(domainObjectKernel, dim3(blocksFor(nn, 256)), dim3(256), 0, s, solid, nn, ubDomain,
                           domainObject ? 1 : 0, nodeObject);
    for (int q = 0; q < n; q++)
        if (!obs[q].inversed) obstacleLevelSet(obs[q], q + 1, g, rj, solid, nodeObject, s);
    bool anyInversed = false;
    for (int q = 0; q < n; q++) anyInversed |= obs[q].inversed;
    if (anyInversed) {
        // union of the inversed obstacles' level sets, negated, then unioned
        // into the solid (fluidsimulation.cpp:5321-5338)
        DevArray<float> inv(nn);
        DevArray<int> invObj(nodeObject ? nn : 0);
        hipLaunchKernelGGL(fillFloatKernel, dim3(blocksFor(nn, 256)), dim3(256), 0, s, inv.p, nn, ubDomain);
        if (nodeObject)
            hipLaunchKernelGGL(domainObjectKernel, dim3(blocksFor(nn, 256)), dim3(256), 0, s, inv.p, nn, ubDomain, 0,
                               invObj.p);
        for (int q = 0; q < n; q++)
            if (obs[q].inversed) obstacleLevelSet(obs[q], q + 1, g, rj, inv.p, invObj.p, s);
        hipLaunchKernelGGL(unionKernel, dim3(blocksFor(nn, 256)), dim3(256), 0, s, inv.p, g.ni + 1, g.nj + 1,
                           g.nk + 1, I3{0, 0, 0}, true, solid, g.ni + 1, g.nj + 1, g.nk + 1, nodeObject,
                           (const int *)invObj.p, 0, 0.0f);
        hipCheck(hipStreamSynchronize(s), "inversed obstacles");
    }
    hipCheck(hipGetLastError(), "mesh obstacle kernels");
    return rj.draws;
}

}  // namespace ff
