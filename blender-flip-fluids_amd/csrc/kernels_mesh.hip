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
//
// With velocity data (moving obstacles) every level set also keeps each
// node's closest triangle and, per face, the weight (the fraction of the face
// inside the level set) times the velocity of the nearest point of the mesh,
// interpolated from the vertex velocities (meshlevelset.cpp:168-205,
// 1320-1425, 1568-1717), and the solid sums these over its parts.
#include <algorithm>
#include <array>
#include <cmath>
#include <cstdlib>
#include <cstring>
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
    int id;  // the triangle's index in its mesh
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

// the band with each node's closest triangle: a 64-bit atomic min over
// (distance bits, triangle index), so the smallest distance wins and, among
// equal distances, the lowest triangle index, as the reference's ascending
// triangle loop with a strict comparison keeps the first
template <bool BLOCK>
__global__ void meshBandTriKernel(const BandTri *tris, int pi, int pj, double dx, unsigned long long *keys) {
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
        const unsigned long long key = ((unsigned long long)__float_as_uint(dist) << 32) | (unsigned)t.id;
        atomicMin(keys + flat(i, j, k, pi, pj), key);
    }
}
__global__ void fillKeysKernel(unsigned long long *keys, size_t n, unsigned long long v) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) keys[i] = v;
}
// keys -> phi (the distance) and the closest triangle (-1: none)
__global__ void splitKeysKernel(const unsigned long long *keys, size_t n, float *phi, int *tri) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const unsigned long long key = keys[i];
    phi[i] = __uint_as_float((unsigned)(key >> 32));
    const unsigned t = (unsigned)(key & 0xffffffffull);
    tri[i] = t == 0xffffffffu ? -1 : (int)t;
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

__global__ void fillIntKernel(int *a, size_t n, int v) {
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
// srcTri/dstTri (optional): the closest triangles, merged as the objects are,
// the source's indices shifted by triOffset (its triangles follow dst's)
__global__ void unionKernel(const float *src, int qi, int qj, int qk, I3 o, bool negate, float *dst, int pi, int pj,
                            int pk, int *dstObj, const int *srcObjArr, int srcObj, float srcUb,
                            const int *srcTri = nullptr, int *dstTri = nullptr, int triOffset = 0) {
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
        if (dstTri && srcTri[idx] != -1 && fabsf(v) < fabsf(d)) dstTri[t] = srcTri[idx] + triOffset;
        d = v;
    }
}

__global__ void negateKernel(float *phi, size_t n) {  // MeshLevelSet::negate (meshlevelset.cpp:702-708)
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) phi[i] = -phi[i];
}

// ---- the velocity data of a level set (meshlevelset.cpp:168-205, 1320-1425, 1568-1717) ----

// one triangle with its vertices' velocities, in world coordinates
struct VelTri {
    V3 a, b, c, va, vb, vc;
};
// a level set's view for the nearest velocity: its grid and offset, phi,
// closest triangles and mesh
struct VelLS {
    int ni, nj, nk;
    I3 off;
    double dx;
    float ub;
    const float *phi;
    const int *tri;
    const VelTri *tris;
    int isStatic;
};

// _pointToSegmentVelocity (meshlevelset.cpp:1697-1717)
__device__ V3 segmentVelocity(V3 x0, V3 x1, V3 x2, V3 v1, V3 v2, float *distance) {
    const V3 d = sub(x2, x1);
    const double m2 = (double)dot3(d, d);
    float s12 = (float)((double)dot3(sub(x2, x0), d) / m2);
    if (s12 < 0)
        s12 = 0;
    else if (s12 > 1)
        s12 = 1;
    *distance = len3(sub(x0, add(scale(s12, x1), scale(1 - s12, x2))));
    return add(scale(s12, v1), scale(1 - s12, v2));
}

// _pointToTriangleVelocity (meshlevelset.cpp:1568-1639), x0 and the vertices
// in the level set's frame
__device__ V3 triangleVelocity(const VelTri &T, V3 po, V3 x0) {
    const V3 v1 = T.va, v2 = T.vb, v3 = T.vc;
    const float eps = 1e-6f;
    if (fabsf(v1.x) < eps && fabsf(v1.y) < eps && fabsf(v1.z) < eps && fabsf(v2.x) < eps && fabsf(v2.y) < eps &&
        fabsf(v2.z) < eps && fabsf(v3.x) < eps && fabsf(v3.y) < eps && fabsf(v3.z) < eps)
        return V3{0.0f, 0.0f, 0.0f};
    const V3 x1 = sub(T.a, po), x2 = sub(T.b, po), x3 = sub(T.c, po);
    const V3 x13 = sub(x1, x3), x23 = sub(x2, x3), x03 = sub(x0, x3);
    const float m13 = dot3(x13, x13), m23 = dot3(x23, x23), d = dot3(x13, x23);
    const float invdet = 1.0f / fmaxf(m13 * m23 - d * d, 1e-30f);
    const float a = dot3(x13, x03), b = dot3(x23, x03);
    const float w23 = invdet * (m23 * a - d * b);
    const float w31 = invdet * (m13 * b - d * a);
    const float w12 = 1 - w23 - w31;
    if (w23 >= 0 && w31 >= 0 && w12 >= 0) return add(add(scale(w23, v1), scale(w31, v2)), scale(w12, v3));
    float d1, d2;
    V3 e1, e2;
    if (w23 > 0) {
        e1 = segmentVelocity(x0, x1, x2, v1, v2, &d1);
        e2 = segmentVelocity(x0, x1, x3, v1, v3, &d2);
    } else if (w31 > 0) {
        e1 = segmentVelocity(x0, x1, x2, v1, v2, &d1);
        e2 = segmentVelocity(x0, x2, x3, v2, v3, &d2);
    } else {
        e1 = segmentVelocity(x0, x1, x3, v1, v3, &d1);
        e2 = segmentVelocity(x0, x2, x3, v2, v3, &d2);
    }
    return d1 < d2 ? e1 : e2;
}

// getNearestVelocity (meshlevelset.cpp:168-205): p in world coordinates
__device__ V3 nearestVelocity(const VelLS &L, V3 p) {
    const V3 po = indexToPos(L.off.i, L.off.j, L.off.k, L.dx);
    p = sub(p, po);
    const I3 g = posToIndex(p, L.dx);
    const int o8[8][3] = {{0, 0, 0}, {1, 0, 0}, {1, 0, 1}, {0, 0, 1}, {0, 1, 0}, {1, 1, 0}, {1, 1, 1}, {0, 1, 1}};
    int nearest = -1;
    float best = L.ub;
    for (int q = 0; q < 8; q++) {
        const int i = g.i + o8[q][0], j = g.j + o8[q][1], k = g.k + o8[q][2];
        if (!inRange(i, j, k, L.ni + 1, L.nj + 1, L.nk + 1)) continue;
        const int t = L.tri[flat(i, j, k, L.ni + 1, L.nj + 1)];
        if (t == -1) continue;
        const VelTri T = L.tris[t];
        const V3 cp = closestPointOnTriangle(p, sub(T.a, po), sub(T.b, po), sub(T.c, po));
        const float d = len3(sub(cp, p));
        if (d < best) {
            best = d;
            nearest = t;
        }
    }
    if (nearest == -1) return V3{0.0f, 0.0f, 0.0f};
    return triangleVelocity(L.tris[nearest], po, p);
}

// _computeVelocityGridThread (meshlevelset.cpp:1320-1374) for one face
// direction, each face's weight x velocity and weight added into the domain
// accumulators at the level set's offset (calculateUnion, 1759-1845: the
// faces inside both arrays)
template <int DIR>
__global__ void velocityGridKernel(VelLS L, Grid g, float *df, float *dw) {
    const int fi = L.ni + (DIR == 0), fj = L.nj + (DIR == 1), fk = L.nk + (DIR == 2);
    const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (size_t)fi * fj * fk) return;
    const int i = (int)(idx % fi), j = (int)((idx / fi) % fj), k = (int)(idx / ((size_t)fi * fj));
    const int a = i + L.off.i, b = j + L.off.j, c = k + L.off.k;
    if (!inRange(a, b, c, g.ni + (DIR == 0), g.nj + (DIR == 1), g.nk + (DIR == 2))) return;
    const int pi = L.ni + 1, pj = L.nj + 1;
    auto P = [&](int x, int y, int z) { return L.phi[flat(x, y, z, pi, pj)]; };
    float weight;
    if (DIR == 0)
        weight = fraction4(P(i, j, k), P(i, j + 1, k), P(i, j, k + 1), P(i, j + 1, k + 1));
    else if (DIR == 1)
        weight = fraction4(P(i, j, k), P(i, j, k + 1), P(i + 1, j, k), P(i + 1, j, k + 1));
    else
        weight = fraction4(P(i, j, k), P(i, j + 1, k), P(i + 1, j, k), P(i + 1, j + 1, k));
    if (!(weight > 0.0f)) return;
    float comp = 0.0f;
    if (!L.isStatic) {
        const double dx = L.dx;
        V3 fp;
        if (DIR == 0)
            fp = mk((float)i * dx, ((float)j + 0.5) * dx, ((float)k + 0.5) * dx);
        else if (DIR == 1)
            fp = mk(((float)i + 0.5) * dx, (float)j * dx, ((float)k + 0.5) * dx);
        else
            fp = mk(((float)i + 0.5) * dx, ((float)j + 0.5) * dx, (float)k * dx);
        const V3 v = nearestVelocity(L, add(fp, indexToPos(L.off.i, L.off.j, L.off.k, dx)));
        comp = DIR == 0 ? v.x : (DIR == 1 ? v.y : v.z);
    }
    const size_t t = flat(a, b, c, g.ni + (DIR == 0), g.nj + (DIR == 1));
    df[t] += weight * comp;
    dw[t] += weight;
}

__global__ void addArraysKernel(float *dst, const float *src, size_t n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] += src[i];
}

// normalizeVelocityGrid's per-face step (meshlevelset.cpp:1739-1757)
__global__ void normalizeVelocityKernel(const float *f, const float *w, size_t n, float *out, uint8_t *valid) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float wt = w[i];
    const bool ok = wt > 1e-6f;
    out[i] = ok ? f[i] / wt : 0.0f;
    valid[i] = ok ? 1 : 0;
}

unsigned blocksFor(size_t n, unsigned t) { return (unsigned)((n + t - 1) / t); }

// ------------------------------------------------------------ host ----------

unsigned floatKeyBits(float f) {
    unsigned u;
    std::memcpy(&u, &f, sizeof(u));
    return u;
}

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
    if (h.size() > d.n) throw std::runtime_error("mesh obstacle: upload larger than its device array");
    if (!h.empty())
        hipCheck(hipMemcpyAsync(d.p, h.data(), sizeof(T) * h.size(), hipMemcpyHostToDevice, s), "upload mesh data");
}

struct HMesh {
    std::vector<V3> v;
    std::vector<std::array<int, 3>> t;
    std::vector<V3> vel;  // per-vertex velocities, or empty: at rest
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
                // the island's vertex velocities follow its vertices (meshutils.cpp:574-593)
                if (!m.vel.empty()) I.vel.push_back(m.vel[t[c]]);
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
// tri (optional, device, nodes): each node's closest triangle, -1 outside the band.
void levelSetFast(const HMesh &m, int isize, int jsize, int ksize, I3 o, double dx, bool block, RandJitter &rj,
                  float *phi, hipStream_t s, int *tri = nullptr) {
    const int pi = isize + 1, pj = jsize + 1, pk = ksize + 1, band = EXACT_BAND;
    const size_t nn = (size_t)pi * pj * pk;
    const float ub = (float)((pi + pj + pk) * dx);  // getDistanceUpperBound (meshlevelset.cpp:769-771)
    hipLaunchKernelGGL(fillFloatKernel, dim3(blocksFor(nn, 256)), dim3(256), 0, s, phi, nn, ub);
    const V3 po = indexToPos(o.i, o.j, o.k, dx);
    std::vector<BandTri> bt;
    bt.reserve(m.t.size());
    for (size_t t = 0; t < m.t.size(); t++) {
        const V3 a = sub(m.v[m.t[t][0]], po), b = sub(m.v[m.t[t][1]], po), c = sub(m.v[m.t[t][2]], po);
        BandTri T{a, b, c, {0, 0, 0}, {0, 0, 0}, (int)t};
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
    if (tri) {
        DevArray<unsigned long long> keys(nn);
        const unsigned long long none = ((unsigned long long)floatKeyBits(ub) << 32) | 0xffffffffull;
        hipLaunchKernelGGL(fillKeysKernel, dim3(blocksFor(nn, 256)), dim3(256), 0, s, keys.p, nn, none);
        if (!bt.empty()) {
            DevArray<BandTri> dbt(bt.size());
            upload(dbt, bt, s);
            if (block)
                hipLaunchKernelGGL(meshBandTriKernel<true>, dim3((unsigned)bt.size()), dim3(256), 0, s, dbt.p, pi, pj,
                                   dx, keys.p);
            else
                hipLaunchKernelGGL(meshBandTriKernel<false>, dim3((unsigned)bt.size()), dim3(256), 0, s, dbt.p, pi,
                                   pj, dx, keys.p);
            hipCheck(hipStreamSynchronize(s), "exact band");
        }
        hipLaunchKernelGGL(splitKeysKernel, dim3(blocksFor(nn, 256)), dim3(256), 0, s, keys.p, nn, phi, tri);
        hipCheck(hipStreamSynchronize(s), "exact band");
    } else if (!bt.empty()) {
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
    m.vel.assign(o.numVertices, V3{0.0f, 0.0f, 0.0f});
    if (o.velocities)
        for (int q = 0; q < o.numVertices; q++)
            m.vel[q] = V3{o.velocities[3 * q], o.velocities[3 * q + 1], o.velocities[3 * q + 2]};
    m.t.resize(o.numTriangles);
    for (int q = 0; q < o.numTriangles; q++)
        for (int c = 0; c < 3; c++) {
            const int v = o.triangles[3 * q + c];
            if (v < 0 || v >= o.numVertices) throw std::runtime_error("mesh obstacle: triangle index out of range");
            m.t[q][c] = v;
        }
    return m;
}

// A domain-sized level set that takes an obstacle's parts with velocity data:
// phi, closest objects and (optional) closest triangles, indices into `tris`,
// the triangles of the parts unioned so far; the face accumulators f / w
// (optional) receive the parts' velocity grids.
struct VelTarget {
    float *phi;
    int *obj;
    int *tri;
    std::vector<VelTri> *tris;
    float *const *f;
    float *const *w;
};

// MeshObject::getMeshLevelSet, unioned into levelset (meshobject.cpp:239-264,
// 633-701, 738-839); obj (optional) tracks the closest object beside it, this
// obstacle being objId.  With vt, the parts go into vt's level set instead,
// with their velocity data.
void obstacleLevelSet(const ObstacleMesh &o, int objId, Grid g, RandJitter &rj, float *levelset, int *obj,
                      hipStream_t s, const VelTarget *vt = nullptr) {
    HMesh m = hostMesh(o);
    // getMesh(frameInterpolation): v + f (v - v), which turns -0 into +0
    for (V3 &p : m.v) p = add(p, V3{0.0f, 0.0f, 0.0f});
    {  // TriangleMesh::removeExtraneousVertices (trianglemesh.cpp:775-822)
        std::vector<char> isUsed(m.v.size(), 0);
        for (const auto &t : m.t) isUsed[t[0]] = isUsed[t[1]] = isUsed[t[2]] = 1;
        std::vector<int> tr(m.v.size(), -1);
        std::vector<V3> kept, keptVel;
        for (size_t q = 0; q < m.v.size(); q++)
            if (isUsed[q]) {
                tr[q] = (int)kept.size();
                kept.push_back(m.v[q]);
                keptVel.push_back(m.vel[q]);
            }
        for (auto &t : m.t)
            for (int c = 0; c < 3; c++) t[c] = tr[t[c]];
        m.v = kept;
        m.vel = keptVel;
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
        const float ub = (float)(((w + 1) + (h + 1) + (d + 1)) * g.dx);  // the island set's upper bound
        if (!vt) {
            // fracture path: per-island level sets with multi-threading off (meshobject.cpp:798-839)
            levelSetFast(part, w, h, d, g0, g.dx, !fracture, rj, phi.p, s);
            hipLaunchKernelGGL(unionKernel, dim3(blocksFor(nn, 256)), dim3(256), 0, s, phi.p, w + 1, h + 1, d + 1,
                               g0, false, levelset, g.ni + 1, g.nj + 1, g.nk + 1, obj, (const int *)nullptr, objId,
                               ub);
            hipCheck(hipStreamSynchronize(s), "obstacle union");
            continue;
        }
        // the part's triangles with their vertices' velocities; a part whose
        // vertices all stand still has zero velocity grids (meshlevelset.cpp:1376-1425)
        std::vector<VelTri> vts(part.t.size());
        bool isStatic = true;
        for (const V3 &v : part.vel)
            if (std::fabs(v.x) > 0.0f || std::fabs(v.y) > 0.0f || std::fabs(v.z) > 0.0f) isStatic = false;
        for (size_t t = 0; t < part.t.size(); t++) {
            const auto &T = part.t[t];
            vts[t] = VelTri{part.v[T[0]], part.v[T[1]], part.v[T[2]], part.vel[T[0]], part.vel[T[1]], part.vel[T[2]]};
        }
        const bool faces = vt->f != nullptr, needTri = vt->tri != nullptr || (faces && !isStatic);
        DevArray<int> tri(needTri ? nn : 0);
        levelSetFast(part, w, h, d, g0, g.dx, !fracture, rj, phi.p, s, needTri ? tri.p : nullptr);
        DevArray<VelTri> dvt(faces && !isStatic ? vts.size() : 0);
        if (faces) {
            if (!isStatic) upload(dvt, vts, s);
            const VelLS L{w, h, d, g0, g.dx, ub, phi.p, tri.p, dvt.p, isStatic ? 1 : 0};
            hipLaunchKernelGGL(velocityGridKernel<0>, dim3(blocksFor((size_t)(w + 1) * h * d, 256)), dim3(256), 0, s,
                               L, g, vt->f[0], vt->w[0]);
            hipLaunchKernelGGL(velocityGridKernel<1>, dim3(blocksFor((size_t)w * (h + 1) * d, 256)), dim3(256), 0, s,
                               L, g, vt->f[1], vt->w[1]);
            hipLaunchKernelGGL(velocityGridKernel<2>, dim3(blocksFor((size_t)w * h * (d + 1), 256)), dim3(256), 0, s,
                               L, g, vt->f[2], vt->w[2]);
        }
        const int triOffset = vt->tris ? (int)vt->tris->size() : 0;
        hipLaunchKernelGGL(unionKernel, dim3(blocksFor(nn, 256)), dim3(256), 0, s, phi.p, w + 1, h + 1, d + 1, g0,
                           false, vt->phi, g.ni + 1, g.nj + 1, g.nk + 1, vt->obj, (const int *)nullptr, objId, ub,
                           vt->tri ? (const int *)tri.p : nullptr, vt->tri, triOffset);
        if (vt->tris) vt->tris->insert(vt->tris->end(), vts.begin(), vts.end());
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

namespace {

size_t faceCount(Grid g, int d) { return (size_t)(g.ni + (d == 0)) * (g.nj + (d == 1)) * (g.nk + (d == 2)); }

void fillF(float *a, size_t n, float v, hipStream_t s) {
    if (n) hipLaunchKernelGGL(fillFloatKernel, dim3(blocksFor(n, 256)), dim3(256), 0, s, a, n, v);
}
void fillI(int *a, size_t n, int v, hipStream_t s) {
    if (n) hipLaunchKernelGGL(fillIntKernel, dim3(blocksFor(n, 256)), dim3(256), 0, s, a, n, v);
}

// the velocity grids of a domain-sized level set (phi, closest triangles into
// tris) added into f / w (zeroed first: the level set's own grids)
void domainVelocityGrids(Grid g, const float *phi, const int *tri, const std::vector<VelTri> &tris, float *const *f,
                         float *const *w, hipStream_t s) {
    bool isStatic = true;
    for (const VelTri &T : tris)
        for (const V3 &v : {T.va, T.vb, T.vc})
            if (std::fabs(v.x) > 0.0f || std::fabs(v.y) > 0.0f || std::fabs(v.z) > 0.0f) isStatic = false;
    DevArray<VelTri> dvt(isStatic ? 0 : tris.size());
    upload(dvt, isStatic ? std::vector<VelTri>() : tris, s);
    const float ub = (float)(((g.ni + 1) + (g.nj + 1) + (g.nk + 1)) * g.dx);
    const VelLS L{g.ni, g.nj, g.nk, I3{0, 0, 0}, g.dx, ub, phi, tri, dvt.p, isStatic ? 1 : 0};
    for (int d = 0; d < 3; d++) {
        fillF(f[d], faceCount(g, d), 0.0f, s);
        fillF(w[d], faceCount(g, d), 0.0f, s);
    }
    hipLaunchKernelGGL(velocityGridKernel<0>, dim3(blocksFor(faceCount(g, 0), 256)), dim3(256), 0, s, L, g, f[0], w[0]);
    hipLaunchKernelGGL(velocityGridKernel<1>, dim3(blocksFor(faceCount(g, 1), 256)), dim3(256), 0, s, L, g, f[1], w[1]);
    hipLaunchKernelGGL(velocityGridKernel<2>, dim3(blocksFor(faceCount(g, 2), 256)), dim3(256), 0, s, L, g, f[2], w[2]);
    hipCheck(hipStreamSynchronize(s), "velocity grids");
}

void addFaces(Grid g, float *const *dst, float *const *src, hipStream_t s) {
    for (int d = 0; d < 3; d++)
        hipLaunchKernelGGL(addArraysKernel, dim3(blocksFor(faceCount(g, d), 256)), dim3(256), 0, s, dst[d], src[d],
                           faceCount(g, d));
}

// FluidSimulation::_updateSolidLevelSet with velocity data
// (fluidsimulation.cpp:5219-5426; MeshLevelSet::calculateUnion, negate and
// the velocity grids, meshlevelset.cpp:577-718, 1320-1425, 1759-1845)
void solidWithVelocity(Grid g, const ObstacleMesh *obs, int n, float *solid, int *nodeObject, RandJitter &rj,
                       SolidVelocityData &V, hipStream_t s) {
    const size_t nn = (size_t)(g.ni + 1) * (g.nj + 1) * (g.nk + 1);
    const float ubDomain = (float)(((g.ni + 1) + (g.nj + 1) + (g.nk + 1)) * g.dx);
    // the domain: its negated box level set's velocity grids (zero velocity, the faces' weights)
    domainVelocityGrids(g, solid, nullptr, {}, V.f, V.w, s);
    for (int pass = 0; pass < 2; pass++) {  // static objects, then animated ones (5387-5426)
        const bool animated = pass == 1;
        for (int q = 0; q < n; q++) {
            if (obs[q].animated != animated || obs[q].inversed) continue;
            // the object's own level set (_tempSolidSDF), reset, then unioned into the solid
            fillF(V.tphi, nn, ubDomain, s);
            fillI(V.tobj, nn, -1, s);
            for (int d = 0; d < 3; d++) {
                fillF(V.tf[d], faceCount(g, d), 0.0f, s);
                fillF(V.tw[d], faceCount(g, d), 0.0f, s);
            }
            const VelTarget vt{V.tphi, V.tobj, nullptr, nullptr, V.tf, V.tw};
            obstacleLevelSet(obs[q], q + 1, g, rj, nullptr, nullptr, s, &vt);
            hipLaunchKernelGGL(unionKernel, dim3(blocksFor(nn, 256)), dim3(256), 0, s, V.tphi, g.ni + 1, g.nj + 1,
                               g.nk + 1, I3{0, 0, 0}, false, solid, g.ni + 1, g.nj + 1, g.nk + 1, nodeObject,
                               (const int *)V.tobj, 0, 0.0f, nullptr, nullptr, 0);
            addFaces(g, V.f, V.tf, s);
            addFaces(g, V.w, V.tw, s);
        }
        bool anyInversed = false;
        for (int q = 0; q < n; q++) anyInversed |= obs[q].animated == animated && obs[q].inversed;
        if (!anyInversed) continue;
        // the inversed objects: their union without velocity data, negated
        // (which computes its velocity grids), then unioned into the solid
        std::vector<VelTri> invTris, tempTris;
        fillF(V.iphi, nn, ubDomain, s);
        fillI(V.iobj, nn, -1, s);
        fillI(V.itri, nn, -1, s);
        for (int q = 0; q < n; q++) {
            if (obs[q].animated != animated || !obs[q].inversed) continue;
            fillF(V.tphi, nn, ubDomain, s);
            fillI(V.tobj, nn, -1, s);
            fillI(V.ttri, nn, -1, s);
            tempTris.clear();
            const VelTarget vt{V.tphi, V.tobj, V.ttri, &tempTris, nullptr, nullptr};
            obstacleLevelSet(obs[q], q + 1, g, rj, nullptr, nullptr, s, &vt);
            hipLaunchKernelGGL(unionKernel, dim3(blocksFor(nn, 256)), dim3(256), 0, s, V.tphi, g.ni + 1, g.nj + 1,
                               g.nk + 1, I3{0, 0, 0}, false, V.iphi, g.ni + 1, g.nj + 1, g.nk + 1, V.iobj,
                               (const int *)V.tobj, 0, 0.0f, (const int *)V.ttri, V.itri, (int)invTris.size());
            invTris.insert(invTris.end(), tempTris.begin(), tempTris.end());
            hipCheck(hipStreamSynchronize(s), "inversed union");
        }
        hipLaunchKernelGGL(negateKernel, dim3(blocksFor(nn, 256)), dim3(256), 0, s, V.iphi, nn);
        domainVelocityGrids(g, V.iphi, V.itri, invTris, V.tf, V.tw, s);
        hipLaunchKernelGGL(unionKernel, dim3(blocksFor(nn, 256)), dim3(256), 0, s, V.iphi, g.ni + 1, g.nj + 1,
                           g.nk + 1, I3{0, 0, 0}, false, solid, g.ni + 1, g.nj + 1, g.nk + 1, nodeObject,
                           (const int *)V.iobj, 0, 0.0f, nullptr, nullptr, 0);
        addFaces(g, V.f, V.tf, s);
        addFaces(g, V.w, V.tw, s);
    }
    hipCheck(hipStreamSynchronize(s), "solid velocity data");
}

}  // namespace

size_t solidVelocityBytes(Grid g) {
    const size_t nn = (size_t)(g.ni + 1) * (g.nj + 1) * (g.nk + 1);
    auto up = [](size_t b) { return (b + 255) / 256 * 256; };
    size_t b = 0;
    for (int d = 0; d < 3; d++) b += 4 * up(sizeof(float) * faceCount(g, d));
    return b + 2 * up(sizeof(float) * nn) + 4 * up(sizeof(int) * nn);
}

SolidVelocityData solidVelocityCarve(Grid g, void *block) {
    const size_t nn = (size_t)(g.ni + 1) * (g.nj + 1) * (g.nk + 1);
    char *p = (char *)block;
    auto take = [&](size_t bytes) {
        char *q = p;
        p += (bytes + 255) / 256 * 256;
        return q;
    };
    SolidVelocityData V{};
    for (int d = 0; d < 3; d++) {
        V.f[d] = (float *)take(sizeof(float) * faceCount(g, d));
        V.w[d] = (float *)take(sizeof(float) * faceCount(g, d));
        V.tf[d] = (float *)take(sizeof(float) * faceCount(g, d));
        V.tw[d] = (float *)take(sizeof(float) * faceCount(g, d));
    }
    V.tphi = (float *)take(sizeof(float) * nn);
    V.iphi = (float *)take(sizeof(float) * nn);
    V.tobj = (int *)take(sizeof(int) * nn);
    V.iobj = (int *)take(sizeof(int) * nn);
    V.ttri = (int *)take(sizeof(int) * nn);
    V.itri = (int *)take(sizeof(int) * nn);
    return V;
}

int meshObstacleSolid(Grid g, const ObstacleMesh *obs, int n, float *solid, int *nodeObject, bool domainObject,
                      hipStream_t s, SolidVelocityData *velocity) {
    RandJitter rj;
    rj.next(g.dx);  // the domain box's own sign test draws first (fluidsimulation.cpp:5277-5287)
    const size_t nn = (size_t)(g.ni + 1) * (g.nj + 1) * (g.nk + 1);
    const float ubDomain = (float)(((g.ni + 1) + (g.nj + 1) + (g.nk + 1)) * g.dx);
    if (nodeObject)
        hipLaunchKernelGGL(domainObjectKernel, dim3(blocksFor(nn, 256)), dim3(256), 0, s, solid, nn, ubDomain,
                           domainObject ? 1 : 0, nodeObject);
    if (velocity) {
        solidWithVelocity(g, obs, n, solid, nodeObject, rj, *velocity, s);
        hipCheck(hipGetLastError(), "mesh obstacle kernels");
        return rj.draws;
    }
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

void normalizeSolidVelocity(Grid g, const SolidVelocityData &d, float *us, float *vs, float *ws, uint8_t *valid,
                            uint8_t *status, hipStream_t s) {
    float *out[3] = {us, vs, ws};
    const int wi[3] = {g.ni + 1, g.ni, g.ni}, wj[3] = {g.nj, g.nj + 1, g.nj}, wk[3] = {g.nk, g.nk, g.nk + 1};
    const int layers = 5;  // MeshLevelSet::_numVelocityExtrapolationLayers (meshlevelset.h:363)
    for (int q = 0; q < 3; q++) {
        const size_t nf = faceCount(g, q);
        hipLaunchKernelGGL(normalizeVelocityKernel, dim3(blocksFor(nf, 256)), dim3(256), 0, s, d.f[q], d.w[q], nf,
                           out[q], valid);
        extrapolateInit(wi[q], wj[q], wk[q], valid, status, s);
        for (int l = 0; l < layers; l++) extrapolateLayer(wi[q], wj[q], wk[q], out[q], status, l, s);
    }
    hipCheck(hipGetLastError(), "normalize solid velocity");
}

}  // namespace ff
