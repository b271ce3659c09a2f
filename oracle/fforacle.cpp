/*
 * fforacle.cpp -- CPU oracle for the FLIP substep path.
 *
 * TEST INFRASTRUCTURE ONLY: a from-scratch restatement of the reference
 * engine's per-substep algorithms (rlguy/Blender-FLIP-Fluids 1.8.5).  Only
 * tests, the smoke check and the benchmark's CPU-baseline leg load it; the
 * shipped HIP engine never does.
 *
 * Threads (oracle_set_threads, default 1) split only loops whose outputs are
 * disjoint and whose per-output operation order is the serial one: the
 * liquid SDF and P2G per 10^3 block (each block's member particles in
 * ascending index, as the serial loop visits them), SpMV and vector updates
 * per row, G2P and advection per particle.  The results are therefore the
 * same bits at any thread count; the MIC(0) factor and sweeps, the dot
 * products and the removal pass stay serial.
 *
 * Every function names the reference lines it restates.  The restatement
 * keeps the reference's float/double promotion points and operation order
 * (x86-64 SSE2, no FMA contraction: built with -ffp-contract=off) so that,
 * given the same inputs, it reproduces the reference bit for bit.  Two facts
 * about the plain configuration are used and stated where they apply:
 *   - solids are static, so solid face velocities are exactly zero and the
 *     divergence's solid terms (pressuresolver.cpp:652-657) add +-0.0 to a
 *     sum that starts at +0.0, which never changes it;
 *   - the 7-point stencil never gives two rows of one MIC(0) column that are
 *     neighbours, so the factorisation reduces to its "pull" form over the
 *     -z, -y, -x neighbours (pcgsolver.h:68-205).
 * Parity of every function is pinned by the reference build in oracle/_ref
 * (tests/test_oracle_vs_reference.py) and by committed golden vectors.
 */
#include "fforacle.h"

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <stdexcept>
#include <vector>

#include <omp.h>

namespace {

// ---------------------------------------------------------------- vec3 ----
// vmath::vec3 semantics: three floats, component-wise float arithmetic
// (vmath.h:37-91, vmath.cpp:51-120).
struct V3 {
    float x, y, z;
};
inline V3 mk(double x, double y, double z) { return V3{(float)x, (float)y, (float)z}; }
inline V3 operator+(V3 a, V3 b) { return V3{a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V3 operator-(V3 a, V3 b) { return V3{a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 operator*(float s, V3 a) { return V3{s * a.x, s * a.y, s * a.z}; }
inline float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline float lengthsq(V3 a) { return a.x * a.x + a.y * a.y + a.z * a.z; }
inline float length(V3 a) { return std::sqrt(lengthsq(a)); }
// vmath::operator/(vec3, float): multiply by the float reciprocal of a double division
inline V3 divide(V3 a, float s) {
    float inv = 1.0 / s;
    return V3{a.x * inv, a.y * inv, a.z * inv};
}
inline V3 normalize(V3 a) { return divide(a, length(a)); }

struct GI {
    int i, j, k;
};

// Grid3d helpers (grid3d.h:32-124)
inline GI posToIndex(V3 p, double dx) {
    double inv = 1.0 / dx;
    return GI{(int)std::floor(p.x * inv), (int)std::floor(p.y * inv), (int)std::floor(p.z * inv)};
}
inline GI posToIndexD(double x, double y, double z, double dx) {
    double inv = 1.0 / dx;
    return GI{(int)std::floor(x * inv), (int)std::floor(y * inv), (int)std::floor(z * inv)};
}
inline V3 indexToPos(int i, int j, int k, double dx) {
    return mk((float)i * dx, (float)j * dx, (float)k * dx);
}
inline V3 indexToCellCenter(int i, int j, int k, double dx) {
    double hw = 0.5 * dx;
    return mk((float)i * dx + hw, (float)j * dx + hw, (float)k * dx + hw);
}
inline bool inRange(int i, int j, int k, int ni, int nj, int nk) {
    return i >= 0 && j >= 0 && k >= 0 && i < ni && j < nj && k < nk;
}
inline size_t flat(int i, int j, int k, int ni, int nj) {
    return (size_t)i + (size_t)ni * ((size_t)j + (size_t)nj * (size_t)k);
}

// ---------------------------------------------------------------- AABB ----
// aabb.cpp:32-38, 122-133, 489-518
struct Box {
    V3 pos;
    double w, h, d;
    void expand(double v) {
        double hh = 0.5 * v;
        pos = pos - mk(hh, hh, hh);
        w += v;
        h += v;
        d += v;
    }
    bool inside(V3 p) const {
        return p.x >= pos.x && p.y >= pos.y && p.z >= pos.z && p.x < pos.x + w &&
               p.y < pos.y + h && p.z < pos.z + d;
    }
    V3 maxPoint() const { return pos + mk(w, h, d); }
    V3 nearestInside(V3 p) const {
        if (inside(p)) return p;
        double eps = 1e-6;
        V3 mn = pos, mx = maxPoint();
        p.x = std::fmax(p.x, mn.x);
        p.y = std::fmax(p.y, mn.y);
        p.z = std::fmax(p.z, mn.z);
        p.x = std::fmin(p.x, mx.x - eps);
        p.y = std::fmin(p.y, mx.y - eps);
        p.z = std::fmin(p.z, mx.z - eps);
        return p;
    }
};

// FluidSimulation::_getBoundaryAABB (fluidsimulation.cpp:5175-5180)
Box boundaryBox(int ni, int nj, int nk, double dx) {
    Box b{V3{0.0f, 0.0f, 0.0f}, ni * dx, nj * dx, nk * dx};
    b.expand(-3 * dx - 1e-4);
    return b;
}

// ------------------------------------------------------- solid box SDF ----
// Collision::findClosestPointOnTriangle (collision.cpp:161-268)
double clampd(double v, double lo, double hi) {
    if (v < lo) v = lo;
    if (v > hi) v = hi;
    return v;
}
V3 closestPointOnTriangle(V3 p0, V3 v0, V3 v1, V3 v2) {
    V3 edge0 = v1 - v0;
    V3 edge1 = v2 - v0;
    V3 pv = v0 - p0;
    double a = dot(edge0, edge0);
    double b = dot(edge0, edge1);
    double c = dot(edge1, edge1);
    double d = dot(edge0, pv);
    double e = dot(edge1, pv);
    double det = a * c - b * b;
    if (std::abs(det) < 1e-9) {
        float feps = 1e-6;
        float len1 = length(v0 - v1), len2 = length(v1 - v2), len3 = length(v2 - v0);
        if (len1 < feps && len2 < feps && len3 < feps) return v0;
        V3 dp, dq;
        if (len1 < feps) {
            dp = v0;
            dq = v2;
        } else if (len2 < feps) {
            dp = v1;
            dq = v0;
        } else {
            dp = v2;
            dq = v1;
        }
        float dotnum = dot(p0 - dp, dq - dp);
        float dotdem = dot(dq - dp, dq - dp);
        if (std::abs(dotdem) < feps) return v0;
        float lambda = dotnum / dotdem;
        if (lambda <= 0.0f) return dp;
        if (lambda >= 1.0f) return dq;
        return dp + lambda * (dq - dp);
    }
    double s = b * e - c * d;
    double t = b * d - a * e;
    if (s + t < det) {
        if (s < 0.0) {
            if (t < 0.0) {
                if (d < 0.0) {
                    s = clampd(-d / a, 0.0, 1.0);
                    t = 0.0;
                } else {
                    s = 0.0;
                    t = clampd(-e / c, 0.0, 1.0);
                }
            } else {
                s = 0.0;
                t = clampd(-e / c, 0.0, 1.0);
            }
        } else if (t < 0.0) {
            s = clampd(-d / a, 0.0, 1.0);
            t = 0.0;
        } else {
            double invDet = 1.0 / det;
            s *= invDet;
            t *= invDet;
        }
    } else {
        if (s < 0.0) {
            double tmp0 = b + d, tmp1 = c + e;
            if (tmp1 > tmp0) {
                s = clampd((tmp1 - tmp0) / (a - 2 * b + c), 0.0, 1.0);
                t = 1 - s;
            } else {
                t = clampd(-e / c, 0.0, 1.0);
                s = 0.0;
            }
        } else if (t < 0.0) {
            if (a + d > b + e) {
                s = clampd((c + e - b - d) / (a - 2 * b + c), 0.0, 1.0);
                t = 1 - s;
            } else {
                s = clampd(-e / c, 0.0, 1.0);
                t = 0.0;
            }
        } else {
            s = clampd((c + e - b - d) / (a - 2 * b + c), 0.0, 1.0);
            t = 1.0 - s;
        }
    }
    return v0 + (float)s * edge0 + (float)t * edge1;
}

}  // namespace

/* Nodal solid SDF of the domain boundary.  Restates
 *   FluidSimulation::_getBoundaryAABB / _getTriangleMeshFromAABB (fluidsimulation.cpp:5148-5186),
 *   _addStaticObjectsToSDF (5277-5287): fastCalculateSignedDistanceField + negate,
 *   MeshLevelSet::_computeExactBandDistanceFieldMultiThreaded (meshlevelset.cpp:781-1163):
 *     per triangle a node band from its AABB (+-3 nodes, _initializeTriangleData 878-918);
 *     distances evaluated in 10^3-node block-local coordinates (1115-1163);
 *   _computeDistanceFieldSigns (1298-1318): nodes inside the box are negated.
 * For the axis-aligned box the inside test is the box test itself: no node lies
 * on the box faces (they sit 1e-4 off the grid planes). */
void oracle_solid_box_phi(int ni, int nj, int nk, double dx, float *phi) {
    const int pi = ni + 1, pj = nj + 1, pk = nk + 1;
    const int band = 3, bw = 10;
    Box box = boundaryBox(ni, nj, nk, dx);
    V3 p = box.pos;
    V3 verts[8] = {
        mk(p.x, p.y, p.z),
        mk(p.x + box.w, p.y, p.z),
        mk(p.x + box.w, p.y, p.z + box.d),
        mk(p.x, p.y, p.z + box.d),
        mk(p.x, p.y + box.h, p.z),
        mk(p.x + box.w, p.y + box.h, p.z),
        mk(p.x + box.w, p.y + box.h, p.z + box.d),
        mk(p.x, p.y + box.h, p.z + box.d),
    };
    const int tris[12][3] = {{0, 1, 2}, {0, 2, 3}, {4, 7, 6}, {4, 6, 5}, {0, 3, 7}, {0, 7, 4},
                             {1, 5, 6}, {1, 6, 2}, {0, 4, 5}, {0, 5, 1}, {3, 2, 6}, {3, 6, 7}};
    float ub = (float)((pi + pj + pk) * dx);  // getDistanceUpperBound (meshlevelset.cpp:769-771)
    size_t nn = (size_t)pi * pj * pk;
    for (size_t n = 0; n < nn; n++) phi[n] = ub;

    for (int t = 0; t < 12; t++) {
        V3 a = verts[tris[t][0]], b = verts[tris[t][1]], c = verts[tris[t][2]];
        // AABB(Triangle, vertices) (aabb.cpp:84-115)
        double minx = a.x, miny = a.y, minz = a.z, maxx = a.x, maxy = a.y, maxz = a.z;
        V3 pts[3] = {a, b, c};
        for (int q = 0; q < 3; q++) {
            minx = std::fmin(pts[q].x, minx);
            miny = std::fmin(pts[q].y, miny);
            minz = std::fmin(pts[q].z, minz);
            maxx = std::fmax(pts[q].x, maxx);
            maxy = std::fmax(pts[q].y, maxy);
            maxz = std::fmax(pts[q].z, maxz);
        }
        Box tb{mk(minx, miny, minz), maxx - minx, maxy - miny, maxz - minz};
        GI gmin = posToIndex(tb.pos, dx);
        GI gmax = posToIndex(tb.maxPoint(), dx);
        gmin.i -= band; gmin.j -= band; gmin.k -= band;
        gmax.i += band + 1; gmax.j += band + 1; gmax.k += band + 1;
        if (gmax.i < 0 || gmin.i >= ni + 1 || gmax.j < 0 || gmin.j >= nj + 1 || gmax.k < 0 ||
            gmin.k >= nk + 1)
            continue;
        gmin.i = std::max(gmin.i, 0); gmin.j = std::max(gmin.j, 0); gmin.k = std::max(gmin.k, 0);
        gmax.i = std::min(gmax.i, ni); gmax.j = std::min(gmax.j, nj); gmax.k = std::min(gmax.k, nk);
        for (int k = gmin.k; k <= gmax.k; k++)
            for (int j = gmin.j; j <= gmax.j; j++)
                for (int i = gmin.i; i <= gmax.i; i++) {
                    int bi = i / bw, bj = j / bw, bk = k / bw;
                    V3 off = indexToPos(bi, bj, bk, bw * dx);
                    V3 gpos = indexToPos(i - bi * bw, j - bj * bw, k - bk * bw, dx);
                    V3 cp = closestPointOnTriangle(gpos, a - off, b - off, c - off);
                    float dist = length(cp - gpos);
                    size_t n = flat(i, j, k, pi, pj);
                    if (dist < phi[n]) phi[n] = dist;
                }
    }
    // sign (inside the box mesh -> negative), then MeshLevelSet::negate (meshlevelset.cpp:702-708)
    for (int k = 0; k < pk; k++)
        for (int j = 0; j < pj; j++)
            for (int i = 0; i < pi; i++) {
                size_t n = flat(i, j, k, pi, pj);
                V3 g = indexToPos(i, j, k, dx);
                bool in = box.inside(g);
                float v = in ? -phi[n] : phi[n];
                phi[n] = -v;
            }
}

// ------------------------------------------------- mesh obstacle SDF ----
namespace {

struct TriMesh {
    std::vector<V3> v;
    std::vector<std::array<int, 3>> t;
    std::vector<V3> vel;  // per-vertex velocities (MeshLevelSet::_vertexVelocities), or empty for zero
};

// AABB(std::vector<vec3>) and AABB(Triangle, vertices): min/max in double,
// widths padded by 1e-9 (aabb.cpp:54-112)
Box aabbOf(const V3 *pts, int n) {
    double minx = pts[0].x, miny = pts[0].y, minz = pts[0].z;
    double maxx = pts[0].x, maxy = pts[0].y, maxz = pts[0].z;
    for (int q = 0; q < n; q++) {
        minx = std::fmin(pts[q].x, minx);
        miny = std::fmin(pts[q].y, miny);
        minz = std::fmin(pts[q].z, minz);
        maxx = std::fmax(pts[q].x, maxx);
        maxy = std::fmax(pts[q].y, maxy);
        maxz = std::fmax(pts[q].z, maxz);
    }
    const double eps = 1e-9;
    return Box{mk(minx, miny, minz), maxx - minx + eps, maxy - miny + eps, maxz - minz + eps};
}
Box aabbTri(const TriMesh &m, int t) {
    V3 p[3] = {m.v[m.t[t][0]], m.v[m.t[t][1]], m.v[m.t[t][2]]};
    return aabbOf(p, 3);
}

void translate(TriMesh &m, V3 d) {  // TriangleMesh::translate (trianglemesh.cpp:869-873)
    for (V3 &p : m.v) p = p + d;
}

V3 cross(V3 a, V3 b) {  // vmath::cross (vmath.h:79-83)
    return V3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}

// Collision::lineIntersectsTriangle (collision.cpp:79-118): float
// Moller-Trumbore; the hit point is p + (float)t * dir.
bool lineIntersectsTriangle(V3 p, V3 dir, V3 v0, V3 v1, V3 v2, V3 *hit) {
    V3 e1 = v1 - v0, e2 = v2 - v0;
    V3 h = cross(dir, e2);
    float a = dot(e1, h);
    double eps = 10e-9;
    if (std::fabs(a) < eps) return false;
    float f = 1 / a;
    V3 s = p - v0;
    float u = f * dot(s, h);
    if (u < 0.0 || u > 1.0) return false;
    V3 q = cross(s, e1);
    float v = f * dot(dir, q);
    if (v < 0.0 || u + v > 1.0) return false;
    double t = f * dot(e2, q);
    *hit = p + (float)t * dir;
    return true;
}

// the three draws of one jitter vector: the reference builds
// vec3(_randomDouble(jit,-jit), x3) and its compiler evaluates the
// arguments right to left, so z takes the first draw (meshutils.cpp:93-117)
struct RandSource {
    const int *r;
    int n, used;
    double randMax;
    double next(double lo, double hi) {
        if (used >= n) throw std::runtime_error("oracle: rand() values exhausted");
        return lo + ((double)r[used++] / randMax) * (hi - lo);
    }
    V3 jitter(double dx) {
        double jit = 0.001 * dx;
        double z = next(jit, -jit), y = next(jit, -jit), x = next(jit, -jit);
        return mk(x, y, z);
    }
};

// The C library's rand() after srand(seed), restated so that the oracle can
// reseed at every substep as FluidSimulation::_stepFluid does
// (fluidsimulation.cpp:10079, srand(frame + substep)).  This is glibc's
// default generator (the TYPE_3 additive feedback generator of random_r.c,
// as published): the state r[0..30] comes from the seed (0 taken as 1) by
// the Park-Miller minimal standard LCG r[i] = 16807 r[i-1] mod 2^31-1,
// r[31..33] = r[0..2], then r[i] = r[i-31] + r[i-3] mod 2^32; the first 310
// sums are discarded and rand() returns each later sum shifted right by one.
struct GlibcRand {
    std::vector<uint32_t> r;
    explicit GlibcRand(unsigned seed) {
        r.assign(34, 0u);
        int32_t s0 = (int32_t)seed;
        if (s0 == 0) s0 = 1;
        r[0] = (uint32_t)s0;
        for (int i = 1; i < 31; i++) {
            const int32_t prev = (int32_t)r[i - 1];
            const int32_t hi = prev / 127773, lo = prev % 127773;
            int32_t word = 16807 * lo - 2836 * hi;  // Schrage's product without overflow
            if (word < 0) word += 2147483647;
            r[i] = (uint32_t)word;
        }
        for (int i = 31; i < 34; i++) r[i] = r[i - 31];
        for (int i = 34; i < 344; i++) r.push_back(r[i - 31] + r[i - 3]);
    }
    int next() {
        const size_t i = r.size();
        r.push_back(r[i - 31] + r[i - 3]);
        return (int)(r.back() >> 1);
    }
};

// MeshUtils::getCellsInsideTriangleMesh (meshutils.cpp:57-224): per column
// (i, j) the z of every crossing of the +z line through the jittered cell
// centre (i, j, -1); a cell is inside when the column has an even number of
// crossings and an odd number of them lie below its centre.
void cellsInside(const TriMesh &m, int isz, int jsz, int ksz, double dx, RandSource &rs,
                 std::vector<GI> &cells) {
    std::vector<std::vector<int>> col((size_t)isz * jsz);
    for (int t = 0; t < (int)m.t.size(); t++) {
        Box b = aabbTri(m, t);
        // Grid3d::getGridIndexBounds(AABB, ...) (grid3d.h:423-436)
        V3 trans = mk(b.w, b.h, b.d);
        GI g1 = posToIndex(b.pos, dx), g2 = posToIndex(b.pos + trans, dx);
        g1.i = std::max(g1.i, 0);
        g1.j = std::max(g1.j, 0);
        g2.i = std::min(g2.i, isz - 1);
        g2.j = std::min(g2.j, jsz - 1);
        for (int j = g1.j; j <= g2.j; j++)
            for (int i = g1.i; i <= g2.i; i++) col[(size_t)i + (size_t)isz * j].push_back(t);
    }
    V3 jitter = rs.jitter(dx);
    V3 dir{0.0f, 0.0f, 1.0f};
    std::vector<std::vector<float>> zs((size_t)isz * jsz);
    for (int j = 0; j < jsz; j++)
        for (int i = 0; i < isz; i++) {
            const std::vector<int> &tl = col[(size_t)i + (size_t)isz * j];
            if (tl.empty()) continue;
            V3 gp = indexToCellCenter(i, j, -1, dx) + jitter;
            for (int t : tl) {
                V3 hit;
                if (lineIntersectsTriangle(gp, dir, m.v[m.t[t][0]], m.v[m.t[t][1]], m.v[m.t[t][2]], &hit))
                    zs[(size_t)i + (size_t)isz * j].push_back(hit.z);
            }
        }
    for (int k = 0; k < ksz; k++)
        for (int j = 0; j < jsz; j++)
            for (int i = 0; i < isz; i++) {
                const std::vector<float> &z = zs[(size_t)i + (size_t)isz * j];
                if (z.size() % 2 != 0) continue;
                float zc = indexToCellCenter(i, j, k, dx).z;
                int less = 0;
                for (float q : z) less += q < zc;
                if (less % 2 == 1) cells.push_back(GI{i, j, k});
            }
}

// MeshUtils::getGridNodesInsideTriangleMesh, list form (meshutils.cpp:356-380):
// nodes are the cells of the grid shifted by half a cell.
void nodesInside(TriMesh m, double dx, RandSource &rs, std::vector<GI> &nodes) {
    if (m.v.empty()) {
        // AABB of no points has zero extent (aabb.h:60-63): a 1x1x1 grid with
        // no triangles, which still draws its jitter and holds no inside cell
        rs.jitter(dx);
        return;
    }
    translate(m, mk(0.5 * dx, 0.5 * dx, 0.5 * dx));
    Box b = aabbOf(m.v.data(), (int)m.v.size());
    GI go = posToIndex(b.pos, dx);
    V3 off = indexToPos(go.i, go.j, go.k, dx);
    translate(m, V3{-off.x, -off.y, -off.z});
    b.pos = b.pos - off;
    V3 trans = mk(b.w, b.h, b.d);
    GI g2 = posToIndex(b.pos + trans, dx);
    std::vector<GI> cells;
    cellsInside(m, g2.i + 1, g2.j + 1, g2.k + 1, dx, rs, cells);
    for (GI c : cells) nodes.push_back(GI{c.i + go.i, c.j + go.j, c.k + go.k});
}

// MeshUtils::_splitIntoMeshIslands (meshutils.cpp:464-573): connected
// components over shared vertices; island vertices in first-use order.
std::vector<TriMesh> splitIslands(const TriMesh &m) {
    int nv = (int)m.v.size();
    std::vector<std::vector<int>> nb(nv);
    for (auto &t : m.t) {
        nb[t[0]].push_back(t[1]); nb[t[0]].push_back(t[2]);
        nb[t[1]].push_back(t[0]); nb[t[1]].push_back(t[2]);
        nb[t[2]].push_back(t[0]); nb[t[2]].push_back(t[1]);
    }
    std::vector<int> group(nv, -1), stack;
    std::vector<char> seen(nv, 0);
    int ng = 0;
    for (int s = 0; s < nv; s++) {
        if (seen[s]) continue;
        stack.assign(1, s);
        seen[s] = 1;
        while (!stack.empty()) {
            int v = stack.back();
            stack.pop_back();
            for (int w : nb[v])
                if (!seen[w]) {
                    stack.push_back(w);
                    seen[w] = 1;
                }
            group[v] = ng;
        }
        ng++;
    }
    int maxg = 0;
    for (int g : group) maxg = std::max(maxg, g);
    std::vector<TriMesh> isl(nv > 0 ? maxg + 1 : 0);
    std::vector<int> tr(nv, -1);
    for (auto t : m.t) {
        int g = group[t[0]];
        for (int c = 0; c < 3; c++)
            if (tr[t[c]] == -1) {
                tr[t[c]] = (int)isl[g].v.size();
                isl[g].v.push_back(m.v[t[c]]);
                // MeshUtils::splitIntoMeshIslands with velocities (meshutils.cpp:574-593)
                if (!m.vel.empty()) isl[g].vel.push_back(m.vel[t[c]]);
            }
        isl[g].t.push_back(t);
    }
    for (auto &I : isl)
        for (auto &t : I.t)
            for (int c = 0; c < 3; c++) t[c] = tr[t[c]];
    return isl;
}

// AABB::getIntersection with a non-empty result (aabb.cpp:170-191)
bool aabbIntersects(const Box &a, const Box &b) {
    V3 mn1 = a.pos, mx1 = a.maxPoint(), mn2 = b.pos, mx2 = b.maxPoint();
    if (mn1.x > mx2.x || mn1.y > mx2.y || mn1.z > mx2.z || mx1.x < mn2.x || mx1.y < mn2.y || mx1.z < mn2.z)
        return false;
    double w = std::fmin(mx1.x, mx2.x) - std::fmax(mn1.x, mn2.x);
    double h = std::fmin(mx1.y, mx2.y) - std::fmax(mn1.y, mn2.y);
    double d = std::fmin(mx1.z, mx2.z) - std::fmax(mn1.z, mn2.z);
    return w > 0.0 || h > 0.0 || d > 0.0;
}
bool aabbContains(const Box &grid, const Box &b) { return grid.inside(b.pos) && grid.inside(b.maxPoint()); }

// MeshUtils::getGridNodesInsideTriangleMesh, grid form (meshutils.cpp:382-437):
// a mesh that leaves the grid is split into islands; those inside are
// tested together, each one crossing the grid border on its own.
void nodesInsideGrid(const TriMesh &m, int pi, int pj, int pk, double dx, RandSource &rs,
                     std::vector<char> &inside) {
    int isz = pi - 1, jsz = pj - 1, ksz = pk - 1;
    Box grid{V3{0.0f, 0.0f, 0.0f}, isz * dx, jsz * dx, ksz * dx};
    bool contained = true;
    for (V3 p : m.v)
        if (!grid.inside(p)) {
            contained = false;
            break;
        }
    std::vector<TriMesh> groups;
    if (contained) {
        groups.push_back(m);
    } else {
        TriMesh in;
        std::vector<TriMesh> out;
        for (TriMesh &I : splitIslands(m)) {
            Box b = aabbOf(I.v.data(), (int)I.v.size());
            if (aabbContains(grid, b)) {
                int base = (int)in.v.size();
                in.v.insert(in.v.end(), I.v.begin(), I.v.end());
                for (auto t : I.t) in.t.push_back({t[0] + base, t[1] + base, t[2] + base});
            } else if (aabbIntersects(grid, b)) {
                out.push_back(I);
            }
        }
        groups.push_back(in);
        for (TriMesh &o : out) groups.push_back(o);
    }
    for (TriMesh &g : groups) {
        std::vector<GI> nodes;
        nodesInside(g, dx, rs, nodes);
        for (GI n : nodes)
            if (inRange(n.i, n.j, n.k, pi, pj, pk)) inside[flat(n.i, n.j, n.k, pi, pj)] = 1;
    }
}

// MeshLevelSet::fastCalculateSignedDistanceField on a level set of
// isize x jsize x ksize cells at node offset (oi, oj, ok) (meshlevelset.cpp:
// 572-601): the exact band (multi-threaded form 781-1163: node range from the
// triangle's AABB +-band, distances in 10^3-node block-local coordinates), then
// the crossing-count signs (1298-1318).  No propagation beyond the band.
//
// singleThreaded: the band as _computeExactBandDistanceFieldSingleThreaded
// computes it (1165-1234), the form the fracture path's per-island level
// sets take: node range from int() of the scaled vertex coordinates,
// clamped to the node grid, distances in the level set's own coordinates.
//
// closest (optional): each node's closest triangle (-1 outside the band): the
// first triangle, in index order, at the node's smallest distance.
void meshLevelSetFast(const TriMesh &m, int isize, int jsize, int ksize, double dx, GI o, int band,
                      RandSource &rs, std::vector<float> &phi, bool singleThreaded = false,
                      std::vector<int> *closest = nullptr) {
    const int pi = isize + 1, pj = jsize + 1, pk = ksize + 1, bw = 10;
    float ub = (float)((pi + pj + pk) * dx);  // getDistanceUpperBound (meshlevelset.cpp:769-771)
    phi.assign((size_t)pi * pj * pk, ub);
    if (closest) closest->assign(phi.size(), -1);
    V3 po = indexToPos(o.i, o.j, o.k, dx);
    auto clampi = [](int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); };
    for (int t = 0; singleThreaded && t < (int)m.t.size(); t++) {
        V3 a = m.v[m.t[t][0]] - po, bb = m.v[m.t[t][1]] - po, c = m.v[m.t[t][2]] - po;
        double invdx = 1.0 / dx;
        double fa[3] = {(double)a.x * invdx, (double)a.y * invdx, (double)a.z * invdx};
        double fb[3] = {(double)bb.x * invdx, (double)bb.y * invdx, (double)bb.z * invdx};
        double fc[3] = {(double)c.x * invdx, (double)c.y * invdx, (double)c.z * invdx};
        int lo[3], hi[3], dim[3] = {pi, pj, pk};
        for (int q = 0; q < 3; q++) {
            lo[q] = clampi(int(std::fmin(fa[q], std::fmin(fb[q], fc[q]))) - band, 0, dim[q] - 1);
            hi[q] = clampi(int(std::fmax(fa[q], std::fmax(fb[q], fc[q]))) + band + 1, 0, dim[q] - 1);
        }
        for (int k = lo[2]; k <= hi[2]; k++)
            for (int j = lo[1]; j <= hi[1]; j++)
                for (int i = lo[0]; i <= hi[0]; i++) {
                    V3 gpos = indexToPos(i, j, k, dx);
                    V3 cp = closestPointOnTriangle(gpos, a, bb, c);
                    float dist = length(cp - gpos);
                    size_t n = flat(i, j, k, pi, pj);
                    if (dist < phi[n]) {
                        phi[n] = dist;
                        if (closest) (*closest)[n] = t;
                    }
                }
    }
    for (int t = 0; !singleThreaded && t < (int)m.t.size(); t++) {
        Box b = aabbTri(m, t);
        b.pos = b.pos - po;
        GI g0 = posToIndex(b.pos, dx), g1 = posToIndex(b.maxPoint(), dx);
        g0.i -= band; g0.j -= band; g0.k -= band;
        g1.i += band + 1; g1.j += band + 1; g1.k += band + 1;
        if (g1.i < 0 || g0.i >= pi || g1.j < 0 || g0.j >= pj || g1.k < 0 || g0.k >= pk) continue;
        g0.i = std::max(g0.i, 0); g0.j = std::max(g0.j, 0); g0.k = std::max(g0.k, 0);
        g1.i = std::min(g1.i, isize); g1.j = std::min(g1.j, jsize); g1.k = std::min(g1.k, ksize);
        V3 a = m.v[m.t[t][0]] - po, bb = m.v[m.t[t][1]] - po, c = m.v[m.t[t][2]] - po;
        for (int k = g0.k; k <= g1.k; k++)
            for (int j = g0.j; j <= g1.j; j++)
                for (int i = g0.i; i <= g1.i; i++) {
                    int bi = i / bw, bj = j / bw, bk = k / bw;
                    V3 off = indexToPos(bi, bj, bk, bw * dx);
                    V3 gpos = indexToPos(i - bi * bw, j - bj * bw, k - bk * bw, dx);
                    V3 cp = closestPointOnTriangle(gpos, a - off, bb - off, c - off);
                    float dist = length(cp - gpos);
                    size_t n = flat(i, j, k, pi, pj);
                    if (dist < phi[n]) {
                        phi[n] = dist;
                        if (closest) (*closest)[n] = t;
                    }
                }
    }
    TriMesh local = m;
    translate(local, V3{-po.x, -po.y, -po.z});
    std::vector<char> inside((size_t)pi * pj * pk, 0);
    nodesInsideGrid(local, pi, pj, pk, dx, rs, inside);
    for (size_t n = 0; n < phi.size(); n++)
        if (inside[n]) phi[n] = -phi[n];
}

// TriangleMesh::removeExtraneousVertices (trianglemesh.cpp:775-822)
void removeUnused(TriMesh &m) {
    std::vector<int> tr(m.v.size(), -1);
    std::vector<char> used(m.v.size(), 0);
    for (auto &t : m.t) used[t[0]] = used[t[1]] = used[t[2]] = 1;
    std::vector<V3> nv, nvel;
    for (size_t q = 0; q < m.v.size(); q++)
        if (used[q]) {
            tr[q] = (int)nv.size();
            nv.push_back(m.v[q]);
            // MeshObject::getMeshLevelSet erases the removed vertices' velocities (meshobject.cpp:244-248)
            if (!m.vel.empty()) nvel.push_back(m.vel[q]);
        }
    m.v = nv;
    m.vel = nvel;
    for (auto &t : m.t)
        for (int c = 0; c < 3; c++) t[c] = tr[t[c]];
}

// MeshLevelSet::calculateUnion, phi part (meshlevelset.cpp:603-638, 1759-1794)
void unionInto(std::vector<float> &phi, int pi, int pj, int pk, const std::vector<float> &other, int qi, int qj,
               int qk, GI o) {
    for (int k = 0; k < qk; k++)
        for (int j = 0; j < qj; j++)
            for (int i = 0; i < qi; i++) {
                int a = i + o.i, b = j + o.j, c = k + o.k;
                if (!inRange(a, b, c, pi, pj, pk)) continue;
                float v = other[flat(i, j, k, qi, qj)];
                float &p = phi[flat(a, b, c, pi, pj)];
                if (v < p) p = v;
            }
}

// MeshObject::getMeshLevelSet for a static mesh (meshobject.cpp:239-264,
// 633-701, 738-754): fewer than 25 islands inside the domain -> the whole
// mesh on its island grid; otherwise each island on its own grid.
// MeshObject::_expandMeshIsland (meshobject.cpp:714-736): every vertex moves
// 0.5 x expansion along its normalized offset from the float centroid
void expandIsland(TriMesh &m, float expansion) {
    if (m.v.empty()) return;
    V3 sum{0.0f, 0.0f, 0.0f};
    for (V3 p : m.v) sum = sum + p;
    V3 centroid = divide(sum, (float)m.v.size());
    float expval = 0.5f * expansion, eps = 1e-9f;
    for (V3 &p : m.v) {
        V3 v = p - centroid;
        if (std::fabs(v.x) < eps && std::fabs(v.y) < eps && std::fabs(v.z) < eps) continue;
        v = normalize(v);
        p = p + expval * v;
    }
}

void obstacleLevelSet(TriMesh m, int ni, int nj, int nk, double dx, int band, RandSource &rs,
                      std::vector<float> &levelset, float expansion = 0.0f) {
    // MeshObject::getMesh(frameInterpolation) (meshobject.cpp:158-177): v + f (v - v)
    // with v - v = +0, so a -0 coordinate becomes +0
    for (V3 &p : m.v) p = p + V3{0.0f, 0.0f, 0.0f};
    removeUnused(m);
    if (m.t.empty()) return;
    Box grid{V3{0.0f, 0.0f, 0.0f}, ni * dx, nj * dx, nk * dx};
    std::vector<TriMesh> islands;
    for (TriMesh &I : splitIslands(m)) {
        Box b = aabbOf(I.v.data(), (int)I.v.size());
        if (aabbContains(grid, b) || aabbIntersects(grid, b)) islands.push_back(I);
    }
    std::vector<TriMesh> parts;
    const bool fracture = (int)islands.size() >= 25;
    if (fracture) {
        // islands expanded (_expandMeshIslands, meshobject.cpp:703-712, a
        // no-op below |expansion| 1e-9); the work queue pops from its back
        // (boundedbuffer.h:119-135): one producer thread takes the islands
        // last-first, and so draws their jitters in that order (with more
        // threads the order varies)
        if (std::fabs(expansion) >= 1e-9f)
            for (TriMesh &I : islands) expandIsland(I, expansion);
        parts.assign(islands.rbegin(), islands.rend());
    }
    else
        parts.push_back(m);
    for (TriMesh &part : parts) {
        Box b = aabbOf(part.v.data(), (int)part.v.size());
        GI g0 = posToIndex(b.pos, dx), g1 = posToIndex(b.maxPoint(), dx);
        g0.i = (int)std::fmax(g0.i - band, 0); g0.j = (int)std::fmax(g0.j - band, 0); g0.k = (int)std::fmax(g0.k - band, 0);
        g1.i = (int)std::fmin(g1.i + band + 1, ni - 1); g1.j = (int)std::fmin(g1.j + band + 1, nj - 1);
        g1.k = (int)std::fmin(g1.k + band + 1, nk - 1);
        int w = g1.i - g0.i, h = g1.j - g0.j, d = g1.k - g0.k;
        if (w <= 0 || h <= 0 || d <= 0) continue;
        std::vector<float> phi;
        // fracture path: one single-threaded level set per island
        // (meshobject.cpp:798-839)
        meshLevelSetFast(part, w, h, d, dx, g0, band, rs, phi, fracture);
        unionInto(levelset, ni + 1, nj + 1, nk + 1, phi, w + 1, h + 1, d + 1, g0);
    }
}

float fractionInside4(float bl, float br, float tl, float tr);  // LevelsetUtils::fractionInside, below

// ---- the solid's velocity grids (moving obstacles) ----
//
// MeshLevelSet with its velocity data (meshlevelset.h:68-76): phi and the
// closest triangle per node, the merged mesh with its vertex velocities, and
// per face the weighted velocity and the weight.  The domain grid has offset
// (0, 0, 0); an island's level set sits at its grid offset.
struct SolidLS {
    int ni = 0, nj = 0, nk = 0;
    GI off{0, 0, 0};
    double dx = 0.0;
    bool vel = true;  // isVelocityDataEnabled
    std::vector<float> phi;
    std::vector<int> tri;
    TriMesh mesh;
    std::vector<float> fu, fv, fw, wu, wv, ww;
    // the constructor and reset() (meshlevelset.cpp:710-718)
    void init(int i, int j, int k, double d, GI o = GI{0, 0, 0}) {
        ni = i; nj = j; nk = k; dx = d; off = o;
        reset();
    }
    void reset() {
        mesh = TriMesh();
        phi.assign((size_t)(ni + 1) * (nj + 1) * (nk + 1), ub());
        tri.assign(phi.size(), -1);
        clearVelocity();
    }
    void clearVelocity() {  // VelocityDataGrid::reset
        fu.assign((size_t)(ni + 1) * nj * nk, 0.0f);
        wu.assign(fu.size(), 0.0f);
        fv.assign((size_t)ni * (nj + 1) * nk, 0.0f);
        wv.assign(fv.size(), 0.0f);
        fw.assign((size_t)ni * nj * (nk + 1), 0.0f);
        ww.assign(fw.size(), 0.0f);
    }
    float ub() const { return (float)((ni + nj + nk + 3) * dx); }  // getDistanceUpperBound
    float P(int i, int j, int k) const { return phi[flat(i, j, k, ni + 1, nj + 1)]; }
    V3 posOff() const { return indexToPos(off.i, off.j, off.k, dx); }
};

// MeshLevelSet::_pointToSegmentVelocity (meshlevelset.cpp:1697-1717)
V3 segmentVelocity(V3 x0, V3 x1, V3 x2, V3 v1, V3 v2, float *distance) {
    V3 d = x2 - x1;
    double m2 = lengthsq(d);
    float s12 = (float)(dot(x2 - x0, d) / m2);
    if (s12 < 0)
        s12 = 0;
    else if (s12 > 1)
        s12 = 1;
    *distance = length(x0 - (s12 * x1 + (1 - s12) * x2));
    return s12 * v1 + (1 - s12) * v2;
}

// MeshLevelSet::_pointToTriangleVelocity (meshlevelset.cpp:1568-1639)
V3 triangleVelocity(const SolidLS &L, V3 x0, int t) {
    const std::array<int, 3> &T = L.mesh.t[t];
    V3 v1 = L.mesh.vel[T[0]], v2 = L.mesh.vel[T[1]], v3 = L.mesh.vel[T[2]];
    float eps = 1e-6f;
    if (std::fabs(v1.x) < eps && std::fabs(v1.y) < eps && std::fabs(v1.z) < eps && std::fabs(v2.x) < eps &&
        std::fabs(v2.y) < eps && std::fabs(v2.z) < eps && std::fabs(v3.x) < eps && std::fabs(v3.y) < eps &&
        std::fabs(v3.z) < eps)
        return V3{0.0f, 0.0f, 0.0f};
    V3 po = L.posOff();
    V3 x1 = L.mesh.v[T[0]] - po, x2 = L.mesh.v[T[1]] - po, x3 = L.mesh.v[T[2]] - po;
    V3 x13 = x1 - x3, x23 = x2 - x3, x03 = x0 - x3;
    float m13 = lengthsq(x13), m23 = lengthsq(x23), d = dot(x13, x23);
    float invdet = 1.0f / std::fmax(m13 * m23 - d * d, 1e-30f);
    float a = dot(x13, x03), b = dot(x23, x03);
    float w23 = invdet * (m23 * a - d * b);
    float w31 = invdet * (m13 * b - d * a);
    float w12 = 1 - w23 - w31;
    if (w23 >= 0 && w31 >= 0 && w12 >= 0) return w23 * v1 + w31 * v2 + w12 * v3;
    float d1, d2;
    V3 e1, e2;
    if (w23 > 0) {  // rules out edge 2-3
        e1 = segmentVelocity(x0, x1, x2, v1, v2, &d1);
        e2 = segmentVelocity(x0, x1, x3, v1, v3, &d2);
    } else if (w31 > 0) {  // rules out edge 1-3
        e1 = segmentVelocity(x0, x1, x2, v1, v2, &d1);
        e2 = segmentVelocity(x0, x2, x3, v2, v3, &d2);
    } else {  // rules out edge 1-2
        e1 = segmentVelocity(x0, x1, x3, v1, v3, &d1);
        e2 = segmentVelocity(x0, x2, x3, v2, v3, &d2);
    }
    return d1 < d2 ? e1 : e2;
}

// MeshLevelSet::getNearestVelocity (meshlevelset.cpp:168-205): among the
// closest triangles of the point's cell's 8 nodes, the nearest one's velocity
V3 nearestVelocity(const SolidLS &L, V3 p) {
    V3 po = L.posOff();
    p = p - po;
    GI g = posToIndex(p, L.dx);
    const int o8[8][3] = {{0, 0, 0}, {1, 0, 0}, {1, 0, 1}, {0, 0, 1}, {0, 1, 0}, {1, 1, 0}, {1, 1, 1}, {0, 1, 1}};
    int nearest = -1;
    float nearestDist = L.ub();
    for (int q = 0; q < 8; q++) {
        int i = g.i + o8[q][0], j = g.j + o8[q][1], k = g.k + o8[q][2];
        if (!inRange(i, j, k, L.ni + 1, L.nj + 1, L.nk + 1)) continue;
        int t = L.tri[flat(i, j, k, L.ni + 1, L.nj + 1)];
        if (t == -1) continue;
        const std::array<int, 3> &T = L.mesh.t[t];
        V3 cp = closestPointOnTriangle(p, L.mesh.v[T[0]] - po, L.mesh.v[T[1]] - po, L.mesh.v[T[2]] - po);
        float d = length(cp - p);
        if (d < nearestDist) {
            nearestDist = d;
            nearest = t;
        }
    }
    if (nearest == -1) return V3{0.0f, 0.0f, 0.0f};
    return triangleVelocity(L, p, nearest);
}

// MeshLevelSet::_computeVelocityGrids (meshlevelset.cpp:1320-1425): every face
// with a positive fraction inside the level set gets weight x the nearest
// velocity (0 for a mesh whose vertices all stand still) and the weight
void velocityGrids(SolidLS &L) {
    L.clearVelocity();
    bool isStatic = true;
    for (V3 v : L.mesh.vel)
        if (std::fabs(v.x) > 0.0f || std::fabs(v.y) > 0.0f || std::fabs(v.z) > 0.0f) {
            isStatic = false;
            break;
        }
    V3 po = L.posOff();
    const double dx = L.dx;
    for (int k = 0; k < L.nk; k++)
        for (int j = 0; j < L.nj; j++)
            for (int i = 0; i < L.ni + 1; i++) {
                float w = fractionInside4(L.P(i, j, k), L.P(i, j + 1, k), L.P(i, j, k + 1), L.P(i, j + 1, k + 1));
                if (!(w > 0.0f)) continue;
                V3 v{0.0f, 0.0f, 0.0f};
                if (!isStatic)
                    v = nearestVelocity(L, mk((float)i * dx, ((float)j + 0.5) * dx, ((float)k + 0.5) * dx) + po);
                size_t f = flat(i, j, k, L.ni + 1, L.nj);
                L.fu[f] = w * v.x;
                L.wu[f] = w;
            }
    for (int k = 0; k < L.nk; k++)
        for (int j = 0; j < L.nj + 1; j++)
            for (int i = 0; i < L.ni; i++) {
                float w = fractionInside4(L.P(i, j, k), L.P(i, j, k + 1), L.P(i + 1, j, k), L.P(i + 1, j, k + 1));
                if (!(w > 0.0f)) continue;
                V3 v{0.0f, 0.0f, 0.0f};
                if (!isStatic)
                    v = nearestVelocity(L, mk(((float)i + 0.5) * dx, (float)j * dx, ((float)k + 0.5) * dx) + po);
                size_t f = flat(i, j, k, L.ni, L.nj + 1);
                L.fv[f] = w * v.y;
                L.wv[f] = w;
            }
    for (int k = 0; k < L.nk + 1; k++)
        for (int j = 0; j < L.nj; j++)
            for (int i = 0; i < L.ni; i++) {
                float w = fractionInside4(L.P(i, j, k), L.P(i, j + 1, k), L.P(i + 1, j, k), L.P(i + 1, j + 1, k));
                if (!(w > 0.0f)) continue;
                V3 v{0.0f, 0.0f, 0.0f};
                if (!isStatic)
                    v = nearestVelocity(L, mk(((float)i + 0.5) * dx, ((float)j + 0.5) * dx, (float)k * dx) + po);
                size_t f = flat(i, j, k, L.ni, L.nj);
                L.fw[f] = w * v.z;
                L.ww[f] = w;
            }
}

// MeshLevelSet::fastCalculateSignedDistanceField with vertex velocities
// (meshlevelset.cpp:577-601)
void fastSDF(SolidLS &L, const TriMesh &m, int band, RandSource &rs, bool singleThreaded) {
    L.mesh = m;
    if (L.mesh.vel.size() != L.mesh.v.size()) L.mesh.vel.assign(L.mesh.v.size(), V3{0.0f, 0.0f, 0.0f});
    meshLevelSetFast(m, L.ni, L.nj, L.nk, L.dx, L.off, band, rs, L.phi, singleThreaded, &L.tri);
    if (L.vel) velocityGrids(L);
}

// MeshLevelSet::negate (meshlevelset.cpp:702-708)
void negateLS(SolidLS &L) {
    for (float &p : L.phi) p = -p;
    if (L.vel) velocityGrids(L);
}

// MeshLevelSet::calculateUnion (meshlevelset.cpp:603-638, 1759-1845): the
// mesh, vertex velocities, phi and closest triangles merged; the velocity data
// added, on the other grid's border only for faces inside both grids
void unionLS(SolidLS &L, const SolidLS &O) {
    const int triOffset = (int)L.mesh.t.size(), vOffset = (int)L.mesh.v.size();
    L.mesh.v.insert(L.mesh.v.end(), O.mesh.v.begin(), O.mesh.v.end());
    for (auto t : O.mesh.t) L.mesh.t.push_back({t[0] + vOffset, t[1] + vOffset, t[2] + vOffset});
    L.mesh.vel.insert(L.mesh.vel.end(), O.mesh.vel.begin(), O.mesh.vel.end());
    auto onBorder = [](int i, int j, int k, int a, int b, int c) {
        return i == 0 || j == 0 || k == 0 || i == a - 1 || j == b - 1 || k == c - 1;
    };
    for (int k = 0; k < O.nk + 1; k++)
        for (int j = 0; j < O.nj + 1; j++)
            for (int i = 0; i < O.ni + 1; i++) {
                int a = i + O.off.i - L.off.i, b = j + O.off.j - L.off.j, c = k + O.off.k - L.off.k;
                if (!inRange(a, b, c, L.ni + 1, L.nj + 1, L.nk + 1)) continue;
                size_t n = flat(a, b, c, L.ni + 1, L.nj + 1);
                float po = O.P(i, j, k);
                if (po < L.phi[n]) {
                    if (std::fabs(po) < std::fabs(L.phi[n])) {
                        int t = O.tri[flat(i, j, k, O.ni + 1, O.nj + 1)];
                        if (t != -1) L.tri[n] = t + triOffset;
                    }
                    L.phi[n] = po;
                }
                if (!L.vel) continue;
                bool border = onBorder(a, b, c, L.ni + 1, L.nj + 1, L.nk + 1) ||
                              onBorder(i, j, k, O.ni + 1, O.nj + 1, O.nk + 1);
                bool uIn = inRange(a, b, c, L.ni + 1, L.nj, L.nk) && inRange(i, j, k, O.ni + 1, O.nj, O.nk);
                bool vIn = inRange(a, b, c, L.ni, L.nj + 1, L.nk) && inRange(i, j, k, O.ni, O.nj + 1, O.nk);
                bool wIn = inRange(a, b, c, L.ni, L.nj, L.nk + 1) && inRange(i, j, k, O.ni, O.nj, O.nk + 1);
                if (!border || uIn) {
                    size_t f = flat(a, b, c, L.ni + 1, L.nj), g = flat(i, j, k, O.ni + 1, O.nj);
                    L.fu[f] += O.fu[g];
                    L.wu[f] += O.wu[g];
                }
                if (!border || vIn) {
                    size_t f = flat(a, b, c, L.ni, L.nj + 1), g = flat(i, j, k, O.ni, O.nj + 1);
                    L.fv[f] += O.fv[g];
                    L.wv[f] += O.wv[g];
                }
                if (!border || wIn) {
                    size_t f = flat(a, b, c, L.ni, L.nj), g = flat(i, j, k, O.ni, O.nj);
                    L.fw[f] += O.fw[g];
                    L.ww[f] += O.ww[g];
                }
            }
}

// MeshLevelSet::normalizeVelocityGrid (meshlevelset.cpp:640-700, 1739-1757):
// velocity = field / weight where the weight exceeds 1e-6 (valid), else 0,
// then 5 extrapolation layers from the valid faces
void normalizeLS(SolidLS &L, std::vector<float> &u, std::vector<float> &v, std::vector<float> &w) {
    auto one = [](const std::vector<float> &f, const std::vector<float> &wt, std::vector<float> &out,
                  std::vector<unsigned char> &valid) {
        out.assign(f.size(), 0.0f);
        valid.assign(f.size(), 0);
        for (size_t q = 0; q < f.size(); q++)
            if (wt[q] > 1e-6f) {
                out[q] = f[q] / wt[q];
                valid[q] = 1;
            }
    };
    std::vector<unsigned char> vu, vv, vw;
    one(L.fu, L.wu, u, vu);
    one(L.fv, L.wv, v, vv);
    one(L.fw, L.ww, w, vw);
    const int layers = 5;  // _numVelocityExtrapolationLayers (meshlevelset.h:363)
    oracle_extrapolate(L.ni + 1, L.nj, L.nk, u.data(), vu.data(), layers);
    oracle_extrapolate(L.ni, L.nj + 1, L.nk, v.data(), vv.data(), layers);
    oracle_extrapolate(L.ni, L.nj, L.nk + 1, w.data(), vw.data(), layers);
}

// One obstacle's level set with velocities on the domain grid L (MeshObject::
// getMeshLevelSet, meshobject.cpp:239-264, with _addMeshIslandsToLevelSet
// 738-754 or the 25-or-more-island path 756-839): m holds the mesh at the
// substep's frame progress and its vertex velocities.
void obstacleLevelSetVel(TriMesh m, int band, RandSource &rs, SolidLS &L, float expansion) {
    for (V3 &p : m.v) p = p + V3{0.0f, 0.0f, 0.0f};
    removeUnused(m);
    if (m.t.empty()) return;
    const int ni = L.ni, nj = L.nj, nk = L.nk;
    const double dx = L.dx;
    Box grid{V3{0.0f, 0.0f, 0.0f}, ni * dx, nj * dx, nk * dx};
    std::vector<TriMesh> islands;
    for (TriMesh &I : splitIslands(m)) {
        Box b = aabbOf(I.v.data(), (int)I.v.size());
        if (aabbContains(grid, b) || aabbIntersects(grid, b)) islands.push_back(I);
    }
    std::vector<TriMesh> parts;
    const bool fracture = (int)islands.size() >= 25;
    if (fracture) {
        if (std::fabs(expansion) >= 1e-9f)
            for (TriMesh &I : islands) expandIsland(I, expansion);
        parts.assign(islands.rbegin(), islands.rend());
    } else {
        parts.push_back(m);
    }
    for (TriMesh &part : parts) {
        Box b = aabbOf(part.v.data(), (int)part.v.size());
        GI g0 = posToIndex(b.pos, dx), g1 = posToIndex(b.maxPoint(), dx);
        g0.i = (int)std::fmax(g0.i - band, 0); g0.j = (int)std::fmax(g0.j - band, 0); g0.k = (int)std::fmax(g0.k - band, 0);
        g1.i = (int)std::fmin(g1.i + band + 1, ni - 1); g1.j = (int)std::fmin(g1.j + band + 1, nj - 1);
        g1.k = (int)std::fmin(g1.k + band + 1, nk - 1);
        int w = g1.i - g0.i, h = g1.j - g0.j, d = g1.k - g0.k;
        if (w <= 0 || h <= 0 || d <= 0) continue;
        SolidLS I;
        I.init(w, h, d, dx, g0);
        fastSDF(I, part, band, rs, fracture);
        unionLS(L, I);
    }
}

TriMesh meshFrom(int nv, const float *verts, int nt, const int *tris) {
    TriMesh m;
    m.v.resize(nv);
    for (int q = 0; q < nv; q++) m.v[q] = V3{verts[3 * q], verts[3 * q + 1], verts[3 * q + 2]};
    m.t.resize(nt);
    for (int q = 0; q < nt; q++) {
        for (int c = 0; c < 3; c++) {
            int v = tris[3 * q + c];
            if (v < 0 || v >= nv) throw std::runtime_error("oracle: triangle index out of range");
            m.t[q][c] = v;
        }
    }
    return m;
}

}  // namespace

int oracle_mesh_levelset_fast(int isize, int jsize, int ksize, double dx, int oi, int oj, int ok, int nv,
                              const float *verts, int nt, const int *tris, int band, const int *rnd, int nrnd,
                              int rand_max, float *phi_nodes) {
    try {
        TriMesh m = meshFrom(nv, verts, nt, tris);
        RandSource rs{rnd, nrnd, 0, (double)rand_max};
        std::vector<float> phi;
        meshLevelSetFast(m, isize, jsize, ksize, dx, GI{oi, oj, ok}, band, rs, phi);
        std::memcpy(phi_nodes, phi.data(), sizeof(float) * phi.size());
        return rs.used;
    } catch (const std::exception &) {
        return -1;
    }
}

namespace {
// One obstacle of the solid update: its meshes (one triangle list) at the
// previous, current and next frame, and its flags.
struct MovingObstacle {
    TriMesh prev, cur, next;
    bool animated, inversed;
    float expansion, velocityScale;
};

// MeshObject::getMesh(frameInterpolation) and getVertexVelocities(dt,
// frameInterpolation) (meshobject.cpp:61-95, 158-220): the mesh moved from the
// current toward the next frame's, and each vertex's translation (current
// minus previous, next minus current) interpolated the same way, times
// (float)(1 / frame dt) and the velocity scale.  A static object's
// translations are zero.  One triangle list: the topology never changes.
TriMesh meshAtProgress(const MovingObstacle &o, float f, double frameDt) {
    f = std::fmax(0.0f, f);
    f = std::fmin(1.0f, f);
    TriMesh m = o.cur;
    m.vel.assign(m.v.size(), V3{0.0f, 0.0f, 0.0f});
    for (size_t q = 0; q < m.v.size(); q++) {
        V3 v1 = o.cur.v[q], v2 = o.next.v[q];
        m.v[q] = v1 + f * (v2 - v1);
    }
    if (!o.animated || frameDt < 1e-10) return m;
    const float invdt = (float)(1.0 / frameDt);
    for (size_t q = 0; q < m.v.size(); q++) {
        V3 p1 = o.cur.v[q] - o.prev.v[q], p2 = o.next.v[q] - o.cur.v[q];
        V3 t = p1 + f * (p2 - p1);
        m.vel[q] = (invdt * t);
        m.vel[q] = o.velocityScale * m.vel[q];
    }
    return m;
}

// FluidSimulation::_updateSolidLevelSet (fluidsimulation.cpp:5387-5426) with
// the default options (no precomputed static level set, no fracture
// optimization): the domain's negated box level set with its velocity
// weights, the static objects (normal, then the negated union of the
// inversed), then the animated ones (5219-5275), then normalizeVelocityGrid.
// Returns the solid phi and the normalized solid face velocities.
void solidUpdate(int ni, int nj, int nk, double dx, const std::vector<MovingObstacle> &obs, double frameDt,
                 float progress, RandSource &rs, std::vector<float> &phi, std::vector<float> &u,
                 std::vector<float> &v, std::vector<float> &w) {
    const int band = 3;
    SolidLS solid;
    solid.init(ni, nj, nk, dx);
    // the domain: fastCalculateSignedDistanceField without velocity data, then
    // negate, which computes the velocity grids of the negated level set
    oracle_solid_box_phi(ni, nj, nk, dx, solid.phi.data());
    rs.jitter(dx);  // the box's sign test
    solid.mesh.v.assign(8, V3{0.0f, 0.0f, 0.0f});
    solid.mesh.vel.assign(8, V3{0.0f, 0.0f, 0.0f});
    velocityGrids(solid);
    solid.mesh = TriMesh();  // the domain's mesh plays no further part
    SolidLS temp;
    temp.init(ni, nj, nk, dx);
    for (int pass = 0; pass < 2; pass++) {  // static objects, then animated ones
        const bool animated = pass == 1;
        for (const MovingObstacle &o : obs) {
            if (o.animated != animated || o.inversed) continue;
            temp.reset();
            obstacleLevelSetVel(meshAtProgress(o, progress, frameDt), band, rs, temp, o.expansion);
            unionLS(solid, temp);
        }
        bool anyInv = false;
        for (const MovingObstacle &o : obs) anyInv = anyInv || (o.animated == animated && o.inversed);
        if (!anyInv) continue;
        SolidLS inv;
        inv.init(ni, nj, nk, dx);
        inv.vel = false;
        for (const MovingObstacle &o : obs) {
            if (o.animated != animated || !o.inversed) continue;
            temp.reset();
            temp.vel = false;
            obstacleLevelSetVel(meshAtProgress(o, progress, frameDt), band, rs, temp, o.expansion);
            unionLS(inv, temp);
        }
        inv.vel = true;
        negateLS(inv);
        unionLS(solid, inv);
        temp.vel = true;
    }
    phi = solid.phi;
    normalizeLS(solid, u, v, w);
}
}  // namespace

/* The solid update with moving obstacles (fluidsimulation.cpp:5219-5426,
 * meshlevelset.cpp velocity data): objects as flat arrays, one triangle list
 * per object and three vertex sets (previous, current, next frame); flags bit
 * 0 animated, bit 1 inversed.  frame_progress is the substep's fraction of the
 * frame (1 - (remaining + step) / frame dt).  Writes the nodal solid phi and
 * the normalized, extrapolated solid face velocities.  Returns the rand()
 * values used, or -1. */
int oracle_solid_update(int ni, int nj, int nk, double dx, int nobj, const int *nv, const int *nt,
                        const float *prev, const float *cur, const float *next, const int *tris, const int *flags,
                        const float *expansion, const float *vscale, double frame_dt, float frame_progress,
                        const int *rnd, int nrnd, int rand_max, float *phi_nodes, float *us, float *vs, float *ws) {
    try {
        std::vector<MovingObstacle> obs;
        size_t vo = 0, to = 0;
        for (int q = 0; q < nobj; q++) {
            MovingObstacle o;
            o.prev = meshFrom(nv[q], prev + 3 * vo, nt[q], tris + 3 * to);
            o.cur = meshFrom(nv[q], cur + 3 * vo, nt[q], tris + 3 * to);
            o.next = meshFrom(nv[q], next + 3 * vo, nt[q], tris + 3 * to);
            o.animated = (flags[q] & 1) != 0;
            o.inversed = (flags[q] & 2) != 0;
            o.expansion = expansion ? expansion[q] : 0.0f;
            o.velocityScale = vscale ? vscale[q] : 1.0f;
            obs.push_back(o);
            vo += nv[q];
            to += nt[q];
        }
        RandSource rs{rnd, nrnd, 0, (double)rand_max};
        std::vector<float> phi, u, v, w;
        solidUpdate(ni, nj, nk, dx, obs, frame_dt, frame_progress, rs, phi, u, v, w);
        std::memcpy(phi_nodes, phi.data(), sizeof(float) * phi.size());
        std::memcpy(us, u.data(), sizeof(float) * u.size());
        std::memcpy(vs, v.data(), sizeof(float) * v.size());
        std::memcpy(ws, w.data(), sizeof(float) * w.size());
        return rs.used;
    } catch (const std::exception &) {
        return -1;
    }
}

/* FluidSimulation::_addStaticObjectsToSDF with static mesh obstacles
 * (fluidsimulation.cpp:5277-5339): the domain box SDF (negated), then the
 * union of every normal obstacle's level set, then the union of the negated
 * union of the inversed obstacles' level sets.  The domain box's own sign
 * test draws one jitter vector (three rand() values) before the obstacles. */
int oracle_solid_phi_meshes(int ni, int nj, int nk, double dx, int nobj, const int *nv, const float *verts,
                            const int *nt, const int *tris, const int *inversed, const float *expansion,
                            const int *rnd, int nrnd, int rand_max, float *phi_nodes) {
    try {
        const int band = 3;
        const size_t nn = (size_t)(ni + 1) * (nj + 1) * (nk + 1);
        oracle_solid_box_phi(ni, nj, nk, dx, phi_nodes);
        RandSource rs{rnd, nrnd, 0, (double)rand_max};
        rs.jitter(dx);  // the box's sign test
        std::vector<float> solid(phi_nodes, phi_nodes + nn);
        const float ub = (float)(((ni + 1) + (nj + 1) + (nk + 1)) * dx);
        std::vector<TriMesh> meshes;
        size_t vo = 0, to = 0;
        for (int q = 0; q < nobj; q++) {
            meshes.push_back(meshFrom(nv[q], verts + 3 * vo, nt[q], tris + 3 * to));
            vo += nv[q];
            to += nt[q];
        }
        for (int q = 0; q < nobj; q++) {
            if (inversed[q]) continue;
            std::vector<float> temp(nn, ub);
            obstacleLevelSet(meshes[q], ni, nj, nk, dx, band, rs, temp, expansion ? expansion[q] : 0.0f);
            unionInto(solid, ni + 1, nj + 1, nk + 1, temp, ni + 1, nj + 1, nk + 1, GI{0, 0, 0});
        }
        bool anyInv = false;
        std::vector<float> inv(nn, ub);
        for (int q = 0; q < nobj; q++) {
            if (!inversed[q]) continue;
            anyInv = true;
            std::vector<float> temp(nn, ub);
            obstacleLevelSet(meshes[q], ni, nj, nk, dx, band, rs, temp, expansion ? expansion[q] : 0.0f);
            unionInto(inv, ni + 1, nj + 1, nk + 1, temp, ni + 1, nj + 1, nk + 1, GI{0, 0, 0});
        }
        if (anyInv) {
            for (float &p : inv) p = -p;
            unionInto(solid, ni + 1, nj + 1, nk + 1, inv, ni + 1, nj + 1, nk + 1, GI{0, 0, 0});
        }
        std::memcpy(phi_nodes, solid.data(), sizeof(float) * nn);
        return rs.used;
    } catch (const std::exception &) {
        return -1;
    }
}

/* Counting sort by the cell index the engine computes for a particle
 * (Grid3d::positionToGridIndex, grid3d.h:53-58; used e.g. at
 * fluidsimulation.cpp:7832).  Stable: within a cell, ascending particle index. */
void oracle_cell_sort(int ni, int nj, int nk, double dx, int np, const float *pos, int *keys,
                      int *counts, int *offsets, int *perm) {
    size_t N = (size_t)ni * nj * nk;
    for (size_t c = 0; c < N; c++) counts[c] = 0;
    for (int p = 0; p < np; p++) {
        V3 x{pos[3 * p], pos[3 * p + 1], pos[3 * p + 2]};
        GI g = posToIndex(x, dx);
        int key = inRange(g.i, g.j, g.k, ni, nj, nk) ? (int)flat(g.i, g.j, g.k, ni, nj) : -1;
        keys[p] = key;
        if (key >= 0) counts[key]++;
    }
    offsets[0] = 0;
    for (size_t c = 0; c < N; c++) offsets[c + 1] = offsets[c] + counts[c];
    std::vector<int> cur(offsets, offsets + N);
    for (int p = 0; p < np; p++)
        if (keys[p] >= 0) perm[cur[keys[p]]++] = p;
}

namespace {

// GridUtils::featherGrid26 (gridutils.cpp:264-300)
void feather26(std::vector<char> &g, int ni, int nj, int nk) {
    std::vector<char> src = g;
    for (int k = 0; k < nk; k++)
        for (int j = 0; j < nj; j++)
            for (int i = 0; i < ni; i++) {
                if (!src[flat(i, j, k, ni, nj)]) continue;
                for (int dk = -1; dk <= 1; dk++)
                    for (int dj = -1; dj <= 1; dj++)
                        for (int di = -1; di <= 1; di++)
                            if (inRange(i + di, j + dj, k + dk, ni, nj, nk))
                                g[flat(i + di, j + dj, k + dk, ni, nj)] = 1;
            }
}

// GridUtils::featherGrid6 (gridutils.cpp:229-262)
void feather6(std::vector<char> &g, int ni, int nj, int nk) {
    std::vector<char> src = g;
    const int d[6][3] = {{-1, 0, 0}, {1, 0, 0}, {0, -1, 0}, {0, 1, 0}, {0, 0, -1}, {0, 0, 1}};
    for (int k = 0; k < nk; k++)
        for (int j = 0; j < nj; j++)
            for (int i = 0; i < ni; i++) {
                if (!src[flat(i, j, k, ni, nj)]) continue;
                for (int q = 0; q < 6; q++)
                    if (inRange(i + d[q][0], j + d[q][1], k + d[q][2], ni, nj, nk))
                        g[flat(i + d[q][0], j + d[q][1], k + d[q][2], ni, nj)] = 1;
            }
}

// Block membership of one point, shared by the particle level set
// (particlelevelset.cpp:466-521) and the velocity advector
// (velocityadvector.cpp:296-353): a point whose +-sr box lies strictly inside
// its own block belongs to that block only, otherwise to every active block
// its box overlaps.  Returns the member block indices.
int blockMembers(V3 p, float sr, float blockdx, int bi, int bj, int bk,
                 const std::vector<char> &active, GI *out) {
    GI b = posToIndex(p, blockdx);
    V3 bp = indexToPos(b.i, b.j, b.k, blockdx);
    if (p.x - sr > bp.x && p.y - sr > bp.y && p.z - sr > bp.z && p.x + sr < bp.x + blockdx &&
        p.y + sr < bp.y + blockdx && p.z + sr < bp.z + blockdx) {
        if (inRange(b.i, b.j, b.k, bi, bj, bk) && active[flat(b.i, b.j, b.k, bi, bj)]) {
            out[0] = b;
            return 1;
        }
        return 0;
    }
    GI g0 = posToIndexD(p.x - sr, p.y - sr, p.z - sr, blockdx);
    GI g1 = posToIndexD(p.x + sr, p.y + sr, p.z + sr, blockdx);
    int n = 0;
    for (int k = g0.k; k <= g1.k; k++)
        for (int j = g0.j; j <= g1.j; j++)
            for (int i = g0.i; i <= g1.i; i++)
                if (inRange(i, j, k, bi, bj, bk) && active[flat(i, j, k, bi, bj)]) out[n++] = GI{i, j, k};
    return n;
}

int g_threads = 1;  // oracle_set_threads

// The member particles of every block, each list in ascending particle
// index (members(p, out) writes p's member blocks' flat indices and returns
// their count): threads count their contiguous particle range's memberships per
// block, the counts are scanned in (block, thread) order and the ranges
// filled in, so every list holds the serial loop's order.
struct BlockLists {
    std::vector<size_t> start;  // nblocks + 1
    std::vector<int> ids;
};
template <class Members>
BlockLists blockLists(int np, size_t nblocks, Members members) {
    const int T = g_threads;
    std::vector<size_t> cnt((size_t)T * nblocks, 0);
    auto range = [&](int t, int &lo, int &hi) {
        lo = (int)((long long)np * t / T);
        hi = (int)((long long)np * (t + 1) / T);
    };
#pragma omp parallel for num_threads(T) schedule(static, 1)
    for (int t = 0; t < T; t++) {
        int lo, hi;
        range(t, lo, hi);
        size_t m[27];
        size_t *c = &cnt[(size_t)t * nblocks];
        for (int p = lo; p < hi; p++) {
            int n = members(p, m);
            for (int q = 0; q < n; q++) c[m[q]]++;
        }
    }
    BlockLists L;
    L.start.assign(nblocks + 1, 0);
    std::vector<size_t> at((size_t)T * nblocks);
    size_t run = 0;
    for (size_t b = 0; b < nblocks; b++) {
        L.start[b] = run;
        for (int t = 0; t < T; t++) {
            at[(size_t)t * nblocks + b] = run;
            run += cnt[(size_t)t * nblocks + b];
        }
    }
    L.start[nblocks] = run;
    L.ids.resize(run);
#pragma omp parallel for num_threads(T) schedule(static, 1)
    for (int t = 0; t < T; t++) {
        int lo, hi;
        range(t, lo, hi);
        size_t m[27];
        size_t *a = &at[(size_t)t * nblocks];
        for (int p = lo; p < hi; p++) {
            int n = members(p, m);
            for (int q = 0; q < n; q++) L.ids[a[m[q]]++] = p;
        }
    }
    return L;
}

}  // namespace

/* ParticleLevelSet::_computeSignedDistanceFromParticles (particlelevelset.cpp:335-411)
 * with _initializeBlockGrid (413-452), _computeGridCountDataThread (484-537) and
 * _computeExactBandProducerThread (620-666).  Min over the member particles of
 * |cell centre - p| - r, in block-local coordinates, clipped at 3*dx. */
void oracle_liquid_sdf(int ni, int nj, int nk, double dx, int np, const float *pos,
                       double radius, float *phi) {
    const int bw = 10;
    float maxd = (float)(3.0 * dx);
    size_t N = (size_t)ni * nj * nk;
    for (size_t c = 0; c < N; c++) phi[c] = maxd;
    if (np == 0) return;
    int bi = (ni + bw - 1) / bw, bj = (nj + bw - 1) / bw, bk = (nk + bw - 1) / bw;
    float blockdx = bw * dx;
    std::vector<char> active((size_t)bi * bj * bk, 0);
    for (int p = 0; p < np; p++) {
        V3 x{pos[3 * p], pos[3 * p + 1], pos[3 * p + 2]};
        GI g = posToIndex(x, blockdx);
        if (inRange(g.i, g.j, g.k, bi, bj, bk)) active[flat(g.i, g.j, g.k, bi, bj)] = 1;
    }
    feather26(active, bi, bj, bk);
    float r = radius;
    float sr = 2.0f * r;
    const size_t nb = (size_t)bi * bj * bk;
    BlockLists L = blockLists(np, nb, [&](int p, size_t *out) {
        V3 x{pos[3 * p], pos[3 * p + 1], pos[3 * p + 2]};
        GI m[27];
        int n = blockMembers(x, sr, blockdx, bi, bj, bk, active, m);
        for (int q = 0; q < n; q++) out[q] = flat(m[q].i, m[q].j, m[q].k, bi, bj);
        return n;
    });
    // each block writes only its own cells, taking its members in ascending index
#pragma omp parallel for num_threads(g_threads) schedule(dynamic, 8)
    for (size_t fb = 0; fb < nb; fb++) {
        const GI b{(int)(fb % bi), (int)((fb / bi) % bj), (int)(fb / ((size_t)bi * bj))};
        for (size_t e = L.start[fb]; e < L.start[fb + 1]; e++) {
            const int p = L.ids[e];
            V3 x{pos[3 * p], pos[3 * p + 1], pos[3 * p + 2]};
            V3 off = indexToPos(b.i, b.j, b.k, bw * dx);
            V3 pl = x - off;
            GI g0 = posToIndex(V3{pl.x - sr, pl.y - sr, pl.z - sr}, dx);
            GI g1 = posToIndex(V3{pl.x + sr, pl.y + sr, pl.z + sr}, dx);
            g0.i = std::max(g0.i, 0); g0.j = std::max(g0.j, 0); g0.k = std::max(g0.k, 0);
            g1.i = std::min(g1.i, bw - 1); g1.j = std::min(g1.j, bw - 1); g1.k = std::min(g1.k, bw - 1);
            for (int k = g0.k; k <= g1.k; k++)
                for (int j = g0.j; j <= g1.j; j++)
                    for (int i = g0.i; i <= g1.i; i++) {
                        V3 gpos = indexToCellCenter(i, j, k, dx);
                        float dist = length(gpos - pl) - r;
                        int ci = b.i * bw + i, cj = b.j * bw + j, ck = b.k * bw + k;
                        if (!inRange(ci, cj, ck, ni, nj, nk)) continue;
                        size_t c = flat(ci, cj, ck, ni, nj);
                        if (dist < phi[c]) phi[c] = dist;
                    }
        }
    }
}

/* ParticleLevelSet::postProcessSignedDistanceField (particlelevelset.cpp:170-194)
 * with MeshLevelSet::getDistanceAtCellCenter (meshlevelset.cpp:152-162). */
void oracle_liquid_sdf_postprocess(int ni, int nj, int nk, double dx, const float *sp,
                                   float *phi) {
    int pi = ni + 1, pj = nj + 1;
    float eps = 0.005 * dx;
    for (int k = 0; k < nk; k++)
        for (int j = 0; j < nj; j++)
            for (int i = 0; i < ni; i++) {
                size_t c = flat(i, j, k, ni, nj);
                if (phi[c] < 0.5 * dx) {
                    float sc = 0.125f * (sp[flat(i, j, k, pi, pj)] + sp[flat(i + 1, j, k, pi, pj)] +
                                         sp[flat(i, j + 1, k, pi, pj)] + sp[flat(i + 1, j + 1, k, pi, pj)] +
                                         sp[flat(i, j, k + 1, pi, pj)] + sp[flat(i + 1, j, k + 1, pi, pj)] +
                                         sp[flat(i, j + 1, k + 1, pi, pj)] +
                                         sp[flat(i + 1, j + 1, k + 1, pi, pj)]);
                    if (sc < 0) phi[c] = -0.5f * dx;
                }
                float val = phi[c];
                if (std::abs(val) < eps) phi[c] = val > 0 ? eps : -eps;
            }
}

/* VelocityAdvector::_advectGrid for one direction with the FLIP producer
 * (velocityadvector.cpp:74-175, 177-251, 296-353, 467-536).  Contributions to a
 * face come from the face's own 10^3 block, whose particle list is in ascending
 * particle index, so the sums here run in ascending particle index too. */
namespace {
void p2gFlip(int ni, int nj, int nk, double dx, int np, const float *pos, const float *vel, double radius, int dir,
             float *field, unsigned char *valid, float *wsumOut);
}  // namespace

void oracle_p2g(int ni, int nj, int nk, double dx, int np, const float *pos, const float *vel,
                double radius, int dir, float *field, unsigned char *valid) {
    p2gFlip(ni, nj, nk, dx, np, pos, vel, radius, dir, field, valid, nullptr);
}

namespace {
// the FLIP P2G of one direction; wsumOut (optional) receives each face's
// weight sum, the quantity the valid flag thresholds
void p2gFlip(int ni, int nj, int nk, double dx, int np, const float *pos, const float *vel, double radius, int dir,
             float *field, unsigned char *valid, float *wsumOut) {
    const int bw = 10;
    int wi = ni + (dir == 0), wj = nj + (dir == 1), wk = nk + (dir == 2);
    size_t F = (size_t)wi * wj * wk;
    std::vector<float> sum(F, 0.0f), wsum(F, 0.0f);
    int bi = (wi + bw - 1) / bw, bj = (wj + bw - 1) / bw, bk = (wk + bw - 1) / bw;
    double chunkdx = dx * bw;
    V3 offset = dir == 0 ? mk(0.0, 0.5 * dx, 0.5 * dx)
                         : (dir == 1 ? mk(0.5 * dx, 0.0, 0.5 * dx) : mk(0.5 * dx, 0.5 * dx, 0.0));
    std::vector<char> active((size_t)bi * bj * bk, 0);
    for (int p = 0; p < np; p++) {
        V3 x = V3{pos[3 * p], pos[3 * p + 1], pos[3 * p + 2]} - offset;
        GI g = posToIndex(x, chunkdx);
        if (inRange(g.i, g.j, g.k, bi, bj, bk)) active[flat(g.i, g.j, g.k, bi, bj)] = 1;
    }
    feather26(active, bi, bj, bk);
    float eps = 1e-6;
    float r = radius;
    float sr = radius + eps;
    float blockdx = chunkdx;
    float rsq = r * r;
    float coef1 = (4.0f / 9.0f) * (1.0f / (r * r * r * r * r * r));
    float coef2 = (17.0f / 9.0f) * (1.0f / (r * r * r * r));
    float coef3 = (22.0f / 9.0f) * (1.0f / (r * r));
    const size_t nb = (size_t)bi * bj * bk;
    BlockLists L = blockLists(np, nb, [&](int p, size_t *out) {
        V3 x = V3{pos[3 * p], pos[3 * p + 1], pos[3 * p + 2]} - offset;
        GI m[27];
        int n = blockMembers(x, sr, blockdx, bi, bj, bk, active, m);
        for (int q = 0; q < n; q++) out[q] = flat(m[q].i, m[q].j, m[q].k, bi, bj);
        return n;
    });
    // each block writes only its own faces, taking its members in ascending index
#pragma omp parallel for num_threads(g_threads) schedule(dynamic, 8)
    for (size_t fb = 0; fb < nb; fb++) {
        const GI b{(int)(fb % bi), (int)((fb / bi) % bj), (int)(fb / ((size_t)bi * bj))};
        for (size_t e = L.start[fb]; e < L.start[fb + 1]; e++) {
            const int p = L.ids[e];
            V3 x = V3{pos[3 * p], pos[3 * p + 1], pos[3 * p + 2]} - offset;
            float velocity = vel[3 * p + dir];
            V3 off = indexToPos(b.i, b.j, b.k, bw * dx);
            V3 pl = x - off;
            GI g0 = posToIndex(V3{pl.x - sr, pl.y - sr, pl.z - sr}, dx);
            GI g1 = posToIndex(V3{pl.x + sr, pl.y + sr, pl.z + sr}, dx);
            g0.i = std::max(g0.i, 0); g0.j = std::max(g0.j, 0); g0.k = std::max(g0.k, 0);
            g1.i = std::min(g1.i, bw - 1); g1.j = std::min(g1.j, bw - 1); g1.k = std::min(g1.k, bw - 1);
            for (int k = g0.k; k <= g1.k; k++)
                for (int j = g0.j; j <= g1.j; j++)
                    for (int i = g0.i; i <= g1.i; i++) {
                        V3 gpos = indexToPos(i, j, k, dx);
                        V3 v = gpos - pl;
                        float d2 = dot(v, v);
                        if (!(d2 < rsq)) continue;
                        float weight = 1.0f - coef1 * d2 * d2 * d2 + coef2 * d2 * d2 - coef3 * d2;
                        int fi = b.i * bw + i, fj = b.j * bw + j, fk = b.k * bw + k;
                        if (!inRange(fi, fj, fk, wi, wj, wk)) continue;
                        size_t f = flat(fi, fj, fk, wi, wj);
                        sum[f] += weight * velocity;
                        wsum[f] += weight;
                    }
        }
    }
    for (size_t f = 0; f < F; f++) {
        float s = sum[f];
        if (wsum[f] > eps) s /= wsum[f];
        field[f] = s;
        valid[f] = wsum[f] > eps;
    }
    if (wsumOut) std::copy(wsum.begin(), wsum.end(), wsumOut);
}
}  // namespace

/* VelocityAdvector::_advectGrid for one direction with the APIC producer
 * (velocityadvector.cpp:74-175, 177-251, 296-465, 543-623).  Block membership
 * is the FLIP path's (the particle radius plus 1e-6 around each block); in
 * each member block the particle deposits onto the 8 nodes of its cell in
 * that block's frame that lie inside the block.  Sums run in ascending
 * particle index. */
void oracle_p2g_apic(int ni, int nj, int nk, double dx, int np, const float *pos, const float *vel,
                     const float *aff, double radius, int dir, float *field, unsigned char *valid) {
    const int bw = 10;
    int wi = ni + (dir == 0), wj = nj + (dir == 1), wk = nk + (dir == 2);
    size_t F = (size_t)wi * wj * wk;
    std::vector<float> sum(F, 0.0f), wsum(F, 0.0f);
    int bi = (wi + bw - 1) / bw, bj = (wj + bw - 1) / bw, bk = (wk + bw - 1) / bw;
    double chunkdx = dx * bw;
    V3 offset = dir == 0 ? mk(0.0, 0.5 * dx, 0.5 * dx)
                         : (dir == 1 ? mk(0.5 * dx, 0.0, 0.5 * dx) : mk(0.5 * dx, 0.5 * dx, 0.0));
    std::vector<char> active((size_t)bi * bj * bk, 0);
    for (int p = 0; p < np; p++) {
        V3 x = V3{pos[3 * p], pos[3 * p + 1], pos[3 * p + 2]} - offset;
        GI g = posToIndex(x, chunkdx);
        if (inRange(g.i, g.j, g.k, bi, bj, bk)) active[flat(g.i, g.j, g.k, bi, bj)] = 1;
    }
    feather26(active, bi, bj, bk);
    float eps = 1e-6;
    float sr = radius + eps;
    float blockdx = chunkdx;
    GI members[27];
    for (int p = 0; p < np; p++) {
        V3 x = V3{pos[3 * p], pos[3 * p + 1], pos[3 * p + 2]} - offset;
        float velocity = vel[3 * p + dir];
        V3 affine{aff[9 * p + 3 * dir], aff[9 * p + 3 * dir + 1], aff[9 * p + 3 * dir + 2]};
        int nm = blockMembers(x, sr, blockdx, bi, bj, bk, active, members);
        for (int m = 0; m < nm; m++) {
            GI b = members[m];
            V3 pl = x - indexToPos(b.i, b.j, b.k, bw * dx);
            GI g = posToIndex(pl, dx);
            V3 ipos = divide(pl - indexToPos(g.i, g.j, g.k, dx), (float)dx);
            float wx[2] = {1.0f - ipos.x, ipos.x}, wy[2] = {1.0f - ipos.y, ipos.y}, wz[2] = {1.0f - ipos.z, ipos.z};
            for (int q = 0; q < 8; q++) {
                int oi = q & 1, oj = (q >> 1) & 1, ok = q >> 2;
                int li = g.i + oi, lj = g.j + oj, lk = g.k + ok;
                if (li < 0 || lj < 0 || lk < 0 || li >= bw || lj >= bw || lk >= bw) continue;
                float weight = wx[oi] * wy[oj] * wz[ok];
                float apicTerm = dot(affine, indexToPos(li, lj, lk, dx) - pl);
                int fi = b.i * bw + li, fj = b.j * bw + lj, fk = b.k * bw + lk;
                if (!inRange(fi, fj, fk, wi, wj, wk)) continue;
                size_t f = flat(fi, fj, fk, wi, wj);
                sum[f] += weight * (velocity + apicTerm);
                wsum[f] += weight;
            }
        }
    }
    for (size_t f = 0; f < F; f++) {
        float s = sum[f];
        if (wsum[f] > eps) s /= wsum[f];
        field[f] = s;
        valid[f] = wsum[f] > eps;
    }
}

/* FluidSimulation::_getIndicesAndGradientWeights (fluidsimulation.cpp:6709-6769)
 * and _updatePICAPICMarkerParticleVelocitiesThread (6791-6843). */
void oracle_g2p_apic(int ni, int nj, int nk, double dx, const float *u, const float *v, const float *w, int np,
                     const float *pos, float *vel, float *aff) {
    const float *arr[3] = {u, v, w};
    const int wi[3] = {ni + 1, ni, ni}, wj[3] = {nj, nj + 1, nj}, wk[3] = {nk, nk, nk + 1};
    float h = 0.5f * dx;
    float invdx = 1.0f / dx;
    for (int p = 0; p < np; p++) {
        V3 x0{pos[3 * p], pos[3 * p + 1], pos[3 * p + 2]};
        for (int d = 0; d < 3; d++) {
            V3 off = d == 0 ? V3{0.0f, h, h} : (d == 1 ? V3{h, 0.0f, h} : V3{h, h, 0.0f});
            V3 x = x0 - off;
            GI g = posToIndex(x, dx);
            V3 ip = divide(x - indexToPos(g.i, g.j, g.k, dx), (float)dx);
            // the reference's eight gradient weights, factor by factor
            V3 gw[8] = {
                V3{-invdx * (1.0f - ip.y) * (1.0f - ip.z), -invdx * (1.0f - ip.x) * (1.0f - ip.z),
                   -invdx * (1.0f - ip.x) * (1.0f - ip.y)},
                V3{invdx * (1.0f - ip.y) * (1.0f - ip.z), ip.x * (-invdx) * (1.0f - ip.z),
                   ip.x * (1.0f - ip.y) * (-invdx)},
                V3{(-invdx) * ip.y * (1.0f - ip.z), (1.0f - ip.x) * invdx * (1.0f - ip.z),
                   (1.0f - ip.x) * ip.y * (-invdx)},
                V3{invdx * ip.y * (1.0f - ip.z), ip.x * invdx * (1.0f - ip.z), ip.x * ip.y * (-invdx)},
                V3{(-invdx) * (1.0f - ip.y) * ip.z, (1.0f - ip.x) * (-invdx) * ip.z,
                   (1.0f - ip.x) * (1.0f - ip.y) * invdx},
                V3{invdx * (1.0f - ip.y) * ip.z, ip.x * (-invdx) * ip.z, ip.x * (1.0f - ip.y) * invdx},
                V3{(-invdx) * ip.y * ip.z, (1.0f - ip.x) * invdx * ip.z, (1.0f - ip.x) * ip.y * invdx},
                V3{invdx * ip.y * ip.z, ip.x * invdx * ip.z, ip.x * ip.y * invdx}};
            V3 a{0.0f, 0.0f, 0.0f};
            for (int q = 0; q < 8; q++) {
                int i = g.i + (q & 1), j = g.j + ((q >> 1) & 1), k = g.k + (q >> 2);
                if (!inRange(i, j, k, wi[d], wj[d], wk[d])) continue;
                float fv = arr[d][flat(i, j, k, wi[d], wj[d])];
                a = a + V3{gw[q].x * fv, gw[q].y * fv, gw[q].z * fv};
            }
            aff[9 * p + 3 * d] = a.x;
            aff[9 * p + 3 * d + 1] = a.y;
            aff[9 * p + 3 * d + 2] = a.z;
        }
    }
    oracle_sample_velocity(ni, nj, nk, dx, u, v, w, np, pos, vel);
}

/* GridUtils::extrapolateGrid (gridutils.h:95-160) with _initializeStatusGridThread
 * (gridutils.cpp:99-118), _findExtrapolationCells (120-175) and
 * _extrapolateCellsThread (gridutils.h:43-92). */
void oracle_extrapolate(int wi, int wj, int wk, float *field, const unsigned char *valid,
                        int layers) {
    const char UNKNOWN = 0, WAITING = 1, KNOWN = 2, DONE = 3;
    size_t F = (size_t)wi * wj * wk;
    std::vector<char> st(F, UNKNOWN);
    for (int k = 0; k < wk; k++)
        for (int j = 0; j < wj; j++)
            for (int i = 0; i < wi; i++) {
                size_t f = flat(i, j, k, wi, wj);
                if (i == 0 || j == 0 || k == 0 || i == wi - 1 || j == wj - 1 || k == wk - 1)
                    st[f] = DONE;
                else if (valid[f])
                    st[f] = KNOWN;
            }
    const long long sx = 1, sy = wi, sz = (long long)wi * wj;
    const long long nb[6] = {sx, -sx, sy, -sy, sz, -sz};
    std::vector<size_t> cells;
    for (int layer = 0; layer < layers; layer++) {
        cells.clear();
        for (size_t f = 0; f < F; f++) {
            if (st[f] != KNOWN) continue;
            for (int q = 0; q < 6; q++) {
                size_t n = (size_t)((long long)f + nb[q]);
                if (st[n] == UNKNOWN) {
                    st[n] = WAITING;
                    cells.push_back(n);
                }
            }
            st[f] = DONE;
        }
        for (size_t c : cells) {
            float s = 0.0f;
            int count = 0;
            for (int q = 0; q < 6; q++) {
                size_t n = (size_t)((long long)c + nb[q]);
                if (st[n] == DONE) {
                    s += field[n];
                    count++;
                }
            }
            field[c] = s / (float)count;
        }
        if (layer != layers - 1)
            for (size_t c : cells) st[c] = KNOWN;
    }
}

namespace {

// LevelsetUtils::fractionInside, two-point form (levelsetutils.cpp:39-51)
float fractionInside2(float l, float r) {
    if (l < 0 && r < 0) return 1;
    if (l < 0 && r >= 0) return l / (l - r);
    if (l >= 0 && r < 0) return r / (r - l);
    return 0;
}
void cycle4(float *a) {
    float t = a[0];
    a[0] = a[1];
    a[1] = a[2];
    a[2] = a[3];
    a[3] = t;
}
// LevelsetUtils::fractionInside, four-corner form (levelsetutils.cpp:62-141)
float fractionInside4(float bl, float br, float tl, float tr) {
    int in = (bl < 0 ? 1 : 0) + (tl < 0 ? 1 : 0) + (br < 0 ? 1 : 0) + (tr < 0 ? 1 : 0);
    float L[4] = {bl, br, tr, tl};
    if (in == 4) return 1;
    if (in == 3) {
        while (L[0] < 0) cycle4(L);
        float s0 = 1 - fractionInside2(L[0], L[3]);
        float s1 = 1 - fractionInside2(L[0], L[1]);
        return 1.0f - 0.5f * s0 * s1;
    }
    if (in == 2) {
        while (L[0] >= 0 || !(L[1] < 0 || L[2] < 0)) cycle4(L);
        if (L[1] < 0) {
            float sl = fractionInside2(L[0], L[3]);
            float sr = fractionInside2(L[1], L[2]);
            return 0.5f * (sl + sr);
        }
        float mid = 0.25f * (L[0] + L[1] + L[2] + L[3]);
        if (mid < 0) {
            float area = 0;
            float s1 = 1 - fractionInside2(L[0], L[3]);
            float s3 = 1 - fractionInside2(L[2], L[3]);
            area += 0.5f * s1 * s3;
            float s2 = 1 - fractionInside2(L[2], L[1]);
            float s0 = 1 - fractionInside2(L[0], L[1]);
            area += 0.5f * s0 * s2;
            return 1.0f - area;
        }
        float area = 0;
        float s0 = fractionInside2(L[0], L[1]);
        float s1 = fractionInside2(L[0], L[3]);
        area += 0.5f * s0 * s1;
        float s2 = fractionInside2(L[2], L[1]);
        float s3 = fractionInside2(L[2], L[3]);
        area += 0.5f * s2 * s3;
        return area;
    }
    if (in == 1) {
        while (L[0] >= 0) cycle4(L);
        float s0 = fractionInside2(L[0], L[3]);
        float s1 = fractionInside2(L[0], L[1]);
        return 0.5f * s0 * s1;
    }
    return 0;
}
float clampf(float v, float lo, float hi) { return std::max(lo, std::min(v, hi)); }

}  // namespace

/* FluidSimulation::_updateWeightGridThread (fluidsimulation.cpp:6110-6150) with
 * MeshLevelSet::getFaceWeightU/V/W (meshlevelset.cpp:357-391). */
void oracle_face_weights(int ni, int nj, int nk, const float *sp, float *wu, float *wv, float *ww) {
    int pi = ni + 1, pj = nj + 1;
    auto P = [&](int i, int j, int k) { return sp[flat(i, j, k, pi, pj)]; };
    for (int k = 0; k < nk; k++)
        for (int j = 0; j < nj; j++)
            for (int i = 0; i < ni + 1; i++)
                wu[flat(i, j, k, ni + 1, nj)] =
                    clampf(1.0f - fractionInside4(P(i, j, k), P(i, j + 1, k), P(i, j, k + 1), P(i, j + 1, k + 1)), 0.0f, 1.0f);
    for (int k = 0; k < nk; k++)
        for (int j = 0; j < nj + 1; j++)
            for (int i = 0; i < ni; i++)
                wv[flat(i, j, k, ni, nj + 1)] =
                    clampf(1.0f - fractionInside4(P(i, j, k), P(i, j, k + 1), P(i + 1, j, k), P(i + 1, j, k + 1)), 0.0f, 1.0f);
    for (int k = 0; k < nk + 1; k++)
        for (int j = 0; j < nj; j++)
            for (int i = 0; i < ni; i++)
                ww[flat(i, j, k, ni, nj)] =
                    clampf(1.0f - fractionInside4(P(i, j, k), P(i, j + 1, k), P(i + 1, j, k), P(i + 1, j + 1, k)), 0.0f, 1.0f);
}

namespace {

float nodalTrilinear(const float *g, int pi, int pj, int pk, double dx, V3 p);
double nodalTrilinearD(const float *g, int pi, int pj, int pk, double dx, V3 p);

// ---- surface tension ----
//
// ParticleLevelSet::calculateCurvatureGrid (particlelevelset.cpp:196-292): the
// liquid SDF reinitialized by LevelSetSolver::reinitializeUpwind
// (levelsetsolver.cpp:82-205, 317-390) over the cells near the surface, the
// mean curvature of that field (particlelevelset.cpp:692-823), extrapolated
// 3 layers.  The solver's second buffer is allocated uninitialized in the
// reference (array3d.h:706): after the first swap its cells outside the
// solver region are read as neighbours.  Here that buffer starts at 0, which
// is what a fresh page-mapped allocation holds.
void curvatureGrid(int ni, int nj, int nk, double dx, const float *phi, std::vector<float> &surface,
                   std::vector<float> &kgrid) {
    const size_t N = (size_t)ni * nj * nk;
    const int band = 3, layers = 3;
    const float outOfRange = 5.0f;
    std::vector<unsigned char> valid(N, 0);
    const float maxSurfaceDist = 2.0f * dx;
    for (size_t c = 0; c < N; c++)
        if (std::abs(phi[c]) < maxSurfaceDist) valid[c] = 1;
    const int bw = 2 * band;
    const int bi = (ni + bw - 1) / bw, bj = (nj + bw - 1) / bw, bk = (nk + bw - 1) / bw;
    std::vector<unsigned char> blocks((size_t)bi * bj * bk, 0);
    for (int k = 0; k < nk; k++)
        for (int j = 0; j < nj; j++)
            for (int i = 0; i < ni; i++)
                if (valid[flat(i, j, k, ni, nj)]) blocks[flat(i / bw, j / bw, k / bw, bi, bj)] = 1;
    {  // GridUtils::featherGrid6 (gridutils.cpp:229-262)
        std::vector<unsigned char> src = blocks;
        const int nb[6][3] = {{-1, 0, 0}, {1, 0, 0}, {0, -1, 0}, {0, 1, 0}, {0, 0, -1}, {0, 0, 1}};
        for (int k = 0; k < bk; k++)
            for (int j = 0; j < bj; j++)
                for (int i = 0; i < bi; i++) {
                    if (!src[flat(i, j, k, bi, bj)]) continue;
                    for (auto &o : nb) {
                        int a = i + o[0], b = j + o[1], c = k + o[2];
                        if (inRange(a, b, c, bi, bj, bk)) blocks[flat(a, b, c, bi, bj)] = 1;
                    }
                }
    }
    size_t numValid = 0;
    for (int k = 0; k < nk; k++)
        for (int j = 0; j < nj; j++)
            for (int i = 0; i < ni; i++)
                if (blocks[flat(i / bw, j / bw, k / bw, bi, bj)]) {
                    valid[flat(i, j, k, ni, nj)] = 1;
                    numValid++;
                }
    // the solver's cell list is sized numValid and then appended to, so it
    // starts with numValid copies of cell (0, 0, 0)
    std::vector<size_t> cells(numValid, 0);
    for (size_t c = 0; c < N; c++)
        if (valid[c]) cells.push_back(c);
    // reinitializeUpwind with max CFL 0.5 and the 0.01 dx stopping threshold
    const float fdx = (float)dx;
    const float width = band * dx;
    auto sign = [&](const std::vector<float> &g, size_t c) {
        double d = g[c];
        return (float)(d / std::sqrt(d * d + fdx * fdx));
    };
    std::vector<float> in(phi, phi + N);
    float maxS = -std::numeric_limits<float>::max();
    for (size_t c = 0; c < N; c++) maxS = std::max(sign(in, c), maxS);
    const float maxCFL = 0.5f;
    float dtau = maxCFL * fdx;
    while (dtau * maxS / fdx > maxCFL) dtau *= 0.5;
    const int iterations = (int)std::ceil(width / dtau);
    std::vector<float> bufA = in, bufB(N, 0.0f);
    std::vector<float> *out = &bufA, *tmp = &bufB;
    float lastMaxDiff = -1.0f;
    const float invdx = 1.0f / fdx;
    auto sq = [](float x) { return x * x; };
    for (int n = 0; n < iterations; n++) {
        const std::vector<float> &o = *out;
        for (size_t c : cells) {
            const int i = (int)(c % ni), j = (int)((c / ni) % nj), k = (int)(c / ((size_t)ni * nj));
            const float s = sign(o, c);
            const int im = i < 1 ? 0 : i - 1, ip = std::min(i + 1, ni - 1);
            const int jm = j < 1 ? 0 : j - 1, jp = std::min(j + 1, nj - 1);
            const int km = k < 1 ? 0 : k - 1, kp = std::min(k + 1, nk - 1);
            const float C0 = o[c];
            const float dx0 = invdx * (C0 - o[flat(im, j, k, ni, nj)]), dx1 = invdx * (o[flat(ip, j, k, ni, nj)] - C0);
            const float dy0 = invdx * (C0 - o[flat(i, jm, k, ni, nj)]), dy1 = invdx * (o[flat(i, jp, k, ni, nj)] - C0);
            const float dz0 = invdx * (C0 - o[flat(i, j, km, ni, nj)]), dz1 = invdx * (o[flat(i, j, kp, ni, nj)] - C0);
            const float val =
                C0 -
                dtau * std::max(s, 0.0f) *
                    (std::sqrt(sq(std::max(dx0, 0.0f)) + sq(std::min(dx1, 0.0f)) + sq(std::max(dy0, 0.0f)) +
                               sq(std::min(dy1, 0.0f)) + sq(std::max(dz0, 0.0f)) + sq(std::min(dz1, 0.0f))) -
                     1.0f) -
                dtau * std::min(s, 0.0f) *
                    (std::sqrt(sq(std::min(dx0, 0.0f)) + sq(std::max(dx1, 0.0f)) + sq(std::min(dy0, 0.0f)) +
                               sq(std::max(dy1, 0.0f)) + sq(std::min(dz0, 0.0f)) + sq(std::max(dz1, 0.0f))) -
                     1.0f);
            (*tmp)[c] = val;
        }
        float maxDiff = 0;
        for (size_t c : cells) maxDiff = std::max(std::abs((*tmp)[c] - (*out)[c]), maxDiff);
        std::swap(tmp, out);
        if (std::abs(maxDiff - lastMaxDiff) < 0.01f * fdx) break;
        lastMaxDiff = maxDiff;
    }
    surface = *out;
    const float outDist = outOfRange * dx;
    for (size_t c = 0; c < N; c++)
        if (!valid[c]) surface[c] = outDist;
    // _getValidCurvatureNodes: within (band - 1) dx, with all six neighbours
    const float upper = (band - 1) * dx;
    std::vector<unsigned char> near(N, 0), cv(N, 0);
    for (size_t c = 0; c < N; c++)
        if (std::abs(surface[c]) < upper) near[c] = 1;
    for (int k = 1; k < nk - 1; k++)
        for (int j = 1; j < nj - 1; j++)
            for (int i = 1; i < ni - 1; i++) {
                auto T = [&](int a, int b, int e) { return near[flat(a, b, e, ni, nj)] != 0; };
                if (T(i, j, k) && T(i + 1, j, k) && T(i - 1, j, k) && T(i, j + 1, k) && T(i, j - 1, k) &&
                    T(i, j, k + 1) && T(i, j, k - 1))
                    cv[flat(i, j, k, ni, nj)] = 1;
            }
    kgrid.assign(N, 0.0f);
    const float kmax = 1.0f / (float)dx;
    for (int k = 0; k < nk; k++)
        for (int j = 0; j < nj; j++)
            for (int i = 0; i < ni; i++) {
                if (!cv[flat(i, j, k, ni, nj)]) continue;
                if (i == 0 || j == 0 || k == 0 || i == ni - 1 || j == nj - 1 || k == nk - 1) continue;
                auto P = [&](int a, int b, int e) { return surface[flat(a, b, e, ni, nj)]; };
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
                den = (float)std::sqrt((double)(den * den * den));
                if (den < 1e-9f) continue;
                float kv = ((xx * (y * y + z * z) + yy * (x * x + z * z) + zz * (x * x + y * y) - 2 * xy * x * y -
                             2 * xz * x * z - 2 * yz * y * z) /
                            den) /
                           dx;
                kv = std::min(kv, kmax);
                kv = std::max(kv, -kmax);
                kgrid[flat(i, j, k, ni, nj)] = kv;
            }
    oracle_extrapolate(ni, nj, nk, kgrid.data(), cv.data(), layers);
}

// PressureSolver::_initializeSurfaceTensionClusterData (pressuresolver.cpp:
// 314-606) with one thread: 4^3 blocks wholly inside or outside the liquid
// mark their cells OK; the domain's border cells are BORDER; from each
// remaining cell beside the surface (flat order) a breadth-first search over
// same-phase cells takes the status of the first marked cell it meets, or OK
// once 36 cells are queued, or BAD when it ends below 36.
enum : char { ST_UNSET = 0, ST_OK_IN = 1, ST_OK_OUT = 2, ST_BAD_IN = 4, ST_BAD_OUT = 8, ST_BORDER = 16 };
void tensionClusterStatus(int ni, int nj, int nk, const float *phi, std::vector<char> &st) {
    const int bw = 4, threshold = 36;
    const int bi = ni / bw, bj = nj / bw, bk = nk / bw;
    const size_t N = (size_t)ni * nj * nk;
    std::vector<char> block((size_t)bi * bj * bk, 0);  // bit 0 inside, bit 1 outside
    for (int k = 0; k < nk; k++)
        for (int j = 0; j < nj; j++)
            for (int i = 0; i < ni; i++) {
                const int a = i / bw, b = j / bw, c = k / bw;
                if (!inRange(a, b, c, bi, bj, bk)) continue;
                block[flat(a, b, c, bi, bj)] |= phi[flat(i, j, k, ni, nj)] < 0 ? 1 : 2;
            }
    st.assign(N, ST_UNSET);
    auto border = [&](int i, int j, int k) {
        return i == 0 || j == 0 || k == 0 || i == ni - 1 || j == nj - 1 || k == nk - 1;
    };
    for (int k = 0; k < nk; k++)
        for (int j = 0; j < nj; j++)
            for (int i = 0; i < ni; i++) {
                const size_t c = flat(i, j, k, ni, nj);
                if (border(i, j, k)) {
                    st[c] = ST_BORDER;
                    continue;
                }
                const int a = i / bw, b = j / bw, e = k / bw;
                if (!inRange(a, b, e, bi, bj, bk)) continue;
                const char bs = block[flat(a, b, e, bi, bj)];
                if ((bs & 1) && !(bs & 2))
                    st[c] = ST_OK_IN;
                else if (!(bs & 1) && (bs & 2))
                    st[c] = ST_OK_OUT;
            }
    const int nb[6][3] = {{1, 0, 0}, {-1, 0, 0}, {0, 1, 0}, {0, -1, 0}, {0, 0, 1}, {0, 0, -1}};
    std::vector<size_t> seeds;
    for (int k = 0; k < nk; k++)
        for (int j = 0; j < nj; j++)
            for (int i = 0; i < ni; i++) {
                const size_t c = flat(i, j, k, ni, nj);
                if (border(i, j, k) || st[c] != ST_UNSET) continue;
                const bool fl = phi[c] < 0.0f;
                for (auto &o : nb) {
                    const float nd = phi[flat(i + o[0], j + o[1], k + o[2], ni, nj)];
                    if (fl ? nd >= 0.0f : nd < 0.0f) {
                        seeds.push_back(c);
                        break;
                    }
                }
            }
    std::vector<unsigned char> processed(N, 0);
    std::vector<size_t> queue;
    for (size_t seed : seeds) {
        if (st[seed] != ST_UNSET) continue;
        queue.clear();
        queue.push_back(seed);
        size_t head = 0;
        processed[seed] = 1;
        const bool fl = phi[seed] < 0.0f;
        char result = ST_UNSET;
        bool stop = false;
        while (head < queue.size()) {
            const size_t g = queue[head++];
            const int i = (int)(g % ni), j = (int)((g / ni) % nj), k = (int)(g / ((size_t)ni * nj));
            for (auto &o : nb) {
                const size_t n = flat(i + o[0], j + o[1], k + o[2], ni, nj);
                if (processed[n]) continue;
                if (!(fl ? phi[n] < 0.0f : phi[n] >= 0.0f)) continue;
                const char ns = st[n];
                if (ns == ST_BORDER) continue;
                if (ns != ST_UNSET) {
                    result = ns;
                    stop = true;
                    break;
                }
                queue.push_back(n);
                processed[n] = 1;
                if ((int)queue.size() == threshold) result = fl ? ST_OK_IN : ST_OK_OUT;
            }
            if (stop) break;
        }
        if (result == ST_UNSET && (int)queue.size() < threshold) result = fl ? ST_BAD_IN : ST_BAD_OUT;
        for (size_t q : queue) {
            st[q] = result;
            processed[q] = 0;
        }
    }
}

// PressureSolver::_getSurfaceTensionTerm (pressuresolver.cpp:715-749): the
// constant times the curvature at the liquid surface's crossing between two
// cells, or 0 when the pair does not cross it or touches a BAD cluster
struct Tension {
    int ni, nj, nk;
    double dx, k;
    const float *phi;
    std::vector<char> status;
    std::vector<float> surface, curvature;
    double term(int i1, int j1, int k1, int i2, int j2, int k2) const {
        const size_t c1 = flat(i1, j1, k1, ni, nj), c2 = flat(i2, j2, k2, ni, nj);
        const char s = status[c1] | status[c2];
        if ((s & ST_BAD_IN) || (s & ST_BAD_OUT)) return 0.0;
        const V3 p1 = indexToCellCenter(i1, j1, k1, dx), p2 = indexToCellCenter(i2, j2, k2, dx);
        const float phi1 = phi[c1], phi2 = phi[c2];
        if ((phi1 < 0.0f && phi2 < 0.0f) || (phi1 >= 0.0f && phi2 >= 0.0f)) return 0.0;
        V3 p;
        const float eps = 1e-6;
        if (std::abs(phi2 - phi1) < eps)
            p = p1 + 0.5f * (p2 - p1);
        else
            p = p1 + (phi1 / (phi1 - phi2)) * (p2 - p1);
        const V3 hdx = mk(0.5 * dx, 0.5 * dx, 0.5 * dx);
        // Interpolation::trilinearInterpolate returns the double sum (interpolation.cpp:72-112)
        const double kv = nodalTrilinearD(curvature.data(), ni, nj, nk, dx, p - hdx);
        return k * kv;
    }
};
void buildTension(int ni, int nj, int nk, double dx, double k, const float *phi, Tension &T) {
    T.ni = ni;
    T.nj = nj;
    T.nk = nk;
    T.dx = dx;
    T.k = k;
    T.phi = phi;
    tensionClusterStatus(ni, nj, nk, phi, T.status);
    curvatureGrid(ni, nj, nk, dx, phi, T.surface, T.curvature);
}

struct PressureSystem {
    int m = 0;
    std::vector<int> rows;                 // flat cell id per row, k,j,i order
    std::vector<int> key;                  // cell -> row or -1
    std::vector<double> rhs, diag, ox, oy, oz, invdiag;
    std::vector<int> nbr;                  // 6 per row: -x,+x,-y,+y,-z,+z
};

/* PressureSolver::_initialize (pressuresolver.cpp:118-159),
 * _calculateNegativeDivergenceVectorThread (624-713; the solid terms are exact
 * no-ops for zero solid velocity, see the file header) and
 * _calculateMatrixCoefficientsThread (767-867) with SparseMatrix::add/set
 * (sparsematrix.h:72-111; add ignores a -1 column). */
//
// us/vs/ws/wc (optional): the solid face velocities and the cell weights (1 -
// the solid's volume fraction) for the divergence's solid terms
// (pressuresolver.cpp:652-657); without them the solid is at rest and those
// terms add +-0.0, which never changes the sum.
// tension (optional): the surface tension terms of the divergence
// (pressuresolver.cpp:659-710)
void buildSystem(int ni, int nj, int nk, double dx, double dt, const float *u, const float *v,
                 const float *w, const float *wu, const float *wv, const float *ww,
                 const float *phi, PressureSystem &S, const float *us = nullptr, const float *vs = nullptr,
                 const float *ws = nullptr, const float *wc = nullptr, const Tension *tension = nullptr) {
    size_t N = (size_t)ni * nj * nk;
    S.key.assign(N, -1);
    S.rows.clear();
    for (int k = 1; k < nk - 1; k++)
        for (int j = 1; j < nj - 1; j++)
            for (int i = 1; i < ni - 1; i++) {
                size_t c = flat(i, j, k, ni, nj);
                if (phi[c] < 0) {
                    S.key[c] = (int)S.rows.size();
                    S.rows.push_back((int)c);
                }
            }
    int m = S.m = (int)S.rows.size();
    S.rhs.assign(m, 0.0);
    S.diag.assign(m, 0.0);
    S.ox.assign(m, 0.0);
    S.oy.assign(m, 0.0);
    S.oz.assign(m, 0.0);
    S.nbr.assign((size_t)6 * m, -1);
    const double factor = 1.0 / dx;
    const double mfactor = dt / (dx * dx);
    const double eps = 1e-9, maxtheta = 25;
    auto U = [&](int i, int j, int k) { return u[flat(i, j, k, ni + 1, nj)]; };
    auto V = [&](int i, int j, int k) { return v[flat(i, j, k, ni, nj + 1)]; };
    auto W = [&](int i, int j, int k) { return w[flat(i, j, k, ni, nj)]; };
    auto WU = [&](int i, int j, int k) { return wu[flat(i, j, k, ni + 1, nj)]; };
    auto WV = [&](int i, int j, int k) { return wv[flat(i, j, k, ni, nj + 1)]; };
    auto WW = [&](int i, int j, int k) { return ww[flat(i, j, k, ni, nj)]; };
    auto PH = [&](int i, int j, int k) { return phi[flat(i, j, k, ni, nj)]; };
    auto KEY = [&](int i, int j, int k) { return S.key[flat(i, j, k, ni, nj)]; };
    for (int r = 0; r < m; r++) {
        int c = S.rows[r];
        int i = c % ni, j = (c / ni) % nj, k = c / (ni * nj);
        double volRight = WU(i + 1, j, k), volLeft = WU(i, j, k);
        double volTop = WV(i, j + 1, k), volBottom = WV(i, j, k);
        double volFront = WW(i, j, k + 1), volBack = WW(i, j, k);
        double div = 0.0;
        div += -factor * volRight * U(i + 1, j, k);
        div += factor * volLeft * U(i, j, k);
        div += -factor * volTop * V(i, j + 1, k);
        div += factor * volBottom * V(i, j, k);
        div += -factor * volFront * W(i, j, k + 1);
        div += factor * volBack * W(i, j, k);
        if (us) {
            double volCenter = wc[c];
            div += factor * (volRight - volCenter) * us[flat(i + 1, j, k, ni + 1, nj)];
            div += -factor * (volLeft - volCenter) * us[flat(i, j, k, ni + 1, nj)];
            div += factor * (volTop - volCenter) * vs[flat(i, j + 1, k, ni, nj + 1)];
            div += -factor * (volBottom - volCenter) * vs[flat(i, j, k, ni, nj + 1)];
            div += factor * (volFront - volCenter) * ws[flat(i, j, k + 1, ni, nj)];
            div += -factor * (volBack - volCenter) * ws[flat(i, j, k, ni, nj)];
        }
        if (tension) {
            const double stfactor = dt / (dx * dx);
            const double pc = PH(i, j, k);
            auto add = [&](double pn, double vol, int a, int b, int e) {
                if (pn >= 0.0) {
                    const double t = tension->term(i, j, k, a, b, e);
                    double theta = (pc - pn) / (pc + eps);
                    theta = std::max(-maxtheta, std::min(theta, maxtheta));
                    div += stfactor * vol * theta * t;
                }
            };
            add(PH(i + 1, j, k), volRight, i + 1, j, k);
            add(PH(i - 1, j, k), volLeft, i - 1, j, k);
            add(PH(i, j + 1, k), volTop, i, j + 1, k);
            add(PH(i, j - 1, k), volBottom, i, j - 1, k);
            add(PH(i, j, k + 1), volFront, i, j, k + 1);
            add(PH(i, j, k - 1), volBack, i, j, k - 1);
        }
        S.rhs[r] = div;

        // density grid is 1.0 everywhere on the plain path (fluidsimulation.cpp:6214)
        double rhoC = 1.0f;
        double rhoR = (rhoC + 1.0f) / 2.0, rhoL = (rhoC + 1.0f) / 2.0, rhoT = (rhoC + 1.0f) / 2.0;
        double rhoB = (rhoC + 1.0f) / 2.0, rhoF = (rhoC + 1.0f) / 2.0, rhoK = (rhoC + 1.0f) / 2.0;
        double phiC = PH(i, j, k);
        double phiR = PH(i + 1, j, k), phiL = PH(i - 1, j, k), phiT = PH(i, j + 1, k);
        double phiB = PH(i, j - 1, k), phiF = PH(i, j, k + 1), phiK = PH(i, j, k - 1);
        double diag = (volRight / rhoR + volLeft / rhoL + volTop / rhoT + volBottom / rhoB +
                       volFront / rhoF + volBack / rhoK) * mfactor;
        int *nb = &S.nbr[(size_t)6 * r];
        auto side = [&](double phiN, double vol, double rho, int nkey, int slot, double *store) {
            if (phiN < 0.0) {
                if (nkey >= 0) {
                    nb[slot] = nkey;
                    if (store) *store = -(vol / rho) * mfactor;
                }
            } else {
                double theta = phiN / (phiC + eps);
                theta = std::max(-maxtheta, std::min(theta, maxtheta));
                diag -= (vol / rho) * mfactor * theta;
            }
        };
        side(phiR, volRight, rhoR, KEY(i + 1, j, k), 1, &S.ox[r]);
        side(phiL, volLeft, rhoL, KEY(i - 1, j, k), 0, nullptr);
        side(phiT, volTop, rhoT, KEY(i, j + 1, k), 3, &S.oy[r]);
        side(phiB, volBottom, rhoB, KEY(i, j - 1, k), 2, nullptr);
        side(phiF, volFront, rhoF, KEY(i, j, k + 1), 5, &S.oz[r]);
        side(phiK, volBack, rhoK, KEY(i, j, k - 1), 4, nullptr);
        diag = std::max(diag, 0.0);
        S.diag[r] = diag;
    }
}

/* factorModifiedIncompleteColesky0 (pcgsolver.h:68-205) in pull form: each row
 * receives its updates from the -z, -y, -x neighbour columns in that (column)
 * order; within a column the "missing" sum runs over the column's other rows
 * in ascending row order. */
void factorMIC0(PressureSystem &S) {
    const double omega = 0.97, ratio = 0.25, eps = 1e-9;
    int m = S.m;
    std::vector<double> adiag(m);
    S.invdiag.assign(m, 0.0);
    for (int r = 0; r < m; r++) {
        double a = S.diag[r];
        if (a < eps) a = 0;
        adiag[r] = a;
        S.invdiag[r] = a;
    }
    // coefficient between row r and its +x/+y/+z neighbour, by slot 1/3/5
    auto up = [&](int r, int slot) { return slot == 1 ? S.ox[r] : (slot == 3 ? S.oy[r] : S.oz[r]); };
    for (int j = 0; j < m; j++) {
        const int *nb = &S.nbr[(size_t)6 * j];
        double d = S.invdiag[j];
        // columns k < j that hold row j: -z (slot 4), -y (slot 2), -x (slot 0)
        const int cslots[3] = {4, 2, 0};
        for (int q = 0; q < 3; q++) {
            int k = nb[cslots[q]];
            if (k < 0 || adiag[k] == 0) continue;
            int jslot = cslots[q] + 1;  // row j is k's +z / +y / +x neighbour
            const int *kb = &S.nbr[(size_t)6 * k];
            double Ljk = up(k, jslot) * S.invdiag[k];
            d -= Ljk * Ljk;
            double missing = 0;
            const int order[3] = {1, 3, 5};  // k's rows in ascending order: +x, +y, +z
            for (int o = 0; o < 3; o++) {
                int s = order[o];
                if (s == jslot || kb[s] < 0) continue;
                missing += up(k, s) * S.invdiag[k];
            }
            d -= omega * Ljk * missing;
        }
        if (adiag[j] == 0) {
            S.invdiag[j] = d;
            continue;
        }
        if (d < ratio * adiag[j])
            S.invdiag[j] = 1 / std::sqrt(adiag[j]);
        else
            S.invdiag[j] = 1 / std::sqrt(d);
    }
}

// L entry of row (k's neighbour in slot) in column k: scaled unless column k is null
inline double Lval(const PressureSystem &S, const std::vector<double> &adiagZero, int k, int slot) {
    double a = slot == 1 ? S.ox[k] : (slot == 3 ? S.oy[k] : S.oz[k]);
    return adiagZero[k] ? a : a * S.invdiag[k];
}

/* solveLower + solveLowerTransposeInPlace (pcgsolver.h:211-241) in pull form. */
void applyMIC0(const PressureSystem &S, const std::vector<double> &nullcol, const double *rv,
               double *z) {
    int m = S.m;
    for (int j = 0; j < m; j++) {
        const int *nb = &S.nbr[(size_t)6 * j];
        double y = rv[j];
        const int cslots[3] = {4, 2, 0};
        for (int q = 0; q < 3; q++) {
            int k = nb[cslots[q]];
            if (k < 0) continue;
            y -= Lval(S, nullcol, k, cslots[q] + 1) * z[k];
        }
        z[j] = y * S.invdiag[j];
    }
    for (int i = m - 1; i >= 0; i--) {
        const int *nb = &S.nbr[(size_t)6 * i];
        double x = z[i];
        const int order[3] = {1, 3, 5};
        for (int o = 0; o < 3; o++) {
            int rr = nb[order[o]];
            if (rr < 0) continue;
            x -= Lval(S, nullcol, i, order[o]) * z[rr];
        }
        z[i] = x * S.invdiag[i];
    }
}

std::vector<double> nullColumns(const PressureSystem &S) {
    std::vector<double> nc(S.m);
    for (int r = 0; r < S.m; r++) nc[r] = (S.diag[r] < 1e-9) ? 1.0 : 0.0;
    return nc;
}

// ThreadUtils::splitRangeIntoIntervals (threadutils.cpp:54-72)
std::vector<int> splitRange(int b, int e, int n) {
    int size = (int)std::floor((double)(e - b) / (double)n);
    int rem = (e - b) - size * n;
    std::vector<int> iv;
    iv.push_back(b);
    int cur = b;
    for (int i = 0; i < n; i++) {
        int end = cur + size;
        if (i < rem) end++;
        iv.push_back(end);
        cur = end;
    }
    return iv;
}

// BLAS::dot (blaswrapper.h:27-57): serial, or per-thread partial sums added in thread order
double blasDot(const std::vector<double> &x, const std::vector<double> &y, int ncpu) {
    int nt = (int)std::fmin(ncpu, std::ceil((float)x.size() / (float)500000));
    if (nt <= 1) {
        double s = 0;
        for (size_t i = 0; i < x.size(); i++) s += x[i] * y[i];
        return s;
    }
    std::vector<int> iv = splitRange(0, (int)x.size(), nt);
    double s = 0;
    for (int t = 0; t < nt; t++) {
        double part = 0;
        for (int i = iv[t]; i < iv[t + 1]; i++) part += x[i] * y[i];
        s += part;
    }
    return s;
}
double absMax(const std::vector<double> &x) {
    double mx = 0;
    for (double v : x)
        if (std::fabs(v) > mx) mx = std::fabs(v);
    return mx;
}

/* FixedSparseMatrix multiply (sparsematrix.h:174-215): each row sums its
 * entries in ascending column order: -z, -y, -x, diag, +x, +y, +z. */
void spmv(const PressureSystem &S, const std::vector<double> &x, std::vector<double> &out) {
#pragma omp parallel for num_threads(g_threads) schedule(static)
    for (int r = 0; r < S.m; r++) {
        const int *nb = &S.nbr[(size_t)6 * r];
        double s = 0;
        if (nb[4] >= 0) s += S.oz[nb[4]] * x[nb[4]];
        if (nb[2] >= 0) s += S.oy[nb[2]] * x[nb[2]];
        if (nb[0] >= 0) s += S.ox[nb[0]] * x[nb[0]];
        s += S.diag[r] * x[r];
        if (nb[1] >= 0) s += S.ox[r] * x[nb[1]];
        if (nb[3] >= 0) s += S.oy[r] * x[nb[3]];
        if (nb[5] >= 0) s += S.oz[r] * x[nb[5]];
        out[r] = s;
    }
}

/* PCGSolver<double>::solve (pcgsolver.h:268-322) with BLAS::addScaled taking a
 * float step length (blaswrapper.h:128-157).  Returns success. */
bool pcg(PressureSystem &S, double tolFactor, int maxIter, int ncpu, std::vector<double> &x,
         double &resid, int &iters) {
    int m = S.m;
    if (tolFactor < 1e-30) tolFactor = 1e-30;
    std::vector<double> r = S.rhs, z(m), s(m);
    x.assign(m, 0.0);
    resid = absMax(r);
    if (resid == 0) {
        iters = 0;
        return true;
    }
    double tol = tolFactor * resid;
    factorMIC0(S);
    std::vector<double> nc = nullColumns(S);
    applyMIC0(S, nc, r.data(), z.data());
    double rho = blasDot(z, r, ncpu);
    if (rho == 0 || rho != rho) {
        iters = 0;
        return false;
    }
    s = z;
    int it;
    for (it = 0; it < maxIter; it++) {
        spmv(S, s, z);
        double alpha = rho / blasDot(s, z, ncpu);
        float fa = alpha, fna = -alpha;
#pragma omp parallel for num_threads(g_threads) schedule(static)
        for (int i = 0; i < m; i++) {
            x[i] += fa * s[i];
            r[i] += fna * z[i];
        }
        resid = absMax(r);
        if (resid <= std::min(tol, 1.0)) {
            iters = it + 1;
            return true;
        }
        applyMIC0(S, nc, r.data(), z.data());
        double rhoNew = blasDot(z, r, ncpu);
        double beta = rhoNew / rho;
        float fb = beta;
#pragma omp parallel for num_threads(g_threads) schedule(static)
        for (int i = 0; i < m; i++) z[i] += fb * s[i];
        s.swap(z);
        rho = rhoNew;
    }
    iters = it;
    return false;
}

}  // namespace

int oracle_pressure_system(int ni, int nj, int nk, double dx, double dt, const float *u,
                           const float *v, const float *w, const float *wu, const float *wv,
                           const float *ww, const float *phi, int *rows, double *rhs,
                           double *diag, double *ox, double *oy, double *oz, int *nbr,
                           double *invdiag) {
    PressureSystem S;
    buildSystem(ni, nj, nk, dx, dt, u, v, w, wu, wv, ww, phi, S);
    factorMIC0(S);
    int m = S.m;
    if (rows) std::copy(S.rows.begin(), S.rows.end(), rows);
    if (rhs) std::copy(S.rhs.begin(), S.rhs.end(), rhs);
    if (diag) std::copy(S.diag.begin(), S.diag.end(), diag);
    if (ox) std::copy(S.ox.begin(), S.ox.end(), ox);
    if (oy) std::copy(S.oy.begin(), S.oy.end(), oy);
    if (oz) std::copy(S.oz.begin(), S.oz.end(), oz);
    if (nbr) std::copy(S.nbr.begin(), S.nbr.end(), nbr);
    if (invdiag) std::copy(S.invdiag.begin(), S.invdiag.end(), invdiag);
    return m;
}

void oracle_apply_mic0(int m, const double *diag, const double *ox, const double *oy,
                       const double *oz, const int *nbr, const double *invdiag, const double *r,
                       double *z) {
    PressureSystem S;
    S.m = m;
    S.diag.assign(diag, diag + m);
    S.ox.assign(ox, ox + m);
    S.oy.assign(oy, oy + m);
    S.oz.assign(oz, oz + m);
    S.nbr.assign(nbr, nbr + (size_t)6 * m);
    S.invdiag.assign(invdiag, invdiag + m);
    std::vector<double> nc = nullColumns(S);
    applyMIC0(S, nc, r, z);
}

/* PressureSolver::solve (pressuresolver.cpp:52-92) and _solveLinearSystem
 * (869-916): early exit when max|rhs| < tol, PCG, float pressure grid, and the
 * acceptable-tolerance rule (maxIterations reached with |r| < 1.0). */
namespace {
// PressureSolver::_conditionSolidVelocityField (pressuresolver.cpp:161-312):
// groups of interior liquid cells connected through faces open to the fluid
// (weight >= 1e-6), never through a border cell; a group of more than one
// cell none of which has an open face to an air cell is an isolated pocket,
// and the solid velocities of all its cells' faces become 0.
void conditionSolidVelocity(int ni, int nj, int nk, const float *phi, const float *wu, const float *wv,
                            const float *ww, float *us, float *vs, float *ws) {
    const float eps = 1e-6f;
    auto U = [&](int i, int j, int k) { return wu[flat(i, j, k, ni + 1, nj)]; };
    auto V = [&](int i, int j, int k) { return wv[flat(i, j, k, ni, nj + 1)]; };
    auto W = [&](int i, int j, int k) { return ww[flat(i, j, k, ni, nj)]; };
    auto L = [&](int i, int j, int k) { return phi[flat(i, j, k, ni, nj)]; };
    auto border = [&](int i, int j, int k) {
        return i == 0 || j == 0 || k == 0 || i == ni - 1 || j == nj - 1 || k == nk - 1;
    };
    std::vector<char> air((size_t)ni * nj * nk, 0), done((size_t)ni * nj * nk, 0);
    for (int k = 0; k < nk; k++)
        for (int j = 0; j < nj; j++)
            for (int i = 0; i < ni; i++) {
                if (border(i, j, k)) continue;
                if ((U(i, j, k) >= 0.0f + eps && L(i - 1, j, k) >= 0.0f) ||
                    (U(i + 1, j, k) >= 0.0f + eps && L(i + 1, j, k) >= 0.0f) ||
                    (V(i, j, k) >= 0.0f + eps && L(i, j - 1, k) >= 0.0f) ||
                    (V(i, j + 1, k) >= 0.0f + eps && L(i, j + 1, k) >= 0.0f) ||
                    (W(i, j, k) >= 0.0f + eps && L(i, j, k - 1) >= 0.0f) ||
                    (W(i, j, k + 1) >= 0.0f + eps && L(i, j, k + 1) >= 0.0f))
                    air[flat(i, j, k, ni, nj)] = 1;
            }
    std::vector<GI> queue, group;
    for (int k = 1; k < nk - 1; k++)
        for (int j = 1; j < nj - 1; j++)
            for (int i = 1; i < ni - 1; i++) {
                size_t c = flat(i, j, k, ni, nj);
                if (L(i, j, k) >= 0.0f) {
                    done[c] = 1;
                    continue;
                }
                if (done[c]) continue;
                queue.assign(1, GI{i, j, k});
                done[c] = 1;
                group.clear();
                while (!queue.empty()) {
                    GI g = queue.back();
                    queue.pop_back();
                    auto visit = [&](int a, int b, int d, float weight) {
                        size_t n = flat(a, b, d, ni, nj);
                        if (!done[n] && L(a, b, d) < 0.0f && weight >= eps) {
                            done[n] = 1;
                            if (!border(a, b, d)) queue.push_back(GI{a, b, d});
                        }
                    };
                    visit(g.i - 1, g.j, g.k, U(g.i, g.j, g.k));
                    visit(g.i + 1, g.j, g.k, U(g.i + 1, g.j, g.k));
                    visit(g.i, g.j - 1, g.k, V(g.i, g.j, g.k));
                    visit(g.i, g.j + 1, g.k, V(g.i, g.j + 1, g.k));
                    visit(g.i, g.j, g.k - 1, W(g.i, g.j, g.k));
                    visit(g.i, g.j, g.k + 1, W(g.i, g.j, g.k + 1));
                    group.push_back(g);
                }
                if (group.size() == 1) continue;
                bool isolated = true;
                for (GI g : group)
                    if (air[flat(g.i, g.j, g.k, ni, nj)]) {
                        isolated = false;
                        break;
                    }
                if (!isolated) continue;
                for (GI g : group) {
                    us[flat(g.i, g.j, g.k, ni + 1, nj)] = 0.0f;
                    us[flat(g.i + 1, g.j, g.k, ni + 1, nj)] = 0.0f;
                    vs[flat(g.i, g.j, g.k, ni, nj + 1)] = 0.0f;
                    vs[flat(g.i, g.j + 1, g.k, ni, nj + 1)] = 0.0f;
                    ws[flat(g.i, g.j, g.k, ni, nj)] = 0.0f;
                    ws[flat(g.i, g.j, g.k + 1, ni, nj)] = 0.0f;
                }
            }
}

// LevelsetUtils volume fractions (levelsetutils.h:60-100, levelsetutils.cpp:
// 213-258) and MeshLevelSet::_getCellWeight (meshlevelset.cpp:1491-1515)
void sort4(float &a, float &b, float &c, float &d) {
    if (a > b) std::swap(a, b);
    if (c > d) std::swap(c, d);
    if (a > c) std::swap(a, c);
    if (b > d) std::swap(b, d);
    if (b > c) std::swap(b, c);
}
float tetFraction(float p0, float p1, float p2, float p3) {
    return p0 * p0 * p0 / ((p0 - p1) * (p0 - p2) * (p0 - p3));
}
float prismFraction(float p0, float p1, float p2, float p3) {
    float a = p0 / (p0 - p2), b = p0 / (p0 - p3), c = p1 / (p1 - p3), d = p1 / (p1 - p2);
    return a * b * (1 - d) + b * (1 - c) * d + c * d;
}
float volumeFraction4(float p0, float p1, float p2, float p3) {
    sort4(p0, p1, p2, p3);
    if (p3 <= 0) return 1;
    if (p2 <= 0) return 1 - tetFraction(p3, p2, p1, p0);
    if (p1 <= 0) return prismFraction(p0, p1, p2, p3);
    if (p0 <= 0) return tetFraction(p0, p1, p2, p3);
    return 0;
}
float volumeFraction8(float p000, float p100, float p010, float p110, float p001, float p101, float p011,
                      float p111) {
    return (volumeFraction4(p000, p001, p101, p011) + volumeFraction4(p000, p101, p100, p110) +
            volumeFraction4(p000, p010, p011, p110) + volumeFraction4(p101, p011, p111, p110) +
            2 * volumeFraction4(p000, p011, p101, p110) + volumeFraction4(p100, p101, p001, p111) +
            volumeFraction4(p100, p001, p000, p010) + volumeFraction4(p100, p110, p111, p010) +
            volumeFraction4(p001, p111, p011, p010) + 2 * volumeFraction4(p100, p111, p001, p010)) /
           12.0f;
}
// the weight grid's cell weights (fluidsimulation.cpp:6140-6147): 1 - the
// solid's volume fraction in each cell, clamped to [0, 1]
void cellWeights(int ni, int nj, int nk, const float *solid, std::vector<float> &wc) {
    wc.assign((size_t)ni * nj * nk, 0.0f);
    const int pi = ni + 1, pj = nj + 1;
    for (int k = 0; k < nk; k++)
        for (int j = 0; j < nj; j++)
            for (int i = 0; i < ni; i++) {
                auto P = [&](int a, int b, int c) { return solid[flat(i + a, j + b, k + c, pi, pj)]; };
                float p000 = P(0, 0, 0), p001 = P(0, 0, 1), p010 = P(0, 1, 0), p011 = P(0, 1, 1);
                float p100 = P(1, 0, 0), p101 = P(1, 0, 1), p110 = P(1, 1, 0), p111 = P(1, 1, 1);
                float weight;
                if (p000 < 0 && p001 < 0 && p010 < 0 && p011 < 0 && p100 < 0 && p101 < 0 && p110 < 0 && p111 < 0)
                    weight = 1.0f;
                else if (p000 >= 0 && p001 >= 0 && p010 >= 0 && p011 >= 0 && p100 >= 0 && p101 >= 0 &&
                         p110 >= 0 && p111 >= 0)
                    weight = 0.0f;
                else
                    weight = volumeFraction8(p000, p100, p010, p110, p001, p101, p011, p111);
                wc[flat(i, j, k, ni, nj)] = clampf(1.0f - weight, 0.0f, 1.0f);
            }
}

// PressureSolver::solve (pressuresolver.cpp:52-92): the moving solid's
// velocities (conditioned in place) enter the divergence when us is given
int pressureSolveImpl(int ni, int nj, int nk, double dx, double dt, const float *u, const float *v,
                      const float *w, const float *wu, const float *wv, const float *ww, const float *phi,
                      double tol, int maxIter, int nthreads, float *pressure, int *status, double *err, float *us,
                      float *vs, float *ws, const float *wc, const Tension *tension = nullptr) {
    PressureSystem S;
    if (us) conditionSolidVelocity(ni, nj, nk, phi, wu, wv, ww, us, vs, ws);
    buildSystem(ni, nj, nk, dx, dt, u, v, w, wu, wv, ww, phi, S, us, vs, ws, wc, tension);
    size_t N = (size_t)ni * nj * nk;
    for (size_t c = 0; c < N; c++) pressure[c] = 0.0f;
    double maxAbs = 0.0;
    for (double x : S.rhs)
        if (std::fabs(x) > maxAbs) maxAbs = std::fabs(x);
    if (maxAbs < tol) {
        status[0] = 0;
        status[1] = 1;
        err[0] = 0.0;
        return S.m;
    }
    std::vector<double> x;
    double resid = -1.0;
    int iters = 0;
    bool ok = pcg(S, tol, maxIter, nthreads, x, resid, iters);
    for (int r = 0; r < S.m; r++) pressure[S.rows[r]] = (float)x[r];
    float e = (float)resid;
    bool ret = ok || (iters == maxIter && e < 1.0);
    status[0] = iters;
    status[1] = ret ? 1 : 0;
    err[0] = e;
    return S.m;
}
}  // namespace

int oracle_pressure_solve(int ni, int nj, int nk, double dx, double dt, const float *u,
                          const float *v, const float *w, const float *wu, const float *wv,
                          const float *ww, const float *phi, double tol, int maxIter,
                          int nthreads, float *pressure, int *status, double *err) {
    return pressureSolveImpl(ni, nj, nk, dx, dt, u, v, w, wu, wv, ww, phi, tol, maxIter, nthreads, pressure,
                             status, err, nullptr, nullptr, nullptr, nullptr);
}

void oracle_curvature_grid(int ni, int nj, int nk, double dx, const float *liquid_phi, float *surface_phi,
                           float *curvature) {
    std::vector<float> surf, kg;
    curvatureGrid(ni, nj, nk, dx, liquid_phi, surf, kg);
    std::copy(surf.begin(), surf.end(), surface_phi);
    std::copy(kg.begin(), kg.end(), curvature);
}
void oracle_tension_cluster_status(int ni, int nj, int nk, const float *liquid_phi, char *status) {
    std::vector<char> st;
    tensionClusterStatus(ni, nj, nk, liquid_phi, st);
    std::copy(st.begin(), st.end(), status);
}
void oracle_cell_weights(int ni, int nj, int nk, const float *solid_phi_nodes, float *wc) {
    std::vector<float> out;
    cellWeights(ni, nj, nk, solid_phi_nodes, out);
    std::copy(out.begin(), out.end(), wc);
}

void oracle_condition_solid_velocity(int ni, int nj, int nk, const float *phi, const float *wu, const float *wv,
                                     const float *ww, float *us, float *vs, float *ws) {
    conditionSolidVelocity(ni, nj, nk, phi, wu, wv, ww, us, vs, ws);
}

namespace {
/* PressureSolver::applySolutionToVelocityField (pressuresolver.cpp:94-112) and
 * _applyPressureToVelocityFieldThread (1047-1205), density 1; with the surface
 * tension term of the face's cell pair when tension is given. */
void applyPressure(int ni, int nj, int nk, double dx, double dt, const float *p, const float *phi,
                   const float *wu, const float *wv, const float *ww, float *u, float *v, float *w,
                   unsigned char *vu, unsigned char *vv, unsigned char *vw, const Tension *T) {
    float factor = dt / dx;
    float eps = 1e-6;
    const double maxtheta = 25;
    auto fluid = [&](int i, int j, int k) { return phi[flat(i, j, k, ni, nj)] < 0.0; };
    auto P = [&](int i, int j, int k) { return p[flat(i, j, k, ni, nj)]; };
    auto PH = [&](int i, int j, int k) { return phi[flat(i, j, k, ni, nj)]; };
    auto clampT = [&](float t) { return (float)std::max(-maxtheta, std::min((double)t, maxtheta)); };
    // one face between cells a (low side) and b (high side); "bordering fluid"
    // is isFaceBorderingMaterialU/V/W (fluidmaterialgrid.cpp:126-146)
    auto face = [&](float wgt, bool bordering, int ai, int aj, int ak, int bi, int bj, int bk,
                    float &val, unsigned char &valid) {
        if (wgt > 0 && bordering) {
            float p1 = 0.0f, p2 = 0.0f;
            if (fluid(ai, aj, ak) && fluid(bi, bj, bk)) {
                p1 = P(ai, aj, ak);
                p2 = P(bi, bj, bk);
            } else {
                float phi1 = PH(ai, aj, ak), phi2 = PH(bi, bj, bk);
                const double tension = T ? T->term(ai, aj, ak, bi, bj, bk) : 0.0;
                if (fluid(ai, aj, ak)) {
                    float thetaP = clampT(phi2 / (phi1 + eps));
                    float thetaT = clampT((phi1 - phi2) / (phi1 + eps));
                    p1 = P(ai, aj, ak);
                    p2 = thetaT * tension + thetaP * p1;
                } else {
                    float thetaP = clampT(phi1 / (phi2 + eps));
                    float thetaT = clampT((phi2 - phi1) / (phi2 + eps));
                    p2 = P(bi, bj, bk);
                    p1 = thetaT * tension + thetaP * p2;
                }
            }
            float rho = (1.0f + 1.0f) / 2.0f;
            val += (float)(-factor * (p2 - p1) / rho);
            valid = 1;
        } else {
            val = 0.0;
        }
    };
    size_t nu = (size_t)(ni + 1) * nj * nk, nv = (size_t)ni * (nj + 1) * nk, nw = (size_t)ni * nj * (nk + 1);
    std::fill(vu, vu + nu, 0);
    std::fill(vv, vv + nv, 0);
    std::fill(vw, vw + nw, 0);
    for (int k = 0; k < nk; k++)
        for (int j = 0; j < nj; j++)
            for (int i = 0; i < ni + 1; i++) {
                if (i == 0 || i == ni - 1) continue;
                size_t f = flat(i, j, k, ni + 1, nj);
                // isFaceBorderingMaterialU (fluidmaterialgrid.cpp:126-130)
                bool lo = i > 0 && fluid(i - 1, j, k);
                bool hi = i < ni && fluid(i, j, k);
                face(wu[f], lo || hi, i - 1, j, k, i, j, k, u[f], vu[f]);
            }
    for (int k = 0; k < nk; k++)
        for (int j = 0; j < nj + 1; j++)
            for (int i = 0; i < ni; i++) {
                if (j == 0 || j == nj - 1) continue;
                size_t f = flat(i, j, k, ni, nj + 1);
                bool lo = j > 0 && fluid(i, j - 1, k);
                bool hi = j < nj && fluid(i, j, k);
                face(wv[f], lo || hi, i, j - 1, k, i, j, k, v[f], vv[f]);
            }
    for (int k = 0; k < nk + 1; k++)
        for (int j = 0; j < nj; j++)
            for (int i = 0; i < ni; i++) {
                if (k == 0 || k == nk - 1) continue;
                size_t f = flat(i, j, k, ni, nj);
                bool lo = k > 0 && fluid(i, j, k - 1);
                bool hi = k < nk && fluid(i, j, k);
                face(ww[f], lo || hi, i, j, k - 1, i, j, k, w[f], vw[f]);
            }
}
}  // namespace

void oracle_apply_pressure(int ni, int nj, int nk, double dx, double dt, const float *p,
                           const float *phi, const float *wu, const float *wv, const float *ww,
                           float *u, float *v, float *w, unsigned char *vu, unsigned char *vv,
                           unsigned char *vw) {
    applyPressure(ni, nj, nk, dx, dt, p, phi, wu, wv, ww, u, v, w, vu, vv, vw, nullptr);
}

int oracle_pressure_solve_tension(int ni, int nj, int nk, double dx, double dt, float *u, float *v, float *w,
                                  const float *wu, const float *wv, const float *ww, const float *liquid_phi,
                                  double k, double tol, int max_iter, float *pressure, int *status, double *err,
                                  unsigned char *vu, unsigned char *vv, unsigned char *vw) {
    Tension T;
    const Tension *tp = nullptr;
    if (k > 1e-6) {
        buildTension(ni, nj, nk, dx, k, liquid_phi, T);
        tp = &T;
    }
    int m = pressureSolveImpl(ni, nj, nk, dx, dt, u, v, w, wu, wv, ww, liquid_phi, tol, max_iter, 1, pressure, status,
                              err, nullptr, nullptr, nullptr, nullptr, tp);
    if (status[1])
        applyPressure(ni, nj, nk, dx, dt, pressure, liquid_phi, wu, wv, ww, u, v, w, vu, vv, vw, tp);
    return m;
}


namespace {

struct Mac {
    int ni, nj, nk;
    double dx;
    const float *u, *v, *w;
};

bool inGridD(double x, double y, double z, double dx, int ni, int nj, int nk) {
    return x >= 0 && y >= 0 && z >= 0 && x < dx * ni && y < dx * nj && z < dx * nk;
}

// Interpolation::trilinearInterpolate(double p[8], x, y, z) (interpolation.cpp:61-70)
double trilin8(const double *p, double x, double y, double z) {
    return p[0] * (1 - x) * (1 - y) * (1 - z) + p[1] * x * (1 - y) * (1 - z) +
           p[2] * (1 - x) * y * (1 - z) + p[3] * (1 - x) * (1 - y) * z + p[4] * x * (1 - y) * z +
           p[5] * (1 - x) * y * z + p[6] * x * y * (1 - z) + p[7] * x * y * z;
}

// MACVelocityField::_interpolateLinearU/V/W (macvelocityfield.cpp:519-613) after the
// in-grid test on the unshifted point; (x, y, z) arrive with the two transverse
// coordinates already shifted by -0.5*dx.
double interpFace(const float *a, int wi, int wj, int wk, double dx, double x, double y, double z) {
    int i, j, k;
    {
        double invdx = 1.0 / dx;
        i = (int)std::floor(x * invdx);
        j = (int)std::floor(y * invdx);
        k = (int)std::floor(z * invdx);
    }
    double gx = (double)i * dx, gy = (double)j * dx, gz = (double)k * dx;
    double inv_dx = 1 / dx;
    double ix = (x - gx) * inv_dx, iy = (y - gy) * inv_dx, iz = (z - gz) * inv_dx;
    double pts[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    const int off[8][3] = {{0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1},
                           {1, 0, 1}, {0, 1, 1}, {1, 1, 0}, {1, 1, 1}};
    for (int q = 0; q < 8; q++) {
        int a0 = i + off[q][0], b0 = j + off[q][1], c0 = k + off[q][2];
        if (inRange(a0, b0, c0, wi, wj, wk)) pts[q] = a[flat(a0, b0, c0, wi, wj)];
    }
    return trilin8(pts, ix, iy, iz);
}

// MACVelocityField::evaluateVelocityAtPositionLinear (macvelocityfield.cpp:631-645)
V3 sampleMac(const Mac &M, V3 p) {
    double x = p.x, y = p.y, z = p.z;
    if (!inGridD(x, y, z, M.dx, M.ni, M.nj, M.nk)) return V3{0, 0, 0};
    double h = 0.5 * M.dx;
    double xu = interpFace(M.u, M.ni + 1, M.nj, M.nk, M.dx, x, y - h, z - h);
    double yv = interpFace(M.v, M.ni, M.nj + 1, M.nk, M.dx, x - h, y, z - h);
    double zw = interpFace(M.w, M.ni, M.nj, M.nk + 1, M.dx, x - h, y - h, z);
    return mk(xu, yv, zw);
}

}  // namespace

void oracle_sample_velocity(int ni, int nj, int nk, double dx, const float *u, const float *v,
                            const float *w, int np, const float *pos, float *out) {
    Mac M{ni, nj, nk, dx, u, v, w};
    for (int p = 0; p < np; p++) {
        V3 r = sampleMac(M, V3{pos[3 * p], pos[3 * p + 1], pos[3 * p + 2]});
        out[3 * p] = r.x;
        out[3 * p + 1] = r.y;
        out[3 * p + 2] = r.z;
    }
}

/* Element actually read at position i of one marker-particle load batch.
 * The batch is a FragmentedVector<MarkerParticle> (fluidsimulation.h:1628-1630)
 * of 5e6-byte fragments, E = 5e6 / 24 = 208333 elements each
 * (fragmentedvector.h:294-301, 320).  operator[] picks the fragment as
 * (int)(i * (1.0 / E)) in double (fragmentedvector.h:140-151), which rounds
 * below c at some exact multiples i = c*E and then reads element (c-1)*E. */
size_t oracle_load_queue_index(size_t i) {
    const int E = (int)(5000000u / 24u);
    const double inv = 1.0 / (double)E;
    int node = (int)((int)i * inv);
    int item = (int)i % E;
    return (size_t)node * E + item;
}

/* ====================================================== whole engine ===== */
namespace {

// Interpolation::trilinearInterpolate(vec3, dx, Array3d<float>) (interpolation.cpp:72-112)
// on the nodal solid SDF, returned through MeshLevelSet::trilinearInterpolate as float
// (meshlevelset.cpp:261-263).
float nodalTrilinear(const float *g, int pi, int pj, int pk, double dx, V3 p) {
    return (float)nodalTrilinearD(g, pi, pj, pk, dx, p);
}
double nodalTrilinearD(const float *g, int pi, int pj, int pk, double dx, V3 p) {
    GI c = posToIndex(p, dx);
    V3 gp = indexToPos(c.i, c.j, c.k, dx);
    double inv_dx = 1.0 / dx;
    double ix = (p.x - gp.x) * inv_dx, iy = (p.y - gp.y) * inv_dx, iz = (p.z - gp.z) * inv_dx;
    double pts[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    const int off[8][3] = {{0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1},
                           {1, 0, 1}, {0, 1, 1}, {1, 1, 0}, {1, 1, 1}};
    for (int q = 0; q < 8; q++) {
        int a = c.i + off[q][0], b = c.j + off[q][1], e = c.k + off[q][2];
        if (inRange(a, b, e, pi, pj, pk)) pts[q] = g[flat(a, b, e, pi, pj)];
    }
    return trilin8(pts, ix, iy, iz);
}

// Interpolation::bilinearInterpolate + trilinearInterpolateGradient (interpolation.cpp:189-260)
double bilin(double v00, double v10, double v01, double v11, double ix, double iy) {
    double l1 = (1 - ix) * v00 + ix * v10;
    double l2 = (1 - ix) * v01 + ix * v11;
    return (1 - iy) * l1 + iy * l2;
}
V3 nodalGradient(const float *g, int pi, int pj, int pk, double dx, V3 p) {
    GI c = posToIndex(p, dx);
    V3 gp = indexToPos(c.i, c.j, c.k, dx);
    double inv_dx = 1.0 / dx;
    double ix = (p.x - gp.x) * inv_dx, iy = (p.y - gp.y) * inv_dx, iz = (p.z - gp.z) * inv_dx;
    auto at = [&](int a, int b, int e) -> float {
        return inRange(a, b, e, pi, pj, pk) ? g[flat(a, b, e, pi, pj)] : 0.0f;
    };
    float v000 = at(c.i, c.j, c.k), v100 = at(c.i + 1, c.j, c.k), v010 = at(c.i, c.j + 1, c.k);
    float v001 = at(c.i, c.j, c.k + 1), v101 = at(c.i + 1, c.j, c.k + 1);
    float v011 = at(c.i, c.j + 1, c.k + 1), v110 = at(c.i + 1, c.j + 1, c.k);
    float v111 = at(c.i + 1, c.j + 1, c.k + 1);
    float ddx00 = v100 - v000, ddx10 = v110 - v010, ddx01 = v101 - v001, ddx11 = v111 - v011;
    float dvdx = bilin(ddx00, ddx10, ddx01, ddx11, iy, iz);
    float ddy00 = v010 - v000, ddy10 = v110 - v100, ddy01 = v011 - v001, ddy11 = v111 - v101;
    float dvdy = bilin(ddy00, ddy10, ddy01, ddy11, ix, iz);
    float ddz00 = v001 - v000, ddz10 = v101 - v100, ddz01 = v011 - v010, ddz11 = v111 - v110;
    float dvdz = bilin(ddz00, ddz10, ddz01, ddz11, ix, iy);
    return V3{dvdx, dvdy, dvdz};
}


// ---- the viscosity solve (ViscositySolver, viscositysolver.cpp) ----
//
// One SPD system over the u, v and w faces together (viscositysolver.cpp:
// 50-79): face states from the solid SDF at the cell centres, control volumes
// from the liquid SDF sampled on a twice finer grid, a matrix index over the
// fluid faces with volume, 15-entry rows of the viscous operator, and the
// reference's PCGSolver<float> with its general-sparse MIC(0) factor.

enum ViscFace : unsigned char { VF_AIR = 0, VF_FLUID = 1, VF_SOLID = 2 };

// SparseMatrix<float> rows (sparsematrix.h:44-112): sorted column lists;
// set/add ignore a -1 index
struct SparseRows {
    std::vector<std::vector<int>> idx;
    std::vector<std::vector<float>> val;
    explicit SparseRows(int n = 0) : idx(n), val(n) {}
    void put(int i, int j, float x, bool add) {
        if (i == -1 || j == -1) return;
        std::vector<int> &I = idx[i];
        std::vector<float> &X = val[i];
        for (size_t k = 0; k < I.size(); k++) {
            if (I[k] == j) {
                if (add)
                    X[k] += x;
                else
                    X[k] = x;
                return;
            }
            if (I[k] > j) {
                I.insert(I.begin() + k, j);
                X.insert(X.begin() + k, x);
                return;
            }
        }
        I.push_back(j);
        X.push_back(x);
    }
};

struct ViscSystem {
    int ni = 0, nj = 0, nk = 0;
    std::vector<unsigned char> su, sv, sw;             // face states
    std::vector<float> center, vu, vv, vw, eu, ev, ew;  // control volumes
    std::vector<int> table;                            // FaceIndexer index -> row, or -1
    int m = 0;
    SparseRows A;
    std::vector<float> rhs;
    size_t U(int i, int j, int k) const { return flat(i, j, k, ni + 1, nj); }
    size_t V(int i, int j, int k) const { return flat(i, j, k, ni, nj + 1); }
    size_t W(int i, int j, int k) const { return flat(i, j, k, ni, nj); }
    // FaceIndexer (viscositysolver.h:130-155): u faces, then v, then w
    int rowU(int i, int j, int k) const { return table[U(i, j, k)]; }
    int rowV(int i, int j, int k) const { return table[(size_t)(ni + 1) * nj * nk + V(i, j, k)]; }
    int rowW(int i, int j, int k) const {
        return table[(size_t)(ni + 1) * nj * nk + (size_t)ni * (nj + 1) * nk + W(i, j, k)];
    }
};

// _computeFaceStateGrid (viscositysolver.cpp:102-181) and the control volumes
// (_computeVolumeGrid, 209-377); liquid (cells) and solid (nodes) SDFs
void viscosityVolumes(int ni, int nj, int nk, double dx, const float *liquid, const float *solid, ViscSystem &S) {
    S.ni = ni; S.nj = nj; S.nk = nk;
    // the solid SDF at the cell centres: the mean of the 8 nodes in the
    // reference's order (MeshLevelSet::getDistanceAtCellCenter, meshlevelset.cpp:150-160)
    std::vector<float> sc((size_t)ni * nj * nk);
    const int pi = ni + 1, pj = nj + 1;
    for (int k = 0; k < nk; k++)
        for (int j = 0; j < nj; j++)
            for (int i = 0; i < ni; i++) {
                auto P = [&](int a, int b, int c) { return solid[flat(i + a, j + b, k + c, pi, pj)]; };
                sc[flat(i, j, k, ni, nj)] = 0.125f * (P(0, 0, 0) + P(1, 0, 0) + P(0, 1, 0) + P(1, 1, 0) +
                                                      P(0, 0, 1) + P(1, 0, 1) + P(0, 1, 1) + P(1, 1, 1));
            }
    auto C = [&](int i, int j, int k) { return sc[flat(i, j, k, ni, nj)]; };
    S.su.assign((size_t)(ni + 1) * nj * nk, VF_AIR);
    S.sv.assign((size_t)ni * (nj + 1) * nk, VF_AIR);
    S.sw.assign((size_t)ni * nj * (nk + 1), VF_AIR);
    for (int k = 0; k < nk; k++)
        for (int j = 0; j < nj; j++)
            for (int i = 0; i < ni + 1; i++) {
                bool edge = i == 0 || i == ni;
                S.su[S.U(i, j, k)] = (edge || C(i - 1, j, k) + C(i, j, k) <= 0) ? VF_SOLID : VF_FLUID;
            }
    for (int k = 0; k < nk; k++)
        for (int j = 0; j < nj + 1; j++)
            for (int i = 0; i < ni; i++) {
                bool edge = j == 0 || j == nj;
                S.sv[S.V(i, j, k)] = (edge || C(i, j - 1, k) + C(i, j, k) <= 0) ? VF_SOLID : VF_FLUID;
            }
    for (int k = 0; k < nk + 1; k++)
        for (int j = 0; j < nj; j++)
            for (int i = 0; i < ni; i++) {
                bool edge = k == 0 || k == nk;
                S.sw[S.W(i, j, k)] = (edge || C(i, j, k - 1) + C(i, j, k) <= 0) ? VF_SOLID : VF_FLUID;
            }
    // cells near the liquid: liquid cells and two layers of 6-neighbours, on
    // an (n+1)^3 grid
    const int vi = ni + 1, vj = nj + 1, vk = nk + 1;
    std::vector<char> valid((size_t)vi * vj * vk, 0);
    for (int k = 0; k < nk; k++)
        for (int j = 0; j < nj; j++)
            for (int i = 0; i < ni; i++)
                if (liquid[flat(i, j, k, ni, nj)] < 0) valid[flat(i, j, k, vi, vj)] = 1;
    for (int layer = 0; layer < 2; layer++) {
        std::vector<char> next = valid;
        const int nb[6][3] = {{-1, 0, 0}, {1, 0, 0}, {0, -1, 0}, {0, 1, 0}, {0, 0, -1}, {0, 0, 1}};
        for (int k = 0; k < vk; k++)
            for (int j = 0; j < vj; j++)
                for (int i = 0; i < vi; i++) {
                    if (!valid[flat(i, j, k, vi, vj)]) continue;
                    for (auto &o : nb) {
                        int a = i + o[0], b = j + o[1], c = k + o[2];
                        if (inRange(a, b, c, vi, vj, vk)) next[flat(a, b, c, vi, vj)] = 1;
                    }
                }
        valid.swap(next);
    }
    // volume fractions of the liquid in the (2n)^3 subcells of the valid cells
    // (_estimateVolumeFractionsThread, 314-348): the liquid SDF trilinearly
    // interpolated at each subcell's corners (ParticleLevelSet::trilinearInterpolate,
    // particlelevelset.cpp:136-138: cell-centred samples)
    const int si = 2 * ni, sj = 2 * nj, sk = 2 * nk;
    std::vector<float> sub((size_t)si * sj * sk, 0.0f);
    const float fdx = (float)dx;  // the solver's float cell width
    const float sdx = 0.5f * fdx;
    const V3 start = V3{0.25f * fdx, 0.25f * fdx, 0.25f * fdx};
    const V3 half = mk(0.5 * dx, 0.5 * dx, 0.5 * dx);
    auto liquidAt = [&](V3 p) { return nodalTrilinear(liquid, ni, nj, nk, dx, p - half); };
#pragma omp parallel for num_threads(g_threads) schedule(static)
    for (int k = 0; k < sk; k++)
        for (int j = 0; j < sj; j++)
            for (int i = 0; i < si; i++) {
                if (!valid[flat(i / 2, j / 2, k / 2, vi, vj)]) continue;
                V3 c = start + V3{i * sdx, j * sdx, k * sdx};
                float h = 0.5f * sdx;
                float p000 = liquidAt(c + V3{-h, -h, -h}), p001 = liquidAt(c + V3{-h, -h, +h});
                float p010 = liquidAt(c + V3{-h, +h, -h}), p011 = liquidAt(c + V3{-h, +h, +h});
                float p100 = liquidAt(c + V3{+h, -h, -h}), p101 = liquidAt(c + V3{+h, -h, +h});
                float p110 = liquidAt(c + V3{+h, +h, -h}), p111 = liquidAt(c + V3{+h, +h, +h});
                sub[flat(i, j, k, si, sj)] = volumeFraction8(p000, p100, p010, p110, p001, p101, p011, p111);
            }
    // each grid's volume: the mean of the 8 subcells around its sample point
    // (_computeVolumeGridThread, 350-377), over cells 1..n-1
    auto volume = [&](std::vector<float> &out, int wi, int wj, int wk, int oi, int oj, int ok) {
        out.assign((size_t)wi * wj * wk, 0.0f);
        for (int k = 1; k < nk; k++)
            for (int j = 1; j < nj; j++)
                for (int i = 1; i < ni; i++) {
                    if (!valid[flat(i, j, k, vi, vj)]) continue;
                    float acc = 0.0f;
                    for (int c = 0; c < 2; c++)
                        for (int b = 0; b < 2; b++)
                            for (int a = 0; a < 2; a++)
                                acc += sub[flat(2 * i + oi + a, 2 * j + oj + b, 2 * k + ok + c, si, sj)];
                    out[flat(i, j, k, wi, wj)] = 0.125f * acc;
                }
    };
    volume(S.center, ni, nj, nk, 0, 0, 0);
    volume(S.vu, ni + 1, nj, nk, -1, 0, 0);
    volume(S.vv, ni, nj + 1, nk, 0, -1, 0);
    volume(S.vw, ni, nj, nk + 1, 0, 0, -1);
    volume(S.eu, ni, nj + 1, nk + 1, 0, -1, -1);
    volume(S.ev, ni + 1, nj, nk + 1, -1, 0, -1);
    volume(S.ew, ni + 1, nj + 1, nk, -1, -1, 0);
}

// _computeMatrixIndexTable (viscositysolver.cpp:383-473) and the three row
// assemblies (_initializeLinearSystemThreadU/V/W, 559-830); visc: the
// (n+1)^3 viscosity grid, uvw the velocity field before the solve
void viscositySystem(double dx, double dt, const float *visc, const float *u, const float *v, const float *w,
                     ViscSystem &S) {
    const int ni = S.ni, nj = S.nj, nk = S.nk;
    const size_t nU = (size_t)(ni + 1) * nj * nk, nV = (size_t)ni * (nj + 1) * nk, nW = (size_t)ni * nj * (nk + 1);
    auto CE = [&](int i, int j, int k) { return S.center[flat(i, j, k, ni, nj)]; };
    auto VU = [&](int i, int j, int k) { return S.vu[flat(i, j, k, ni + 1, nj)]; };
    auto VV = [&](int i, int j, int k) { return S.vv[flat(i, j, k, ni, nj + 1)]; };
    auto VW = [&](int i, int j, int k) { return S.vw[flat(i, j, k, ni, nj)]; };
    auto EU = [&](int i, int j, int k) { return S.eu[flat(i, j, k, ni, nj + 1)]; };
    auto EV = [&](int i, int j, int k) { return S.ev[flat(i, j, k, ni + 1, nj)]; };
    auto EW = [&](int i, int j, int k) { return S.ew[flat(i, j, k, ni + 1, nj + 1)]; };
    auto SU = [&](int i, int j, int k) { return S.su[S.U(i, j, k)]; };
    auto SV = [&](int i, int j, int k) { return S.sv[S.V(i, j, k)]; };
    auto SW = [&](int i, int j, int k) { return S.sw[S.W(i, j, k)]; };
    // A row's neighbour can lie outside its face grid (U(i, nj, k),
    // U(i, j, nk), V(ni, j, k), V(i, j, nk), W(ni, j, k), W(i, nj, k)).  The
    // reference then reads the state and matrix index at the wrapped flat
    // index, which is a border face (never a row) or lies past the array, and
    // the velocity as 0 (MACVelocityField::U, macvelocityfield.cpp:112-118):
    // the neighbour adds nothing.  Here it is air.
    auto stU = [&](int i, int j, int k) -> unsigned char {
        return inRange(i, j, k, ni + 1, nj, nk) ? SU(i, j, k) : (unsigned char)VF_AIR;
    };
    auto stV = [&](int i, int j, int k) -> unsigned char {
        return inRange(i, j, k, ni, nj + 1, nk) ? SV(i, j, k) : (unsigned char)VF_AIR;
    };
    auto stW = [&](int i, int j, int k) -> unsigned char {
        return inRange(i, j, k, ni, nj, nk + 1) ? SW(i, j, k) : (unsigned char)VF_AIR;
    };
    std::vector<char> in(nU + nV + nW, 0);
    for (int k = 1; k < nk; k++)
        for (int j = 1; j < nj; j++)
            for (int i = 1; i < ni; i++) {
                if (SU(i, j, k) == VF_FLUID &&
                    (VU(i, j, k) > 0.0 || CE(i, j, k) > 0.0 || CE(i - 1, j, k) > 0.0 || EW(i, j + 1, k) > 0.0 ||
                     EW(i, j, k) > 0.0 || EV(i, j, k + 1) > 0.0 || EV(i, j, k) > 0.0))
                    in[S.U(i, j, k)] = 1;
            }
    for (int k = 1; k < nk; k++)
        for (int j = 1; j < nj; j++)
            for (int i = 1; i < ni; i++) {
                if (SV(i, j, k) == VF_FLUID &&
                    (VV(i, j, k) > 0.0 || EW(i + 1, j, k) > 0.0 || EW(i, j, k) > 0.0 || CE(i, j, k) > 0.0 ||
                     CE(i, j - 1, k) > 0.0 || EU(i, j, k + 1) > 0.0 || EU(i, j, k) > 0.0))
                    in[nU + S.V(i, j, k)] = 1;
            }
    for (int k = 1; k < nk; k++)
        for (int j = 1; j < nj; j++)
            for (int i = 1; i < ni; i++) {
                if (SW(i, j, k) == VF_FLUID &&
                    (VW(i, j, k) > 0.0 || EV(i + 1, j, k) > 0.0 || EV(i, j, k) > 0.0 || EU(i, j + 1, k) > 0.0 ||
                     EU(i, j, k) > 0.0 || CE(i, j, k) > 0.0 || CE(i, j, k - 1) > 0.0))
                    in[nU + nV + S.W(i, j, k)] = 1;
            }
    S.table.assign(nU + nV + nW, -1);
    S.m = 0;
    for (size_t q = 0; q < in.size(); q++)
        if (in[q]) S.table[q] = S.m++;
    S.A = SparseRows(S.m);
    S.rhs.assign(S.m, 0.0f);
    const int vi = ni + 1, vj = nj + 1;
    auto VS = [&](int i, int j, int k) { return visc[flat(i, j, k, vi, vj)]; };
    auto Uv = [&](int i, int j, int k) { return u[flat(i, j, k, ni + 1, nj)]; };
    auto Vv = [&](int i, int j, int k) { return v[flat(i, j, k, ni, nj + 1)]; };
    auto Wv = [&](int i, int j, int k) { return w[flat(i, j, k, ni, nj)]; };
    // the solver keeps the cell width and time step as float (viscositysolver.h:36-39)
    const float fdx = (float)dx, fdt = (float)dt;
    const float invdx = 1.0f / fdx;
    const float factor = fdt * invdx * invdx;
    auto rU = [&](int i, int j, int k) { return inRange(i, j, k, ni + 1, nj, nk) ? S.rowU(i, j, k) : -1; };
    auto rV = [&](int i, int j, int k) { return inRange(i, j, k, ni, nj + 1, nk) ? S.rowV(i, j, k) : -1; };
    auto rW = [&](int i, int j, int k) { return inRange(i, j, k, ni, nj, nk + 1) ? S.rowW(i, j, k) : -1; };
    auto Uo = [&](int i, int j, int k) { return inRange(i, j, k, ni + 1, nj, nk) ? Uv(i, j, k) : 0.0f; };
    auto Vo = [&](int i, int j, int k) { return inRange(i, j, k, ni, nj + 1, nk) ? Vv(i, j, k) : 0.0f; };
    auto Wo = [&](int i, int j, int k) { return inRange(i, j, k, ni, nj, nk + 1) ? Wv(i, j, k) : 0.0f; };
    // the u rows (559-648); the v and w rows are the same operator with the
    // axes cycled (650-830)
    for (int k = 1; k < nk; k++)
        for (int j = 1; j < nj; j++)
            for (int i = 1; i < ni; i++) {
                if (SU(i, j, k) != VF_FLUID) continue;
                const int row = rU(i, j, k);
                if (row == -1) continue;
                float vR = VS(i, j, k), vL = VS(i - 1, j, k);
                float vT = 0.25f * (VS(i - 1, j + 1, k) + VS(i - 1, j, k) + VS(i, j + 1, k) + VS(i, j, k));
                float vB = 0.25f * (VS(i - 1, j, k) + VS(i - 1, j - 1, k) + VS(i, j, k) + VS(i, j - 1, k));
                float vF = 0.25f * (VS(i - 1, j, k + 1) + VS(i - 1, j, k) + VS(i, j, k + 1) + VS(i, j, k));
                float vK = 0.25f * (VS(i - 1, j, k) + VS(i - 1, j, k - 1) + VS(i, j, k) + VS(i, j, k - 1));
                float fR = 2 * factor * vR * CE(i, j, k), fL = 2 * factor * vL * CE(i - 1, j, k);
                float fT = factor * vT * EW(i, j + 1, k), fB = factor * vB * EW(i, j, k);
                float fF = factor * vF * EV(i, j, k + 1), fK = factor * vK * EV(i, j, k);
                float diag = VU(i, j, k) + fR + fL + fT + fB + fF + fK;
                S.A.put(row, row, diag, false);
                struct E { unsigned char st; int col; float coef; float vel; };
                const E es[14] = {
                    {stU(i + 1, j, k), rU(i + 1, j, k), -fR, Uo(i + 1, j, k)},
                    {stU(i - 1, j, k), rU(i - 1, j, k), -fL, Uo(i - 1, j, k)},
                    {stU(i, j + 1, k), rU(i, j + 1, k), -fT, Uo(i, j + 1, k)},
                    {stU(i, j - 1, k), rU(i, j - 1, k), -fB, Uo(i, j - 1, k)},
                    {stU(i, j, k + 1), rU(i, j, k + 1), -fF, Uo(i, j, k + 1)},
                    {stU(i, j, k - 1), rU(i, j, k - 1), -fK, Uo(i, j, k - 1)},
                    {stV(i, j + 1, k), rV(i, j + 1, k), -fT, Vo(i, j + 1, k)},
                    {stV(i - 1, j + 1, k), rV(i - 1, j + 1, k), fT, Vo(i - 1, j + 1, k)},
                    {stV(i, j, k), rV(i, j, k), fB, Vo(i, j, k)},
                    {stV(i - 1, j, k), rV(i - 1, j, k), -fB, Vo(i - 1, j, k)},
                    {stW(i, j, k + 1), rW(i, j, k + 1), -fF, Wo(i, j, k + 1)},
                    {stW(i - 1, j, k + 1), rW(i - 1, j, k + 1), fF, Wo(i - 1, j, k + 1)},
                    {stW(i, j, k), rW(i, j, k), fK, Wo(i, j, k)},
                    {stW(i - 1, j, k), rW(i - 1, j, k), -fK, Wo(i - 1, j, k)}};
                for (const E &e : es)
                    if (e.st == VF_FLUID) S.A.put(row, e.col, e.coef, true);
                float rval = VU(i, j, k) * Uo(i, j, k);
                for (const E &e : es)
                    if (e.st == VF_SOLID) rval -= e.coef * e.vel;
                S.rhs[row] = rval;
            }
    for (int k = 1; k < nk; k++)
        for (int j = 1; j < nj; j++)
            for (int i = 1; i < ni; i++) {
                if (SV(i, j, k) != VF_FLUID) continue;
                const int row = rV(i, j, k);
                if (row == -1) continue;
                float vT = VS(i, j, k), vB = VS(i, j - 1, k);
                float vR = 0.25f * (VS(i, j - 1, k) + VS(i + 1, j - 1, k) + VS(i, j, k) + VS(i + 1, j, k));
                float vL = 0.25f * (VS(i, j - 1, k) + VS(i - 1, j - 1, k) + VS(i, j, k) + VS(i - 1, j, k));
                float vF = 0.25f * (VS(i, j - 1, k) + VS(i, j - 1, k + 1) + VS(i, j, k) + VS(i, j, k + 1));
                float vK = 0.25f * (VS(i, j - 1, k) + VS(i, j - 1, k - 1) + VS(i, j, k) + VS(i, j, k - 1));
                float fR = factor * vR * EW(i + 1, j, k), fL = factor * vL * EW(i, j, k);
                float fT = 2 * factor * vT * CE(i, j, k), fB = 2 * factor * vB * CE(i, j - 1, k);
                float fF = factor * vF * EU(i, j, k + 1), fK = factor * vK * EU(i, j, k);
                float diag = VV(i, j, k) + fR + fL + fT + fB + fF + fK;
                S.A.put(row, row, diag, false);
                struct E { unsigned char st; int col; float coef; float vel; };
                const E es[14] = {
                    {stV(i + 1, j, k), rV(i + 1, j, k), -fR, Vo(i + 1, j, k)},
                    {stV(i - 1, j, k), rV(i - 1, j, k), -fL, Vo(i - 1, j, k)},
                    {stV(i, j + 1, k), rV(i, j + 1, k), -fT, Vo(i, j + 1, k)},
                    {stV(i, j - 1, k), rV(i, j - 1, k), -fB, Vo(i, j - 1, k)},
                    {stV(i, j, k + 1), rV(i, j, k + 1), -fF, Vo(i, j, k + 1)},
                    {stV(i, j, k - 1), rV(i, j, k - 1), -fK, Vo(i, j, k - 1)},
                    {stU(i + 1, j, k), rU(i + 1, j, k), -fR, Uo(i + 1, j, k)},
                    {stU(i + 1, j - 1, k), rU(i + 1, j - 1, k), fR, Uo(i + 1, j - 1, k)},
                    {stU(i, j, k), rU(i, j, k), fL, Uo(i, j, k)},
                    {stU(i, j - 1, k), rU(i, j - 1, k), -fL, Uo(i, j - 1, k)},
                    {stW(i, j, k + 1), rW(i, j, k + 1), -fF, Wo(i, j, k + 1)},
                    {stW(i, j - 1, k + 1), rW(i, j - 1, k + 1), fF, Wo(i, j - 1, k + 1)},
                    {stW(i, j, k), rW(i, j, k), fK, Wo(i, j, k)},
                    {stW(i, j - 1, k), rW(i, j - 1, k), -fK, Wo(i, j - 1, k)}};
                for (const E &e : es)
                    if (e.st == VF_FLUID) S.A.put(row, e.col, e.coef, true);
                float rval = VV(i, j, k) * Vo(i, j, k);
                for (const E &e : es)
                    if (e.st == VF_SOLID) rval -= e.coef * e.vel;
                S.rhs[row] = rval;
            }
    for (int k = 1; k < nk; k++)
        for (int j = 1; j < nj; j++)
            for (int i = 1; i < ni; i++) {
                if (SW(i, j, k) != VF_FLUID) continue;
                const int row = rW(i, j, k);
                if (row == -1) continue;
                float vF = VS(i, j, k), vK = VS(i, j, k - 1);
                float vR = 0.25f * (VS(i, j, k) + VS(i, j, k - 1) + VS(i + 1, j, k) + VS(i + 1, j, k - 1));
                float vL = 0.25f * (VS(i, j, k) + VS(i, j, k - 1) + VS(i - 1, j, k) + VS(i - 1, j, k - 1));
                float vT = 0.25f * (VS(i, j, k) + VS(i, j, k - 1) + VS(i, j + 1, k) + VS(i, j + 1, k - 1));
                float vB = 0.25f * (VS(i, j, k) + VS(i, j, k - 1) + VS(i, j - 1, k) + VS(i, j - 1, k - 1));
                float fR = factor * vR * EV(i + 1, j, k), fL = factor * vL * EV(i, j, k);
                float fT = factor * vT * EU(i, j + 1, k), fB = factor * vB * EU(i, j, k);
                float fF = 2 * factor * vF * CE(i, j, k), fK = 2 * factor * vK * CE(i, j, k - 1);
                float diag = VW(i, j, k) + fR + fL + fT + fB + fF + fK;
                S.A.put(row, row, diag, false);
                struct E { unsigned char st; int col; float coef; float vel; };
                const E es[14] = {
                    {stW(i + 1, j, k), rW(i + 1, j, k), -fR, Wo(i + 1, j, k)},
                    {stW(i - 1, j, k), rW(i - 1, j, k), -fL, Wo(i - 1, j, k)},
                    {stW(i, j + 1, k), rW(i, j + 1, k), -fT, Wo(i, j + 1, k)},
                    {stW(i, j - 1, k), rW(i, j - 1, k), -fB, Wo(i, j - 1, k)},
                    {stW(i, j, k + 1), rW(i, j, k + 1), -fF, Wo(i, j, k + 1)},
                    {stW(i, j, k - 1), rW(i, j, k - 1), -fK, Wo(i, j, k - 1)},
                    {stU(i + 1, j, k), rU(i + 1, j, k), -fR, Uo(i + 1, j, k)},
                    {stU(i + 1, j, k - 1), rU(i + 1, j, k - 1), fR, Uo(i + 1, j, k - 1)},
                    {stU(i, j, k), rU(i, j, k), fL, Uo(i, j, k)},
                    {stU(i, j, k - 1), rU(i, j, k - 1), -fL, Uo(i, j, k - 1)},
                    {stV(i, j + 1, k), rV(i, j + 1, k), -fT, Vo(i, j + 1, k)},
                    {stV(i, j + 1, k - 1), rV(i, j + 1, k - 1), fT, Vo(i, j + 1, k - 1)},
                    {stV(i, j, k), rV(i, j, k), fB, Vo(i, j, k)},
                    {stV(i, j, k - 1), rV(i, j, k - 1), -fB, Vo(i, j, k - 1)}};
                for (const E &e : es)
                    if (e.st == VF_FLUID) S.A.put(row, e.col, e.coef, true);
                float rval = VW(i, j, k) * Wo(i, j, k);
                for (const E &e : es)
                    if (e.st == VF_SOLID) rval -= e.coef * e.vel;
                S.rhs[row] = rval;
            }
}

// SparseColumnLowerFactor + factorModifiedIncompleteColesky0 (pcgsolver.h:
// 27-190) in float: the lower triangle copied column by column, diagonals
// below 1e-9 taken as 0 (null rows), then the push-form elimination with the
// dropped fill-in moved onto the diagonal (modification 0.97, pivots below
// 0.25 x the original diagonal replaced by it).  The pivot's inverse square
// root is taken in double and rounded once: the template's unqualified sqrt
// of a float resolves to the C library's double sqrt under the reference's
// includes.
struct ViscFactor {
    std::vector<float> invdiag, value, adiag;
    std::vector<unsigned> rowindex, colstart;
};
void viscFactor(const SparseRows &A, ViscFactor &F) {
    const unsigned n = (unsigned)A.idx.size();
    F.invdiag.assign(n, 0.0f);
    F.adiag.assign(n, 0.0f);
    F.colstart.assign(n + 1, 0);
    F.value.clear();
    F.rowindex.clear();
    const float eps = 1e-9f;
    for (unsigned i = 0; i < n; i++) {
        F.colstart[i] = (unsigned)F.rowindex.size();
        for (size_t q = 0; q < A.idx[i].size(); q++) {
            const unsigned c = (unsigned)A.idx[i][q];
            if (c > i) {
                F.rowindex.push_back(c);
                F.value.push_back(A.val[i][q]);
            } else if (c == i) {
                float x = A.val[i][q];
                if (x < eps) x = 0.0f;
                F.invdiag[i] = F.adiag[i] = x;
            }
        }
    }
    F.colstart[n] = (unsigned)F.rowindex.size();
    const float omega = 0.97f, ratio = 0.25f;
    for (unsigned k = 0; k < n; k++) {
        if (F.adiag[k] == 0) continue;
        if (F.invdiag[k] < ratio * F.adiag[k])
            F.invdiag[k] = (float)(1 / std::sqrt((double)F.adiag[k]));
        else
            F.invdiag[k] = (float)(1 / std::sqrt((double)F.invdiag[k]));
        const unsigned c0 = F.colstart[k], c1 = F.colstart[k + 1];
        for (unsigned p = c0; p < c1; p++) F.value[p] *= F.invdiag[k];
        for (unsigned p = c0; p < c1; p++) {
            const unsigned j = F.rowindex[p];
            const float mult = F.value[p];
            float missing = 0;
            unsigned a = c0;
            // entries of column k above row j that column j of A lacks
            size_t b = 0;
            const std::vector<int> &Aj = A.idx[j];
            while (a < c1 && F.rowindex[a] < j) {
                while (b < Aj.size()) {
                    if ((unsigned)Aj[b] < F.rowindex[a]) {
                        b++;
                    } else if ((unsigned)Aj[b] == F.rowindex[a]) {
                        break;
                    } else {
                        missing += F.value[a];
                        break;
                    }
                }
                a++;
            }
            if (a < c1 && F.rowindex[a] == j) F.invdiag[j] -= mult * F.value[a];
            a++;
            // below the diagonal: update matching entries, the rest is missing
            unsigned bb = F.colstart[j];
            while (a < c1 && bb < F.colstart[j + 1]) {
                if (F.rowindex[bb] < F.rowindex[a]) {
                    bb++;
                } else if (F.rowindex[bb] == F.rowindex[a]) {
                    F.value[bb] -= mult * F.value[a];
                    a++;
                    bb++;
                } else {
                    missing += F.value[a];
                    a++;
                }
            }
            while (a < c1) missing += F.value[a++];
            F.invdiag[j] -= omega * mult * missing;
        }
    }
}

// solveLower then solveLowerTransposeInPlace (pcgsolver.h:196-229)
void viscPrecondition(const ViscFactor &F, const std::vector<float> &x, std::vector<float> &out) {
    const unsigned n = (unsigned)F.invdiag.size();
    out = x;
    for (unsigned i = 0; i < n; i++) {
        out[i] *= F.invdiag[i];
        for (unsigned q = F.colstart[i]; q < F.colstart[i + 1]; q++) out[F.rowindex[q]] -= F.value[q] * out[i];
    }
    for (unsigned i = n; i-- > 0;) {
        for (unsigned q = F.colstart[i]; q < F.colstart[i + 1]; q++) out[i] -= F.value[q] * out[F.rowindex[q]];
        out[i] *= F.invdiag[i];
    }
}

// BLAS::dot<float> (blaswrapper.h:27-57): a float sum, per thread over
// 500000-element chunks when more than one thread is allowed
float viscDot(const std::vector<float> &x, const std::vector<float> &y, int ncpu) {
    int nt = (int)std::fmin(ncpu, std::ceil((float)x.size() / (float)500000));
    if (nt <= 1) {
        float s = 0;
        for (size_t i = 0; i < x.size(); i++) s += x[i] * y[i];
        return s;
    }
    std::vector<int> iv = splitRange(0, (int)x.size(), nt);
    float s = 0;
    for (int t = 0; t < nt; t++) {
        float part = 0;
        for (int i = iv[t]; i < iv[t + 1]; i++) part += x[i] * y[i];
        s += part;
    }
    return s;
}
float viscAbsMax(const std::vector<float> &x) {
    float mx = 0;
    for (float v : x)
        if (std::fabs(v) > mx) mx = std::fabs(v);
    return mx;
}
// BLAS::addScaled (blaswrapper.h:128-157): y += alpha x with a float alpha
void viscAddScaled(float alpha, const std::vector<float> &x, std::vector<float> &y) {
    for (size_t i = 0; i < x.size(); i++) y[i] += alpha * x[i];
}

// PCGSolver<float>::solve (pcgsolver.h:246-300) with the viscosity solver's
// parameters (viscositysolver.cpp:832-868): relative tolerance tol on the
// residual's infinity norm, capped at 1.0, and success also when the
// iteration limit is reached below an estimated error of 10.
struct ViscResult {
    int iterations = 0;
    float error = 0.0f;
    bool success = true;
};
ViscResult viscPCG(const SparseRows &A, const std::vector<float> &rhs, double tol, int maxIter, int ncpu,
                   std::vector<float> &x) {
    ViscResult res;
    const size_t n = rhs.size();
    x.assign(n, 0.0f);
    float tolFactor = (float)tol;
    if (tolFactor < 1e-30) tolFactor = 1e-30f;
    const float maxErr = 1.0f;
    std::vector<float> r = rhs, z(n), s(n);
    float resid = viscAbsMax(r);
    if (resid == 0) {
        res.iterations = 0;
        res.error = 0.0f;
        res.success = true;
        return res;
    }
    const double tolAbs = tolFactor * resid;
    ViscFactor F;
    viscFactor(A, F);
    viscPrecondition(F, r, z);
    double rho = viscDot(z, r, ncpu);
    if (rho == 0 || rho != rho) {
        res.iterations = 0;
        res.error = resid;
        res.success = false;
        return res;
    }
    s = z;
    bool ok = false;
    int it;
    for (it = 0; it < maxIter; it++) {
        for (size_t i = 0; i < n; i++) {  // FixedSparseMatrix rows in column order
            float acc = 0;
            for (size_t q = 0; q < A.idx[i].size(); q++) acc += A.val[i][q] * s[A.idx[i][q]];
            z[i] = acc;
        }
        const double alpha = rho / viscDot(s, z, ncpu);
        viscAddScaled((float)alpha, s, x);
        viscAddScaled((float)-alpha, z, r);
        resid = viscAbsMax(r);
        if (resid <= std::min(tolAbs, (double)maxErr)) {
            ok = true;
            it++;
            break;
        }
        viscPrecondition(F, r, z);
        const double rhoNew = viscDot(z, r, ncpu);
        const double beta = rhoNew / rho;
        viscAddScaled((float)beta, s, z);
        s.swap(z);
        rho = rhoNew;
    }
    res.iterations = it;
    res.error = resid;
    res.success = ok || (it == maxIter && resid < 10.0);
    return res;
}

// ViscositySolver::applyViscosityToVelocityField (viscositysolver.cpp:50-79,
// 870-905): on success the velocity field is cleared and the matrix faces take
// the solution; on failure it is left as it was.  Returns the matrix size.
int viscositySolve(int ni, int nj, int nk, double dx, double dt, const float *liquid, const float *solid,
                   const float *visc, double tol, int maxIter, int ncpu, float *u, float *v, float *w,
                   ViscResult &res) {
    ViscSystem S;
    viscosityVolumes(ni, nj, nk, dx, liquid, solid, S);
    viscositySystem(dx, dt, visc, u, v, w, S);
    res = ViscResult();
    if (S.m == 0) return 0;
    std::vector<float> x;
    res = viscPCG(S.A, S.rhs, tol, maxIter, ncpu, x);
    if (!res.success) return S.m;
    const size_t nU = (size_t)(ni + 1) * nj * nk, nV = (size_t)ni * (nj + 1) * nk, nW = (size_t)ni * nj * (nk + 1);
    std::fill(u, u + nU, 0.0f);
    std::fill(v, v + nV, 0.0f);
    std::fill(w, w + nW, 0.0f);
    for (size_t q = 0; q < nU; q++)
        if (S.table[q] != -1) u[q] = x[S.table[q]];
    for (size_t q = 0; q < nV; q++)
        if (S.table[nU + q] != -1) v[q] = x[S.table[nU + q]];
    for (size_t q = 0; q < nW; q++)
        if (S.table[nU + nV + q] != -1) w[q] = x[S.table[nU + nV + q]];
    return S.m;
}
}  // namespace

struct OracleViscSystem {
    ViscSystem S;
    int nnz = 0;
};
void *oracle_viscosity_system_new(int ni, int nj, int nk, double dx, double dt, const float *liquid_phi,
                                  const float *solid_phi_nodes, const float *viscosity, const float *u,
                                  const float *v, const float *w, int *m, int *nnz) {
    OracleViscSystem *h = new OracleViscSystem();
    viscosityVolumes(ni, nj, nk, dx, liquid_phi, solid_phi_nodes, h->S);
    viscositySystem(dx, dt, viscosity, u, v, w, h->S);
    for (const auto &r : h->S.A.idx) h->nnz += (int)r.size();
    *m = h->S.m;
    *nnz = h->nnz;
    return h;
}
void oracle_viscosity_system_get(void *handle, unsigned char *states, float *volumes, int *table, int *rowstart,
                                 int *cols, float *vals, float *rhs) {
    const ViscSystem &S = ((OracleViscSystem *)handle)->S;
    if (states) {
        states = std::copy(S.su.begin(), S.su.end(), states);
        states = std::copy(S.sv.begin(), S.sv.end(), states);
        std::copy(S.sw.begin(), S.sw.end(), states);
    }
    if (volumes)
        for (const std::vector<float> *g : {&S.center, &S.vu, &S.vv, &S.vw, &S.eu, &S.ev, &S.ew})
            volumes = std::copy(g->begin(), g->end(), volumes);
    if (table) std::copy(S.table.begin(), S.table.end(), table);
    if (rowstart) {
        int q = 0;
        rowstart[0] = 0;
        for (int r = 0; r < S.m; r++) {
            if (cols) std::copy(S.A.idx[r].begin(), S.A.idx[r].end(), cols + q);
            if (vals) std::copy(S.A.val[r].begin(), S.A.val[r].end(), vals + q);
            q += (int)S.A.idx[r].size();
            rowstart[r + 1] = q;
        }
    }
    if (rhs) std::copy(S.rhs.begin(), S.rhs.end(), rhs);
}
void oracle_viscosity_system_free(void *handle) { delete (OracleViscSystem *)handle; }

int oracle_viscosity_pcg(int m, const int *rowstart, const int *cols, const float *vals, const float *rhs,
                         double tol, int max_iter, int nthreads, float *x, int *iterations, float *error) {
    SparseRows A(m);
    for (int r = 0; r < m; r++) {
        A.idx[r].assign(cols + rowstart[r], cols + rowstart[r + 1]);
        A.val[r].assign(vals + rowstart[r], vals + rowstart[r + 1]);
    }
    std::vector<float> b(rhs, rhs + m), sol;
    ViscResult res = viscPCG(A, b, tol, max_iter, nthreads, sol);
    std::copy(sol.begin(), sol.end(), x);
    *iterations = res.iterations;
    *error = res.error;
    return res.success ? 1 : 0;
}

int oracle_viscosity_solve(int ni, int nj, int nk, double dx, double dt, const float *liquid_phi,
                           const float *solid_phi_nodes, const float *viscosity, double tol, int max_iter,
                           int nthreads, float *u, float *v, float *w, int *iterations, float *error,
                           int *success) {
    ViscResult res;
    int m = viscositySolve(ni, nj, nk, dx, dt, liquid_phi, solid_phi_nodes, viscosity, tol, max_iter, nthreads, u,
                           v, w, res);
    *iterations = res.iterations;
    *error = res.error;
    *success = res.success ? 1 : 0;
    return m;
}

struct OracleSim {
    int ni, nj, nk;
    double dx;
    // settings (fluidsimulation.h:2158-2470 defaults)
    double cfl = 5.0;
    double picflip = 0.05;
    int minSteps = 1, maxSteps = 6;
    double ptol = 1e-9;
    int maxPIters = 900;
    int chunks = 1;
    bool extremeRemoval = true;
    bool apic = false;  // velocity transfer method (fluidsimulation.cpp:2837-2857)
    // constant viscosity (setViscosity, fluidsimulation.cpp:2569-2585): off
    // until set; the solver's tolerance and iteration limit (fluidsimulation.h:2364-2365)
    bool viscEnabled = false;
    double viscosity = 0.0;
    // surface tension (setSurfaceTension, fluidsimulation.cpp:2608-2621;
    // fluidsimulation.h:2191, 2236)
    bool tensionEnabled = false, smoothTension = false;
    double tensionK = 0.0, tensionCondition = 10.0;
    double viscTol = 1e-4;
    int viscMaxIters = 900;
    std::vector<V3> bodyForces;
    // Hooks of the differentiable rollout (no counterpart in the reference):
    // particle removal can be switched off; a substep can record its liquid
    // SDF and P2G flags, or reuse recorded ones (then P2G normalises a face by
    // its weight sum exactly where the recorded flag is set).
    bool removal = true;
    const float *frozenPhi = nullptr;            // N
    const unsigned char *frozenValid = nullptr;  // U, V, W flags, concatenated
    std::vector<float> *recordPhi = nullptr;
    std::vector<unsigned char> *recordValid = nullptr;
    std::vector<float> wsum[3];                  // weight sums of the last FLIP P2G
    // state
    bool initialized = false;
    int frame = 0, stepNumber = 0;
    double frameDt = 0, remaining = 0;
    std::vector<V3> pos, vel;
    std::vector<float> aff;  // APIC: 9 per particle (AFFINEX, AFFINEY, AFFINEZ), zero on load
    std::vector<std::vector<V3>> loadPos, loadVel;  // one batch per load call
    bool solidGiven = false;
    std::vector<float> solid;         // nodal
    std::vector<float> wu, wv, ww;    // face weights
    std::vector<char> nearSolid;
    int nsi = 0, nsj = 0, nsk = 0;
    double nsCell = 0;
    std::vector<float> phi, u, v, w, su, sv, sw, pres;
    std::vector<unsigned char> vu, vv, vw;
    // mesh obstacles with their velocities: the solid is rebuilt at a
    // substep's start when the list changed or an animated mesh moves
    // (fluidsimulation.cpp:5387-5426, 5550-5570); the sign tests take their
    // jitter from the given rand() values in order
    std::vector<MovingObstacle> movers;
    bool moversChanged = false;
    size_t rndUsed = 0;  // rand() draws of the solid updates so far
    std::vector<float> us, vs, ws, wc;  // solid face velocities, cell weights
    int lastSubsteps = 0, lastIters = 0, pIters = 0;
    bool pSuccess = true;
    double lastErr = 0, pErr = 0;
    // viscosity solve statistics of the frame, kept like the pressure's
    // (fluidsimulation.cpp:6039-6053)
    int vIters = 0, lastVIters = 0, vSize = 0;
    bool vSuccess = true, lastVSuccess = true;
    float vErr = 0.0f, lastVErr = 0.0f;

    size_t nU() const { return (size_t)(ni + 1) * nj * nk; }
    size_t nV() const { return (size_t)ni * (nj + 1) * nk; }
    size_t nW() const { return (size_t)ni * nj * (nk + 1); }
    size_t nN() const { return (size_t)ni * nj * nk; }
    size_t nNodes() const { return (size_t)(ni + 1) * (nj + 1) * (nk + 1); }

    V3 bodyForce() const {
        V3 b{0, 0, 0};
        for (const V3 &f : bodyForces) b = b + f;
        return b;
    }

    // FluidSimulation::_initializeSimulation / _loadMarkerParticles (fluidsimulation.cpp:4362-4398, 5010-5016)
    void initialize() {
        if (initialized) return;
        Box bounds{V3{0, 0, 0}, ni * dx, nj * dx, nk * dx};
        for (size_t b = 0; b < loadPos.size(); b++) {
            const std::vector<V3> &P = loadPos[b], &Vl = loadVel[b];
            for (size_t p = 0; p < P.size(); p++) {
                size_t q = oracle_load_queue_index(p);
                if (bounds.inside(P[q])) {
                    pos.push_back(P[q]);
                    vel.push_back(Vl[q]);
                    if (apic) aff.insert(aff.end(), 9, 0.0f);
                }
            }
        }
        loadPos.clear();
        loadVel.clear();
        if (!solidGiven) {
            solid.resize(nNodes());
            oracle_solid_box_phi(ni, nj, nk, dx, solid.data());
        }
        wu.resize(nU());
        wv.resize(nV());
        ww.resize(nW());
        oracle_face_weights(ni, nj, nk, solid.data(), wu.data(), wv.data(), ww.data());
        initialized = true;
    }

    // fluidsimulation.cpp:4351; scaled by 2 for the smooth surface tension
    // kernel (5593-5597, 5635-5638, fluidsimulation.h:2236-2239)
    double liquidRadius() const {
        const double r = 0.5 * 1.0 * dx * std::sqrt(3.0);
        return tensionEnabled && smoothTension ? 2.0 * r : r;
    }

    // _updateNearSolidGrid (fluidsimulation.cpp:5436-5480)
    void updateNearSolid() {
        nsCell = 3 * dx;
        nsi = (int)std::ceil(ni * dx / nsCell);
        nsj = (int)std::ceil(nj * dx / nsCell);
        nsk = (int)std::ceil(nk * dx / nsCell);
        nearSolid.assign((size_t)nsi * nsj * nsk, 0);
        float maxd = 3 * dx;
        for (int k = 0; k < nk; k++)
            for (int j = 0; j < nj; j++)
                for (int i = 0; i < ni; i++)
                    if (std::abs(solid[flat(i, j, k, ni + 1, nj + 1)]) < maxd)
                        nearSolid[flat(i / 3, j / 3, k / 3, nsi, nsj)] = 1;
        int layers = (int)std::ceil((float)cfl / (float)3);
        for (int l = 0; l < layers; l++) feather6(nearSolid, nsi, nsj, nsk);
    }

    // _getMaximumMarkerParticleSpeed / _predictMaximumMarkerParticleSpeed /
    // _calculateNextTimeStep (fluidsimulation.cpp:10165-10263)
    double nextTimeStep(double dt) {
        double maxu = 0.0;
        if (frame == 0 && stepNumber == 0) {
            maxu += length(bodyForce()) * dt;
        } else {
            double maxsq = 0.0;
            for (const V3 &v0 : vel) {
                double d = dot(v0, v0);
                if (d > maxsq) maxsq = d;
            }
            maxu = std::sqrt(maxsq);
        }
        if (!pos.empty()) maxu = std::max(0.0, maxu);
        double eps = 1e-6;
        double timeStep = cfl * dx / (maxu + eps);
        if (tensionEnabled && !pos.empty()) {  // fluidsimulation.cpp:10248-10251
            double restriction = std::sqrt(dx * dx * dx) * std::sqrt(1.0 / (tensionK + eps));
            timeStep = std::min(timeStep, tensionCondition * restriction);
        }
        int nsub = std::max((int)std::ceil(dt / timeStep), 1);
        return dt / nsub;
    }

    // FluidSimulation::update (fluidsimulation.cpp:10410-10501)
    void update(double dt) {
        if (!initialized) throw std::runtime_error("Error: FluidSimulation must be initialized before update.\n");
        if (dt < 0.0) throw std::domain_error("Error: delta time must be greater than or equal to 0.\n");
        double epsdt = 1e-6;
        bool zeroLen = dt < epsdt;
        dt = std::max(dt, epsdt);
        frameDt = dt;
        remaining = dt;
        stepNumber = 0;
        bool skipped = zeroLen && frame > 0;
        double substepTime = frameDt / (double)minSteps;
        pSuccess = true;
        pIters = 0;
        pErr = 0.0;
        double eps = 1e-9;
        do {
            double step = std::fmin(nextTimeStep(dt), remaining);
            double done = frameDt - remaining;
            double limit = (stepNumber + 1) * substepTime;
            if (done + step > limit) step = std::fmin(substepTime, remaining);
            if (stepNumber == maxSteps - 1) step = remaining;
            remaining -= step;
            if (!skipped) stepFluid(step);
            stepNumber++;
        } while (remaining > eps);
        lastSubsteps = stepNumber;
        lastIters = pIters;
        lastErr = pErr;
        lastVIters = vIters;
        lastVErr = vErr;
        lastVSuccess = vSuccess;
        frame++;
    }

    // FluidSimulation::_stepFluid (fluidsimulation.cpp:10078-10121), plain configuration
    // MeshObject::_isMeshChanged (meshobject.cpp:893-912)
    static bool meshChanged(const MovingObstacle &o) {
        if (!o.animated) return false;
        if (o.prev.v.size() != o.cur.v.size()) return true;
        for (size_t q = 0; q < o.prev.v.size(); q++)
            if (length(o.prev.v[q] - o.cur.v[q]) > 1e-5f) return true;
        return false;
    }
    // _updateSolidLevelSet (fluidsimulation.cpp:5387-5426) at the substep's
    // frame progress, then the weight grid (6066-6148)
    void updateMovers(double dt) {
        bool rebuild = moversChanged;
        for (const MovingObstacle &o : movers) rebuild = rebuild || meshChanged(o);
        if (!rebuild) return;
        moversChanged = false;
        float progress = 1.0f - (float)(remaining + dt) / (float)frameDt;
        // _stepFluid seeds rand() with frame + substep before the obstacle
        // update (fluidsimulation.cpp:10079)
        GlibcRand g((unsigned)(frame + stepNumber));
        std::vector<int> vals(4096);
        for (int &x : vals) x = g.next();
        RandSource rs{vals.data(), (int)vals.size(), 0, 2147483647.0};
        solidUpdate(ni, nj, nk, dx, movers, frameDt, progress, rs, solid, us, vs, ws);
        rndUsed += rs.used;
        oracle_face_weights(ni, nj, nk, solid.data(), wu.data(), wv.data(), ww.data());
        cellWeights(ni, nj, nk, solid.data(), wc);
    }

    void stepFluid(double dt) {
        bool fluid = !pos.empty();
        if (fluid && !movers.empty()) updateMovers(dt);
        if (fluid) updateNearSolid();
        // liquid level set (5586-5617)
        phi.assign(nN(), 0.0f);
        std::vector<float> P(3 * pos.size()), Vv(3 * pos.size());
        for (size_t p = 0; p < pos.size(); p++) {
            P[3 * p] = pos[p].x; P[3 * p + 1] = pos[p].y; P[3 * p + 2] = pos[p].z;
            Vv[3 * p] = vel[p].x; Vv[3 * p + 1] = vel[p].y; Vv[3 * p + 2] = vel[p].z;
        }
        int np = (int)pos.size();
        if (fluid)
            oracle_liquid_sdf(ni, nj, nk, dx, np, P.data(), liquidRadius(), phi.data());
        else
            std::fill(phi.begin(), phi.end(), (float)(3.0 * dx));
        oracle_liquid_sdf_postprocess(ni, nj, nk, dx, solid.data(), phi.data());
        if (frozenPhi) std::copy(frozenPhi, frozenPhi + nN(), phi.begin());
        if (recordPhi) *recordPhi = phi;
        // advect velocity field (5623-5661)
        u.assign(nU(), 0.0f); v.assign(nV(), 0.0f); w.assign(nW(), 0.0f);
        vu.assign(nU(), 0); vv.assign(nV(), 0); vw.assign(nW(), 0);
        int layers = (int)std::ceil(std::sqrt(3) * cfl) + 3;  // _extrapolateFluidVelocities 6282-6286
        if (fluid && apic) {
            oracle_p2g_apic(ni, nj, nk, dx, np, P.data(), Vv.data(), aff.data(), liquidRadius(), 0, u.data(),
                            vu.data());
            oracle_p2g_apic(ni, nj, nk, dx, np, P.data(), Vv.data(), aff.data(), liquidRadius(), 1, v.data(),
                            vv.data());
            oracle_p2g_apic(ni, nj, nk, dx, np, P.data(), Vv.data(), aff.data(), liquidRadius(), 2, w.data(),
                            vw.data());
            extrapolateAll(layers);
        } else if (fluid) {
            wsum[0].resize(nU());
            wsum[1].resize(nV());
            wsum[2].resize(nW());
            p2gFlip(ni, nj, nk, dx, np, P.data(), Vv.data(), liquidRadius(), 0, u.data(), vu.data(), wsum[0].data());
            p2gFlip(ni, nj, nk, dx, np, P.data(), Vv.data(), liquidRadius(), 1, v.data(), vv.data(), wsum[1].data());
            p2gFlip(ni, nj, nk, dx, np, P.data(), Vv.data(), liquidRadius(), 2, w.data(), vw.data(), wsum[2].data());
            if (frozenValid) freezeP2G();
            if (recordValid) {
                recordValid->assign(vu.begin(), vu.end());
                recordValid->insert(recordValid->end(), vv.begin(), vv.end());
                recordValid->insert(recordValid->end(), vw.begin(), vw.end());
            }
            extrapolateAll(layers);
        }
        if (!fluid) return;
        // save (5671-5685)
        su = u; sv = v; sw = w;
        // body forces (5837-5876): every face, no inflow constraints on this path
        V3 bf = bodyForce();
        float beps = 1e-6;
        if (std::fabs(bf.x) > beps) for (float &x : u) x += (float)(bf.x * dt);
        if (std::fabs(bf.y) > beps) for (float &x : v) x += (float)(bf.y * dt);
        if (std::fabs(bf.z) > beps) for (float &x : w) x += (float)(bf.z * dt);
        applyViscosity(dt);
        // pressure (6152-6276)
        pres.assign(nN(), 0.0f);
        int status[2];
        double err = 0;
        const bool moving = !movers.empty() && !us.empty();
        // the curvature of the liquid surface and the clusters of the surface
        // tension terms (fluidsimulation.cpp:5704-5730, pressuresolver.cpp:314-606)
        Tension T;
        if (tensionEnabled) buildTension(ni, nj, nk, dx, tensionK, phi.data(), T);
        const Tension *tp = tensionEnabled ? &T : nullptr;
        pressureSolveImpl(ni, nj, nk, dx, dt, u.data(), v.data(), w.data(), wu.data(), wv.data(), ww.data(),
                          phi.data(), ptol, maxPIters, chunks, pres.data(), status, &err,
                          moving ? us.data() : nullptr, moving ? vs.data() : nullptr, moving ? ws.data() : nullptr,
                          moving ? wc.data() : nullptr, tp);
        bool success = status[1] != 0;
        if (success)
            applyPressure(ni, nj, nk, dx, dt, pres.data(), phi.data(), wu.data(), wv.data(), ww.data(), u.data(),
                          v.data(), w.data(), vu.data(), vv.data(), vw.data(), tp);
        if (stepNumber == 0) {
            pSuccess = success;
            pIters = status[0];
            pErr = (float)err;
        } else if (pSuccess && (!success || status[0] > pIters)) {
            pSuccess = success;
            pIters = status[0];
            pErr = (float)err;
        }
        extrapolateAll(layers);
        // constrain saved then current (6444-6459)
        constrain(su, sv, sw);
        constrain(u, v, w);
        // G2P: FLIP (6771-6784) or APIC (6791-6843)
        Mac cur{ni, nj, nk, dx, u.data(), v.data(), w.data()};
        Mac sav{ni, nj, nk, dx, su.data(), sv.data(), sw.data()};
        if (apic) {
            for (size_t p = 0; p < pos.size(); p++) P[3 * p] = pos[p].x, P[3 * p + 1] = pos[p].y, P[3 * p + 2] = pos[p].z;
            oracle_g2p_apic(ni, nj, nk, dx, u.data(), v.data(), w.data(), np, P.data(), Vv.data(), aff.data());
            for (size_t p = 0; p < pos.size(); p++) vel[p] = V3{Vv[3 * p], Vv[3 * p + 1], Vv[3 * p + 2]};
        } else {
            float r1 = (float)picflip, r2 = (float)(1 - picflip);
#pragma omp parallel for num_threads(g_threads) schedule(static)
            for (size_t p = 0; p < pos.size(); p++) {
                V3 vPIC = sampleMac(cur, pos[p]);
                V3 vFLIP = vel[p] + vPIC - sampleMac(sav, pos[p]);
                vel[p] = r1 * vPIC + r2 * vFLIP;
            }
        }
        advance(dt);
    }

    // _applyViscosityToVelocityField (fluidsimulation.cpp:5983-6056) with a
    // constant viscosity: the field constrained, then the viscosity solve
    void applyViscosity(double dt) {
        if (!viscEnabled || pos.empty() || viscosity == 0.0) return;
        constrain(u, v, w);
        std::vector<float> grid(nNodes(), (float)viscosity);
        ViscResult res;
        vSize = viscositySolve(ni, nj, nk, dx, dt, phi.data(), solid.data(), grid.data(), viscTol, viscMaxIters,
                               chunks, u.data(), v.data(), w.data(), res);
        if (stepNumber == 0) {
            vSuccess = res.success;
            vIters = res.iterations;
            vErr = res.error;
        } else if (vSuccess && (!res.success || res.iterations > vIters)) {
            vSuccess = res.success;
            vIters = res.iterations;
            vErr = res.error;
        }
    }

    // P2G with the recorded flags: S/W where a flag is recorded (0 without
    // weight), the plain sum S elsewhere; the flags become the recorded ones
    void freezeP2G() {
        std::vector<float> *f[3] = {&u, &v, &w};
        std::vector<unsigned char> *fl[3] = {&vu, &vv, &vw};
        const unsigned char *rec = frozenValid;
        for (int d = 0; d < 3; d++) {
            std::vector<float> &a = *f[d];
            for (size_t q = 0; q < a.size(); q++) {
                const float W = wsum[d][q];
                const bool want = rec[q] != 0, got = W > 1e-6f;
                if (want && !got)
                    a[q] = W > 0.0f ? a[q] / W : 0.0f;
                else if (!want && got)
                    a[q] = a[q] * W;
                (*fl[d])[q] = rec[q];
            }
            rec += a.size();
        }
    }

    void extrapolateAll(int layers) {
        oracle_extrapolate(ni + 1, nj, nk, u.data(), vu.data(), layers);
        oracle_extrapolate(ni, nj + 1, nk, v.data(), vv.data(), layers);
        oracle_extrapolate(ni, nj, nk + 1, w.data(), vw.data(), layers);
    }

    // _constrainVelocityFieldThread (fluidsimulation.cpp:6388-6442): domain friction 0.0
    // (fluidsimulation.h:2227), solid face velocity 0 for static solids.
    // _constrainVelocityFieldThread (fluidsimulation.cpp:6392-6442) without
    // friction; the solid face velocities of moving obstacles, else 0
    void constrain(std::vector<float> &a, std::vector<float> &b, std::vector<float> &c) {
        const bool moving = !movers.empty() && !us.empty();
        float f = 0.0f;
        auto one = [&](std::vector<float> &x, const std::vector<float> &wt, const std::vector<float> &sv) {
            for (size_t q = 0; q < x.size(); q++) {
                float sface = moving ? sv[q] : 0.0f;
                if (wt[q] == 0)
                    x[q] = sface;
                else if (wt[q] < 1.0f)
                    x[q] = f * sface + (1.0f - f) * x[q];
            }
        };
        one(a, wu, us);
        one(b, wv, vs);
        one(c, ww, ws);
    }

    // _RK3 (fluidsimulation.cpp:7616-7623)
    V3 rk3(const Mac &M, V3 p0, double dt) {
        V3 k1 = sampleMac(M, p0);
        V3 k2 = sampleMac(M, p0 + (float)(0.5 * dt) * k1);
        V3 k3 = sampleMac(M, p0 + (float)(0.75 * dt) * k2);
        return p0 + (float)(dt / 9.0f) * (2.0f * k1 + 3.0f * k2 + 4.0f * k3);
    }

    float solidAt(V3 p) const { return nodalTrilinear(solid.data(), ni + 1, nj + 1, nk + 1, dx, p); }
    bool nearSolidAt(V3 p) const {
        GI g = posToIndex(p, nsCell);
        return nearSolid[flat(g.i, g.j, g.k, nsi, nsj)] != 0;
    }

    // _resolveCollision (fluidsimulation.cpp:7646-7721)
    V3 resolveCollision(V3 oldp, V3 newp, const Box &boundary) {
        GI gg = posToIndex(newp, dx);
        if (!inRange(gg.i, gg.j, gg.k, ni, nj, nk)) newp = boundary.nearestInside(newp);
        if (!nearSolidAt(oldp) && !nearSolidAt(newp)) return newp;
        float eps = 1e-6;
        float stepDistance = 0.1f * (float)dx;
        float travel = length(newp - oldp);
        if (travel < eps) return newp;
        int numSteps = (int)std::ceil(travel / stepDistance);
        V3 stepdir = normalize(newp - oldp);
        V3 last = oldp, current{0, 0, 0};
        bool found = false;
        float collisionPhi = 0.0f;
        for (int s = 0; s < numSteps; s++) {
            if (s == numSteps - 1)
                current = newp;
            else
                current = oldp + (float)(s + 1) * stepDistance * stepdir;
            float ph = solidAt(current);
            if (ph < 0.0f || !boundary.inside(current)) {
                collisionPhi = ph;
                found = true;
                break;
            }
            last = current;
        }
        if (!found) return newp;
        V3 resolved;
        float maxResolved = cfl * dx;
        V3 grad = nodalGradient(solid.data(), ni + 1, nj + 1, nk + 1, dx, current);
        if (length(grad) > eps) {
            grad = normalize(grad);
            resolved = current - (float)(collisionPhi - 0.2f * dx) * grad;
            float rphi = solidAt(resolved);
            float rdist = length(resolved - current);
            if (rphi < 0 || rdist > maxResolved) resolved = last;
        } else {
            resolved = last;
        }
        if (!boundary.inside(resolved)) {
            V3 orig = resolved;
            resolved = boundary.nearestInside(resolved);
            float rphi = solidAt(resolved);
            float rdist = length(resolved - orig);
            if (rphi < 0.0f || rdist > maxResolved) resolved = last;
        }
        return resolved;
    }

    // _getMarkerParticleSpeedLimit (fluidsimulation.cpp:7723-7771)
    float speedLimit(double dt) {
        double maxSpeed = 0.0;
        double stepv = cfl * dx / dt;
        std::vector<int> counts(maxSteps, 0);
        for (const V3 &v0 : vel) {
            double speed = (double)length(v0);
            int idx = std::fmin(std::floor(speed / stepv), maxSteps - 1);
            counts[idx]++;
            maxSpeed = std::max(speed, maxSpeed);
        }
        int maxRemoval = std::fmin((int)((double)vel.size() * 0.0005), 35);
        double maxspeed = maxSteps * stepv;
        int cur = 0;
        for (int i = (int)counts.size() - 1; i > 0; i--) {
            if (cur + counts[i] > maxRemoval) break;
            cur += counts[i];
            maxspeed = std::max(i + 4, maxSteps) * stepv;
        }
        double lower = 0.90 * maxSpeed, thr = 0.99999 * maxSpeed;
        int nLower = 0, nOut = 0;
        for (const V3 &v0 : vel) {
            double speed = (double)length(v0);
            if (speed >= lower && speed < thr) nLower++;
            if (speed >= thr) nOut++;
        }
        if (nOut <= 6 && nLower <= 6) maxspeed = std::min(thr, maxspeed);
        return (float)maxspeed;
    }

    // _advanceMarkerParticles + _removeMarkerParticles (fluidsimulation.cpp:7626-7900)
    void advance(double dt) {
        Mac M{ni, nj, nk, dx, u.data(), v.data(), w.data()};
        Box boundary = boundaryBox(ni, nj, nk, dx);
        boundary.expand(-0.2f * dx);
        std::vector<V3> out(pos.size());
#pragma omp parallel for num_threads(g_threads) schedule(static)
        for (size_t p = 0; p < pos.size(); p++) out[p] = resolveCollision(pos[p], rk3(M, pos[p], dt), boundary);
        pos = out;
        if (!removal) return;
        // removal uses the frame delta time (7896)
        float maxspeed = speedLimit(frameDt);
        double maxspeedsq = maxspeed * maxspeed;
        std::vector<int> count(nN(), 0);
        std::vector<char> removed(pos.size(), 0);
#pragma omp parallel for num_threads(g_threads) schedule(static)
        for (size_t p = 0; p < pos.size(); p++) removed[p] = solidAt(pos[p]) < 0.0f;
        for (size_t p = 0; p < pos.size(); p++) {
            if (removed[p]) continue;
            GI g = posToIndex(pos[p], dx);
            int &c = count[flat(g.i, g.j, g.k, ni, nj)];
            if (c >= 250) {
                removed[p] = 1;
                continue;
            }
            c++;
            if (extremeRemoval && dot(vel[p], vel[p]) > maxspeedsq) removed[p] = 1;
        }
        size_t q = 0;
        for (size_t p = 0; p < pos.size(); p++)
            if (!removed[p]) {
                pos[q] = pos[p];
                vel[q] = vel[p];
                if (apic)
                    for (int e = 0; e < 9; e++) aff[9 * q + e] = aff[9 * p + e];
                q++;
            }
        pos.resize(q);
        vel.resize(q);
        if (apic) aff.resize(9 * q);
    }
};

void oracle_set_threads(int n) { g_threads = n < 1 ? 1 : n; }
int oracle_get_threads(void) { return g_threads; }

OracleSim *oracle_sim_new(int ni, int nj, int nk, double dx) {
    OracleSim *s = new OracleSim();
    s->ni = ni;
    s->nj = nj;
    s->nk = nk;
    s->dx = dx;
    return s;
}
void oracle_sim_destroy(OracleSim *s) { delete s; }
void oracle_sim_add_body_force(OracleSim *s, double fx, double fy, double fz) {
    s->bodyForces.push_back(mk(fx, fy, fz));
}
void oracle_sim_set_picflip_ratio(OracleSim *s, double r) { s->picflip = r; }
void oracle_sim_set_apic(OracleSim *s, int on) { s->apic = on != 0; }
void oracle_sim_get_affine(OracleSim *s, float *aff) {
    for (size_t e = 0; e < s->aff.size(); e++) aff[e] = s->aff[e];
}
void oracle_sim_set_cfl(OracleSim *s, int cfl) { s->cfl = cfl; }
void oracle_sim_set_pressure_tolerance(OracleSim *s, double tol) { s->ptol = tol; }
void oracle_sim_set_max_pressure_iterations(OracleSim *s, int n) { s->maxPIters = n; }
void oracle_sim_set_thread_chunks(OracleSim *s, int n) { s->chunks = n < 1 ? 1 : n; }
void oracle_sim_set_viscosity(OracleSim *s, double visc, double tol, int max_iter) {
    s->viscEnabled = true;
    s->viscosity = (float)visc;
    s->viscTol = tol;
    s->viscMaxIters = max_iter;
}
void oracle_sim_set_surface_tension(OracleSim *s, double k, double condition, int smooth) {
    s->tensionEnabled = k > 1e-6;
    s->tensionK = k;
    s->tensionCondition = condition;
    s->smoothTension = smooth != 0;
}
int oracle_sim_last_viscosity_stats(OracleSim *s, int *iterations, float *error) {
    *iterations = s->lastVIters;
    *error = s->lastVErr;
    return s->lastVSuccess ? 1 : 0;
}
int oracle_sim_add_moving_obstacle(OracleSim *s, int nv, const float *prev, const float *cur, const float *next,
                                   int nt, const int *tris, int flags, float expansion, float vscale) {
    try {
        MovingObstacle o;
        o.prev = meshFrom(nv, prev, nt, tris);
        o.cur = meshFrom(nv, cur, nt, tris);
        o.next = meshFrom(nv, next, nt, tris);
        o.animated = (flags & 1) != 0;
        o.inversed = (flags & 2) != 0;
        o.expansion = expansion;
        o.velocityScale = vscale;
        s->movers.push_back(o);
        s->moversChanged = true;
        return (int)s->movers.size() - 1;
    } catch (const std::exception &) {
        return -1;
    }
}
int oracle_sim_update_moving_obstacle(OracleSim *s, int index, const float *prev, const float *cur,
                                      const float *next) {
    if (index < 0 || index >= (int)s->movers.size()) return -1;
    MovingObstacle &o = s->movers[index];
    for (size_t q = 0; q < o.cur.v.size(); q++) {
        o.prev.v[q] = V3{prev[3 * q], prev[3 * q + 1], prev[3 * q + 2]};
        o.cur.v[q] = V3{cur[3 * q], cur[3 * q + 1], cur[3 * q + 2]};
        o.next.v[q] = V3{next[3 * q], next[3 * q + 1], next[3 * q + 2]};
    }
    return 0;
}
int oracle_glibc_rand(unsigned seed, int n, int *out) {
    GlibcRand g(seed);
    for (int q = 0; q < n; q++) out[q] = g.next();
    return n;
}
int oracle_sim_rand_used(OracleSim *s) { return (int)s->rndUsed; }
void oracle_sim_get_solid_velocity(OracleSim *s, float *u, float *v, float *w) {
    if (s->us.empty()) {
        std::fill(u, u + s->nU(), 0.0f);
        std::fill(v, v + s->nV(), 0.0f);
        std::fill(w, w + s->nW(), 0.0f);
        return;
    }
    std::copy(s->us.begin(), s->us.end(), u);
    std::copy(s->vs.begin(), s->vs.end(), v);
    std::copy(s->ws.begin(), s->ws.end(), w);
}
void oracle_sim_get_solid_phi(OracleSim *s, float *phi) { std::copy(s->solid.begin(), s->solid.end(), phi); }

void oracle_sim_set_solid_phi(OracleSim *s, const float *phi) {
    s->solid.assign(phi, phi + s->nNodes());
    s->solidGiven = true;
}
void oracle_sim_load_particles(OracleSim *s, int np, const float *pos, const float *vel) {
    if (np <= 0) return;
    s->loadPos.emplace_back();
    s->loadVel.emplace_back();
    for (int p = 0; p < np; p++) {
        s->loadPos.back().push_back(V3{pos[3 * p], pos[3 * p + 1], pos[3 * p + 2]});
        s->loadVel.back().push_back(V3{vel[3 * p], vel[3 * p + 1], vel[3 * p + 2]});
    }
}
void oracle_sim_initialize(OracleSim *s) { s->initialize(); }
void oracle_sim_update(OracleSim *s, double dt) { s->update(dt); }
void oracle_sim_step(OracleSim *s, double dt) {
    s->frameDt = dt;
    s->remaining = 0.0;  // one substep is the whole frame: progress 0
    s->stepFluid(dt);
    s->stepNumber++;
}
int oracle_sim_num_particles(OracleSim *s) { return (int)s->pos.size(); }
void oracle_sim_get_particles(OracleSim *s, float *pos, float *vel) {
    for (size_t p = 0; p < s->pos.size(); p++) {
        if (pos) {
            pos[3 * p] = s->pos[p].x; pos[3 * p + 1] = s->pos[p].y; pos[3 * p + 2] = s->pos[p].z;
        }
        if (vel) {
            vel[3 * p] = s->vel[p].x; vel[3 * p + 1] = s->vel[p].y; vel[3 * p + 2] = s->vel[p].z;
        }
    }
}
int oracle_sim_last_substeps(OracleSim *s) { return s->lastSubsteps; }
int oracle_sim_last_pressure_iterations(OracleSim *s) { return s->lastIters; }
double oracle_sim_last_pressure_error(OracleSim *s) { return s->lastErr; }
void oracle_sim_get_grids(OracleSim *s, float *phi, float *u, float *v, float *w, float *p) {
    if (phi && !s->phi.empty()) std::copy(s->phi.begin(), s->phi.end(), phi);
    if (u && !s->u.empty()) std::copy(s->u.begin(), s->u.end(), u);
    if (v && !s->v.empty()) std::copy(s->v.begin(), s->v.end(), v);
    if (w && !s->w.empty()) std::copy(s->w.begin(), s->w.end(), w);
    if (p && !s->pres.empty()) std::copy(s->pres.begin(), s->pres.end(), p);
}

/* ========================= differentiable rollout and its FP64 adjoint ===== */
/* The build's differentiable rollout (no counterpart in the reference; the HIP
 * version is blender-flip-fluids_amd/csrc/rollout.hip): `substeps` substeps of
 * length dt of _stepFluid (fluidsimulation.cpp:10078-10121) without particle
 * removal, from particles at fixed positions whose velocities are the
 * trilinear sample of an initial MAC field, returning the particle density on
 * the cells.
 *
 * The forward is the oracle engine's own substep in float.  The backward is
 * the adjoint of that map computed in double: each stage's vector-Jacobian
 * product from its own derivative, evaluated at the forward's states, with
 * the liquid SDF and P2G flags of the recording forward held fixed and the
 * collision response taken as the identity.  The forward's float roundings
 * are not differentiated; the P2G derivative takes the forward's own float
 * transfer weights, weight sums and face values (the quantities it divided by
 * and stored), and each pressure solve's A^-1 is applied by an FP64
 * MIC(0)-CG run to a relative residual of 1e-13. */
namespace {

constexpr unsigned char XS_BORDER = 254, XS_UNKNOWN = 255;

// The status every face ends with in GridUtils::extrapolateGrid
// (gridutils.h:95-160): 0 valid, L+1 filled in layer L (the mean of its
// neighbours that were done by then), XS_BORDER the array border (always
// done), XS_UNKNOWN never reached.
std::vector<unsigned char> extrapolationStatus(int wi, int wj, int wk, const unsigned char *valid, int layers) {
    const size_t n = (size_t)wi * wj * wk;
    std::vector<unsigned char> st(n);
    for (int k = 0; k < wk; k++)
        for (int j = 0; j < wj; j++)
            for (int i = 0; i < wi; i++) {
                const size_t f = flat(i, j, k, wi, wj);
                const bool border = i == 0 || j == 0 || k == 0 || i == wi - 1 || j == wj - 1 || k == wk - 1;
                st[f] = border ? XS_BORDER : (valid[f] ? 0 : XS_UNKNOWN);
            }
    const long long nb[6] = {1, -1, wi, -(long long)wi, (long long)wi * wj, -(long long)wi * wj};
    std::vector<size_t> filled;
    for (int L = 0; L < layers; L++) {
        filled.clear();
        for (size_t f = 0; f < n; f++) {
            if (st[f] != XS_UNKNOWN) continue;  // not on the border, so all six neighbours exist
            for (int q = 0; q < 6; q++)
                if (st[(long long)f + nb[q]] == (unsigned char)L) {
                    filled.push_back(f);
                    break;
                }
        }
        for (size_t f : filled) st[f] = (unsigned char)(L + 1);
    }
    return st;
}

bool doneAt(unsigned char s, int L) { return s == XS_BORDER || s <= (unsigned char)L; }

// transpose of the layered extrapolation for the statuses it left: the
// layers in reverse, each filled face handing its cotangent in equal shares
// to the neighbours it averaged; filled faces lose their own (their input
// was overwritten)
void extrapolationAdjoint(int wi, int wj, int wk, const std::vector<unsigned char> &st, int layers, double *bar) {
    const size_t n = (size_t)wi * wj * wk;
    const long long nb[6] = {1, -1, wi, -(long long)wi, (long long)wi * wj, -(long long)wi * wj};
    for (int L = layers - 1; L >= 0; L--)
        for (size_t f = 0; f < n; f++) {
            if (st[f] != (unsigned char)(L + 1)) continue;
            int count = 0;
            for (int q = 0; q < 6; q++) count += doneAt(st[(long long)f + nb[q]], L) ? 1 : 0;
            const double share = bar[f] / count;
            for (int q = 0; q < 6; q++)
                if (doneAt(st[(long long)f + nb[q]], L)) bar[(long long)f + nb[q]] += share;
        }
    for (size_t f = 0; f < n; f++)
        if (st[f] != 0 && st[f] != XS_BORDER && st[f] != XS_UNKNOWN) bar[f] = 0.0;
}

struct FaceDims {
    int wi, wj, wk;
    size_t n() const { return (size_t)wi * wj * wk; }
};
FaceDims faceDims(int ni, int nj, int nk, int d) { return FaceDims{ni + (d == 0), nj + (d == 1), nk + (d == 2)}; }

// The trilinear stencil of component d at p (macvelocityfield.cpp:519-645):
// base corner and fractions; false outside the domain (the sample is 0 there)
bool sampleStencil(int ni, int nj, int nk, double dx, int d, V3 p, int &i, int &j, int &k, double &fx, double &fy,
                   double &fz) {
    double x = p.x, y = p.y, z = p.z;
    if (!inGridD(x, y, z, dx, ni, nj, nk)) return false;
    const double h = 0.5 * dx;
    if (d != 0) x -= h;
    if (d != 1) y -= h;
    if (d != 2) z -= h;
    const double inv = 1.0 / dx;
    i = (int)std::floor(x * inv);
    j = (int)std::floor(y * inv);
    k = (int)std::floor(z * inv);
    fx = (x - (double)i * dx) * inv;
    fy = (y - (double)j * dx) * inv;
    fz = (z - (double)k * dx) * inv;
    return true;
}

// bar_f += d(sample)/d(face f) * cot for the three components of one point
void sampleAdjointAdd(int ni, int nj, int nk, double dx, V3 p, const double cot[3], std::vector<double> *bar) {
    for (int d = 0; d < 3; d++) {
        if (cot[d] == 0.0) continue;
        const FaceDims D = faceDims(ni, nj, nk, d);
        int i, j, k;
        double fx, fy, fz;
        if (!sampleStencil(ni, nj, nk, dx, d, p, i, j, k, fx, fy, fz)) continue;
        for (int q = 0; q < 8; q++) {
            const int ox = q & 1, oy = (q >> 1) & 1, oz = q >> 2;
            if (!inRange(i + ox, j + oy, k + oz, D.wi, D.wj, D.wk)) continue;
            const double wgt = (ox ? fx : 1 - fx) * (oy ? fy : 1 - fy) * (oz ? fz : 1 - fz);
            bar[d][flat(i + ox, j + oy, k + oz, D.wi, D.wj)] += wgt * cot[d];
        }
    }
}

// G[d][e] = d V_d / d p_e of the trilinear sample inside the cell the
// forward picks; 0 outside the domain
void sampleGradient(const Mac &M, V3 p, double G[3][3]) {
    for (int d = 0; d < 3; d++)
        for (int e = 0; e < 3; e++) G[d][e] = 0.0;
    const float *arr[3] = {M.u, M.v, M.w};
    for (int d = 0; d < 3; d++) {
        const FaceDims D = faceDims(M.ni, M.nj, M.nk, d);
        int i, j, k;
        double fx, fy, fz;
        if (!sampleStencil(M.ni, M.nj, M.nk, M.dx, d, p, i, j, k, fx, fy, fz)) return;
        for (int q = 0; q < 8; q++) {
            const int ox = q & 1, oy = (q >> 1) & 1, oz = q >> 2;
            if (!inRange(i + ox, j + oy, k + oz, D.wi, D.wj, D.wk)) continue;
            const double val = arr[d][flat(i + ox, j + oy, k + oz, D.wi, D.wj)];
            const double ax = ox ? fx : 1 - fx, ay = oy ? fy : 1 - fy, az = oz ? fz : 1 - fz;
            G[d][0] += val * (ox ? 1.0 : -1.0) * ay * az / M.dx;
            G[d][1] += val * ax * (oy ? 1.0 : -1.0) * az / M.dx;
            G[d][2] += val * ax * ay * (oz ? 1.0 : -1.0) / M.dx;
        }
    }
}

// The transfer kernel w(d2) = 1 - c1 d2^3 + c2 d2^2 - c3 d2 for d2 < r^2
// (velocityadvector.cpp:467-536), its constants as the forward's floats
struct Kernel64 {
    double rsq, c1, c2, c3;
    explicit Kernel64(double radius) {
        const float r = (float)radius;
        rsq = (float)(r * r);
        c1 = (4.0f / 9.0f) * (1.0f / (r * r * r * r * r * r));
        c2 = (17.0f / 9.0f) * (1.0f / (r * r * r * r));
        c3 = (22.0f / 9.0f) * (1.0f / (r * r));
    }
    double w(double d2) const { return 1.0 - c1 * d2 * d2 * d2 + c2 * d2 * d2 - c3 * d2; }
    double dw(double d2) const { return -3.0 * c1 * d2 * d2 + 2.0 * c2 * d2 - c3; }
};

// faces of direction d whose transfer support can hold p (shifted frame)
void p2gFaceRange(const FaceDims &D, double dx, double r, const double ps[3], int lo[3], int hi[3]) {
    const int w[3] = {D.wi, D.wj, D.wk};
    for (int e = 0; e < 3; e++) {
        lo[e] = std::max(0, (int)std::floor((ps[e] - r) / dx));
        hi[e] = std::min(w[e] - 1, (int)std::floor((ps[e] + r) / dx) + 1);
    }
}

// applyPressure's rule for one face (pressuresolver.cpp:1047-1205): kind 0
// kept (index 0 or n-1 along its axis), 1 updated by the pressure gradient
// with dLo / dHi its derivatives in the low / high cell's pressure, 2 zeroed
struct FaceRule64 {
    int kind = 0;
    double dLo = 0, dHi = 0;
    long long la = -1, hb = -1;  // cells, -1 outside the grid
};
FaceRule64 faceRule64(int ni, int nj, int nk, double dx, double dt, const float *phi, const float *wgt, int d, int i,
                      int j, int k) {
    const FaceDims D = faceDims(ni, nj, nk, d);
    FaceRule64 r;
    const int c = d == 0 ? i : (d == 1 ? j : k);
    const int nc = d == 0 ? ni : (d == 1 ? nj : nk);
    if (c > 0) r.la = (long long)flat(i - (d == 0), j - (d == 1), k - (d == 2), ni, nj);
    if (c < nc) r.hb = (long long)flat(i, j, k, ni, nj);
    if (c == 0 || c == nc - 1) return r;
    const float factor = (float)(dt / dx);
    const float eps = 1e-6f;
    const double maxtheta = 25;
    const bool lo = r.la >= 0 && phi[r.la] < 0.0f;
    const bool hi = r.hb >= 0 && phi[r.hb] < 0.0f;
    if (!(wgt[flat(i, j, k, D.wi, D.wj)] > 0 && (lo || hi))) {
        r.kind = 2;
        return r;
    }
    r.kind = 1;
    if (lo && hi) {
        r.dLo = (double)factor;
        r.dHi = -(double)factor;
    } else if (lo) {
        const float phi1 = phi[r.la], phi2 = r.hb >= 0 ? phi[r.hb] : phi1;
        const float thetaP = (float)std::fmax(-maxtheta, std::fmin((double)(phi2 / (phi1 + eps)), maxtheta));
        r.dLo = -(double)factor * ((double)thetaP - 1.0);
    } else {
        const float phi1 = r.la >= 0 ? phi[r.la] : phi[r.hb], phi2 = phi[r.hb];
        const float thetaP = (float)std::fmax(-maxtheta, std::fmin((double)(phi1 / (phi2 + eps)), maxtheta));
        r.dHi = -(double)factor * (1.0 - (double)thetaP);
    }
    return r;
}

}  // namespace

struct OracleRollout {
    OracleSim sim;
    double dt = 0;
    int K = 0;
    std::vector<V3> pos0;
    std::vector<std::vector<V3>> P, Vel;       // the state before substep k, k = 0..K
    std::vector<std::vector<float>> recPhi;
    std::vector<std::vector<unsigned char>> recValid;
    std::vector<float> rho;
    bool recorded = false;
    long long solveIterations = 0;

    size_t nFaces(int d) const { return faceDims(sim.ni, sim.nj, sim.nk, d).n(); }
    int layers() const { return (int)std::ceil(std::sqrt(3) * sim.cfl) + 3; }

    // rho_c = sum of w(|c - x_p|^2) over the particles within r of the cell
    // centre, gathered from the 27 surrounding cells in cell order and each
    // cell's particles in index order (the HIP density kernel's order)
    void density(const std::vector<V3> &x) {
        const int ni = sim.ni, nj = sim.nj, nk = sim.nk;
        const double dx = sim.dx;
        const size_t N = sim.nN();
        std::vector<int> cnt(N + 1, 0), start(N + 1, 0), order(x.size());
        std::vector<long long> key(x.size(), -1);
        for (size_t p = 0; p < x.size(); p++) {
            GI g = posToIndex(x[p], dx);
            if (inRange(g.i, g.j, g.k, ni, nj, nk)) key[p] = (long long)flat(g.i, g.j, g.k, ni, nj), cnt[key[p]]++;
        }
        for (size_t c = 0; c < N; c++) start[c + 1] = start[c] + cnt[c];
        std::vector<int> fillp(start.begin(), start.end() - 1);
        for (size_t p = 0; p < x.size(); p++)
            if (key[p] >= 0) order[fillp[key[p]]++] = (int)p;
        const float r = (float)sim.liquidRadius();
        const float rsq = r * r, c1 = (4.0f / 9.0f) * (1.0f / (r * r * r * r * r * r)),
                    c2 = (17.0f / 9.0f) * (1.0f / (r * r * r * r)), c3 = (22.0f / 9.0f) * (1.0f / (r * r));
        const double h = 0.5 * dx;
        rho.assign(N, 0.0f);
        for (int k = 0; k < nk; k++)
            for (int j = 0; j < nj; j++)
                for (int i = 0; i < ni; i++) {
                    const double cx = i * dx + h, cy = j * dx + h, cz = k * dx + h;
                    double acc = 0.0;
                    for (int dk = -1; dk <= 1; dk++)
                        for (int dj = -1; dj <= 1; dj++)
                            for (int di = -1; di <= 1; di++) {
                                if (!inRange(i + di, j + dj, k + dk, ni, nj, nk)) continue;
                                const size_t nc = flat(i + di, j + dj, k + dk, ni, nj);
                                for (int s = start[nc]; s < start[nc + 1]; s++) {
                                    const V3 q = x[order[s]];
                                    const double ex = cx - q.x, ey = cy - q.y, ez = cz - q.z;
                                    const double d2 = ex * ex + ey * ey + ez * ez;
                                    if (d2 < (double)rsq) acc += 1.0 - c1 * d2 * d2 * d2 + c2 * d2 * d2 - c3 * d2;
                                }
                            }
                    rho[flat(i, j, k, ni, nj)] = (float)acc;
                }
    }

    // one substep of the working state (sim.pos, sim.vel); k indexes the
    // recorded masks
    void substep(int k, bool freeze, bool record) {
        sim.frozenPhi = freeze ? recPhi[k].data() : nullptr;
        sim.frozenValid = freeze ? recValid[k].data() : nullptr;
        sim.recordPhi = record ? &recPhi[k] : nullptr;
        sim.recordValid = record ? &recValid[k] : nullptr;
        sim.frameDt = dt;
        sim.stepNumber = 0;  // the solve statistics below are this substep's
        sim.stepFluid(dt);
        solveIterations += sim.pIters;
        sim.frozenPhi = nullptr;
        sim.frozenValid = nullptr;
        sim.recordPhi = nullptr;
        sim.recordValid = nullptr;
        if (!sim.pSuccess) throw std::runtime_error("rollout: a pressure solve did not converge");
    }

    void forward(const float *u0, const float *v0, const float *w0, bool freeze) {
        if (freeze && !recorded) throw std::runtime_error("rollout: no recorded forward to freeze the masks of");
        Mac init{sim.ni, sim.nj, sim.nk, sim.dx, u0, v0, w0};
        sim.pos = pos0;
        sim.vel.resize(pos0.size());
        for (size_t p = 0; p < pos0.size(); p++) sim.vel[p] = sampleMac(init, pos0[p]);
        if (!freeze) {
            recPhi.assign(K, {});
            recValid.assign(K, {});
        }
        P.assign(K + 1, {});
        Vel.assign(K + 1, {});
        solveIterations = 0;
        for (int k = 0; k < K; k++) {
            P[k] = sim.pos;
            Vel[k] = sim.vel;
            substep(k, freeze, !freeze);
        }
        P[K] = sim.pos;
        Vel[K] = sim.vel;
        density(sim.pos);
        if (!freeze) recorded = true;
    }

    // adjoint of substep k: in (xb, vb) for the state after it, out for the
    // state before it
    void adjointSubstep(int k, std::vector<double> &xb, std::vector<double> &vb) {
        const int ni = sim.ni, nj = sim.nj, nk = sim.nk, np = (int)pos0.size();
        const double dx = sim.dx;
        // the forward substep again, for its grids
        sim.pos = P[k];
        sim.vel = Vel[k];
        substep(k, true, false);
        const std::vector<V3> &x0 = P[k], &v0 = Vel[k];
        Mac cur{ni, nj, nk, dx, sim.u.data(), sim.v.data(), sim.w.data()};
        Mac sav{ni, nj, nk, dx, sim.su.data(), sim.sv.data(), sim.sw.data()};  // constrained copy
        std::vector<double> ub[3], sb[3];
        for (int d = 0; d < 3; d++) {
            ub[d].assign(nFaces(d), 0.0);
            sb[d].assign(nFaces(d), 0.0);
        }
        // G2P blend and RK3 (collision response as the identity):
        // v' = r1 V(x) + r2 ((v + V(x)) - Vs(x)), x' = RK3(x; V)
        const float r1 = (float)sim.picflip, r2 = (float)(1 - sim.picflip);
        const float c2 = (float)(0.5 * dt), c3 = (float)(0.75 * dt), a = (float)(dt / 9.0f);
        std::vector<double> xPrev(3 * (size_t)np), vPrev(3 * (size_t)np);
        for (int p = 0; p < np; p++) {
            const V3 p0 = x0[p];
            const V3 k1 = sampleMac(cur, p0);
            const V3 y2 = p0 + c2 * k1;
            const V3 k2 = sampleMac(cur, y2);
            const V3 y3 = p0 + c3 * k2;
            double kb1[3], kb2[3], kb3[3], pb[3], s[3], G[3][3];
            for (int e = 0; e < 3; e++) {
                const double x = xb[3 * (size_t)p + e], v = vb[3 * (size_t)p + e];
                kb1[e] = 2.0 * (double)a * x + ((double)r1 + (double)r2) * v;
                kb2[e] = 3.0 * (double)a * x;
                kb3[e] = 4.0 * (double)a * x;
                pb[e] = x;
                s[e] = -(double)r2 * v;
                vPrev[3 * (size_t)p + e] = (double)r2 * v;
            }
            sampleGradient(sav, p0, G);
            for (int e = 0; e < 3; e++) pb[e] += G[0][e] * s[0] + G[1][e] * s[1] + G[2][e] * s[2];
            sampleGradient(cur, y3, G);
            for (int e = 0; e < 3; e++) {
                const double yb = G[0][e] * kb3[0] + G[1][e] * kb3[1] + G[2][e] * kb3[2];
                pb[e] += yb;
                kb2[e] += (double)c3 * yb;
            }
            sampleGradient(cur, y2, G);
            for (int e = 0; e < 3; e++) {
                const double yb = G[0][e] * kb2[0] + G[1][e] * kb2[1] + G[2][e] * kb2[2];
                pb[e] += yb;
                kb1[e] += (double)c2 * yb;
            }
            sampleGradient(cur, p0, G);
            for (int e = 0; e < 3; e++) pb[e] += G[0][e] * kb1[0] + G[1][e] * kb1[1] + G[2][e] * kb1[2];
            for (int e = 0; e < 3; e++) xPrev[3 * (size_t)p + e] = pb[e];
            sampleAdjointAdd(ni, nj, nk, dx, p0, kb1, ub);
            sampleAdjointAdd(ni, nj, nk, dx, y2, kb2, ub);
            sampleAdjointAdd(ni, nj, nk, dx, y3, kb3, ub);
            sampleAdjointAdd(ni, nj, nk, dx, p0, s, sb);
        }
        // the solid constraint: faces of weight 0 take the solid's velocity
        const float *wt[3] = {sim.wu.data(), sim.wv.data(), sim.ww.data()};
        for (int d = 0; d < 3; d++)
            for (size_t f = 0; f < nFaces(d); f++)
                if (wt[d][f] == 0) ub[d][f] = sb[d][f] = 0.0;
        // the second extrapolation, from the flags the projection left
        const unsigned char *pv[3] = {sim.vu.data(), sim.vv.data(), sim.vw.data()};
        for (int d = 0; d < 3; d++) {
            const FaceDims D = faceDims(ni, nj, nk, d);
            extrapolationAdjoint(D.wi, D.wj, D.wk, extrapolationStatus(D.wi, D.wj, D.wk, pv[d], layers()), layers(),
                                 ub[d].data());
        }
        // the projection u' = K u + C A^-1 B u for the recorded liquid SDF:
        // ubar <- K ubar + B^T A^-1 C^T ubar
        projectionAdjoint(recPhi[k].data(), ub);
        // the body force adds a constant; the saved field is u before it
        for (int d = 0; d < 3; d++)
            for (size_t f = 0; f < nFaces(d); f++) ub[d][f] += sb[d][f];
        // the first extrapolation, from the P2G flags
        const unsigned char *rv = recValid[k].data();
        for (int d = 0; d < 3; d++) {
            const FaceDims D = faceDims(ni, nj, nk, d);
            extrapolationAdjoint(D.wi, D.wj, D.wk, extrapolationStatus(D.wi, D.wj, D.wk, rv, layers()), layers(),
                                 ub[d].data());
            rv += D.n();
        }
        // P2G: u_f = S_f / W_f on flagged faces, S_f elsewhere
        p2gAdjoint(k, x0, v0, ub, xPrev, vPrev);
        xb.swap(xPrev);
        vb.swap(vPrev);
    }

    void projectionAdjoint(const float *phi, std::vector<double> *ub) {
        const int ni = sim.ni, nj = sim.nj, nk = sim.nk;
        const double dx = sim.dx;
        const size_t N = sim.nN();
        const float *wt[3] = {sim.wu.data(), sim.wv.data(), sim.ww.data()};
        std::vector<double> cellBar(N, 0.0);
        std::vector<FaceRule64> rules[3];
        for (int d = 0; d < 3; d++) {
            const FaceDims D = faceDims(ni, nj, nk, d);
            rules[d].resize(D.n());
            for (int k = 0; k < D.wk; k++)
                for (int j = 0; j < D.wj; j++)
                    for (int i = 0; i < D.wi; i++) {
                        const size_t f = flat(i, j, k, D.wi, D.wj);
                        const FaceRule64 r = faceRule64(ni, nj, nk, dx, dt, phi, wt[d], d, i, j, k);
                        rules[d][f] = r;
                        if (r.kind != 1) continue;
                        if (r.la >= 0) cellBar[r.la] += r.dLo * ub[d][f];
                        if (r.hb >= 0) cellBar[r.hb] += r.dHi * ub[d][f];
                    }
        }
        // the system the forward solved (its rhs is not needed)
        std::vector<float> zero[3];
        for (int d = 0; d < 3; d++) zero[d].assign(nFaces(d), 0.0f);
        PressureSystem S;
        buildSystem(ni, nj, nk, dx, dt, zero[0].data(), zero[1].data(), zero[2].data(), wt[0], wt[1], wt[2], phi, S);
        factorMIC0(S);
        // rows the factorisation leaves null (pcgsolver.h:103-108) keep a zero solution
        std::vector<double> b(S.m), lam(S.m, 0.0);
        for (int r = 0; r < S.m; r++) b[r] = S.diag[r] < 1e-9 ? 0.0 : cellBar[S.rows[r]];
        solveFP64(S, b, lam);
        std::vector<double> lamCell(N, 0.0);
        for (int r = 0; r < S.m; r++) lamCell[S.rows[r]] = lam[r];
        for (int d = 0; d < 3; d++) {
            for (size_t f = 0; f < nFaces(d); f++) {
                const FaceRule64 &r = rules[d][f];
                double v = r.kind == 2 ? 0.0 : ub[d][f];
                const double lHi = r.hb >= 0 ? lamCell[r.hb] : 0.0, lLo = r.la >= 0 ? lamCell[r.la] : 0.0;
                v += (1.0 / dx) * (double)wt[d][f] * (lHi - lLo);
                ub[d][f] = v;
            }
        }
    }

    // A x = b by MIC(0)-preconditioned CG in double to |r|_inf <= 1e-13 |b|_inf
    void solveFP64(PressureSystem &S, const std::vector<double> &b, std::vector<double> &x) {
        const int m = S.m;
        x.assign(m, 0.0);
        double bmax = 0.0;
        for (double v : b) bmax = std::max(bmax, std::fabs(v));
        if (m == 0 || bmax == 0.0) return;
        std::vector<double> nc = nullColumns(S);
        std::vector<double> r = b, z(m), s(m), q(m);
        applyMIC0(S, nc, r.data(), z.data());
        s = z;
        double rho = 0.0;
        for (int i = 0; i < m; i++) rho += z[i] * r[i];
        for (int it = 0; it < 20000; it++) {
            spmv(S, s, q);
            double sq = 0.0;
            for (int i = 0; i < m; i++) sq += s[i] * q[i];
            const double alpha = rho / sq;
            double rmax = 0.0;
            for (int i = 0; i < m; i++) {
                x[i] += alpha * s[i];
                r[i] -= alpha * q[i];
                rmax = std::max(rmax, std::fabs(r[i]));
            }
            if (rmax <= 1e-13 * bmax) return;
            applyMIC0(S, nc, r.data(), z.data());
            double rhoNew = 0.0;
            for (int i = 0; i < m; i++) rhoNew += z[i] * r[i];
            const double beta = rhoNew / rho;
            for (int i = 0; i < m; i++) s[i] = z[i] + beta * s[i];
            rho = rhoNew;
        }
        throw std::runtime_error("rollout adjoint: the FP64 CG did not reach 1e-13");
    }

    // P2G's adjoint for the particles x (velocities v): adds to vOut the
    // velocity cotangent sum_f (w_pf / W_f or w_pf) ubar_f and to xOut the
    // position cotangent sum_f ubar_f ((v_p - u_f) / W_f or v_p) w'(d2) (-2)(f - x_p)
    // The P2G derivative at the forward's own transfer: each particle-face
    // weight is the float weight the forward summed (p2gFlip's expression in
    // the face block's frame), and W_f, u_f the float weight sum and value the
    // forward divided by and stored.  Near-empty faces (W_f a few 1e-5) take
    // weights from float expressions with heavy cancellation; the derivative
    // is that of the map the forward actually evaluated, in double.
    void p2gAdjoint(int k, const std::vector<V3> &x, const std::vector<V3> &v, const std::vector<double> *ub,
                    std::vector<double> &xOut, std::vector<double> &vOut) {
        const int ni = sim.ni, nj = sim.nj, nk = sim.nk, np = (int)x.size();
        const double dx = sim.dx, h = 0.5 * dx;
        const Kernel64 Kw(sim.liquidRadius());
        const double r = std::sqrt(Kw.rsq);
        const int bw = 10;
        const float rf = (float)sim.liquidRadius();
        const float rsq = rf * rf;
        const float coef1 = (4.0f / 9.0f) * (1.0f / (rf * rf * rf * rf * rf * rf));
        const float coef2 = (17.0f / 9.0f) * (1.0f / (rf * rf * rf * rf));
        const float coef3 = (22.0f / 9.0f) * (1.0f / (rf * rf));
        const unsigned char *rv = recValid[k].data();
        std::vector<float> pos(3 * (size_t)np), vel(3 * (size_t)np);
        for (int p = 0; p < np; p++) {
            pos[3 * (size_t)p] = x[p].x; pos[3 * (size_t)p + 1] = x[p].y; pos[3 * (size_t)p + 2] = x[p].z;
            vel[3 * (size_t)p] = v[p].x; vel[3 * (size_t)p + 1] = v[p].y; vel[3 * (size_t)p + 2] = v[p].z;
        }
        for (int d = 0; d < 3; d++) {
            const FaceDims D = faceDims(ni, nj, nk, d);
            const double off[3] = {d == 0 ? 0.0 : h, d == 1 ? 0.0 : h, d == 2 ? 0.0 : h};
            const V3 offset = mk(off[0], off[1], off[2]);
            // the forward's float weight sums and values
            std::vector<float> Wf(D.n()), uf(D.n());
            std::vector<unsigned char> vf(D.n());
            p2gFlip(ni, nj, nk, dx, np, pos.data(), vel.data(), sim.liquidRadius(), d, uf.data(), vf.data(),
                    Wf.data());
            for (int p = 0; p < np; p++) {
                const double vd = d == 0 ? v[p].x : (d == 1 ? v[p].y : v[p].z);
                const double ps[3] = {(double)x[p].x - off[0], (double)x[p].y - off[1], (double)x[p].z - off[2]};
                const V3 xs = x[p] - offset;
                int lo[3], hi[3];
                p2gFaceRange(D, dx, r, ps, lo, hi);
                double vbAcc = 0.0, xbAcc[3] = {0.0, 0.0, 0.0};
                for (int fk = lo[2]; fk <= hi[2]; fk++)
                    for (int fj = lo[1]; fj <= hi[1]; fj++)
                        for (int fi = lo[0]; fi <= hi[0]; fi++) {
                            // the forward's frame: the face's 10^3 block
                            const V3 bo = indexToPos(fi / bw, fj / bw, fk / bw, bw * dx);
                            const V3 pl = xs - bo;
                            const V3 e = indexToPos(fi - fi / bw * bw, fj - fj / bw * bw, fk - fk / bw * bw, dx) - pl;
                            const float d2 = dot(e, e);
                            if (!(d2 < rsq)) continue;
                            const float wgt = 1.0f - coef1 * d2 * d2 * d2 + coef2 * d2 * d2 - coef3 * d2;
                            const size_t f = flat(fi, fj, fk, D.wi, D.wj);
                            const bool norm = rv[f] != 0 && Wf[f] > 0.0f;
                            const double W = Wf[f];
                            vbAcc += (norm ? (double)wgt / W : (double)wgt) * ub[d][f];
                            const double sl = (norm ? (vd - (double)uf[f]) / W : vd) * ub[d][f] * Kw.dw((double)d2) *
                                              -2.0;
                            xbAcc[0] += sl * (double)e.x;
                            xbAcc[1] += sl * (double)e.y;
                            xbAcc[2] += sl * (double)e.z;
                        }
                vOut[3 * (size_t)p + d] += vbAcc;
                for (int c = 0; c < 3; c++) xOut[3 * (size_t)p + c] += xbAcc[c];
            }
            rv += D.n();
        }
    }

    void backward(const double *rhoBar, double *gu, double *gv, double *gw) {
        if (!recorded) throw std::runtime_error("rollout: backward needs a recording forward");
        const int ni = sim.ni, nj = sim.nj, nk = sim.nk, np = (int)pos0.size();
        const double dx = sim.dx, h = 0.5 * dx;
        // the density of the final state: xbar_p = sum_c rhobar_c w'(d2) (-2) (c - x_p)
        std::vector<double> xb(3 * (size_t)np, 0.0), vb(3 * (size_t)np, 0.0);
        const Kernel64 Kw(sim.liquidRadius());
        for (int p = 0; p < np; p++) {
            const V3 q = P[K][p];
            const GI c0 = posToIndex(q, dx);
            for (int dk = -1; dk <= 1; dk++)
                for (int dj = -1; dj <= 1; dj++)
                    for (int di = -1; di <= 1; di++) {
                        const int i = c0.i + di, j = c0.j + dj, k = c0.k + dk;
                        if (!inRange(i, j, k, ni, nj, nk)) continue;
                        const double e[3] = {i * dx + h - q.x, j * dx + h - q.y, k * dx + h - q.z};
                        const double d2 = e[0] * e[0] + e[1] * e[1] + e[2] * e[2];
                        if (!(d2 < Kw.rsq)) continue;
                        const double s = rhoBar[flat(i, j, k, ni, nj)] * Kw.dw(d2) * -2.0;
                        for (int c = 0; c < 3; c++) xb[3 * (size_t)p + c] += s * e[c];
                    }
        }
        for (int k = K - 1; k >= 0; k--) adjointSubstep(k, xb, vb);
        // the initial velocities are the trilinear sample of the initial field
        std::vector<double> g[3];
        for (int d = 0; d < 3; d++) g[d].assign(nFaces(d), 0.0);
        for (int p = 0; p < np; p++) sampleAdjointAdd(ni, nj, nk, dx, pos0[p], &vb[3 * (size_t)p], g);
        std::copy(g[0].begin(), g[0].end(), gu);
        std::copy(g[1].begin(), g[1].end(), gv);
        std::copy(g[2].begin(), g[2].end(), gw);
    }
};

OracleRollout *oracle_rollout_new(int ni, int nj, int nk, double dx, int np, const float *pos,
                                  const float *solid_nodes, double dt, int substeps, double picflip, double gx,
                                  double gy, double gz, double cfl, double tol, int max_iter) {
    OracleRollout *r = new OracleRollout();
    OracleSim &s = r->sim;
    s.ni = ni;
    s.nj = nj;
    s.nk = nk;
    s.dx = dx;
    s.picflip = picflip;
    s.cfl = cfl;
    s.ptol = tol;
    s.maxPIters = max_iter;
    s.removal = false;
    s.bodyForces.push_back(mk(gx, gy, gz));
    if (solid_nodes) oracle_sim_set_solid_phi(&s, solid_nodes);
    s.initialize();
    r->dt = dt;
    r->K = substeps;
    for (int p = 0; p < np; p++) r->pos0.push_back(V3{pos[3 * p], pos[3 * p + 1], pos[3 * p + 2]});
    return r;
}
void oracle_rollout_destroy(OracleRollout *r) { delete r; }
int oracle_rollout_forward(OracleRollout *r, const float *u0, const float *v0, const float *w0, int freeze,
                           float *density) {
    try {
        r->forward(u0, v0, w0, freeze != 0);
    } catch (const std::exception &) {
        return -1;
    }
    std::copy(r->rho.begin(), r->rho.end(), density);
    return 0;
}
int oracle_rollout_backward(OracleRollout *r, const double *density_bar, double *gu, double *gv, double *gw) {
    try {
        r->backward(density_bar, gu, gv, gw);
    } catch (const std::exception &) {
        return -1;
    }
    return 0;
}
void oracle_rollout_positions(OracleRollout *r, float *pos) {
    const std::vector<V3> &x = r->P.back();
    for (size_t p = 0; p < x.size(); p++) {
        pos[3 * p] = x[p].x;
        pos[3 * p + 1] = x[p].y;
        pos[3 * p + 2] = x[p].z;
    }
}
long long oracle_rollout_iterations(OracleRollout *r) { return r->solveIterations; }
