// ffdevice.h -- device-side arithmetic shared by the HIP kernels of the FLIP
// substep.  Every helper keeps the float/double promotion points of the engine
// it replaces (rlguy/Blender-FLIP-Fluids 1.8.5, src/engine/): vmath::vec3 is
// three floats with component-wise float arithmetic (vmath.h:37-91), grid
// index/position conversions are grid3d.h:32-124.  All HIP sources are
// compiled with -ffp-contract=off and IEEE-rounded division and square root,
// so the same inputs give the same bits as the x86-64 (SSE2, no FMA) engine.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define FF_HD __host__ __device__ __forceinline__

namespace ff {

struct V3 {
    float x, y, z;
};
FF_HD V3 mk(double x, double y, double z) { return V3{(float)x, (float)y, (float)z}; }
FF_HD V3 add(V3 a, V3 b) { return V3{a.x + b.x, a.y + b.y, a.z + b.z}; }
FF_HD V3 sub(V3 a, V3 b) { return V3{a.x - b.x, a.y - b.y, a.z - b.z}; }
FF_HD V3 scale(float s, V3 a) { return V3{s * a.x, s * a.y, s * a.z}; }
FF_HD float dot3(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
FF_HD float len3(V3 a) { return sqrtf(a.x * a.x + a.y * a.y + a.z * a.z); }
// vmath::operator/(vec3, float): multiply by the float reciprocal of a double division
FF_HD V3 divide3(V3 a, float s) {
    float inv = (float)(1.0 / (double)s);
    return V3{a.x * inv, a.y * inv, a.z * inv};
}
FF_HD V3 normalize3(V3 a) { return divide3(a, len3(a)); }

struct I3 {
    int i, j, k;
};
FF_HD I3 posToIndex(V3 p, double dx) {
    double inv = 1.0 / dx;
    return I3{(int)floor((double)p.x * inv), (int)floor((double)p.y * inv),
              (int)floor((double)p.z * inv)};
}
FF_HD I3 posToIndexD(double x, double y, double z, double dx) {
    double inv = 1.0 / dx;
    return I3{(int)floor(x * inv), (int)floor(y * inv), (int)floor(z * inv)};
}
FF_HD V3 indexToPos(int i, int j, int k, double dx) {
    return mk((double)(float)i * dx, (double)(float)j * dx, (double)(float)k * dx);
}
FF_HD V3 indexToCellCenter(int i, int j, int k, double dx) {
    double hw = 0.5 * dx;
    return mk((double)(float)i * dx + hw, (double)(float)j * dx + hw, (double)(float)k * dx + hw);
}
FF_HD bool inRange(int i, int j, int k, int ni, int nj, int nk) {
    return i >= 0 && j >= 0 && k >= 0 && i < ni && j < nj && k < nk;
}
FF_HD size_t flat(int i, int j, int k, int ni, int nj) {
    return (size_t)i + (size_t)ni * ((size_t)j + (size_t)nj * (size_t)k);
}

// AABB with float position and double extents (aabb.cpp:32-38, 122-133, 489-518)
struct Box {
    V3 pos;
    double w, h, d;
    FF_HD void expand(double v) {
        double hh = 0.5 * v;
        pos = sub(pos, mk(hh, hh, hh));
        w += v;
        h += v;
        d += v;
    }
    FF_HD bool inside(V3 p) const {
        return p.x >= pos.x && p.y >= pos.y && p.z >= pos.z && (double)p.x < (double)pos.x + w &&
               (double)p.y < (double)pos.y + h && (double)p.z < (double)pos.z + d;
    }
    FF_HD V3 maxPoint() const { return add(pos, mk(w, h, d)); }
    FF_HD V3 nearestInside(V3 p) const {
        if (inside(p)) return p;
        double eps = 1e-6;
        V3 mn = pos, mx = maxPoint();
        p.x = (float)fmax((double)p.x, (double)mn.x);
        p.y = (float)fmax((double)p.y, (double)mn.y);
        p.z = (float)fmax((double)p.z, (double)mn.z);
        p.x = (float)fmin((double)p.x, (double)mx.x - eps);
        p.y = (float)fmin((double)p.y, (double)mx.y - eps);
        p.z = (float)fmin((double)p.z, (double)mx.z - eps);
        return p;
    }
};

FF_HD double clampd(double v, double lo, double hi) {
    if (v < lo) v = lo;
    if (v > hi) v = hi;
    return v;
}

// Collision::findClosestPointOnTriangle (collision.cpp:161-268)
FF_HD V3 closestPointOnTriangle(V3 p0, V3 v0, V3 v1, V3 v2) {
    V3 edge0 = sub(v1, v0), edge1 = sub(v2, v0), pv = sub(v0, p0);
    double a = dot3(edge0, edge0), b = dot3(edge0, edge1), c = dot3(edge1, edge1);
    double d = dot3(edge0, pv), e = dot3(edge1, pv);
    double det = a * c - b * b;
    if (fabs(det) < 1e-9) {
        float feps = 1e-6f;
        float len1 = len3(sub(v0, v1)), len2 = len3(sub(v1, v2)), len3v = len3(sub(v2, v0));
        if (len1 < feps && len2 < feps && len3v < feps) return v0;
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
        float dotnum = dot3(sub(p0, dp), sub(dq, dp));
        float dotdem = dot3(sub(dq, dp), sub(dq, dp));
        if (fabsf(dotdem) < feps) return v0;
        float lambda = dotnum / dotdem;
        if (lambda <= 0.0f) return dp;
        if (lambda >= 1.0f) return dq;
        return add(dp, scale(lambda, sub(dq, dp)));
    }
    double s = b * e - c * d, t = b * d - a * e;
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
    return add(add(v0, scale((float)s, edge0)), scale((float)t, edge1));
}

// FluidSimulation::_getBoundaryAABB (fluidsimulation.cpp:5175-5180)
// LevelsetUtils::fractionInside (levelsetutils.cpp:39-141): the fraction of
// a segment, or of a square with the four corner values, below zero
FF_HD float fraction2(float l, float r) {
    if (l < 0 && r < 0) return 1;
    if (l < 0 && r >= 0) return l / (l - r);
    if (l >= 0 && r < 0) return r / (r - l);
    return 0;
}
FF_HD void cycle4(float *a) {
    float t = a[0];
    a[0] = a[1];
    a[1] = a[2];
    a[2] = a[3];
    a[3] = t;
}
FF_HD float fraction4(float bl, float br, float tl, float tr) {
    int in = (bl < 0 ? 1 : 0) + (tl < 0 ? 1 : 0) + (br < 0 ? 1 : 0) + (tr < 0 ? 1 : 0);
    float L[4] = {bl, br, tr, tl};
    if (in == 4) return 1;
    if (in == 3) {
        while (L[0] < 0) cycle4(L);
        float s0 = 1 - fraction2(L[0], L[3]);
        float s1 = 1 - fraction2(L[0], L[1]);
        return 1.0f - 0.5f * s0 * s1;
    }
    if (in == 2) {
        while (L[0] >= 0 || !(L[1] < 0 || L[2] < 0)) cycle4(L);
        if (L[1] < 0) {
            float sl = fraction2(L[0], L[3]);
            float sr = fraction2(L[1], L[2]);
            return 0.5f * (sl + sr);
        }
        float mid = 0.25f * (L[0] + L[1] + L[2] + L[3]);
        if (mid < 0) {
            float area = 0;
            float s1 = 1 - fraction2(L[0], L[3]);
            float s3 = 1 - fraction2(L[2], L[3]);
            area += 0.5f * s1 * s3;
            float s2 = 1 - fraction2(L[2], L[1]);
            float s0 = 1 - fraction2(L[0], L[1]);
            area += 0.5f * s0 * s2;
            return 1.0f - area;
        }
        float area = 0;
        float s0 = fraction2(L[0], L[1]);
        float s1 = fraction2(L[0], L[3]);
        area += 0.5f * s0 * s1;
        float s2 = fraction2(L[2], L[1]);
        float s3 = fraction2(L[2], L[3]);
        area += 0.5f * s2 * s3;
        return area;
    }
    if (in == 1) {
        while (L[0] >= 0) cycle4(L);
        float s0 = fraction2(L[0], L[3]);
        float s1 = fraction2(L[0], L[1]);
        return 0.5f * s0 * s1;
    }
    return 0;
}

FF_HD Box boundaryBox(int ni, int nj, int nk, double dx) {
    Box b{V3{0.0f, 0.0f, 0.0f}, ni * dx, nj * dx, nk * dx};
    b.expand(-3 * dx - 1e-4);
    return b;
}

// Order-preserving map of a float onto an unsigned key, so that an unsigned
// min over keys is the float min (used for deterministic LDS min-reductions).
FF_HD uint32_t floatKey(float f) {
    uint32_t u = __builtin_bit_cast(uint32_t, f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
FF_HD float keyFloat(uint32_t k) {
    uint32_t u = (k & 0x80000000u) ? (k & 0x7fffffffu) : ~k;
    return __builtin_bit_cast(float, u);
}

// Interpolation::trilinearInterpolate(double p[8], x, y, z) (interpolation.cpp:61-70)
FF_HD double trilin8(const double *p, double x, double y, double z) {
    return p[0] * (1 - x) * (1 - y) * (1 - z) + p[1] * x * (1 - y) * (1 - z) +
           p[2] * (1 - x) * y * (1 - z) + p[3] * (1 - x) * (1 - y) * z + p[4] * x * (1 - y) * z +
           p[5] * (1 - x) * y * z + p[6] * x * y * (1 - z) + p[7] * x * y * z;
}

// Grid shape of one simulation (cells ni x nj x nk, cell size dx).
struct Grid {
    int ni, nj, nk;
    double dx;
};

// The FLIP transfer's frame of face (fi, fj, fk) of direction dir
// (velocityadvector.cpp:177-188, 467-536): particles are shifted by the
// direction's half-cell offset and taken relative to the face's 10^3 block,
// in which the face sits at gpos.
constexpr int P2G_BLOCK = 10;
struct FaceFrame {
    V3 offd, offb, gpos;
};
FF_HD FaceFrame p2gFrame(Grid g, int dir, int fi, int fj, int fk) {
    const double h = 0.5 * g.dx;
    const int bi = fi / P2G_BLOCK, bj = fj / P2G_BLOCK, bk = fk / P2G_BLOCK;
    return FaceFrame{dir == 0 ? mk(0.0, h, h) : (dir == 1 ? mk(h, 0.0, h) : mk(h, h, 0.0)),
                     indexToPos(bi, bj, bk, P2G_BLOCK * g.dx),
                     indexToPos(fi - bi * P2G_BLOCK, fj - bj * P2G_BLOCK, fk - bk * P2G_BLOCK, g.dx)};
}
// squared distance of a particle from the face in that frame
FF_HD float p2gDist2(const FaceFrame &F, float x, float y, float z) {
    V3 p = sub(sub(V3{x, y, z}, F.offd), F.offb);
    V3 v = sub(F.gpos, p);
    return dot3(v, v);
}
// the transfer kernel's weight at squared distance d2 < r^2
struct P2GWeight {
    float rsq, coef1, coef2, coef3;
};
FF_HD P2GWeight p2gWeight(double radius) {
    const float r = (float)radius;
    return P2GWeight{r * r, (4.0f / 9.0f) * (1.0f / (r * r * r * r * r * r)), (17.0f / 9.0f) * (1.0f / (r * r * r * r)),
                     (22.0f / 9.0f) * (1.0f / (r * r))};
}
FF_HD float p2gWeightAt(const P2GWeight &W, float d2) {
    return 1.0f - W.coef1 * d2 * d2 * d2 + W.coef2 * d2 * d2 - W.coef3 * d2;
}

// ---- volume fractions of level sets ----
// LevelsetUtils::volumeFraction of a tetrahedron (levelsetutils.cpp:213-258)
__device__ inline void sort4(float &a, float &b, float &c, float &d) {
    float t;
    if (a > b) { t = a; a = b; b = t; }
    if (c > d) { t = c; c = d; d = t; }
    if (a > c) { t = a; a = c; c = t; }
    if (b > d) { t = b; b = d; d = t; }
    if (b > c) { t = b; b = c; c = t; }
}
__device__ inline float tetFraction(float p0, float p1, float p2, float p3) {
    return p0 * p0 * p0 / ((p0 - p1) * (p0 - p2) * (p0 - p3));
}
__device__ inline float prismFraction(float p0, float p1, float p2, float p3) {
    const float a = p0 / (p0 - p2), b = p0 / (p0 - p3), c = p1 / (p1 - p3), d = p1 / (p1 - p2);
    return a * b * (1 - d) + b * (1 - c) * d + c * d;
}
__device__ inline float volumeFraction4(float p0, float p1, float p2, float p3) {
    sort4(p0, p1, p2, p3);
    if (p3 <= 0) return 1;
    if (p2 <= 0) return 1 - tetFraction(p3, p2, p1, p0);
    if (p1 <= 0) return prismFraction(p0, p1, p2, p3);
    if (p0 <= 0) return tetFraction(p0, p1, p2, p3);
    return 0;
}
// the cube's ten tetrahedra, the two central ones counted twice, / 12
// (levelsetutils.h:60-100)
__device__ inline float volumeFraction8(float p000, float p100, float p010, float p110, float p001, float p101, float p011,
                                 float p111) {
    return (volumeFraction4(p000, p001, p101, p011) + volumeFraction4(p000, p101, p100, p110) +
            volumeFraction4(p000, p010, p011, p110) + volumeFraction4(p101, p011, p111, p110) +
            2 * volumeFraction4(p000, p011, p101, p110) + volumeFraction4(p100, p101, p001, p111) +
            volumeFraction4(p100, p001, p000, p010) + volumeFraction4(p100, p110, p111, p010) +
            volumeFraction4(p001, p111, p011, p010) + 2 * volumeFraction4(p100, p111, p001, p010)) /
           12.0f;
}
// Interpolation::trilinearInterpolate (interpolation.cpp:72-112) of a float
// grid of pi x pj x pk samples at spacing dx whose sample (0,0,0) sits at the
// origin; samples outside the grid count as 0
__device__ inline double gridTrilinearD(const float *a, int pi, int pj, int pk, double dx, V3 p);
__device__ inline float gridTrilinear(const float *a, int pi, int pj, int pk, double dx, V3 p) {
    return (float)gridTrilinearD(a, pi, pj, pk, dx, p);
}
// the same, keeping the double sum
__device__ inline double gridTrilinearD(const float *a, int pi, int pj, int pk, double dx, V3 p) {
    const I3 c = posToIndex(p, dx);
    const V3 gp = indexToPos(c.i, c.j, c.k, dx);
    const double inv_dx = 1.0 / dx;
    const double ix = (double)(p.x - gp.x) * inv_dx, iy = (double)(p.y - gp.y) * inv_dx,
                 iz = (double)(p.z - gp.z) * inv_dx;
    double pts[8];
    const int o[8][3] = {{0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1}, {1, 0, 1}, {0, 1, 1}, {1, 1, 0}, {1, 1, 1}};
#pragma unroll
    for (int q = 0; q < 8; q++) {
        const int x = c.i + o[q][0], y = c.j + o[q][1], z = c.k + o[q][2];
        pts[q] = inRange(x, y, z, pi, pj, pk) ? (double)a[flat(x, y, z, pi, pj)] : 0.0;
    }
    return trilin8(pts, ix, iy, iz);
}

// ---- surface tension ----
// The liquid surface's curvature grid and cluster status for the pressure
// solve's tension terms; status null: no surface tension.
struct TensionView {
    const int8_t *status = nullptr;    // per cell; bits 4 and 8 mark BAD clusters
    const float *curvature = nullptr;  // per cell
    const float *phi = nullptr;        // the liquid SDF, cell-centred
    double k = 0.0;                    // the surface tension constant
};
// PressureSolver::_getSurfaceTensionTerm (pressuresolver.cpp:715-749): k times
// the curvature, trilinearly interpolated (as a double) at the surface's
// crossing between the two cells' centres; 0 when both cells are on one
// side or one of them is in a BAD cluster
__device__ inline double tensionTerm(Grid g, const TensionView &T, int i1, int j1, int k1, int i2, int j2,
                                     int k2) {
    const size_t c1 = flat(i1, j1, k1, g.ni, g.nj), c2 = flat(i2, j2, k2, g.ni, g.nj);
    const int8_t s = T.status[c1] | T.status[c2];
    if ((s & 4) || (s & 8)) return 0.0;
    const float phi1 = T.phi[c1], phi2 = T.phi[c2];
    if ((phi1 < 0.0f && phi2 < 0.0f) || (phi1 >= 0.0f && phi2 >= 0.0f)) return 0.0;
    const V3 p1 = indexToCellCenter(i1, j1, k1, g.dx), p2 = indexToCellCenter(i2, j2, k2, g.dx);
    const float eps = 1e-6f;
    const V3 p = fabsf(phi2 - phi1) < eps ? add(p1, scale(0.5f, sub(p2, p1)))
                                         : add(p1, scale(phi1 / (phi1 - phi2), sub(p2, p1)));
    const V3 hdx = mk(0.5 * g.dx, 0.5 * g.dx, 0.5 * g.dx);
    return T.k * gridTrilinearD(T.curvature, g.ni, g.nj, g.nk, g.dx, sub(p, hdx));
}

}  // namespace ff
