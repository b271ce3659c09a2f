// viscosity.hip -- the viscosity solve of the FLIP substep on the GPU.
//
// ViscositySolver::applyViscosityToVelocityField (viscositysolver.cpp:50-79)
// as a chain of grid-parallel kernels:
//
//   face states      one thread per u/v/w face: the solid SDF's cell-centre
//                    values (the mean of the cell's eight nodes,
//                    meshlevelset.cpp:152-162) of the two cells beside the
//                    face; domain-border faces and faces whose two centre
//                    values sum to <= 0 are solid (viscositysolver.cpp:142-181)
//   valid cells      liquid cells on an (n+1)^3 grid, dilated twice through
//                    the six face neighbours (209-240)
//   subcells         the liquid's volume fraction in each of the (2n)^3
//                    subcells of a valid cell, from the liquid SDF
//                    trilinearly interpolated at the subcell's corners
//                    (cell-centred samples, particlelevelset.cpp:136-138) with
//                    the ten-tetrahedron split (314-348)
//   volumes          the seven control-volume grids, each the mean of the
//                    eight subcells around its sample point (350-377)
//   matrix index     faces that are fluid with some volume around them,
//                    numbered u faces first, then v, then w (383-473): a
//                    flag per face and a prefix sum
//   assembly         one thread per matrix face: its 15-entry row of the
//                    viscous operator and right-hand side (559-830), the
//                    entries sorted by column into a 16-slot ELL block
//   solve            CG with the inverse diagonal as preconditioner, float
//                    vectors, FP64 reductions in a fixed order (the same bits
//                    every run), stopped by the reference's rule
//                    (pcgsolver.h:268-300, viscositysolver.cpp:832-868)
//                    checked on the device, the host reading it once per
//                    batch of iterations
//   apply            on success the field is cleared and the matrix faces
//                    take the solution (870-905)
//
// Every step up to the assembled system computes the reference's float and
// double operations in its order, so states, volumes, index, matrix and
// right-hand side are the reference's bit for bit.  The solve reaches the
// reference's stopping rule with its own preconditioner, so its solution
// agrees with the reference's to the solver tolerance.
#include "viscosity.h"

#include <algorithm>
#include <cmath>
#include <stdexcept>

namespace ff {

void hipCheck(hipError_t e, const char *what);

namespace {

constexpr int TPB = 256;
constexpr int ELL = 16;    // slots per row: 15 entries at most, one padding
constexpr int BATCH = 8;   // CG iterations per host check of the stopping rule
constexpr int RED_BLOCKS = 1024;
enum : unsigned char { VF_AIR = 0, VF_FLUID = 1, VF_SOLID = 2 };
enum { S_RHO = 0, S_ALPHA, S_BETA, S_RESID, S_DONE, S_ITERS, S_NAN, S_COUNT = 8 };

unsigned blocksFor(size_t n, unsigned t) { return (unsigned)((n + t - 1) / t); }

struct Dims {
    int ni, nj, nk;
    size_t nU, nV, nW;
};
Dims dimsOf(Grid g) {
    return Dims{g.ni, g.nj, g.nk, (size_t)(g.ni + 1) * g.nj * g.nk, (size_t)g.ni * (g.nj + 1) * g.nk,
                (size_t)g.ni * g.nj * (g.nk + 1)};
}
// face q of the concatenated u, v, w faces -> direction and indices
__device__ inline int decodeFace(const Dims &D, size_t q, int &i, int &j, int &k) {
    int dir, wi, wj;
    if (q < D.nU) {
        dir = 0, wi = D.ni + 1, wj = D.nj;
    } else if (q < D.nU + D.nV) {
        q -= D.nU;
        dir = 1, wi = D.ni, wj = D.nj + 1;
    } else {
        q -= D.nU + D.nV;
        dir = 2, wi = D.ni, wj = D.nj;
    }
    i = (int)(q % wi);
    j = (int)((q / wi) % wj);
    k = (int)(q / ((size_t)wi * wj));
    return dir;
}

// ---------------------------------------------------------- face states ---
__device__ inline float centerPhi(const float *S, int i, int j, int k, int pi, int pj) {
    auto P = [&](int a, int b, int c) { return S[flat(i + a, j + b, k + c, pi, pj)]; };
    return 0.125f * (P(0, 0, 0) + P(1, 0, 0) + P(0, 1, 0) + P(1, 1, 0) + P(0, 0, 1) + P(1, 0, 1) + P(0, 1, 1) +
                     P(1, 1, 1));
}
__global__ void stateKernel(Dims D, const float *solid, unsigned char *state) {
    const size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= D.nU + D.nV + D.nW) return;
    int i, j, k;
    const int dir = decodeFace(D, q, i, j, k);
    const int pi = D.ni + 1, pj = D.nj + 1;
    bool edge;
    int a = i, b = j, c = k;  // the cell on the low side
    if (dir == 0) {
        edge = i == 0 || i == D.ni;
        a = i - 1;
    } else if (dir == 1) {
        edge = j == 0 || j == D.nj;
        b = j - 1;
    } else {
        edge = k == 0 || k == D.nk;
        c = k - 1;
    }
    bool solid_ = edge;
    if (!edge) solid_ = centerPhi(solid, a, b, c, pi, pj) + centerPhi(solid, i, j, k, pi, pj) <= 0;
    state[q] = solid_ ? VF_SOLID : VF_FLUID;
}

// ---------------------------------------------------------- valid cells ---
__global__ void validInitKernel(Grid g, const float *liquid, unsigned char *valid) {
    const int vi = g.ni + 1, vj = g.nj + 1, vk = g.nk + 1;
    const size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= (size_t)vi * vj * vk) return;
    const int i = (int)(c % vi), j = (int)((c / vi) % vj), k = (int)(c / ((size_t)vi * vj));
    valid[c] = (i < g.ni && j < g.nj && k < g.nk && liquid[flat(i, j, k, g.ni, g.nj)] < 0) ? 1 : 0;
}
__global__ void dilateKernel(Grid g, const unsigned char *in, unsigned char *out) {
    const int vi = g.ni + 1, vj = g.nj + 1, vk = g.nk + 1;
    const size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= (size_t)vi * vj * vk) return;
    const int i = (int)(c % vi), j = (int)((c / vi) % vj), k = (int)(c / ((size_t)vi * vj));
    unsigned char v = in[c];
    if (!v) {
        const int nb[6][3] = {{-1, 0, 0}, {1, 0, 0}, {0, -1, 0}, {0, 1, 0}, {0, 0, -1}, {0, 0, 1}};
        for (int q = 0; q < 6 && !v; q++) {
            const int a = i + nb[q][0], b = j + nb[q][1], e = k + nb[q][2];
            if (inRange(a, b, e, vi, vj, vk) && in[flat(a, b, e, vi, vj)]) v = 1;
        }
    }
    out[c] = v;
}

// ------------------------------------------------------------- subcells ---
__global__ void subcellKernel(Grid g, const float *liquid, const unsigned char *valid, float *frac) {
    const int si = 2 * g.ni, sj = 2 * g.nj, sk = 2 * g.nk;
    const size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= (size_t)si * sj * sk) return;
    const int i = (int)(c % si), j = (int)((c / si) % sj), k = (int)(c / ((size_t)si * sj));
    if (!valid[flat(i / 2, j / 2, k / 2, g.ni + 1, g.nj + 1)]) {
        frac[c] = 0.0f;
        return;
    }
    // the solver's float cell width; the level set's own width is double
    const float fdx = (float)g.dx;
    const float sdx = 0.5f * fdx;
    const V3 start = V3{0.25f * fdx, 0.25f * fdx, 0.25f * fdx};
    const V3 half = mk(0.5 * g.dx, 0.5 * g.dx, 0.5 * g.dx);
    const V3 ctr = add(start, V3{i * sdx, j * sdx, k * sdx});
    const float h = 0.5f * sdx;
    auto at = [&](float a, float b, float e) {
        return gridTrilinear(liquid, g.ni, g.nj, g.nk, g.dx, sub(add(ctr, V3{a, b, e}), half));
    };
    const float p000 = at(-h, -h, -h), p001 = at(-h, -h, +h), p010 = at(-h, +h, -h), p011 = at(-h, +h, +h);
    const float p100 = at(+h, -h, -h), p101 = at(+h, -h, +h), p110 = at(+h, +h, -h), p111 = at(+h, +h, +h);
    frac[c] = volumeFraction8(p000, p100, p010, p110, p001, p101, p011, p111);
}

// -------------------------------------------------------------- volumes ---
struct VolumeGrids {
    float *g[7];
};
__global__ void volumeKernel(Grid g, const unsigned char *valid, const float *sub, VolumeGrids V) {
    const int ni = g.ni, nj = g.nj, nk = g.nk;
    const int vi = ni + 1, vj = nj + 1, vk = nk + 1;
    const size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= (size_t)vi * vj * vk) return;
    const int i = (int)(c % vi), j = (int)((c / vi) % vj), k = (int)(c / ((size_t)vi * vj));
    // dims and subcell offsets of center, U, V, W, edgeU, edgeV, edgeW
    const int dim[7][3] = {{ni, nj, nk},         {ni + 1, nj, nk},     {ni, nj + 1, nk},    {ni, nj, nk + 1},
                           {ni, nj + 1, nk + 1}, {ni + 1, nj, nk + 1}, {ni + 1, nj + 1, nk}};
    const int off[7][3] = {{0, 0, 0}, {-1, 0, 0}, {0, -1, 0}, {0, 0, -1}, {0, -1, -1}, {-1, 0, -1}, {-1, -1, 0}};
    const bool inner = i >= 1 && j >= 1 && k >= 1 && i < ni && j < nj && k < nk && valid[c];
    const int si = 2 * ni, sj = 2 * nj;
#pragma unroll
    for (int q = 0; q < 7; q++) {
        if (i >= dim[q][0] || j >= dim[q][1] || k >= dim[q][2]) continue;
        float val = 0.0f;
        if (inner) {
            float acc = 0.0f;
            for (int e = 0; e < 2; e++)
                for (int b = 0; b < 2; b++)
                    for (int a = 0; a < 2; a++)
                        acc += sub[flat(2 * i + off[q][0] + a, 2 * j + off[q][1] + b, 2 * k + off[q][2] + e, si, sj)];
            val = 0.125f * acc;
        }
        V.g[q][flat(i, j, k, dim[q][0], dim[q][1])] = val;
    }
}

// accessors of the seven grids
struct VolView {
    const float *center, *U, *V, *W, *eU, *eV, *eW;
    int ni, nj, nk;
    __device__ float CE(int i, int j, int k) const { return center[flat(i, j, k, ni, nj)]; }
    __device__ float VU(int i, int j, int k) const { return U[flat(i, j, k, ni + 1, nj)]; }
    __device__ float VV(int i, int j, int k) const { return V[flat(i, j, k, ni, nj + 1)]; }
    __device__ float VW(int i, int j, int k) const { return W[flat(i, j, k, ni, nj)]; }
    __device__ float EU(int i, int j, int k) const { return eU[flat(i, j, k, ni, nj + 1)]; }
    __device__ float EV(int i, int j, int k) const { return eV[flat(i, j, k, ni + 1, nj)]; }
    __device__ float EW(int i, int j, int k) const { return eW[flat(i, j, k, ni + 1, nj + 1)]; }
};
VolView volView(Grid g, float *const *vol) {
    return VolView{vol[0], vol[1], vol[2], vol[3], vol[4], vol[5], vol[6], g.ni, g.nj, g.nk};
}

// --------------------------------------------------------- matrix index ---
__global__ void flagKernel(Dims D, const unsigned char *state, VolView Vo, int *flag) {
    const size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= D.nU + D.nV + D.nW) return;
    int i, j, k;
    const int dir = decodeFace(D, q, i, j, k);
    int f = 0;
    if (i >= 1 && j >= 1 && k >= 1 && i < D.ni && j < D.nj && k < D.nk && state[q] == VF_FLUID) {
        if (dir == 0)
            f = Vo.VU(i, j, k) > 0.0 || Vo.CE(i, j, k) > 0.0 || Vo.CE(i - 1, j, k) > 0.0 || Vo.EW(i, j + 1, k) > 0.0 ||
                Vo.EW(i, j, k) > 0.0 || Vo.EV(i, j, k + 1) > 0.0 || Vo.EV(i, j, k) > 0.0;
        else if (dir == 1)
            f = Vo.VV(i, j, k) > 0.0 || Vo.EW(i + 1, j, k) > 0.0 || Vo.EW(i, j, k) > 0.0 || Vo.CE(i, j, k) > 0.0 ||
                Vo.CE(i, j - 1, k) > 0.0 || Vo.EU(i, j, k + 1) > 0.0 || Vo.EU(i, j, k) > 0.0;
        else
            f = Vo.VW(i, j, k) > 0.0 || Vo.EV(i + 1, j, k) > 0.0 || Vo.EV(i, j, k) > 0.0 || Vo.EU(i, j + 1, k) > 0.0 ||
                Vo.EU(i, j, k) > 0.0 || Vo.CE(i, j, k) > 0.0 || Vo.CE(i, j, k - 1) > 0.0;
    }
    flag[q] = f;
}
__global__ void tableKernel(size_t F, const int *flag, const int *rowOf, int *table, int *rowFace) {
    const size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= F) return;
    const int r = flag[q] ? rowOf[q] : -1;
    table[q] = r;
    if (r >= 0) rowFace[r] = (int)q;
}

// ------------------------------------------------------------- assembly ---
// The 14 neighbours of a row, in the order the reference lists them
// (viscositysolver.cpp:610-650 for u rows, 700-740 for v, 790-830 for w):
// the face grid (0 u, 1 v, 2 w) and the offset from the row's (i, j, k).
struct Nb {
    signed char dir, di, dj, dk;
};
__constant__ Nb NB[3][14] = {
    {{0, 1, 0, 0}, {0, -1, 0, 0}, {0, 0, 1, 0}, {0, 0, -1, 0}, {0, 0, 0, 1}, {0, 0, 0, -1}, {1, 0, 1, 0},
     {1, -1, 1, 0}, {1, 0, 0, 0}, {1, -1, 0, 0}, {2, 0, 0, 1}, {2, -1, 0, 1}, {2, 0, 0, 0}, {2, -1, 0, 0}},
    {{1, 1, 0, 0}, {1, -1, 0, 0}, {1, 0, 1, 0}, {1, 0, -1, 0}, {1, 0, 0, 1}, {1, 0, 0, -1}, {0, 1, 0, 0},
     {0, 1, -1, 0}, {0, 0, 0, 0}, {0, 0, -1, 0}, {2, 0, 0, 1}, {2, 0, -1, 1}, {2, 0, 0, 0}, {2, 0, -1, 0}},
    {{2, 1, 0, 0}, {2, -1, 0, 0}, {2, 0, 1, 0}, {2, 0, -1, 0}, {2, 0, 0, 1}, {2, 0, 0, -1}, {0, 1, 0, 0},
     {0, 1, 0, -1}, {0, 0, 0, 0}, {0, 0, 0, -1}, {1, 0, 1, 0}, {1, 0, 1, -1}, {1, 0, 0, 0}, {1, 0, 0, -1}}};
// The concatenated face index of neighbour e of row (dir, i, j, k), or -1
// when it lies outside its face grid (U(i, nj, k), U(i, j, nk), V(ni, j, k),
// V(i, j, nk), W(ni, j, k), W(i, nj, k)).  The reference then reads the
// state and row at the wrapped flat index, a border face that is never a
// row or a place past the array, and the velocity as 0
// (macvelocityfield.cpp:112-118): such a neighbour adds nothing, and here it
// is air.
__device__ inline long long nbFaceOf(const Nb n, const Dims &D, int i, int j, int k) {
    const int a = i + n.di, b = j + n.dj, c = k + n.dk;
    if (n.dir == 0) return inRange(a, b, c, D.ni + 1, D.nj, D.nk) ? (long long)flat(a, b, c, D.ni + 1, D.nj) : -1;
    if (n.dir == 1)
        return inRange(a, b, c, D.ni, D.nj + 1, D.nk) ? (long long)(D.nU + flat(a, b, c, D.ni, D.nj + 1)) : -1;
    return inRange(a, b, c, D.ni, D.nj, D.nk + 1) ? (long long)(D.nU + D.nV + flat(a, b, c, D.ni, D.nj)) : -1;
}
__device__ inline long long nbFace(const Dims &D, int dir, int i, int j, int k, int e) {
    return nbFaceOf(NB[dir][e], D, i, j, k);
}
struct FaceView {
    const unsigned char *state;
    const int *table;
    const float *u, *v, *w, *visc;
    Dims D;
    __device__ float VS(int i, int j, int k) const { return visc[flat(i, j, k, D.ni + 1, D.nj + 1)]; }
    // the velocity of concatenated face f
    __device__ float vel(long long f) const {
        if (f < (long long)D.nU) return u[f];
        if (f < (long long)(D.nU + D.nV)) return v[f - D.nU];
        return w[f - D.nU - D.nV];
    }
};

// One thread per row: the row of the viscous operator and its right-hand
// side in the reference's arithmetic (viscositysolver.cpp:559-830).  The
// entries (the diagonal and the fluid neighbours that are rows) are sorted
// by column into a 16-slot ELL block stored slot-major (padding: column =
// row, value 0); solid neighbours move to the right-hand side in list order.
__global__ void assembleKernel(FaceView Fv, VolView Vo, float fdx, float fdt, int m, const int *rowFace,
                               int *ellCol, float *ellVal, int *nnz, float *rhs, float *dinv) {
    const int row = blockIdx.x * blockDim.x + threadIdx.x;
    if (row >= m) return;
    int i, j, k;
    const Dims &D = Fv.D;
    const int dir = decodeFace(D, (size_t)rowFace[row], i, j, k);
    const float invdx = 1.0f / fdx;
    const float factor = fdt * invdx * invdx;
    auto VS = [&](int a, int b, int c) { return Fv.VS(a, b, c); };
    float c[14];
    float diag, rval;
    if (dir == 0) {
        const float vR = VS(i, j, k), vL = VS(i - 1, j, k);
        const float vT = 0.25f * (VS(i - 1, j + 1, k) + VS(i - 1, j, k) + VS(i, j + 1, k) + VS(i, j, k));
        const float vB = 0.25f * (VS(i - 1, j, k) + VS(i - 1, j - 1, k) + VS(i, j, k) + VS(i, j - 1, k));
        const float vF = 0.25f * (VS(i - 1, j, k + 1) + VS(i - 1, j, k) + VS(i, j, k + 1) + VS(i, j, k));
        const float vK = 0.25f * (VS(i - 1, j, k) + VS(i - 1, j, k - 1) + VS(i, j, k) + VS(i, j, k - 1));
        const float fR = 2 * factor * vR * Vo.CE(i, j, k), fL = 2 * factor * vL * Vo.CE(i - 1, j, k);
        const float fT = factor * vT * Vo.EW(i, j + 1, k), fB = factor * vB * Vo.EW(i, j, k);
        const float fF = factor * vF * Vo.EV(i, j, k + 1), fK = factor * vK * Vo.EV(i, j, k);
        diag = Vo.VU(i, j, k) + fR + fL + fT + fB + fF + fK;
        const float cs[14] = {-fR, -fL, -fT, -fB, -fF, -fK, -fT, fT, fB, -fB, -fF, fF, fK, -fK};
        for (int e = 0; e < 14; e++) c[e] = cs[e];
        rval = Vo.VU(i, j, k) * Fv.u[flat(i, j, k, D.ni + 1, D.nj)];
    } else if (dir == 1) {
        const float vT = VS(i, j, k), vB = VS(i, j - 1, k);
        const float vR = 0.25f * (VS(i, j - 1, k) + VS(i + 1, j - 1, k) + VS(i, j, k) + VS(i + 1, j, k));
        const float vL = 0.25f * (VS(i, j - 1, k) + VS(i - 1, j - 1, k) + VS(i, j, k) + VS(i - 1, j, k));
        const float vF = 0.25f * (VS(i, j - 1, k) + VS(i, j - 1, k + 1) + VS(i, j, k) + VS(i, j, k + 1));
        const float vK = 0.25f * (VS(i, j - 1, k) + VS(i, j - 1, k - 1) + VS(i, j, k) + VS(i, j, k - 1));
        const float fR = factor * vR * Vo.EW(i + 1, j, k), fL = factor * vL * Vo.EW(i, j, k);
        const float fT = 2 * factor * vT * Vo.CE(i, j, k), fB = 2 * factor * vB * Vo.CE(i, j - 1, k);
        const float fF = factor * vF * Vo.EU(i, j, k + 1), fK = factor * vK * Vo.EU(i, j, k);
        diag = Vo.VV(i, j, k) + fR + fL + fT + fB + fF + fK;
        const float cs[14] = {-fR, -fL, -fT, -fB, -fF, -fK, -fR, fR, fL, -fL, -fF, fF, fK, -fK};
        for (int e = 0; e < 14; e++) c[e] = cs[e];
        rval = Vo.VV(i, j, k) * Fv.v[flat(i, j, k, D.ni, D.nj + 1)];
    } else {
        const float vF = VS(i, j, k), vK = VS(i, j, k - 1);
        const float vR = 0.25f * (VS(i, j, k) + VS(i, j, k - 1) + VS(i + 1, j, k) + VS(i + 1, j, k - 1));
        const float vL = 0.25f * (VS(i, j, k) + VS(i, j, k - 1) + VS(i - 1, j, k) + VS(i - 1, j, k - 1));
        const float vT = 0.25f * (VS(i, j, k) + VS(i, j, k - 1) + VS(i, j + 1, k) + VS(i, j + 1, k - 1));
        const float vB = 0.25f * (VS(i, j, k) + VS(i, j, k - 1) + VS(i, j - 1, k) + VS(i, j - 1, k - 1));
        const float fR = factor * vR * Vo.EV(i + 1, j, k), fL = factor * vL * Vo.EV(i, j, k);
        const float fT = factor * vT * Vo.EU(i, j + 1, k), fB = factor * vB * Vo.EU(i, j, k);
        const float fF = 2 * factor * vF * Vo.CE(i, j, k), fK = 2 * factor * vK * Vo.CE(i, j, k - 1);
        diag = Vo.VW(i, j, k) + fR + fL + fT + fB + fF + fK;
        const float cs[14] = {-fR, -fL, -fT, -fB, -fF, -fK, -fR, fR, fL, -fL, -fT, fT, fB, -fB};
        for (int e = 0; e < 14; e++) c[e] = cs[e];
        rval = Vo.VW(i, j, k) * Fv.w[flat(i, j, k, D.ni, D.nj)];
    }
    int cols[15];
    float vals[15];
    cols[0] = row;
    vals[0] = diag;
    int cnt = 1;
    for (int e = 0; e < 14; e++) {
        const long long f = nbFace(D, dir, i, j, k, e);
        const unsigned char st = f >= 0 ? Fv.state[f] : (unsigned char)VF_AIR;
        if (st == VF_SOLID) rval -= c[e] * Fv.vel(f);
        const int col = f >= 0 ? Fv.table[f] : -1;
        if (st == VF_FLUID && col >= 0) {
            int p = cnt++;
            while (p > 0 && cols[p - 1] > col) {
                cols[p] = cols[p - 1];
                vals[p] = vals[p - 1];
                p--;
            }
            cols[p] = col;
            vals[p] = c[e];
        }
    }
    for (int sl = 0; sl < ELL; sl++) {
        ellCol[(size_t)sl * m + row] = sl < cnt ? cols[sl] : row;
        ellVal[(size_t)sl * m + row] = sl < cnt ? vals[sl] : 0.0f;
    }
    nnz[row] = cnt;
    rhs[row] = rval;
    // null rows (diagonal below 1e-9, pcgsolver.h:88-104) get no correction
    dinv[row] = diag >= 1e-9f ? 1.0f / diag : 0.0f;
}

// --------------------------------------------------------------- solve ---
__device__ inline double blockSum(double v, double *sh) {
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int o = TPB / 2; o > 0; o >>= 1) {
        if (threadIdx.x < (unsigned)o) sh[threadIdx.x] += sh[threadIdx.x + o];
        __syncthreads();
    }
    return sh[0];
}
__device__ inline double blockMax(double v, double *sh) {
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int o = TPB / 2; o > 0; o >>= 1) {
        if (threadIdx.x < (unsigned)o) sh[threadIdx.x] = fmax(sh[threadIdx.x], sh[threadIdx.x + o]);
        __syncthreads();
    }
    return sh[0];
}
// The CG iterations run in batches without a host round trip: betaKernel
// sets scal[S_DONE] when the residual meets the stopping rule (or is NaN)
// and records the iteration, and every later kernel of the batch returns at
// once.
__device__ inline bool finished(const double *scal) { return scal[S_DONE] != 0.0; }

// x = 0, r = b, z = D^-1 r, s = z; partials of |r|_inf and z.r
__global__ __launch_bounds__(TPB) void initKernel(int m, const float *rhs, const float *dinv, float *x, float *r,
                                                  float *z, float *s, double *pmax, double *pdot) {
    __shared__ double sh[TPB];
    double mx = 0.0, acc = 0.0;
    for (size_t q = (size_t)blockIdx.x * TPB + threadIdx.x; q < (size_t)m; q += (size_t)gridDim.x * TPB) {
        const float b = rhs[q];
        const float zz = dinv[q] * b;
        x[q] = 0.0f;
        r[q] = b;
        z[q] = zz;
        s[q] = zz;
        mx = fmax(mx, (double)fabsf(b));
        acc += (double)zz * (double)b;
    }
    mx = blockMax(mx, sh);
    __syncthreads();
    acc = blockSum(acc, sh);
    if (threadIdx.x == 0) {
        pmax[blockIdx.x] = mx;
        pdot[blockIdx.x] = acc;
    }
}
__global__ __launch_bounds__(TPB) void initFinishKernel(int nb, const double *pmax, const double *pdot,
                                                        double *scal) {
    __shared__ double sh[TPB];
    double mx = 0.0, acc = 0.0;
    for (int q = threadIdx.x; q < nb; q += TPB) {
        mx = fmax(mx, pmax[q]);
        acc += pdot[q];
    }
    mx = blockMax(mx, sh);
    __syncthreads();
    acc = blockSum(acc, sh);
    if (threadIdx.x == 0) {
        scal[S_RESID] = mx;
        scal[S_RHO] = acc;
        scal[S_DONE] = 0.0;
        scal[S_ITERS] = 0.0;
        scal[S_NAN] = 0.0;
    }
}
// q = A s and partials of s.q
__global__ __launch_bounds__(TPB) void spmvKernel(int m, const int *ellCol, const float *ellVal, const float *s,
                                                  float *qv, const double *scal, double *pdot) {
    if (finished(scal)) return;
    __shared__ double sh[TPB];
    double acc = 0.0;
    for (size_t q = (size_t)blockIdx.x * TPB + threadIdx.x; q < (size_t)m; q += (size_t)gridDim.x * TPB) {
        float y = 0.0f;
#pragma unroll
        for (int e = 0; e < ELL; e++) y += ellVal[(size_t)e * m + q] * s[ellCol[(size_t)e * m + q]];
        qv[q] = y;
        acc += (double)s[q] * (double)y;
    }
    acc = blockSum(acc, sh);
    if (threadIdx.x == 0) pdot[blockIdx.x] = acc;
}
__global__ __launch_bounds__(TPB) void alphaKernel(int nb, const double *pdot, double *scal) {
    if (finished(scal)) return;
    __shared__ double sh[TPB];
    double acc = 0.0;
    for (int q = threadIdx.x; q < nb; q += TPB) acc += pdot[q];
    acc = blockSum(acc, sh);
    if (threadIdx.x == 0) scal[S_ALPHA] = scal[S_RHO] / acc;
}
// x += a s, r -= a q, z = D^-1 r; partials of |r|_inf and z.r
__global__ __launch_bounds__(TPB) void updateKernel(int m, const float *s, const float *qv, const float *dinv,
                                                    float *x, float *r, float *z, const double *scal, double *pmax,
                                                    double *pdot) {
    if (finished(scal)) return;
    __shared__ double sh[TPB];
    const float a = (float)scal[S_ALPHA];
    double mx = 0.0, acc = 0.0;
    for (size_t q = (size_t)blockIdx.x * TPB + threadIdx.x; q < (size_t)m; q += (size_t)gridDim.x * TPB) {
        x[q] += a * s[q];
        const float rr = r[q] - a * qv[q];
        r[q] = rr;
        const float zz = dinv[q] * rr;
        z[q] = zz;
        mx = fmax(mx, (double)fabsf(rr));
        acc += (double)zz * (double)rr;
    }
    mx = blockMax(mx, sh);
    __syncthreads();
    acc = blockSum(acc, sh);
    if (threadIdx.x == 0) {
        pmax[blockIdx.x] = mx;
        pdot[blockIdx.x] = acc;
    }
}
// the residual's norm, the stopping rule |r|_inf <= min(tol |b|_inf, 1) of
// iteration it, and beta
__global__ __launch_bounds__(TPB) void betaKernel(int nb, const double *pmax, const double *pdot, double *scal,
                                                  double stopAt, int it) {
    if (finished(scal)) return;
    __shared__ double sh[TPB];
    double mx = 0.0, acc = 0.0;
    for (int q = threadIdx.x; q < nb; q += TPB) {
        mx = fmax(mx, pmax[q]);
        acc += pdot[q];
    }
    mx = blockMax(mx, sh);
    __syncthreads();
    acc = blockSum(acc, sh);
    if (threadIdx.x == 0) {
        scal[S_RESID] = mx;
        scal[S_BETA] = acc / scal[S_RHO];
        scal[S_RHO] = acc;
        if (mx <= stopAt || mx != mx) {
            scal[S_DONE] = 1.0;
            scal[S_ITERS] = it + 1;
            scal[S_NAN] = mx != mx ? 1.0 : 0.0;
        }
    }
}
// s = z + beta s
__global__ __launch_bounds__(TPB) void directionKernel(int m, const float *z, float *s, const double *scal) {
    if (finished(scal)) return;
    const float b = (float)scal[S_BETA];
    for (size_t q = (size_t)blockIdx.x * TPB + threadIdx.x; q < (size_t)m; q += (size_t)gridDim.x * TPB)
        s[q] = z[q] + b * s[q];
}

// -------------------------------------------------------------- apply ---
__global__ void scatterKernel(Dims D, const int *table, const float *x, float *u, float *v, float *w) {
    const size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= D.nU + D.nV + D.nW) return;
    const int r = table[q];
    const float val = r >= 0 ? x[r] : 0.0f;
    if (q < D.nU)
        u[q] = val;
    else if (q < D.nU + D.nV)
        v[q - D.nU] = val;
    else
        w[q - D.nU - D.nV] = val;
}

}  // namespace

ViscositySolver::~ViscositySolver() {
    if (_gridBlock) (void)hipFree(_gridBlock);
    if (_rowBlock) (void)hipFree(_rowBlock);
    if (_hostScal) (void)hipHostFree(_hostScal);
}

namespace {
size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }
}  // namespace

void ViscositySolver::allocateGrid(Grid g) {
    if (_gridBlock && g.ni == _g.ni && g.nj == _g.nj && g.nk == _g.nk && g.dx == _g.dx) return;
    if ((size_t)g.ni * g.nj * g.nk * 8 > (size_t)0x7fffffff)
        throw std::runtime_error("viscosity solve: the grid's subcells exceed the 32-bit row range");
    if (_gridBlock) hipCheck(hipFree(_gridBlock), "viscosity free");
    _gridBlock = nullptr;
    _g = g;
    const Dims D = dimsOf(g);
    _faces = D.nU + D.nV + D.nW;
    _nodes = (size_t)(g.ni + 1) * (g.nj + 1) * (g.nk + 1);
    _subcells = (size_t)8 * g.ni * g.nj * g.nk;
    const size_t vs[7] = {(size_t)g.ni * g.nj * g.nk,       D.nU, D.nV, D.nW,
                          (size_t)g.ni * (g.nj + 1) * (g.nk + 1), (size_t)(g.ni + 1) * g.nj * (g.nk + 1),
                          (size_t)(g.ni + 1) * (g.nj + 1) * g.nk};
    size_t bytes = align256(_faces) + 2 * align256(_nodes) + align256(4 * _subcells);
    for (size_t n : vs) bytes += align256(4 * n);
    bytes += align256(4 * _faces) + align256(4 * (_faces + 1)) + align256(4 * _faces);
    bytes += align256(4 * scanScratchInts(_faces));
    hipCheck(hipMalloc(&_gridBlock, bytes), "viscosity alloc");
    char *p = (char *)_gridBlock;
    auto take = [&](size_t b) {
        char *q = p;
        p += align256(b);
        return q;
    };
    _state = (unsigned char *)take(_faces);
    _valid = (unsigned char *)take(_nodes);
    _valid2 = (unsigned char *)take(_nodes);
    _sub = (float *)take(4 * _subcells);
    for (int q = 0; q < 7; q++) _vol[q] = (float *)take(4 * vs[q]);
    _flag = (int *)take(4 * _faces);
    _rowOf = (int *)take(4 * (_faces + 1));
    _table = (int *)take(4 * _faces);
    _scanScratch = (int *)take(4 * scanScratchInts(_faces));
    if (!_hostScal) hipCheck(hipHostMalloc((void **)&_hostScal, sizeof(double) * S_COUNT), "viscosity pinned");
}

void ViscositySolver::allocateRows(int m) {
    if (m <= _rowCap && _rowBlock) return;
    if (_rowBlock) hipCheck(hipFree(_rowBlock), "viscosity free");
    _rowBlock = nullptr;
    const int cap = std::max(m + m / 4, 1024);
    const size_t n = (size_t)cap;
    size_t bytes = 2 * align256(4 * ELL * n) + 10 * align256(4 * n) + align256(8 * 2 * RED_BLOCKS) +
                   align256(8 * S_COUNT);
    hipCheck(hipMalloc(&_rowBlock, bytes), "viscosity alloc");
    char *p = (char *)_rowBlock;
    auto take = [&](size_t b) {
        char *q = p;
        p += align256(b);
        return q;
    };
    _ellCol = (int *)take(4 * ELL * n);
    _ellVal = (float *)take(4 * ELL * n);
    _rowFace = (int *)take(4 * n);
    _nnz = (int *)take(4 * n);
    _rhs = (float *)take(4 * n);
    _dinv = (float *)take(4 * n);
    _x = (float *)take(4 * n);
    _r = (float *)take(4 * n);
    _z = (float *)take(4 * n);
    _s = (float *)take(4 * n);
    _q = (float *)take(4 * n);
    _partial = (double *)take(8 * 2 * RED_BLOCKS);
    _scal = (double *)take(8 * S_COUNT);
    _rowCap = cap;
}

int ViscositySolver::build(Grid g, double dt, const float *liquidPhi, const float *solidNodes,
                           const float *viscNodes, const float *u, const float *v, const float *w,
                           hipStream_t s) {
    allocateGrid(g);
    const Dims D = dimsOf(g);
    stateKernel<<<blocksFor(_faces, TPB), TPB, 0, s>>>(D, solidNodes, _state);
    validInitKernel<<<blocksFor(_nodes, TPB), TPB, 0, s>>>(g, liquidPhi, _valid);
    dilateKernel<<<blocksFor(_nodes, TPB), TPB, 0, s>>>(g, _valid, _valid2);
    dilateKernel<<<blocksFor(_nodes, TPB), TPB, 0, s>>>(g, _valid2, _valid);
    subcellKernel<<<blocksFor(_subcells, TPB), TPB, 0, s>>>(g, liquidPhi, _valid, _sub);
    VolumeGrids V;
    for (int q = 0; q < 7; q++) V.g[q] = _vol[q];
    volumeKernel<<<blocksFor(_nodes, TPB), TPB, 0, s>>>(g, _valid, _sub, V);
    const VolView Vo = volView(g, _vol);
    flagKernel<<<blocksFor(_faces, TPB), TPB, 0, s>>>(D, _state, Vo, _flag);
    hipCheck(hipGetLastError(), "viscosity index launch");
    exclusiveScan(_flag, _rowOf, _faces, _scanScratch, s);
    int m = 0;
    hipCheck(hipMemcpyAsync(&m, _rowOf + _faces, sizeof(int), hipMemcpyDeviceToHost, s), "viscosity D2H");
    hipCheck(hipStreamSynchronize(s), "viscosity sync");
    _m = m;
    if (m == 0) return 0;
    allocateRows(m);
    tableKernel<<<blocksFor(_faces, TPB), TPB, 0, s>>>(_faces, _flag, _rowOf, _table, _rowFace);
    const FaceView Fv{_state, _table, u, v, w, viscNodes, D};
    assembleKernel<<<blocksFor((size_t)m, TPB), TPB, 0, s>>>(Fv, Vo, (float)g.dx, (float)dt, m, _rowFace, _ellCol,
                                                            _ellVal, _nnz, _rhs, _dinv);
    hipCheck(hipGetLastError(), "viscosity assembly launch");
    return m;
}

ViscosityResult ViscositySolver::solve(Grid g, double dt, const float *liquidPhi, const float *solidNodes,
                                       const float *viscNodes, double tol, int maxIter, float *u, float *v,
                                       float *w, hipStream_t s) {
    ViscosityResult res;
    const int m = build(g, dt, liquidPhi, solidNodes, viscNodes, u, v, w, s);
    res.unknowns = m;
    if (m == 0) return res;  // nothing to solve: the field is kept
    const unsigned nb = std::min<unsigned>(blocksFor((size_t)m, TPB), RED_BLOCKS);
    double *pmax = _partial, *pdot = _partial + RED_BLOCKS;
    const Dims D = dimsOf(g);
    initKernel<<<nb, TPB, 0, s>>>(m, _rhs, _dinv, _x, _r, _z, _s, pmax, pdot);
    initFinishKernel<<<1, TPB, 0, s>>>((int)nb, pmax, pdot, _scal);
    hipCheck(hipMemcpyAsync(_hostScal, _scal, sizeof(double) * S_COUNT, hipMemcpyDeviceToHost, s), "viscosity D2H");
    hipCheck(hipStreamSynchronize(s), "viscosity sync");
    float tolFactor = (float)tol;
    if (tolFactor < 1e-30) tolFactor = 1e-30f;
    const float resid0 = (float)_hostScal[S_RESID];
    res.error = resid0;
    const double tolAbs = tolFactor * resid0;
    const double rho0 = _hostScal[S_RHO];
    bool ok = resid0 == 0;  // a zero right-hand side: the zero solution
    int it = 0;
    if (!ok && (rho0 == 0 || rho0 != rho0)) {
        res.success = false;
        res.iterations = 0;
        return res;
    }
    const double stopAt = std::min(tolAbs, 1.0);
    bool nan = false;
    while (!ok && !nan && it < maxIter) {
        const int end = std::min(it + BATCH, maxIter);
        for (int k = it; k < end; k++) {
            spmvKernel<<<nb, TPB, 0, s>>>(m, _ellCol, _ellVal, _s, _q, _scal, pdot);
            alphaKernel<<<1, TPB, 0, s>>>((int)nb, pdot, _scal);
            updateKernel<<<nb, TPB, 0, s>>>(m, _s, _q, _dinv, _x, _r, _z, _scal, pmax, pdot);
            betaKernel<<<1, TPB, 0, s>>>((int)nb, pmax, pdot, _scal, stopAt, k);
            directionKernel<<<nb, TPB, 0, s>>>(m, _z, _s, _scal);
        }
        hipCheck(hipGetLastError(), "viscosity CG launch");
        hipCheck(hipMemcpyAsync(_hostScal, _scal, sizeof(double) * S_COUNT, hipMemcpyDeviceToHost, s),
                 "viscosity D2H");
        hipCheck(hipStreamSynchronize(s), "viscosity sync");
        it = end;
        res.error = (float)_hostScal[S_RESID];
        if (_hostScal[S_DONE] != 0.0) {
            it = (int)_hostScal[S_ITERS];
            if (_hostScal[S_NAN] != 0.0)
                nan = true;  // the reference would run to the limit and fail
            else
                ok = true;
        }
    }
    if (nan) it = maxIter;
    res.iterations = it;
    res.success = ok || (it == maxIter && !nan && res.error < 10.0f);
    if (!res.success) return res;
    scatterKernel<<<blocksFor(_faces, TPB), TPB, 0, s>>>(D, _table, _x, u, v, w);
    hipCheck(hipGetLastError(), "viscosity apply launch");
    return res;
}

void ViscositySolver::copySystem(ViscosityHostSystem &out, hipStream_t s) {
    const Dims D = dimsOf(_g);
    out.states.resize(_faces);
    hipCheck(hipMemcpyAsync(out.states.data(), _state, _faces, hipMemcpyDeviceToHost, s), "viscosity D2H");
    const size_t vs[7] = {(size_t)_g.ni * _g.nj * _g.nk,         D.nU, D.nV, D.nW,
                          (size_t)_g.ni * (_g.nj + 1) * (_g.nk + 1), (size_t)(_g.ni + 1) * _g.nj * (_g.nk + 1),
                          (size_t)(_g.ni + 1) * (_g.nj + 1) * _g.nk};
    size_t total = 0;
    for (size_t n : vs) total += n;
    out.volumes.resize(total);
    size_t o = 0;
    for (int q = 0; q < 7; q++) {
        hipCheck(hipMemcpyAsync(out.volumes.data() + o, _vol[q], 4 * vs[q], hipMemcpyDeviceToHost, s),
                 "viscosity D2H");
        o += vs[q];
    }
    out.table.assign(_faces, -1);
    out.rowstart.assign((size_t)_m + 1, 0);
    out.rhs.assign(_m, 0.0f);
    out.cols.clear();
    out.vals.clear();
    if (_m == 0) {
        hipCheck(hipStreamSynchronize(s), "viscosity sync");
        return;
    }
    hipCheck(hipMemcpyAsync(out.table.data(), _table, 4 * _faces, hipMemcpyDeviceToHost, s), "viscosity D2H");
    hipCheck(hipMemcpyAsync(out.rhs.data(), _rhs, 4 * (size_t)_m, hipMemcpyDeviceToHost, s), "viscosity D2H");
    std::vector<int> nnz(_m), ec((size_t)ELL * _m);
    std::vector<float> ev((size_t)ELL * _m);
    hipCheck(hipMemcpyAsync(nnz.data(), _nnz, 4 * (size_t)_m, hipMemcpyDeviceToHost, s), "viscosity D2H");
    hipCheck(hipMemcpyAsync(ec.data(), _ellCol, 4 * (size_t)ELL * _m, hipMemcpyDeviceToHost, s), "viscosity D2H");
    hipCheck(hipMemcpyAsync(ev.data(), _ellVal, 4 * (size_t)ELL * _m, hipMemcpyDeviceToHost, s), "viscosity D2H");
    hipCheck(hipStreamSynchronize(s), "viscosity sync");
    for (int r = 0; r < _m; r++) {
        for (int e = 0; e < nnz[r]; e++) {
            out.cols.push_back(ec[(size_t)e * _m + r]);
            out.vals.push_back(ev[(size_t)e * _m + r]);
        }
        out.rowstart[r + 1] = (int)out.cols.size();
    }
}

void ViscositySolver::copySolution(std::vector<float> &x, hipStream_t s) {
    x.assign(_m, 0.0f);
    if (_m == 0) return;
    hipCheck(hipMemcpyAsync(x.data(), _x, 4 * (size_t)_m, hipMemcpyDeviceToHost, s), "viscosity D2H");
    hipCheck(hipStreamSynchronize(s), "viscosity sync");
}

}  // namespace ff
