// kernels_pressure.hip -- the variational pressure solve of the FLIP substep:
// pressure-cell rows, divergence and 7-point ghost-fluid assembly, the MIC(0)
// factor, its hyperplane-scheduled triangular solves, SpMV and the PCG vector
// ops.
//
// Rows keep the reference's order (interior liquid cells in k,j,i order,
// pressuresolver.cpp:118-159).  The factor and both triangular sweeps depend
// only on the -x/-y/-z (or +x/+y/+z) neighbours, which sit on the previous
// (next) i+j+k hyperplane, so running the rows hyperplane by hyperplane
// performs the reference's own serial recurrences (pcgsolver.h:68-241) with
// identical operations per row.
#include "kernels.h"

namespace ff {

namespace {

unsigned blocksFor(size_t n, unsigned t) { return (unsigned)((n + t - 1) / t); }
constexpr int TPB = 256;

__global__ void maskKernel(Grid g, const float *phi, int *mask) {
    size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t n = (size_t)g.ni * g.nj * g.nk;
    if (c > n) return;
    if (c == n) {
        mask[n] = 0;
        return;
    }
    int i = (int)(c % g.ni), j = (int)((c / g.ni) % g.nj), k = (int)(c / ((size_t)g.ni * g.nj));
    bool interior = i >= 1 && j >= 1 && k >= 1 && i < g.ni - 1 && j < g.nj - 1 && k < g.nk - 1;
    mask[c] = (interior && phi[c] < 0) ? 1 : 0;
}

__global__ void rowsKernel(size_t n, const int *mask, const int *rowOf, int *rows, int *cellRow) {
    size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n) return;
    int row = mask[c] ? rowOf[c] : -1;
    cellRow[c] = row;
    if (row >= 0) rows[row] = (int)c;
}

// _calculateNegativeDivergenceVectorThread (pressuresolver.cpp:624-713) and
// _calculateMatrixCoefficientsThread (767-867).  Static solids have zero face
// velocity, so the divergence's solid terms add +-0.0 to a sum that is never
// -0.0 and are left out; a moving solid's (T.us given) are added with the
// cell weight.  Density is 1 (fluidsimulation.cpp:6214).  An
// off-diagonal entry exists when the neighbour is liquid and a pressure row
// (SparseMatrix::add ignores -1, sparsematrix.h:94-96).
__global__ void assembleKernel(Grid g, double dt, const float *u, const float *v, const float *w,
                               const float *wu, const float *wv, const float *ww, const float *phi,
                               const int *mask, const int *rowOf, PressureSys S, SolidTerms T,
                               TensionView tv) {
    int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= S.m) return;
    const int ni = g.ni, nj = g.nj;
    int c = S.rows[r];
    int i = c % ni, j = (c / ni) % nj, k = c / (ni * nj);
    const double factor = 1.0 / g.dx;
    const double mfactor = dt / (g.dx * g.dx);
    const double eps = 1e-9, maxtheta = 25;
    double volRight = wu[flat(i + 1, j, k, ni + 1, nj)], volLeft = wu[flat(i, j, k, ni + 1, nj)];
    double volTop = wv[flat(i, j + 1, k, ni, nj + 1)], volBottom = wv[flat(i, j, k, ni, nj + 1)];
    double volFront = ww[flat(i, j, k + 1, ni, nj)], volBack = ww[flat(i, j, k, ni, nj)];
    double div = 0.0;
    div += -factor * volRight * (double)u[flat(i + 1, j, k, ni + 1, nj)];
    div += factor * volLeft * (double)u[flat(i, j, k, ni + 1, nj)];
    div += -factor * volTop * (double)v[flat(i, j + 1, k, ni, nj + 1)];
    div += factor * volBottom * (double)v[flat(i, j, k, ni, nj + 1)];
    div += -factor * volFront * (double)w[flat(i, j, k + 1, ni, nj)];
    div += factor * volBack * (double)w[flat(i, j, k, ni, nj)];
    if (T.us) {
        const double volCenter = T.wc[c];
        div += factor * (volRight - volCenter) * (double)T.us[flat(i + 1, j, k, ni + 1, nj)];
        div += -factor * (volLeft - volCenter) * (double)T.us[flat(i, j, k, ni + 1, nj)];
        div += factor * (volTop - volCenter) * (double)T.vs[flat(i, j + 1, k, ni, nj + 1)];
        div += -factor * (volBottom - volCenter) * (double)T.vs[flat(i, j, k, ni, nj + 1)];
        div += factor * (volFront - volCenter) * (double)T.ws[flat(i, j, k + 1, ni, nj)];
        div += -factor * (volBack - volCenter) * (double)T.ws[flat(i, j, k, ni, nj)];
    }
    if (tv.status) {  // surface tension (pressuresolver.cpp:659-710)
        const double stfactor = dt / (g.dx * g.dx);
        const double pc = phi[c];
        auto add = [&](int a, int b, int e, double vol) {
            const double pn = phi[flat(a, b, e, ni, nj)];
            if (pn >= 0.0) {
                const double t = tensionTerm(g, tv, i, j, k, a, b, e);
                double theta = (pc - pn) / (pc + eps);
                theta = fmax(-maxtheta, fmin(theta, maxtheta));
                div += stfactor * vol * theta * t;
            }
        };
        add(i + 1, j, k, volRight);
        add(i - 1, j, k, volLeft);
        add(i, j + 1, k, volTop);
        add(i, j - 1, k, volBottom);
        add(i, j, k + 1, volFront);
        add(i, j, k - 1, volBack);
    }
    S.rhs[r] = div;

    const double rho = (1.0 + 1.0) / 2.0;
    double phiC = phi[c];
    const size_t sx = 1, sy = (size_t)ni, sz = (size_t)ni * nj;
    double diag = (volRight / rho + volLeft / rho + volTop / rho + volBottom / rho + volFront / rho +
                   volBack / rho) * mfactor;
    int *nb = S.nbr + (size_t)6 * r;
    double off[6] = {0, 0, 0, 0, 0, 0};
    const double vols[6] = {volLeft, volRight, volBottom, volTop, volBack, volFront};
    const long long steps[6] = {-(long long)sx, (long long)sx, -(long long)sy, (long long)sy,
                                -(long long)sz, (long long)sz};
    // reference neighbour order: +x, -x, +y, -y, +z, -z
    const int order[6] = {1, 0, 3, 2, 5, 4};
#pragma unroll
    for (int q = 0; q < 6; q++) {
        int sl = order[q];
        size_t nc = (size_t)((long long)c + steps[sl]);
        double phiN = phi[nc];
        nb[sl] = -1;
        if (phiN < 0.0) {
            if (mask[nc]) {
                nb[sl] = rowOf[nc];
                off[sl] = -(vols[sl] / rho) * mfactor;
            }
        } else {
            double theta = phiN / (phiC + eps);
            theta = fmax(-maxtheta, fmin(theta, maxtheta));
            diag -= (vols[sl] / rho) * mfactor * theta;
        }
    }
    diag = fmax(diag, 0.0);
    S.diag[r] = diag;
    S.mx[r] = off[0];
    S.ox[r] = off[1];
    S.my[r] = off[2];
    S.oy[r] = off[3];
    S.mz[r] = off[4];
    S.oz[r] = off[5];
}

__global__ void levelCountKernel(Grid g, PressureSys S) {
    int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= S.m) return;
    int c = S.rows[r];
    int l = c % g.ni + (c / g.ni) % g.nj + c / (g.ni * g.nj);
    atomicAdd(&S.levelCount[l], 1);
}
__global__ void levelScatterKernel(Grid g, PressureSys S) {
    int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= S.m) return;
    int c = S.rows[r];
    int l = c % g.ni + (c / g.ni) % g.nj + c / (g.ni * g.nj);
    int slot = atomicAdd(&S.levelFill[l], 1);
    S.levelRows[S.levelStart[l] + slot] = r;
}

// factorModifiedIncompleteColesky0 (pcgsolver.h:68-205), omega 0.97, min
// diagonal ratio 0.25, one hyperplane of rows: every row pulls its updates
// from the -z, -y, -x neighbour columns (ascending column order); within a
// column the "missing" sum runs over the column's other rows in ascending row
// order (+x, +y, +z).  A row whose diagonal is below 1e-9 is null: it keeps
// its updated diagonal and contributes nothing as a column.
__global__ void mic0LevelKernel(PressureSys S, const int *rows, int count) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= count) return;
    int j = rows[t];
    const int *nb = S.nbr + (size_t)6 * j;
    double a = S.diag[j];
    if (a < 1e-9) a = 0;
    double d = a;
    const int cslots[3] = {4, 2, 0};
#pragma unroll
    for (int q = 0; q < 3; q++) {
        int k = nb[cslots[q]];
        if (k < 0) continue;
        if (S.diag[k] < 1e-9) continue;
        int jslot = cslots[q] + 1;
        const int *kb = S.nbr + (size_t)6 * k;
        double ik = S.invdiag[k];
        double okx = S.ox[k], oky = S.oy[k], okz = S.oz[k];
        double Ljk = (jslot == 1 ? okx : (jslot == 3 ? oky : okz)) * ik;
        d -= Ljk * Ljk;
        double missing = 0;
        if (jslot != 1 && kb[1] >= 0) missing += okx * ik;
        if (jslot != 3 && kb[3] >= 0) missing += oky * ik;
        if (jslot != 5 && kb[5] >= 0) missing += okz * ik;
        d -= 0.97 * Ljk * missing;
    }
    double inv;
    if (a == 0)
        inv = d;
    else if (d < 0.25 * a)
        inv = 1 / sqrt(a);
    else
        inv = 1 / sqrt(d);
    S.invdiag[j] = inv;
}

// L entries toward +x/+y/+z: A * invdiag, or A itself for a null column
__global__ void lfactorKernel(PressureSys S) {
    int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= S.m) return;
    bool null = S.diag[r] < 1e-9;
    double iv = S.invdiag[r];
    S.lx[r] = null ? S.ox[r] : S.ox[r] * iv;
    S.ly[r] = null ? S.oy[r] : S.oy[r] * iv;
    S.lz[r] = null ? S.oz[r] : S.oz[r] * iv;
}

// solveLower (pcgsolver.h:211-224), one hyperplane: subtractions in column
// order -z, -y, -x, then the scaling by invdiag
template <class T>
__global__ void forwardLevelKernel(const int *nbr, const T *invdiag, const T *lx, const T *ly,
                                   const T *lz, const int *rows, int count, const T *rv, T *z) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= count) return;
    int j = rows[t];
    const int *nb = nbr + (size_t)6 * j;
    T y = rv[j];
    int k = nb[4];
    if (k >= 0) y -= lz[k] * z[k];
    k = nb[2];
    if (k >= 0) y -= ly[k] * z[k];
    k = nb[0];
    if (k >= 0) y -= lx[k] * z[k];
    z[j] = y * invdiag[j];
}
// solveLowerTransposeInPlace (pcgsolver.h:227-241), one hyperplane
template <class T>
__global__ void backwardLevelKernel(const int *nbr, const T *invdiag, const T *lx, const T *ly,
                                    const T *lz, const int *rows, int count, T *z) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= count) return;
    int i = rows[t];
    const int *nb = nbr + (size_t)6 * i;
    T x = z[i];
    int r = nb[1];
    if (r >= 0) x -= lx[i] * z[r];
    r = nb[3];
    if (r >= 0) x -= ly[i] * z[r];
    r = nb[5];
    if (r >= 0) x -= lz[i] * z[r];
    z[i] = x * invdiag[i];
}

// FixedSparseMatrix multiply (sparsematrix.h:174-215): ascending column order
// (A x)[r] with the terms added in ascending column order: -z, -y, -x, diag,
// +x, +y, +z.  Every load is issued up front (an absent neighbour reads x[r]
// and its term is skipped by a select), so the row costs one round trip.
// The -x/-y/-z coefficients are read as the neighbour rows' +x/+y/+z ones:
// the assembly computes both from the same face by the same expression, so
// they are the same bits, and those rows were just read by nearby threads
// (mx, my, mz are kept for the other users of the matrix).
// Row r's product with the 7-point matrix, (A x)_r, in the reference's order
// (sparsematrix.h:174-215 over the row's entries by column: -z, -y, -x,
// diagonal, +x, +y, +z).  The -x/-y/-z coefficients are the neighbour rows'
// +x/+y/+z entries (the same bits: both sides assemble -(weight / rho) * dt /
// dx^2 from the shared face).  A missing neighbour's term is left out, as
// SparseMatrix leaves out absent entries.
struct RowTopo {
    int m;
    const int *nbr;  // 6 per row: -x, +x, -y, +y, -z, +z rows or -1
};

template <class T>
__device__ inline T rowProduct(int r, const RowTopo &t, const T *diag, const T *ox, const T *oy, const T *oz,
                               const T *x) {
    const int *nb = t.nbr + 6 * (size_t)r;
    const int n0 = nb[0], n1 = nb[1], n2 = nb[2], n3 = nb[3], n4 = nb[4], n5 = nb[5];
    const int a0 = n0 >= 0 ? n0 : r, a1 = n1 >= 0 ? n1 : r, a2 = n2 >= 0 ? n2 : r, a3 = n3 >= 0 ? n3 : r,
              a4 = n4 >= 0 ? n4 : r, a5 = n5 >= 0 ? n5 : r;
    const T cmz = oz[a4], cmy = oy[a2], cmx = ox[a0], cd = diag[r], cox = ox[r], coy = oy[r], coz = oz[r];
    const T x4 = x[a4], x2 = x[a2], x0 = x[a0], xr = x[r], x1 = x[a1], x3 = x[a3], x5 = x[a5];
    T s = 0;
    if (n4 >= 0) s += cmz * x4;
    if (n2 >= 0) s += cmy * x2;
    if (n0 >= 0) s += cmx * x0;
    s += cd * xr;
    if (n1 >= 0) s += cox * x1;
    if (n3 >= 0) s += coy * x3;
    if (n5 >= 0) s += coz * x5;
    return s;
}

// one row per thread; blocks renumbered so that each die streams one
// contiguous eighth of the rows (its neighbour rows then sit in its L2)
__device__ inline int dieRow() {
    const unsigned b = blockIdx.x, g = gridDim.x;
    const unsigned lb = g % XCDS == 0 ? (b % XCDS) * (g / XCDS) + b / XCDS : b;
    return (int)(lb * blockDim.x + threadIdx.x);
}

template <class T>
__global__ void spmvKernel(RowTopo t, const T *diag, const T *ox, const T *oy, const T *oz, const T *x, T *y) {
    const int r = dieRow();
    if (r >= t.m) return;
    y[r] = rowProduct<T>(r, t, diag, ox, oy, oz, x);
}

// U rows' products at once: the neighbour lists of all U rows are loaded
// before their values, so a thread keeps U lookups in flight.  Indices of
// absent neighbours are clamped to the row itself (a valid address) and their
// terms left out, as in rowProduct.
template <class T, int U>
__device__ inline void rowProducts(const int (&r)[U], const RowTopo &t, const T *__restrict__ diag,
                                   const T *__restrict__ ox, const T *__restrict__ oy, const T *__restrict__ oz,
                                   const T *__restrict__ x, T (&out)[U]) {
    int nb[U][6];
#pragma unroll
    for (int u = 0; u < U; u++)
#pragma unroll
        for (int q = 0; q < 6; q++) nb[u][q] = t.nbr[6 * (size_t)r[u] + q];
#pragma unroll
    for (int u = 0; u < U; u++) {
        const int q = r[u];
        const int n0 = nb[u][0], n1 = nb[u][1], n2 = nb[u][2], n3 = nb[u][3], n4 = nb[u][4], n5 = nb[u][5];
        const int a0 = n0 >= 0 ? n0 : q, a1 = n1 >= 0 ? n1 : q, a2 = n2 >= 0 ? n2 : q, a3 = n3 >= 0 ? n3 : q,
                  a4 = n4 >= 0 ? n4 : q, a5 = n5 >= 0 ? n5 : q;
        const T cmz = oz[a4], cmy = oy[a2], cmx = ox[a0], cd = diag[q], cox = ox[q], coy = oy[q], coz = oz[q];
        const T x4 = x[a4], x2 = x[a2], x0 = x[a0], xr = x[q], x1 = x[a1], x3 = x[a3], x5 = x[a5];
        T s = 0;
        if (n4 >= 0) s += cmz * x4;
        if (n2 >= 0) s += cmy * x2;
        if (n0 >= 0) s += cmx * x0;
        s += cd * xr;
        if (n1 >= 0) s += cox * x1;
        if (n3 >= 0) s += coy * x3;
        if (n5 >= 0) s += coz * x5;
        out[u] = s;
    }
}

// Neighbour lookup of the fused SpMV: the rows' six neighbour ids (RowTopo,
// the default) or, built with -DFF_SPMV_WORDS=1, the rows' cells and the line
// words (WordTopo: 4 B per row instead of 24, one more dependent load; it
// measured slower at 256^3, 148 ms against 107 ms of SpMV per 10 substeps).
#ifndef FF_SPMV_WORDS
#define FF_SPMV_WORDS 0
#endif
// The rows' neighbours found from the rows' cells and the line words (no
// per-row neighbour ids): bit i%32 of word i/32 of line (j, k) marks a
// pressure cell, whose row is the word's base plus the marked cells below it
// in the word.  Rows lie on interior cells, so every neighbour cell is inside
// the grid.
struct WordTopo {
    int m, ni, nj, nw;
    const int *rows;  // flat cell of each row, (k*nj + j)*ni + i
    const LineWord *words;
};

__device__ inline int wordRow(const LineWord *__restrict__ words, size_t line, int nw, int i) {
    const LineWord w = words[line * nw + (i >> 5)];
    const unsigned bit = 1u << (i & 31);
    return (w.bits & bit) ? w.base + __builtin_popcount(w.bits & (bit - 1u)) : -1;
}

template <class T, int U>
__device__ inline void rowProducts(const int (&r)[U], const WordTopo &t, const T *__restrict__ diag,
                                    const T *__restrict__ ox, const T *__restrict__ oy, const T *__restrict__ oz,
                                    const T *__restrict__ x, T (&out)[U]) {
    int cell[U];
#pragma unroll
    for (int u = 0; u < U; u++) cell[u] = t.rows[r[u]];
    int nb[U][6];
#pragma unroll
    for (int u = 0; u < U; u++) {
        const int c = cell[u], i = c % t.ni, line = c / t.ni;
        const size_t nj = (size_t)t.nj;
        nb[u][0] = wordRow(t.words, line, t.nw, i - 1);
        nb[u][1] = wordRow(t.words, line, t.nw, i + 1);
        nb[u][2] = wordRow(t.words, line - 1, t.nw, i);
        nb[u][3] = wordRow(t.words, line + 1, t.nw, i);
        nb[u][4] = wordRow(t.words, line - nj, t.nw, i);
        nb[u][5] = wordRow(t.words, line + nj, t.nw, i);
    }
#pragma unroll
    for (int u = 0; u < U; u++) {
        const int q = r[u];
        const int n0 = nb[u][0], n1 = nb[u][1], n2 = nb[u][2], n3 = nb[u][3], n4 = nb[u][4], n5 = nb[u][5];
        const int a0 = n0 >= 0 ? n0 : q, a1 = n1 >= 0 ? n1 : q, a2 = n2 >= 0 ? n2 : q, a3 = n3 >= 0 ? n3 : q,
                  a4 = n4 >= 0 ? n4 : q, a5 = n5 >= 0 ? n5 : q;
        const T cmz = oz[a4], cmy = oy[a2], cmx = ox[a0], cd = diag[q], cox = ox[q], coy = oy[q], coz = oz[q];
        const T x4 = x[a4], x2 = x[a2], x0 = x[a0], xr = x[q], x1 = x[a1], x3 = x[a3], x5 = x[a5];
        T s = 0;
        if (n4 >= 0) s += cmz * x4;
        if (n2 >= 0) s += cmy * x2;
        if (n0 >= 0) s += cmx * x0;
        s += cd * xr;
        if (n1 >= 0) s += cox * x1;
        if (n3 >= 0) s += coy * x3;
        if (n5 >= 0) s += coz * x5;
        out[u] = s;
    }
}

// SpMV over the rows of the reduction spans, with each thread's partial of
// s . q in the order of the standalone dot product (its rows in ascending
// order, SPMV_ROWS at a time)
constexpr int SPMV_ROWS = 4;
template <class T, class Topo>
__global__ __launch_bounds__(REDUCE_THREADS) void spmvDotKernel(Topo t, const T *__restrict__ diag,
                                                                const T *__restrict__ ox, const T *__restrict__ oy,
                                                                const T *__restrict__ oz, const T *__restrict__ x,
                                                                T *__restrict__ y, double *partial) {
    const unsigned L = reduceLogicalBlock();
    const size_t span = reduceSpan((size_t)t.m, gridDim.x);
    const size_t lo = (size_t)L * span, hi = lo + span < (size_t)t.m ? lo + span : (size_t)t.m;
    double acc = 0.0;
    for (size_t b = lo + threadIdx.x; b < hi; b += SPMV_ROWS * REDUCE_THREADS) {
        int r[SPMV_ROWS];
        bool ok[SPMV_ROWS];
#pragma unroll
        for (int u = 0; u < SPMV_ROWS; u++) {
            const size_t q = b + (size_t)u * REDUCE_THREADS;
            ok[u] = q < hi;
            r[u] = (int)(ok[u] ? q : b);
        }
        T s[SPMV_ROWS];
        rowProducts<T, SPMV_ROWS>(r, t, diag, ox, oy, oz, x, s);
#pragma unroll
        for (int u = 0; u < SPMV_ROWS; u++)
            if (ok[u]) {
                y[r[u]] = s[u];
                acc += (double)x[r[u]] * (double)s[u];
            }
    }
    acc = reduceFold<0>(acc);
    if (threadIdx.x == 0) partial[L] = acc;
}

// x += a s; r -= a q with a the float-rounded step, and the partial max |r|
template <class TX, class T>
__global__ __launch_bounds__(REDUCE_THREADS) void pcgStepKernel(TX *x, const T *s, T *r, const T *q, size_t n,
                                                                const double *num, const double *den,
                                                                double *partial, float *rf) {
    const double a = *num / *den;
    const float fa = (float)a, fna = (float)(-a);
    const unsigned L = reduceLogicalBlock();
    const size_t span = reduceSpan(n, gridDim.x);
    const size_t lo = (size_t)L * span, hi = lo + span < n ? lo + span : n;
    double acc = 0.0;
    for (size_t i = lo + threadIdx.x; i < hi; i += REDUCE_THREADS) {
        x[i] += (TX)fa * (TX)s[i];
        const T ri = r[i] + (T)fna * q[i];
        r[i] = ri;
        if (rf) rf[i] = (float)ri;  // FP32 copy for an FP32 preconditioner
        acc = fmax(acc, fabs((double)ri));
    }
    acc = reduceFold<1>(acc);
    if (threadIdx.x == 0) partial[L] = acc;
}

// r = b - A x
__global__ void residualKernel(RowTopo t, const double *diag, const double *ox, const double *oy,
                               const double *oz, const double *b, const double *x, double *r) {
    const int q = dieRow();
    if (q >= t.m) return;
    r[q] = b[q] - rowProduct<double>(q, t, diag, ox, oy, oz, x);
}

// BLAS::addScaled(float alpha, ...) (blaswrapper.h:128-157): the step length
// is rounded to float before it scales the FP64 vector
template <class T>
__global__ void axpyStepKernel(T *y, const T *x, size_t n, const double *num, const double *den,
                               double sign) {
    size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    double a = den ? *num / *den : *num;
    float fa = (float)(sign * a);
    y[q] += (T)fa * x[q];
}

// z = (double)zf + (double)(float)(num/den) * s: the FP32 preconditioner's
// result widened, then the search-direction update of axpyStepKernel
__global__ void widenAxpyKernel(double *z, const float *zf, const double *s, size_t n, const double *num,
                                const double *den) {
    size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    const float fa = (float)(*num / *den);
    double zq = (double)zf[q];
    zq += (double)fa * s[q];
    z[q] = zq;
}

// x += (double)(float)(sign * num/den) * (double)p: the FP64 solution takes
// the FP32 direction with the same float-rounded step length
__global__ void axpyMixedKernel(double *x, const float *p, size_t n, const double *num, const double *den,
                                double sign) {
    size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    double a = den ? *num / *den : *num;
    float fa = (float)(sign * a);
    x[q] += (double)fa * (double)p[q];
}

__global__ void gatherRowsKernel(int m, const int *rows, const double *cells, double *rhs) {
    int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r < m) rhs[r] = cells[rows[r]];
}
__global__ void gatherRowsFKernel(int m, const int *rows, const float *cells, double *x) {
    int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r < m) x[r] = (double)cells[rows[r]];
}
__global__ void scatterRowsKernel(int m, const int *rows, const double *x, double *cells) {
    int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r < m) cells[rows[r]] = x[r];
}
__global__ void toGridKernel(int m, const int *rows, const double *x, float *p) {
    int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r < m) p[rows[r]] = (float)x[r];
}
__global__ void castKernel(const double *a, float *b, size_t n) {
    size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q < n) b[q] = (float)a[q];
}
__global__ void addFloatKernel(double *x, const float *d, size_t n) {
    size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q < n) x[q] += (double)d[q];
}

}  // namespace

void pressureMask(Grid g, const float *phi, int *mask, hipStream_t s) {
    size_t n = (size_t)g.ni * g.nj * g.nk;
    hipLaunchKernelGGL(maskKernel, dim3(blocksFor(n + 1, TPB)), dim3(TPB), 0, s, g, phi, mask);
}

void pressureRows(Grid g, const int *mask, const int *rowOf, int *rows, int *cellRow, hipStream_t s) {
    size_t n = (size_t)g.ni * g.nj * g.nk;
    hipLaunchKernelGGL(rowsKernel, dim3(blocksFor(n, TPB)), dim3(TPB), 0, s, n, mask, rowOf, rows, cellRow);
}

void pressureAssemble(Grid g, double dt, const float *u, const float *v, const float *w,
                      const float *wu, const float *wv, const float *ww, const float *phi,
                      PressureSys &S, hipStream_t s, SolidTerms solid, TensionView tension) {
    if (S.m == 0) return;
    hipLaunchKernelGGL(assembleKernel, dim3(blocksFor(S.m, TPB)), dim3(TPB), 0, s, g, dt, u, v, w, wu,
                       wv, ww, phi, S.mask, S.rowOf, S, solid, tension);
}

void pressureLevels(Grid g, PressureSys &S, int *scanScratch, hipStream_t s) {
    S.nLevels = g.ni + g.nj + g.nk - 2;
    hipCheck(hipMemsetAsync(S.levelCount, 0, sizeof(int) * (S.nLevels + 1), s), "hipMemsetAsync");
    hipCheck(hipMemsetAsync(S.levelFill, 0, sizeof(int) * S.nLevels, s), "hipMemsetAsync");
    if (S.m > 0)
        hipLaunchKernelGGL(levelCountKernel, dim3(blocksFor(S.m, TPB)), dim3(TPB), 0, s, g, S);
    exclusiveScan(S.levelCount, S.levelStart, S.nLevels, scanScratch, s);
    if (S.m > 0)
        hipLaunchKernelGGL(levelScatterKernel, dim3(blocksFor(S.m, TPB)), dim3(TPB), 0, s, g, S);
}

void mic0Factor(PressureSys &S, const int *hs, hipStream_t s) {
    for (int l = 0; l < S.nLevels; l++) {
        int cnt = hs[l + 1] - hs[l];
        if (cnt == 0) continue;
        hipLaunchKernelGGL(mic0LevelKernel, dim3(blocksFor(cnt, TPB)), dim3(TPB), 0, s, S,
                           S.levelRows + hs[l], cnt);
    }
    mic0LFactor(S, s);
}

void mic0LFactor(PressureSys &S, hipStream_t s) {
    if (S.m > 0) hipLaunchKernelGGL(lfactorKernel, dim3(blocksFor(S.m, TPB)), dim3(TPB), 0, s, S);
}

template <class T>
static void micApply(const PressureSys &S, const int *hs, const T *invdiag, const T *lx, const T *ly,
                     const T *lz, const T *r, T *z, hipStream_t s) {
    for (int l = 0; l < S.nLevels; l++) {
        int cnt = hs[l + 1] - hs[l];
        if (cnt == 0) continue;
        hipLaunchKernelGGL(forwardLevelKernel<T>, dim3(blocksFor(cnt, TPB)), dim3(TPB), 0, s, S.nbr,
                           invdiag, lx, ly, lz, S.levelRows + hs[l], cnt, r, z);
    }
    for (int l = S.nLevels - 1; l >= 0; l--) {
        int cnt = hs[l + 1] - hs[l];
        if (cnt == 0) continue;
        hipLaunchKernelGGL(backwardLevelKernel<T>, dim3(blocksFor(cnt, TPB)), dim3(TPB), 0, s, S.nbr,
                           invdiag, lx, ly, lz, S.levelRows + hs[l], cnt, z);
    }
}

void mic0Apply(const PressureSys &S, const int *hs, const double *r, double *z, hipStream_t s) {
    micApply<double>(S, hs, S.invdiag, S.lx, S.ly, S.lz, r, z, s);
}
void mic0ApplyF(const PressureSys &S, const int *hs, const float *r, float *z, hipStream_t s) {
    micApply<float>(S, hs, S.finvdiag, S.flx, S.fly, S.flz, r, z, s);
}

static RowTopo topo(const PressureSys &S) { return RowTopo{S.m, S.nbr}; }
// one row per thread over whole runs of XCDS blocks (so dieRow renumbers them)
static unsigned dieBlocks(size_t m) {
    const unsigned b = blocksFor(m, TPB);
    return (b + XCDS - 1) / XCDS * XCDS;
}

void spmv(const PressureSys &S, const double *x, double *y, hipStream_t s) {
    if (S.m == 0) return;
    hipLaunchKernelGGL(spmvKernel<double>, dim3(dieBlocks(S.m)), dim3(TPB), 0, s, topo(S), S.diag, S.ox, S.oy,
                       S.oz, x, y);
}
void spmvF(const PressureSys &S, const float *x, float *y, hipStream_t s) {
    if (S.m == 0) return;
    hipLaunchKernelGGL(spmvKernel<float>, dim3(dieBlocks(S.m)), dim3(TPB), 0, s, topo(S), S.fdiag, S.fox, S.foy,
                       S.foz, x, y);
}

void axpyFloatStep(double *y, const double *x, size_t n, const double *num, const double *den,
                   double sign, hipStream_t s) {
    if (n == 0) return;
    hipLaunchKernelGGL(axpyStepKernel<double>, dim3(blocksFor(n, TPB)), dim3(TPB), 0, s, y, x, n, num, den, sign);
}
void axpyFloatStepF(float *y, const float *x, size_t n, const double *num, const double *den,
                    double sign, hipStream_t s) {
    if (n == 0) return;
    hipLaunchKernelGGL(axpyStepKernel<float>, dim3(blocksFor(n, TPB)), dim3(TPB), 0, s, y, x, n, num, den, sign);
}

template <class T>
static void launchSpmvDot(const PressureSys &S, const T *diag, const T *ox, const T *oy, const T *oz, const T *x,
                          T *y, double *partial, double *sq, hipStream_t st) {
    const int nb = reduceBlocks(S.m);
#if FF_SPMV_WORDS
    WordTopo t{S.m, S.ni, S.nj, S.nw, S.rows, S.words};
#else
    RowTopo t = topo(S);
#endif
    hipLaunchKernelGGL((spmvDotKernel<T, decltype(t)>), dim3(nb), dim3(REDUCE_THREADS), 0, st, t, diag, ox, oy, oz,
                       x, y, partial);
    reduceFinish(partial, nb, sq, false, st);
}
void spmvDot(const PressureSys &S, const double *x, double *y, double *partial, double *sq, hipStream_t st) {
    if (S.m == 0) return;
    launchSpmvDot<double>(S, S.diag, S.ox, S.oy, S.oz, x, y, partial, sq, st);
}
void spmvDotF(const PressureSys &S, const float *x, float *y, double *partial, double *sq, hipStream_t st) {
    if (S.m == 0) return;
    launchSpmvDot<float>(S, S.fdiag, S.fox, S.foy, S.foz, x, y, partial, sq, st);
}

template <class TX, class T>
static void launchPcgStep(TX *x, const T *s, T *r, const T *q, size_t n, const double *num, const double *den,
                          double *partial, double *rmax, hipStream_t st, float *rf = nullptr,
                          double *rmaxHost = nullptr) {
    if (n == 0) return;
    const int nb = reduceBlocks(n);
    hipLaunchKernelGGL((pcgStepKernel<TX, T>), dim3(nb), dim3(REDUCE_THREADS), 0, st, x, s, r, q, n, num, den,
                       partial, rf);
    reduceFinish(partial, nb, rmax, true, st, rmaxHost);
}
void pcgStep(double *x, const double *s, double *r, const double *q, size_t n, const double *num,
             const double *den, double *partial, double *rmax, hipStream_t st, float *rf, double *rmaxHost) {
    launchPcgStep<double, double>(x, s, r, q, n, num, den, partial, rmax, st, rf, rmaxHost);
}
void pcgStepF(float *x, const float *s, float *r, const float *q, size_t n, const double *num,
              const double *den, double *partial, double *rmax, hipStream_t st, double *rmaxHost) {
    launchPcgStep<float, float>(x, s, r, q, n, num, den, partial, rmax, st, nullptr, rmaxHost);
}
void pcgStepMixed(double *x, const float *s, float *r, const float *q, size_t n, const double *num,
                  const double *den, double *partial, double *rmax, hipStream_t st, double *rmaxHost) {
    launchPcgStep<double, float>(x, s, r, q, n, num, den, partial, rmax, st, nullptr, rmaxHost);
}

void widenAxpyStep(double *z, const float *zf, const double *s, size_t n, const double *num, const double *den,
                   hipStream_t st) {
    if (n == 0) return;
    hipLaunchKernelGGL(widenAxpyKernel, dim3(blocksFor(n, TPB)), dim3(TPB), 0, st, z, zf, s, n, num, den);
}

void axpyMixed(double *x, const float *p, size_t n, const double *num, const double *den, double sign,
               hipStream_t s) {
    if (n == 0) return;
    hipLaunchKernelGGL(axpyMixedKernel, dim3(blocksFor(n, TPB)), dim3(TPB), 0, s, x, p, n, num, den, sign);
}

void residual(const PressureSys &S, const double *b, const double *x, double *r, hipStream_t s) {
    if (S.m == 0) return;
    hipLaunchKernelGGL(residualKernel, dim3(dieBlocks(S.m)), dim3(TPB), 0, s, topo(S), S.diag, S.ox, S.oy, S.oz, b,
                       x, r);
}

void pressureToGrid(Grid g, const PressureSys &S, const double *x, float *p, hipStream_t s) {
    size_t n = (size_t)g.ni * g.nj * g.nk;
    hipCheck(hipMemsetAsync(p, 0, sizeof(float) * n, s), "hipMemsetAsync");
    if (S.m > 0) hipLaunchKernelGGL(toGridKernel, dim3(blocksFor(S.m, TPB)), dim3(TPB), 0, s, S.m, S.rows, x, p);
}

void gatherRows(const PressureSys &S, const double *cells, double *rhs, hipStream_t s) {
    if (S.m == 0) return;
    hipLaunchKernelGGL(gatherRowsKernel, dim3(blocksFor(S.m, TPB)), dim3(TPB), 0, s, S.m, S.rows, cells, rhs);
}
void gatherRowsF(const PressureSys &S, const float *cells, double *x, hipStream_t s) {
    if (S.m == 0) return;
    hipLaunchKernelGGL(gatherRowsFKernel, dim3(blocksFor(S.m, TPB)), dim3(TPB), 0, s, S.m, S.rows, cells, x);
}
void scatterRows(const PressureSys &S, const double *x, double *cells, hipStream_t s) {
    if (S.m == 0) return;
    hipLaunchKernelGGL(scatterRowsKernel, dim3(blocksFor(S.m, TPB)), dim3(TPB), 0, s, S.m, S.rows, x, cells);
}
void castToFloat(const double *a, float *b, size_t n, hipStream_t s) {
    if (n == 0) return;
    hipLaunchKernelGGL(castKernel, dim3(blocksFor(n, TPB)), dim3(TPB), 0, s, a, b, n);
}
void addFloatToDouble(double *x, const float *d, size_t n, hipStream_t s) {
    if (n == 0) return;
    hipLaunchKernelGGL(addFloatKernel, dim3(blocksFor(n, TPB)), dim3(TPB), 0, s, x, d, n);
}
void mic0ToFloat(PressureSys &S, hipStream_t s) {
    size_t m = S.m;
    castToFloat(S.invdiag, S.finvdiag, m, s);
    castToFloat(S.lx, S.flx, m, s);
    castToFloat(S.ly, S.fly, m, s);
    castToFloat(S.lz, S.flz, m, s);
    castToFloat(S.diag, S.fdiag, m, s);
    castToFloat(S.ox, S.fox, m, s);
    castToFloat(S.oy, S.foy, m, s);
    castToFloat(S.oz, S.foz, m, s);
    castToFloat(S.mx, S.fmx, m, s);
    castToFloat(S.my, S.fmy, m, s);
    castToFloat(S.mz, S.fmz, m, s);
}

}  // namespace ff
