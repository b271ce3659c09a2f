// kernels_band.hip -- the MIC(0) preconditioner application z = (L L^T)^-1 r
// (solveLower + solveLowerTransposeInPlace, pcgsolver.h:211-241) in bands of
// K consecutive i+j+k hyperplanes per launch.
//
// The forward solve of row (i,j,k) needs the results of (i,j,k-1), (i,j-1,k)
// and (i-1,j,k), all on the previous hyperplane.  A launch handles the band
// of hyperplanes [l0, l0+K).  Each workgroup owns a BT x BT tile of (j, k)
// lines along i and also recomputes the K-1 lines before the tile in -y and
// -z (its halo): a row on band level q depends, through the band, only on
// rows at most K-1-q lines away in -y/-z, and on rows of hyperplane l0-1,
// which earlier launches have finished.  So the workgroup needs no other
// workgroup's results from the same launch: every band level is one barrier,
// the levels exchange results through LDS, and only the tile's own rows are
// written.  Rows recomputed in the halo perform the same operations on the
// same inputs as in their own tile, so they are bit-identical.
//
// The band's cells are staged in LDS first, numbered line by line so that
// consecutive threads read consecutive cells along i: cell ids, row ids and
// the row-indexed operands (rows are in k,j,i order) are read as contiguous
// runs.  Then a thread follows one line along i through the band.
//
// The backward solve mirrors this (halo in +y/+z, hyperplanes descending).
// It reads the forward result from one array and writes to another, so no
// workgroup reads a value another one has already overwritten.
//
// Every row performs the reference's operations in the reference's order
// (forward: subtract the -z, -y, -x terms, then scale by invdiag; backward:
// +x, +y, +z, then scale), so the result equals the serial sweeps bit for bit.
#include "kernels.h"

#include <stdexcept>

namespace ff {

namespace {

template <class T, int K, int BT>
struct Band {
    static constexpr int E = BT + K - 1;                // lines per side, halo included
    static constexpr int LINES = E * E;
    static constexpr int THREADS = (LINES + 63) / 64 * 64;
    static constexpr int P = K + 1;                      // LDS stride per line (odd: no bank conflicts)
    static constexpr int CELLS = LINES * K;
    static constexpr int PER_THREAD = (CELLS + THREADS - 1) / THREADS;
};

// FWD: levels l0 .. l0+K-1, in = r, out = y, rec = invdiag and the L entries
// of the row's -x/-y/-z neighbour rows toward it.  !FWD: levels l0 .. l0-K+1,
// in = y, out = x, rec = invdiag and the row's own L entries toward +x/+y/+z.
// The four operands of a row sit in one record, so a row costs one 16-byte
// (FP32) or two 16-byte (FP64) loads besides its input value.
//
// Two phases.  Staging: the workgroup's band cells are numbered line by line
// (K consecutive cells of one line along i), so consecutive threads read
// consecutive cells and rows; every load of the phase is issued before any is
// used.  Sweep: one thread per line runs the K band levels from LDS.
template <class T, int K, int BT, bool FWD>
__global__ __launch_bounds__((Band<T, K, BT>::THREADS)) void micBandKernel(Grid g, const int *cellRow,
                                                                         const T *in, T *out,
                                                                         const MicRec<T> *rec,
                                                                         const int *tileLo, const int *tileHi,
                                                                         int l0, int tilesJ, TileList list) {
    using B = Band<T, K, BT>;
    constexpr int E = B::E, LINES = B::LINES, P = B::P;
    __shared__ T sIn[LINES * P], sInv[LINES * P], sA[LINES * P], sB[LINES * P], sC[LINES * P];
    __shared__ int sRow[LINES * P];
    __shared__ T val[2][LINES];
    __shared__ int has[2][LINES];
    const int ni = g.ni, nj = g.nj, nk = g.nk;
    // the launch names the tiles with rows in the band; without a list every
    // tile has a workgroup, which exits (uniformly) when none of the tile's own
    // rows lies in the band
    const int lo = FWD ? l0 : l0 - (K - 1), hi = FWD ? l0 + (K - 1) : l0;
    int tile = blockIdx.x;
    if (list.n > 0) {
        tile = list.id[blockIdx.x];
    } else if (tileHi[tile] < lo || tileLo[tile] > hi) {
        return;
    }
    const int J0 = (tile % tilesJ) * BT, K0 = (tile / tilesJ) * BT;
    const int s = FWD ? 1 : -1;  // direction of the sweep along the hyperplanes
    const int jBase = FWD ? J0 - (K - 1) : J0, kBase = FWD ? K0 - (K - 1) : K0;

    // ---- staging: rows, then operands, of every band cell of the workgroup ----
    // A halo line dj + dk lines away from the tile feeds the tile's rows only
    // through band levels q <= K-1-dj-dk; its later cells are left unloaded
    // (absent) and are never read by a row that is.
    int crow[B::PER_THREAD];
#pragma unroll
    for (int t = 0; t < B::PER_THREAD; t++) {
        const int c = threadIdx.x + t * B::THREADS;
        const int line = c / K, q = c % K;
        const int lj = line % E, lk = line / E;
        const int j = jBase + lj, k = kBase + lk;
        const int i = (l0 + s * q) - j - k;
        const int dj = FWD ? (K - 1 - lj > 0 ? K - 1 - lj : 0) : (lj - (BT - 1) > 0 ? lj - (BT - 1) : 0);
        const int dk = FWD ? (K - 1 - lk > 0 ? K - 1 - lk : 0) : (lk - (BT - 1) > 0 ? lk - (BT - 1) : 0);
        const bool ok = c < B::CELLS && q <= K - 1 - dj - dk && j >= 0 && j < nj && k >= 0 && k < nk &&
                        i >= 0 && i < ni;
        const int r = cellRow[ok ? i + (size_t)ni * (j + (size_t)nj * k) : 0];
        crow[t] = ok ? r : -1;
    }
    // the line's neighbours on the hyperplane just outside the band, for its
    // first band level (finished by earlier launches)
    const int tid = threadIdx.x;
    const bool active = tid < LINES;
    const int ej = tid % E, ek = tid / E;
    const int j = jBase + ej, k = kBase + ek;
    const bool lineIn = active && j >= 0 && j < nj && k >= 0 && k < nk;
    const int i0 = l0 - j - k;
    const bool ok0 = lineIn && i0 >= 0 && i0 < ni;
    const size_t c0 = ok0 ? i0 + (size_t)ni * (j + (size_t)nj * k) : 0;
    const int ii = i0 - s, jj = j - s, kk = k - s;
    const bool okx = ok0 && ii >= 0 && ii < ni, oky = ok0 && jj >= 0 && jj < nj, okz = ok0 && kk >= 0 && kk < nk;
    const int r0 = cellRow[c0];
    const int rx = cellRow[okx ? c0 - s : 0];
    const int ry = cellRow[oky ? c0 - s * (size_t)ni : 0];
    const int rz = cellRow[okz ? c0 - s * (size_t)ni * nj : 0];
    const bool first = ok0 && r0 >= 0;
    const bool hx = first && okx && rx >= 0, hy = first && oky && ry >= 0, hz = first && okz && rz >= 0;
    T opIn[B::PER_THREAD];
    MicRec<T> op[B::PER_THREAD];
#pragma unroll
    for (int t = 0; t < B::PER_THREAD; t++) {
        const size_t r = crow[t] >= 0 ? (size_t)crow[t] : 0;  // row 0 stands in for an absent row
        opIn[t] = in[r];
        op[t] = rec[r];
    }
    const T nx = out[hx ? rx : 0], ny = out[hy ? ry : 0], nz = out[hz ? rz : 0];
#pragma unroll
    for (int t = 0; t < B::PER_THREAD; t++) {
        const int c = threadIdx.x + t * B::THREADS;
        if (c < B::CELLS) {
            const int a = (c / K) * P + c % K;
            sRow[a] = crow[t];
            sIn[a] = opIn[t];
            sInv[a] = op[t].inv;
            sA[a] = op[t].a;
            sB[a] = op[t].b;
            sC[a] = op[t].c;
        }
    }
    __syncthreads();

    // ---- sweep: one thread per line, one barrier per band level ----
    const bool owned = FWD ? (ej >= K - 1 && ek >= K - 1) : (ej < BT && ek < BT);
    const bool yin = FWD ? ej > 0 : ej < E - 1;
    const bool zin = FWD ? ek > 0 : ek < E - 1;
    const int cy = tid - s, cz = tid - s * E;
    const int base = tid * P;
    // the line's own operands, read before the first level: a barrier orders
    // LDS accesses, so reads inside the level loop would sit on its critical path
    int lRow[K];
    T lIn[K], lInv[K], lA[K], lB[K], lC[K];
#pragma unroll
    for (int q = 0; q < K; q++) {
        lRow[q] = active ? sRow[base + q] : -1;
        lIn[q] = sIn[base + q];
        lInv[q] = sInv[base + q];
        lA[q] = sA[base + q];
        lB[q] = sB[base + q];
        lC[q] = sC[base + q];
    }
    T prev = T(0);  // this line's result on the previous band level (the x neighbour)
    bool prevHas = false;
#pragma unroll
    for (int q = 0; q < K; q++) {
        const int buf = q & 1, pb = buf ^ 1;
        const int row = lRow[q];
        T v = T(0);
        if (row >= 0) {
            T y = lIn[q];
            const T a = lA[q], b = lB[q], c = lC[q];
            if (FWD) {
                if (q == 0) {
                    if (hz) y -= c * nz;
                    if (hy) y -= b * ny;
                    if (hx) y -= a * nx;
                } else {
                    if (zin && has[pb][cz]) y -= c * val[pb][cz];
                    if (yin && has[pb][cy]) y -= b * val[pb][cy];
                    if (prevHas) y -= a * prev;
                }
            } else {
                if (q == 0) {
                    if (hx) y -= a * nx;
                    if (hy) y -= b * ny;
                    if (hz) y -= c * nz;
                } else {
                    if (prevHas) y -= a * prev;
                    if (yin && has[pb][cy]) y -= b * val[pb][cy];
                    if (zin && has[pb][cz]) y -= c * val[pb][cz];
                }
            }
            v = y * lInv[q];
            if (owned) out[row] = v;
        }
        if (active) {
            val[buf][tid] = v;
            has[buf][tid] = row >= 0;
        }
        prev = v;
        prevHas = row >= 0;
        __syncthreads();
    }
}

// factorModifiedIncompleteColesky0 (pcgsolver.h:68-205) in bands of K
// hyperplanes, with the tiles and halos of the forward sweep above: a row's
// diagonal needs the invdiag of its -z, -y, -x neighbour rows, all on the
// previous hyperplane, so the band scheme applies as it is.  One thread
// follows one line along i; the rows of the previous band level come from
// LDS, those of the hyperplane before the band (earlier launches) from
// global memory.  Each row performs mic0LevelKernel's operations in its
// order, so invdiag is bit-identical to the hyperplane-by-hyperplane factor.
template <int K, int BT>
__global__ __launch_bounds__(((BT + K - 1) * (BT + K - 1) + 63) / 64 * 64) void micFactorBandKernel(
    Grid g, PressureSys S, const int *tileLo, const int *tileHi, int l0, int tilesJ, TileList list) {
    constexpr int E = BT + K - 1, LINES = E * E;
    __shared__ double val[2][LINES];
    __shared__ int rowv[2][LINES];
    const int ni = g.ni, nj = g.nj, nk = g.nk;
    int tile = blockIdx.x;
    if (list.n > 0) {
        tile = list.id[blockIdx.x];
    } else if (tileHi[tile] < l0 || tileLo[tile] > l0 + K - 1) {
        return;
    }
    const int J0 = (tile % tilesJ) * BT, K0 = (tile / tilesJ) * BT;
    const int tid = threadIdx.x;
    const bool active = tid < LINES;
    const int ej = tid % E, ek = tid / E;
    const int j = J0 - (K - 1) + ej, k = K0 - (K - 1) + ek;
    const bool lineIn = active && j >= 0 && j < nj && k >= 0 && k < nk;
    // a halo line dj + dk lines from the tile feeds the tile's rows only
    // through band levels q <= K-1-dj-dk; its later cells are left absent
    const int dj = K - 1 - ej > 0 ? K - 1 - ej : 0, dk = K - 1 - ek > 0 ? K - 1 - ek : 0;
    const bool owned = ej >= K - 1 && ek >= K - 1;
    double prev = 0.0;
    int prevRow = -1;
    for (int q = 0; q < K; q++) {
        const int buf = q & 1, pb = buf ^ 1;
        const int i = l0 + q - j - k;
        const bool ok = lineIn && q <= K - 1 - dj - dk && i >= 0 && i < ni;
        const size_t c = ok ? i + (size_t)ni * (j + (size_t)nj * k) : 0;
        const int row = ok ? S.cellRow[c] : -1;
        double inv = 0.0;
        if (row >= 0) {
            // the -z, -y, -x neighbour rows and their invdiag; a present row
            // past the band's first level has both neighbour lines inside the
            // workgroup (ek, ej >= 1)
            int rz, ry, rx;
            double vz = 0.0, vy = 0.0, vx = 0.0;
            if (q == 0) {
                rz = k > 0 ? S.cellRow[c - (size_t)ni * nj] : -1;
                ry = j > 0 ? S.cellRow[c - ni] : -1;
                rx = i > 0 ? S.cellRow[c - 1] : -1;
                if (rz >= 0) vz = S.invdiag[rz];
                if (ry >= 0) vy = S.invdiag[ry];
                if (rx >= 0) vx = S.invdiag[rx];
            } else {
                rz = rowv[pb][tid - E];
                vz = val[pb][tid - E];
                ry = rowv[pb][tid - 1];
                vy = val[pb][tid - 1];
                rx = prevRow;
                vx = prev;
            }
            double a = S.diag[row];
            if (a < 1e-9) a = 0;
            double d = a;
            const int kr[3] = {rz, ry, rx};
            const double kv[3] = {vz, vy, vx};
            const int jslots[3] = {5, 3, 1};
#pragma unroll
            for (int t = 0; t < 3; t++) {
                const int kk = kr[t];
                if (kk < 0) continue;
                if (S.diag[kk] < 1e-9) continue;
                const int jslot = jslots[t];
                const int *kb = S.nbr + (size_t)6 * kk;
                const double ik = kv[t];
                const double okx = S.ox[kk], oky = S.oy[kk], okz = S.oz[kk];
                const double Ljk = (jslot == 1 ? okx : (jslot == 3 ? oky : okz)) * ik;
                d -= Ljk * Ljk;
                double missing = 0;
                if (jslot != 1 && kb[1] >= 0) missing += okx * ik;
                if (jslot != 3 && kb[3] >= 0) missing += oky * ik;
                if (jslot != 5 && kb[5] >= 0) missing += okz * ik;
                d -= 0.97 * Ljk * missing;
            }
            if (a == 0)
                inv = d;
            else if (d < 0.25 * a)
                inv = 1 / sqrt(a);
            else
                inv = 1 / sqrt(d);
            if (owned) S.invdiag[row] = inv;
        }
        if (active) {
            val[buf][tid] = inv;
            rowv[buf][tid] = row;
        }
        prev = inv;
        prevRow = row;
        __syncthreads();
    }
}

// L entry of each row's -x/-y/-z neighbour row toward the row (0 when
// absent), and the band sweeps' records: forward {invdiag, those three},
// backward {invdiag, the row's own lx, ly, lz}
template <class T>
__global__ void lowerEntriesKernel(int m, const int *nbr, const T *inv, const T *lx, const T *ly, const T *lz,
                                   T *lmx, T *lmy, T *lmz, MicRec<T> *recFwd, MicRec<T> *recBwd) {
    int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= m) return;
    const int *nb = nbr + (size_t)6 * r;
    const T a = nb[0] >= 0 ? lx[nb[0]] : T(0);
    const T b = nb[2] >= 0 ? ly[nb[2]] : T(0);
    const T c = nb[4] >= 0 ? lz[nb[4]] : T(0);
    lmx[r] = a;
    lmy[r] = b;
    lmz[r] = c;
    const T d = inv[r];
    recFwd[r] = MicRec<T>{d, a, b, c};
    recBwd[r] = MicRec<T>{d, lx[r], ly[r], lz[r]};
}

// lowest and highest hyperplane of the rows in each tile of BT x BT lines:
// one workgroup per tile reads the tile's cells line by line (no atomics)
__global__ __launch_bounds__(256) void tileRangeKernel(Grid g, const int *cellRow, int bt, int tilesJ, int *tileLo,
                                                       int *tileHi) {
    __shared__ int sLo[256], sHi[256];
    const int ni = g.ni, nj = g.nj, nk = g.nk;
    const int J0 = (blockIdx.x % tilesJ) * bt, K0 = (blockIdx.x / tilesJ) * bt;
    const int lj = nj - J0 < bt ? nj - J0 : bt, lk = nk - K0 < bt ? nk - K0 : bt;
    const long long cells = (long long)lj * lk * ni;
    int lo = 0x7fffffff, hi = -1;
    for (long long q = threadIdx.x; q < cells; q += blockDim.x) {
        const int i = (int)(q % ni);
        const int line = (int)(q / ni);
        const int j = J0 + line % lj, k = K0 + line / lj;
        if (cellRow[i + (size_t)ni * (j + (size_t)nj * k)] >= 0) {
            lo = min(lo, i + j + k);
            hi = max(hi, i + j + k);
        }
    }
    sLo[threadIdx.x] = lo;
    sHi[threadIdx.x] = hi;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < (unsigned)o) {
            sLo[threadIdx.x] = min(sLo[threadIdx.x], sLo[threadIdx.x + o]);
            sHi[threadIdx.x] = max(sHi[threadIdx.x], sHi[threadIdx.x + o]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        tileLo[blockIdx.x] = sLo[0];
        tileHi[blockIdx.x] = sHi[0];
    }
}

template <class T, int K, int BT>
void bandSweeps(Grid g, const PressureSys &S, const BandPlan &plan, const T *r, T *y, T *z,
                const MicRec<T> *recFwd, const MicRec<T> *recBwd, const int *tileLo, const int *tileHi,
                hipStream_t st) {
    const int tilesJ = (g.nj + BT - 1) / BT, tilesK = (g.nk + BT - 1) / BT;
    const dim3 block(Band<T, K, BT>::THREADS);
    for (size_t q = 0; q < plan.l0.size(); q++) {
        const TileList &tl = plan.list[q];
        const dim3 grid(tl.n > 0 ? tl.n : tilesJ * tilesK);
        if ((int)q < plan.nFwd)
            hipLaunchKernelGGL((micBandKernel<T, K, BT, true>), grid, block, 0, st, g, S.cellRow, r, y, recFwd,
                               tileLo, tileHi, plan.l0[q], tilesJ, tl);
        else
            hipLaunchKernelGGL((micBandKernel<T, K, BT, false>), grid, block, 0, st, g, S.cellRow, y, z, recBwd,
                               tileLo, tileHi, plan.l0[q], tilesJ, tl);
    }
    hipCheck(hipGetLastError(), "MIC(0) band kernel launch");
}

// launches of the two sweeps: the bands of K hyperplanes that hold rows of a
// tile (all bands that hold rows when the tile ranges are unknown)
template <int K, int BT>
void bandPlan(Grid g, const PressureSys &S, const int *hs, const int *hLo, const int *hHi, BandPlan &plan) {
    const int tilesJ = (g.nj + BT - 1) / BT, tilesK = (g.nk + BT - 1) / BT, nt = tilesJ * tilesK;
    const bool ranges = hLo && hHi && nt <= 65535;
    const int nl = S.nLevels;
    plan.l0.clear();
    plan.list.clear();
    auto add = [&](int l0, int lo, int hi) {
        TileList tl;
        tl.n = 0;
        if (ranges) {
            int count = 0;
            for (int t = 0; t < nt; t++) count += !(hHi[t] < lo || hLo[t] > hi);
            if (count == 0) return;
            if (count <= TileList::MAX)
                for (int t = 0; t < nt; t++)
                    if (!(hHi[t] < lo || hLo[t] > hi)) tl.id[tl.n++] = (uint16_t)t;
        }
        plan.l0.push_back(l0);
        plan.list.push_back(tl);
    };
    // forward, bands [l0, l0+K) that hold rows
    for (int l0 = 0; l0 < nl; l0 += K) {
        const int l1 = l0 + K < nl ? l0 + K : nl;
        if (hs[l1] > hs[l0]) add(l0, l0, l0 + K - 1);
    }
    plan.nFwd = (int)plan.l0.size();
    // backward, bands (l0-K, l0] from the top
    for (int l0 = nl - 1; l0 >= 0; l0 -= K) {
        const int lb = l0 - K + 1 > 0 ? l0 - K + 1 : 0;
        if (hs[l0 + 1] > hs[lb]) add(l0, l0 - K + 1, l0);
    }
}

// band depth K and tile side BT: the staged band, (BT+K-1)^2 lines of K
// cells, must fit in LDS
#ifndef FF_BAND64_K
#define FF_BAND64_K 8
#endif
#ifndef FF_BAND64_BT
#define FF_BAND64_BT 11
#endif
#ifndef FF_BAND32_K
#define FF_BAND32_K 12
#endif
#ifndef FF_BAND32_BT
#define FF_BAND32_BT 11
#endif
constexpr int K64 = FF_BAND64_K, BT64 = FF_BAND64_BT, K32 = FF_BAND32_K, BT32 = FF_BAND32_BT;

}  // namespace

void mic0LowerEntries(const PressureSys &S, bool fp64, hipStream_t s) {
    if (S.m == 0) return;
    dim3 grid((S.m + 255) / 256), block(256);
    if (fp64)
        hipLaunchKernelGGL(lowerEntriesKernel<double>, grid, block, 0, s, S.m, S.nbr, S.invdiag, S.lx, S.ly, S.lz,
                           S.lmx, S.lmy, S.lmz, S.recFwd, S.recBwd);
    else
        hipLaunchKernelGGL(lowerEntriesKernel<float>, grid, block, 0, s, S.m, S.nbr, S.finvdiag, S.flx, S.fly,
                           S.flz, S.flmx, S.flmy, S.flmz, S.frecFwd, S.frecBwd);
}

void mic0BandPlan(Grid g, const PressureSys &S, const int *hLevelStart, bool fp64, const int *hTileLo,
                  const int *hTileHi, BandPlan &plan) {
    plan.fp64 = fp64;
    if (fp64)
        bandPlan<K64, BT64>(g, S, hLevelStart, hTileLo, hTileHi, plan);
    else
        bandPlan<K32, BT32>(g, S, hLevelStart, hTileLo, hTileHi, plan);
}

void mic0BandShape(bool fp64, int &levels, int &tile) {
    levels = fp64 ? K64 : K32;
    tile = fp64 ? BT64 : BT32;
}

int mic0BandTiles(Grid g, bool fp64) {
    const int bt = fp64 ? BT64 : BT32;
    return ((g.nj + bt - 1) / bt) * ((g.nk + bt - 1) / bt);
}

void mic0BandTileRanges(Grid g, const PressureSys &S, bool fp64, int *tileLo, int *tileHi, hipStream_t s) {
    const int bt = fp64 ? BT64 : BT32, tilesJ = (g.nj + bt - 1) / bt;
    const int nt = mic0BandTiles(g, fp64);
    hipLaunchKernelGGL(tileRangeKernel, dim3(nt), dim3(256), 0, s, g, S.cellRow, bt, tilesJ, tileLo, tileHi);
}

void mic0ApplyBands(Grid g, const PressureSys &S, const BandPlan &plan, const int *tileLo, const int *tileHi,
                    const double *r, double *y, double *z, hipStream_t s) {
    if (S.m == 0) return;
    if (!plan.fp64) throw std::runtime_error("MIC(0) bands: the launch plan was built for FP32");
    bandSweeps<double, K64, BT64>(g, S, plan, r, y, z, S.recFwd, S.recBwd, tileLo, tileHi, s);
}

void mic0ApplyBandsF(Grid g, const PressureSys &S, const BandPlan &plan, const int *tileLo, const int *tileHi,
                     const float *r, float *y, float *z, hipStream_t s) {
    if (S.m == 0) return;
    if (plan.fp64) throw std::runtime_error("MIC(0) bands: the launch plan was built for FP64");
    bandSweeps<float, K32, BT32>(g, S, plan, r, y, z, S.frecFwd, S.frecBwd, tileLo, tileHi, s);
}

constexpr int KF = 13;  // factor band depth: 59 launches for the 766 hyperplanes of 256^3

void mic0FactorBands(Grid g, const PressureSys &S, const int *hs, const int *hTileLo, const int *hTileHi,
                     const int *tileLo, const int *tileHi, hipStream_t s) {
    if (S.m == 0) return;
    // the tile ranges are those of the sweeps' tiles (both precisions use BT64)
    static_assert(BT64 == BT32, "one set of tile ranges serves both sweeps and the factor");
    const int tilesJ = (g.nj + BT64 - 1) / BT64, tilesK = (g.nk + BT64 - 1) / BT64, nt = tilesJ * tilesK;
    const bool ranges = hTileLo && hTileHi && nt <= 65535;
    const dim3 block(((BT64 + KF - 1) * (BT64 + KF - 1) + 63) / 64 * 64);
    for (int l0 = 0; l0 < S.nLevels; l0 += KF) {
        const int l1 = l0 + KF < S.nLevels ? l0 + KF : S.nLevels;
        if (hs[l1] == hs[l0]) continue;
        TileList tl;
        tl.n = 0;
        if (ranges) {
            int count = 0;
            for (int t = 0; t < nt; t++) count += !(hTileHi[t] < l0 || hTileLo[t] > l0 + KF - 1);
            if (count == 0) continue;
            if (count <= TileList::MAX)
                for (int t = 0; t < nt; t++)
                    if (!(hTileHi[t] < l0 || hTileLo[t] > l0 + KF - 1)) tl.id[tl.n++] = (uint16_t)t;
        }
        const dim3 grid(tl.n > 0 ? tl.n : nt);
        hipLaunchKernelGGL((micFactorBandKernel<KF, BT64>), grid, block, 0, s, g, S, tileLo, tileHi, l0, tilesJ, tl);
    }
    hipCheck(hipGetLastError(), "MIC(0) factor band launch");
}

int mic0FactorBandLaunches(const PressureSys &S, const int *hs) {
    int n = 0;
    for (int l0 = 0; l0 < S.nLevels; l0 += KF) n += hs[(l0 + KF < S.nLevels ? l0 + KF : S.nLevels)] > hs[l0];
    return n;
}

}  // namespace ff
