// kernels_mic.hip -- the MIC(0) preconditioner application z = (L L^T)^-1 r
// (solveLower + solveLowerTransposeInPlace, pcgsolver.h:211-241) as ONE
// persistent, tile-pipelined launch.
//
// Dependency structure.  Row (i,j,k) of the forward solve needs the results of
// its -x, -y, -z neighbours; the backward solve needs +x, +y, +z.  The grid is
// cut into tiles of 16 x 16 x TK cells.  A tile's forward solve needs only
// the forward results of its -x, -y, -z neighbour tiles (their face cells);
// its backward solve needs its own forward results and the backward results
// of its +x, +y, +z neighbour tiles.
//
// Schedule.  Non-empty tiles are listed by tile level I+J+K.  A launch hands
// out 2T tickets from one device counter: ticket t < T is the forward solve of
// tile order[t], ticket T+q the backward solve of order[T-1-q].  A workgroup
// only ever waits for tiles with smaller tickets, which were taken earlier by
// workgroups that are running, so the launch drains for any grid size and any
// number of resident workgroups.  Every wait is bounded in wall-clock time:
// past the limit the workgroup raises the error word, every other wait sees
// it and gives up, and the host turns it into an error.
//
// Inside a tile.  One thread per (i, j) column; the tile's inputs are staged
// in LDS, then the tile runs its local hyperplanes li+lj+lk in order (reverse
// order for the backward solve), one barrier per hyperplane, exchanging the
// previous hyperplane's results through a double-buffered LDS row.  Results
// go to global memory as they are produced; an agent-scope release and an
// atomic exchange of the tile's flag publish the tile, and the waiting side's
// agent-scope acquire makes the tile's results visible to it.
//
// Every row performs the reference's operations in the reference's order
// (forward: subtract the -z, -y, -x terms, then scale by invdiag; backward:
// +x, +y, +z, then scale), so the result equals the serial sweeps bit for bit.
#include <stdexcept>

#include "kernels.h"

namespace ff {

namespace {

constexpr int TW = 16;         // tile extent along i and j
constexpr int TPT = TW * TW;   // threads per workgroup: one per (i, j) column

template <class T>
struct MicArgs {
    Grid g;
    const int *cellRow;           // N: pressure row of the cell, or -1
    const int *nbr;               // 6 per row: -x,+x,-y,+y,-z,+z rows or -1
    const T *r;                   // forward right-hand side
    T *z;                         // forward result, then backward result in place
    const T *invdiag;
    const T *lx, *ly, *lz;        // L entries of the row toward its +x/+y/+z rows
    const T *lmx, *lmy, *lmz;     // L entries of the row's -x/-y/-z rows toward it
    const int *order;             // non-empty tiles by tile level
    const int *tileRows;          // per tile: non-zero when it holds rows
    int *flags;                   // per tile: last published target
    int *counter;                 // ticket counter
    int *error;                   // raised when a wait exceeds the spin limit
    int nActive, ti, tj, tk;
    int base;                     // counter value at the start of this launch
    int fwdTarget, bwdTarget;
    long long waitTicks;          // longest wait, in wall-clock ticks
};

// Flags and the error word are read and written only with read-modify-write
// atomics, which the device performs at one coherence point for all XCDs;
// a plain load could keep returning a stale copy held in another XCD's L2.
// Both are never negative, so max(x, 0) reads without changing them (an
// add of 0 would be folded into a plain load by the compiler).
__device__ inline int pollFlag(int *p) {
    return __hip_atomic_fetch_max(p, 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

template <class T, int TK>
__global__ __launch_bounds__(TPT) void micTileKernel(MicArgs<T> a) {
    __shared__ T sv[TK * TPT], sinv[TK * TPT], sa[TK * TPT], sb[TK * TPT], sc[TK * TPT];
    __shared__ int srow[TK * TPT];
    __shared__ T zb[2][TPT];
    __shared__ unsigned smask[TPT];
    __shared__ int bnx[TW * TK], bny[TW * TK], bnz[TPT];
    __shared__ T bzx[TW * TK], bzy[TW * TK], bzz[TPT];
    __shared__ int sTicket, sAbort;

    const int col = threadIdx.x, li = col % TW, lj = col / TW;
    const int ni = a.g.ni, nj = a.g.nj, nk = a.g.nk;
    const int T2 = 2 * a.nActive;
    const int nlev = 2 * (TW - 1) + TK;
    const int s = li + lj;

    for (;;) {
        if (col == 0) {
            sTicket = atomicAdd(a.counter, 1) - a.base;
            sAbort = 0;
        }
        __syncthreads();
        // The ticket and the abort word are the same for the whole workgroup;
        // readfirstlane makes that visible to the compiler, so every exit of
        // this loop is a uniform branch.  Read as per-lane values, the loop
        // exits would look divergent and the compiler may let every lane but
        // the one that takes the next ticket run on with the old one.
        const int t = __builtin_amdgcn_readfirstlane(sTicket);
        // t < 0 only after an aborted launch left the counter short; the
        // error word is already raised then
        if (t < 0 || t >= T2) break;
        const bool fwd = t < a.nActive;
        const int tile = a.order[fwd ? t : T2 - 1 - t];
        const int I = tile % a.ti, J = (tile / a.ti) % a.tj, K = tile / (a.ti * a.tj);
        const int gi = I * TW + li, gj = J * TW + lj, k0 = K * TK;
        const bool colIn = gi < ni && gj < nj;

        // ---- stage this tile's own inputs (independent of other tiles) ----
        unsigned mask = 0;
        for (int lk = 0; lk < TK; lk++) {
            const int gk = k0 + lk;
            int row = -1;
            if (colIn && gk < nk) row = a.cellRow[gi + (size_t)ni * (gj + (size_t)nj * gk)];
            const int c = lk * TPT + col;
            srow[c] = row;
            if (row >= 0) {
                mask |= 1u << lk;
                sinv[c] = a.invdiag[row];
                if (fwd) {
                    sv[c] = a.r[row];
                    sa[c] = a.lmx[row];
                    sb[c] = a.lmy[row];
                    sc[c] = a.lmz[row];
                } else {
                    sa[c] = a.lx[row];
                    sb[c] = a.ly[row];
                    sc[c] = a.lz[row];
                }
            }
            const int *nb = a.nbr + (size_t)6 * (row >= 0 ? row : 0);
            if (fwd) {
                if (li == 0) bnx[lj * TK + lk] = row >= 0 ? nb[0] : -1;
                if (lj == 0) bny[li * TK + lk] = row >= 0 ? nb[2] : -1;
                if (lk == 0) bnz[col] = row >= 0 ? nb[4] : -1;
            } else {
                if (li == TW - 1) bnx[lj * TK + lk] = row >= 0 ? nb[1] : -1;
                if (lj == TW - 1) bny[li * TK + lk] = row >= 0 ? nb[3] : -1;
                if (lk == TK - 1) bnz[col] = row >= 0 ? nb[5] : -1;
            }
        }
        smask[col] = mask;

        // ---- wait for the tiles this one reads ----
        if (col < 4) {
            int dep = -1, target = fwd ? a.fwdTarget : a.bwdTarget;
            if (fwd) {
                if (col == 0 && I > 0) dep = tile - 1;
                if (col == 1 && J > 0) dep = tile - a.ti;
                if (col == 2 && K > 0) dep = tile - a.ti * a.tj;
            } else {
                if (col == 0 && I < a.ti - 1) dep = tile + 1;
                if (col == 1 && J < a.tj - 1) dep = tile + a.ti;
                if (col == 2 && K < a.tk - 1) dep = tile + a.ti * a.tj;
                if (col == 3) {
                    dep = tile;  // this tile's own forward solve
                    target = a.fwdTarget;
                }
            }
            if (dep >= 0 && a.tileRows[dep] != 0) {
                const long long start = (long long)wall_clock64();
                while (pollFlag(&a.flags[dep]) < target) {
                    if (pollFlag(a.error) != 0 || (long long)wall_clock64() - start > a.waitTicks) {
                        __hip_atomic_fetch_or(a.error, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                        sAbort = 1;
                        break;
                    }
                    __builtin_amdgcn_s_sleep(2);
                }
            }
        }
        __syncthreads();
        if (__builtin_amdgcn_readfirstlane(sAbort)) break;
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");

        // ---- inputs produced by other workgroups ----
        if (!fwd) {
            for (int lk = 0; lk < TK; lk++) {
                const int c = lk * TPT + col;
                if ((mask >> lk) & 1u) sv[c] = a.z[srow[c]];
            }
        }
        for (int q = col; q < 2 * TW * TK + TPT; q += TPT) {
            if (q < TW * TK) {
                const int id = bnx[q];
                bzx[q] = id >= 0 ? a.z[id] : T(0);
            } else if (q < 2 * TW * TK) {
                const int id = bny[q - TW * TK];
                bzy[q - TW * TK] = id >= 0 ? a.z[id] : T(0);
            } else {
                const int id = bnz[q - 2 * TW * TK];
                bzz[q - 2 * TW * TK] = id >= 0 ? a.z[id] : T(0);
            }
        }
        __syncthreads();

        // ---- the tile's hyperplanes ----
        T own = T(0);  // this column's result on the previous hyperplane
        if (fwd) {
            for (int l = 0; l < nlev; l++) {
                const int lk = l - s;
                if (lk >= 0 && lk < TK && ((mask >> lk) & 1u)) {
                    const int c = lk * TPT + col;
                    const T *zp = zb[(l + 1) & 1];
                    T y = sv[c];
                    if (lk > 0) {
                        if ((mask >> (lk - 1)) & 1u) y -= sc[c] * own;
                    } else if (bnz[col] >= 0) {
                        y -= sc[c] * bzz[col];
                    }
                    if (lj > 0) {
                        if ((smask[col - TW] >> lk) & 1u) y -= sb[c] * zp[col - TW];
                    } else if (bny[li * TK + lk] >= 0) {
                        y -= sb[c] * bzy[li * TK + lk];
                    }
                    if (li > 0) {
                        if ((smask[col - 1] >> lk) & 1u) y -= sa[c] * zp[col - 1];
                    } else if (bnx[lj * TK + lk] >= 0) {
                        y -= sa[c] * bzx[lj * TK + lk];
                    }
                    const T zv = y * sinv[c];
                    a.z[srow[c]] = zv;
                    zb[l & 1][col] = zv;
                    own = zv;
                }
                __syncthreads();
            }
        } else {
            for (int l = nlev - 1; l >= 0; l--) {
                const int lk = l - s;
                if (lk >= 0 && lk < TK && ((mask >> lk) & 1u)) {
                    const int c = lk * TPT + col;
                    const T *zp = zb[(l + 1) & 1];
                    T x = sv[c];
                    if (li < TW - 1) {
                        if ((smask[col + 1] >> lk) & 1u) x -= sa[c] * zp[col + 1];
                    } else if (bnx[lj * TK + lk] >= 0) {
                        x -= sa[c] * bzx[lj * TK + lk];
                    }
                    if (lj < TW - 1) {
                        if ((smask[col + TW] >> lk) & 1u) x -= sb[c] * zp[col + TW];
                    } else if (bny[li * TK + lk] >= 0) {
                        x -= sb[c] * bzy[li * TK + lk];
                    }
                    if (lk < TK - 1) {
                        if ((mask >> (lk + 1)) & 1u) x -= sc[c] * own;
                    } else if (bnz[col] >= 0) {
                        x -= sc[c] * bzz[col];
                    }
                    const T zv = x * sinv[c];
                    a.z[srow[c]] = zv;
                    zb[l & 1][col] = zv;
                    own = zv;
                }
                __syncthreads();
            }
        }

        // ---- publish ----
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        __syncthreads();
        if (col == 0)
            __hip_atomic_exchange(&a.flags[tile], fwd ? a.fwdTarget : a.bwdTarget, __ATOMIC_RELEASE,
                                  __HIP_MEMORY_SCOPE_AGENT);
    }
}

__global__ void tileMarkKernel(Grid g, int m, const int *rows, int tk, int ti, int tj, int *tileRows) {
    int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= m) return;
    int c = rows[r];
    int i = c % g.ni, j = (c / g.ni) % g.nj, k = c / (g.ni * g.nj);
    tileRows[i / TW + ti * (j / TW + tj * (k / tk))] = 1;
}

}  // namespace

int micTileDepth(bool fp64) { return fp64 ? 12 : 16; }

// LDS of one workgroup: five staged values and a row id per cell, the
// double-buffered hyperplane, column masks and the three face buffers
size_t micTileLds(bool fp64) {
    const size_t tk = micTileDepth(fp64), v = fp64 ? 8 : 4;
    return 5 * v * tk * TPT + 4 * tk * TPT + 2 * v * TPT + 4 * TPT + (4 + v) * (2 * TW * tk + TPT) + 8;
}

void micTileCounts(Grid g, int tk, int &ti, int &tj, int &tkc) {
    ti = (g.ni + TW - 1) / TW;
    tj = (g.nj + TW - 1) / TW;
    tkc = (g.nk + tk - 1) / tk;
}

void micMarkTiles(Grid g, const PressureSys &S, int tk, int *tileRows, hipStream_t s) {
    int ti, tj, tkc;
    micTileCounts(g, tk, ti, tj, tkc);
    hipCheck(hipMemsetAsync(tileRows, 0, sizeof(int) * (size_t)ti * tj * tkc, s), "memset tiles");
    if (S.m > 0)
        hipLaunchKernelGGL(tileMarkKernel, dim3((S.m + 255) / 256), dim3(256), 0, s, g, S.m, S.rows, tk, ti, tj,
                           tileRows);
}

template <class T, int TK>
static void launchTiles(const PressureSys &S, const MicSchedule &ms, const T *r, T *z, const T *inv,
                        const T *lx, const T *ly, const T *lz, const T *lmx, const T *lmy, const T *lmz,
                        hipStream_t s) {
    MicArgs<T> a;
    a.g = ms.g;
    a.cellRow = S.cellRow;
    a.nbr = S.nbr;
    a.r = r;
    a.z = z;
    a.invdiag = inv;
    a.lx = lx;
    a.ly = ly;
    a.lz = lz;
    a.lmx = lmx;
    a.lmy = lmy;
    a.lmz = lmz;
    a.order = ms.order;
    a.tileRows = ms.tileRows;
    a.flags = ms.flags;
    a.counter = ms.counter;
    a.error = ms.error;
    a.nActive = ms.nActive;
    a.ti = ms.ti;
    a.tj = ms.tj;
    a.tk = ms.tk;
    a.base = ms.base;
    a.fwdTarget = 2 * ms.epoch - 1;
    a.bwdTarget = 2 * ms.epoch;
    a.waitTicks = ms.waitTicks;
    hipLaunchKernelGGL((micTileKernel<T, TK>), dim3(ms.grid), dim3(TPT), 0, s, a);
    hipCheck(hipGetLastError(), "MIC(0) tile kernel launch");
}

// The tile-wavefront application is switched off.  Under test it never
// finished: the compiler split the ticket loop so that after a tile's publish
// barrier only the lane that takes the next ticket left the inner loop, and
// the workgroup's other lanes ran on with the previous ticket.  It stays off
// until a version whose compiled loop is checked to be uniform has run.
static void tilesOff() {
    throw std::runtime_error("the tile-wavefront MIC(0) kernel is switched off; use the hyperplane schedule");
}

void micApplyTiles(const PressureSys &S, MicSchedule &ms, const double *r, double *z, hipStream_t s) {
    tilesOff();
    if (ms.nActive == 0) return;
    ms.epoch++;
    launchTiles<double, 12>(S, ms, r, z, S.invdiag, S.lx, S.ly, S.lz, S.lmx, S.lmy, S.lmz, s);
    ms.base += 2 * ms.nActive + ms.grid;
}

void micApplyTilesF(const PressureSys &S, MicSchedule &ms, const float *r, float *z, hipStream_t s) {
    tilesOff();
    if (ms.nActive == 0) return;
    ms.epoch++;
    launchTiles<float, 16>(S, ms, r, z, S.finvdiag, S.flx, S.fly, S.flz, S.flmx, S.flmy, S.flmz, s);
    ms.base += 2 * ms.nActive + ms.grid;
}

}  // namespace ff
