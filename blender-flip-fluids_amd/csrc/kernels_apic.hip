// kernels_apic.hip -- the APIC particle-to-grid transfer as a deterministic
// gather.
//
// VelocityAdvector with the APIC producer (velocityadvector.cpp:38-465,
// 543-623) works per 10^3 block of one face array.  A particle belongs to
// every active block within its radius (+1e-6) and, in the frame of each
// such block, deposits w (v + a . (node - p)) onto the 8 nodes of its cell
// that lie inside the block, w being the trilinear weight.  Here each
// (particle, block) membership becomes a record keyed by the block and the
// particle's cell in that block's frame (computed with the reference's own
// float arithmetic), the records are counting-sorted by key (ascending
// particle index inside a key), and one thread per face merges its 8 cells'
// lists in particle order (staged in LDS per quarter block).  The sums
// therefore run in the reference's order and no float atomics are used.
#include "kernels.h"

namespace ff {

namespace {

constexpr int TPB = 256;
constexpr int BW = P2G_BLOCK;          // block width in faces
constexpr int LC = BW + 2;             // local cells per axis: -1 .. BW
constexpr int CELLS_PER_BLOCK = LC * LC * LC;
constexpr int MAX_MEMBERS = 8;         // the radius is below a block width: two blocks per axis at most

unsigned blocksFor(size_t n, unsigned t) { return (unsigned)((n + t - 1) / t); }

struct FaceDims {
    int wi, wj, wk, bi, bj, bk;
};
FF_HD FaceDims faceDims(Grid g, int dir) {
    FaceDims d;
    d.wi = g.ni + (dir == 0);
    d.wj = g.nj + (dir == 1);
    d.wk = g.nk + (dir == 2);
    d.bi = (d.wi + BW - 1) / BW;
    d.bj = (d.wj + BW - 1) / BW;
    d.bk = (d.wk + BW - 1) / BW;
    return d;
}
FF_HD V3 dirOffset(Grid g, int dir) {
    const double h = 0.5 * g.dx;
    return dir == 0 ? mk(0.0, h, h) : (dir == 1 ? mk(h, 0.0, h) : mk(h, h, 0.0));
}

// _initializeActiveBlocksThread (velocityadvector.cpp:239-251): the block of
// each particle (shifted by the direction offset) is active
__global__ void markBlocksKernel(Grid g, int dir, int np, const float *px, const float *py, const float *pz,
                                 uint8_t *mark) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= np) return;
    const FaceDims D = faceDims(g, dir);
    const V3 x = sub(V3{px[q], py[q], pz[q]}, dirOffset(g, dir));
    const I3 b = posToIndex(x, g.dx * BW);
    if (inRange(b.i, b.j, b.k, D.bi, D.bj, D.bk)) mark[flat(b.i, b.j, b.k, D.bi, D.bj)] = 1;
}

// GridUtils::featherGrid26: active when any of the 27 blocks around is marked
__global__ void featherKernel(int bi, int bj, int bk, const uint8_t *mark, uint8_t *active) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= bi * bj * bk) return;
    const int i = b % bi, j = (b / bi) % bj, k = b / (bi * bj);
    uint8_t a = 0;
    for (int dk = -1; dk <= 1; dk++)
        for (int dj = -1; dj <= 1; dj++)
            for (int di = -1; di <= 1; di++)
                if (inRange(i + di, j + dj, k + dk, bi, bj, bk) && mark[flat(i + di, j + dj, k + dk, bi, bj)]) a = 1;
    active[b] = a;
}

// _computeGridCountDataThread (velocityadvector.cpp:296-353): the member
// blocks of particle q, then its cell in each member block's frame.  Record
// 8q + m holds the key of the m-th member or -1.
__global__ void apicKeysKernel(Grid g, int dir, int np, const float *px, const float *py, const float *pz,
                               double radius, const uint8_t *active, int *keys) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= np) return;
    const FaceDims D = faceDims(g, dir);
    const V3 x = sub(V3{px[q], py[q], pz[q]}, dirOffset(g, dir));
    const float eps = 1e-6f;
    const float sr = (float)(radius + (double)eps);
    const float blockdx = (float)(g.dx * BW);
    const I3 b0 = posToIndex(x, blockdx);
    const V3 bp = indexToPos(b0.i, b0.j, b0.k, blockdx);
    int lo[3], hi[3];
    if (x.x - sr > bp.x && x.y - sr > bp.y && x.z - sr > bp.z && x.x + sr < bp.x + blockdx &&
        x.y + sr < bp.y + blockdx && x.z + sr < bp.z + blockdx) {
        lo[0] = hi[0] = b0.i;
        lo[1] = hi[1] = b0.j;
        lo[2] = hi[2] = b0.k;
    } else {
        const I3 g0 = posToIndexD(x.x - sr, x.y - sr, x.z - sr, blockdx);
        const I3 g1 = posToIndexD(x.x + sr, x.y + sr, x.z + sr, blockdx);
        lo[0] = g0.i; hi[0] = g1.i;
        lo[1] = g0.j; hi[1] = g1.j;
        lo[2] = g0.k; hi[2] = g1.k;
    }
    int m = 0;
    for (int k = lo[2]; k <= hi[2]; k++)
        for (int j = lo[1]; j <= hi[1]; j++)
            for (int i = lo[0]; i <= hi[0]; i++) {
                if (!inRange(i, j, k, D.bi, D.bj, D.bk) || m >= MAX_MEMBERS) continue;
                const size_t bid = flat(i, j, k, D.bi, D.bj);
                if (!active[bid]) continue;
                const V3 pl = sub(x, indexToPos(i, j, k, BW * g.dx));
                const I3 c = posToIndex(pl, g.dx);
                // the radius keeps a member particle within one cell of its block
                if (c.i < -1 || c.j < -1 || c.k < -1 || c.i > BW || c.j > BW || c.k > BW) continue;
                const int key = (int)(bid * CELLS_PER_BLOCK) + (c.i + 1) + LC * ((c.j + 1) + LC * (c.k + 1));
                keys[(size_t)q * MAX_MEMBERS + m] = key;
                m++;
            }
    for (; m < MAX_MEMBERS; m++) keys[(size_t)q * MAX_MEMBERS + m] = -1;
}

// One workgroup per 10 x 5 x 5 quarter of a block: it stages the records of
// the 11 x 6 x 6 cells its faces read (36 rows of 11 cells, each row one
// contiguous run of the sorted records) in LDS with the data the deposit
// needs, then one thread per face merges its 8 cells' lists in particle
// order (velocityadvector.cpp:559-611).  A quarter whose records exceed LDS
// reads them from global memory instead: same merge, same bits.
constexpr int AT_J = 5, AT_K = 5;                          // quarter of a block in j and k
constexpr int AT_FACES = BW * AT_J * AT_K;                 // 250
constexpr int AT_THREADS = 256;
constexpr int AT_ROWS = (AT_J + 1) * (AT_K + 1);           // 36 rows of cells ci = -1 .. BW-1
constexpr int AT_CAP = 3968;                               // staged records: 8 arrays, 127 KB
static_assert(AT_FACES <= AT_THREADS && AT_ROWS <= AT_THREADS, "one face and one row per thread");

template <int dir>
__global__ __launch_bounds__(AT_THREADS) void apicTileKernel(Grid g, const float *px, const float *py, const float *pz,
                                                             const float *vel, const float *ax, const float *ay,
                                                             const float *az, const uint8_t *active,
                                                             const int *offsets, const int *perm, float *field,
                                                             uint8_t *valid) {
    __shared__ int sRowStart[AT_ROWS + 1], sRowGlob[AT_ROWS];
    __shared__ int sQ[AT_CAP];
    __shared__ float sX[AT_CAP], sY[AT_CAP], sZ[AT_CAP], sV[AT_CAP], sAx[AT_CAP], sAy[AT_CAP], sAz[AT_CAP];
    const FaceDims D = faceDims(g, dir);
    const int t = threadIdx.x;
    const int blk = blockIdx.x / 4, quarter = blockIdx.x % 4;
    const int bi = blk % D.bi, bj = (blk / D.bi) % D.bj, bk = blk / (D.bi * D.bj);
    const int lj0 = (quarter & 1) * AT_J, lk0 = (quarter >> 1) * AT_K;
    const size_t bid = (size_t)blk;
    // this thread's face
    const int li = t % BW, lj = lj0 + (t / BW) % AT_J, lk = lk0 + t / (BW * AT_J);
    const int fi = bi * BW + li, fj = bj * BW + lj, fk = bk * BW + lk;
    const bool mine = t < AT_FACES && fi < D.wi && fj < D.wj && fk < D.wk;
    const size_t f = mine ? flat(fi, fj, fk, D.wi, D.wj) : 0;
    if (!active[bid]) {  // workgroup-uniform: the block received no particles
        if (mine) {
            field[f] = 0.0f;
            valid[f] = 0;
        }
        return;
    }
    const int rowKey0 = (int)(bid * CELLS_PER_BLOCK);
    auto rowKey = [&](int cj, int ck) { return rowKey0 + LC * ((cj + 1) + LC * (ck + 1)); };  // cell ci = -1
    int cnt = 0;
    if (t < AT_ROWS) {
        const int k0 = rowKey(lj0 - 1 + t % (AT_J + 1), lk0 - 1 + t / (AT_J + 1));
        const int b = offsets[k0];
        sRowGlob[t] = b;
        cnt = offsets[k0 + BW + 1] - b;  // the row's 11 cells are contiguous keys
        sRowStart[t + 1] = cnt;
    }
    __syncthreads();
    if (t == 0) {
        int acc = 0;
        sRowStart[0] = 0;
        for (int r = 0; r < AT_ROWS; r++) {
            acc += sRowStart[r + 1];
            sRowStart[r + 1] = acc;
        }
    }
    __syncthreads();
    const int total = sRowStart[AT_ROWS];
    const bool staged = total <= AT_CAP;
    const V3 off = dirOffset(g, dir);
    const V3 bpos = indexToPos(bi, bj, bk, BW * g.dx);
    if (staged) {
        for (int q = t; q < total; q += AT_THREADS) {
            int lo = 0, hi = AT_ROWS - 1;  // last row whose start is <= q
            while (lo < hi) {
                const int mid = (lo + hi + 1) >> 1;
                if (sRowStart[mid] <= q) lo = mid;
                else hi = mid - 1;
            }
            const int p = perm[sRowGlob[lo] + (q - sRowStart[lo])] / MAX_MEMBERS;
            const V3 pl = sub(sub(V3{px[p], py[p], pz[p]}, off), bpos);
            sQ[q] = p;
            sX[q] = pl.x;
            sY[q] = pl.y;
            sZ[q] = pl.z;
            sV[q] = vel[p];
            sAx[q] = ax[p];
            sAy[q] = ay[p];
            sAz[q] = az[p];
        }
    }
    __syncthreads();
    if (!mine) return;
    const V3 npos = indexToPos(li, lj, lk, g.dx);
    // cell o = node - corner o; its list starts at the row's start plus the
    // cells before it in the row
    constexpr int NONE = 0x7fffffff;
    int cur[8], end[8], head[8];
#pragma unroll
    for (int o = 0; o < 8; o++) {
        const int ci = li - (o & 1), cj = lj - ((o >> 1) & 1), ck = lk - (o >> 2);
        const int row = (cj - (lj0 - 1)) + (AT_J + 1) * (ck - (lk0 - 1));
        const int key = rowKey(cj, ck) + (ci + 1);
        const int b = offsets[key], e = offsets[key + 1];
        cur[o] = staged ? sRowStart[row] + (b - sRowGlob[row]) : b;
        end[o] = cur[o] + (e - b);
        head[o] = cur[o] < end[o] ? (staged ? sQ[cur[o]] : perm[cur[o]] / MAX_MEMBERS) : NONE;
    }
    float sum = 0.0f, wsum = 0.0f;
    while (true) {
        int best = 0, pbest = head[0];
#pragma unroll
        for (int o = 1; o < 8; o++)
            if (head[o] < pbest) {
                pbest = head[o];
                best = o;
            }
        if (pbest == NONE) break;
        int s = 0;
#pragma unroll
        for (int o = 0; o < 8; o++)
            if (o == best) {
                s = cur[o];
                cur[o] = s + 1;
                head[o] = s + 1 < end[o] ? (staged ? sQ[s + 1] : perm[s + 1] / MAX_MEMBERS) : NONE;
            }
        V3 pl, a;
        float v;
        if (staged) {
            pl = V3{sX[s], sY[s], sZ[s]};
            a = V3{sAx[s], sAy[s], sAz[s]};
            v = sV[s];
        } else {
            pl = sub(sub(V3{px[pbest], py[pbest], pz[pbest]}, off), bpos);
            a = V3{ax[pbest], ay[pbest], az[pbest]};
            v = vel[pbest];
        }
        const int oi = best & 1, oj = (best >> 1) & 1, ok = best >> 2;
        const V3 ip = divide3(sub(pl, indexToPos(li - oi, lj - oj, lk - ok, g.dx)), (float)g.dx);
        const float wx = oi ? ip.x : 1.0f - ip.x, wy = oj ? ip.y : 1.0f - ip.y, wz = ok ? ip.z : 1.0f - ip.z;
        const float weight = wx * wy * wz;
        const float apicTerm = dot3(a, sub(npos, pl));
        sum += weight * (v + apicTerm);
        wsum += weight;
    }
    const float eps = 1e-6f;
    if (wsum > eps) sum /= wsum;
    field[f] = sum;
    valid[f] = wsum > eps ? 1 : 0;
}

}  // namespace

size_t apicRecords(int np) { return (size_t)np * MAX_MEMBERS; }
size_t apicBuckets(Grid g) {
    size_t nb = 0;
    for (int d = 0; d < 3; d++) {
        const FaceDims D = faceDims(g, d);
        nb = std::max(nb, (size_t)D.bi * D.bj * D.bk * CELLS_PER_BLOCK);
    }
    return nb;
}
size_t apicBlocks(Grid g) {
    size_t nb = 0;
    for (int d = 0; d < 3; d++) {
        const FaceDims D = faceDims(g, d);
        nb = std::max(nb, (size_t)D.bi * D.bj * D.bk);
    }
    return nb;
}

void p2gApic(Grid g, int np, const float *px, const float *py, const float *pz, const float *const vel[3],
             const float *const aff[9], double radius, float *const field[3], uint8_t *const valid[3],
             ApicScratch &A, hipStream_t s) {
    if (apicBuckets(g) >= (size_t)0x7fffffff || apicRecords(np) >= (size_t)0x7fffffff)
        throw std::domain_error("APIC transfer: grid or particle count too large for 32-bit record keys");
    for (int dir = 0; dir < 3; dir++) {
        const FaceDims D = faceDims(g, dir);
        const size_t nblk = (size_t)D.bi * D.bj * D.bk, nb = nblk * CELLS_PER_BLOCK;
        const size_t nf = (size_t)D.wi * D.wj * D.wk, R = apicRecords(np);
        hipCheck(hipMemsetAsync(A.mark, 0, nblk, s), "hipMemsetAsync");
        if (np > 0) {
            hipLaunchKernelGGL(markBlocksKernel, dim3(blocksFor(np, TPB)), dim3(TPB), 0, s, g, dir, np, px, py, pz,
                               A.mark);
        }
        hipLaunchKernelGGL(featherKernel, dim3(blocksFor(nblk, TPB)), dim3(TPB), 0, s, D.bi, D.bj, D.bk, A.mark,
                           A.active);
        if (np > 0)
            hipLaunchKernelGGL(apicKeysKernel, dim3(blocksFor(np, TPB)), dim3(TPB), 0, s, g, dir, np, px, py, pz,
                               radius, A.active, A.keys);
        sortKeys(A.keys, R, nb, A.slot, A.counts, A.offsets, A.perm, A.scan, s);
        const float *a3[3] = {aff[3 * dir], aff[3 * dir + 1], aff[3 * dir + 2]};
        const dim3 grid((unsigned)(4 * nblk)), block(AT_THREADS);
        (void)nf;
        if (dir == 0)
            hipLaunchKernelGGL(apicTileKernel<0>, grid, block, 0, s, g, px, py, pz, vel[0], a3[0], a3[1], a3[2],
                               A.active, A.offsets, A.perm, field[0], valid[0]);
        else if (dir == 1)
            hipLaunchKernelGGL(apicTileKernel<1>, grid, block, 0, s, g, px, py, pz, vel[1], a3[0], a3[1], a3[2],
                               A.active, A.offsets, A.perm, field[1], valid[1]);
        else
            hipLaunchKernelGGL(apicTileKernel<2>, grid, block, 0, s, g, px, py, pz, vel[2], a3[0], a3[1], a3[2],
                               A.active, A.offsets, A.perm, field[2], valid[2]);
    }
}

}  // namespace ff
