// kernels_particles.hip -- particle-side kernels of the FLIP substep:
// cell-key counting sort, liquid SDF, P2G gather, G2P + RK3 + collision,
// removal flags and stable compaction.
#include "kernels.h"

namespace ff {

namespace {

unsigned blocksFor(size_t n, unsigned t) { return (unsigned)((n + t - 1) / t); }
constexpr int TPB = 256;

// ------------------------------------------------------------ cell sort ----
// Key = flat cell of Grid3d::positionToGridIndex(p, dx) (grid3d.h:53-58).
// The arrival slot from the atomic counter gives an unordered placement;
// sorting every cell's segment by particle index afterwards makes the whole
// permutation the stable counting sort (ascending index inside a cell).
__global__ void cellKeysKernel(Grid g, const float *px, const float *py, const float *pz, int np,
                               int *keys, int *slot, int *counts) {
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= np) return;
    I3 c = posToIndex(V3{px[p], py[p], pz[p]}, g.dx);
    int key = inRange(c.i, c.j, c.k, g.ni, g.nj, g.nk) ? (int)flat(c.i, c.j, c.k, g.ni, g.nj) : -1;
    keys[p] = key;
    if (key >= 0) slot[p] = atomicAdd(&counts[key], 1);
}

__global__ void scatterKernel(const int *keys, const int *slot, const int *offsets, int np, int *perm) {
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= np) return;
    int key = keys[p];
    if (key >= 0) perm[offsets[key] + slot[p]] = p;
}

__global__ void segmentSortKernel(const int *offsets, size_t ncells, int *perm) {
    size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= ncells) return;
    int b = offsets[c], e = offsets[c + 1];
    for (int a = b + 1; a < e; a++) {
        int v = perm[a];
        int q = a - 1;
        while (q >= b && perm[q] > v) {
            perm[q + 1] = perm[q];
            q--;
        }
        perm[q + 1] = v;
    }
}

// sorted copies; particles outside the grid have no key and no sorted slot,
// so only the first offsets[N] entries of perm are filled
__global__ void gatherSortedKernel(const int *perm, const int *total, Particles p, float *spx, float *spy,
                                   float *spz, float *svx, float *svy, float *svz) {
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= p.n || s >= *total) return;
    int q = perm[s];
    spx[s] = p.px[q];
    spy[s] = p.py[q];
    spz[s] = p.pz[q];
    svx[s] = p.vx[q];
    svy[s] = p.vy[q];
    svz[s] = p.vz[q];
}

// --------------------------------------------------------------- SDF ------
// ParticleLevelSet::_computeSignedDistanceFromParticles (particlelevelset.cpp:335-666)
// as one workgroup per 10^3 block: the block's min-distance tile lives in LDS
// and every particle that belongs to the block (the count-data membership rule,
// 484-537) lowers the cells of its block-local search range with an unsigned
// atomic min on an order-preserving float key -- a min is order independent,
// so the result is exact and deterministic.
constexpr int BW = 10;
__global__ __launch_bounds__(256) void liquidSDFKernel(Grid g, const int *offsets, const float *spx,
                                                       const float *spy, const float *spz,
                                                       double radius, float *phi, int nbi, int nbj) {
    __shared__ uint32_t tile[BW * BW * BW];
    const int b = blockIdx.x;
    const int bi = b % nbi, bj = (b / nbi) % nbj, bk = b / (nbi * nbj);
    const float maxd = (float)(3.0 * g.dx);
    for (int c = threadIdx.x; c < BW * BW * BW; c += blockDim.x) tile[c] = floatKey(maxd);
    __syncthreads();
    const float blockdx = (float)(BW * g.dx);
    const float r = (float)radius;
    const float sr = 2.0f * r;
    const V3 off = indexToPos(bi, bj, bk, BW * g.dx);
    const int i0 = max(bi * BW - 2, 0), i1 = min(bi * BW + BW + 1, g.ni - 1);
    const int j0 = max(bj * BW - 2, 0), j1 = min(bj * BW + BW + 1, g.nj - 1);
    const int k0 = max(bk * BW - 2, 0), k1 = min(bk * BW + BW + 1, g.nk - 1);
    const int nrowj = j1 - j0 + 1, nrows = nrowj * (k1 - k0 + 1);
    const int wave = threadIdx.x / 64, lane = threadIdx.x % 64, nwaves = blockDim.x / 64;
    for (int row = wave; row < nrows; row += nwaves) {
        int j = j0 + row % nrowj, k = k0 + row / nrowj;
        int s0 = offsets[flat(i0, j, k, g.ni, g.nj)];
        int s1 = offsets[flat(i1, j, k, g.ni, g.nj) + 1];
        for (int s = s0 + lane; s < s1; s += 64) {
            V3 p{spx[s], spy[s], spz[s]};
            // membership of block (bi, bj, bk)
            I3 ob = posToIndex(p, blockdx);
            V3 bp = indexToPos(ob.i, ob.j, ob.k, blockdx);
            bool member;
            if (p.x - sr > bp.x && p.y - sr > bp.y && p.z - sr > bp.z && p.x + sr < bp.x + blockdx &&
                p.y + sr < bp.y + blockdx && p.z + sr < bp.z + blockdx) {
                member = ob.i == bi && ob.j == bj && ob.k == bk;
            } else {
                I3 a = posToIndexD(p.x - sr, p.y - sr, p.z - sr, blockdx);
                I3 c = posToIndexD(p.x + sr, p.y + sr, p.z + sr, blockdx);
                member = bi >= a.i && bi <= c.i && bj >= a.j && bj <= c.j && bk >= a.k && bk <= c.k;
            }
            if (!member) continue;
            V3 pl = sub(p, off);
            I3 a = posToIndex(V3{pl.x - sr, pl.y - sr, pl.z - sr}, g.dx);
            I3 c = posToIndex(V3{pl.x + sr, pl.y + sr, pl.z + sr}, g.dx);
            a.i = max(a.i, 0); a.j = max(a.j, 0); a.k = max(a.k, 0);
            c.i = min(c.i, BW - 1); c.j = min(c.j, BW - 1); c.k = min(c.k, BW - 1);
            for (int kk = a.k; kk <= c.k; kk++)
                for (int jj = a.j; jj <= c.j; jj++)
                    for (int ii = a.i; ii <= c.i; ii++) {
                        V3 gpos = indexToCellCenter(ii, jj, kk, g.dx);
                        float dist = len3(sub(gpos, pl)) - r;
                        atomicMin(&tile[ii + BW * (jj + BW * kk)], floatKey(dist));
                    }
        }
    }
    __syncthreads();
    for (int c = threadIdx.x; c < BW * BW * BW; c += blockDim.x) {
        int ii = bi * BW + c % BW, jj = bj * BW + (c / BW) % BW, kk = bk * BW + c / (BW * BW);
        if (ii < g.ni && jj < g.nj && kk < g.nk) phi[flat(ii, jj, kk, g.ni, g.nj)] = keyFloat(tile[c]);
    }
}

// ParticleLevelSet::postProcessSignedDistanceField (particlelevelset.cpp:170-194)
__global__ void sdfPostKernel(Grid g, const float *sp, float *phi) {
    size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t n = (size_t)g.ni * g.nj * g.nk;
    if (idx >= n) return;
    int i = (int)(idx % g.ni), j = (int)((idx / g.ni) % g.nj), k = (int)(idx / ((size_t)g.ni * g.nj));
    const int pi = g.ni + 1, pj = g.nj + 1;
    float eps = (float)(0.005 * g.dx);
    float val = phi[idx];
    if ((double)val < 0.5 * g.dx) {
        float sc = 0.125f * (sp[flat(i, j, k, pi, pj)] + sp[flat(i + 1, j, k, pi, pj)] +
                             sp[flat(i, j + 1, k, pi, pj)] + sp[flat(i + 1, j + 1, k, pi, pj)] +
                             sp[flat(i, j, k + 1, pi, pj)] + sp[flat(i + 1, j, k + 1, pi, pj)] +
                             sp[flat(i, j + 1, k + 1, pi, pj)] + sp[flat(i + 1, j + 1, k + 1, pi, pj)]);
        if (sc < 0) val = (float)(-0.5f * g.dx);
    }
    if (fabsf(val) < eps) val = val > 0 ? eps : -eps;
    phi[idx] = val;
}

// --------------------------------------------------------------- P2G ------
// VelocityAdvector FLIP transfer (velocityadvector.cpp:467-536) as a gather:
// a face sums w*v and w over the particles within r of it, with the distance
// taken in the local coordinates of the face's own 10^3 block (the block that
// holds the face in the reference, whose particle list is in ascending
// particle index).  The candidate cells' sorted lists are merged by particle
// index, so the float sums run in the reference's order and no atomics are
// used.
//
// One workgroup handles a tile of 8 x 8 x 4 faces of each of the three
// directions.  The particles of the 10 x 10 x 6 cells those faces can reach
// (the tile grown by one cell on each side) are staged in LDS once, with
// their positions, three velocity components and particle index; then one
// thread per face merges its 18 candidate cells from LDS.  A tile whose cells
// hold more particles than LDS does reads them from the sorted arrays instead
// (same merge, same order, same bits).
constexpr int P2T_I = 8, P2T_J = 8, P2T_K = 4;  // faces of a tile per direction
constexpr int P2R_I = P2T_I + 2, P2R_J = P2T_J + 2, P2R_K = P2T_K + 2;
constexpr int P2R_CELLS = P2R_I * P2R_J * P2R_K;  // staged cells
constexpr int P2_FACES = P2T_I * P2T_J * P2T_K;   // faces per direction and tile
constexpr int P2_THREADS = 768;                   // one face per thread
constexpr int P2G_CAP = 5500;                     // staged particles (LDS: 28 bytes each)

// Face (fi, fj, fk) of direction dir: sum and weight sum over the particles
// of its 18 candidate cells (range(c, b, e) gives cell c's particles [b, e)
// as indices into x/y/z/vel/pidx), merged in ascending particle index.
// The merge takes a full minimum over the 18 list heads only when the list it
// draws from stops being the smallest: it then keeps drawing from that list
// while its next head is below the second smallest head, which gives the same
// order with one scan per run of a list.  The 18-entry cursor and head arrays
// are only addressed by unrolled constant indices (selects), so the merge
// uses no dynamically indexed registers or private memory.
template <int dir, class F, class I, class R>
__device__ __forceinline__ void p2gFace(Grid g, int fi, int fj, int fk, R range, F x, F y, F z, F vel,
                                        I pidx, double radius, float *field, uint8_t *valid, float *wsumOut,
                                        size_t f) {
    const float eps = 1e-6f;
    const P2GWeight W = p2gWeight(radius);
    const float rsq = W.rsq;
    const FaceFrame frame = p2gFrame(g, dir, fi, fj, fk);
    auto d2of = [&](int s) { return p2gDist2(frame, x[s], y[s], z[s]); };
    // first particle at or after c within r: its particle index (or none) and distance
    auto next = [&](int &c, int e, float &d2) {
        while (c < e) {
            d2 = d2of(c);
            if (d2 < rsq) return pidx[c];
            c++;
        }
        return 0x7fffffff;
    };
    auto put = [](int(&a)[18], int c, int v) {
#pragma unroll
        for (int q = 0; q < 18; q++) a[q] = q == c ? v : a[q];
    };
    int cur[18], end[18], head[18];
#pragma unroll 1
    for (int c = 0; c < 18; c++) {
        int b, e;
        range(c, b, e);
        float d2;
        const int hh = next(b, e, d2);
        put(cur, c, b);
        put(end, c, e);
        put(head, c, hh);
    }
    float sum = 0.0f, wsum = 0.0f;
    while (true) {
        int best = -1, bidx = 0x7fffffff, second = 0x7fffffff, bcur = 0, bend = 0;
#pragma unroll
        for (int c = 0; c < 18; c++) {
            const int hc = head[c];
            if (hc < bidx) {
                second = bidx;
                bidx = hc;
                best = c;
                bcur = cur[c];
                bend = end[c];
            } else if (hc < second) {
                second = hc;
            }
        }
        if (best < 0) break;
        float d2 = d2of(bcur);
        do {
            const float weight = p2gWeightAt(W, d2);
            sum += weight * vel[bcur];
            wsum += weight;
            bcur++;
            bidx = next(bcur, bend, d2);
        } while (bidx < second);
        put(cur, best, bcur);
        put(head, best, bidx);
    }
    if (wsumOut) wsumOut[f] = wsum;
    if (wsum > eps) sum /= wsum;
    field[f] = sum;
    valid[f] = wsum > eps ? 1 : 0;
}

struct P2GOut {
    float *f[3];
    uint8_t *valid[3];
    float *wsum[3];  // optional: each face's weight sum
};

// face a of the tile at (I0, J0, K0) in direction dir, from the staged
// particles (or the sorted arrays when the tile's particles overflowed LDS)
template <int dir>
__device__ __forceinline__ void p2gTileFace(Grid g, int a, int I0, int J0, int K0, int total, bool staged,
                                            const int *sStart, const int *sGlob, const float *sx, const float *sy,
                                            const float *sz, const float *sv, const int *sp, const float *spx,
                                            const float *spy, const float *spz, const float *svel, const int *perm,
                                            double radius, P2GOut out) {
    const int fi = I0 + a % P2T_I, fj = J0 + (a / P2T_I) % P2T_J, fk = K0 + a / (P2T_I * P2T_J);
    const int wi = g.ni + (dir == 0), wj = g.nj + (dir == 1), wk = g.nk + (dir == 2);
    if (fi >= wi || fj >= wj || fk >= wk) return;
    const size_t f = flat(fi, fj, fk, wi, wj);
    if (total == 0) {  // no particle can reach the tile's faces
        out.f[dir][f] = 0.0f;
        out.valid[dir][f] = 0;
        if (out.wsum[dir]) out.wsum[dir][f] = 0.0f;
        return;
    }
    // candidate cells: 2 along the face normal, 3 across; candidate c of the
    // box, i fastest
    constexpr int NI = dir == 0 ? 2 : 3, NJ = dir == 1 ? 2 : 3;
    const int ci0 = fi - 1, cj0 = fj - 1, ck0 = fk - 1;
    auto range = [&](int c, int &b, int &e) {
        const int i = ci0 + c % NI, j = cj0 + (c / NI) % NJ, k = ck0 + c / (NI * NJ);
        b = 0;
        e = 0;
        if (inRange(i, j, k, g.ni, g.nj, g.nk)) {
            const int rc = (i - (I0 - 1)) + P2R_I * ((j - (J0 - 1)) + P2R_J * (k - (K0 - 1)));
            b = staged ? sStart[rc] : sGlob[rc];
            e = b + (sStart[rc + 1] - sStart[rc]);
        }
    };
    if (staged)
        p2gFace<dir>(g, fi, fj, fk, range, sx, sy, sz, sv, sp, radius, out.f[dir], out.valid[dir], out.wsum[dir], f);
    else
        p2gFace<dir>(g, fi, fj, fk, range, spx, spy, spz, svel, perm, radius, out.f[dir], out.valid[dir],
                     out.wsum[dir], f);
}

__global__ __launch_bounds__(P2_THREADS) void p2gTileKernel(Grid g, const int *offsets, const int *perm,
                                                            const float *spx, const float *spy, const float *spz,
                                                            const float *svx, const float *svy, const float *svz,
                                                            double radius, P2GOut out, int tilesI, int tilesJ) {
    __shared__ int sStart[P2R_CELLS + 1], sGlob[P2R_CELLS], sWave[P2_THREADS / 64];
    __shared__ float sx[P2G_CAP], sy[P2G_CAP], sz[P2G_CAP], sv[3][P2G_CAP];
    __shared__ int sp[P2G_CAP];
    const int t = threadIdx.x;
    const int I0 = (blockIdx.x % tilesI) * P2T_I, J0 = ((blockIdx.x / tilesI) % tilesJ) * P2T_J,
              K0 = (blockIdx.x / (tilesI * tilesJ)) * P2T_K;
    // particle counts of the staged cells (cells 2t and 2t+1 for thread t),
    // then an exclusive scan over them in cell order
    static_assert(2 * P2_THREADS >= P2R_CELLS, "two staged cells per thread");
    int cnt[2] = {0, 0};
#pragma unroll
    for (int h = 0; h < 2; h++) {
        const int rc = 2 * t + h;
        if (rc < P2R_CELLS) {
            const int i = I0 - 1 + rc % P2R_I, j = J0 - 1 + (rc / P2R_I) % P2R_J, k = K0 - 1 + rc / (P2R_I * P2R_J);
            int b = 0;
            if (inRange(i, j, k, g.ni, g.nj, g.nk)) {
                const size_t c = flat(i, j, k, g.ni, g.nj);
                b = offsets[c];
                cnt[h] = offsets[c + 1] - b;
            }
            sGlob[rc] = b;
        }
    }
    const int pair = cnt[0] + cnt[1];
    int incl = pair;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const int v = __shfl_up(incl, o, 64);
        if ((t & 63) >= o) incl += v;
    }
    if ((t & 63) == 63) sWave[t >> 6] = incl;
    __syncthreads();
    int waveBase = 0;
    for (int w = 0; w < (t >> 6); w++) waveBase += sWave[w];
    const int excl = waveBase + incl - pair;
    if (2 * t < P2R_CELLS) sStart[2 * t] = excl;
    if (2 * t + 1 < P2R_CELLS) sStart[2 * t + 1] = excl + cnt[0];
    if (2 * t + 2 == P2R_CELLS || 2 * t + 1 == P2R_CELLS) sStart[P2R_CELLS] = excl + pair;
    __syncthreads();
    const int total = sStart[P2R_CELLS];
    const bool staged = total <= P2G_CAP;
    if (staged) {
        for (int q = t; q < total; q += P2_THREADS) {
            int lo = 0, hi = P2R_CELLS - 1;  // last cell whose start is <= q
            while (lo < hi) {
                const int mid = (lo + hi + 1) >> 1;
                if (sStart[mid] <= q) lo = mid;
                else hi = mid - 1;
            }
            const int gs = sGlob[lo] + (q - sStart[lo]);
            sx[q] = spx[gs];
            sy[q] = spy[gs];
            sz[q] = spz[gs];
            sv[0][q] = svx[gs];
            sv[1][q] = svy[gs];
            sv[2][q] = svz[gs];
            sp[q] = perm[gs];
        }
    }
    __syncthreads();
    // one face per thread (direction = face / P2_FACES, wave-uniform)
    for (int task = t; task < 3 * P2_FACES; task += P2_THREADS) {
        const int dir = task / P2_FACES;
        if (dir == 0) p2gTileFace<0>(g, task % P2_FACES, I0, J0, K0, total, staged, sStart, sGlob, sx, sy, sz, sv[0], sp,
                                     spx, spy, spz, svx, perm, radius, out);
        else if (dir == 1) p2gTileFace<1>(g, task % P2_FACES, I0, J0, K0, total, staged, sStart, sGlob, sx, sy, sz, sv[1],
                                          sp, spx, spy, spz, svy, perm, radius, out);
        else p2gTileFace<2>(g, task % P2_FACES, I0, J0, K0, total, staged, sStart, sGlob, sx, sy, sz, sv[2], sp, spx, spy,
                            spz, svz, perm, radius, out);
    }
}

// ------------------------------------------------- max |v|^2 over particles
__global__ __launch_bounds__(256) void maxSpeedStage1(Particles p, double *partial) {
    __shared__ double s[256];
    double acc = 0.0;
    for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < p.n; q += gridDim.x * blockDim.x) {
        V3 v{p.vx[q], p.vy[q], p.vz[q]};
        acc = fmax(acc, (double)dot3(v, v));
    }
    s[threadIdx.x] = acc;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < (unsigned)o) s[threadIdx.x] = fmax(s[threadIdx.x], s[threadIdx.x + o]);
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = s[0];
}
__global__ __launch_bounds__(256) void maxStage2(const double *partial, int nb, double *out) {
    __shared__ double s[256];
    double acc = 0.0;
    for (int q = threadIdx.x; q < nb; q += 256) acc = fmax(acc, partial[q]);
    s[threadIdx.x] = acc;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < (unsigned)o) s[threadIdx.x] = fmax(s[threadIdx.x], s[threadIdx.x + o]);
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = s[0];
}

// ------------------------------------------------- G2P + RK3 + collision ---
struct FaceArr {
    const float *a;
    int wi, wj, wk;
};
// MACVelocityField::_interpolateLinearU/V/W after the in-grid test
// (macvelocityfield.cpp:519-613); transverse coordinates already shifted.
__device__ double interpFace(FaceArr F, double dx, double x, double y, double z) {
    double invdx = 1.0 / dx;
    int i = (int)floor(x * invdx), j = (int)floor(y * invdx), k = (int)floor(z * invdx);
    double gx = (double)i * dx, gy = (double)j * dx, gz = (double)k * dx;
    double inv_dx = 1 / dx;
    double ix = (x - gx) * inv_dx, iy = (y - gy) * inv_dx, iz = (z - gz) * inv_dx;
    double pts[8];
    const int o[8][3] = {{0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1}, {1, 0, 1}, {0, 1, 1}, {1, 1, 0}, {1, 1, 1}};
#pragma unroll
    for (int q = 0; q < 8; q++) {
        int a = i + o[q][0], b = j + o[q][1], c = k + o[q][2];
        pts[q] = inRange(a, b, c, F.wi, F.wj, F.wk) ? (double)F.a[flat(a, b, c, F.wi, F.wj)] : 0.0;
    }
    return trilin8(pts, ix, iy, iz);
}
// MACVelocityField::evaluateVelocityAtPositionLinear (macvelocityfield.cpp:631-645)
__device__ V3 sampleMac(Grid g, MacView M, V3 p) {
    double x = p.x, y = p.y, z = p.z;
    if (!(x >= 0 && y >= 0 && z >= 0 && x < g.dx * g.ni && y < g.dx * g.nj && z < g.dx * g.nk))
        return V3{0.0f, 0.0f, 0.0f};
    double h = 0.5 * g.dx;
    double xu = interpFace(FaceArr{M.u, g.ni + 1, g.nj, g.nk}, g.dx, x, y - h, z - h);
    double yv = interpFace(FaceArr{M.v, g.ni, g.nj + 1, g.nk}, g.dx, x - h, y, z - h);
    double zw = interpFace(FaceArr{M.w, g.ni, g.nj, g.nk + 1}, g.dx, x - h, y - h, z);
    return mk(xu, yv, zw);
}

// Gradient of the trilinear sample of each component at p: grad[d][e] =
// d V_d / d p_e inside the cell the forward picks (floor of the shifted
// coordinate), 0 outside the domain, where the forward returns 0.
__device__ void sampleMacGrad(Grid g, MacView M, V3 p, double grad[3][3]) {
    for (int d = 0; d < 3; d++)
        for (int e = 0; e < 3; e++) grad[d][e] = 0.0;
    double x = p.x, y = p.y, z = p.z;
    if (!(x >= 0 && y >= 0 && z >= 0 && x < g.dx * g.ni && y < g.dx * g.nj && z < g.dx * g.nk)) return;
    const double h = 0.5 * g.dx;
    const double invdx = 1.0 / g.dx;
    for (int d = 0; d < 3; d++) {
        const FaceArr F = d == 0 ? FaceArr{M.u, g.ni + 1, g.nj, g.nk}
                                 : (d == 1 ? FaceArr{M.v, g.ni, g.nj + 1, g.nk} : FaceArr{M.w, g.ni, g.nj, g.nk + 1});
        const double xs = d == 0 ? x : x - h, ys = d == 1 ? y : y - h, zs = d == 2 ? z : z - h;
        const int i = (int)floor(xs * invdx), j = (int)floor(ys * invdx), k = (int)floor(zs * invdx);
        const double ix = (xs - (double)i * g.dx) * invdx, iy = (ys - (double)j * g.dx) * invdx,
                     iz = (zs - (double)k * g.dx) * invdx;
        for (int q = 0; q < 8; q++) {
            const int ox = q & 1, oy = (q >> 1) & 1, oz = q >> 2;
            const int a = i + ox, b = j + oy, c = k + oz;
            if (!inRange(a, b, c, F.wi, F.wj, F.wk)) continue;
            const double val = (double)F.a[flat(a, b, c, F.wi, F.wj)];
            const double ax = ox ? ix : 1 - ix, ay = oy ? iy : 1 - iy, az = oz ? iz : 1 - iz;
            grad[d][0] += val * (ox ? 1.0 : -1.0) * ay * az * invdx;
            grad[d][1] += val * ax * (oy ? 1.0 : -1.0) * az * invdx;
            grad[d][2] += val * ax * ay * (oz ? 1.0 : -1.0) * invdx;
        }
    }
}

// _RK3 (fluidsimulation.cpp:7616-7623) without the collision step
__global__ void advectRK3Kernel(Grid g, int np, const float *px, const float *py, const float *pz, MacView cur,
                                double dt, float *ox, float *oy, float *oz) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= np) return;
    const V3 p0{px[q], py[q], pz[q]};
    const V3 k1 = sampleMac(g, cur, p0);
    const V3 k2 = sampleMac(g, cur, add(p0, scale((float)(0.5 * dt), k1)));
    const V3 k3 = sampleMac(g, cur, add(p0, scale((float)(0.75 * dt), k2)));
    const V3 p1 = add(p0, scale((float)(dt / (double)9.0f), add(add(scale(2.0f, k1), scale(3.0f, k2)), scale(4.0f, k3))));
    ox[q] = p1.x;
    oy[q] = p1.y;
    oz[q] = p1.z;
}

// Adjoint of advectRK3Kernel: pbar (P x 3) for xbar (P x 3), and the three
// sample points and their velocity cotangents (k1 at p0, k2 at y2, k3 at y3)
// whose sampling adjoints give the face-grid cotangent.  With B.vbarOut set
// it is the adjoint of g2pAdvectKernel with the collision response taken as
// the identity: the new velocity vn = r1 k1 + r2 ((v0 + k1) - V_sav(p0))
// adds (r1 + r2) vbarOut to k1's cotangent, r2 vbarOut to v0's (B.vbar) and
// -r2 vbarOut to the saved field's sample at p0 (B.cotSav), whose position
// gradient enters pbar.
__global__ void advectRK3AdjKernel(Grid g, int np, const float *px, const float *py, const float *pz, MacView cur,
                                   double dt, const float *xbar, float *pbar, float *pts, float *cot, BlendAdj B) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= np) return;
    const V3 p0{px[q], py[q], pz[q]};
    const float c2 = (float)(0.5 * dt), c3 = (float)(0.75 * dt), a = (float)(dt / (double)9.0f);
    const V3 k1 = sampleMac(g, cur, p0);
    const V3 y2 = add(p0, scale(c2, k1));
    const V3 k2 = sampleMac(g, cur, y2);
    const V3 y3 = add(p0, scale(c3, k2));
    double xb[3] = {xbar[3 * (size_t)q], xbar[3 * (size_t)q + 1], xbar[3 * (size_t)q + 2]};
    double kb1[3], kb2[3], kb3[3], pb[3], G[3][3];
    for (int e = 0; e < 3; e++) {
        kb1[e] = 2.0 * (double)a * xb[e];
        kb2[e] = 3.0 * (double)a * xb[e];
        kb3[e] = 4.0 * (double)a * xb[e];
        pb[e] = xb[e];
    }
    if (B.vbarOut) {
        double sb[3];
        for (int e = 0; e < 3; e++) {
            const double vb = (double)B.vbarOut[3 * (size_t)q + e];
            kb1[e] += ((double)B.r1 + (double)B.r2) * vb;
            sb[e] = -(double)B.r2 * vb;
            B.vbar[3 * (size_t)q + e] = (float)((double)B.r2 * vb);
            B.cotSav[3 * (size_t)q + e] = (float)sb[e];
        }
        sampleMacGrad(g, B.sav, p0, G);
        for (int e = 0; e < 3; e++) pb[e] += G[0][e] * sb[0] + G[1][e] * sb[1] + G[2][e] * sb[2];
    }
    // k3 = V(y3), y3 = p0 + c3 k2
    sampleMacGrad(g, cur, y3, G);
    for (int e = 0; e < 3; e++) {
        const double yb = G[0][e] * kb3[0] + G[1][e] * kb3[1] + G[2][e] * kb3[2];
        pb[e] += yb;
        kb2[e] += (double)c3 * yb;
    }
    // k2 = V(y2), y2 = p0 + c2 k1
    sampleMacGrad(g, cur, y2, G);
    for (int e = 0; e < 3; e++) {
        const double yb = G[0][e] * kb2[0] + G[1][e] * kb2[1] + G[2][e] * kb2[2];
        pb[e] += yb;
        kb1[e] += (double)c2 * yb;
    }
    // k1 = V(p0)
    sampleMacGrad(g, cur, p0, G);
    for (int e = 0; e < 3; e++) pb[e] += G[0][e] * kb1[0] + G[1][e] * kb1[1] + G[2][e] * kb1[2];
    const V3 at[3] = {p0, y2, y3};
    const double *kb[3] = {kb1, kb2, kb3};
    for (int e = 0; e < 3; e++) pbar[3 * (size_t)q + e] = (float)pb[e];
    const size_t n3 = 3 * (size_t)np;  // pts: x of all 3P points, then y, then z
    for (int s = 0; s < 3; s++) {
        const size_t o = (size_t)s * np + q;
        pts[o] = at[s].x;
        pts[n3 + o] = at[s].y;
        pts[2 * n3 + o] = at[s].z;
        for (int e = 0; e < 3; e++) cot[3 * o + e] = (float)kb[s][e];
    }
}

// Interpolation::trilinearInterpolate on the nodal solid SDF (interpolation.cpp:72-112)
__device__ float nodalTrilinear(Grid g, const float *sp, V3 p) {
    const int pi = g.ni + 1, pj = g.nj + 1, pk = g.nk + 1;
    I3 c = posToIndex(p, g.dx);
    V3 gp = indexToPos(c.i, c.j, c.k, g.dx);
    double inv_dx = 1.0 / g.dx;
    double ix = (double)(p.x - gp.x) * inv_dx, iy = (double)(p.y - gp.y) * inv_dx,
           iz = (double)(p.z - gp.z) * inv_dx;
    double pts[8];
    const int o[8][3] = {{0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1}, {1, 0, 1}, {0, 1, 1}, {1, 1, 0}, {1, 1, 1}};
#pragma unroll
    for (int q = 0; q < 8; q++) {
        int a = c.i + o[q][0], b = c.j + o[q][1], e = c.k + o[q][2];
        pts[q] = inRange(a, b, e, pi, pj, pk) ? (double)sp[flat(a, b, e, pi, pj)] : 0.0;
    }
    return (float)trilin8(pts, ix, iy, iz);
}
// Interpolation::trilinearInterpolateGradient (interpolation.cpp:189-260)
__device__ double bilin(double v00, double v10, double v01, double v11, double ix, double iy) {
    double l1 = (1 - ix) * v00 + ix * v10;
    double l2 = (1 - ix) * v01 + ix * v11;
    return (1 - iy) * l1 + iy * l2;
}
__device__ V3 nodalGradient(Grid g, const float *sp, V3 p) {
    const int pi = g.ni + 1, pj = g.nj + 1, pk = g.nk + 1;
    I3 c = posToIndex(p, g.dx);
    V3 gp = indexToPos(c.i, c.j, c.k, g.dx);
    double inv_dx = 1.0 / g.dx;
    double ix = (double)(p.x - gp.x) * inv_dx, iy = (double)(p.y - gp.y) * inv_dx,
           iz = (double)(p.z - gp.z) * inv_dx;
    auto at = [&](int a, int b, int e) -> float {
        return inRange(a, b, e, pi, pj, pk) ? sp[flat(a, b, e, pi, pj)] : 0.0f;
    };
    float v000 = at(c.i, c.j, c.k), v100 = at(c.i + 1, c.j, c.k), v010 = at(c.i, c.j + 1, c.k);
    float v001 = at(c.i, c.j, c.k + 1), v101 = at(c.i + 1, c.j, c.k + 1);
    float v011 = at(c.i, c.j + 1, c.k + 1), v110 = at(c.i + 1, c.j + 1, c.k);
    float v111 = at(c.i + 1, c.j + 1, c.k + 1);
    float dvdx = (float)bilin(v100 - v000, v110 - v010, v101 - v001, v111 - v011, iy, iz);
    float dvdy = (float)bilin(v010 - v000, v110 - v100, v011 - v001, v111 - v101, ix, iz);
    float dvdz = (float)bilin(v001 - v000, v101 - v100, v011 - v010, v111 - v110, ix, iy);
    return V3{dvdx, dvdy, dvdz};
}

struct NearGrid {
    const uint8_t *a;
    int ci, cj, ck;
    double cell;
};
__device__ bool nearAt(NearGrid N, V3 p) {
    I3 c = posToIndex(p, N.cell);
    if (!inRange(c.i, c.j, c.k, N.ci, N.cj, N.ck)) return false;
    return N.a[flat(c.i, c.j, c.k, N.ci, N.cj)] != 0;
}

// FluidSimulation::_resolveCollision (fluidsimulation.cpp:7646-7721)
__device__ V3 resolveCollision(Grid g, const float *sp, NearGrid N, double cfl, V3 oldp, V3 newp,
                               const Box &boundary) {
    I3 gg = posToIndex(newp, g.dx);
    if (!inRange(gg.i, gg.j, gg.k, g.ni, g.nj, g.nk)) newp = boundary.nearestInside(newp);
    if (!nearAt(N, oldp) && !nearAt(N, newp)) return newp;
    const float eps = 1e-6f;
    float stepDistance = 0.1f * (float)g.dx;
    float travel = len3(sub(newp, oldp));
    if (travel < eps) return newp;
    // A non-finite travel converts to INT_MIN steps on x86 and the reference
    // returns newp; the step count is also capped so that a runaway particle
    // cannot stall the wave (the cap is 100 grid widths of march).
    if (!(travel < 3.0e38f)) return newp;
    float fsteps = ceilf(travel / stepDistance);
    int numSteps = fsteps > 1.0e7f ? 10000000 : (int)fsteps;
    V3 stepdir = normalize3(sub(newp, oldp));
    V3 last = oldp, current{0, 0, 0};
    bool found = false;
    float collisionPhi = 0.0f;
    for (int st = 0; st < numSteps; st++) {
        if (st == numSteps - 1)
            current = newp;
        else
            current = add(oldp, scale((float)(st + 1) * stepDistance, stepdir));
        float ph = nodalTrilinear(g, sp, current);
        if (ph < 0.0f || !boundary.inside(current)) {
            collisionPhi = ph;
            found = true;
            break;
        }
        last = current;
    }
    if (!found) return newp;
    V3 resolved;
    float maxResolved = (float)(cfl * g.dx);
    V3 grad = nodalGradient(g, sp, current);
    if (len3(grad) > eps) {
        grad = normalize3(grad);
        resolved = sub(current, scale((float)((double)collisionPhi - (double)0.2f * g.dx), grad));
        float rphi = nodalTrilinear(g, sp, resolved);
        float rdist = len3(sub(resolved, current));
        if (rphi < 0 || rdist > maxResolved) resolved = last;
    } else {
        resolved = last;
    }
    if (!boundary.inside(resolved)) {
        V3 orig = resolved;
        resolved = boundary.nearestInside(resolved);
        float rphi = nodalTrilinear(g, sp, resolved);
        float rdist = len3(sub(resolved, orig));
        if (rphi < 0.0f || rdist > maxResolved) resolved = last;
    }
    return resolved;
}

// G2P (fluidsimulation.cpp:6771-6784), then _RK3 (7616-7623) on the unchanged
// current field and the collision of the old/new pair (7626-7644).
// FluidSimulation::_getIndicesAndGradientWeights (fluidsimulation.cpp:6709-6769)
// and the affine sums of _updatePICAPICMarkerParticleVelocitiesThread
// (6791-6843): the gradient of component d of the field at p, in the
// reference's float arithmetic, written to out[3d .. 3d+2]
__device__ void apicAffine(Grid g, MacView M, V3 p0, float *const *out, int q) {
    const float h = (float)(0.5f * g.dx);
    const float invdx = (float)(1.0f / g.dx);
    for (int d = 0; d < 3; d++) {
        const FaceArr F = d == 0 ? FaceArr{M.u, g.ni + 1, g.nj, g.nk}
                                 : (d == 1 ? FaceArr{M.v, g.ni, g.nj + 1, g.nk} : FaceArr{M.w, g.ni, g.nj, g.nk + 1});
        const V3 off = d == 0 ? V3{0.0f, h, h} : (d == 1 ? V3{h, 0.0f, h} : V3{h, h, 0.0f});
        const V3 x = sub(p0, off);
        const I3 c = posToIndex(x, g.dx);
        const V3 ip = divide3(sub(x, indexToPos(c.i, c.j, c.k, g.dx)), (float)g.dx);
        const float mx = 1.0f - ip.x, my = 1.0f - ip.y, mz = 1.0f - ip.z, ninv = -invdx;
        // the eight gradient weights with the reference's factor order
        const V3 gw[8] = {V3{ninv * my * mz, ninv * mx * mz, ninv * mx * my},
                          V3{invdx * my * mz, ip.x * ninv * mz, ip.x * my * ninv},
                          V3{ninv * ip.y * mz, mx * invdx * mz, mx * ip.y * ninv},
                          V3{invdx * ip.y * mz, ip.x * invdx * mz, ip.x * ip.y * ninv},
                          V3{ninv * my * ip.z, mx * ninv * ip.z, mx * my * invdx},
                          V3{invdx * my * ip.z, ip.x * ninv * ip.z, ip.x * my * invdx},
                          V3{ninv * ip.y * ip.z, mx * invdx * ip.z, mx * ip.y * invdx},
                          V3{invdx * ip.y * ip.z, ip.x * invdx * ip.z, ip.x * ip.y * invdx}};
        V3 a{0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int o = 0; o < 8; o++) {
            const int i = c.i + (o & 1), j = c.j + ((o >> 1) & 1), k = c.k + (o >> 2);
            if (!inRange(i, j, k, F.wi, F.wj, F.wk)) continue;
            const float fv = F.a[flat(i, j, k, F.wi, F.wj)];
            a = add(a, V3{gw[o].x * fv, gw[o].y * fv, gw[o].z * fv});
        }
        out[3 * d][q] = a.x;
        out[3 * d + 1][q] = a.y;
        out[3 * d + 2][q] = a.z;
    }
}

__global__ __launch_bounds__(256) void g2pAdvectKernel(Grid g, Particles p, MacView cur, MacView sav,
                                                       const float *sp, NearGrid N, float r1,
                                                       float r2, double dt, double cfl, uint8_t *moved,
                                                       AffineArrays aff) {
    int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= p.n) return;
    V3 p0{p.px[q], p.py[q], p.pz[q]};
    V3 v0{p.vx[q], p.vy[q], p.vz[q]};
    V3 vPIC = sampleMac(g, cur, p0);
    V3 vn;
    if (aff.a[0]) {
        // APIC: the particle takes the sampled velocity and the field's gradients
        vn = vPIC;
        apicAffine(g, cur, p0, aff.a, q);
    } else {
        V3 vFLIP = sub(add(v0, vPIC), sampleMac(g, sav, p0));
        vn = add(scale(r1, vPIC), scale(r2, vFLIP));
    }
    p.vx[q] = vn.x;
    p.vy[q] = vn.y;
    p.vz[q] = vn.z;
    V3 k1 = vPIC;
    V3 k2 = sampleMac(g, cur, add(p0, scale((float)(0.5 * dt), k1)));
    V3 k3 = sampleMac(g, cur, add(p0, scale((float)(0.75 * dt), k2)));
    V3 p1 = add(p0, scale((float)(dt / (double)9.0f), add(add(scale(2.0f, k1), scale(3.0f, k2)), scale(4.0f, k3))));
    Box boundary = boundaryBox(g.ni, g.nj, g.nk, g.dx);
    boundary.expand((double)(-0.2f) * g.dx);
    V3 pn = resolveCollision(g, sp, N, cfl, p0, p1, boundary);
    if (moved) moved[q] = pn.x != p1.x || pn.y != p1.y || pn.z != p1.z;
    p.px[q] = pn.x;
    p.py[q] = pn.y;
    p.pz[q] = pn.z;
}

// ------------------------------------------------------------ removal -----
// _getMarkerParticleSpeedLimit (fluidsimulation.cpp:7723-7771): speed bins and max speed
__global__ __launch_bounds__(256) void speedHistKernel(Particles p, double stepv, int nbins,
                                                       int *counts, double *partial) {
    __shared__ double s[256];
    __shared__ int bins[1024];
    const bool local = nbins <= 1024;
    if (local)
        for (int b = threadIdx.x; b < nbins; b += blockDim.x) bins[b] = 0;
    __syncthreads();
    double acc = 0.0;
    for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < p.n; q += gridDim.x * blockDim.x) {
        double speed = (double)len3(V3{p.vx[q], p.vy[q], p.vz[q]});
        int idx = (int)fmin(floor(speed / stepv), (double)(nbins - 1));
        atomicAdd(local ? &bins[idx] : &counts[idx], 1);
        acc = fmax(acc, speed);
    }
    __syncthreads();
    if (local)
        for (int b = threadIdx.x; b < nbins; b += blockDim.x)
            if (bins[b]) atomicAdd(&counts[b], bins[b]);
    s[threadIdx.x] = acc;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < (unsigned)o) s[threadIdx.x] = fmax(s[threadIdx.x], s[threadIdx.x + o]);
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = s[0];
}
__global__ void speedOutlierKernel(Particles p, double lower, double thr, int *counts2) {
    int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= p.n) return;
    double speed = (double)len3(V3{p.vx[q], p.vy[q], p.vz[q]});
    if (speed >= lower && speed < thr) atomicAdd(&counts2[0], 1);
    if (speed >= thr) atomicAdd(&counts2[1], 1);
}

// _removeMarkerParticles (fluidsimulation.cpp:7773-7851), first pass: solid
// test (meshlevelset.h:333-339) and the cell key of every surviving particle
__global__ void removalKeysKernel(Grid g, Particles p, const float *sp, int *keys, int *slot,
                                  int *counts, int *maxCount) {
    int q = blockIdx.x * blockDim.x + threadIdx.x;
    int myCount = 0;
    if (q < p.n) {
        V3 x{p.px[q], p.py[q], p.pz[q]};
        int key = -1;
        if (!(nodalTrilinear(g, sp, x) < 0.0f)) {
            I3 c = posToIndex(x, g.dx);
            if (inRange(c.i, c.j, c.k, g.ni, g.nj, g.nk)) key = (int)flat(c.i, c.j, c.k, g.ni, g.nj);
        }
        keys[q] = key;
        if (key >= 0) {
            slot[q] = atomicAdd(&counts[key], 1);
            myCount = slot[q] + 1;
        }
    }
    // block max of the counts seen, then one atomic per block
    __shared__ int s[256];
    s[threadIdx.x] = myCount;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < (unsigned)o) s[threadIdx.x] = max(s[threadIdx.x], s[threadIdx.x + o]);
        __syncthreads();
    }
    if (threadIdx.x == 0) atomicMax(maxCount, s[0]);
}
// second pass: per-cell cap on the rank in ascending particle index (rank is
// null when no cell exceeds the cap), then the extreme-speed rule
__global__ void removalFlagsKernel(Particles p, const int *keys, const int *rank, int maxPerCell,
                                   int extreme, double maxSpeedSq, int *keep) {
    int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= p.n) return;
    int k = 1;
    if (keys[q] < 0)
        k = 0;
    else if (rank && rank[q] >= maxPerCell)
        k = 0;
    else if (extreme) {
        V3 v{p.vx[q], p.vy[q], p.vz[q]};
        if ((double)dot3(v, v) > maxSpeedSq) k = 0;
    }
    keep[q] = k;
}
__global__ void rankInCellKernel(const int *perm, const int *offsets, const int *keys, int ns, int *rank) {
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= ns) return;
    int q = perm[s];
    rank[q] = s - offsets[keys[q]];
}

// ParticleSystem::removeParticles (particlesystem.cpp:88-99): stable compaction
__global__ void compactKernel(Particles src, Particles dst, const int *keep, const int *rank) {
    int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= src.n || !keep[q]) return;
    int d = rank[q];
    dst.px[d] = src.px[q];
    dst.py[d] = src.py[q];
    dst.pz[d] = src.pz[q];
    dst.vx[d] = src.vx[q];
    dst.vy[d] = src.vy[q];
    dst.vz[d] = src.vz[q];
}

}  // namespace

void cellSort(Grid g, const Particles &p, CellSort &cs, int *scanScratch, hipStream_t s) {
    size_t N = (size_t)g.ni * g.nj * g.nk;
    hipCheck(hipMemsetAsync(cs.counts, 0, sizeof(int) * (N + 1), s), "hipMemsetAsync");
    if (p.n > 0)
        hipLaunchKernelGGL(cellKeysKernel, dim3(blocksFor(p.n, TPB)), dim3(TPB), 0, s, g, p.px, p.py,
                           p.pz, p.n, cs.keys, cs.slot, cs.counts);
    exclusiveScan(cs.counts, cs.offsets, N, scanScratch, s);
    if (p.n > 0) {
        hipLaunchKernelGGL(scatterKernel, dim3(blocksFor(p.n, TPB)), dim3(TPB), 0, s, cs.keys, cs.slot,
                           cs.offsets, p.n, cs.perm);
        hipLaunchKernelGGL(segmentSortKernel, dim3(blocksFor(N, TPB)), dim3(TPB), 0, s, cs.offsets, N,
                           cs.perm);
        hipLaunchKernelGGL(gatherSortedKernel, dim3(blocksFor(p.n, TPB)), dim3(TPB), 0, s, cs.perm,
                           cs.offsets + N, p, cs.spx, cs.spy, cs.spz, cs.svx, cs.svy, cs.svz);
    }
}

namespace {
__global__ void countKeysKernel(const int *keys, size_t n, int *slot, int *counts) {
    const size_t r = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    const int key = keys[r];
    // arrival slots are reordered by the segment sort below, so the order
    // of these integer atomics does not reach the result
    if (key >= 0) slot[r] = atomicAdd(&counts[key], 1);
}
}  // namespace

void sortKeys(const int *keys, size_t n, size_t nb, int *slot, int *counts, int *offsets, int *perm,
              int *scanScratch, hipStream_t s) {
    hipCheck(hipMemsetAsync(counts, 0, sizeof(int) * (nb + 1), s), "hipMemsetAsync");
    if (n > 0)
        hipLaunchKernelGGL(countKeysKernel, dim3(blocksFor(n, TPB)), dim3(TPB), 0, s, keys, n, slot, counts);
    exclusiveScan(counts, offsets, nb, scanScratch, s);
    if (n > 0) {
        hipLaunchKernelGGL(scatterKernel, dim3(blocksFor(n, TPB)), dim3(TPB), 0, s, keys, slot, offsets, (int)n,
                           perm);
        hipLaunchKernelGGL(segmentSortKernel, dim3(blocksFor(nb, TPB)), dim3(TPB), 0, s, offsets, nb, perm);
    }
}

void liquidSDF(Grid g, const CellSort &cs, int np, double radius, float *phi, hipStream_t s) {
    int nbi = (g.ni + BW - 1) / BW, nbj = (g.nj + BW - 1) / BW, nbk = (g.nk + BW - 1) / BW;
    (void)np;
    hipLaunchKernelGGL(liquidSDFKernel, dim3(nbi * nbj * nbk), dim3(256), 0, s, g, cs.offsets, cs.spx,
                       cs.spy, cs.spz, radius, phi, nbi, nbj);
}

void liquidSDFPostprocess(Grid g, const float *sp, float *phi, hipStream_t s) {
    size_t n = (size_t)g.ni * g.nj * g.nk;
    hipLaunchKernelGGL(sdfPostKernel, dim3(blocksFor(n, TPB)), dim3(TPB), 0, s, g, sp, phi);
}

void p2g(Grid g, const CellSort &cs, double radius, float *u, float *v, float *w, uint8_t *vu,
         uint8_t *vv, uint8_t *vw, hipStream_t s, float *wsu, float *wsv, float *wsw) {
    const int tilesI = (g.ni + 1 + P2T_I - 1) / P2T_I, tilesJ = (g.nj + 1 + P2T_J - 1) / P2T_J,
              tilesK = (g.nk + 1 + P2T_K - 1) / P2T_K;
    P2GOut out{{u, v, w}, {vu, vv, vw}, {wsu, wsv, wsw}};
    hipLaunchKernelGGL(p2gTileKernel, dim3((unsigned)tilesI * tilesJ * tilesK), dim3(P2_THREADS), 0, s, g,
                       cs.offsets, cs.perm, cs.spx, cs.spy, cs.spz, cs.svx, cs.svy, cs.svz, radius, out, tilesI,
                       tilesJ);
}

void maxSpeedSq(const Particles &p, double *partial, double *out, hipStream_t s) {
    int nb = (int)blocksFor(p.n > 0 ? p.n : 1, TPB * 8);
    if (nb > 1024) nb = 1024;
    hipLaunchKernelGGL(maxSpeedStage1, dim3(nb), dim3(256), 0, s, p, partial);
    hipLaunchKernelGGL(maxStage2, dim3(1), dim3(256), 0, s, partial, nb, out);
}

void g2pAdvect(Grid g, Particles p, MacView cur, MacView saved, const float *solidNodes,
               const uint8_t *nearSolid, int ci, int cj, int ck, float r1, float r2, double dt,
               double cfl, hipStream_t s, uint8_t *moved, const AffineArrays *affine) {
    if (p.n == 0) return;
    NearGrid N{nearSolid, ci, cj, ck, 3 * g.dx};
    const AffineArrays A = affine ? *affine : AffineArrays{};
    hipLaunchKernelGGL(g2pAdvectKernel, dim3(blocksFor(p.n, TPB)), dim3(TPB), 0, s, g, p, cur, saved,
                       solidNodes, N, r1, r2, dt, cfl, moved, A);
}

void speedHistogram(const Particles &p, double stepv, int nbins, int *counts, double *partial,
                    double *maxSpeed, hipStream_t s) {
    hipCheck(hipMemsetAsync(counts, 0, sizeof(int) * nbins, s), "hipMemsetAsync");
    int nb = (int)blocksFor(p.n > 0 ? p.n : 1, TPB * 8);
    if (nb > 1024) nb = 1024;
    hipLaunchKernelGGL(speedHistKernel, dim3(nb), dim3(256), 0, s, p, stepv, nbins, counts, partial);
    hipLaunchKernelGGL(maxStage2, dim3(1), dim3(256), 0, s, partial, nb, maxSpeed);
}

void speedOutliers(const Particles &p, double lower, double thr, int *counts2, hipStream_t s) {
    hipCheck(hipMemsetAsync(counts2, 0, sizeof(int) * 2, s), "hipMemsetAsync");
    if (p.n > 0)
        hipLaunchKernelGGL(speedOutlierKernel, dim3(blocksFor(p.n, TPB)), dim3(TPB), 0, s, p, lower, thr,
                           counts2);
}

void removalFlags(Grid g, const Particles &p, const float *sp, int *keys, int *counts, int *slot,
                  int *offsets, int *perm, int *scanScratch, int maxPerCell, bool extreme,
                  double maxSpeedSq, int *keep, int *maxCount, hipStream_t s) {
    if (p.n == 0) return;
    size_t N = (size_t)g.ni * g.nj * g.nk;
    hipCheck(hipMemsetAsync(counts, 0, sizeof(int) * (N + 1), s), "hipMemsetAsync");
    hipCheck(hipMemsetAsync(maxCount, 0, sizeof(int), s), "hipMemsetAsync");
    hipLaunchKernelGGL(removalKeysKernel, dim3(blocksFor(p.n, TPB)), dim3(TPB), 0, s, g, p, sp, keys,
                       slot, counts, maxCount);
    int hmax = 0;
    hipCheck(hipMemcpyAsync(&hmax, maxCount, sizeof(int), hipMemcpyDeviceToHost, s), "hipMemcpyAsync");
    hipCheck(hipStreamSynchronize(s), "hipStreamSynchronize");
    const int *rank = nullptr;
    if (hmax > maxPerCell) {
        // exact cap: stable sort of the survivors by cell, rank inside the cell
        exclusiveScan(counts, offsets, N, scanScratch, s);
        hipLaunchKernelGGL(scatterKernel, dim3(blocksFor(p.n, TPB)), dim3(TPB), 0, s, keys, slot, offsets,
                           p.n, perm);
        hipLaunchKernelGGL(segmentSortKernel, dim3(blocksFor(N, TPB)), dim3(TPB), 0, s, offsets, N, perm);
        int ns = 0;
        hipCheck(hipMemcpyAsync(&ns, offsets + N, sizeof(int), hipMemcpyDeviceToHost, s), "hipMemcpyAsync");
        hipCheck(hipStreamSynchronize(s), "hipStreamSynchronize");
        if (ns > 0)
            hipLaunchKernelGGL(rankInCellKernel, dim3(blocksFor(ns, TPB)), dim3(TPB), 0, s, perm, offsets,
                               keys, ns, slot);
        rank = slot;
    }
    hipLaunchKernelGGL(removalFlagsKernel, dim3(blocksFor(p.n, TPB)), dim3(TPB), 0, s, p, keys, rank,
                       maxPerCell, extreme ? 1 : 0, maxSpeedSq, keep);
}

namespace {
__global__ void compactArraysKernel(AffineArrays src, AffineArrays dst, int n, const int *keep, const int *rank) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n || !keep[q]) return;
    const int r = rank[q];
#pragma unroll
    for (int e = 0; e < 9; e++) dst.a[e][r] = src.a[e][q];
}
}  // namespace

void compactAffine(const AffineArrays &src, const AffineArrays &dst, int n, const int *keep, const int *rank,
                   hipStream_t s) {
    if (n > 0)
        hipLaunchKernelGGL(compactArraysKernel, dim3(blocksFor(n, TPB)), dim3(TPB), 0, s, src, dst, n, keep, rank);
}

void compactParticles(Particles &src, Particles &dst, const int *keep, int *rank, int *scanScratch,
                      hipStream_t s) {
    exclusiveScan(keep, rank, src.n, scanScratch, s);
    if (src.n > 0)
        hipLaunchKernelGGL(compactKernel, dim3(blocksFor(src.n, TPB)), dim3(TPB), 0, s, src, dst, keep, rank);
    int total = 0;
    hipCheck(hipMemcpyAsync(&total, rank + src.n, sizeof(int), hipMemcpyDeviceToHost, s), "hipMemcpyAsync");
    hipCheck(hipStreamSynchronize(s), "hipStreamSynchronize");
    dst.n = total;
}

void advectRK3(Grid g, int np, const float *px, const float *py, const float *pz, MacView cur, double dt, float *ox,
               float *oy, float *oz, hipStream_t s) {
    if (np == 0) return;
    hipLaunchKernelGGL(advectRK3Kernel, dim3(blocksFor(np, TPB)), dim3(TPB), 0, s, g, np, px, py, pz, cur, dt, ox, oy,
                       oz);
}

void advectRK3Adjoint(Grid g, int np, const float *px, const float *py, const float *pz, MacView cur, double dt,
                      const float *xbar, float *pbar, float *pts, float *cot, hipStream_t s, const BlendAdj *blend) {
    if (np == 0) return;
    const BlendAdj B = blend ? *blend : BlendAdj{};
    hipLaunchKernelGGL(advectRK3AdjKernel, dim3(blocksFor(np, TPB)), dim3(TPB), 0, s, g, np, px, py, pz, cur, dt, xbar,
                       pbar, pts, cot, B);
}

// G2P of the MAC field at the particle positions (the sample the G2P blend
// calls vPIC): v = V(p)
__global__ void sampleParticlesKernel(Grid g, int np, const float *px, const float *py, const float *pz, MacView M,
                                      float *vx, float *vy, float *vz) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= np) return;
    const V3 v = sampleMac(g, M, V3{px[q], py[q], pz[q]});
    vx[q] = v.x;
    vy[q] = v.y;
    vz[q] = v.z;
}

void sampleParticles(Grid g, int np, const float *px, const float *py, const float *pz, MacView M, float *vx,
                     float *vy, float *vz, hipStream_t s) {
    if (np == 0) return;
    hipLaunchKernelGGL(sampleParticlesKernel, dim3(blocksFor(np, TPB)), dim3(TPB), 0, s, g, np, px, py, pz, M, vx, vy,
                       vz);
}

// Particle density on the cells: rho_c = sum over particles within r of the
// cell centre of the transfer kernel's weight w(|x_p - c|^2), gathered from
// the 27 cells around c in cell order, then the cells' sorted order.
__global__ void densityKernel(Grid g, const int *offsets, const float *spx, const float *spy, const float *spz,
                              double radius, float *rho) {
    const size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t N = (size_t)g.ni * g.nj * g.nk;
    if (c >= N) return;
    const int i = (int)(c % g.ni), j = (int)((c / g.ni) % g.nj), k = (int)(c / ((size_t)g.ni * g.nj));
    const P2GWeight W = p2gWeight(radius);
    const double h = 0.5 * g.dx;
    const double cx = i * g.dx + h, cy = j * g.dx + h, cz = k * g.dx + h;
    double acc = 0.0;
    for (int dk = -1; dk <= 1; dk++)
        for (int dj = -1; dj <= 1; dj++)
            for (int di = -1; di <= 1; di++) {
                const int a = i + di, b = j + dj, e = k + dk;
                if (!inRange(a, b, e, g.ni, g.nj, g.nk)) continue;
                const size_t nc = flat(a, b, e, g.ni, g.nj);
                for (int s = offsets[nc]; s < offsets[nc + 1]; s++) {
                    const double dx = cx - spx[s], dy = cy - spy[s], dz = cz - spz[s];
                    const double d2 = dx * dx + dy * dy + dz * dz;
                    if (d2 < (double)W.rsq) acc += 1.0 - W.coef1 * d2 * d2 * d2 + W.coef2 * d2 * d2 - W.coef3 * d2;
                }
            }
    rho[c] = (float)acc;
}

// xbar_p = sum_c rhobar_c w'(d2) (-2) (c - x_p) over the cells within r
__global__ void densityAdjKernel(Grid g, int np, const float *px, const float *py, const float *pz, double radius,
                                 const float *rhobar, float *xbar) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= np) return;
    const P2GWeight W = p2gWeight(radius);
    const double x = px[q], y = py[q], z = pz[q];
    const I3 c0 = posToIndex(V3{px[q], py[q], pz[q]}, g.dx);
    const double h = 0.5 * g.dx;
    double xb[3] = {0.0, 0.0, 0.0};
    for (int dk = -1; dk <= 1; dk++)
        for (int dj = -1; dj <= 1; dj++)
            for (int di = -1; di <= 1; di++) {
                const int a = c0.i + di, b = c0.j + dj, e = c0.k + dk;
                if (!inRange(a, b, e, g.ni, g.nj, g.nk)) continue;
                const double ex = a * g.dx + h - x, ey = b * g.dx + h - y, ez = e * g.dx + h - z;
                const double d2 = ex * ex + ey * ey + ez * ez;
                if (!(d2 < (double)W.rsq)) continue;
                const double dw = -3.0 * W.coef1 * d2 * d2 + 2.0 * W.coef2 * d2 - W.coef3;
                const double s = (double)rhobar[flat(a, b, e, g.ni, g.nj)] * dw * -2.0;
                xb[0] += s * ex;
                xb[1] += s * ey;
                xb[2] += s * ez;
            }
    for (int e = 0; e < 3; e++) xbar[3 * (size_t)q + e] = (float)xb[e];
}

void densityGrid(Grid g, const CellSort &cs, double radius, float *rho, hipStream_t s) {
    const size_t N = (size_t)g.ni * g.nj * g.nk;
    hipLaunchKernelGGL(densityKernel, dim3(blocksFor(N, TPB)), dim3(TPB), 0, s, g, cs.offsets, cs.spx, cs.spy, cs.spz,
                       radius, rho);
}

void densityAdjoint(Grid g, int np, const float *px, const float *py, const float *pz, double radius,
                    const float *rhobar, float *xbar, hipStream_t s) {
    if (np == 0) return;
    hipLaunchKernelGGL(densityAdjKernel, dim3(blocksFor(np, TPB)), dim3(TPB), 0, s, g, np, px, py, pz, radius, rhobar,
                       xbar);
}

}  // namespace ff
