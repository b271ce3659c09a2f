// kernels.h -- host-side launchers of the HIP kernels of the FLIP substep.
// Every launcher enqueues on the given stream and returns without waiting.
// Device pointers only; sizes in elements.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <vector>

#include "ffdevice.h"
#include "pencil_types.h"

namespace ff {

// throws std::runtime_error naming `what` when e is not hipSuccess
void hipCheck(hipError_t e, const char *what);

// ---- static mesh obstacles (kernels_mesh.hip) ----
// A triangle mesh in host memory: 3 floats per vertex, 3 vertex indices per
// triangle.
struct ObstacleMesh {
    const float *vertices;
    int numVertices;
    const int *triangles;
    int numTriangles;
    bool inversed;
    float meshExpansion;  // used by the 25-or-more-island path only (the reference ignores it otherwise)
    const float *velocities = nullptr;  // per vertex (3 floats), or null: at rest
    bool animated = false;  // animated objects are unioned after the static ones
};
// The solid's velocity data (MeshLevelSet velocity grids, meshlevelset.cpp:
// 1320-1425, 603-700): per face the weighted velocity f and the weight w
// summed over the domain and the obstacles, with the domain-sized scratch of
// one rebuild (an obstacle's level set and its faces; the union of the
// inversed obstacles).  All device memory, carved from one block of
// solidVelocityBytes(g) bytes by solidVelocityCarve.
struct SolidVelocityData {
    float *f[3], *w[3];    // accumulated over the solid (u, v, w faces)
    float *tf[3], *tw[3];  // one obstacle's, or the inversed union's
    float *tphi, *iphi;    // nodal: one obstacle's level set; the inversed union
    int *tobj, *iobj;      // nodal: their closest mesh objects
    int *ttri, *itri;      // nodal: their closest triangles
};
size_t solidVelocityBytes(Grid g);
SolidVelocityData solidVelocityCarve(Grid g, void *block);
// Unions the obstacles' signed distance level sets into the nodal solid SDF
// `solid` (device, (ni+1)(nj+1)(nk+1)), which holds the domain SDF: normal
// obstacles in order, then the negated union of the inversed ones.  Takes the
// sign tests' jitter from rand() (three draws per test, the domain box's
// first) and returns the number of draws.  Synchronizes the stream.
// nodeObject (optional, device, nodal): each node's closest mesh object
// (MeshLevelSet::getClosestMeshObject): -1 none, 0 the domain object (on
// the domain SDF's exact band when domainObject), q + 1 obstacle q.
// velocity (optional): the solid update with velocity data
// (FluidSimulation::_updateSolidLevelSet, fluidsimulation.cpp:5219-5426): the
// domain's weights, then the static obstacles and then the animated ones,
// each group's normal obstacles one by one and then the negated union of its
// inversed ones, every obstacle built on its own domain-sized level set
// (meshes and vertex velocities as given, at the substep's frame progress);
// the velocity data is left in velocity->f / w.
int meshObstacleSolid(Grid g, const ObstacleMesh *obstacles, int n, float *solid, int *nodeObject, bool domainObject,
                      hipStream_t s, SolidVelocityData *velocity = nullptr);
// normalizeVelocityGrid (meshlevelset.cpp:640-700, 1739-1757): velocity =
// f / w where w exceeds 1e-6, else 0, then 5 extrapolation layers from those
// faces; writes us, vs, ws (device face arrays).  valid and status: scratch of
// the largest face array's size.
void normalizeSolidVelocity(Grid g, const SolidVelocityData &d, float *us, float *vs, float *ws, uint8_t *valid,
                            uint8_t *status, hipStream_t s);
// The weight grid's cell weights (fluidsimulation.cpp:6140-6147,
// MeshLevelSet::_getCellWeight meshlevelset.cpp:1491-1515): 1 - the solid's
// volume fraction in each cell, clamped to [0, 1] (device, ni nj nk).
void cellWeights(Grid g, const float *solid, float *wc, hipStream_t s);
// PressureSolver::_conditionSolidVelocityField (pressuresolver.cpp:161-312):
// the interior liquid cells (phi < 0) connected through faces of weight
// >= 1e-6, never through a border cell, form groups; in a group of more than
// one cell none of which has such a face to an air cell the solid velocities
// of all its cells' faces become 0 (us, vs, ws in place).  scratch: 2 ni nj nk
// ints.
void conditionSolidVelocity(Grid g, const float *phi, const float *wu, const float *wv, const float *ww, float *us,
                            float *vs, float *ws, int *scratch, hipStream_t s);

// One mesh's fast signed distance level set on `cells` cells at node offset
// `offset` (device phi, (ni+1)(nj+1)(nk+1)); returns the rand() draws.
int meshLevelSetFastDevice(Grid cells, I3 offset, const ObstacleMesh &mesh, float *phi, hipStream_t s);

// ---- scan / reductions (kernels_scan.hip) ----
// Exclusive prefix sum of n int32 values; out may alias in; out[n] = total.
// scratch must hold scanScratchInts(n) ints.
size_t scanScratchInts(size_t n);
void exclusiveScan(const int *in, int *out, size_t n, int *scratch, hipStream_t s);
// Deterministic FP64 dot product: fixed two-stage tree.  partial must hold
// reduceBlocks(n) doubles; result is written to *out (device).
int reduceBlocks(size_t n);
void dot(const double *a, const double *b, size_t n, double *partial, double *out, hipStream_t s);
void dotf(const float *a, const float *b, size_t n, double *partial, double *out, hipStream_t s);
// sum of (double)a[i] * b[i] (FP32 a, FP64 b), same tree
void dotfd(const float *a, const double *b, size_t n, double *partial, double *out, hipStream_t s);
void absMax(const double *a, size_t n, double *partial, double *out, hipStream_t s);
// BLAS::dot in the reference's order (blaswrapper.h:27-63): `threads` (the
// engine's thread count) decides the number of intervals, each summed from
// zero in index order, then the interval sums in order.  partial must hold
// referenceDotChunks(n, threads) doubles.
int referenceDotChunks(size_t n, int threads);
void dotReference(const double *a, const double *b, size_t n, int threads, double *partial, double *out,
                  hipStream_t s);
void absMaxf(const float *a, size_t n, double *partial, double *out, hipStream_t s);
// The reductions above run reduceBlocks(n) blocks of REDUCE_THREADS threads.
// Logical block L takes the contiguous span of rows [L*sp, (L+1)*sp),
// sp = reduceSpan(n, blocks), thread t its rows L*sp + t + q*REDUCE_THREADS in
// order; the block folds its threads' values (reduceFold), and one block folds
// the partials in order.  Fused kernels that use the same blocks and spans
// produce the same bits.
constexpr int REDUCE_THREADS = 256;
constexpr int XCDS = 8;  // MI355X: 8 accelerator dies, workgroup b runs on die b % 8
__host__ __device__ inline size_t reduceSpan(size_t n, unsigned blocks) {
    const size_t per = (n + blocks - 1) / blocks;
    return (per + REDUCE_THREADS - 1) / REDUCE_THREADS * REDUCE_THREADS;
}
// the logical block of this workgroup: blocks are renumbered so that each die
// runs one contiguous run of spans (blocks is a multiple of XCDS or below it),
// and the rows next to a block's rows sit in the same die's L2
__device__ inline unsigned reduceLogicalBlock() {
    const unsigned b = blockIdx.x, g = gridDim.x;
    return g % XCDS == 0 ? (b % XCDS) * (g / XCDS) + b / XCDS : b;
}
// fold one double per thread of a REDUCE_THREADS block, OP 0 sum, 1 max, in a
// fixed order: a wave64 butterfly (every lane ends with the same bits), then
// the four wave results in wave order.  The result is valid in thread 0.
template <int OP>
__device__ inline double reduceFold(double v) {
    __shared__ double waves[REDUCE_THREADS / 64];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        const double o = __shfl_xor(v, off, 64);
        v = OP == 0 ? v + o : fmax(v, o);
    }
    if ((threadIdx.x & 63) == 0) waves[threadIdx.x >> 6] = v;
    __syncthreads();
    double acc = waves[0];
#pragma unroll
    for (int q = 1; q < REDUCE_THREADS / 64; q++) acc = OP == 0 ? acc + waves[q] : fmax(acc, waves[q]);
    return acc;
}
// outHost (optional): device alias of mapped pinned host memory that also
// receives the result
void reduceFinish(const double *partial, int nb, double *out, bool isMax, hipStream_t s, double *outHost = nullptr);

// ---- setup (kernels_setup.hip) ----
void solidBoxPhi(Grid g, float *phiNodes, hipStream_t s);
void faceWeights(Grid g, const float *phiNodes, float *wu, float *wv, float *ww, hipStream_t s);
// near-solid coarse grid (cells of 3dx), marked then feathered `layers` times
void nearSolidGrid(Grid g, const float *phiNodes, uint8_t *near, uint8_t *tmp, int ci, int cj,
                   int ck, int layers, hipStream_t s);

// ---- particles (kernels_particles.hip) ----
struct Particles {
    float *px, *py, *pz, *vx, *vy, *vz;
    int n;
};
struct CellSort {
    int *keys;     // P: flat cell of floor(p/dx) or -1
    int *slot;     // P: arrival slot inside the cell (scratch)
    int *counts;   // N+1
    int *offsets;  // N+1 (exclusive scan of counts)
    int *perm;     // P: sorted position -> particle index (stable)
    // sorted copies
    float *spx, *spy, *spz, *svx, *svy, *svz;
};
void cellSort(Grid g, const Particles &p, CellSort &cs, int *scanScratch, hipStream_t s);
void liquidSDF(Grid g, const CellSort &cs, int np, double radius, float *phi, hipStream_t s);
void liquidSDFPostprocess(Grid g, const float *solidNodes, float *phi, hipStream_t s);
// P2G of all three components into u, v, w with validity flags; wsu/wsv/wsw
// (optional) receive each face's weight sum
void p2g(Grid g, const CellSort &cs, double radius, float *u, float *v, float *w, uint8_t *vu,
         uint8_t *vv, uint8_t *vw, hipStream_t s, float *wsu = nullptr, float *wsv = nullptr,
         float *wsw = nullptr);
// stable counting sort of n keys (-1: no bucket) into nb buckets: offsets
// (nb + 1, exclusive) and perm (the ids of each bucket in ascending order);
// slot holds n ints, counts nb + 1, scanScratch scanScratchInts(nb + 1)
void sortKeys(const int *keys, size_t n, size_t nb, int *slot, int *counts, int *offsets, int *perm,
              int *scanScratch, hipStream_t s);
// APIC P2G (kernels_apic.hip): per direction, records of (particle, member
// block) keyed by the particle's cell in the block's frame, sorted, then
// gathered per face in particle order.  vel: the three velocity components
// (SoA); aff: AFFINEX.xyz, AFFINEY.xyz, AFFINEZ.xyz (SoA, nine arrays).
struct ApicScratch {
    uint8_t *mark, *active;          // apicBlocks(g) bytes each
    int *keys, *slot, *perm;         // apicRecords(np) ints each
    int *counts, *offsets;           // apicBuckets(g) + 1 ints each
    int *scan;                       // scanScratchInts(apicBuckets(g) + 1) ints
};
size_t apicRecords(int np);
size_t apicBuckets(Grid g);
size_t apicBlocks(Grid g);
void p2gApic(Grid g, int np, const float *px, const float *py, const float *pz, const float *const vel[3],
             const float *const aff[9], double radius, float *const field[3], uint8_t *const valid[3],
             ApicScratch &A, hipStream_t s);
// max over particles of dot(v, v) (float) as double
void maxSpeedSq(const Particles &p, double *partial, double *out, hipStream_t s);

// G2P blend + RK3 + collision; writes new positions in place.  moved
// (optional, P bytes) flags the particles whose position the collision
// response changed.  With affine arrays (AFFINEX.xyz, AFFINEY.xyz,
// AFFINEZ.xyz, SoA) the G2P is APIC's: the velocity becomes the sample and
// the affine vectors the field's gradients.
struct MacView {
    const float *u, *v, *w;
};
struct AffineArrays {
    float *a[9] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
};
void g2pAdvect(Grid g, Particles p, MacView cur, MacView saved, const float *solidNodes,
               const uint8_t *nearSolid, int ci, int cj, int ck, float picRatio, float flipRatio,
               double dt, double cfl, hipStream_t s, uint8_t *moved = nullptr,
               const AffineArrays *affine = nullptr);
// the affine arrays of the particles compactParticles just kept (same keep
// flags and ranks; n = particle count before the compaction)
void compactAffine(const AffineArrays &src, const AffineArrays &dst, int n, const int *keep, const int *rank,
                   hipStream_t s);
// v = trilinear sample of the MAC field at the positions (SoA)
void sampleParticles(Grid g, int np, const float *px, const float *py, const float *pz, MacView M, float *vx,
                     float *vy, float *vz, hipStream_t s);
// particle density on the cells: the transfer kernel's weights summed at the
// cell centres; and its adjoint in the positions (xbar: P x 3)
void densityGrid(Grid g, const CellSort &cs, double radius, float *rho, hipStream_t s);
void densityAdjoint(Grid g, int np, const float *px, const float *py, const float *pz, double radius,
                    const float *rhobar, float *xbar, hipStream_t s);
// RK3 advection through the MAC field without collision (positions out), and
// its adjoint: pbar (P x 3) for xbar (P x 3), plus the 3P sample points
// (pts: their x, then y, then z coordinates) with their velocity cotangents
// (cot: 3P x 3) for the sampling adjoint
void advectRK3(Grid g, int np, const float *px, const float *py, const float *pz, MacView cur, double dt, float *ox,
               float *oy, float *oz, hipStream_t s);
// blend: the G2P blend's terms (vbarOut, P x 3: the new velocities'
// cotangent; vbar, P x 3: the old velocities'; cotSav, P x 3: the saved
// field's samples at the positions); with it the adjoint is that of
// g2pAdvect with the collision response taken as the identity
struct BlendAdj {
    MacView sav{nullptr, nullptr, nullptr};
    float r1 = 0.0f, r2 = 0.0f;
    const float *vbarOut = nullptr;
    float *vbar = nullptr, *cotSav = nullptr;
};
void advectRK3Adjoint(Grid g, int np, const float *px, const float *py, const float *pz, MacView cur, double dt,
                      const float *xbar, float *pbar, float *pts, float *cot, hipStream_t s,
                      const BlendAdj *blend = nullptr);
// speed statistics for the extreme-velocity rule: counts[nbins] and max speed
void speedHistogram(const Particles &p, double stepv, int nbins, int *counts, double *partial,
                    double *maxSpeed, hipStream_t s);
void speedOutliers(const Particles &p, double lower, double thr, int *counts2, hipStream_t s);
// removal flags: solid test, per-cell cap, extreme speed.  keep[P] = 1 to keep.
void removalFlags(Grid g, const Particles &p, const float *solidNodes, int *keys, int *counts,
                  int *slot, int *offsets, int *perm, int *scanScratch, int maxPerCell,
                  bool extremeRemoval, double maxSpeedSq, int *keep, int *maxCount,
                  hipStream_t s);
// stable compaction of the six particle arrays by keep flags (scan in rank)
void compactParticles(Particles &src, Particles &dst, const int *keep, int *rank,
                      int *scanScratch, hipStream_t s);

// ---- grid stencils (kernels_grid.hip) ----
void extrapolateInit(int wi, int wj, int wk, const uint8_t *valid, uint8_t *stat, hipStream_t s);
void extrapolateLayer(int wi, int wj, int wk, float *field, uint8_t *stat, int layer, hipStream_t s);
// adjoint of the layered extrapolation (for the status the forward's
// `layers` layers left), in place on the face cotangent bar
void extrapolateAdjoint(int wi, int wj, int wk, float *bar, const uint8_t *stat, int layers, hipStream_t s);
void addToFaces(float *a, size_t n, float inc, hipStream_t s);
// _constrainVelocityFieldThread (fluidsimulation.cpp:6391-6442) with friction
// 0: closed faces (weight 0) take the solid's face velocity sv (device, the
// face array's size; null: a static solid, velocity 0)
void constrainFaces(float *a, const float *w, size_t n, hipStream_t s, const float *sv = nullptr);
// the same with friction: f = the mean of the face's four nodes' closest
// objects' friction (fluidsimulation.cpp:6292-6353), objFriction[object]
// (none: 0); partly open faces become f sv + (1 - f) a; dir 0/1/2 = u/v/w
void constrainFacesFriction(Grid g, int dir, float *a, const float *w, const int *nodeObject, const float *objFriction,
                            hipStream_t s, const float *sv = nullptr);
void applyPressure(Grid g, double dt, const float *p, const float *phi, const float *wu,
                   const float *wv, const float *ww, float *u, float *v, float *w, uint8_t *vu,
                   uint8_t *vv, uint8_t *vw, hipStream_t s, TensionView tension = TensionView{});

// ---- surface tension (particlelevelset.cpp:196-292, pressuresolver.cpp:314-606) ----
// Device buffers of the curvature grid and the cluster status for one grid,
// carved from one allocation of tensionBytes(g) bytes.
struct TensionBuffers {
    float *a = nullptr, *b = nullptr;  // the reinitialization's two fields
    float *curvature = nullptr;        // N
    int8_t *status = nullptr;          // N
    uint8_t *valid = nullptr, *solver = nullptr, *cv = nullptr, *ext = nullptr;  // N
    int *parent = nullptr, *count = nullptr, *rootFlags = nullptr;             // N
    uint8_t *blocks = nullptr, *blocks2 = nullptr, *blocks4 = nullptr;
    float *partial = nullptr;        // per-block maxima
    int *any = nullptr;              // 1 int
    const float *surface = nullptr;  // the reinitialized field after curvatureGrid (a or b)
};
size_t tensionBytes(Grid g);
void tensionCarve(Grid g, void *block, TensionBuffers &B);
// the reinitialized surface field (B.surface) and the extrapolated curvature
// (B.curvature) of the cell-centred liquid phi; waits for the stream once per
// reinitialization step (the reference's convergence test)
void curvatureGrid(Grid g, const float *phi, TensionBuffers &B, hipStream_t s);
// B.status: 16 on the border, 1 / 2 in 4^3 blocks wholly inside / outside the
// liquid, 4 / 8 in the BAD clusters (connected same-phase regions of the
// remaining cells with fewer than 36 cells and no same-phase neighbour in
// such a block), 0 elsewhere
void tensionClusterStatus(Grid g, const float *phi, TensionBuffers &B, hipStream_t s);

// ---- pressure (kernels_pressure.hip) ----
// MIC(0) operands of one row as the band sweeps read them: invdiag and the
// three L entries the sweep multiplies neighbour results by (forward: those of
// the -x/-y/-z neighbour rows toward the row; backward: the row's own toward
// +x/+y/+z)
template <class T>
struct alignas(4 * sizeof(T)) MicRec {
    T inv, a, b, c;
};
struct PressureSys {
    int *mask;      // N+1: pressure-cell flag, then scanned into row ids
    int *rowOf;     // N+1 (exclusive scan of mask)
    int *rows;      // M: flat cell of each row (k,j,i order)
    int *cellRow;   // N: row of each cell, or -1
    double *rhs, *diag, *ox, *oy, *oz, *mx, *my, *mz;  // M
    int *nbr;       // 6M: -x,+x,-y,+y,-z,+z row ids or -1
    double *invdiag, *lx, *ly, *lz;  // M: MIC(0) factor (L entries toward +x/+y/+z)
    double *lmx, *lmy, *lmz;         // M: L entries of the -x/-y/-z neighbour rows toward the row
    float *finvdiag, *flx, *fly, *flz;  // FP32 copies for the mixed-precision inner solve
    float *flmx, *flmy, *flmz;
    MicRec<double> *recFwd, *recBwd;  // M: band-sweep records (FP64)
    MicRec<float> *frecFwd, *frecBwd;  // M: band-sweep records (FP32 copies)
    float *fdiag, *fox, *foy, *foz, *fmx, *fmy, *fmz;
    int *levelCount;  // L+1
    int *levelStart;  // L+1
    int *levelFill;   // L
    int *levelRows;   // M: rows grouped by i+j+k
    // row bits of each 32-cell word of every (j, k) line along i, and the row
    // of the word's first liquid cell (nj*nk lines x nw words; SpMV and the
    // line-pipeline sweeps find neighbour rows in them)
    LineWord *words;
    int nw, ni, nj, nk;
    int m;
    int nLevels;
};
void pressureMask(Grid g, const float *phi, int *mask, hipStream_t s);
void pressureRows(Grid g, const int *mask, const int *rowOf, int *rows, int *cellRow, hipStream_t s);
// the moving solid's terms of the divergence: face velocities and cell weights
// (all null for a static solid)
struct SolidTerms {
    const float *us = nullptr, *vs = nullptr, *ws = nullptr, *wc = nullptr;
};
// tension: the surface tension terms of the divergence (status null: none)
void pressureAssemble(Grid g, double dt, const float *u, const float *v, const float *w,
                      const float *wu, const float *wv, const float *ww, const float *phi,
                      PressureSys &S, hipStream_t s, SolidTerms solid = SolidTerms{},
                      TensionView tension = TensionView{});
void pressureLevels(Grid g, PressureSys &S, int *scanScratch, hipStream_t s);
// MIC(0) factor and application z = (L L^T)^-1 r; hyperplane-by-hyperplane
// launches driven by the host copy of the level table
void mic0Factor(PressureSys &S, const int *hLevelStart, hipStream_t s);
// the same invdiag in bands of 13 hyperplanes over the sweeps' tiles (host and
// device tile ranges from mic0BandTileRanges); lfactor entries not included
void mic0FactorBands(Grid g, const PressureSys &S, const int *hLevelStart, const int *hTileLo, const int *hTileHi,
                     const int *tileLo, const int *tileHi, hipStream_t s);
int mic0FactorBandLaunches(const PressureSys &S, const int *hLevelStart);
void mic0LFactor(PressureSys &S, hipStream_t s);
void mic0Apply(const PressureSys &S, const int *hLevelStart, const double *r, double *z,
               hipStream_t s);
void mic0ApplyF(const PressureSys &S, const int *hLevelStart, const float *r, float *z,
                hipStream_t s);

// ---- MIC(0) application in bands of hyperplanes (kernels_band.hip) ----
// L entries of the lower neighbour rows (lmx.. from lx..) and the band-sweep
// records (FP64 or the FP32 copies)
void mic0LowerEntries(const PressureSys &S, bool fp64, hipStream_t s);
// tiles of the band schedule and, per tile, the lowest and highest i+j+k of
// its rows (tileLo/tileHi hold mic0BandTiles ints; once per system)
int mic0BandTiles(Grid g, bool fp64);
// hyperplanes per launch and tile side of the band schedule
void mic0BandShape(bool fp64, int &levels, int &tile);
void mic0BandTileRanges(Grid g, const PressureSys &S, bool fp64, int *tileLo, int *tileHi, hipStream_t s);
// the tiles that hold rows of a launch's band, passed as a launch argument:
// workgroup b works on tile id[b]
struct TileList {
    static constexpr int MAX = 1900;  // keeps the launch arguments under 4 KB
    int n;  // 0: one workgroup per tile, each reading its tile's range from tileLo/tileHi
    uint16_t id[MAX];
};
// the launches of one application: the first nFwd run the forward sweep
struct BandPlan {
    bool fp64 = true;
    int nFwd = 0;
    std::vector<int> l0;         // first hyperplane of each launch's band
    std::vector<TileList> list;  // tiles with rows in the band
};
// builds the plan from the host copies of the level table and the tile ranges
// (hTileLo/hTileHi may be null: then every band that holds rows is launched
// and the workgroups check their tile's range on the device)
void mic0BandPlan(Grid g, const PressureSys &S, const int *hLevelStart, bool fp64, const int *hTileLo,
                  const int *hTileHi, BandPlan &plan);
// z = (L L^T)^-1 r with y as the forward result (M values, distinct from r and z)
void mic0ApplyBands(Grid g, const PressureSys &S, const BandPlan &plan, const int *tileLo, const int *tileHi,
                    const double *r, double *y, double *z, hipStream_t s);
void mic0ApplyBandsF(Grid g, const PressureSys &S, const BandPlan &plan, const int *tileLo, const int *tileHi,
                     const float *r, float *y, float *z, hipStream_t s);
void spmv(const PressureSys &S, const double *x, double *y, hipStream_t s);
void spmvF(const PressureSys &S, const float *x, float *y, hipStream_t s);
// y += (float)(sign * num/den) * x   (device scalars; den may be null -> scalar = num)
void axpyFloatStep(double *y, const double *x, size_t n, const double *num, const double *den,
                   double sign, hipStream_t s);
void axpyFloatStepF(float *y, const float *x, size_t n, const double *num, const double *den,
                    double sign, hipStream_t s);
// q = A s and *sq = s . q (the dot in the dot()/dotf() order)
void spmvDot(const PressureSys &S, const double *s, double *q, double *partial, double *sq, hipStream_t st);
void spmvDotF(const PressureSys &S, const float *s, float *q, double *partial, double *sq, hipStream_t st);
// a = (float)(num/den); x += a * s; r -= a * q; *rmax = |r|_inf (the absMax order);
// rf (optional) receives (float)r; rmaxHost (optional) a mapped host copy of *rmax
void pcgStep(double *x, const double *s, double *r, const double *q, size_t n, const double *num,
             const double *den, double *partial, double *rmax, hipStream_t st, float *rf = nullptr,
             double *rmaxHost = nullptr);
void pcgStepF(float *x, const float *s, float *r, const float *q, size_t n, const double *num,
              const double *den, double *partial, double *rmax, hipStream_t st, double *rmaxHost = nullptr);
void pcgStepMixed(double *x, const float *s, float *r, const float *q, size_t n, const double *num,
                  const double *den, double *partial, double *rmax, hipStream_t st, double *rmaxHost = nullptr);
// z = (double)zf + (double)(float)(num/den) * s   (FP32 zf, FP64 z and s)
void widenAxpyStep(double *z, const float *zf, const double *s, size_t n, const double *num, const double *den,
                   hipStream_t st);
// x += (double)(float)(sign * num/den) * (double)p   (FP64 x, FP32 p)
void axpyMixed(double *x, const float *p, size_t n, const double *num, const double *den, double sign,
               hipStream_t s);
// r = b - A x (FP64 true residual)
void residual(const PressureSys &S, const double *b, const double *x, double *r, hipStream_t s);
void pressureToGrid(Grid g, const PressureSys &S, const double *x, float *p, hipStream_t s);
void castToFloat(const double *a, float *b, size_t n, hipStream_t s);
void addFloatToDouble(double *x, const float *d, size_t n, hipStream_t s);
void mic0ToFloat(PressureSys &S, hipStream_t s);

// ---- adjoint of the pressure projection (kernels_adjoint.hip) ----
// pbar (N cells) = C^T u'bar: the pressure-gradient update's transpose
void projectionAdjointCells(Grid g, double dt, const float *phi, const float *wu, const float *wv, const float *ww,
                            const float *gu, const float *gv, const float *gw, double *pbar, hipStream_t s);
// ubar = K u'bar + B^T lambda (lambda on cells, 0 off the pressure rows)
void projectionAdjointFaces(Grid g, double dt, const float *phi, const float *wu, const float *wv, const float *ww,
                            const float *gu, const float *gv, const float *gw, const double *lambda, float *ou,
                            float *ov, float *ow, hipStream_t s);
// adjoint of P2G: vbar (per particle, SoA) from ubar on the faces and the
// forward's per-face weight sums; with the velocities vx/vy/vz and the
// forward's face values fu/fv/fw also xbar (P x 3), the positions' cotangent
// (vx and xbar may be null)
void p2gAdjoint(Grid g, int np, const float *px, const float *py, const float *pz, const float *vx, const float *vy,
                const float *vz, double radius, const float *gu, const float *gv, const float *gw, const float *wsu,
                const float *wsv, const float *wsw, const float *fu, const float *fv, const float *fw, float *vbx,
                float *vby, float *vbz, float *xbar, hipStream_t s);
// adjoint of trilinear sampling of component dir: ubar on that face grid from
// vbar (P x 3, interleaved).  key/slot/perm hold P ints, counts/offsets
// sampleBuckets(g, dir) + 1 ints, scanScratch scanScratchInts of that.
size_t sampleBuckets(Grid g, int dir);
void sampleAdjoint(Grid g, int dir, int np, const float *px, const float *py, const float *pz, const float *vbar,
                   int *key, int *slot, int *counts, int *offsets, int *perm, int *scanScratch, float *out,
                   hipStream_t s);
// rhs[r] = cells[rows[r]];  cells[rows[r]] = x[r] (other cells untouched)
void gatherRows(const PressureSys &S, const double *cells, double *rhs, hipStream_t s);
// x[r] = (double)cells[rows[r]] (a float cell grid onto the rows)
void gatherRowsF(const PressureSys &S, const float *cells, double *x, hipStream_t s);
void scatterRows(const PressureSys &S, const double *x, double *cells, hipStream_t s);

// ---- MIC(0) application as a line pipeline (kernels_pencil.hip) ----
// Pencils of 16 x 16 (j, k) lines along i, one workgroup each, one launch per
// triangular sweep; neighbour pencils hand results over through epoch-tagged
// exchange words.  The plan's device buffers are owned by the caller.
struct PencilPlan {
    int npj = 0, npk = 0, np = 0, nw = 0;  // pencils along j, k; total; 32-cell words per line
    int *order = nullptr;                  // np pencils by PJ+PK (device)
    LineWord *words = nullptr;             // nj*nk lines x nw words of the current system (device)
    unsigned long long *ex = nullptr;      // pencilExchangeWords(g) words, cleared once (device)
    unsigned *counter = nullptr;           // ticket counter, cleared once (device)
    int *error = nullptr;                  // raised by a wait past waitTicks (device)
    unsigned base = 0, epoch = 0;          // counter value and last epoch used
    long long waitTicks = 0;               // longest wait in wall-clock ticks
};
void pencilShape(Grid g, int &npj, int &npk, int &nw);
size_t pencilExchangeWords(Grid g);
size_t pencilLdsBytes(Grid g);  // dynamic LDS per workgroup
// pencils by PJ+PK, then PJ (the forward sweep's dependency order)
void pencilOrder(Grid g, std::vector<int> &order);
// line words of the current system (after pressureRows)
void pencilLineWords(Grid g, const PressureSys &S, PencilPlan &pp, hipStream_t s);
// z = (L L^T)^-1 r, y = the forward result (M values, distinct from r and z)
void pencilApply(Grid g, const PressureSys &S, PencilPlan &pp, const double *r, double *y, double *z,
                 hipStream_t s);
void pencilApplyF(Grid g, const PressureSys &S, PencilPlan &pp, const float *r, float *y, float *z,
                  hipStream_t s);

// ---- MIC(0) application, persistent tile wavefront (kernels_mic.hip) ----
// Tiles are 16 x 16 x depth cells; the schedule lists the non-empty tiles by
// tile level i/16 + j/16 + k/depth.  base and epoch advance with every launch.
struct MicSchedule {
    Grid g{};
    int depth = 0;               // cells per tile along k
    int ti = 0, tj = 0, tk = 0;  // tiles per axis
    int nTiles = 0, nActive = 0;
    int *order = nullptr;        // nActive tile ids (device)
    int *tileRows = nullptr;     // nTiles: non-zero when the tile holds rows (device)
    int *flags = nullptr;        // nTiles (device)
    int *counter = nullptr;      // ticket counter (device)
    int *error = nullptr;        // set when a wait exceeded spinLimit (device)
    int base = 0, epoch = 0;
    int grid = 0;                // persistent workgroups per launch
    long long waitTicks = 0;     // longest tile wait in wall-clock ticks (set by the owner)
};
int micTileDepth(bool fp64);
size_t micTileLds(bool fp64);  // bytes of LDS per workgroup
void micTileCounts(Grid g, int depth, int &ti, int &tj, int &tk);
void micMarkTiles(Grid g, const PressureSys &S, int depth, int *tileRows, hipStream_t s);
void micApplyTiles(const PressureSys &S, MicSchedule &ms, const double *r, double *z, hipStream_t s);
void micApplyTilesF(const PressureSys &S, MicSchedule &ms, const float *r, float *z, hipStream_t s);

}  // namespace ff
