// kernels_scan.hip -- prefix sums and deterministic reductions.
//
// The scan orders pressure rows and compacts particles; the reductions are the
// FP64 dot products and inf-norms of the pressure solve.  Reductions run a
// fixed two-stage tree (fixed grid, fixed per-thread stride, fixed LDS tree),
// so a given vector always reduces to the same bits.
#include "kernels.h"

#include <cmath>
#include <stdexcept>

namespace ff {

namespace {
constexpr int SCAN_THREADS = 256;
constexpr int SCAN_ITEMS = 4;
constexpr int SCAN_TILE = SCAN_THREADS * SCAN_ITEMS;
constexpr int RED_THREADS = 256;
constexpr int RED_MAX_BLOCKS = 1024;

__global__ __launch_bounds__(SCAN_THREADS) void scanTile(const int *in, int *out, size_t n,
                                                         int *blockSums) {
    __shared__ int s[SCAN_THREADS];
    size_t base = (size_t)blockIdx.x * SCAN_TILE + (size_t)threadIdx.x * SCAN_ITEMS;
    int v[SCAN_ITEMS];
#pragma unroll
    for (int q = 0; q < SCAN_ITEMS; q++) v[q] = (base + q < n) ? in[base + q] : 0;
    int run = 0;
#pragma unroll
    for (int q = 0; q < SCAN_ITEMS; q++) {
        int t = v[q];
        v[q] = run;
        run += t;
    }
    s[threadIdx.x] = run;
    __syncthreads();
    for (int off = 1; off < SCAN_THREADS; off <<= 1) {
        int add = threadIdx.x >= (unsigned)off ? s[threadIdx.x - off] : 0;
        __syncthreads();
        s[threadIdx.x] += add;
        __syncthreads();
    }
    int excl = s[threadIdx.x] - run;
#pragma unroll
    for (int q = 0; q < SCAN_ITEMS; q++)
        if (base + q < n) out[base + q] = v[q] + excl;
    if (threadIdx.x == SCAN_THREADS - 1) blockSums[blockIdx.x] = s[threadIdx.x];
}

__global__ __launch_bounds__(SCAN_THREADS) void scanAdd(int *out, size_t n, const int *blockOffsets) {
    size_t base = (size_t)blockIdx.x * SCAN_TILE + (size_t)threadIdx.x * SCAN_ITEMS;
    int add = blockOffsets[blockIdx.x];
#pragma unroll
    for (int q = 0; q < SCAN_ITEMS; q++)
        if (base + q < n) out[base + q] += add;
}

__global__ void writeTotal(int *out, size_t n, const int *blockSums, const int *blockOffsets,
                           int nb) {
    out[n] = blockOffsets[nb - 1] + blockSums[nb - 1];
}

size_t tiles(size_t n) { return (n + SCAN_TILE - 1) / SCAN_TILE; }

}  // namespace

size_t scanScratchInts(size_t n) {
    size_t total = 0;
    while (true) {
        size_t nb = tiles(n < 1 ? 1 : n);
        total += 2 * nb + 2;
        if (nb <= 1) break;
        n = nb;
    }
    return total;
}

// out[0..n) = exclusive scan of in, out[n] = total.  Recursive over tile sums.
void exclusiveScan(const int *in, int *out, size_t n, int *scratch, hipStream_t s) {
    if (n == 0) {
        hipCheck(hipMemsetAsync(out, 0, sizeof(int), s), "hipMemsetAsync");
        return;
    }
    size_t nb = tiles(n);
    int *sums = scratch;
    int *offs = scratch + nb + 1;
    hipLaunchKernelGGL(scanTile, dim3((unsigned)nb), dim3(SCAN_THREADS), 0, s, in, out, n, sums);
    if (nb > 1) {
        exclusiveScan(sums, offs, nb, scratch + 2 * nb + 2, s);
        hipLaunchKernelGGL(scanAdd, dim3((unsigned)nb), dim3(SCAN_THREADS), 0, s, out, n, offs);
        hipLaunchKernelGGL(writeTotal, dim3(1), dim3(1), 0, s, out, n, sums, offs, (int)nb);
    } else {
        hipCheck(hipMemsetAsync(offs, 0, sizeof(int), s), "hipMemsetAsync");
        hipLaunchKernelGGL(writeTotal, dim3(1), dim3(1), 0, s, out, n, sums, offs, 1);
    }
}

int reduceBlocks(size_t n) {
    size_t b = (n + RED_THREADS * 8 - 1) / (RED_THREADS * 8);
    if (b < 1) b = 1;
    if (b > RED_MAX_BLOCKS) b = RED_MAX_BLOCKS;
    if (b >= XCDS) b -= b % XCDS;  // whole runs of blocks per die
    return (int)b;
}

namespace {

template <class T, int OP, class TB = T>  // OP 0: sum of a*b, 1: max |a|
__global__ __launch_bounds__(RED_THREADS) void reduceStage1(const T *a, const TB *b, size_t n,
                                                            double *partial) {
    const unsigned L = reduceLogicalBlock();
    const size_t span = reduceSpan(n, gridDim.x);
    const size_t lo = (size_t)L * span, hi = lo + span < n ? lo + span : n;
    double acc = 0.0;
    for (size_t i = lo + threadIdx.x; i < hi; i += RED_THREADS) {
        if (OP == 0)
            acc += (double)a[i] * (double)b[i];
        else
            acc = fmax(acc, fabs((double)a[i]));
    }
    acc = reduceFold<OP>(acc);
    if (threadIdx.x == 0) partial[L] = acc;
}

// outHost (optional): a mapped pinned host copy of the result, which the
// host reads after a stream synchronize instead of copying it back
template <int OP>
__global__ __launch_bounds__(RED_THREADS) void reduceStage2(const double *partial, int nb,
                                                            double *out, double *outHost) {
    double acc = 0.0;
    for (int i = threadIdx.x; i < nb; i += RED_THREADS) {
        if (OP == 0)
            acc += partial[i];
        else
            acc = fmax(acc, partial[i]);
    }
    acc = reduceFold<OP>(acc);
    if (threadIdx.x == 0) {
        *out = acc;
        if (outHost) *outHost = acc;
    }
}

template <class T, int OP, class TB = T>
void reduce(const T *a, const TB *b, size_t n, double *partial, double *out, hipStream_t s) {
    int nb = reduceBlocks(n);
    hipLaunchKernelGGL((reduceStage1<T, OP, TB>), dim3(nb), dim3(RED_THREADS), 0, s, a, b, n, partial);
    hipLaunchKernelGGL((reduceStage2<OP>), dim3(1), dim3(RED_THREADS), 0, s, partial, nb, out, nullptr);
}

}  // namespace

void reduceFinish(const double *partial, int nb, double *out, bool isMax, hipStream_t s, double *outHost) {
    if (isMax)
        hipLaunchKernelGGL((reduceStage2<1>), dim3(1), dim3(RED_THREADS), 0, s, partial, nb, out, outHost);
    else
        hipLaunchKernelGGL((reduceStage2<0>), dim3(1), dim3(RED_THREADS), 0, s, partial, nb, out, outHost);
}

void dot(const double *a, const double *b, size_t n, double *partial, double *out, hipStream_t s) {
    reduce<double, 0>(a, b, n, partial, out, s);
}
void dotf(const float *a, const float *b, size_t n, double *partial, double *out, hipStream_t s) {
    reduce<float, 0>(a, b, n, partial, out, s);
}
void dotfd(const float *a, const double *b, size_t n, double *partial, double *out, hipStream_t s) {
    reduce<float, 0, double>(a, b, n, partial, out, s);
}
void absMax(const double *a, size_t n, double *partial, double *out, hipStream_t s) {
    reduce<double, 1>(a, a, n, partial, out, s);
}
void absMaxf(const float *a, size_t n, double *partial, double *out, hipStream_t s) {
    reduce<float, 1>(a, a, n, partial, out, s);
}

// ---- dot products in the reference's summation order ----
namespace {

// v of lane l (l the same in every lane), through two scalar reads
__device__ __forceinline__ double laneValue(double v, int l) {
    const unsigned long long u = (unsigned long long)__double_as_longlong(v);
    const unsigned lo = __builtin_amdgcn_readlane((unsigned)u, l);
    const unsigned hi = __builtin_amdgcn_readlane((unsigned)(u >> 32), l);
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// interval c of [0, n) split into `chunks` (ThreadUtils::splitRangeIntoIntervals,
// threadutils.cpp:54-72), summed from zero in index order by one wave: the
// lanes form 64 products at a time, lane 0's running sum takes them in order
__global__ __launch_bounds__(64) void dotChunkKernel(const double *a, const double *b, int n, int chunks,
                                                     double *partial) {
    const int c = blockIdx.x, lane = threadIdx.x;
    const int size = (int)floor((double)n / (double)chunks);
    const int rem = n - size * chunks;
    const int lo = c * size + (c < rem ? c : rem), hi = lo + size + (c < rem ? 1 : 0);
    double s = 0.0;
    for (int base = lo; base < hi; base += 64) {
        const int i = base + lane;
        const double p = i < hi ? a[i] * b[i] : 0.0;
        const int cnt = hi - base < 64 ? hi - base : 64;
        for (int l = 0; l < cnt; l++) s += laneValue(p, l);
    }
    if (lane == 0) partial[c] = s;
}

__global__ void dotChunkSumKernel(const double *partial, int chunks, double *out) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    double s = 0.0;
    for (int c = 0; c < chunks; c++) s += partial[c];
    *out = s;
}

}  // namespace

int referenceDotChunks(size_t n, int threads) {
    // BLAS::dot: numthreads = min(numCPU, ceil((float)size / 500000)) (blaswrapper.h:39-40)
    const int c = (int)fmin((double)threads, std::ceil((float)n / (float)500000));
    return c < 1 ? 1 : c;
}

void dotReference(const double *a, const double *b, size_t n, int threads, double *partial, double *out,
                  hipStream_t s) {
    const int chunks = referenceDotChunks(n, threads);
    if (chunks > RED_MAX_BLOCKS) throw std::runtime_error("reference-order dot: too many intervals");
    hipLaunchKernelGGL(dotChunkKernel, dim3(chunks), dim3(64), 0, s, a, b, (int)n, chunks, partial);
    hipLaunchKernelGGL(dotChunkSumKernel, dim3(1), dim3(1), 0, s, partial, chunks, out);
}

}  // namespace ff
