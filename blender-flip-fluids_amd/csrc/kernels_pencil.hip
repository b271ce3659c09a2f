// kernels_pencil.hip -- the MIC(0) preconditioner application z = (L L^T)^-1 r
// (solveLower + solveLowerTransposeInPlace, pcgsolver.h:211-241) as two
// launches, one per triangular sweep: the line pipeline of pencil_sweep.h.
#include <stdexcept>
#include <vector>

#include "kernels.h"

#define FF_PS_DEVICE __device__ __forceinline__
#define FF_PS_SHARED __shared__
#define FF_PS_BARRIER() __syncthreads()
#define FF_PS_LOAD64(p) \
    __hip_atomic_load((unsigned long long *)(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)
#define FF_PS_STORE64(p, v) __hip_atomic_store((p), (v), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)
#define FF_PS_LOAD32(p) __hip_atomic_load((p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)
#define FF_PS_RAISE(p) __hip_atomic_fetch_or((p), 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)
#define FF_PS_TICKET(p) atomicAdd((p), 1u)
#define FF_PS_CLOCK() ((long long)wall_clock64())
#define FF_PS_PAUSE() __builtin_amdgcn_s_sleep(2)
#define FF_PS_UNIFORM(x) __builtin_amdgcn_readfirstlane(x)
#define FF_PS_F2U(v) __float_as_uint(v)
#define FF_PS_U2F(u) __uint_as_float(u)
#define FF_PS_D2U(v) ((unsigned long long)__double_as_longlong(v))
#define FF_PS_U2D(u) __longlong_as_double((long long)(u))
#include "pencil_sweep.h"

namespace ff {

namespace {

// pencil shapes (lines per side) and chunk lengths: FP32 and FP64
constexpr int PTJ = 16, PTK = 16, PC32 = 8, PC64 = 4;

// dynamic LDS: the staged records (16-byte aligned), inputs and outputs, then
// the line words and the face lines' row bits
template <class T, int TJ, int TK, int C, bool FWD>
__global__ __launch_bounds__(TJ *TK) void pencilSweepKernel(PencilArgs<T, MicRec<T>> a) {
    extern __shared__ unsigned long long pencilShared[];
    constexpr int NT = TJ * TK;
    MicRec<T> *sRec = (MicRec<T> *)pencilShared;
    T *sIn = (T *)(sRec + C * NT);
    T *sOut = sIn + C * NT;
    LineWord *sW = (LineWord *)(sOut + C * NT);
    unsigned *sN = (unsigned *)(sW + NT * a.nw);
    pencilSweep<T, MicRec<T>, TJ, TK, C, FWD>(a, (int)threadIdx.x, sW, sN, sRec, sIn, sOut);
}

// one thread per 32-cell word of a line: its row bits and the row of its
// first liquid cell (the exclusive scan of the row mask)
__global__ void lineWordsKernel(Grid g, int nw, const int *cellRow, const int *rowOf, LineWord *words) {
    const size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t lines = (size_t)g.nj * g.nk;
    if (q >= lines * nw) return;
    const size_t line = q / nw;
    const int w = (int)(q % nw);
    const size_t c0 = line * g.ni + 32 * (size_t)w;
    const int cnt = g.ni - 32 * w < 32 ? g.ni - 32 * w : 32;
    unsigned bits = 0u;
    for (int b = 0; b < cnt; b++) bits |= (cellRow[c0 + b] >= 0 ? 1u : 0u) << b;
    words[q] = LineWord{rowOf[c0], bits};
}

template <class T, int C>
size_t pencilLds(int nw) {
    return (size_t)PencilLds<T, MicRec<T>, PTJ, PTK, C>::bytes(nw);
}

template <class T, int C>
void pencilSweeps(Grid g, const PressureSys &S, PencilPlan &pp, const T *r, T *y, T *z, const MicRec<T> *recFwd,
                  const MicRec<T> *recBwd, hipStream_t st) {
    PencilArgs<T, MicRec<T>> a;
    a.ni = g.ni;
    a.nj = g.nj;
    a.nk = g.nk;
    a.nw = pp.nw;
    a.npj = pp.npj;
    a.npk = pp.npk;
    a.words = pp.words;
    a.order = pp.order;
    a.ex = pp.ex;
    a.counter = pp.counter;
    a.error = pp.error;
    a.waitTicks = pp.waitTicks;
    const dim3 grid(pp.np), block(PTJ * PTK);
    const size_t lds = pencilLds<T, C>(pp.nw);
    for (int sweep = 0; sweep < 2; sweep++) {
        a.base = pp.base;
        if (++pp.epoch == 0) pp.epoch = 1;  // 0 tags the cleared buffer
        a.epoch = pp.epoch;
        pp.base += (unsigned)pp.np;
        if (sweep == 0) {
            a.in = r;
            a.out = y;
            a.rec = recFwd;
            hipLaunchKernelGGL((pencilSweepKernel<T, PTJ, PTK, C, true>), grid, block, lds, st, a);
        } else {
            a.in = y;
            a.out = z;
            a.rec = recBwd;
            hipLaunchKernelGGL((pencilSweepKernel<T, PTJ, PTK, C, false>), grid, block, lds, st, a);
        }
    }
    hipCheck(hipGetLastError(), "MIC(0) line-pipeline launch");
}

}  // namespace

void pencilShape(Grid g, int &npj, int &npk, int &nw) {
    npj = (g.nj + PTJ - 1) / PTJ;
    npk = (g.nk + PTK - 1) / PTK;
    nw = (g.ni + 31) / 32;
}

size_t pencilExchangeWords(Grid g) {
    int npj, npk, nw;
    pencilShape(g, npj, npk, nw);
    const int tm = PTJ > PTK ? PTJ : PTK;
    return (size_t)npj * npk * 2 * tm * g.ni * 2;  // two words per value covers FP64
}

size_t pencilLdsBytes(Grid g) {
    int npj, npk, nw;
    pencilShape(g, npj, npk, nw);
    const size_t a = pencilLds<double, PC64>(nw), b = pencilLds<float, PC32>(nw);
    return a > b ? a : b;
}

void pencilOrder(Grid g, std::vector<int> &order) {
    int npj, npk, nw;
    pencilShape(g, npj, npk, nw);
    order.clear();
    for (int d = 0; d <= npj + npk - 2; d++)
        for (int pj = 0; pj < npj; pj++) {
            const int pk = d - pj;
            if (pk >= 0 && pk < npk) order.push_back(pj + pk * npj);
        }
}

void pencilLineWords(Grid g, const PressureSys &S, PencilPlan &pp, hipStream_t s) {
    const size_t n = (size_t)g.nj * g.nk * pp.nw;
    hipLaunchKernelGGL(lineWordsKernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, g, pp.nw, S.cellRow,
                       S.rowOf, pp.words);
}

void pencilApply(Grid g, const PressureSys &S, PencilPlan &pp, const double *r, double *y, double *z,
                 hipStream_t s) {
    if (S.m == 0) return;
    pencilSweeps<double, PC64>(g, S, pp, r, y, z, S.recFwd, S.recBwd, s);
}

void pencilApplyF(Grid g, const PressureSys &S, PencilPlan &pp, const float *r, float *y, float *z,
                  hipStream_t s) {
    if (S.m == 0) return;
    pencilSweeps<float, PC32>(g, S, pp, r, y, z, S.frecFwd, S.frecBwd, s);
}

}  // namespace ff
