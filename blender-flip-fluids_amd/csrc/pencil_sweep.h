// pencil_sweep.h -- one triangular sweep of the MIC(0) preconditioner
// application z = (L L^T)^-1 r (solveLower / solveLowerTransposeInPlace,
// pcgsolver.h:211-241) as ONE launch: the line pipeline.
//
// Work split.  The (j, k) lines along i are cut into pencils of TJ x TK lines.
// One workgroup runs one pencil, one lane per line.  At step s lane (lj, lk)
// handles cell i = s - lj - lk of its line, so a step is one i+j+k hyperplane
// of the pencil:
//   * the -x (+x) neighbour is the lane's own result of the previous step, in a
//     register;
//   * the -y/-z (+y/+z) neighbours inside the pencil are the neighbour lanes'
//     results of the previous step, exchanged through a double-buffered LDS
//     plane, one barrier per step;
//   * across the pencil's faces the neighbour pencil's lanes publish their
//     results into an exchange buffer as they go, and this pencil's boundary
//     lanes read them from there.
// The steps are limited to the range that holds the pencil's rows; a pencil
// without rows takes no part.
//
// Hand-off.  A published value is one 64-bit word {epoch, 32 value bits} (two
// words for a double), written and read with relaxed agent-scope atomics: the
// value and its validity travel in one single-copy-atomic access, so no fence
// orders them.  The epoch changes with every launch, so the buffer is never
// cleared.  A reader re-polls a word until it carries the launch's epoch.
//
// Progress.  Each workgroup first takes a ticket from a device counter and
// runs the pencil of that rank in dependency order (pencils by PJ+PK, reversed
// for the backward sweep).  A pencil reads only from pencils of a lower rank,
// whose workgroups took their tickets earlier and are running or done, so the
// launch drains for any number of resident workgroups.  Every wait is bounded
// in wall-clock time: past the limit the lane raises the error word (which
// every other waiting lane also gives up on) and the launch still finishes;
// the host turns the word into an error.
//
// Latency.  The operands of the next chunk of C steps (rows from the pencil's
// line words in LDS, then the right-hand side and the MIC(0) record of each
// row, and the exchange words) are loaded while the current chunk runs, so a
// step costs one LDS exchange and one barrier.
//
// Arithmetic.  Every row performs the reference's operations in the
// reference's order (forward: subtract the -z, -y, -x terms, then scale by
// invdiag; backward: +x, +y, +z, then scale), so the result equals the serial
// sweeps bit for bit.
//
// This header holds only the sweep's body.  Its includer supplies the
// hardware operations as macros (the HIP kernel file on the GPU, and a
// thread-per-lane CPU emulation used to check the indexing):
//   FF_PS_DEVICE        qualifier of the functions below
//   FF_PS_SHARED        storage class of workgroup-shared variables
//   FF_PS_BARRIER()     workgroup barrier (orders LDS)
//   FF_PS_LOAD64(p)     relaxed agent-scope atomic load of a uint64
//   FF_PS_STORE64(p, v) relaxed agent-scope atomic store of a uint64
//   FF_PS_LOAD32(p)     relaxed agent-scope atomic load of an int
//   FF_PS_RAISE(p)      atomic or of 1 into an int
//   FF_PS_TICKET(p)     atomic add of 1 to a uint32, returning the old value
//   FF_PS_CLOCK()       constant-rate clock as long long
//   FF_PS_PAUSE()       short sleep inside a poll
//   FF_PS_UNIFORM(x)    an int the whole workgroup holds (readfirstlane)
//   FF_PS_F2U / FF_PS_U2F / FF_PS_D2U / FF_PS_U2D   bit casts
#pragma once

#include "pencil_types.h"

namespace ff {

template <class T>
struct PencilValue;  // 64-bit exchange words per value
template <>
struct PencilValue<float> {
    static constexpr int W = 1;
    static FF_PS_DEVICE void pack(float v, unsigned (&b)[1]) { b[0] = FF_PS_F2U(v); }
    static FF_PS_DEVICE float unpack(const unsigned long long (&w)[1]) { return FF_PS_U2F((unsigned)w[0]); }
};
template <>
struct PencilValue<double> {
    static constexpr int W = 2;
    static FF_PS_DEVICE void pack(double v, unsigned (&b)[2]) {
        const unsigned long long u = FF_PS_D2U(v);
        b[0] = (unsigned)u;
        b[1] = (unsigned)(u >> 32);
    }
    static FF_PS_DEVICE double unpack(const unsigned long long (&w)[2]) {
        return FF_PS_U2D((unsigned long long)(unsigned)w[0] | ((unsigned long long)(unsigned)w[1] << 32));
    }
};

// Rec: {inv, a, b, c} -- invdiag and the L entries the sweep multiplies the
// -x/-y/-z (forward) or +x/+y/+z (backward) neighbours' results by
template <class T, class Rec>
struct PencilArgs {
    int ni, nj, nk;
    int nw;                 // 32-cell words per line, (ni + 31) / 32
    int npj, npk;           // pencils along j and k
    const LineWord *words;  // nj * nk lines (k-major) x nw
    const int *order;       // pencils by PJ + PK (forward ranks)
    const T *in;            // M: right-hand side (forward) or forward result (backward)
    T *out;                 // M
    const Rec *rec;         // M
    unsigned long long *ex; // exchange words: pencil x 2 faces x max(TJ,TK) lines x ni x W
    unsigned *counter;      // ticket counter
    int *error;             // raised when a wait exceeded waitTicks
    unsigned base;          // counter value when this launch started
    unsigned epoch;         // tag of this launch's published words (never 0)
    long long waitTicks;    // longest wait, FF_PS_CLOCK() ticks
};

// exchange words of one face line of pencil p: face 0 = the y face (line
// index lk), 1 = the z face (line index lj)
template <int TM, int W>
FF_PS_DEVICE unsigned long long *pencilFace(unsigned long long *ex, int p, int face, int line, int ni) {
    return ex + ((((unsigned long long)p * 2 + face) * TM + line) * ni) * W;
}

template <class T, class Rec, int TJ, int TK, int C, bool FWD>
FF_PS_DEVICE void pencilSweep(const PencilArgs<T, Rec> &a, const int tid, LineWord *sW, unsigned *sN) {
    static_assert(C % 2 == 0, "the LDS plane parity follows the step within a chunk");
    constexpr int NT = TJ * TK, TM = TJ > TK ? TJ : TK;
    constexpr int W = PencilValue<T>::W;
    FF_PS_SHARED T sv[2][NT];
    FF_PS_SHARED int sh[2][NT];
    FF_PS_SHARED int sLo[NT], sHi[NT];
    FF_PS_SHARED unsigned sTicket;

    const int ni = a.ni, nj = a.nj, nk = a.nk, nw = a.nw;
    const int lj = tid % TJ, lk = tid / TJ, off = lj + lk;
    if (tid == 0) sTicket = FF_PS_TICKET(a.counter) - a.base;
    sv[0][tid] = T(0);
    sv[1][tid] = T(0);
    sh[0][tid] = 0;
    sh[1][tid] = 0;
    FF_PS_BARRIER();
    const int np = a.npj * a.npk;
    const int ticket = FF_PS_UNIFORM((int)sTicket);
    if (ticket < 0 || ticket >= np) return;  // the grid has np workgroups: never taken
    const int p = FF_PS_UNIFORM(a.order[FWD ? ticket : np - 1 - ticket]);
    const int PJ = p % a.npj, PK = p / a.npj;
    const int j = PJ * TJ + lj, k = PK * TK + lk;
    const bool lineIn = j < nj && k < nk;

    // ---- the pencil's line words, and the row bits of the lines across its
    // y and z faces that this sweep reads (forward: -y/-z, backward: +y/+z) ----
    for (int q = tid; q < NT * nw; q += NT) {
        const int line = q / nw, w = q % nw;
        const int jj = PJ * TJ + line % TJ, kk = PK * TK + line / TJ;
        LineWord lw;
        lw.base = 0;
        lw.bits = 0u;
        if (jj < nj && kk < nk) lw = a.words[((unsigned long long)kk * nj + jj) * nw + w];
        sW[q] = lw;
    }
    const int yNbJ = FWD ? PJ * TJ - 1 : (PJ + 1) * TJ;
    const int zNbK = FWD ? PK * TK - 1 : (PK + 1) * TK;
    for (int q = tid; q < (TK + TJ) * nw; q += NT) {
        const int line = q / nw, w = q % nw;
        const int jj = line < TK ? yNbJ : PJ * TJ + (line - TK);
        const int kk = line < TK ? PK * TK + line : zNbK;
        const bool ok = jj >= 0 && jj < nj && kk >= 0 && kk < nk;
        sN[q] = ok ? a.words[((unsigned long long)kk * nj + jj) * nw + w].bits : 0u;
    }
    FF_PS_BARRIER();

    // ---- the steps that hold the pencil's rows ----
    int first = 0x7fffffff, last = -1;
    if (lineIn)
        for (int w = 0; w < nw; w++) {
            const unsigned b = sW[tid * nw + w].bits;
            if (b) {
                if (first == 0x7fffffff) first = 32 * w + __builtin_ctz(b);
                last = 32 * w + 31 - __builtin_clz(b);
            }
        }
    sLo[tid] = last >= 0 ? first + off : 0x7fffffff;
    sHi[tid] = last >= 0 ? last + off : -1;
    FF_PS_BARRIER();
    for (int o = NT / 2; o > 0; o >>= 1) {
        if (tid < o) {
            sLo[tid] = sLo[tid + o] < sLo[tid] ? sLo[tid + o] : sLo[tid];
            sHi[tid] = sHi[tid + o] > sHi[tid] ? sHi[tid + o] : sHi[tid];
        }
        FF_PS_BARRIER();
    }
    const int lo = FF_PS_UNIFORM(sLo[0]), hi = FF_PS_UNIFORM(sHi[0]);
    if (hi < lo) return;  // no rows: nothing is published and nobody reads this pencil
    const int nc = (hi - lo + 1 + C - 1) / C;

    // ---- roles across the faces ----
    const bool pubY = FWD ? (lj == TJ - 1 && PJ + 1 < a.npj) : (lj == 0 && PJ > 0);
    const bool pubZ = FWD ? (lk == TK - 1 && PK + 1 < a.npk) : (lk == 0 && PK > 0);
    const bool conY = FWD ? lj == 0 : lj == TJ - 1;  // reads the y face of the neighbour pencil
    const bool conZ = FWD ? lk == 0 : lk == TK - 1;
    // a face line's row bits are non-zero only where the neighbour pencil exists
    const int pY = FWD ? (PJ > 0 ? p - 1 : p) : (PJ + 1 < a.npj ? p + 1 : p);
    const int pZ = FWD ? (PK > 0 ? p - a.npj : p) : (PK + 1 < a.npk ? p + a.npj : p);
    unsigned long long *const ownY = pencilFace<TM, W>(a.ex, p, 0, lk, ni);
    unsigned long long *const ownZ = pencilFace<TM, W>(a.ex, p, 1, lj, ni);
    const unsigned long long *const nbY = pencilFace<TM, W>(a.ex, pY, 0, lk, ni);
    const unsigned long long *const nbZ = pencilFace<TM, W>(a.ex, pZ, 1, lj, ni);
    const unsigned *const bitsY = sN + lk * nw;         // the y face line next to this lane's line
    const unsigned *const bitsZ = sN + (TK + lj) * nw;  // the z face line

    struct Chunk {
        int row[C];
        T in[C];
        Rec rec[C];
        unsigned long long ey[C][W], ez[C][W];
        unsigned need;  // bit q: the y face value of step q is read; bit C+q: the z face value
    };
    auto cellOf = [&](int t) { return (FWD ? lo + t : hi - t) - off; };
    auto rowAt = [&](int i) {
        if (!lineIn || i < 0 || i >= ni) return -1;
        const LineWord w = sW[tid * nw + (i >> 5)];
        const unsigned bit = 1u << (i & 31);
        return (w.bits & bit) ? w.base + __builtin_popcount(w.bits & (bit - 1u)) : -1;
    };
    auto faceBit = [&](const unsigned *bits, int i) {
        return i >= 0 && i < ni && ((bits[i >> 5] >> (i & 31)) & 1u) != 0u;
    };
    // loads of a chunk's operands and exchange words (used one chunk later)
    auto issue = [&](Chunk &ch, int t0) {
        ch.need = 0u;
#pragma unroll
        for (int q = 0; q < C; q++) {
            const int i = cellOf(t0 + q);
            const int row = rowAt(i);
            ch.row[q] = row;
            // row 0 stands in for a cell without a row (its operands are not used)
            const int rr = row >= 0 ? row : 0;
            ch.in[q] = a.in[rr];
            ch.rec[q] = a.rec[rr];
            if (row >= 0) {
                if (conY && faceBit(bitsY, i)) {
                    ch.need |= 1u << q;
#pragma unroll
                    for (int w = 0; w < W; w++) ch.ey[q][w] = FF_PS_LOAD64(nbY + (unsigned long long)i * W + w);
                }
                if (conZ && faceBit(bitsZ, i)) {
                    ch.need |= 1u << (C + q);
#pragma unroll
                    for (int w = 0; w < W; w++) ch.ez[q][w] = FF_PS_LOAD64(nbZ + (unsigned long long)i * W + w);
                }
            }
        }
    };
    // re-polls the exchange words that do not yet carry this launch's epoch
    auto await = [&](const unsigned long long *src, unsigned long long (&w)[W]) {
        auto fresh = [&]() {
            bool ok = true;
#pragma unroll
            for (int q = 0; q < W; q++) ok = ok && (unsigned)(w[q] >> 32) == a.epoch;
            return ok;
        };
        if (fresh()) return;
        const long long start = FF_PS_CLOCK();
        for (;;) {
            FF_PS_PAUSE();
#pragma unroll
            for (int q = 0; q < W; q++) w[q] = FF_PS_LOAD64(src + q);
            if (fresh()) return;
            if (FF_PS_LOAD32(a.error) != 0 || FF_PS_CLOCK() - start > a.waitTicks) {
                FF_PS_RAISE(a.error);
                return;
            }
        }
    };
    auto verify = [&](Chunk &ch, int t0) {
#pragma unroll
        for (int q = 0; q < C; q++) {
            const int i = cellOf(t0 + q);
            if (ch.need & (1u << q)) await(nbY + (unsigned long long)i * W, ch.ey[q]);
            if (ch.need & (1u << (C + q))) await(nbZ + (unsigned long long)i * W, ch.ez[q]);
        }
    };
    T prev = T(0);  // this line's result of the previous step (the x neighbour)
    bool prevHas = false;
    auto publish = [&](unsigned long long *dst, T v) {
        unsigned b[W];
        PencilValue<T>::pack(v, b);
#pragma unroll
        for (int q = 0; q < W; q++) FF_PS_STORE64(dst + q, ((unsigned long long)a.epoch << 32) | b[q]);
    };
    auto run = [&](const Chunk &ch, int t0) {
#pragma unroll
        for (int q = 0; q < C; q++) {
            const int buf = q & 1, pb = buf ^ 1;  // t0 is even
            const int row = ch.row[q];
            T v = T(0);
            if (row >= 0) {
                const Rec rc = ch.rec[q];
                T y = ch.in[q];
                if (FWD) {
                    if (lk > 0) {
                        if (sh[pb][tid - TJ]) y -= rc.c * sv[pb][tid - TJ];
                    } else if (ch.need & (1u << (C + q))) {
                        y -= rc.c * PencilValue<T>::unpack(ch.ez[q]);
                    }
                    if (lj > 0) {
                        if (sh[pb][tid - 1]) y -= rc.b * sv[pb][tid - 1];
                    } else if (ch.need & (1u << q)) {
                        y -= rc.b * PencilValue<T>::unpack(ch.ey[q]);
                    }
                    if (prevHas) y -= rc.a * prev;
                } else {
                    if (prevHas) y -= rc.a * prev;
                    if (lj < TJ - 1) {
                        if (sh[pb][tid + 1]) y -= rc.b * sv[pb][tid + 1];
                    } else if (ch.need & (1u << q)) {
                        y -= rc.b * PencilValue<T>::unpack(ch.ey[q]);
                    }
                    if (lk < TK - 1) {
                        if (sh[pb][tid + TJ]) y -= rc.c * sv[pb][tid + TJ];
                    } else if (ch.need & (1u << (C + q))) {
                        y -= rc.c * PencilValue<T>::unpack(ch.ez[q]);
                    }
                }
                v = y * rc.inv;
                a.out[row] = v;
                const int i = cellOf(t0 + q);
                if (pubY) publish(ownY + (unsigned long long)i * W, v);
                if (pubZ) publish(ownZ + (unsigned long long)i * W, v);
            }
            sv[buf][tid] = v;
            sh[buf][tid] = row >= 0 ? 1 : 0;
            prev = v;
            prevHas = row >= 0;
            FF_PS_BARRIER();
        }
    };

    // ---- the chunks, two register sets in turn: a chunk's loads are issued
    // before the previous chunk runs and checked after it ----
    Chunk A, B;
    issue(A, 0);
    for (int c = 0; c < nc; c += 2) {
        verify(A, c * C);
        if (c + 1 < nc) issue(B, (c + 1) * C);
        run(A, c * C);
        if (c + 1 >= nc) break;
        verify(B, (c + 1) * C);
        if (c + 2 < nc) issue(A, (c + 2) * C);
        run(B, (c + 1) * C);
    }
}

}  // namespace ff
