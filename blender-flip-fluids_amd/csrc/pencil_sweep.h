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
// Staging.  The steps run in chunks of C.  A chunk's operands (the
// right-hand side and the MIC(0) record of each lane's C cells) are loaded
// one chunk ahead by a coalesced mapping (C consecutive threads take one
// line's C consecutive cells, whose rows are consecutive), written into LDS
// step-major, and read back by each lane at its step; the chunk's results go
// to LDS and out to memory by the same coalesced mapping when the chunk ends.
// A step therefore touches only LDS and registers (plus the fire-and-forget
// publication of a face value), and costs one LDS exchange and one barrier.
// The exchange words a chunk reads are loaded one chunk ahead too and checked
// when the chunk starts.
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

template <class T, class Rec, int TJ, int TK, int C>
struct PencilLds {
    static constexpr unsigned long long bytes(int nw) {
        return (unsigned long long)TJ * TK * nw * sizeof(LineWord) + (unsigned long long)(TJ + TK) * nw * 4 +
               (unsigned long long)C * TJ * TK * (sizeof(Rec) + 2 * sizeof(T));
    }
};

// sW: TJ*TK*nw line words; sN: (TJ+TK)*nw face-line row bits; sRec, sIn, sOut:
// C*TJ*TK step-major staging (entry q*TJ*TK + lane)
template <class T, class Rec, int TJ, int TK, int C, bool FWD>
FF_PS_DEVICE void pencilSweep(const PencilArgs<T, Rec> &a, const int tid, LineWord *sW, unsigned *sN, Rec *sRec,
                              T *sIn, T *sOut) {
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

    // the cell of lane `lane` at step t (t counts from lo forward, from hi
    // backward), and its row (-1: no row, or outside the grid)
    auto cellOfLane = [&](int lane, int t) {
        const int o = lane % TJ + lane / TJ;
        return (FWD ? lo + t : hi - t) - o;
    };
    auto rowOfLane = [&](int lane, int i) {
        const int jl = PJ * TJ + lane % TJ, kl = PK * TK + lane / TJ;
        if (jl >= nj || kl >= nk || i < 0 || i >= ni) return -1;
        const LineWord w = sW[lane * nw + (i >> 5)];
        const unsigned bit = 1u << (i & 31);
        return (w.bits & bit) ? w.base + __builtin_popcount(w.bits & (bit - 1u)) : -1;
    };
    auto faceBit = [&](const unsigned *bits, int i) {
        return i >= 0 && i < ni && ((bits[i >> 5] >> (i & 31)) & 1u) != 0u;
    };

    // operands of one chunk in the coalesced mapping: item e = tid + u*NT is
    // lane e / C's step e % C
    struct Stage {  // the record's four values kept apart, so they stay in registers
        int row[C];
        T in[C], inv[C], ra[C], rb[C], rc[C];
    };
    auto load = [&](Stage &st, int t0) {
#pragma unroll
        for (int u = 0; u < C; u++) {
            const int e = tid + u * NT, lane = e / C, q = e % C;
            const int row = rowOfLane(lane, cellOfLane(lane, t0 + q));
            st.row[u] = row;
            // row 0 stands in for a cell without a row (its operands are not used)
            const int rr = row >= 0 ? row : 0;
            st.in[u] = a.in[rr];
            const Rec rec = a.rec[rr];
            st.inv[u] = rec.inv;
            st.ra[u] = rec.a;
            st.rb[u] = rec.b;
            st.rc[u] = rec.c;
        }
    };
    auto stage = [&](const Stage &st) {
#pragma unroll
        for (int u = 0; u < C; u++) {
            const int e = tid + u * NT, lane = e / C, q = e % C;
            sIn[q * NT + lane] = st.in[u];
            Rec rec;
            rec.inv = st.inv[u];
            rec.a = st.ra[u];
            rec.b = st.rb[u];
            rec.c = st.rc[u];
            sRec[q * NT + lane] = rec;
        }
    };
    auto writeOut = [&](const int (&rows)[C]) {
#pragma unroll
        for (int u = 0; u < C; u++) {
            const int e = tid + u * NT, lane = e / C, q = e % C;
            if (rows[u] >= 0) a.out[rows[u]] = sOut[q * NT + lane];
        }
    };

    // exchange words of one chunk (this lane's face cells)
    struct Faces {
        unsigned long long ey[C][W], ez[C][W];
        unsigned need;  // bit q: the y face value of step q is read; bit C+q: the z face value
    };
    auto fetch = [&](Faces &fc, int t0) {
        fc.need = 0u;
#pragma unroll
        for (int q = 0; q < C; q++)
#pragma unroll
            for (int w = 0; w < W; w++) fc.ey[q][w] = fc.ez[q][w] = 0ull;
#pragma unroll
        for (int q = 0; q < C; q++) {
            const int i = cellOfLane(tid, t0 + q);
            if (rowOfLane(tid, i) >= 0) {
                if (conY && faceBit(bitsY, i)) {
                    fc.need |= 1u << q;
#pragma unroll
                    for (int w = 0; w < W; w++) fc.ey[q][w] = FF_PS_LOAD64(nbY + (unsigned long long)i * W + w);
                }
                if (conZ && faceBit(bitsZ, i)) {
                    fc.need |= 1u << (C + q);
#pragma unroll
                    for (int w = 0; w < W; w++) fc.ez[q][w] = FF_PS_LOAD64(nbZ + (unsigned long long)i * W + w);
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
    auto verify = [&](Faces &fc, int t0) {
#pragma unroll
        for (int q = 0; q < C; q++) {
            const int i = cellOfLane(tid, t0 + q);
            if (fc.need & (1u << q)) await(nbY + (unsigned long long)i * W, fc.ey[q]);
            if (fc.need & (1u << (C + q))) await(nbZ + (unsigned long long)i * W, fc.ez[q]);
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
    // the C steps of a chunk: LDS and registers only
    auto run = [&](const Faces &fc, int t0) {
#pragma unroll
        for (int q = 0; q < C; q++) {
            const int buf = q & 1, pb = buf ^ 1;  // t0 is even
            const int i = cellOfLane(tid, t0 + q);
            const bool has = rowOfLane(tid, i) >= 0;
            T v = T(0);
            if (has) {
                const Rec rc = sRec[q * NT + tid];
                T y = sIn[q * NT + tid];
                if (FWD) {
                    if (lk > 0) {
                        if (sh[pb][tid - TJ]) y -= rc.c * sv[pb][tid - TJ];
                    } else if (fc.need & (1u << (C + q))) {
                        y -= rc.c * PencilValue<T>::unpack(fc.ez[q]);
                    }
                    if (lj > 0) {
                        if (sh[pb][tid - 1]) y -= rc.b * sv[pb][tid - 1];
                    } else if (fc.need & (1u << q)) {
                        y -= rc.b * PencilValue<T>::unpack(fc.ey[q]);
                    }
                    if (prevHas) y -= rc.a * prev;
                } else {
                    if (prevHas) y -= rc.a * prev;
                    if (lj < TJ - 1) {
                        if (sh[pb][tid + 1]) y -= rc.b * sv[pb][tid + 1];
                    } else if (fc.need & (1u << q)) {
                        y -= rc.b * PencilValue<T>::unpack(fc.ey[q]);
                    }
                    if (lk < TK - 1) {
                        if (sh[pb][tid + TJ]) y -= rc.c * sv[pb][tid + TJ];
                    } else if (fc.need & (1u << (C + q))) {
                        y -= rc.c * PencilValue<T>::unpack(fc.ez[q]);
                    }
                }
                v = y * rc.inv;
                sOut[q * NT + tid] = v;
                if (pubY) publish(ownY + (unsigned long long)i * W, v);
                if (pubZ) publish(ownZ + (unsigned long long)i * W, v);
            }
            sv[buf][tid] = v;
            sh[buf][tid] = has ? 1 : 0;
            prev = v;
            prevHas = has;
            FF_PS_BARRIER();
        }
    };

    // ---- the chunks: a chunk's operands and exchange words are loaded while
    // the previous chunk runs, staged and checked when it has ended, then run;
    // its results go out when the next chunk starts ----
    Stage S;
    Faces FA, FB;  // the current chunk's exchange words and the next one's, in turn
    int outRows[C];
    auto chunk = [&](int c, Faces &cur, Faces &next) {
        const int t0 = c * C;
        // the previous chunk's last barrier has passed: every lane has read its
        // operands of that chunk and written its results into sOut
        if (c > 0) writeOut(outRows);
        stage(S);
#pragma unroll
        for (int u = 0; u < C; u++) outRows[u] = S.row[u];
        verify(cur, t0);
        FF_PS_BARRIER();
        if (c + 1 < nc) {
            load(S, t0 + C);
            fetch(next, t0 + C);
        }
        run(cur, t0);
    };
    load(S, 0);
    fetch(FA, 0);
    for (int c = 0; c < nc; c += 2) {
        chunk(c, FA, FB);
        if (c + 1 >= nc) break;
        chunk(c + 1, FB, FA);
    }
    writeOut(outRows);
}

}  // namespace ff
