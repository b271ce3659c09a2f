// rollout.hip -- forward and adjoint of a run of FLIP substeps on one stream.
#include "rollout.h"

#include <algorithm>
#include <cmath>

namespace ff {

namespace {

constexpr int TPB = 256;
unsigned blocksFor(size_t n, unsigned t) { return (unsigned)((n + t - 1) / t); }

__global__ void addKernel(float *a, const float *b, size_t n) {
    size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q < n) a[q] += b[q];
}

// pbar = pbarNew + xbarP2G;  vbar = vbarNew + (vbx, vby, vbz)  (P x 3 each)
__global__ void combineKernel(int np, const float *pbarNew, const float *xbarP2G, const float *vbarNew,
                              const float *vb, float *pbar, float *vbar) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= np) return;
    for (int e = 0; e < 3; e++) {
        const size_t o = 3 * (size_t)q + e;
        pbar[o] = pbarNew[o] + xbarP2G[o];
        vbar[o] = vbarNew[o] + vb[(size_t)e * np + q];
    }
}

// P2G normalises a face by its weight sum W when W > 1e-6 and keeps the
// plain sum S otherwise (velocityadvector.cpp:467-536), a jump when W
// crosses the threshold.  With the recorded flags held fixed, the decision
// follows the record instead: S/W on recorded faces (0 when no weight is
// left), S elsewhere.  Where the decisions agree nothing changes.
__global__ void freezeNormKernel(float *field, const float *wsum, const uint8_t *rec, size_t n) {
    size_t f = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= n) return;
    const float W = wsum[f];
    const bool want = rec[f] != 0, got = W > 1e-6f;
    if (want && !got)
        field[f] = W > 0.0f ? field[f] / W : 0.0f;
    else if (!want && got)
        field[f] = field[f] * W;
}

__global__ void countKernel(const uint8_t *a, int n, int *count) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    // one integer atomic per flagged particle: a count, summed in any order
    if (q < n && a[q]) atomicAdd(count, 1);
}

void copyD(void *dst, const void *src, size_t bytes, hipStream_t s) {
    if (bytes) hipCheck(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, s), "device copy");
}

}  // namespace

Rollout::Rollout(Grid g, int np, const float *pos, const float *solidNodes, const RolloutSettings &st)
    : _g(g), _st(st), _np(np) {
    if (np <= 0) throw std::domain_error("the rollout needs particles");
    if (st.substeps < 1 || st.checkpointEvery < 1) throw std::domain_error("substeps and checkpoint spacing must be >= 1");
    if (!(st.dt > 0.0)) throw std::domain_error("the substep length must be > 0");
    if (st.picflip < 0.0 || st.picflip > 1.0) throw std::domain_error("the PIC share must lie in [0, 1]");
    // the extrapolation's status bytes hold layer numbers below 254
    if (!(st.cfl > 0.0) || std::ceil(std::sqrt(3.0) * st.cfl) + 3 > 250)
        throw std::domain_error("the CFL number must lie in (0, 142]");
    if (!(st.tol > 0.0) || st.maxIter < 1) throw std::domain_error("the solver tolerance and iteration cap must be > 0");
    for (size_t q = 0; q < (size_t)np; q++) {
        const double x = pos[3 * q], y = pos[3 * q + 1], z = pos[3 * q + 2];
        if (!(x >= 0.0 && y >= 0.0 && z >= 0.0 && x < g.ni * g.dx && y < g.nj * g.dx && z < g.nk * g.dx))
            throw std::domain_error("rollout particles must lie inside the grid");
    }
    _N = (size_t)g.ni * g.nj * g.nk;
    _nf[0] = (size_t)(g.ni + 1) * g.nj * g.nk;
    _nf[1] = (size_t)g.ni * (g.nj + 1) * g.nk;
    _nf[2] = (size_t)g.ni * g.nj * (g.nk + 1);
    const size_t F = _nf[0] + _nf[1] + _nf[2], P = (size_t)np, NN = (size_t)(g.ni + 1) * (g.nj + 1) * (g.nk + 1);
    _layers = (int)std::ceil(std::sqrt(3.0) * st.cfl) + 3;   // fluidsimulation.cpp:6282-6286
    _radius = 0.5 * 1.0 * g.dx * std::sqrt(3.0);             // fluidsimulation.cpp:4351
    hipCheck(hipStreamCreateWithFlags(&_s, hipStreamNonBlocking), "hipStreamCreate");

    _pos0.alloc(3 * P);
    _state.alloc(6 * P);
    _final.alloc(6 * P);
    _ckpt.alloc((size_t)(st.substeps / st.checkpointEvery + 1) * 6 * P);
    _seg.alloc((size_t)st.checkpointEvery * 6 * P);
    _phiRec.alloc((size_t)st.substeps * _N);
    _validRec.alloc((size_t)st.substeps * F);
    _keys.alloc(P); _slot.alloc(P); _perm.alloc(P);
    _counts.alloc(_N + 1); _offsets.alloc(_N + 1);
    _scan.alloc(scanScratchInts(std::max(_N + 1, P + 1)));
    for (auto &b : _sorted) b.alloc(P);
    _cs = CellSort{_keys, _slot, _counts, _offsets, _perm, _sorted[0], _sorted[1], _sorted[2],
                   _sorted[3], _sorted[4], _sorted[5]};
    _phi.alloc(_N); _pressure.alloc(_N); _rho.alloc(_N); _rhoBar.alloc(_N); _solid.alloc(NN);
    _moved.alloc(P); _movedCount.alloc(1);
    for (int d = 0; d < 3; d++) {
        _init[d].alloc(_nf[d]); _wt[d].alloc(_nf[d]);
        _u[d].alloc(_nf[d]); _su[d].alloc(_nf[d]); _cu[d].alloc(_nf[d]); _ws[d].alloc(_nf[d]);
        _valid[d].alloc(_nf[d]); _st0[d].alloc(_nf[d]); _st1[d].alloc(_nf[d]);
        _ub[d].alloc(_nf[d]); _sb[d].alloc(_nf[d]); _tmp[d].alloc(_nf[d]); _zero[d].alloc(_nf[d]);
        hipCheck(hipMemsetAsync(_zero[d], 0, sizeof(float) * _nf[d], _s), "memset");
    }
    _ci = (int)std::ceil(g.ni * g.dx / (3 * g.dx));
    _cj = (int)std::ceil(g.nj * g.dx / (3 * g.dx));
    _ck = (int)std::ceil(g.nk * g.dx / (3 * g.dx));
    _near.alloc((size_t)_ci * _cj * _ck);
    _nearTmp.alloc((size_t)_ci * _cj * _ck);
    _pbar.alloc(3 * P); _vbar.alloc(3 * P); _pbarNew.alloc(3 * P); _vbarNew.alloc(3 * P);
    _xbarP2G.alloc(3 * P); _vb3.alloc(3 * P); _cotSav.alloc(3 * P);
    _pts.alloc(9 * P); _cot.alloc(9 * P);
    size_t nbMax = 0;
    for (int d = 0; d < 3; d++) nbMax = std::max(nbMax, sampleBuckets(g, d));
    _akey.alloc(3 * P); _aslot.alloc(3 * P); _aperm.alloc(3 * P);
    _acounts.alloc(nbMax + 1); _aoffsets.alloc(nbMax + 1);
    _ascan.alloc(scanScratchInts(nbMax + 1));
    _cellBar.alloc(_N); _lambda.alloc(_N);
    _ps.allocate(g);

    // seeded positions (SoA), solid SDF, face weights and the near-solid grid
    std::vector<float> col(P);
    for (int e = 0; e < 3; e++) {
        for (size_t q = 0; q < P; q++) col[q] = pos[3 * q + e];
        hipCheck(hipMemcpyAsync(_pos0.p + e * P, col.data(), sizeof(float) * P, hipMemcpyHostToDevice, _s), "upload");
        hipCheck(hipStreamSynchronize(_s), "sync");
    }
    if (solidNodes)
        hipCheck(hipMemcpyAsync(_solid, solidNodes, sizeof(float) * NN, hipMemcpyHostToDevice, _s), "upload");
    else
        solidBoxPhi(g, _solid, _s);
    faceWeights(g, _solid, _wt[0], _wt[1], _wt[2], _s);
    nearSolidGrid(g, _solid, _near, _nearTmp, _ci, _cj, _ck, (int)std::ceil((float)st.cfl / (float)3), _s);
    hipCheck(hipStreamSynchronize(_s), "sync");
}

Rollout::~Rollout() {
    if (_s) (void)hipStreamDestroy(_s);
}

Particles Rollout::particlesAt(float *base) const {
    const size_t P = (size_t)_np;
    return Particles{base, base + P, base + 2 * P, base + 3 * P, base + 4 * P, base + 5 * P, _np};
}

void Rollout::setInitialState() {
    const size_t P = (size_t)_np;
    copyD(_state.p, _pos0.p, sizeof(float) * 3 * P, _s);
    Particles S = particlesAt(_state.p);
    sampleParticles(_g, _np, S.px, S.py, S.pz, MacView{_init[0], _init[1], _init[2]}, S.vx, S.vy, S.vz, _s);
}

// FluidSimulation::_stepFluid (fluidsimulation.cpp:10078-10121) on the working
// state, in the engine's order and with the engine's kernels
void Rollout::substep(int k, bool freeze, bool record) {
    const Grid g = _g;
    const size_t F = _nf[0] + _nf[1] + _nf[2];
    Particles S = particlesAt(_state.p);
    cellSort(g, S, _cs, _scan, _s);
    if (freeze) {
        copyD(_phi.p, _phiRec.p + (size_t)k * _N, sizeof(float) * _N, _s);
    } else {
        liquidSDF(g, _cs, _np, _radius, _phi, _s);
        liquidSDFPostprocess(g, _solid, _phi, _s);
        if (record) copyD(_phiRec.p + (size_t)k * _N, _phi.p, sizeof(float) * _N, _s);
    }
    p2g(g, _cs, _radius, _u[0], _u[1], _u[2], _valid[0], _valid[1], _valid[2], _s, _ws[0], _ws[1], _ws[2]);
    size_t off = (size_t)k * F;
    for (int d = 0; d < 3; d++) {
        if (freeze) {
            freezeNormKernel<<<blocksFor(_nf[d], TPB), TPB, 0, _s>>>(_u[d], _ws[d], _validRec.p + off, _nf[d]);
            copyD(_valid[d].p, _validRec.p + off, _nf[d], _s);
        } else if (record)
            copyD(_validRec.p + off, _valid[d].p, _nf[d], _s);
        off += _nf[d];
    }
    const int wi[3] = {g.ni + 1, g.ni, g.ni}, wj[3] = {g.nj, g.nj + 1, g.nj}, wk[3] = {g.nk, g.nk, g.nk + 1};
    for (int d = 0; d < 3; d++) extrapolateInit(wi[d], wj[d], wk[d], _valid[d], _st0[d], _s);
    for (int l = 0; l < _layers; l++)
        for (int d = 0; d < 3; d++) extrapolateLayer(wi[d], wj[d], wk[d], _u[d], _st0[d], l, _s);
    for (int d = 0; d < 3; d++) copyD(_su[d].p, _u[d].p, sizeof(float) * _nf[d], _s);
    // the body force is a float vector scaled by the double substep length
    // (fluidsimulation.cpp:5837-5876)
    const float bf[3] = {(float)_st.gx, (float)_st.gy, (float)_st.gz};
    for (int d = 0; d < 3; d++)
        if (std::fabs(bf[d]) > 1e-6f) addToFaces(_u[d], _nf[d], (float)(bf[d] * _st.dt), _s);
    PressureResult r = _ps.solve(_st.dt, _u[0], _u[1], _u[2], _wt[0], _wt[1], _wt[2], _phi, _st.tol, _st.maxIter, 0,
                                 1e-4, _pressure, _scan, _s, nullptr, nullptr, _refDots);
    _iters += r.iterations;
    if (!r.success) throw std::runtime_error("rollout: a pressure solve did not converge");
    applyPressure(g, _st.dt, _pressure, _phi, _wt[0], _wt[1], _wt[2], _u[0], _u[1], _u[2], _valid[0], _valid[1],
                  _valid[2], _s);
    for (int d = 0; d < 3; d++) extrapolateInit(wi[d], wj[d], wk[d], _valid[d], _st1[d], _s);
    for (int l = 0; l < _layers; l++)
        for (int d = 0; d < 3; d++) extrapolateLayer(wi[d], wj[d], wk[d], _u[d], _st1[d], l, _s);
    for (int d = 0; d < 3; d++) {
        copyD(_cu[d].p, _su[d].p, sizeof(float) * _nf[d], _s);
        constrainFaces(_cu[d], _wt[d], _nf[d], _s);
        constrainFaces(_u[d], _wt[d], _nf[d], _s);
    }
    g2pAdvect(g, S, MacView{_u[0], _u[1], _u[2]}, MacView{_cu[0], _cu[1], _cu[2]}, _solid, _near, _ci, _cj, _ck,
              (float)_st.picflip, (float)(1 - _st.picflip), _st.dt, _st.cfl, _s, _moved);
    countKernel<<<blocksFor(_np, TPB), TPB, 0, _s>>>(_moved, _np, _movedCount);
}

void Rollout::forward(const float *u0, const float *v0, const float *w0, float *density, bool freeze) {
    if (freeze && !_recorded) throw std::runtime_error("rollout: no recorded forward to freeze the masks of");
    const float *src[3] = {u0, v0, w0};
    for (int d = 0; d < 3; d++)
        hipCheck(hipMemcpyAsync(_init[d], src[d], sizeof(float) * _nf[d], hipMemcpyHostToDevice, _s), "upload");
    hipCheck(hipMemsetAsync(_movedCount, 0, sizeof(int), _s), "memset");
    _iters = 0;
    setInitialState();
    const size_t S6 = 6 * (size_t)_np;
    const int C = _st.checkpointEvery;
    copyD(_ckpt.p, _state.p, sizeof(float) * S6, _s);
    for (int k = 0; k < _st.substeps; k++) {
        substep(k, freeze, !freeze);
        if ((k + 1) % C == 0 && k + 1 < _st.substeps)
            copyD(_ckpt.p + (size_t)((k + 1) / C) * S6, _state.p, sizeof(float) * S6, _s);
    }
    copyD(_final.p, _state.p, sizeof(float) * S6, _s);
    Particles S = particlesAt(_final.p);
    cellSort(_g, S, _cs, _scan, _s);
    densityGrid(_g, _cs, _radius, _rho, _s);
    int moved = 0;
    hipCheck(hipMemcpyAsync(&moved, _movedCount, sizeof(int), hipMemcpyDeviceToHost, _s), "D2H");
    hipCheck(hipMemcpyAsync(density, _rho, sizeof(float) * _N, hipMemcpyDeviceToHost, _s), "D2H");
    hipCheck(hipStreamSynchronize(_s), "sync");
    _collided = moved;
    _fwdIters = _iters;
    if (!freeze) _recorded = true;
    _haveForward = true;
}

void Rollout::sampleAdjointInto(DevBuf<float> *out, int npts, const float *px, const float *py, const float *pz,
                                const float *cot, bool accumulate) {
    for (int d = 0; d < 3; d++) {
        float *dst = accumulate ? _tmp[d].p : out[d].p;
        sampleAdjoint(_g, d, npts, px, py, pz, cot, _akey, _aslot, _acounts, _aoffsets, _aperm, _ascan, dst, _s);
        if (accumulate) addKernel<<<blocksFor(_nf[d], TPB), TPB, 0, _s>>>(out[d], _tmp[d], _nf[d]);
    }
}

// Adjoint of substep(k) whose grids are in place (recomputed just before)
// and whose starting state is `start` (6P).  In: _pbar, _vbar (cotangents of
// the state after the substep).  Out: the same for the state before it.
void Rollout::adjointSubstep(const float *start) {
    const Grid g = _g;
    const size_t P = (size_t)_np, P3 = 3 * P;
    const float *px = start, *py = start + P, *pz = start + 2 * P;
    const float *vx = start + 3 * P, *vy = start + 4 * P, *vz = start + 5 * P;
    // G2P blend and RK3 (collision response as the identity)
    BlendAdj B;
    B.sav = MacView{_cu[0], _cu[1], _cu[2]};
    B.r1 = (float)_st.picflip;
    B.r2 = (float)(1 - _st.picflip);
    B.vbarOut = _vbar;
    B.vbar = _vbarNew;
    B.cotSav = _cotSav;
    advectRK3Adjoint(g, _np, px, py, pz, MacView{_u[0], _u[1], _u[2]}, _st.dt, _pbar, _pbarNew, _pts, _cot, _s, &B);
    // field cotangents: the three RK3 samples (k1's carries the blend's
    // vPIC term) on the final field, the saved-field sample on the saved one
    sampleAdjointInto(_ub, (int)P3, _pts.p, _pts.p + P3, _pts.p + 2 * P3, _cot, false);
    sampleAdjointInto(_sb, _np, px, py, pz, _cotSav, false);
    // solid constraint (a mask) on both, then the second extrapolation
    const int wi[3] = {g.ni + 1, g.ni, g.ni}, wj[3] = {g.nj, g.nj + 1, g.nj}, wk[3] = {g.nk, g.nk, g.nk + 1};
    for (int d = 0; d < 3; d++) {
        constrainFaces(_ub[d], _wt[d], _nf[d], _s);
        constrainFaces(_sb[d], _wt[d], _nf[d], _s);
        extrapolateAdjoint(wi[d], wj[d], wk[d], _ub[d], _st1[d], _layers, _s);
    }
    // pressure projection: K ubar + B^T A^-1 (C^T ubar), the same FP64 solve
    projectionAdjointCells(g, _st.dt, _phi, _wt[0], _wt[1], _wt[2], _ub[0], _ub[1], _ub[2], _cellBar, _s);
    PressureResult r = _ps.solve(_st.dt, _zero[0], _zero[1], _zero[2], _wt[0], _wt[1], _wt[2], _phi, _st.tol,
                                 _st.maxIter, 0, 1e-4, _pressure, _scan, _s, nullptr, nullptr, _refDots, _cellBar,
                                 _lambda);
    _bwdIters += r.iterations;  // the adjoint solves only; recomputed forward solves are not counted
    if (!r.success) throw std::runtime_error("rollout: an adjoint pressure solve did not converge");
    projectionAdjointFaces(g, _st.dt, _phi, _wt[0], _wt[1], _wt[2], _ub[0], _ub[1], _ub[2], _lambda, _tmp[0], _tmp[1],
                           _tmp[2], _s);
    // body force adds a constant; the saved copy adds its cotangent; then the
    // first extrapolation
    for (int d = 0; d < 3; d++) {
        copyD(_ub[d].p, _tmp[d].p, sizeof(float) * _nf[d], _s);
        addKernel<<<blocksFor(_nf[d], TPB), TPB, 0, _s>>>(_ub[d], _sb[d], _nf[d]);
        extrapolateAdjoint(wi[d], wj[d], wk[d], _ub[d], _st0[d], _layers, _s);
    }
    // P2G in velocities and positions (the saved field holds P2G's values on
    // every face that had weight)
    p2gAdjoint(g, _np, px, py, pz, vx, vy, vz, _radius, _ub[0], _ub[1], _ub[2], _ws[0], _ws[1], _ws[2], _su[0],
               _su[1], _su[2], _vb3.p, _vb3.p + P, _vb3.p + 2 * P, _xbarP2G, _s);
    combineKernel<<<blocksFor(P, TPB), TPB, 0, _s>>>(_np, _pbarNew, _xbarP2G, _vbarNew, _vb3, _pbar, _vbar);
}

void Rollout::backward(const float *densityBar, float *gu, float *gv, float *gw) {
    if (!_haveForward) throw std::runtime_error("rollout: backward needs a forward");
    if (!_recorded) throw std::runtime_error("rollout: backward needs a recording forward");
    const size_t P = (size_t)_np, S6 = 6 * P;
    const int C = _st.checkpointEvery, K = _st.substeps;
    _bwdIters = 0;
    // density of the final state: its positions' cotangent; velocities get none
    hipCheck(hipMemcpyAsync(_rhoBar, densityBar, sizeof(float) * _N, hipMemcpyHostToDevice, _s), "upload");
    Particles S = particlesAt(_final.p);
    densityAdjoint(_g, _np, S.px, S.py, S.pz, _radius, _rhoBar, _pbar, _s);
    hipCheck(hipMemsetAsync(_vbar, 0, sizeof(float) * 3 * P, _s), "memset");
    const int segments = (K + C - 1) / C;
    for (int sgi = segments - 1; sgi >= 0; sgi--) {
        const int k0 = sgi * C, k1 = std::min(k0 + C, K);
        copyD(_state.p, _ckpt.p + (size_t)sgi * S6, sizeof(float) * S6, _s);
        for (int k = k0; k < k1; k++) {
            copyD(_seg.p + (size_t)(k - k0) * S6, _state.p, sizeof(float) * S6, _s);
            if (k + 1 < k1) substep(k, true, false);
        }
        for (int k = k1 - 1; k >= k0; k--) {
            const float *start = _seg.p + (size_t)(k - k0) * S6;
            copyD(_state.p, start, sizeof(float) * S6, _s);
            substep(k, true, false);
            adjointSubstep(start);
        }
    }
    // initial velocities are G2P of the initial field at the seeded positions
    sampleAdjointInto(_ub, _np, _pos0.p, _pos0.p + P, _pos0.p + 2 * P, _vbar, false);
    float *dst[3] = {gu, gv, gw};
    for (int d = 0; d < 3; d++)
        hipCheck(hipMemcpyAsync(dst[d], _ub[d], sizeof(float) * _nf[d], hipMemcpyDeviceToHost, _s), "D2H");
    hipCheck(hipStreamSynchronize(_s), "sync");
}

void Rollout::positions(float *out) const {
    const size_t P = (size_t)_np;
    std::vector<float> c(3 * P);
    hipCheck(hipMemcpyAsync(c.data(), _final.p, sizeof(float) * 3 * P, hipMemcpyDeviceToHost, _s), "D2H");
    hipCheck(hipStreamSynchronize(_s), "sync");
    for (size_t q = 0; q < P; q++)
        for (int e = 0; e < 3; e++) out[3 * q + e] = c[e * P + q];
}

}  // namespace ff
