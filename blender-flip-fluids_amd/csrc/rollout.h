// rollout.h -- a differentiable run of FLIP substeps on the device.
//
// Rollout runs `substeps` substeps of a fixed length from particles at fixed
// seeded positions whose velocities are the trilinear sample (G2P) of an
// initial MAC velocity field, and returns the particle density on the cells
// after the last substep.  Each substep is the engine's _stepFluid
// (fluidsimulation.cpp:10078-10121: liquid SDF, P2G, extrapolation, saved
// field, body force, pressure projection, extrapolation, solid constraint,
// G2P blend, RK3 advection with the collision response) without particle
// removal.  backward() returns the gradient of <densityBar, density> with
// respect to the initial MAC field.  The forward keeps the particle state
// every `checkpointEvery` substeps in HBM; the backward recomputes each
// segment from its checkpoint, then each substep's grids right before its
// adjoint.  The liquid SDF and the P2G validity flags of every substep are
// recorded by the forward and held fixed in the backward (the adjoint is that
// of the substep for fixed cell classification), and the collision response
// is taken as the identity.
#pragma once

#include <vector>

#include "engine.h"
#include "kernels.h"
#include "pressure.h"

namespace ff {

struct RolloutSettings {
    double dt = 1.0 / 120.0;   // substep length
    int substeps = 1;
    int checkpointEvery = 1;
    double picflip = 0.05;     // PIC share of the G2P blend (the engine's PIC/FLIP ratio)
    double gx = 0.0, gy = -9.81, gz = 0.0;
    double cfl = 5.0;          // sets the extrapolation layers and the collision search, as in the engine
    double tol = 1e-9;         // pressure solves (forward and adjoint), FP64 MIC(0)-PCG
    int maxIter = 900;
};

class Rollout {
public:
    // pos: np x 3 host positions; solidNodes: (ni+1)(nj+1)(nk+1) host nodal
    // solid SDF or null for the domain walls
    Rollout(Grid g, int np, const float *pos, const float *solidNodes, const RolloutSettings &st);
    ~Rollout();
    Rollout(const Rollout &) = delete;
    Rollout &operator=(const Rollout &) = delete;

    // u0/v0/w0: host face grids of the initial MAC field; density: N host
    // floats.  freeze = false records the liquid SDF and P2G flags of every
    // substep; freeze = true reuses those of the last recording forward.
    void forward(const float *u0, const float *v0, const float *w0, float *density, bool freeze);
    // gradient of <densityBar, density> in the initial field of the last
    // forward (host face grids)
    void backward(const float *densityBar, float *gu, float *gv, float *gw);
    void positions(float *out) const;  // np x 3 after the last forward
    // particles whose position the collision response changed, summed over
    // the substeps of the last forward; pressure iterations of the last
    // forward and backward
    long long collided() const { return _collided; }
    long long forwardIterations() const { return _fwdIters; }
    long long backwardIterations() const { return _bwdIters; }
    int particles() const { return _np; }
    // FP64 dot products of every pressure solve in the reference's summation
    // order for `threads` threads (blaswrapper.h:27-57), or 0 for the
    // parallel tree (the default)
    void setReferenceSummation(int threads) { _refDots = threads < 0 ? 0 : threads; }

private:
    Grid _g;
    RolloutSettings _st;
    int _np;
    size_t _N, _nf[3];
    int _layers;
    double _radius;
    hipStream_t _s = nullptr;
    bool _recorded = false, _haveForward = false;
    long long _collided = 0, _fwdIters = 0, _bwdIters = 0, _iters = 0;
    int _refDots = 0;

    DevBuf<float> _pos0;            // 3P: seeded positions (SoA)
    DevBuf<float> _state;           // 6P: working particle state (px py pz vx vy vz)
    DevBuf<float> _final;           // 6P: state after the last forward
    DevBuf<float> _ckpt;            // (substeps / checkpointEvery + 1) x 6P
    DevBuf<float> _seg;             // checkpointEvery x 6P: states of one segment
    DevBuf<float> _phiRec;          // substeps x N
    DevBuf<uint8_t> _validRec;      // substeps x (faces)
    DevBuf<float> _init[3];         // initial MAC field of the last forward
    CellSort _cs{};
    DevBuf<int> _keys, _slot, _perm, _counts, _offsets, _scan;
    DevBuf<float> _sorted[6];
    DevBuf<float> _phi, _solid, _wt[3], _pressure, _rho, _rhoBar;
    DevBuf<uint8_t> _near, _nearTmp, _moved;
    int _ci = 0, _cj = 0, _ck = 0;
    DevBuf<float> _u[3], _su[3], _cu[3], _ws[3];  // field, saved (unconstrained), saved (constrained), weight sums
    DevBuf<uint8_t> _valid[3], _st0[3], _st1[3];
    PressureSolver _ps;
    // adjoint
    DevBuf<float> _ub[3], _sb[3], _tmp[3], _zero[3];
    DevBuf<float> _pbar, _vbar, _pbarNew, _vbarNew, _xbarP2G, _vb3, _pts, _cot, _cotSav;
    DevBuf<int> _akey, _aslot, _aperm, _acounts, _aoffsets, _ascan;
    DevBuf<double> _cellBar, _lambda;
    DevBuf<int> _movedCount;

    Particles particlesAt(float *base) const;
    void setInitialState();
    // one substep on the working state; k indexes the recorded masks
    void substep(int k, bool freeze, bool record);
    void adjointSubstep(const float *start);
    void sampleAdjointInto(DevBuf<float> *out, int npts, const float *px, const float *py, const float *pz,
                           const float *cot, bool accumulate);
};

}  // namespace ff
