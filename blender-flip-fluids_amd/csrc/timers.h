// timers.h -- optional HIP-event timing of kernel groups on a stream.
//
// When enabled, begin()/end() record events around a kernel group on the
// stream the kernels run on; collect() waits for them and accumulates the
// elapsed milliseconds and launch counts per group.  Disabled, the calls are
// no-ops, so the substep pays nothing.
#pragma once

#include <hip/hip_runtime.h>

#include <vector>

namespace ff {

enum TimerId {
    T_CELLSORT = 0,
    T_LIQUID_SDF,
    T_P2G,
    T_EXTRAPOLATE,
    T_ASSEMBLE,
    T_MIC0_FACTOR,
    T_SPMV,
    T_MIC0_APPLY,
    T_VECTOR_OPS,
    T_G2P_ADVECT,
    T_REMOVAL,
    T_APPLY_PRESSURE,
    T_VISCOSITY,
    T_COUNT
};

class KernelTimers {
public:
    ~KernelTimers() {
        for (hipEvent_t e : _pool) (void)hipEventDestroy(e);
    }
    void enable(bool on) {
        _on = on;
        reset();
    }
    bool enabled() const { return _on; }
    void reset() {
        for (int q = 0; q < T_COUNT; q++) {
            _ms[q] = 0.0;
            _count[q] = 0;
        }
        _pending.clear();
        _used = 0;
    }
    void begin(int id, hipStream_t s, int launches = 1) {
        if (!_on) return;
        _cur[id] = take();
        _launches[id] = launches;
        (void)hipEventRecord(_cur[id], s);
    }
    void end(int id, hipStream_t s) {
        if (!_on) return;
        hipEvent_t e = take();
        (void)hipEventRecord(e, s);
        _pending.push_back(Pending{id, _cur[id], e, _launches[id]});
    }
    // wait for all recorded pairs and accumulate
    void collect() {
        if (!_on) return;
        for (const Pending &p : _pending) {
            (void)hipEventSynchronize(p.b);
            float ms = 0.0f;
            (void)hipEventElapsedTime(&ms, p.a, p.b);
            _ms[p.id] += ms;
            _count[p.id] += p.launches;
        }
        _pending.clear();
        _used = 0;
    }
    double ms(int id) const { return _ms[id]; }
    long long launches(int id) const { return _count[id]; }

private:
    struct Pending {
        int id;
        hipEvent_t a, b;
        int launches;
    };
    bool _on = false;
    double _ms[T_COUNT] = {};
    long long _count[T_COUNT] = {};
    hipEvent_t _cur[T_COUNT] = {};
    int _launches[T_COUNT] = {};
    std::vector<Pending> _pending;
    std::vector<hipEvent_t> _pool;
    size_t _used = 0;
    hipEvent_t take() {
        if (_used == _pool.size()) {
            hipEvent_t e;
            (void)hipEventCreate(&e);
            _pool.push_back(e);
        }
        return _pool[_used++];
    }
};

}  // namespace ff
