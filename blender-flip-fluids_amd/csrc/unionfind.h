// unionfind.h -- lock-free union-find over flat cell indices on the device.
//
// parent[x] is x for a root, a smaller index otherwise, -1 for a cell outside
// the graph.  Every hook points a root at a smaller root, so parent indices
// only decrease, every find ends, and concurrent unions from all threads of a
// launch leave one root per connected component.
#pragma once

#include <hip/hip_runtime.h>

namespace ff {

__device__ inline int loadParent(const int *parent, int x) {
    return __hip_atomic_load(parent + x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// root of x with path halving; parent[y] <= y for every member y, so the walk
// ends, and the halving stores only ever lower an entry to an ancestor
__device__ inline int findRoot(int *parent, int x) {
    int p = loadParent(parent, x);
    while (p != x) {
        const int gp = loadParent(parent, p);
        if (gp != p) __hip_atomic_store(parent + x, gp, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        x = p;
        p = gp;
    }
    return x;
}

// hook the larger root onto the smaller; a failed swap returns the larger
// one's new (smaller) parent, so each retry starts lower and the loop ends
__device__ inline void unite(int *parent, int a, int b) {
    int ra = findRoot(parent, a), rb = findRoot(parent, b);
    while (ra != rb) {
        if (ra < rb) {
            const int t = ra;
            ra = rb;
            rb = t;
        }
        const int old = atomicCAS(parent + ra, ra, rb);
        if (old == ra) return;
        ra = findRoot(parent, old);
        rb = findRoot(parent, rb);
    }
}

}  // namespace ff
