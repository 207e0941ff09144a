// labels.hpp -- connected-component labelling of the activated bonds, the
// whole-lattice way of building a Wolff cluster: every bond's activation is a
// pure function of its keyed uniform and the pre-flip spins, so the cluster of
// the seed is its component in the graph of activated bonds whatever order the
// bonds are examined in.  A grid-wide lock-free union-find (hooks a root under
// a smaller root with a CAS; finds halve paths) labels all components in
// parallel; the seed's label then marks its cluster.
//
// Loads and stores of the parent array go through agent-scope atomics so that
// every CU sees the links made by the others (plain loads may be served from a
// CU's own L1).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace ucmc {

// Per-flip control block written by the first kernel of a flip and read by the
// following ones.
struct ClusterState {
    double r[4];                  // Wolff axis (O(N)); r[0] = 1 for Ising
    std::uint32_t seed;           // seed site
    std::uint32_t skip;           // 1: the touch target is reached, flip not run
    unsigned long long touched;   // sites flipped so far in this call
    unsigned long long count[8];  // per-flip integer sums (cluster size, bond counts ...)
    double sum[8];                // per-flip floating-point sums
    std::uint32_t blocks_done;    // last-block detection
    std::uint32_t accept;         // SLMC: the Metropolis-Hastings decision
};

__device__ __forceinline__ std::uint32_t lab_load(const std::uint32_t* p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void lab_store(std::uint32_t* p, std::uint32_t v)
{
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Root of x, halving the path on the way (every store points a non-root at an
// ancestor, so concurrent finds and hooks keep each tree a tree).
__device__ __forceinline__ std::uint32_t lab_find(std::uint32_t* parent, std::uint32_t x)
{
    for (;;) {
        const std::uint32_t p = lab_load(parent + x);
        if (p == x) return x;
        const std::uint32_t g = lab_load(parent + p);
        if (g == p) return p;
        lab_store(parent + x, g);
        x = g;
    }
}

// Merge the components of a and b: the larger root is hooked under the
// smaller one, retried until the hook lands on a root.
__device__ __forceinline__ void lab_unite(std::uint32_t* parent, std::uint32_t a, std::uint32_t b)
{
    for (;;) {
        a = lab_find(parent, a);
        b = lab_find(parent, b);
        if (a == b) return;
        if (a < b) {
            const std::uint32_t t = a;
            a = b;
            b = t;
        }
        std::uint32_t expected = a;
        if (__hip_atomic_compare_exchange_strong(parent + a, &expected, b, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                                 __HIP_MEMORY_SCOPE_AGENT))
            return;
    }
}

namespace {

// parent[i] = root of i for every site (one thread per site, after all unites).
// The walk to the root only reads: a halving store here could replace a root
// pointer another thread has just written with a non-root ancestor.
__global__ void lab_flatten(std::uint32_t* parent, std::uint32_t n, const ClusterState* st)
{
    const std::uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || st->skip) return;
    std::uint32_t x = i;
    for (std::uint32_t p = lab_load(parent + x); p != x; p = lab_load(parent + x)) x = p;
    lab_store(parent + i, x);
}

}  // namespace

}  // namespace ucmc
