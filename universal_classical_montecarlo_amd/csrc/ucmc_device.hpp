// ucmc_device.hpp -- device-side building blocks shared by every kernel:
// the per-site counter-based PCG32 stream, uniform conversions and wave
// helpers.  gfx950 only.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace ucmc {

constexpr std::uint64_t kPcgMult = 6364136223846793005ull;

// SplitMix64 output number (counter + 1) for state run_seed.
__host__ __device__ inline std::uint64_t key_of(std::uint64_t run_seed, std::uint64_t counter)
{
    std::uint64_t z = run_seed + (counter + 1u) * 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

constexpr std::uint64_t kWolffCounterBase = 1ull << 62;

// The state whose XSH-RR output is the first draw of Pcg32(seed, stream)
// (rng.hpp:14-20 seeding, rng.hpp:31-40 step): (seed + inc) M + inc with
// inc = 2 stream + 1.  It is affine in the stream: A + stream * D with
// A = (seed + 1) M + 1 and D = 2 M + 2.
__host__ __device__ inline std::uint64_t first_state(std::uint64_t seed, std::uint64_t stream)
{
    const std::uint64_t inc = (stream << 1u) | 1u;
    return (seed + inc) * kPcgMult + inc;
}
constexpr std::uint64_t kStreamStride = 2ull * kPcgMult + 2ull;

// XSH-RR output of a 64-bit state: rotr32(((old >> 18) ^ old) >> 27, old >> 59).
// ((old >> 18) ^ old) >> 27 == (old >> 27) ^ (old >> 45), whose low 32 bits are
// alignbit(hi, lo, 27) ^ (hi >> 13); the rotate is alignbit(x, x, rot).
__device__ __forceinline__ std::uint32_t pcg_out(std::uint64_t old)
{
    const std::uint32_t hi = static_cast<std::uint32_t>(old >> 32);
    const std::uint32_t lo = static_cast<std::uint32_t>(old);
    const std::uint32_t xs = __builtin_amdgcn_alignbit(hi, lo, 27) ^ (hi >> 13);
    return __builtin_amdgcn_alignbit(xs, xs, hi >> 27);
}

// A full Pcg32 for kernels that need several draws per site.
struct Pcg {
    std::uint64_t state;
    std::uint64_t inc;

    __device__ __forceinline__ void seed(std::uint64_t s, std::uint64_t stream)
    {
        inc = (stream << 1u) | 1u;
        state = (s + inc) * kPcgMult + inc; // two seeding steps folded
    }
    __device__ __forceinline__ std::uint32_t next()
    {
        const std::uint64_t old = state;
        state = old * kPcgMult + inc;
        return pcg_out(old);
    }
    // rng.hpp:24-27: exactly k * 2^-32, in double.
    __device__ __forceinline__ double uni() { return static_cast<double>(next()) * 0x1.0p-32; }
};

// The accept test of driver.hpp:84, log(u) < log_p with u = k 2^-32: decided
// in fp32 when the two sides are more than 1e-3 apart (the fp32 logarithm is
// within ~3e-6 of the double one for every k, so the outcome is the same), and
// with the double expression otherwise.  k = 0 gives log 0 = -inf (accept);
// a NaN log_p rejects, as in the reference.
__device__ __forceinline__ bool accept_draw(std::uint32_t k, double log_p)
{
    const float lf = __logf(static_cast<float>(k) * 0x1.0p-32f);
    const float lp = static_cast<float>(log_p);
    if (lf < lp - 1e-3f) return true;
    if (lf > lp + 1e-3f) return false;
    return log(static_cast<double>(k) * 0x1.0p-32) < log_p;
}

// Wave-wide shifts by one lane (DPP wave_shr:1 / wave_shl:1).  Lane 0
// (resp. 63) receives `edge`.
__device__ __forceinline__ std::uint32_t from_prev_lane(std::uint32_t v, std::uint32_t edge)
{
    return static_cast<std::uint32_t>(
        __builtin_amdgcn_update_dpp(static_cast<int>(edge), static_cast<int>(v), 0x138, 0xf, 0xf, false));
}
__device__ __forceinline__ std::uint32_t from_next_lane(std::uint32_t v, std::uint32_t edge)
{
    return static_cast<std::uint32_t>(
        __builtin_amdgcn_update_dpp(static_cast<int>(edge), static_cast<int>(v), 0x130, 0xf, 0xf, false));
}

// Wave-level double sum via shuffles (ds_swizzle-free butterfly using __shfl_xor).
__device__ __forceinline__ double wave_sum(double v)
{
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}
__device__ __forceinline__ unsigned long long wave_sum_u64(unsigned long long v)
{
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

}  // namespace ucmc
