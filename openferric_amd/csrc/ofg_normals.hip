// The reference's per-stream standard normals on the GPU: stream i is the
// generator of resolve_stream_seed(seed, first + i) (src/math/fast_rng.rs:
// 202-214; xoshiro256++ or the Pcg64 kind) and yields n normals by
// fill_standard_normals / sample_standard_normal (fast_rng.rs:224-240):
// Acklam's inverse CDF of the clamped uniform. One lane per stream, draws
// inverted 8 at a time with the tail values of a block compacted per wave
// (acklam_block_wave); values written stream-major or step-major.
#include "ofg_normals.hpp"
#include "ofg_math.hpp"

#include <type_traits>

namespace ofg {

namespace {

constexpr int NB = 256;   // threads per block
constexpr int NK = 8;     // draws per inverse-CDF block

template <int RNG>
__global__ __launch_bounds__(NB) void normals_kernel(NormalsArgs a) {
    __shared__ double tail_slots[NB * NK];
    double *slots = tail_slots + (threadIdx.x / 64) * 64 * NK;
    const int64_t i = (int64_t)blockIdx.x * NB + threadIdx.x;
    if (i >= a.n_streams) return;
    using Rng = typename std::conditional<RNG == OFG_RNG_PCG64, Pcg64, Xoshiro256pp>::type;
    Rng rng;
    rng.seed(stream_seed(a.seed, a.first + (uint64_t)i));
    const int64_t n = a.n_per_stream;
    for (int64_t j = 0; j < n; j += NK) {
        double z[NK];
        acklam_block_wave<NK>(
            [&](int k) { return j + k < n ? uniform_open01(rng.next_f64()) : 0.5; }, z, slots);
#pragma unroll
        for (int k = 0; k < NK; ++k) {
            if (j + k < n) {
                const int64_t at = a.step_major ? (j + k) * a.n_streams + i : i * n + j + k;
                a.out[at] = z[k];
            }
        }
    }
}

}  // namespace

hipError_t launch_normals(const NormalsArgs &a, int32_t rng_mode, hipStream_t stream) {
    const int64_t blocks = (a.n_streams + NB - 1) / NB;
    if (rng_mode == OFG_RNG_PCG64)
        hipLaunchKernelGGL(normals_kernel<OFG_RNG_PCG64>, dim3((uint32_t)blocks), dim3(NB), 0,
                           stream, a);
    else
        hipLaunchKernelGGL(normals_kernel<OFG_RNG_REFERENCE>, dim3((uint32_t)blocks), dim3(NB), 0,
                           stream, a);
    return hipGetLastError();
}

}  // namespace ofg
