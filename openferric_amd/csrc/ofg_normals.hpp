// The reference's per-stream normals on the GPU (ofg_normals.hip): launch
// interface shared with the C ABI.
#pragma once

#include "ofg_internal.hpp"

namespace ofg {

struct NormalsArgs {
    double *out;            // n_streams * n_per_stream values
    uint64_t seed;          // base seed of stream_seed
    uint64_t first;         // stream index of out's first stream
    int64_t n_streams;
    int64_t n_per_stream;
    int32_t step_major;     // out[j * n_streams + i], else out[i * n_per_stream + j]
};

// Fills a.out on `stream`; rng_mode OFG_RNG_REFERENCE or OFG_RNG_PCG64.
hipError_t launch_normals(const NormalsArgs &a, int32_t rng_mode, hipStream_t stream);

}  // namespace ofg
