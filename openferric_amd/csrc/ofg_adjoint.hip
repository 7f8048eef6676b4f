// Pathwise adjoint (reverse-mode) Greeks kernel for MI355X (gfx950), FP64.
//
// Per sample (and per antithetic leg):
//   forward sweep   the path exactly as the pricing kernel simulates it (same
//                   draws, same operations), storing the state (S, and v for
//                   Heston) every adj_segment(model) steps into this wave's scratch
//                   slot, laid out [checkpoint][field][lane] so every store is
//                   a coalesced 512-byte row;
//   payoff          value and the payoff's derivative seeds (terminal,
//                   arithmetic and geometric observation weights);
//   backward sweep  segments in reverse: restore the checkpoint, regenerate the
//                   segment's normals by stepping the xoshiro256 stream
//                   backwards (its state transition is invertible), recompute
//                   the segment's states into registers (the variance path is
//                   recomputed, never stored whole), then sweep the adjoints of
//                   S and v back through each step, accumulating the parameter
//                   adjoints.
//
// Semantics follow the pathwise reverse sweep of mc_european_pathwise_aad
// (src/engines/monte_carlo/mc_aad.rs:16-72: positive_part derivative 1{x > 0},
// d/dT through dt = T/n and the discount) generalised to Asian and barrier
// payoffs and to the Heston Euler step (src/models/stochastic.rs:56-71) with
// the generator clamps: indicator derivatives for max(., 0) and max(., 1e-12),
// and d sqrt(v)/dv taken as 0 at v = 0. Barrier indicators have zero
// pathwise derivative.
#include "ofg_adjoint.hpp"
#include "ofg_math.hpp"
#include "ofg_paths.hpp"

#include <type_traits>

namespace ofg {

namespace {

constexpr int WAVE = 64;

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) v += __shfl_xor(v, m, WAVE);
    return v;
}

template <int RNG>
using RngOf = typename std::conditional<RNG == OFG_RNG_FAST, Xoroshiro128p, Xoshiro256pp>::type;

// ---------------------------------------------------------------------------
// Normals of one segment [j0, j0 + len), len <= K, forward or reverse.
// The grouping matches the pricing kernel: fast mode draws three words per
// four normals (GBM: four steps, Heston: two steps); reference draws are one
// word per normal (Heston: z1 then z2 of a step).
// ---------------------------------------------------------------------------
template <int RNG, int MODEL, int K = adj_segment(MODEL)>
__device__ __forceinline__ void seg_forward(RngOf<RNG> &rng, const ExternalNormals &ext, int j0,
                                            int len, double sign, double (&z1)[K],
                                            double (&z2)[K], const double *tab) {
    if constexpr (RNG == OFG_RNG_FAST) {
        if constexpr (MODEL == OFG_MODEL_GBM) {
#pragma unroll
            for (int g = 0; g < K / 4; ++g) {
                if (4 * g < len) {
                    double a, b, c, d;
                    normals4(rng, a, b, c, d, tab);
                    z1[4 * g] = sign * a;
                    z1[4 * g + 1] = sign * b;
                    z1[4 * g + 2] = sign * c;
                    z1[4 * g + 3] = sign * d;
                }
            }
        } else {
#pragma unroll
            for (int g = 0; g < K / 2; ++g) {
                if (2 * g < len) {
                    double a, b, c, d;
                    normals4(rng, a, b, c, d, tab);
                    z1[2 * g] = sign * a;
                    z2[2 * g] = sign * b;
                    z1[2 * g + 1] = sign * c;
                    z2[2 * g + 1] = sign * d;
                }
            }
        }
    } else if constexpr (RNG == OFG_RNG_REFERENCE) {
        // the segment's uniforms, then the inverse CDF with its tail region
        // deferred over the whole segment (acklam_block)
        constexpr int PER = MODEL == OFG_MODEL_HESTON ? 2 : 1;
        double p[PER * K], z[PER * K];
#pragma unroll
        for (int k = 0; k < K; ++k)
#pragma unroll
            for (int e = 0; e < PER; ++e)
                p[PER * k + e] = k < len ? uniform_open01(rng.next_f64()) : 0.5;
        acklam_block<PER * K>(p, z);
#pragma unroll
        for (int k = 0; k < K; ++k) {
            if (k < len) {
                z1[k] = sign * z[PER * k];
                if constexpr (MODEL == OFG_MODEL_HESTON) z2[k] = sign * z[PER * k + 1];
            }
        }
    } else {
#pragma unroll
        for (int k = 0; k < K; ++k) {
            if (k < len) {
                z1[k] = sign * ext.z1(j0 + k);
                if constexpr (MODEL == OFG_MODEL_HESTON) z2[k] = sign * ext.z2(j0 + k);
            }
        }
    }
}

// The same normals regenerated by stepping the stream backwards over the
// segment's draws (leaves the stream at the segment start).
template <int RNG, int MODEL, int K = adj_segment(MODEL)>
__device__ __forceinline__ void seg_reverse(RngOf<RNG> &rng, const ExternalNormals &ext, int j0,
                                            int len, double sign, double (&z1)[K],
                                            double (&z2)[K], const double *tab) {
    if constexpr (RNG == OFG_RNG_FAST) {
        constexpr int G = MODEL == OFG_MODEL_GBM ? K / 4 : K / 2;
        constexpr int SPG = MODEL == OFG_MODEL_GBM ? 4 : 2;  // steps per group
#pragma unroll
        for (int g = G - 1; g >= 0; --g) {
            if (SPG * g < len) {
                const uint64_t wc = rng.prev_u64();
                const uint64_t wb = rng.prev_u64();
                const uint64_t wa = rng.prev_u64();
                double a, b, c, d;
                box_muller_pair(wa, (uint32_t)(wc >> 32), a, b, tab);
                box_muller_pair(wb, (uint32_t)wc, c, d, tab);
                if constexpr (MODEL == OFG_MODEL_GBM) {
                    z1[4 * g] = sign * a;
                    z1[4 * g + 1] = sign * b;
                    z1[4 * g + 2] = sign * c;
                    z1[4 * g + 3] = sign * d;
                } else {
                    z1[2 * g] = sign * a;
                    z2[2 * g] = sign * b;
                    z1[2 * g + 1] = sign * c;
                    z2[2 * g + 1] = sign * d;
                }
            }
        }
    } else if constexpr (RNG == OFG_RNG_REFERENCE) {
        constexpr int PER = MODEL == OFG_MODEL_HESTON ? 2 : 1;
        double p[PER * K], z[PER * K];
#pragma unroll
        for (int k = K - 1; k >= 0; --k)
#pragma unroll
            for (int e = PER - 1; e >= 0; --e)
                p[PER * k + e] =
                    k < len ? uniform_open01((double)(rng.prev_u64() >> 11) *
                                             (1.0 / 9007199254740992.0))
                            : 0.5;
        acklam_block<PER * K>(p, z);
#pragma unroll
        for (int k = 0; k < K; ++k) {
            if (k < len) {
                z1[k] = sign * z[PER * k];
                if constexpr (MODEL == OFG_MODEL_HESTON) z2[k] = sign * z[PER * k + 1];
            }
        }
    } else {
        seg_forward<RNG, MODEL, K>(rng, ext, j0, len, sign, z1, z2, tab);
    }
}

struct Grad {
    double g[ADJ_NG];
};

// One leg: forward sweep with checkpoints, payoff, backward sweep.
// Returns pv = df x and adds the leg's gradient to gr.
template <int MODEL, int PK, int RNG>
__device__ double adjoint_leg(const OptDev &o, int64_t sample, double sign,
                              const double *__restrict__ w, const AdjArgs &a, double *ck,
                              int lane, const double *tab, Grad &gr) {
    constexpr int K = adj_segment(MODEL);
    const int steps = o.steps;
    const int nseg = (steps + K - 1) / K;
    constexpr bool ASIAN = PK == PK_ASIAN || PK == PK_ASIAN_LOG;
    constexpr int NV = MODEL == OFG_MODEL_HESTON ? 2 : 1;

    RngOf<RNG> rng;
    ExternalNormals ext;
    if constexpr (RNG == OFG_RNG_EXTERNAL) {
        ext.p1 = a.normals1 + o.normals_offset + sample;
        ext.p2 = a.normals2 ? a.normals2 + o.normals_offset + sample : nullptr;
        ext.stride = o.samples;
    } else {
        rng.seed(stream_seed(o.seed, (uint64_t)sample));
    }

    // ---- forward sweep -------------------------------------------------
    Leg g;
    leg_init<PK>(g, o, ASIAN ? w[0] : 0.0);
    double v = o.v0;
    double z1[K], z2[K];
    for (int c = 0; c < nseg; ++c) {
        double *row = ck + (int64_t)c * NV * WAVE + lane;
        row[0] = g.s;
        if constexpr (NV == 2) row[WAVE] = v;
        const int j0 = c * K;
        const int len = min(K, steps - j0);
        seg_forward<RNG, MODEL>(rng, ext, j0, len, sign, z1, z2, tab);
#pragma unroll
        for (int k = 0; k < K; ++k) {
            if (k < len) {
                if constexpr (MODEL == OFG_MODEL_GBM) gbm_step<PK, RNG == OFG_RNG_FAST>(g, o, z1[k], tab);
                else heston_step<PK, RNG != OFG_RNG_FAST>(g, v, o, z1[k], z2[k], tab);
                leg_observe<PK>(g, ASIAN ? w[j0 + k + 1] : 0.0, o.barrier);
            }
        }
    }

    // ---- payoff ---------------------------------------------------------
    double ct, ca, cg;
    const double x = leg_payoff_seeds<PK>(g, o, ct, ca, cg);
    const double pv = o.df * x;
    ca *= o.df;
    cg *= o.df;
    double sb = o.df * ct;   // d pv / d S_{j+1} while sweeping back
    double vb = 0.0;          // d pv / d v_{j+1}
    double mub = 0.0, sigb = 0.0, tb = 0.0, kb = 0.0, thb = 0.0, xib = 0.0, rhob = 0.0;
    const double T = o.dt * (double)steps;
    const double inv_n = 1.0 / (double)steps;
    const double half_inv_sdt = 0.5 / o.sqrt_dt;
    const double rho_ratio = o.rho / o.rho_c;

    // ---- backward sweep -------------------------------------------------
    for (int c = nseg - 1; c >= 0; --c) {
        const double *row = ck + (int64_t)c * NV * WAVE + lane;
        const int j0 = c * K;
        const int len = min(K, steps - j0);
        seg_reverse<RNG, MODEL>(rng, ext, j0, len, sign, z1, z2, tab);
        double S[K + 1], V[K + 1], E[K];
        S[0] = row[0];
        if constexpr (NV == 2) V[0] = row[WAVE];
#pragma unroll
        for (int k = 0; k < K; ++k) {
            if (k < len) {
                if constexpr (MODEL == OFG_MODEL_GBM) {
                    double x_unused;
                    E[k] = gbm_growth<RNG == OFG_RNG_FAST>(o, z1[k], tab, x_unused);
                    S[k + 1] = S[k] * E[k];
                } else {
                    double e_unused;
                    heston_euler<RNG != OFG_RNG_FAST>(S[k], V[k], o, z1[k], z2[k], S[k + 1], V[k + 1], E[k], e_unused, tab);
                }
            }
        }
#pragma unroll
        for (int k = K - 1; k >= 0; --k) {
            if (k < len) {
                if constexpr (ASIAN) {
                    const double wj = w[j0 + k + 1];
                    sb = fma(ca, wj, sb);
                    if constexpr (PK == PK_ASIAN_LOG) sb = fma(cg * wj, 1.0 / S[k + 1], sb);
                }
                if constexpr (MODEL == OFG_MODEL_GBM) {
                    const double z = z1[k];
                    const double eb = sb * S[k + 1];
                    sb *= E[k];
                    sigb = fma(eb, fma(o.sqrt_dt, z, -o.sigma * o.dt), sigb);
                    mub = fma(eb, o.dt, mub);
                    tb = fma(eb, fma(0.5 * o.diffusion, z, o.drift), tb);
                } else {
                    const double s = S[k], vv = V[k];
                    const double sn = s * E[k];
                    const double snb = (sn >= 1e-12) ? sb : 0.0;
                    const double vnb = (V[k + 1] > 0.0) ? vb : 0.0;
                    const double eb = snb * sn;
                    const double v_pos = fmax(vv, 0.0);
                    const double sv = sqrt(v_pos);
                    const double zs = o.rho * z1[k] + o.rho_c * z2[k];
                    mub = fma(eb, o.dt, mub);
                    double vpb = fma(-0.5 * o.dt, eb, -o.kappa * o.dt * vnb);
                    const double svb = fma(eb * o.sqrt_dt, zs, vnb * o.xi * o.sqrt_dt * z1[k]);
                    const double zsb = eb * sv * o.sqrt_dt;
                    rhob = fma(zsb, fma(-rho_ratio, z2[k], z1[k]), rhob);
                    kb = fma(vnb * (o.theta - v_pos), o.dt, kb);
                    thb = fma(vnb * o.kappa, o.dt, thb);
                    xib = fma(vnb * sv * o.sqrt_dt, z1[k], xib);
                    if (sv > 0.0) vpb = fma(svb, 0.5 / sv, vpb);
                    // d/d dt of e and vn, mapped to T through dt = T / n
                    const double de_ddt = fma(sv * zs, half_inv_sdt, o.mu - 0.5 * v_pos);
                    const double dvn_ddt = fma(o.xi * sv * z1[k], half_inv_sdt, o.kappa * (o.theta - v_pos));
                    tb = fma(fma(eb, de_ddt, vnb * dvn_ddt), inv_n * T, tb);
                    sb = snb * E[k];
                    vb = vnb + ((vv > 0.0) ? vpb : 0.0);
                }
            }
        }
    }
    if constexpr (ASIAN) {
        sb = fma(ca, w[0], sb);
        if constexpr (PK == PK_ASIAN_LOG) sb = fma(cg * w[0], 1.0 / o.s0, sb);
    }
    // tb is T d pv / dT through dt = T / n (GBM: sum of eb (drift + diff z / 2);
    // Heston: the same through e and vn); the discount adds -r pv.
    gr.g[0] += sb;
    gr.g[2] += mub - T * pv;     // mu = r - q, and the discount exp(-r T)
    gr.g[3] += tb / T - o.rate * pv;
    if constexpr (MODEL == OFG_MODEL_GBM) {
        gr.g[1] += sigb;
    } else {
        gr.g[1] += vb;
        gr.g[4] += kb;
        gr.g[5] += thb;
        gr.g[6] += xib;
        gr.g[7] += rhob;
    }
    return pv;
}

template <int MODEL, int PK, int RNG, bool ANTI>
__device__ void adjoint_item(const OptDev &o, const WorkItem &wi, int lane, const AdjArgs &a,
                             double *ck, const double *tab, double &spv, double &spv2, Grad &acc) {
    const double *__restrict__ w = a.obs_w + o.obs_offset;
    for (int64_t base = wi.begin; base < wi.end; base += WAVE) {
        const int64_t sample = base + lane;
        if (sample < wi.end) {
            Grad g;
#pragma unroll
            for (int k = 0; k < ADJ_NG; ++k) g.g[k] = 0.0;
            double pv = adjoint_leg<MODEL, PK, RNG>(o, sample, 1.0, w, a, ck, lane, tab, g);
            if constexpr (ANTI) {
                pv += adjoint_leg<MODEL, PK, RNG>(o, sample, -1.0, w, a, ck, lane, tab, g);
                pv *= 0.5;
#pragma unroll
                for (int k = 0; k < ADJ_NG; ++k) g.g[k] *= 0.5;
            }
            spv += pv;
            spv2 += pv * pv;
#pragma unroll
            for (int k = 0; k < ADJ_NG; ++k) acc.g[k] += g.g[k];
        }
    }
}

// Waves per SIMD asked of the register allocator: the fast-draw GBM sweep
// takes about 191 VGPRs left free (2 waves) and is held to 3.
template <int RNG, int MODEL, bool ANTI>
constexpr int adjoint_waves_target() {
    return (RNG == OFG_RNG_FAST && MODEL == OFG_MODEL_GBM && !ANTI) ? 3 : 1;
}

template <int RNG, int MODEL, bool ANTI, int BLOCK>
__global__ __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(
    adjoint_waves_target<RNG, MODEL, ANTI>()))) void adjoint_kernel(AdjArgs a) {
    const int lane = threadIdx.x & (WAVE - 1);
    __shared__ double tab[RNG == OFG_RNG_FAST ? OFG_FAST_TABLE_LEN : 1];
    if constexpr (RNG == OFG_RNG_FAST) {
        for (int i = threadIdx.x; i < OFG_FAST_TABLE_LEN; i += BLOCK) tab[i] = OFG_FAST_TABLE[i];
        __syncthreads();
    }
    const int64_t slot = (int64_t)blockIdx.x * (BLOCK / WAVE) + threadIdx.x / WAVE;
    double *ck = a.scratch + slot * a.wave_stride;
    const uint32_t span = (uint32_t)(a.order_end - a.order_begin);
    for (;;) {
        // every lane takes part in the atomic (see the pricing kernel)
        const uint32_t q = __builtin_amdgcn_readfirstlane(
            atomicAdd(a.queue_head, lane == 0 ? 1u : 0u));
        if (q >= span) break;
        const int32_t it = __builtin_amdgcn_readfirstlane(a.order[a.order_begin + (int32_t)q]);
        const WorkItem wi = a.items[it];
        const OptDev o = a.opts[wi.option];
        double spv = 0.0, spv2 = 0.0;
        Grad acc;
#pragma unroll
        for (int k = 0; k < ADJ_NG; ++k) acc.g[k] = 0.0;
        switch (o.kind) {
        case PK_VANILLA: adjoint_item<MODEL, PK_VANILLA, RNG, ANTI>(o, wi, lane, a, ck, tab, spv, spv2, acc); break;
        case PK_ASIAN: adjoint_item<MODEL, PK_ASIAN, RNG, ANTI>(o, wi, lane, a, ck, tab, spv, spv2, acc); break;
        case PK_ASIAN_LOG: adjoint_item<MODEL, PK_ASIAN_LOG, RNG, ANTI>(o, wi, lane, a, ck, tab, spv, spv2, acc); break;
        case PK_BARRIER: adjoint_item<MODEL, PK_BARRIER, RNG, ANTI>(o, wi, lane, a, ck, tab, spv, spv2, acc); break;
        default: adjoint_item<MODEL, PK_BARRIER_DOWN, RNG, ANTI>(o, wi, lane, a, ck, tab, spv, spv2, acc); break;
        }
        AdjPartial p;
        p.spv = wave_sum(spv);
        p.spv2 = wave_sum(spv2);
#pragma unroll
        for (int k = 0; k < ADJ_NG; ++k) p.g[k] = wave_sum(acc.g[k]);
        p.n = (double)(wi.end - wi.begin);
        if (lane == 0) a.partials[it] = p;
    }
}

__global__ __launch_bounds__(256) void adjoint_reduce_kernel(const OptDev *__restrict__ opts,
                                                             const int64_t *__restrict__ item_first,
                                                             const int32_t *__restrict__ item_count,
                                                             const AdjPartial *__restrict__ partials,
                                                             AdjOut *__restrict__ out, int32_t n_opts) {
    const int gw = (int)((blockIdx.x * blockDim.x + threadIdx.x) / WAVE);
    const int lane = threadIdx.x & (WAVE - 1);
    if (gw >= n_opts) return;
    const int64_t first = item_first[gw];
    const int32_t count = item_count[gw];
    double spv = 0, spv2 = 0, n = 0, g[ADJ_NG];
#pragma unroll
    for (int k = 0; k < ADJ_NG; ++k) g[k] = 0.0;
    for (int32_t i = lane; i < count; i += WAVE) {
        const AdjPartial &p = partials[first + i];
        spv += p.spv;
        spv2 += p.spv2;
        n += p.n;
#pragma unroll
        for (int k = 0; k < ADJ_NG; ++k) g[k] += p.g[k];
    }
    spv = wave_sum(spv);
    spv2 = wave_sum(spv2);
    n = wave_sum(n);
#pragma unroll
    for (int k = 0; k < ADJ_NG; ++k) g[k] = wave_sum(g[k]);
    if (lane == 0) {
        // mc_aad.rs:193-200: mean, sample variance (n - 1), stderr
        AdjOut r;
        r.price = spv / n;
        const double var = n > 1.0 ? (spv2 - spv * spv / n) / (n - 1.0) : 0.0;
        r.stderr_ = sqrt(var / n);
#pragma unroll
        for (int k = 0; k < ADJ_NG; ++k) r.g[k] = g[k] / n;
        r.n = n;
        out[gw] = r;
    }
}

template <int RNG, int MODEL, bool ANTI>
hipError_t launch_v(const AdjArgs &a, hipStream_t s) {
    switch (a.block) {
    case 64: adjoint_kernel<RNG, MODEL, ANTI, 64><<<a.grid, 64, 0, s>>>(a); break;
    case 128: adjoint_kernel<RNG, MODEL, ANTI, 128><<<a.grid, 128, 0, s>>>(a); break;
    default: adjoint_kernel<RNG, MODEL, ANTI, 256><<<a.grid, 256, 0, s>>>(a); break;
    }
    return hipGetLastError();
}

template <int RNG, int MODEL, bool ANTI>
const void *ptr_v(int block) {
    switch (block) {
    case 64: return (const void *)adjoint_kernel<RNG, MODEL, ANTI, 64>;
    case 128: return (const void *)adjoint_kernel<RNG, MODEL, ANTI, 128>;
    default: return (const void *)adjoint_kernel<RNG, MODEL, ANTI, 256>;
    }
}

template <int RNG>
hipError_t launch_r(const AdjArgs &a, hipStream_t s) {
    if (a.model == OFG_MODEL_GBM)
        return a.antithetic ? launch_v<RNG, OFG_MODEL_GBM, true>(a, s) : launch_v<RNG, OFG_MODEL_GBM, false>(a, s);
    return a.antithetic ? launch_v<RNG, OFG_MODEL_HESTON, true>(a, s) : launch_v<RNG, OFG_MODEL_HESTON, false>(a, s);
}

template <int RNG>
const void *ptr_r(int block, int model, int anti) {
    if (model == OFG_MODEL_GBM)
        return anti ? ptr_v<RNG, OFG_MODEL_GBM, true>(block) : ptr_v<RNG, OFG_MODEL_GBM, false>(block);
    return anti ? ptr_v<RNG, OFG_MODEL_HESTON, true>(block) : ptr_v<RNG, OFG_MODEL_HESTON, false>(block);
}

}  // namespace

hipError_t launch_adjoint(const AdjArgs &a, hipStream_t s) {
    switch (a.rng_mode) {
    case OFG_RNG_EXTERNAL: return launch_r<OFG_RNG_EXTERNAL>(a, s);
    case OFG_RNG_FAST: return launch_r<OFG_RNG_FAST>(a, s);
    default: return launch_r<OFG_RNG_REFERENCE>(a, s);
    }
}

hipError_t launch_adjoint_reduce(const OptDev *opts, const int64_t *item_first,
                                 const int32_t *item_count, const AdjPartial *partials,
                                 AdjOut *out, int32_t n_opts, hipStream_t s) {
    if (n_opts <= 0) return hipSuccess;
    adjoint_reduce_kernel<<<(n_opts + 3) / 4, 256, 0, s>>>(opts, item_first, item_count, partials,
                                                            out, n_opts);
    return hipGetLastError();
}

int adjoint_grid(int block, int rng_mode, int model, int antithetic, int device) {
    const void *k = rng_mode == OFG_RNG_EXTERNAL ? ptr_r<OFG_RNG_EXTERNAL>(block, model, antithetic)
                    : rng_mode == OFG_RNG_FAST   ? ptr_r<OFG_RNG_FAST>(block, model, antithetic)
                                                 : ptr_r<OFG_RNG_REFERENCE>(block, model, antithetic);
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k, block, 0) != hipSuccess || per_cu <= 0)
        per_cu = 1;
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess || cus <= 0)
        cus = 256;
    return per_cu * cus;
}

}  // namespace ofg
