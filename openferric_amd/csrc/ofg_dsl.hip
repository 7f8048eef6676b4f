// Structured-product Monte Carlo on MI355X (gfx950), FP64 on the VALU:
// correlated multi-asset paths (DslMonteCarloEngine::price_path_chunk_scalar,
// src/dsl/engine.rs:413-495) and the compiled product evaluated per path
// (evaluate_product_with_plan_in_place and execute_program, src/dsl/
// eval.rs:383-500, 1774-1980).
//
// One thread simulates one path at a time: per step it draws one normal per
// asset (the reference's single stream positioned by jump-ahead, caller
// normals [step][asset][path], or fast per-path xoroshiro128+ Box-Muller
// draws), correlates them with the scenario's lower Cholesky factor (running
// sum from 0 in column order, correlate_normals) and moves every asset with
// its stepper (GBM / Schwartz one-factor / Vasicek, engine.rs:95-122). The
// spots at snapshot steps are stored, then the product runs on them: the
// schedules in order, each observation with cleared locals, the schedule's
// register program under predication (ofg_dsl.hpp). Split form (the
// default): dsl_sim_kernel writes every path's snapshots, [snap][asset][path]
// per scenario, and dsl_eval_kernel runs the programs -- the simulation then
// carries none of the interpreter's registers. Fused form (when all
// snapshots would not fit the buffer budget): one kernel per tile with a
// reused [snap][asset][thread] buffer. Per tile (sum, sum of squares) of the
// path values and the first out-of-range price() index go to the tile's
// slot; one block per scenario sums the slots in tile order.
#include "ofg_dsl.hpp"
#include "ofg_internal.hpp"
#include "ofg_math.hpp"

namespace ofg {

namespace {

constexpr int WAVE = 64;
constexpr int NWAVES = DSL_BLOCK / WAVE;

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) v += __shfl_xor(v, m, WAVE);
    return v;
}

__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) v = fmax(v, __shfl_xor(v, m, WAVE));
    return v;
}

// A GBM step's factor exp(p0 + p1 z): the reference's exp on its own and
// external draws; the table-driven 2^(y/256) of the fast mode otherwise
// (fast draws are checked statistically, like every fast-mode kernel).
// With TABLE the reference's exponent p0 + p1 z goes through the same
// 2^(j/256) table (tab: its 256 entries), within about an ulp of exp.
template <bool FAST, bool TABLE = false>
__device__ __forceinline__ double gbm_factor(double p0, double p1, double z, const double *tab) {
    if constexpr (FAST) {
        double y = (p0 + p1 * z) * OFG_INV_LN2_256;
        y = fmin(fmax(y, -OFG_EXP2_256_YMAX), OFG_EXP2_256_YMAX);
        return exp2_256(y, tab);
    } else if constexpr (TABLE) {
        double y = (p0 + p1 * z) * OFG_INV_LN2_256;
        y = fmin(fmax(y, -OFG_EXP2_256_YMAX), OFG_EXP2_256_YMAX);
        return exp2_256t(y, tab);
    } else {
        return exp_fp64(p0 + p1 * z);
    }
}

// Fast mode of the GBM-only kernels: a spot is carried as its log growth
// Y = 256 log2(S / s0), advanced by gbm_factor's exponent, and formed as
// s0 2^(Y/256) only where a snapshot is stored (the same discretisation, one
// exponential per snapshot instead of one per step).
__device__ __forceinline__ double gbm_log_growth(double p0, double p1, double z) {
    return (p0 + p1 * z) * OFG_INV_LN2_256;
}

__device__ __forceinline__ double gbm_spot(double s0, double y, const double *tab) {
    return s0 * exp2_256(fmin(fmax(y, -OFG_EXP2_256_YMAX), OFG_EXP2_256_YMAX), tab);
}

// AssetStepper::step (engine.rs:95-122)
template <bool FAST, bool TABLE = false>
__device__ __forceinline__ double step_asset(int kind, double prev, double p0, double p1, double p2,
                                             double z, const double *tab) {
    if (kind == DSL_GBM) return prev * gbm_factor<FAST, TABLE>(p0, p1, z, tab);
    if (kind == DSL_SCHWARTZ) {
        const double log_prev = log(prev);
        const double log_next = p0 + p1 * (log_prev - p0) + p2 * z;
        return exp_fp64(log_next);
    }
    return p0 + p1 * (prev - p0) + p2 * z;
}

// `v as usize` of an f64 (saturating; NaN -> 0), compared with n; `shown`
// is that index as a double for the error message.
__device__ __forceinline__ bool index_in_range(double v, int n, int &idx, double &shown) {
    if (!(v >= 1.0)) {   // NaN, negatives and [0, 1) truncate to 0
        idx = 0;
        shown = 0.0;
        return n > 0;
    }
    if (v >= (double)n) {
        idx = 0;
        shown = v >= 1.8446744073709552e19 ? 1.8446744073709552e19 : trunc(v);
        return false;
    }
    idx = (int)v;
    shown = (double)idx;
    return true;
}

// One market's step parameters as the simulation reads them. Up to four
// assets they are copied out of the scenario once per kernel, before any
// snapshot store: the stores could alias the scenario as far as the compiler
// knows, so a read inside the step loop would be a vector reload with a
// memory wait every step, where the copies stay in scalar registers. Wider
// markets read the scenario in place (their copies would not fit).
template <int MAXA, bool COPY = (MAXA <= 4)>
struct StepParams;

template <int MAXA>
struct StepParams<MAXA, true> {
    int kind_[MAXA];
    double p0_[MAXA], p1_[MAXA], p2_[MAXA], ch_[MAXA * (MAXA + 1) / 2];
    __device__ explicit StepParams(const DslScen &o) {
#pragma unroll
        for (int i = 0; i < MAXA; ++i) {
            kind_[i] = o.kind[i];
            p0_[i] = o.p0[i];
            p1_[i] = o.p1[i];
            p2_[i] = o.p2[i];
        }
#pragma unroll
        for (int q = 0; q < MAXA * (MAXA + 1) / 2; ++q) ch_[q] = o.chol[q];
    }
    __device__ int kind(int i) const { return kind_[i]; }
    __device__ double p0(int i) const { return p0_[i]; }
    __device__ double p1(int i) const { return p1_[i]; }
    __device__ double p2(int i) const { return p2_[i]; }
    __device__ double chol(int q) const { return ch_[q]; }
};

template <int MAXA>
struct StepParams<MAXA, false> {
    const DslScen &o;
    __device__ explicit StepParams(const DslScen &scen) : o(scen) {}
    __device__ int kind(int i) const { return o.kind[i]; }
    __device__ double p0(int i) const { return o.p0[i]; }
    __device__ double p1(int i) const { return o.p1[i]; }
    __device__ double p2(int i) const { return o.p2[i]; }
    __device__ double chol(int q) const { return o.chol[q]; }
};

// MAXA: the asset count rounded up to 1, 2, 4, 8 or 16 (registers per path).
// Snapshot k of asset i goes to snap[(k n + i) stride].
// With WTAIL (reference draws) slots is this wave's 64 max(8, MAXA) doubles
// of LDS for the draws' tail values (acklam_block_wave) and tab the 256
// entries 2^(j/256) of the GBM factors; without (a kernel whose LDS holds
// the programs' register files) each lane inverts its own tail values and
// GBM factors take exp_fp64.
// Reference draws: the paths of one thread are consecutive in the engine's
// single stream (thread_stream_start), so the seed kernel positions one
// stream per thread and the thread draws on from path to path. g: the
// thread's index among its scenario's tiles.
template <int RNG>
__device__ __forceinline__ Xoshiro256pp thread_stream(const DslArgs &a, const DslScen &o,
                                                      int64_t g) {
    Xoshiro256pp ref;
    ref.s0 = ref.s1 = ref.s2 = ref.s3 = 0;
    if constexpr (RNG == OFG_RNG_REFERENCE) {
        const uint64_t *sd = a.seeds + o.seed_offset + 4 * g;
        ref.s0 = sd[0];
        ref.s1 = sd[1];
        ref.s2 = sd[2];
        ref.s3 = sd[3];
    }
    return ref;
}

// NA > 0: exactly NA assets (NA = MAXA), every one stepping as GBM, so the
// per-asset kind tests and asset-count guards are gone from the step.
// The snapshot steps are walked with a cursor over a.snap_list (a scalar
// load per snapshot, not per step).
// ref: the reference stream, positioned at this path's first draw and left
// after its last (the thread's next path follows it in the stream).
template <int RNG, int MAXA, bool WTAIL = true, int NA = 0>
__device__ __forceinline__ void simulate_path(const DslArgs &a, const DslScen &o,
                                              const StepParams<MAXA> &sp, int64_t p,
                                              double *snap, int64_t stride, const double *tab,
                                              double *slots, Xoshiro256pp &ref) {
    static_assert(NA == 0 || NA == MAXA, "exact asset count");
    // log growths instead of spots (gbm_log_growth)
    constexpr bool LOGS = NA > 0 && RNG == OFG_RNG_FAST;
    const int n = NA > 0 ? NA : o.n_assets;
    double s[MAXA];
#pragma unroll
    for (int i = 0; i < MAXA; ++i) s[i] = LOGS ? 0.0 : o.s0[i];
    const int32_t *sl = a.snap_list;
    int next_step = sl[0], next_k = sl[1];
    if (next_step == 0) {
#pragma unroll
        for (int i = 0; i < MAXA; ++i)
            if (i < n) snap[(int64_t)(next_k * n + i) * stride] = o.s0[i];
        sl += 2;
        next_step = sl[0];
        next_k = sl[1];
    }
    Xoroshiro128p fast;
    if constexpr (RNG == OFG_RNG_FAST) fast.seed(stream_seed(o.seed, (uint64_t)p));
    const double *zp = a.normals + o.normals_offset + p;
    // fast draws: Box-Muller pairs run across steps, so an odd asset count
    // wastes no normal
    double zpair = 0.0;
    bool have_pair = false;
    // one step from its normals (zr: the reference draws of the step)
    auto advance = [&](int step, const double *zr) {
        double c[MAXA];
#pragma unroll
        for (int i = 0; i < MAXA; ++i) c[i] = 0.0;
#pragma unroll
        for (int j = 0; j < MAXA; ++j) {
            if (j < n) {
                double z;
                if constexpr (RNG == OFG_RNG_REFERENCE) {
                    z = zr[j];
                } else if constexpr (RNG == OFG_RNG_EXTERNAL) {
                    z = zp[((int64_t)(step - 1) * n + j) * o.paths];
                } else {
                    if (!have_pair) {
                        const uint64_t wu = fast.next_u64();
                        const uint64_t wt = fast.next_u64();
                        box_muller_pair(wu, (uint32_t)(wt >> 32), z, zpair, tab);
                    } else {
                        z = zpair;
                    }
                    have_pair = !have_pair;
                }
#pragma unroll
                for (int i = j; i < MAXA; ++i)
                    if (i < n) c[i] += sp.chol(i * (i + 1) / 2 + j) * z;
            }
        }
#pragma unroll
        for (int i = 0; i < MAXA; ++i) {
            if (i < n) {
                if constexpr (LOGS)
                    s[i] += gbm_log_growth(sp.p0(i), sp.p1(i), c[i]);
                else if constexpr (NA > 0)
                    s[i] = s[i] * gbm_factor<RNG == OFG_RNG_FAST, WTAIL && RNG == OFG_RNG_REFERENCE>(
                                      sp.p0(i), sp.p1(i), c[i], tab);
                else
                    s[i] = step_asset<RNG == OFG_RNG_FAST, WTAIL && RNG == OFG_RNG_REFERENCE>(
                        sp.kind(i), s[i], sp.p0(i), sp.p1(i), sp.p2(i), c[i], tab);
            }
        }
        if (step == next_step) {
#pragma unroll
            for (int i = 0; i < MAXA; ++i)
                if (i < n)
                    snap[(int64_t)(next_k * n + i) * stride] =
                        LOGS ? gbm_spot(o.s0[i], s[i], tab) : s[i];
            sl += 2;
            next_step = sl[0];
            next_k = sl[1];
        }
    };
    if constexpr (RNG == OFG_RNG_REFERENCE && WTAIL) {
        // the draws of SPB steps per inverse-CDF block (in the reference's
        // order: step by step, asset by asset), tail values inverted per wave
        constexpr int SPB = MAXA <= 4 ? 8 / MAXA : 1;
        for (int step = 1; step <= a.steps; step += SPB) {
            double zr[SPB * MAXA];
            acklam_block_wave<SPB * MAXA>(
                [&](int j) {
                    return (j % MAXA < n && step + j / MAXA <= a.steps)
                               ? uniform_open01(ref.next_f64())
                               : 0.5;
                },
                zr, slots);
#pragma unroll
            for (int q = 0; q < SPB; ++q)
                if (step + q <= a.steps) advance(step + q, zr + q * MAXA);
        }
    } else if constexpr (RNG == OFG_RNG_REFERENCE) {
        for (int step = 1; step <= a.steps; ++step) {
            double pr[MAXA], zr[MAXA];   // each lane inverts its own tail values
#pragma unroll
            for (int j = 0; j < MAXA; ++j) pr[j] = j < n ? uniform_open01(ref.next_f64()) : 0.5;
            acklam_block<MAXA>(pr, zr);
            advance(step, zr);
        }
    } else {
        for (int step = 1; step <= a.steps; ++step) advance(step, nullptr);
    }
}

// The product's value on one simulated path; err receives an out-of-range
// price() index of an active lane. The register file lives in LDS, register
// i of thread t at rf[i * DSL_BLOCK] with rf = base + t (uniform register
// indices: conflict-free 8-byte rows); the if / else mask stack is two bit
// fields.
#define R(i) rf[(int)(i) * DSL_BLOCK]
__device__ double evaluate_path(const DslArgs &a, const DslScen &o, const double *snap,
                                int64_t stride, double *rf, double &err) {
    const int n = o.n_assets;
    uint32_t saved = 0, cond = 0;
    for (int k = 0; k < a.n_state; ++k) R(a.n_locals + k) = a.state_init[k];
    bool alive = true;
    double pv = 0.0;
    for (int si = 0; si < a.n_sched; ++si) {
        const DslSched sc = a.sched[si];
        for (int oi = 0; oi < sc.n_obs; ++oi) {
            if (!__any(alive)) break;   // every path of the wave has ended
            const DslObs ob = a.obs[sc.obs_first + oi];
            const double df = a.df[o.df_off + sc.obs_first + oi];
            const double *row = snap + (int64_t)ob.snap * n * stride;
            for (int k = 0; k < a.n_locals; ++k) R(k) = 0.0;
            bool active = alive;
            for (int pc = sc.code_off; pc < sc.code_off + sc.code_len; ++pc) {
                const DslIns in = a.code[pc];
                const int d = in.dst;
                switch (in.op) {
                case DOP_CONST: R(d) = a.consts[in.aux]; break;
                case DOP_NOTIONAL: R(d) = a.notional; break;
                case DOP_DATE: R(d) = ob.date; break;
                case DOP_ISFINAL: R(d) = ob.is_final ? 1.0 : 0.0; break;
                case DOP_SET:
                    if (in.b || active) R(d) = R(in.a);
                    break;
                case DOP_ADD: R(d) = R(in.a) + R(in.b); break;
                case DOP_SUB: R(d) = R(in.a) - R(in.b); break;
                case DOP_MUL: R(d) = R(in.a) * R(in.b); break;
                case DOP_DIV: {
                    const double den = R(in.b);
                    R(d) = den == 0.0 ? NAN : R(in.a) / den;
                    break;
                }
                case DOP_MIN: R(d) = fmin(R(in.a), R(in.b)); break;
                case DOP_MAX: R(d) = fmax(R(in.a), R(in.b)); break;
                case DOP_EQ: R(d) = fabs(R(in.a) - R(in.b)) < 2.220446049250313e-16 ? 1.0 : 0.0; break;
                case DOP_NE: R(d) = fabs(R(in.a) - R(in.b)) >= 2.220446049250313e-16 ? 1.0 : 0.0; break;
                case DOP_LT: R(d) = R(in.a) < R(in.b) ? 1.0 : 0.0; break;
                case DOP_LE: R(d) = R(in.a) <= R(in.b) ? 1.0 : 0.0; break;
                case DOP_GT: R(d) = R(in.a) > R(in.b) ? 1.0 : 0.0; break;
                case DOP_GE: R(d) = R(in.a) >= R(in.b) ? 1.0 : 0.0; break;
                case DOP_AND: R(d) = (R(in.a) != 0.0 && R(in.b) != 0.0) ? 1.0 : 0.0; break;
                case DOP_OR: R(d) = (R(in.a) != 0.0 || R(in.b) != 0.0) ? 1.0 : 0.0; break;
                case DOP_NEG: R(d) = -R(in.a); break;
                case DOP_ABS: R(d) = fabs(R(in.a)); break;
                case DOP_EXP: R(d) = exp(R(in.a)); break;
                case DOP_LOG: R(d) = log(R(in.a)); break;
                case DOP_NOT: R(d) = R(in.a) != 0.0 ? 0.0 : 1.0; break;
                case DOP_WOF_PERF:
                case DOP_BOF_PERF: {
                    const bool worst = in.op == DOP_WOF_PERF;
                    double v = worst ? INFINITY : -INFINITY;
                    for (int i = 0; i < n; ++i) {
                        if (o.s0[i] > 0.0) {
                            const double perf = row[i * stride] / o.s0[i];
                            if (worst ? perf < v : perf > v) v = perf;
                        }
                    }
                    R(d) = v;
                    break;
                }
                case DOP_WORST:
                case DOP_BEST: {
                    const bool worst = in.op == DOP_WORST;
                    double v = worst ? INFINITY : -INFINITY;
                    for (int k = (int)in.aux - 1; k >= 0; --k) {   // popped last first
                        const double x = R(in.a + k);
                        if (worst ? x < v : x > v) v = x;
                    }
                    R(d) = v;
                    break;
                }
                case DOP_PRICE: {
                    int idx;
                    double shown;
                    const bool ok = index_in_range(R(in.a), n, idx, shown);
                    if (!ok && active) err = fmax(err, shown);
                    R(d) = ok ? row[idx * stride] : 0.0;
                    break;
                }
                case DOP_PAY:
                    if (active) pv += R(in.a) * df;
                    break;
                case DOP_REDEEM:
                    if (active) {
                        pv += R(in.a) * df;
                        alive = false;
                        active = false;
                    }
                    break;
                case DOP_SKIP:
                    if (active) {
                        alive = false;
                        active = false;
                    }
                    break;
                case DOP_IF: {
                    const uint32_t bit = 1u << in.aux;
                    const bool c = R(in.a) != 0.0;
                    saved = active ? (saved | bit) : (saved & ~bit);
                    cond = c ? (cond | bit) : (cond & ~bit);
                    active = active && c;
                    break;
                }
                case DOP_ELSE:
                    active = ((saved >> in.aux) & 1u) && !((cond >> in.aux) & 1u) && alive;
                    break;
                case DOP_ENDIF: active = ((saved >> in.aux) & 1u) && alive; break;
                default: break;
                }
            }
        }
    }
    return pv;
}
#undef R

template <int RNG, int MAXA>
__global__ __launch_bounds__(DSL_BLOCK) void dsl_kernel(DslArgs a) {
    extern __shared__ double regs[];   // the programs' register files, [reg][thread]
    __shared__ double tab[RNG == OFG_RNG_FAST ? OFG_FAST_TABLE_LEN : 1];
    __shared__ double red[NWAVES][3];
    if constexpr (RNG == OFG_RNG_FAST) {
        for (int i = threadIdx.x; i < OFG_FAST_TABLE_LEN; i += DSL_BLOCK) tab[i] = OFG_FAST_TABLE[i];
        __syncthreads();
    }
    const int tile = blockIdx.x;
    const DslScen &o = a.scen[a.tile_scen[tile]];
    const StepParams<MAXA> sp(o);   // read once, before any snapshot store
    double *snap = a.snaps + (int64_t)tile * a.n_snaps * o.n_assets * DSL_BLOCK + threadIdx.x;
    const int64_t base = (int64_t)(tile - o.tile_first) * DSL_BLOCK * o.ppt;
    double sum = 0.0, sq = 0.0, err = -1.0;
    Xoshiro256pp ref =
        thread_stream<RNG>(a, o, (int64_t)(tile - o.tile_first) * DSL_BLOCK + threadIdx.x);
    for (int j = 0; j < o.ppt; ++j) {
        const int64_t p = base + (int64_t)j * DSL_BLOCK + threadIdx.x;
        if (p >= o.paths) break;
        simulate_path<RNG, MAXA, false>(a, o, sp, p, snap, DSL_BLOCK, tab, nullptr, ref);
        const double x = evaluate_path(a, o, snap, DSL_BLOCK, regs + threadIdx.x, err);
        sum += x;
        sq += x * x;
    }
    sum = wave_sum(sum);
    sq = wave_sum(sq);
    err = wave_max(err);
    const int lane = threadIdx.x & (WAVE - 1), wave = threadIdx.x / WAVE;
    if (lane == 0) {
        red[wave][0] = sum;
        red[wave][1] = sq;
        red[wave][2] = err;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double s = red[0][0], q = red[0][1], e = red[0][2];
        for (int w = 1; w < NWAVES; ++w) {
            s += red[w][0];
            q += red[w][1];
            e = fmax(e, red[w][2]);
        }
        a.partials[tile].sum = s;
        a.partials[tile].sum_sq = q;
        a.partials[tile].err_idx = e;
    }
}

// Split form, pass 1: the paths' snapshots, [snap][asset][path] per scenario.
template <int RNG, int MAXA, int NA = 0>
__global__ __launch_bounds__(DSL_BLOCK) void dsl_sim_kernel(DslArgs a) {
    // the fast mode's tables, or the reference draws' 2^(j/256) table and
    // each wave's slots for the tail values of an inverse-CDF block
    constexpr bool REF = RNG == OFG_RNG_REFERENCE;
    constexpr int AK = MAXA <= 4 ? 8 : MAXA;
    __shared__ double tab[RNG == OFG_RNG_FAST ? OFG_FAST_TABLE_LEN : REF ? 256 : 1];
    __shared__ double tail_slots[REF ? DSL_BLOCK * AK : 1];
    double *slots = tail_slots + (REF ? (threadIdx.x / WAVE) * WAVE * AK : 0);
    if constexpr (RNG == OFG_RNG_FAST) {
        for (int i = threadIdx.x; i < OFG_FAST_TABLE_LEN; i += DSL_BLOCK) tab[i] = OFG_FAST_TABLE[i];
        __syncthreads();
    } else if constexpr (REF) {
        for (int i = threadIdx.x; i < 256; i += DSL_BLOCK) tab[i] = OFG_FAST_TABLE[OFG_TAB_EXP2 + i];
        __syncthreads();
    }
    const int tile = blockIdx.x;
    const DslScen &o = a.scen[a.tile_scen[tile]];
    const StepParams<MAXA> sp(o);   // read once, before any snapshot store
    const int64_t base = (int64_t)(tile - o.tile_first) * DSL_BLOCK * o.ppt;
    Xoshiro256pp ref =
        thread_stream<RNG>(a, o, (int64_t)(tile - o.tile_first) * DSL_BLOCK + threadIdx.x);
    for (int j = 0; j < o.ppt; ++j) {
        const int64_t p = base + (int64_t)j * DSL_BLOCK + threadIdx.x;
        if (p >= o.paths) break;
        simulate_path<RNG, MAXA, true, NA>(a, o, sp, p, a.snaps + o.snap_off + p, o.paths, tab,
                                           slots, ref);
    }
}

// Split form, pass 1 for several markets at once (their Greeks bumps): one
// thread per path draws each step's normals once and moves every market's
// spots; a market's GBM asset whose step equals market 0's (share_mask)
// multiplies by market 0's factor -- the reference forms the same product --
// and a market with its own Cholesky factor correlates the same draws.
// Launched over market 0's tiles; writes every market's snapshots.
// Held to 3 waves per SIMD (168 VGPRs): left free, four-asset markets take
// about 195 and run at 2; the few registers spilled cost less than the lost
// wave (five-market Greeks 16.5 -> 15.7 ms on reference draws).
// NA > 0: every market has exactly NA = MAXA assets, all stepping as GBM.
template <int RNG, int MAXA, int NA = 0>
__global__ __launch_bounds__(DSL_BLOCK) __attribute__((amdgpu_waves_per_eu(3))) void
dsl_sim_markets_kernel(DslArgs a) {
    static_assert(NA == 0 || NA == MAXA, "exact asset count");
    // log growths instead of spots (gbm_log_growth), as simulate_path
    constexpr bool LOGS = NA > 0 && RNG == OFG_RNG_FAST;
    __shared__ double tab[RNG == OFG_RNG_FAST ? OFG_FAST_TABLE_LEN : 1];
    if constexpr (RNG == OFG_RNG_FAST) {
        for (int i = threadIdx.x; i < OFG_FAST_TABLE_LEN; i += DSL_BLOCK) tab[i] = OFG_FAST_TABLE[i];
        __syncthreads();
    }
    const DslScen &o = a.scen[0];
    const StepParams<MAXA> sp0(o);   // market 0's, read once, before any snapshot store
    const int n = NA > 0 ? NA : o.n_assets, ns = a.n_scen;
    uint32_t share[DSL_FUSE_MAXS];     // every market's share mask and own-factor flag, likewise
    bool own[DSL_FUSE_MAXS];
#pragma unroll
    for (int m = 0; m < DSL_FUSE_MAXS; ++m) {
        share[m] = m < ns ? a.scen[m].share_mask : 0u;
        own[m] = m < ns && a.scen[m].own_corr;
    }
    const int64_t base = (int64_t)blockIdx.x * DSL_BLOCK * o.ppt;
    Xoshiro256pp ref = thread_stream<RNG>(a, o, (int64_t)blockIdx.x * DSL_BLOCK + threadIdx.x);
    for (int j = 0; j < o.ppt; ++j) {
        const int64_t p = base + (int64_t)j * DSL_BLOCK + threadIdx.x;
        if (p >= o.paths) break;
        double s[DSL_FUSE_MAXS][MAXA];
#pragma unroll
        for (int m = 0; m < DSL_FUSE_MAXS; ++m)
#pragma unroll
            for (int i = 0; i < MAXA; ++i) s[m][i] = m < ns && !LOGS ? a.scen[m].s0[i] : 0.0;
        // snapshot steps by a cursor over a.snap_list
        const int32_t *sl = a.snap_list;
        int next_step = sl[0], next_k = sl[1];
        if (next_step == 0) {
#pragma unroll
            for (int m = 0; m < DSL_FUSE_MAXS; ++m)
                if (m < ns)
#pragma unroll
                    for (int i = 0; i < MAXA; ++i)
                        if (i < n)
                            a.snaps[a.scen[m].snap_off + (int64_t)(next_k * n + i) * o.paths + p] =
                                a.scen[m].s0[i];
            sl += 2;
            next_step = sl[0];
            next_k = sl[1];
        }
        Xoroshiro128p fast;
        if constexpr (RNG == OFG_RNG_FAST) fast.seed(stream_seed(o.seed, (uint64_t)p));
        const double *zp = a.normals + o.normals_offset + p;
        double zpair = 0.0;   // fast draws: pairs run across steps, as simulate_path
        bool have_pair = false;
        for (int step = 1; step <= a.steps; ++step) {
            double z[MAXA], c0[MAXA], g0[MAXA];
            // reference draws of the step, each lane inverting its own tail
            // values (this kernel's registers are spoken for; the wave form's
            // slot indices spill)
            double zr[MAXA];
            if constexpr (RNG == OFG_RNG_REFERENCE) {
                double pr[MAXA];
#pragma unroll
                for (int j2 = 0; j2 < MAXA; ++j2)
                    pr[j2] = j2 < n ? uniform_open01(ref.next_f64()) : 0.5;
                acklam_block<MAXA>(pr, zr);
            }
#pragma unroll
            for (int j2 = 0; j2 < MAXA; ++j2) {
                if (j2 < n) {
                    if constexpr (RNG == OFG_RNG_REFERENCE) {
                        z[j2] = zr[j2];
                    } else if constexpr (RNG == OFG_RNG_EXTERNAL) {
                        z[j2] = zp[((int64_t)(step - 1) * n + j2) * o.paths];
                    } else {
                        if (!have_pair) {
                            const uint64_t wu = fast.next_u64();
                            const uint64_t wt = fast.next_u64();
                            box_muller_pair(wu, (uint32_t)(wt >> 32), z[j2], zpair, tab);
                        } else {
                            z[j2] = zpair;
                        }
                        have_pair = !have_pair;
                    }
                } else {
                    z[j2] = 0.0;
                }
            }
            // market 0: correlated normals and GBM factors (log growths when
            // LOGS)
#pragma unroll
            for (int i = 0; i < MAXA; ++i) c0[i] = 0.0;
#pragma unroll
            for (int j2 = 0; j2 < MAXA; ++j2)
#pragma unroll
                for (int i = j2; i < MAXA; ++i)
                    if (i < n && j2 < n) c0[i] += sp0.chol(i * (i + 1) / 2 + j2) * z[j2];
#pragma unroll
            for (int i = 0; i < MAXA; ++i)
                g0[i] = !(i < n && (NA > 0 || sp0.kind(i) == DSL_GBM)) ? 1.0
                        : LOGS ? gbm_log_growth(sp0.p0(i), sp0.p1(i), c0[i])
                               : gbm_factor<RNG == OFG_RNG_FAST>(sp0.p0(i), sp0.p1(i), c0[i], tab);
            const bool snap_now = step == next_step;
            const int ks = next_k;
#pragma unroll
            for (int m = 0; m < DSL_FUSE_MAXS; ++m) {
                if (m >= ns) continue;
                const DslScen &q = a.scen[m];
                double c[MAXA];
                if (own[m]) {
#pragma unroll
                    for (int i = 0; i < MAXA; ++i) c[i] = 0.0;
#pragma unroll
                    for (int j2 = 0; j2 < MAXA; ++j2)
#pragma unroll
                        for (int i = j2; i < MAXA; ++i)
                            if (i < n && j2 < n) c[i] += q.chol[i * (i + 1) / 2 + j2] * z[j2];
                } else {
#pragma unroll
                    for (int i = 0; i < MAXA; ++i) c[i] = c0[i];
                }
#pragma unroll
                for (int i = 0; i < MAXA; ++i) {
                    if (i < n) {
                        if constexpr (LOGS)
                            s[m][i] += ((share[m] >> i) & 1u)
                                           ? g0[i]
                                           : gbm_log_growth(q.p0[i], q.p1[i], c[i]);
                        else if ((share[m] >> i) & 1u) s[m][i] = s[m][i] * g0[i];
                        else if constexpr (NA > 0)
                            s[m][i] = s[m][i] * gbm_factor<RNG == OFG_RNG_FAST>(q.p0[i], q.p1[i],
                                                                                c[i], tab);
                        else
                            s[m][i] = step_asset<RNG == OFG_RNG_FAST>(q.kind[i], s[m][i], q.p0[i],
                                                                      q.p1[i], q.p2[i], c[i], tab);
                    }
                }
                if (snap_now) {
#pragma unroll
                    for (int i = 0; i < MAXA; ++i)
                        if (i < n)
                            a.snaps[q.snap_off + (int64_t)(ks * n + i) * o.paths + p] =
                                LOGS ? gbm_spot(q.s0[i], s[m][i], tab) : s[m][i];
                }
            }
            if (snap_now) {
                sl += 2;
                next_step = sl[0];
                next_k = sl[1];
            }
        }
    }
}

// Split form, pass 2: the product on every path; tile sums as the fused form.
__global__ __launch_bounds__(DSL_BLOCK) void dsl_eval_kernel(DslArgs a) {
    extern __shared__ double regs[];   // the programs' register files, [reg][thread]
    __shared__ double red[NWAVES][3];
    const int tile = blockIdx.x;
    const DslScen &o = a.scen[a.tile_scen[tile]];
    const int64_t base = (int64_t)(tile - o.tile_first) * DSL_BLOCK * o.ppt;
    double sum = 0.0, sq = 0.0, err = -1.0;
    for (int j = 0; j < o.ppt; ++j) {
        const int64_t p = base + (int64_t)j * DSL_BLOCK + threadIdx.x;
        if (p >= o.paths) break;
        const double x = evaluate_path(a, o, a.snaps + o.snap_off + p, o.paths, regs + threadIdx.x, err);
        sum += x;
        sq += x * x;
    }
    sum = wave_sum(sum);
    sq = wave_sum(sq);
    err = wave_max(err);
    const int lane = threadIdx.x & (WAVE - 1), wave = threadIdx.x / WAVE;
    if (lane == 0) {
        red[wave][0] = sum;
        red[wave][1] = sq;
        red[wave][2] = err;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double s = red[0][0], q = red[0][1], e = red[0][2];
        for (int w = 1; w < NWAVES; ++w) {
            s += red[w][0];
            q += red[w][1];
            e = fmax(e, red[w][2]);
        }
        a.partials[tile].sum = s;
        a.partials[tile].sum_sq = q;
        a.partials[tile].err_idx = e;
    }
}

// One block per scenario: the tile slots in tile order.
__global__ __launch_bounds__(DSL_BLOCK) void dsl_close_kernel(DslArgs a) {
    __shared__ double red[NWAVES][3];
    const DslScen &o = a.scen[blockIdx.x];
    double s = 0.0, q = 0.0, e = -1.0;
    for (int t = threadIdx.x; t < o.n_tiles; t += DSL_BLOCK) {
        const DslPartial pt = a.partials[o.tile_first + t];
        s += pt.sum;
        q += pt.sum_sq;
        e = fmax(e, pt.err_idx);
    }
    s = wave_sum(s);
    q = wave_sum(q);
    e = wave_max(e);
    const int lane = threadIdx.x & (WAVE - 1), wave = threadIdx.x / WAVE;
    if (lane == 0) {
        red[wave][0] = s;
        red[wave][1] = q;
        red[wave][2] = e;
    }
    __syncthreads();
    if (threadIdx.x != 0) return;
    double sum = red[0][0], sq = red[0][1], err = red[0][2];
    for (int w = 1; w < NWAVES; ++w) {
        sum += red[w][0];
        sq += red[w][1];
        err = fmax(err, red[w][2]);
    }
    a.out[blockIdx.x].sum = sum;
    a.out[blockIdx.x].sum_sq = sq;
    a.out[blockIdx.x].err_idx = err;
}

}  // namespace

size_t lds_bytes(const DslArgs &a) {
    return (size_t)(a.n_regs > 0 ? a.n_regs : 1) * DSL_BLOCK * sizeof(double);
}

// The split form's simulation (one market, or several in one pass) of
// markets whose assets all step as GBM, at their exact asset count (up to
// four).
template <int RNG>
void launch_sim_gbm(const DslArgs &a, hipStream_t stream) {
    if (a.fuse_markets) {
        switch (a.n_assets) {
        case 1: dsl_sim_markets_kernel<RNG, 1, 1><<<a.fuse_tiles, DSL_BLOCK, 0, stream>>>(a); break;
        case 2: dsl_sim_markets_kernel<RNG, 2, 2><<<a.fuse_tiles, DSL_BLOCK, 0, stream>>>(a); break;
        case 3: dsl_sim_markets_kernel<RNG, 3, 3><<<a.fuse_tiles, DSL_BLOCK, 0, stream>>>(a); break;
        default: dsl_sim_markets_kernel<RNG, 4, 4><<<a.fuse_tiles, DSL_BLOCK, 0, stream>>>(a); break;
        }
        return;
    }
    switch (a.n_assets) {
    case 1: dsl_sim_kernel<RNG, 1, 1><<<a.n_tiles, DSL_BLOCK, 0, stream>>>(a); break;
    case 2: dsl_sim_kernel<RNG, 2, 2><<<a.n_tiles, DSL_BLOCK, 0, stream>>>(a); break;
    case 3: dsl_sim_kernel<RNG, 3, 3><<<a.n_tiles, DSL_BLOCK, 0, stream>>>(a); break;
    default: dsl_sim_kernel<RNG, 4, 4><<<a.n_tiles, DSL_BLOCK, 0, stream>>>(a); break;
    }
}

template <int RNG, int MAXA>
hipError_t launch_rng(const DslArgs &a, hipStream_t stream, hipFunction_t jit_eval,
                      const DslJitArgs *jit_args) {
    if (a.split) {
        if constexpr (MAXA <= DSL_FUSE_MAXA) {
            if (a.gbm_only && a.n_assets >= 1 && a.n_assets <= 4)
                launch_sim_gbm<RNG>(a, stream);
            else if (a.fuse_markets)
                dsl_sim_markets_kernel<RNG, MAXA><<<a.fuse_tiles, DSL_BLOCK, 0, stream>>>(a);
            else
                dsl_sim_kernel<RNG, MAXA><<<a.n_tiles, DSL_BLOCK, 0, stream>>>(a);
        } else {
            dsl_sim_kernel<RNG, MAXA><<<a.n_tiles, DSL_BLOCK, 0, stream>>>(a);
        }
        if (jit_eval) {
            DslJitArgs ja = *jit_args;
            void *params[] = {&ja};
            return hipModuleLaunchKernel(jit_eval, (unsigned)a.n_tiles, 1, 1, DSL_BLOCK, 1, 1, 0,
                                         stream, params, nullptr);
        }
        dsl_eval_kernel<<<a.n_tiles, DSL_BLOCK, lds_bytes(a), stream>>>(a);
    } else {
        dsl_kernel<RNG, MAXA><<<a.n_tiles, DSL_BLOCK, lds_bytes(a), stream>>>(a);
    }
    return hipSuccess;
}

template <int MAXA>
hipError_t launch_tiles(const DslArgs &a, hipStream_t stream, hipFunction_t jit_eval,
                        const DslJitArgs *jit_args) {
    switch (a.rng_mode) {
    case OFG_RNG_FAST: return launch_rng<OFG_RNG_FAST, MAXA>(a, stream, jit_eval, jit_args);
    case OFG_RNG_EXTERNAL: return launch_rng<OFG_RNG_EXTERNAL, MAXA>(a, stream, jit_eval, jit_args);
    default: return launch_rng<OFG_RNG_REFERENCE, MAXA>(a, stream, jit_eval, jit_args);
    }
}

hipError_t launch_dsl(const DslArgs &a, hipStream_t stream, hipFunction_t jit_eval,
                      const DslJitArgs *jit_args) {
    if (a.n_tiles <= 0) return hipSuccess;
    if (jit_eval && (!jit_args || !a.split)) return hipErrorInvalidValue;
    hipError_t le;
    if (a.n_assets <= 1) le = launch_tiles<1>(a, stream, jit_eval, jit_args);
    else if (a.n_assets <= 2) le = launch_tiles<2>(a, stream, jit_eval, jit_args);
    else if (a.n_assets <= 4) le = launch_tiles<4>(a, stream, jit_eval, jit_args);
    else if (a.n_assets <= 8) le = launch_tiles<8>(a, stream, jit_eval, jit_args);
    else le = launch_tiles<DSL_MAX_ASSETS>(a, stream, jit_eval, jit_args);
    if (le != hipSuccess) return le;
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    dsl_close_kernel<<<a.n_scen, DSL_BLOCK, 0, stream>>>(a);
    return hipGetLastError();
}

}  // namespace ofg
