// Least-squares Monte Carlo (Longstaff-Schwartz) kernels for MI355X (gfx950),
// FP64 on the VALU.
//
// Reference semantics: src/engines/lsm/longstaff_schwartz.rs -- path
// simulation (:195-283, :506-556), backward induction with the quadratic
// regression over in-the-money paths (:291-417, :558-660), the 3x3 normal
// equations solved by nalgebra's LU with partial pivoting (zeros when
// singular), mean_and_stderr (:128-143) and the engine's plain Monte Carlo
// barrier pricer (:674-735).
//
// Work decomposition (ofg_lsm.hpp): a tile is one block of 256 threads over
// 256 * ppt consecutive paths of one option, so every launch parameter is
// block-uniform and the per-path loads and stores of the [date][path] spot
// matrix and the value vector are coalesced. Launch 0 simulates the paths and
// stores the spot at each exercise date; launch r >= 1 is one backward round.
// A launch's reductions are deterministic: each tile writes its partial sums
// to its own slot, and a close kernel (one block per option) sums the slots in
// tile order and closes the round (solve, exercise statistics, price). The
// kernel boundary orders the slots' writes before their reads; no fences or
// atomics across workgroups are needed.
#include "ofg_lsm.hpp"
#include "ofg_math.hpp"
#include "ofg_paths.hpp"

#include <type_traits>

namespace ofg {

namespace {

constexpr int WAVE = 64;
constexpr int NWAVES = LSM_BLOCK / WAVE;
constexpr int LSM_AK = 8;   // reference draws per inverse-CDF block

__device__ __forceinline__ double lsm_intrinsic(bool put, double s, double k) {
    return put ? fmax(k - s, 0.0) : fmax(s - k, 0.0);
}

// Per-thread accumulators of one launch.
struct LsmAcc {
    double v[LSM_NACC];   // reg[0..7], ex_n, ex_ext, term_itm, fsum, fsq
};
enum { A_EXN = 8, A_EXT = 9, A_TERM = 10, A_FSUM = 11, A_FSQ = 12 };

__device__ __forceinline__ void acc_init(LsmAcc &a, bool put) {
#pragma unroll
    for (int i = 0; i < LSM_NACC; ++i) a.v[i] = 0.0;
    a.v[A_EXT] = put ? -INFINITY : INFINITY;
}

// Regression sums at one date (longstaff_schwartz.rs:598-609).
__device__ __forceinline__ void acc_regress(LsmAcc &a, bool put, double s, double strike,
                                            double y) {
    if (lsm_intrinsic(put, s, strike) > 0.0) {
        const double s2 = s * s;
        a.v[0] += 1.0;
        a.v[1] += s;
        a.v[2] += s2;
        a.v[3] += s2 * s;
        a.v[4] += s2 * s2;
        a.v[5] += y;
        a.v[6] += s * y;
        a.v[7] += s2 * y;
    }
}

__device__ __forceinline__ double combine(int i, bool put, double x, double y) {
    if (i == A_EXT) return put ? fmax(x, y) : fmin(x, y);
    return x + y;
}

// nalgebra 0.34 LU::new + LU::solve on a 3x3 system (partial pivoting by the
// first largest |entry|, multipliers scaled by the pivot's reciprocal,
// unit-lower forward and upper back substitution by columns). Returns false
// when U has an exactly zero pivot.
__device__ bool lu3_solve(double a[3][3], double b[3], double x[3]) {
    int from[3], to[3], np = 0;
    for (int i = 0; i < 3; ++i) {
        int piv = i;
        double best = fabs(a[i][i]);
        for (int r = i + 1; r < 3; ++r) {
            const double v = fabs(a[r][i]);
            if (v > best) {
                best = v;
                piv = r;
            }
        }
        const double diag = a[piv][i];
        if (diag == 0.0) continue;
        if (piv != i) {
            from[np] = i;
            to[np] = piv;
            ++np;
            for (int c = 0; c < 3; ++c) {
                const double t = a[i][c];
                a[i][c] = a[piv][c];
                a[piv][c] = t;
            }
        }
        const double inv = 1.0 / diag;
        for (int r = i + 1; r < 3; ++r) a[r][i] = a[r][i] * inv;
        for (int c = i + 1; c < 3; ++c)
            for (int r = i + 1; r < 3; ++r) a[r][c] = (-a[i][c]) * a[r][i] + a[r][c];
    }
    for (int k = 0; k < np; ++k) {
        const double t = b[from[k]];
        b[from[k]] = b[to[k]];
        b[to[k]] = t;
    }
    for (int i = 0; i < 2; ++i) {
        const double coeff = b[i];
        for (int r = i + 1; r < 3; ++r) b[r] = (-coeff) * a[r][i] + b[r];
    }
    for (int i = 2; i >= 0; --i) {
        const double diag = a[i][i];
        if (diag == 0.0) return false;
        const double coeff = b[i] / diag;
        b[i] = coeff;
        for (int r = 0; r < i; ++r) b[r] = (-coeff) * a[r][i] + b[r];
    }
    x[0] = b[0];
    x[1] = b[1];
    x[2] = b[2];
    return true;
}

// Block reduction of the accumulators into `out` (valid in thread 0): a
// fixed butterfly per wave, then the waves in order.
__device__ void block_reduce(LsmAcc &acc, bool put, double (*lds)[LSM_NACC], double *out) {
    const int lane = threadIdx.x & (WAVE - 1);
    const int wave = threadIdx.x / WAVE;
#pragma unroll
    for (int i = 0; i < LSM_NACC; ++i) {
        double v = acc.v[i];
#pragma unroll
        for (int m = 32; m >= 1; m >>= 1) v = combine(i, put, v, __shfl_xor(v, m, WAVE));
        if (lane == 0) lds[wave][i] = v;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int i = 0; i < LSM_NACC; ++i) {
            double v = lds[0][i];
            for (int w = 1; w < NWAVES; ++w) v = combine(i, put, v, lds[w][i]);
            out[i] = v;
        }
    }
    __syncthreads();
}

// What the current launch computes for one option.
struct RoundPlan {
    int k_update;   // date whose exercise decision is applied (-1: none)
    int k_regress;  // date whose regression sums are accumulated (-1: none)
    bool final_;    // values reach time 0: price and stderr
};

__device__ __forceinline__ RoundPlan round_plan(const LsmDev &o, int launch) {
    RoundPlan p;
    if (o.barrier) {
        p.k_update = -1;
        p.k_regress = -1;
        p.final_ = true;
        return p;
    }
    if (launch == 0) {
        p.k_update = -1;
        p.k_regress = o.n_ex - 1;
        p.final_ = o.n_ex == 0;
    } else {
        p.k_update = o.n_ex - launch;
        p.k_regress = p.k_update - 1;
        p.final_ = p.k_update == 0;
    }
    return p;
}

// Tile epilogue: the block's partial sums to the tile's own slot.
__device__ void finish_tile(const LsmArgs &a, const LsmDev &o, int tile, LsmAcc &acc) {
    __shared__ double lds[NWAVES][LSM_NACC];
    __shared__ double tot[LSM_NACC];
    block_reduce(acc, o.put != 0, lds, tot);
    if (threadIdx.x == 0) {
        LsmPartial &p = a.partials[tile];
#pragma unroll
        for (int i = 0; i < 8; ++i) p.reg[i] = tot[i];
        p.ex_n = tot[A_EXN];
        p.ex_ext = tot[A_EXT];
        p.term_itm = tot[A_TERM];
        p.fsum = tot[A_FSUM];
        p.fsq = tot[A_FSQ];
    }
}

// Closes one launch for one option (one block per option, after the tile
// kernel): sums the option's tile slots in tile order, then solves the
// regression of the date just accumulated, records the exercise statistics of
// the date just updated, and at time 0 the price and standard error.
__global__ __launch_bounds__(LSM_BLOCK) void lsm_close_kernel(LsmArgs a) {
    __shared__ double lds[NWAVES][LSM_NACC];
    __shared__ double tot[LSM_NACC];
    const int opt = blockIdx.x;
    const LsmDev o = a.opts[opt];
    if (o.barrier ? a.launch > 0 : a.launch > o.n_ex) return;
    const RoundPlan rp = round_plan(o, a.launch);
    const bool put = o.put != 0;
    LsmAcc s;
    acc_init(s, put);
    for (int t = threadIdx.x; t < o.n_tiles; t += LSM_BLOCK) {
        const LsmPartial &p = a.partials[o.tile_first + t];
#pragma unroll
        for (int i = 0; i < 8; ++i) s.v[i] += p.reg[i];
        s.v[A_EXN] += p.ex_n;
        s.v[A_EXT] = combine(A_EXT, put, s.v[A_EXT], p.ex_ext);
        s.v[A_TERM] += p.term_itm;
        s.v[A_FSUM] += p.fsum;
        s.v[A_FSQ] += p.fsq;
    }
    block_reduce(s, put, lds, tot);
    if (threadIdx.x != 0) return;
    if (rp.k_regress >= 0) {
        LsmDateOut &d = a.date_out[o.date_offset + rp.k_regress];
        const double itm = tot[0];
        d.itm = itm;
        d.valid = itm >= 3.0;
        d.beta[0] = d.beta[1] = d.beta[2] = 0.0;
        d.exercised = 0.0;
        d.extreme = 0.0;
        if (d.valid) {
            double m[3][3] = {{tot[0], tot[1], tot[2]}, {tot[1], tot[2], tot[3]},
                              {tot[2], tot[3], tot[4]}};
            double rhs[3] = {tot[5], tot[6], tot[7]};
            double beta[3];
            if (lu3_solve(m, rhs, beta)) {
                d.beta[0] = beta[0];
                d.beta[1] = beta[1];
                d.beta[2] = beta[2];
            }
        }
    }
    if (rp.k_update >= 0) {
        LsmDateOut &d = a.date_out[o.date_offset + rp.k_update];
        d.exercised = tot[A_EXN];
        d.extreme = tot[A_EXT];
    }
    LsmOptOut &out = a.opt_out[opt];
    if (a.launch == 0) out.term_itm = tot[A_TERM];
    if (rp.final_) {
        const double n = (double)o.paths;
        const double sum = tot[A_FSUM], sum_sq = tot[A_FSQ];
        const double mean = sum / n;
        const double var = o.paths > 1 ? (sum_sq - sum * sum / n) / (n - 1.0) : 0.0;
        out.price = mean;
        out.stderr_ = sqrt(var / n);
    }
}

// ------------------------------------------------------------------------
// Path simulation (launch 0)
// ------------------------------------------------------------------------

// Drives one path's draws: step(z1, z2) once per path step. Reference draws
// walk the path's slice of the engine's single stream (z1 then z2 per Heston
// step) from the thread's stream state `ref`, 8 at a time with the tail values inverted per wave (slots: the
// wave's LDS for them; a tile is one option, so every lane has the same step
// count); external normals read [step][path], the fast mode runs Box-Muller
// on a per-path xoroshiro128+ stream.
template <int RNG, bool TWO, class Step>
__device__ __forceinline__ void drive(const LsmDev &o, const LsmArgs &a, int64_t p,
                                      const double *tab, double *slots, Xoshiro256pp &ref,
                                      Step &&step) {
    const int n = o.m.steps;
    if constexpr (RNG == OFG_RNG_REFERENCE) {
        // ref: this thread's stream, at the path's first draw (the thread's
        // paths follow one another in the option's stream)
        reference_steps_wave<LSM_AK, TWO ? 2 : 1>(
            ref, n, [&](int j, double z1, double z2) { step(z1, z2, j + 1); }, slots);
    } else if constexpr (RNG == OFG_RNG_EXTERNAL) {
        const double *e1 = a.normals1 + o.normals_offset + p;
        const double *e2 = TWO ? a.normals2 + o.normals_offset + p : nullptr;
        for (int t = 1; t <= n; ++t) {
            const int64_t off = (int64_t)(t - 1) * o.paths;
            step(e1[off], TWO ? e2[off] : 0.0, t);
        }
    } else {
        Xoroshiro128p rng;
        rng.seed(stream_seed(o.m.seed, (uint64_t)p));
        int t = 1;
        if constexpr (TWO) {
            for (; t + 1 <= n; t += 2) {
                double z0, z1, z2, z3;
                normals4(rng, z0, z1, z2, z3, tab);
                step(z0, z1, t);
                step(z2, z3, t + 1);
            }
            if (t <= n) {
                double z0, z1, z2, z3;
                normals4(rng, z0, z1, z2, z3, tab);
                step(z0, z1, t);
            }
        } else {
            for (; t + 3 <= n; t += 4) {
                double z0, z1, z2, z3;
                normals4(rng, z0, z1, z2, z3, tab);
                step(z0, 0.0, t);
                step(z1, 0.0, t + 1);
                step(z2, 0.0, t + 2);
                step(z3, 0.0, t + 3);
            }
            if (t <= n) {
                double z0, z1, z2, z3;
                normals4(rng, z0, z1, z2, z3, tab);
                step(z0, 0.0, t);
                if (t + 1 <= n) step(z1, 0.0, t + 1);
                if (t + 2 <= n) step(z2, 0.0, t + 2);
            }
        }
    }
}

// One exercise-date path: simulate, store the spot at each date, then the
// terminal cash flow discounted to the latest date.
template <int RNG, int DYN>
__device__ __forceinline__ void lsm_path(const LsmDev &o, const LsmArgs &a, int64_t p,
                                         const double *tab, double *slots, Xoshiro256pp &ref,
                                         const RoundPlan &rp, LsmAcc &acc) {
    constexpr bool FAST = RNG == OFG_RNG_FAST;
    const LsmDate *dates = a.dates + o.date_offset;
    double *sp = a.spots + o.s_offset + p;
    int k = 0;
    int next = o.n_ex > 0 ? dates[0].step : 0x7fffffff;
    double s = o.m.s0, v = o.m.v0, s_last = o.m.s0;
    auto record = [&](int t) {
        if (t == next) {
            sp[(int64_t)k * o.paths] = s;
            s_last = s;
            ++k;
            next = k < o.n_ex ? dates[k].step : 0x7fffffff;
        }
    };
    // reference draws: growth factors by the 2^(j/256) table (tab), as the
    // pricing kernel's reference mode
    constexpr bool TABLE = RNG == OFG_RNG_REFERENCE;
    if constexpr (DYN == LSM_HESTON) {
        drive<RNG, true>(o, a, p, tab, slots, ref, [&](double z1, double z2, int t) {
            double sn, vn, growth, e;
            heston_euler<!FAST, TABLE>(s, v, o.m, z1, z2, sn, vn, growth, e, tab);
            s = sn;   // fmax(s_next, 1e-12) (:272)
            v = vn;
            record(t);
        });
    } else if constexpr (DYN == LSM_LOCAL_VOL) {
        // flat surface: sigma(S, t) = vol; s *= exp(drift + sigma sqrt(dt) z),
        // floored at 1e-12 (:231-240)
        drive<RNG, false>(o, a, p, tab, slots, ref, [&](double z, double, int t) {
            double g;
            if constexpr (FAST) {
                g = exp2_256(o.m.drift2 + o.m.diffusion2 * z, tab);
            } else if constexpr (TABLE) {
                const double x = o.m.drift + o.m.diffusion * z;
                g = exp2_256t(fmin(fmax(x * OFG_INV_LN2_256, -OFG_EXP2_256_YMAX),
                                   OFG_EXP2_256_YMAX), tab);
            } else {
                g = exp_fp64(o.m.drift + o.m.diffusion * z);
            }
            s = fmax(s * g, 1.0e-12);
            record(t);
        });
    } else {
        drive<RNG, false>(o, a, p, tab, slots, ref, [&](double z, double, int t) {
            double x;
            s *= gbm_growth<FAST, TABLE>(o.m, z, tab, x);
            record(t);
        });
    }
    const bool put = o.put != 0;
    double val = lsm_intrinsic(put, s, o.terminal_strike);
    acc.v[A_TERM] += val > 0.0 ? 1.0 : 0.0;
    for (int g = 0; g < o.term_gap; ++g) val *= o.disc;
    if (rp.final_) {
        acc.v[A_FSUM] += val;
        acc.v[A_FSQ] += val * val;
    } else {
        acc_regress(acc, put, s_last, dates[o.n_ex - 1].strike, val);
        a.values[o.v_offset + p] = val;
    }
}

// Barrier instrument: plain Monte Carlo with inclusive monitoring of every
// path point (:704-729).
template <int RNG>
__device__ __forceinline__ void barrier_path(const LsmDev &o, const LsmArgs &a, int64_t p,
                                             const double *tab, double *slots,
                                             Xoshiro256pp &ref, LsmAcc &acc) {
    constexpr bool FAST = RNG == OFG_RNG_FAST;
    const double level = o.barrier_level;
    const bool down = o.barrier_down != 0;
    double s = o.m.s0;
    bool hit = down ? s <= level : s >= level;
    drive<RNG, false>(o, a, p, tab, slots, ref, [&](double z, double, int) {
        double x;
        s *= gbm_growth<FAST, RNG == OFG_RNG_REFERENCE>(o.m, z, tab, x);
        hit = hit || (down ? s <= level : s >= level);
    });
    const bool active = o.knock_in ? hit : !hit;
    const double payoff = active ? lsm_intrinsic(o.put != 0, s, o.m.strike) : o.rebate;
    const double pv = o.df_total * payoff;
    acc.v[A_FSUM] += pv;
    acc.v[A_FSQ] += pv * pv;
}

template <int RNG>
__global__ __launch_bounds__(LSM_BLOCK) void lsm_paths_kernel(LsmArgs a) {
    constexpr bool REF = RNG == OFG_RNG_REFERENCE;
    // the fast mode's tables, or the reference draws' 2^(j/256) table and
    // each wave's slots for the tail values of an inverse-CDF block
    __shared__ double tab[RNG == OFG_RNG_FAST ? OFG_FAST_TABLE_LEN : REF ? 256 : 1];
    __shared__ double tail_slots[REF ? LSM_BLOCK * LSM_AK : 1];
    double *slots = tail_slots + (REF ? (threadIdx.x / WAVE) * WAVE * LSM_AK : 0);
    if constexpr (RNG == OFG_RNG_FAST) {
        for (int i = threadIdx.x; i < OFG_FAST_TABLE_LEN; i += LSM_BLOCK) tab[i] = OFG_FAST_TABLE[i];
        __syncthreads();
    } else if constexpr (REF) {
        for (int i = threadIdx.x; i < 256; i += LSM_BLOCK) tab[i] = OFG_FAST_TABLE[OFG_TAB_EXP2 + i];
        __syncthreads();
    }
    const int tile = blockIdx.x;
    const int opt = a.tile_opt[tile];
    const LsmDev o = a.opts[opt];
    const RoundPlan rp = round_plan(o, 0);
    LsmAcc acc;
    acc_init(acc, o.put != 0);
    const int64_t base = (int64_t)(tile - o.tile_first) * LSM_BLOCK * o.ppt;
    // reference draws: the thread's paths are consecutive in the option's
    // stream (thread_stream_start); its state comes from the seed kernel
    Xoshiro256pp ref;
    ref.s0 = ref.s1 = ref.s2 = ref.s3 = 0;
    if constexpr (REF) {
        const uint64_t *sd =
            a.seeds + o.seed_offset + 4 * ((int64_t)(tile - o.tile_first) * LSM_BLOCK + threadIdx.x);
        ref.s0 = sd[0];
        ref.s1 = sd[1];
        ref.s2 = sd[2];
        ref.s3 = sd[3];
    }
    for (int j = 0; j < o.ppt; ++j) {
        const int64_t p = base + (int64_t)j * LSM_BLOCK + threadIdx.x;
        if (p >= o.paths) break;
        if (o.barrier) barrier_path<RNG>(o, a, p, tab, slots, ref, acc);
        else if (o.dynamics == LSM_HESTON)
            lsm_path<RNG, LSM_HESTON>(o, a, p, tab, slots, ref, rp, acc);
        else if (o.dynamics == LSM_LOCAL_VOL)
            lsm_path<RNG, LSM_LOCAL_VOL>(o, a, p, tab, slots, ref, rp, acc);
        else lsm_path<RNG, LSM_GBM>(o, a, p, tab, slots, ref, rp, acc);
    }
    finish_tile(a, o, tile, acc);
}

// ------------------------------------------------------------------------
// Backward round r >= 1
// ------------------------------------------------------------------------

__global__ __launch_bounds__(LSM_BLOCK) void lsm_round_kernel(LsmArgs a) {
    const int tile = blockIdx.x;
    const int opt = a.tile_opt[tile];
    const LsmDev o = a.opts[opt];
    if (o.barrier || a.launch > o.n_ex) return;   // this option is already priced
    const RoundPlan rp = round_plan(o, a.launch);
    const bool put = o.put != 0;
    const LsmDate du = a.dates[o.date_offset + rp.k_update];
    const LsmDateOut dout = a.date_out[o.date_offset + rp.k_update];
    const double strike_r = rp.k_regress >= 0 ? a.dates[o.date_offset + rp.k_regress].strike : 0.0;
    const double b0 = dout.beta[0], b1 = dout.beta[1], b2 = dout.beta[2];
    const double *s_u = a.spots + o.s_offset + (int64_t)rp.k_update * o.paths;
    const double *s_r = a.spots + o.s_offset + (int64_t)(rp.k_regress >= 0 ? rp.k_regress : 0) * o.paths;
    double *vals = a.values + o.v_offset;
    LsmAcc acc;
    acc_init(acc, put);
    const int64_t base = (int64_t)(tile - o.tile_first) * LSM_BLOCK * o.ppt;
    // RU paths per thread per pass, all loads issued before any use
    constexpr int RU = 4;
    for (int j0 = 0; j0 < o.ppt; j0 += RU) {
        double v[RU], su[RU], sr[RU];
        bool ok[RU];
#pragma unroll
        for (int u = 0; u < RU; ++u) {
            const int64_t p = base + (int64_t)(j0 + u) * LSM_BLOCK + threadIdx.x;
            ok[u] = j0 + u < o.ppt && p < o.paths;
            v[u] = su[u] = sr[u] = 0.0;
            if (ok[u]) {
                v[u] = vals[p];
                su[u] = s_u[p];
                if (!rp.final_) sr[u] = s_r[p];
            }
        }
#pragma unroll
        for (int u = 0; u < RU; ++u) {
            if (!ok[u]) continue;
            double val = v[u];
            if (dout.valid) {
                // exercise decision (:618-626): continuation from the regression
                const double s = su[u];
                const double ex = lsm_intrinsic(put, s, du.strike);
                if (ex > 0.0) {
                    const double cont = b0 + b1 * s + b2 * s * s;
                    if (ex > cont) {
                        val = ex;
                        acc.v[A_EXN] += 1.0;
                        acc.v[A_EXT] = put ? fmax(acc.v[A_EXT], s) : fmin(acc.v[A_EXT], s);
                    }
                }
            }
            for (int g = 0; g < du.gap; ++g) val *= o.disc;
            if (rp.final_) {
                acc.v[A_FSUM] += val;
                acc.v[A_FSQ] += val * val;
            } else {
                acc_regress(acc, put, sr[u], strike_r, val);
                vals[base + (int64_t)(j0 + u) * LSM_BLOCK + threadIdx.x] = val;
            }
        }
    }
    finish_tile(a, o, tile, acc);
}

// ------------------------------------------------------------------------
// Reference-stream start states
// ------------------------------------------------------------------------

__global__ __launch_bounds__(256) void lsm_seed_kernel(const LsmSeedJob *__restrict__ jobs,
                                                       int32_t n_jobs, int64_t total,
                                                       const uint64_t *__restrict__ polys,
                                                       uint64_t *__restrict__ seeds) {
    const int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (gid >= total) return;
    int lo = 0, hi = n_jobs - 1;   // last job with thread_first <= gid
    while (lo < hi) {
        const int mid = (lo + hi + 1) / 2;
        if (jobs[mid].thread_first <= gid) lo = mid;
        else hi = mid - 1;
    }
    const LsmSeedJob &j = jobs[lo];
    const int64_t g = gid - j.thread_first;
    Xoshiro256pp st;
    st.s0 = j.base[0];
    st.s1 = j.base[1];
    st.s2 = j.base[2];
    st.s3 = j.base[3];
    if (j.tile_ppt > 0) {
        // one state per thread of the consuming kernel, at its first path:
        // a jump to the first of `chunk` consecutive consumer threads, then
        // their paths' draws walked (their streams follow one another)
        const int64_t per_tile = 256 * (int64_t)j.tile_ppt;
        const int64_t n_cons = (j.paths + per_tile - 1) / per_tile * 256;
        const int64_t g0 = g * j.chunk;
        if (g0 >= n_cons) return;
        int64_t count;
        const int64_t start = thread_stream_start(g0, j.paths, j.tile_ppt, &count);
        if (count <= 0) return;
        for (int k = 0; k < j.n_polys; ++k)
            if ((start >> k) & 1) xoshiro_apply_poly(polys + 4 * (j.poly_offset + k), st);
        const int64_t g_end = g0 + j.chunk < n_cons ? g0 + j.chunk : n_cons;
        for (int64_t gg = g0; gg < g_end; ++gg) {
            (void)thread_stream_start(gg, j.paths, j.tile_ppt, &count);
            if (count <= 0) break;   // the rest of the last tile has no paths
            uint64_t *out = seeds + j.seed_offset + 4 * gg;
            out[0] = st.s0;
            out[1] = st.s1;
            out[2] = st.s2;
            out[3] = st.s3;
            if (gg + 1 < g_end)
                for (int64_t d = 0; d < count * j.draws; ++d) (void)st.next_u64();
        }
        return;
    }
    const int64_t first = g * j.chunk;
    if (first >= j.paths) return;
    for (int k = 0; k < j.n_polys; ++k)
        if ((g >> k) & 1) xoshiro_apply_poly(polys + 4 * (j.poly_offset + k), st);
    uint64_t *out = seeds + j.seed_offset;
    for (int q = 0; q < j.chunk; ++q) {
        const int64_t p = first + q;
        if (p >= j.paths) break;
        // one path's state is one contiguous 32-byte sector
        out[4 * p] = st.s0;
        out[4 * p + 1] = st.s1;
        out[4 * p + 2] = st.s2;
        out[4 * p + 3] = st.s3;
        for (int64_t d = 0; d < j.draws; ++d) (void)st.next_u64();
    }
}


}  // namespace

hipError_t launch_lsm_seeds(const LsmSeedJob *jobs, int32_t n_jobs, int64_t total_threads,
                            const uint64_t *polys, uint64_t *seeds, hipStream_t stream) {
    if (total_threads <= 0 || n_jobs <= 0) return hipSuccess;
    const int64_t grid = (total_threads + 255) / 256;
    lsm_seed_kernel<<<(unsigned)grid, 256, 0, stream>>>(jobs, n_jobs, total_threads, polys, seeds);
    return hipGetLastError();
}

hipError_t launch_lsm(const LsmArgs &a, hipStream_t stream) {
    if (a.n_tiles <= 0) return hipSuccess;
    if (a.launch == 0) {
        switch (a.rng_mode) {
        case OFG_RNG_EXTERNAL:
            lsm_paths_kernel<OFG_RNG_EXTERNAL><<<a.n_tiles, LSM_BLOCK, 0, stream>>>(a);
            break;
        case OFG_RNG_FAST:
            lsm_paths_kernel<OFG_RNG_FAST><<<a.n_tiles, LSM_BLOCK, 0, stream>>>(a);
            break;
        default:
            lsm_paths_kernel<OFG_RNG_REFERENCE><<<a.n_tiles, LSM_BLOCK, 0, stream>>>(a);
            break;
        }
    } else {
        lsm_round_kernel<<<a.n_tiles, LSM_BLOCK, 0, stream>>>(a);
    }
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    lsm_close_kernel<<<a.n_opts, LSM_BLOCK, 0, stream>>>(a);
    return hipGetLastError();
}

}  // namespace ofg
