// Jump-ahead of the xoshiro256++ stream on the host: the state transition T
// is linear over GF(2), so T^k = r(T) with r(x) = x^k mod P(x), P the
// characteristic polynomial of T (found once by Berlekamp-Massey on the
// sequence of one state bit). Used to split an engine's single stream across
// paths (ofg_lsm_api.cpp, ofg_basket_api.cpp) and by ofg_stream_jump.
#pragma once

#include "ofg_host.hpp"
#include "ofg_lsm.hpp"
#include "ofg_math.hpp"

#include <algorithm>
#include <cstdint>
#include <cstring>
#include <map>
#include <vector>

namespace ofg_jump {

using ofg_host::fail;

// ---------------------------------------------------------------------------
// GF(2) polynomials of degree < 256 (bit i = coefficient of x^i)
// ---------------------------------------------------------------------------
struct Poly {
    uint64_t w[4] = {0, 0, 0, 0};
    bool bit(int i) const { return (w[i >> 6] >> (i & 63)) & 1u; }
};

// Low part of the characteristic polynomial P(x) = x^256 + low(x) of the
// xoshiro256 state transition.
inline const Poly &xoshiro_charpoly() {
    static const Poly P = [] {
        // the sequence of bit 0 of s0 satisfies the minimal recurrence of T
        ofg::Xoshiro256pp st;
        st.seed(0x2545F4914F6CDD1Dull);
        const int N = 512;
        std::vector<uint8_t> s(N);
        for (int j = 0; j < N; ++j) {
            s[j] = (uint8_t)(st.s0 & 1u);
            (void)st.next_u64();
        }
        // Berlekamp-Massey over GF(2): connection polynomial C, length L
        std::vector<uint8_t> C(N + 1, 0), B(N + 1, 0), T;
        C[0] = B[0] = 1;
        int L = 0, m = 1;
        for (int n = 0; n < N; ++n) {
            uint8_t d = s[n];
            for (int i = 1; i <= L; ++i) d ^= (uint8_t)(C[i] & s[n - i]);
            if (!d) {
                ++m;
                continue;
            }
            T = C;
            for (int i = 0; i + m <= N; ++i) C[i + m] ^= B[i];
            if (2 * L <= n) {
                L = n + 1 - L;
                B = T;
                m = 1;
            } else {
                ++m;
            }
        }
        if (L != 256) fail(OFG_NUMERICAL_ERROR, "xoshiro256 recurrence has unexpected degree");
        // P(x) = x^L C(1/x): coefficient of x^(L - i) is C[i]
        Poly p;
        for (int i = 1; i <= L; ++i)
            if (C[i]) p.w[(L - i) >> 6] |= 1ull << ((L - i) & 63);
        return p;
    }();
    return P;
}

inline Poly mulmod(const Poly &a, const Poly &b) {
    uint64_t prod[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (int i = 0; i < 256; ++i) {
        if (!a.bit(i)) continue;
        const int ws = i >> 6, bs = i & 63;
        for (int k = 0; k < 4; ++k) {
            prod[k + ws] ^= b.w[k] << bs;
            if (bs) prod[k + ws + 1] ^= b.w[k] >> (64 - bs);
        }
    }
    const Poly &low = xoshiro_charpoly();
    for (int j = 511; j >= 256; --j) {
        if (!((prod[j >> 6] >> (j & 63)) & 1u)) continue;
        prod[j >> 6] ^= 1ull << (j & 63);   // x^j = x^(j-256) x^256 = x^(j-256) low(x)
        const int sh = j - 256, ws = sh >> 6, bs = sh & 63;
        for (int k = 0; k < 4; ++k) {
            prod[k + ws] ^= low.w[k] << bs;
            if (bs && k + ws + 1 < 8) prod[k + ws + 1] ^= low.w[k] >> (64 - bs);
        }
    }
    Poly r;
    for (int k = 0; k < 4; ++k) r.w[k] = prod[k];
    return r;
}

// x^k mod P(x)
inline Poly x_pow(uint64_t k) {
    Poly result;
    result.w[0] = 1;
    Poly base;
    base.w[0] = 2;   // x
    while (k) {
        if (k & 1u) result = mulmod(result, base);
        base = mulmod(base, base);
        k >>= 1;
    }
    return result;
}

inline void apply_poly(const Poly &r, uint64_t s[4]) {
    ofg::Xoshiro256pp cur;
    cur.s0 = s[0];
    cur.s1 = s[1];
    cur.s2 = s[2];
    cur.s3 = s[3];
    uint64_t acc[4] = {0, 0, 0, 0};
    for (int i = 0; i < 256; ++i) {
        if (r.bit(i)) {
            acc[0] ^= cur.s0;
            acc[1] ^= cur.s1;
            acc[2] ^= cur.s2;
            acc[3] ^= cur.s3;
        }
        (void)cur.next_u64();
    }
    std::memcpy(s, acc, sizeof acc);
}

// Seed job of the seed kernel for one engine stream: path p starts at draw
// p * draws of the stream seeded by seed_from_u64(seed). A thread walks
// `chunk` paths (64, fewer for long per-path streams) after one jump of
// g * chunk * draws draws, composed from the
// polynomials x^(chunk * draws * 2^k) (cached per distance, appended to
// polys).
inline ofg::LsmSeedJob make_seed_job(uint64_t seed, int64_t paths, int64_t draws,
                                     int64_t seed_offset, int64_t thread_first,
                                     std::vector<uint64_t> &polys,
                                     std::map<uint64_t, std::vector<Poly>> &cache) {
    ofg::LsmSeedJob j{};
    ofg::Xoshiro256pp st;
    st.seed(seed);   // Xoshiro256PlusPlus::seed_from_u64 (fast_rng.rs:35-48)
    j.base[0] = st.s0;
    j.base[1] = st.s1;
    j.base[2] = st.s2;
    j.base[3] = st.s3;
    j.paths = paths;
    j.draws = draws;
    // paths per thread: long enough to amortise the jump, short enough that
    // long per-path streams still give many threads
    j.chunk = (int32_t)std::max<int64_t>(1, std::min<int64_t>(64, 8192 / std::max<int64_t>(1, draws)));
    const int64_t threads = (paths + j.chunk - 1) / j.chunk;
    int32_t bits = 0;
    while (bits < 63 && (int64_t(1) << bits) < threads) ++bits;
    j.n_polys = bits;
    j.seed_offset = seed_offset;
    j.thread_first = thread_first;
    const uint64_t key = (uint64_t)j.chunk * (uint64_t)j.draws;
    std::vector<Poly> &qs = cache[key];
    if ((int32_t)qs.size() < bits) {
        if (qs.empty()) qs.push_back(x_pow(key));
        while ((int32_t)qs.size() < bits) qs.push_back(mulmod(qs.back(), qs.back()));
    }
    j.poly_offset = (int32_t)(polys.size() / 4);
    for (int32_t k = 0; k < bits; ++k) polys.insert(polys.end(), qs[(size_t)k].w, qs[(size_t)k].w + 4);
    return j;
}

inline int64_t seed_job_threads(const ofg::LsmSeedJob &j) {
    if (j.tile_ppt > 0) {   // every `chunk` threads of the consuming kernel's tiles
        const int64_t per_tile = 256 * (int64_t)j.tile_ppt;
        const int64_t n_cons = (j.paths + per_tile - 1) / per_tile * 256;
        return (n_cons + j.chunk - 1) / j.chunk;
    }
    return (j.paths + j.chunk - 1) / j.chunk;
}

// Seed job for a kernel whose threads each simulate consecutive paths of the
// stream (ofg::thread_stream_start): one state per consumer thread. A seed
// thread jumps to the first of `chunk` consecutive consumer threads (start
// * draws draws, composed from the polynomials x^(draws 2^k)) and walks the
// draws of their paths to the next ones' starts; chunk shrinks to 1 when a
// consumer thread's draws outnumber a jump's cost.
inline ofg::LsmSeedJob make_thread_seed_job(uint64_t seed, int64_t paths, int64_t draws,
                                            int32_t tile_ppt, int64_t seed_offset,
                                            int64_t thread_first, std::vector<uint64_t> &polys,
                                            std::map<uint64_t, std::vector<Poly>> &cache) {
    ofg::LsmSeedJob j{};
    ofg::Xoshiro256pp st;
    st.seed(seed);   // Xoshiro256PlusPlus::seed_from_u64 (fast_rng.rs:35-48)
    j.base[0] = st.s0;
    j.base[1] = st.s1;
    j.base[2] = st.s2;
    j.base[3] = st.s3;
    j.paths = paths;
    j.draws = draws;
    j.chunk = (int32_t)std::max<int64_t>(
        1, std::min<int64_t>(64, 8192 / std::max<int64_t>(1, (int64_t)tile_ppt * draws)));
    j.tile_ppt = tile_ppt;
    int32_t bits = 0;   // start < paths
    while (bits < 63 && (int64_t(1) << bits) < paths) ++bits;
    j.n_polys = bits;
    j.seed_offset = seed_offset;
    j.thread_first = thread_first;
    const uint64_t key = (uint64_t)draws;
    std::vector<Poly> &qs = cache[key];
    if ((int32_t)qs.size() < bits) {
        if (qs.empty()) qs.push_back(x_pow(key));
        while ((int32_t)qs.size() < bits) qs.push_back(mulmod(qs.back(), qs.back()));
    }
    j.poly_offset = (int32_t)(polys.size() / 4);
    for (int32_t k = 0; k < bits; ++k) polys.insert(polys.end(), qs[(size_t)k].w, qs[(size_t)k].w + 4);
    return j;
}

}  // namespace ofg_jump
