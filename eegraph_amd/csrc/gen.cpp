// Seeded synthetic graph generators (host side, multi-threaded, counter-based).
// The reference reads SNAP downloads only (datasets/*/Makefile); these generators produce the
// RMAT and grid workloads the measurements are quoted on.
#include <algorithm>
#include <cmath>
#include <thread>
#include <vector>

#include "eeg_internal.hpp"

namespace {

template <class F>
void parallel_for(uint64_t n, int threads, F&& body) {
    if (threads <= 0) {
        unsigned hc = std::thread::hardware_concurrency();
        threads = hc ? (int)std::min(hc, 16u) : 1;
    }
    if (n < 65536 || threads == 1) {
        body(0, n);
        return;
    }
    std::vector<std::thread> pool;
    const uint64_t chunk = (n + threads - 1) / threads;
    for (int t = 0; t < threads; ++t) {
        const uint64_t lo = chunk * t, hi = std::min(n, lo + chunk);
        if (lo >= hi) break;
        pool.emplace_back([=, &body] { body(lo, hi); });
    }
    for (auto& th : pool) th.join();
}

bool bad_weights(uint32_t wmin, uint32_t wmax) { return wmin > wmax; }

}  // namespace

extern "C" eeg_status eeg_gen_rmat(uint32_t scale, uint32_t edge_factor, double a, double b,
                                   double c, uint64_t seed, uint32_t wmin, uint32_t wmax,
                                   uint32_t* src, uint32_t* dst, uint32_t* w, int threads) {
    eeg::clear_error();
    if (scale == 0 || scale > 31 || edge_factor == 0 || !src || !dst)
        return eeg::fail(EEG_ERR_INVALID_ARG, "rmat: scale must be 1..31, edge_factor > 0");
    if (a < 0 || b < 0 || c < 0 || a + b + c > 1.0 + 1e-12)
        return eeg::fail(EEG_ERR_INVALID_ARG, "rmat: probabilities must be >= 0 with a+b+c <= 1");
    if (w && bad_weights(wmin, wmax)) return eeg::fail(EEG_ERR_INVALID_ARG, "rmat: wmin > wmax");
    const uint64_t m = (uint64_t)edge_factor << scale;
    const auto th = eeg::rmat_thresholds(a, b, c);
    const uint64_t key = eeg::stream_key(seed, eeg::STREAM_RMAT);
    const uint64_t wkey = eeg::stream_key(seed, eeg::STREAM_WEIGHT);
    parallel_for(m, threads, [&](uint64_t lo, uint64_t hi) {
        for (uint64_t i = lo; i < hi; ++i) {
            eeg::rmat_edge(key, scale, th.ta, th.tab, th.tabc, i, &src[i], &dst[i]);
            if (w) w[i] = eeg::edge_weight(wkey, i, wmin, wmax);
        }
    });
    return EEG_OK;
}

// Grid layout of the edge list (all pairs are emitted as (u->v, v->u)):
//   horizontal pairs k = y*(W-1) + x, x < W-1: edges 2k, 2k+1 between y*W+x and y*W+x+1
//   vertical pairs   k = y*W + x, y < H-1: edges base_v + 2k, +1 between y*W+x and (y+1)*W+x
//   shortcut pairs   s < S: edges base_s + 2s, +1 between u_s and v_s, where
//                    u_s = draw 2s, v_s = draw 2s+1 of STREAM_SHORTCUT, mapped onto [0, W*H)
//   S = llround(frac * (H*(W-1) + (H-1)*W)); weights: draw i of STREAM_WEIGHT per edge i.
namespace {
struct GridShape {
    uint64_t nh, nv, ns, m;
};
bool grid_shape(uint32_t W, uint32_t H, double frac, GridShape* g) {
    if (W == 0 || H == 0 || (uint64_t)W * H > 0xffffffffULL || frac < 0 || frac > 1) return false;
    g->nh = (uint64_t)H * (W - 1);
    g->nv = (uint64_t)(H - 1) * W;
    g->ns = (uint64_t)std::llround(frac * (double)(g->nh + g->nv));
    g->m = 2 * (g->nh + g->nv + g->ns);
    return true;
}
}  // namespace

extern "C" eeg_status eeg_gen_grid_size(uint32_t width, uint32_t height, double shortcut_frac,
                                        uint64_t* m) {
    eeg::clear_error();
    GridShape g;
    if (!m || !grid_shape(width, height, shortcut_frac, &g))
        return eeg::fail(EEG_ERR_INVALID_ARG, "grid: bad dimensions or shortcut fraction");
    *m = g.m;
    return EEG_OK;
}

extern "C" eeg_status eeg_gen_grid(uint32_t width, uint32_t height, double shortcut_frac,
                                   uint64_t seed, uint32_t wmin, uint32_t wmax, uint32_t* src,
                                   uint32_t* dst, uint32_t* w, int threads) {
    eeg::clear_error();
    GridShape g;
    if (!src || !dst || !grid_shape(width, height, shortcut_frac, &g))
        return eeg::fail(EEG_ERR_INVALID_ARG, "grid: bad dimensions or shortcut fraction");
    if (w && bad_weights(wmin, wmax)) return eeg::fail(EEG_ERR_INVALID_ARG, "grid: wmin > wmax");
    const uint64_t W = width, N = (uint64_t)width * height;
    const uint64_t skey = eeg::stream_key(seed, eeg::STREAM_SHORTCUT);
    const uint64_t npairs = g.nh + g.nv + g.ns;
    parallel_for(npairs, threads, [&](uint64_t lo, uint64_t hi) {
        for (uint64_t k = lo; k < hi; ++k) {
            uint32_t u, v;
            if (k < g.nh) {
                const uint64_t y = k / (W - 1), x = k % (W - 1);
                u = (uint32_t)(y * W + x);
                v = u + 1;
            } else if (k < g.nh + g.nv) {
                const uint64_t kk = k - g.nh;
                u = (uint32_t)kk;
                v = (uint32_t)(kk + W);
            } else {
                const uint64_t s = k - g.nh - g.nv;
                u = eeg::draw_below(eeg::stream_draw(skey, 2 * s), N);
                v = eeg::draw_below(eeg::stream_draw(skey, 2 * s + 1), N);
            }
            src[2 * k] = u;
            dst[2 * k] = v;
            src[2 * k + 1] = v;
            dst[2 * k + 1] = u;
        }
    });
    if (w) {
        const uint64_t wkey = eeg::stream_key(seed, eeg::STREAM_WEIGHT);
        parallel_for(g.m, threads, [&](uint64_t lo, uint64_t hi) {
            for (uint64_t i = lo; i < hi; ++i) w[i] = eeg::edge_weight(wkey, i, wmin, wmax);
        });
    }
    return EEG_OK;
}
