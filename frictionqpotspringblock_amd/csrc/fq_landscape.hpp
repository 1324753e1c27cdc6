// Device-side chunked yield-position landscape: the role of prrng's
// pcg32_tensor_cumsum in the reference (type Line1d.h:51-52, construction
// Line1d.h:148-157, used through align(u) in every potential force, e.g.
// detail.h:144).
//
// Each particle owns one row of `nchunk` yield positions in HBM. The row is a
// window [start, start + nchunk) of that particle's infinite sequence
//   y_g = cumsum(increments)_g + offset          for the first window,
//   y_g = y_{g-1} + d_g                          beyond it (forward refill),
//   y_{g-1} = y_g - d_g                          before it (backward refill),
// where d_g is drawn from the particle's own pcg32 stream. Only the thread
// that owns a particle touches its row, so refills are done inline by that
// thread without synchronisation; they are rare (a particle has to drift
// close to the edge of its window).
#pragma once

#include "fq_common.hpp"

namespace fq {

constexpr int64_t MAX_SEARCH_WELLS = int64_t(1) << 20;

// Move particle p's window so that it starts at global index ns.
__device__ __noinline__ void chunk_shift_to(const Landscape& L, int64_t p, int64_t ns)
{
    const int64_t n = L.nchunk;
    double* y = L.y + p * n;
    const int64_t s = L.start[p];
    const int64_t d = ns - s;
    if (d == 0) {
        return;
    }
    if (d > 0) {
        double val = y[n - 1];
        uint64_t st = L.st_end[p];
        if (d < n) {
            for (int64_t k = 0; k < n - d; ++k) {
                y[k] = y[k + d];
            }
        }
        for (int64_t g = s + n; g < ns + n; ++g) {
            val = val + draw_increment(L, st);
            if (g >= ns) {
                y[g - ns] = val;
            }
        }
        L.st_end[p] = st;
        L.st_begin[p] = pcg_advance(L.st_begin[p], L.inc, d);
    }
    else {
        const int64_t e = -d;
        double val = y[0];
        uint64_t st = L.st_begin[p];
        if (e < n) {
            for (int64_t k = n - 1; k >= e; --k) {
                y[k] = y[k - e];
            }
        }
        for (int64_t g = s; g > ns; --g) {
            st = pcg_step_back(st, L.inc);
            uint64_t tmp = st;
            double dg = draw_increment(L, tmp);
            val = val - dg;
            if (g - 1 - ns < n) {
                y[g - 1 - ns] = val;
            }
        }
        L.st_begin[p] = pcg_advance(L.st_begin[p], L.inc, d);
        L.st_end[p] = pcg_advance(L.st_end[p], L.inc, d);
    }
    L.start[p] = ns;
}

// Slow path of the alignment of particle p to slip u (u outside its current
// well, or the well too close to the window edge). Returns the new local
// index; afterwards y[j] <= u < y[j+1] and buffer <= j <= n - 2 - buffer.
__device__ __forceinline__ int32_t chunk_align_slow_inl(const Landscape& L, int64_t p, double u, int32_t j0)
{
    const int64_t n = L.nchunk;
    const double* y = L.y + p * n;
    int64_t j = j0;
    if (!__builtin_isfinite(u)) {
        return j0;
    }
    const int64_t b = L.buffer, m = L.margin;
    // A slip further than MAX_SEARCH_WELLS wells from the window is not
    // searched (bounded work per call); the particle is left at the window edge.
    int64_t budget = MAX_SEARCH_WELLS;
    while (!(u < y[n - 1])) {
        if (budget <= 0) {
            L.jloc[p] = (int32_t)(n - 2);
            return (int32_t)(n - 2);
        }
        budget -= n - 1 - m;
        chunk_shift_to(L, p, L.start[p] + (n - 1 - m));
        j = m;
    }
    while (u < y[0]) {
        if (budget <= 0) {
            L.jloc[p] = 0;
            return 0;
        }
        budget -= n - 1 - m;
        chunk_shift_to(L, p, L.start[p] - (n - 1 - m));
        j = n - 1 - m;
    }
    if (j < 0) {
        j = 0;
    }
    if (j > n - 2) {
        j = n - 2;
    }
    // bounded walk: a row that is not increasing (which only corrupt data can
    // produce) never leads outside the window
    while (j < n - 2 && u >= y[j + 1]) {
        ++j;
    }
    while (j > 0 && u < y[j]) {
        --j;
    }
    if (j < b || j > n - 2 - b) {
        chunk_shift_to(L, p, L.start[p] + j - m);
        j = m;
    }
    L.jloc[p] = (int32_t)j;
    return (int32_t)j;
}

__device__ __noinline__ int32_t chunk_align_slow(const Landscape& L, int64_t p, double u, int32_t j0)
{
    return chunk_align_slow_inl(L, p, u, j0);
}

// Wavefront-cooperative form of chunk_shift_to for a window that starts at
// `s` with pcg32 states (stb, ste), called by all 64 lanes with the same
// arguments: lane l draws the increments l, l + 64, ... in parallel (each from
// the stream state advanced to it), and the running sum is then taken in the
// sequential order, so the row is bit-identical to chunk_shift_to's. Writes
// the row and updates stb, ste; the caller stores start and the states.
__device__ __forceinline__ void chunk_shift_wave(const Landscape& L, int64_t p, int64_t s, int64_t ns, uint64_t& stb,
                                                 uint64_t& ste)
{
    const int lane = threadIdx.x & 63;
    const int64_t n = L.nchunk;
    double* y = L.y + p * n;
    const int64_t d = ns - s;
    if (d == 0) {
        return;
    }
    if (d > 0) {
        double val = y[n - 1];
        if (d < n) {
            // y[k] = y[k + d]; in each block of 64 the loads precede the stores,
            // and later blocks read above everything written before
            for (int64_t c = 0; c < n - d; c += 64) {
                const int64_t k = c + lane;
                const double x = k < n - d ? y[k + d] : 0.0;
                if (k < n - d) {
                    y[k] = x;
                }
            }
        }
        for (int64_t c = 0; c < d; c += 64) {
            const int64_t k = c + lane; // increment of global index s + n + k
            double dk = 0.0;
            if (k < d) {
                uint64_t st = pcg_advance(ste, L.inc, k);
                dk = draw_increment(L, st);
            }
            const int nc = (int)(d - c < 64 ? d - c : 64);
            double mine = 0.0;
            for (int q = 0; q < nc; ++q) {
                val = val + __shfl(dk, q, 64);
                if (q == lane) {
                    mine = val;
                }
            }
            const int64_t g = s + n + k;
            if (k < d && g >= ns) {
                y[g - ns] = mine;
            }
        }
    }
    else {
        const int64_t e = -d;
        double val = y[0];
        if (e < n) {
            // y[k] = y[k - e] from the top down, mirrored
            for (int64_t c = n - 1; c >= e; c -= 64) {
                const int64_t k = c - lane;
                const double x = k >= e ? y[k - e] : 0.0;
                if (k >= e) {
                    y[k] = x;
                }
            }
        }
        for (int64_t c = 0; c < e; c += 64) {
            const int64_t k = c + lane; // the draw of global index s - k, from s - k - 1
            double dk = 0.0;
            if (k < e) {
                uint64_t st = pcg_advance(stb, L.inc, -(k + 1));
                dk = draw_increment(L, st);
            }
            const int nc = (int)(e - c < 64 ? e - c : 64);
            double mine = 0.0;
            for (int q = 0; q < nc; ++q) {
                val = val - __shfl(dk, q, 64);
                if (q == lane) {
                    mine = val;
                }
            }
            const int64_t o = e - k - 1; // local index of global s - k - 1
            if (k < e && o < n) {
                y[o] = mine;
            }
        }
    }
    stb = pcg_advance(stb, L.inc, d);
    ste = pcg_advance(ste, L.inc, d);
}

// Wavefront-cooperative chunk_align_slow: all 64 lanes call it with the same
// particle p, slip u and local index j0 and follow the same control flow;
// the window shifts are drawn by all lanes (chunk_shift_wave) and lane 0
// stores the new start, states and index. Returns the new local index.
__device__ __forceinline__ int32_t chunk_align_wave(const Landscape& L, int64_t p, double u, int32_t j0)
{
    if (!__builtin_isfinite(u)) {
        return j0;
    }
    const int64_t n = L.nchunk;
    const double* y = L.y + p * n;
    const int64_t b = L.buffer, m = L.margin;
    int64_t start = L.start[p];
    uint64_t stb = L.st_begin[p], ste = L.st_end[p];
    auto shift = [&](int64_t ns) {
        chunk_shift_wave(L, p, start, ns, stb, ste);
        start = ns;
        __threadfence_block(); // the row written by all lanes is read back below
    };
    int64_t j = j0;
    int64_t res = -1;
    int64_t budget = MAX_SEARCH_WELLS;
    while (!(u < y[n - 1])) {
        if (budget <= 0) {
            res = n - 2;
            break;
        }
        budget -= n - 1 - m;
        shift(start + (n - 1 - m));
        j = m;
    }
    if (res < 0) {
        while (u < y[0]) {
            if (budget <= 0) {
                res = 0;
                break;
            }
            budget -= n - 1 - m;
            shift(start - (n - 1 - m));
            j = n - 1 - m;
        }
    }
    if (res < 0) {
        if (j < 0) {
            j = 0;
        }
        if (j > n - 2) {
            j = n - 2;
        }
        while (j < n - 2 && u >= y[j + 1]) {
            ++j;
        }
        while (j > 0 && u < y[j]) {
            --j;
        }
        if (j < b || j > n - 2 - b) {
            shift(start + j - m);
            j = m;
        }
        res = j;
    }
    if ((threadIdx.x & 63) == 0) {
        L.start[p] = start;
        L.st_begin[p] = stb;
        L.st_end[p] = ste;
        L.jloc[p] = (int32_t)res;
    }
    return (int32_t)res;
}

// Align particle p (registers yl, yr, j hold its current well). Returns the
// change of the global well index (for event and activity counting).
__device__ __forceinline__ int64_t align_particle(const Landscape& L, int64_t p, double u, double& yl, double& yr,
                                                  int32_t& j)
{
    if (yl <= u && u < yr) {
        return 0;
    }
    const int64_t g0 = L.start[p] + j;
    j = chunk_align_slow(L, p, u, j);
    const double* y = L.y + p * L.nchunk;
    yl = y[j];
    yr = y[j + 1];
    return L.start[p] + j - g0;
}

} // namespace fq
