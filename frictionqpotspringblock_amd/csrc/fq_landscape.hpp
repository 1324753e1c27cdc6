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
