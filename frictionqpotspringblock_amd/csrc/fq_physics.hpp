// Per-particle force laws of the reference's policies, written as functions
// of register values so that every kernel (resident, service, no-passing)
// evaluates exactly the same expression tree as the reference.
#pragma once

#include "fq_common.hpp"

namespace fq {

constexpr double FQ_PI = 3.14159265358979323846;

// Cuspy::force detail.h:164-169: f = mu * (0.5 * (y_l + y_r) - u)
// SemiSmooth::force detail.h:261-276
// Smooth::force detail.h:402-408 (sin evaluated with fq_sin)
template <int POT>
FQ_HD double potential_force(const SysView& S, double u, double yl, double yr)
{
    if constexpr (POT == POT_CUSPY) {
        return (0.5 * (yl + yr) - u) * S.mu;
    }
    else if constexpr (POT == POT_SEMISMOOTH) {
        const double mu = S.mu, kappa = S.kappa;
        double xi = 0.5 * (yl + yr);
        double u_r = (mu * xi + kappa * yr) / (mu + kappa);
        double u_l = (mu * xi + kappa * yl) / (mu + kappa);
        if (u < u_l) {
            return kappa * (u - yl);
        }
        if (u <= u_r) {
            return mu * (0.5 * (yl + yr) - u);
        }
        return kappa * (u - yr);
    }
    else {
        double umin = 0.5 * (yl + yr);
        double dy = 0.5 * (yr - yl);
        return -S.mu * dy / FQ_PI * fq_sin(FQ_PI * (u - umin) / dy);
    }
}

// The same forces with the constants of the current well computed once per
// well (when it changes) instead of every step: Cuspy its mid-point (c0);
// SemiSmooth the edges u_l (c0), u_r (c1) of the parabolic part (two
// divisions); Smooth the prefactor (-mu dy) / pi of the sine (c0). Each
// constant is the subexpression of potential_force above, so the force is
// bit-identical.
template <int POT, class Sys>
FQ_HD void well_consts(const Sys& S, double yl, double yr, double& c0, double& c1)
{
    if constexpr (POT == POT_CUSPY) {
        c0 = 0.5 * (yl + yr);
        c1 = 0.0;
    }
    else if constexpr (POT == POT_SEMISMOOTH) {
        const double mu = S.mu, kappa = S.kappa;
        double xi = 0.5 * (yl + yr);
        c0 = (mu * xi + kappa * yl) / (mu + kappa);
        c1 = (mu * xi + kappa * yr) / (mu + kappa);
    }
    else {
        double dy = 0.5 * (yr - yl);
        c0 = -S.mu * dy / FQ_PI;
        c1 = 0.0;
    }
}

template <int POT, class Sys>
FQ_HD double potential_force_wc(const Sys& S, double u, double yl, double yr, double c0, double c1)
{
    if constexpr (POT == POT_CUSPY) {
        return (c0 - u) * S.mu;
    }
    else if constexpr (POT == POT_SEMISMOOTH) {
        if (u < c0) {
            return S.kappa * (u - yl);
        }
        if (u <= c1) {
            return S.mu * (0.5 * (yl + yr) - u);
        }
        return S.kappa * (u - yr);
    }
    else {
        double umin = 0.5 * (yl + yr);
        double dy = 0.5 * (yr - yl);
        return c0 * fq_sin(FQ_PI * (u - umin) / dy);
    }
}

// maxUniformDisplacement of one particle in its current well (yl <= u < yr):
//   Cuspy      detail.h:177-187  distance to the well's edge in `direction`
//   SemiSmooth detail.h:283-330  distance to the edge of the parabolic part;
//              `outside` is set (and +inf returned) when u is outside it
template <int POT>
FQ_HD double max_uniform_particle(const SysView& S, double up, double yl, double yr, int direction, int& outside)
{
    if constexpr (POT == POT_CUSPY) {
        return direction > 0 ? yr - up : up - yl;
    }
    else {
        const double mu = S.mu, kappa = S.kappa;
        double xi = 0.5 * (yl + yr);
        double u_r = (mu * xi + kappa * yr) / (mu + kappa);
        double u_l = (mu * xi + kappa * yl) / (mu + kappa);
        if (up < u_l || !(up <= u_r)) {
            outside = 1;
            return __builtin_inf();
        }
        return direction > 0 ? u_r - up : up - u_l;
    }
}

// Nearest-neighbour chain interactions of particle with slip u and
// neighbours um (p-1) and up (p+1), periodic.
//   Laplace1d         detail.h:480-486
//   Quartic1d         detail.h:784-803
//   QuarticGradient1d detail.h:642-651
template <int INT>
FQ_HD double chain_force(const SysView& S, double um, double u, double up)
{
    if constexpr (INT == INT_LAPLACE1D) {
        return (um - 2 * u + up) * S.k;
    }
    else if constexpr (INT == INT_QUARTIC1D) {
        double dup = up - u;
        double dun = um - u;
        return S.a1 * (um - 2 * u + up) + S.a2 * (dup * dup * dup + dun * dun * dun);
    }
    else if constexpr (INT == INT_QUARTICGRADIENT1D) {
        double du = up - um;
        return (um - 2 * u + up) * (S.k2 + S.k4_4 * du * du);
    }
    else {
        return 0.0;
    }
}

FQ_HD int64_t wrap(int64_t i, int64_t n)
{
    return i < 0 ? i + n : (i >= n ? i - n : i);
}

// Lattice interactions of site (i, j) read from a row-major slip field U;
// S supplies rows, cols (the periodic wrap) and k, k2, k4 (the system, or the
// extended rows of a tile).
//   Laplace2d         detail.h:545-583
//   QuarticGradient2d detail.h:682-750
template <int INT, class Sys, class Acc>
FQ_HD double lattice_force(const Sys& S, const Acc& U, int64_t i, int64_t j)
{
    const int64_t R = S.rows, C = S.cols;
    const int64_t im = wrap(i - 1, R), ip = wrap(i + 1, R), jm = wrap(j - 1, C), jp = wrap(j + 1, C);
    const double c0 = U(i * C + j);
    if constexpr (INT == INT_LAPLACE2D) {
        return (U(im * C + j) + U(ip * C + j) + U(i * C + jm) + U(i * C + jp) - 4 * c0) * S.k;
    }
    else {
        const double mk4_3 = S.k4 / 3.0;
        const double mk4_23 = 2.0 * mk4_3;
        const double uip = U(ip * C + j), uim = U(im * C + j), ujp = U(i * C + jp), ujm = U(i * C + jm);
        double l = uip + uim + ujp + ujm - 4 * c0;
        double dudx = 0.5 * (uip - uim);
        double dudy = 0.5 * (ujp - ujm);
        double d2udxdy = 0.25 * (U(ip * C + jp) - U(ip * C + jm) - U(im * C + jp) + U(im * C + jm));
        double d2udx2 = uip - 2 * c0 + uim;
        double d2udy2 = ujp - 2 * c0 + ujm;
        return l * (S.k2 + mk4_3) + mk4_23 * (dudx * dudx * d2udx2 + dudy * dudy * d2udy2 + 2.0 * dudx * dudy * d2udxdy);
    }
}

// LongRange1d::force detail.h:852-866: sequential sum over i = 0..N-1.
template <class Acc, class Pref>
FQ_HD double longrange_force(const SysView& S, const Acc& U, const Pref& P, int64_t p)
{
    const int64_t n = S.N;
    double fp = 0.0;
    const double up = U(p);
    for (int64_t i = 0; i < n; ++i) {
        if (i == p) {
            continue;
        }
        int64_t d = i > p ? i - p : p - i;
        if (d > S.lr_m) {
            d = n - d;
        }
        fp += (U(i) - up) * P(d);
    }
    return fp;
}

} // namespace fq
