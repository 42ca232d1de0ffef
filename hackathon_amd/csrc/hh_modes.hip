// Host-side classification of the per-mode tridiagonal systems and the
// reference-order factors of the ill-conditioned ones.
//
// Every mode m of an operator o owns a tridiagonal system T (sub, dia, sup
// along y in 2D, z in 3D).  The partitioned elimination of the device kernel
// (hh_tri.hip) restarts the pivot recurrence at every 16-row chunk and
// combines the chunks through a reduced system; its relative error on a mode
// grows like c eps kappa(T) with c <= 0.05 measured over the C2/C4 operators
// (FP64, white-noise right-hand sides, against 40-digit solutions), which is
// up to 10x the error of the reference's sequential Thomas elimination on
// the nearly singular low modes.  Modes are therefore split:
//
//  * well separated: strictly diagonally dominant with Varah's bound
//      kappa_inf(T) <= max_q (|sub|+|dia|+|sup|) / min_q (|dia|-|sub|-|sup|)
//    at most kappa_flag (default 4096): partitioned elimination on the
//    device (relative error <= ~2e-14 in FP64; FP32 in the mixed path);
//  * near-resonant (everything else, including every indefinite mode with
//    |lambda_j - k^2| small): the factors of the reference's LU() are
//    computed here in its operation order (2D_Program/main_functions/
//    LU.c:16-40, divide form; 3D_Program/main_functions/LU.c:34-72,
//    reciprocal form), with contraction off, so they are bit-identical to
//    the reference's for real operators, and the device runs the
//    reference's substitution (2D_Program/solver_functions/solver.c:71-86,
//    3D_Program/solver_functions/solver.c:356-369) in sequential order with
//    the same operations, so these modes differ from the reference's only
//    through their input.  (An exact double-double substitution, evaluated
//    as a parallel scan, was measured too: more accurate than the
//    reference, it disagreed with it by up to 1.3e-12 on a manufactured
//    solution at N = 511, k = 39, where the reference's own FP64
//    substitution loses 4e-13 to pivots of 4e-3; the sequential order stays
//    within 3e-14 there.)
//
// A near-resonant mode whose estimated condition number exceeds
// HH_SINGULAR_KAPPA (or whose pivot is zero or not finite) makes its
// operator singular to working precision: the plan reports HH_ERR_SINGULAR
// for solves that use it.  The reference divides by such pivots unchecked
// (2D LU.c:36, 3D LU.c:68).
#include <cmath>
#include <complex>
#include <cstdint>
#include <vector>

#include "hh_internal.h"

namespace hh {

using cd = std::complex<double>;

namespace {

// Bands of row q of mode (lx, ly) in the reference's operation order:
// 2D  sub = bm lam + am, dia = b lam + a, sup = bp lam + ap  (LU.c:16-23),
// 3D  e3 = a lxy + b lx + c ly + d, and the P / M variants  (LU.c:34-47),
// complex coefficient times real eigenvalue componentwise.  Table rows:
// 2D (bm, am, b, a, bp, ap); 3D (a, b, c, d, ap, bp, cp, dp, am, bm, cm, dm).
struct RowBands {
    cd sub, dia, sup;
};

#pragma clang fp contract(off)
inline cd cr(cd a, double s) { return cd(a.real() * s, a.imag() * s); }
inline cd add(cd a, cd b) { return cd(a.real() + b.real(), a.imag() + b.imag()); }

inline RowBands row_bands(int dim, const cd *row, double lx, double ly, double lxy) {
    RowBands r;
    if (dim == 2) {
        r.sub = add(cr(row[0], lx), row[1]);
        r.dia = add(cr(row[2], lx), row[3]);
        r.sup = add(cr(row[4], lx), row[5]);
    } else {
        auto e = [&](const cd *c) { return add(add(add(cr(c[0], lxy), cr(c[1], lx)), cr(c[2], ly)), c[3]); };
        r.dia = e(row);
        r.sup = e(row + 4);
        r.sub = e(row + 8);
    }
    return r;
}

// Real-operator factors, reference order, FP64 with no fused operations.
// out rows: (L_q, P_q, R_q, Up_q).  2D: the reference divides by its pivot
// U_q; P_q = U_q and R_q = RN(1/U_q), with which the device reproduces the
// correctly rounded quotient (one fused correction step, Markstein).  3D:
// the reference multiplies by its stored reciprocal pivot U_q; R_q = U_q and
// P_q = 1/U_q (unused).  Returns false when a pivot is zero or not finite.
bool factors_real(int dim, int n, const cd *tab, int nc, double lx, double ly, double lxy, double *out) {
    bool ok = true;
    double Uprev = 0, Upprev = 0;
    for (int q = 0; q < n; q++) {
        const cd *row = tab + (long)q * nc;
        double sub, dia, sup;
        if (dim == 2) {
            sub = row[0].real() * lx + row[1].real();
            dia = row[2].real() * lx + row[3].real();
            sup = row[4].real() * lx + row[5].real();
        } else {
            auto e = [&](const cd *c) {
                return ((c[0].real() * lxy + c[1].real() * lx) + c[2].real() * ly) + c[3].real();
            };
            dia = e(row);
            sup = e(row + 4);
            sub = e(row + 8);
        }
        double L, U, R, P;
        if (dim == 2) {
            // LU.c:32-40: L = sub / U_{q-1}, U = dia - L Up_{q-1}
            L = q == 0 ? 0.0 : sub / Uprev;
            U = q == 0 ? dia : dia - L * Upprev;
            R = 1.0 / U;
            P = U;
            Uprev = U;
        } else {
            // LU.c:61-72: L = sub U_{q-1} (U reciprocal), U = 1 / (dia - Up_{q-1} L)
            L = q == 0 ? 0.0 : sub * Uprev;
            U = q == 0 ? 1.0 / dia : 1.0 / (dia - Upprev * L);
            R = U;
            P = 1.0 / U;
            Uprev = U;
        }
        if (!std::isfinite(L) || !std::isfinite(R) || !std::isfinite(P) || U == 0.0) ok = false;
        Upprev = sup;
        out[4 * q] = L;
        out[4 * q + 1] = P;
        out[4 * q + 2] = R;
        out[4 * q + 3] = q == n - 1 ? 0.0 : sup;   // the last row has no superdiagonal
    }
    return ok;
}

bool factors_complex(int dim, int n, const cd *tab, int nc, double lx, double ly, double lxy, cd *out) {
    bool ok = true;
    cd Uprev = 0, Upprev = 0;
    for (int q = 0; q < n; q++) {
        const RowBands b = row_bands(dim, tab + (long)q * nc, lx, ly, lxy);
        cd L, U, R, P;
        if (dim == 2) {
            L = q == 0 ? cd(0) : b.sub / Uprev;
            U = q == 0 ? b.dia : b.dia - L * Upprev;
            R = cd(1) / U;
            P = U;
        } else {
            L = q == 0 ? cd(0) : b.sub * Uprev;
            U = q == 0 ? cd(1) / b.dia : cd(1) / (b.dia - Upprev * L);
            R = U;
            P = cd(1) / U;
        }
        Uprev = U;
        if (!std::isfinite(std::abs(L)) || !std::isfinite(std::abs(R)) || U == cd(0)) ok = false;
        Upprev = b.sup;
        out[4 * q] = L;
        out[4 * q + 1] = P;
        out[4 * q + 2] = R;
        out[4 * q + 3] = q == n - 1 ? cd(0) : b.sup;
    }
    return ok;
}
#pragma clang fp contract(on)

// ||T^{-1} e||_inf for e = all ones and e = a fixed +-1 pattern, with the
// factors (L, P, R, Up) of the reference substitution.
template <typename V> double inv_norm_est(int n, const V *f) {
    double best = 0;
    std::vector<V> y(n);
    for (int pass = 0; pass < 2; pass++) {
        uint32_t s = 0x9e3779b9u;
        for (int q = 0; q < n; q++) {
            double e = 1.0;
            if (pass) { s = s * 1664525u + 1013904223u; e = (s >> 31) ? 1.0 : -1.0; }
            y[q] = q == 0 ? V(e) : V(e) - f[4 * q] * y[q - 1];
        }
        V x = 0;
        double mx = 0;
        for (int q = n - 1; q >= 0; q--) {
            x = q == n - 1 ? y[q] * f[4 * q + 2] : (y[q] - f[4 * q + 3] * x) * f[4 * q + 2];
            mx = std::max(mx, std::abs(x));
        }
        best = std::max(best, mx);
    }
    return best;
}

}  // namespace

void classify_modes(int dim, int Nx, int Ny, int Nz, int nop, const cd *hcoef, const double *lamx,
                    const double *lamy, bool real, double kappa_flag, ModeSet &ms) {
    const int n = dim == 2 ? Ny : Nz;
    const int nc = dim == 2 ? 6 : 12;
    const long nmodes = dim == 2 ? Nx : (long)Nx * Ny;
    ms.n = n;
    ms.nmodes = nmodes;
    ms.nop = nop;
    ms.words = (int)((nmodes + 31) / 32);
    ms.foff.assign(nop + 1, 0);
    ms.fmode.clear();
    ms.bits.assign((size_t)nop * ms.words, 0u);
    ms.facc.clear();
    ms.facr.clear();
    ms.singular.assign(nop, 0);
    ms.kappa_worst.assign(nop, 0.0);
    ms.nfx = 0;
    std::vector<RowBands> rows(n);
    for (int o = 0; o < nop; o++) {
        const cd *tab = hcoef + (size_t)o * n * nc;
        // a table whose rows are all equal (constant k) gives every interior
        // row the same bands: one row decides the mode
        bool uniform = true;
        for (int q = 1; q < n && uniform; q++)
            for (int t = 0; t < nc; t++)
                if (tab[(long)q * nc + t] != tab[t]) { uniform = false; break; }
        ms.foff[o] = (int)ms.fmode.size();
        for (long m = 0; m < nmodes; m++) {
            const double lx = dim == 2 ? lamx[m] : lamx[m % Nx];
            const double ly = dim == 2 ? 0.0 : lamy[m / Nx];
            const double lxy = lx * ly;
            double margin = INFINITY, rowsum = 0;
            const int qend = uniform ? std::min(n, 3) : n;
            for (int q = 0; q < qend; q++) {
                // magnitudes as sqrt(re^2 + im^2) (the bands are far from
                // overflow; std::abs's hypot costs several times more, and a
                // serving loop classifies every request's operator)
                double s, d, u;
                if (real) {
                    const cd *row = tab + (long)q * nc;
                    double sub, dia, sup;
                    if (dim == 2) {
                        sub = row[0].real() * lx + row[1].real();
                        dia = row[2].real() * lx + row[3].real();
                        sup = row[4].real() * lx + row[5].real();
                    } else {
                        auto e = [&](const cd *c) {
                            return c[0].real() * lxy + c[1].real() * lx + c[2].real() * ly + c[3].real();
                        };
                        dia = e(row);
                        sup = e(row + 4);
                        sub = e(row + 8);
                    }
                    s = std::fabs(sub); d = std::fabs(dia); u = std::fabs(sup);
                } else {
                    const RowBands b = row_bands(dim, tab + (long)q * nc, lx, ly, lxy);
                    auto mag = [](cd z) { return std::sqrt(z.real() * z.real() + z.imag() * z.imag()); };
                    s = mag(b.sub); d = mag(b.dia); u = mag(b.sup);
                }
                if (q == 0) s = 0.0;
                if (q == n - 1) u = 0.0;
                margin = std::min(margin, d - s - u);
                rowsum = std::max(rowsum, d + s + u);
            }
            const bool flag = !(margin > 0) || rowsum > kappa_flag * margin;
            if (!flag) continue;
            const long e = (long)ms.fmode.size();
            ms.fmode.push_back((int)m);
            ms.bits[(size_t)o * ms.words + m / 32] |= 1u << (m % 32);
            ms.facc.resize((size_t)(e + 1) * n * 4);
            cd *fc = ms.facc.data() + (size_t)e * n * 4;
            bool ok;
            double inv;
            if (real) {
                std::vector<double> fr(4 * (size_t)n);
                ok = factors_real(dim, n, tab, nc, lx, ly, lxy, fr.data());
                for (size_t i = 0; i < fr.size(); i++) fc[i] = fr[i];
                ms.facr.insert(ms.facr.end(), fr.begin(), fr.end());
                inv = ok ? inv_norm_est<double>(n, fr.data()) : INFINITY;
            } else {
                ok = factors_complex(dim, n, tab, nc, lx, ly, lxy, fc);
                inv = ok ? inv_norm_est<cd>(n, fc) : INFINITY;
            }
            const double kap = rowsum * inv;
            if (!ok || !std::isfinite(kap) || kap > HH_SINGULAR_KAPPA) ms.singular[o] = 1;
            if (!(kap <= ms.kappa_worst[o])) ms.kappa_worst[o] = kap;
        }
        const int cnt = (int)ms.fmode.size() - ms.foff[o];
        if (cnt > ms.nfx) ms.nfx = cnt;
    }
    ms.foff[nop] = (int)ms.fmode.size();
}

}  // namespace hh
