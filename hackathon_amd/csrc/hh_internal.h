// Internal launch interface between the C-ABI layer and the kernel files.
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include <complex>
#include <vector>

namespace hh {

enum Prec { F64 = 0, F32 = 1 };

// Element types of a DST pass: input, compute, output (0 = double, 1 = float)
// and, for power-of-two N+1, the transform: 0 = (N+1)-point FFT with
// symmetric pre/post-processing (table: M twiddles then M sines), 1 =
// 2(N+1)-point FFT of the odd extension (table: 2(N+1) twiddles).
struct DstTypes {
    int in, comp, out;
    int algo = 0;
};

// Lines of a DST-I pass.  Line l = (plane p, column c), p = l / ncols,
// c = l % ncols; element q of the line sits at
//   p*plane_stride + c*col_stride + q*elem_stride.
struct Lines {
    int N;              // line length
    long ncols;         // lines per plane
    long nplanes;       // planes (batch folded in)
    long plane_stride;
    long col_stride;
    long elem_stride;
    // Row passes only: input / output in the mode-tiled layout (tile width,
    // a power of two; 0 = row-major), ny lines per right-hand side of `field`
    // elements.  Tile t0 holds modes t0..t0+w-1 of all ny lines, element
    // (j, k) at t0*ny + j*w + (k - t0), w = min(tw, N - t0).
    int in_tw = 0, out_tw = 0, ny = 1;
    long field = 0;
    // Contiguous lines only: acc != 0 adds acc times the transform to the
    // (FP64) output instead of storing it, out += acc X (the update of an
    // iterative refinement applied as the last pass stores).
    double acc = 0;
};

bool fft_supported(int N);  // N+1 a power of two, 2(N+1) <= HH_FFT_LMAX

// Orthonormal DST-I of every line.  FFT path (in == out allowed) when
// fft_supported(N), else the dense-matrix path (in != out required).
// tw: device twiddle table e^{-2 pi i q/L}, L = 2(N+1), in the compute type.
// S: device DST matrix (N*N, compute type) for the dense path.
hipError_t launch_dst(const Lines &ln, DstTypes ty, const void *in, void *out,
                      const void *tw, const void *S, hipStream_t st);

// Per-mode tridiagonal solves of the Helmholtz operator along the last axis.
struct TriHelm {
    int dim;            // 2 or 3
    int Nx, Ny, Nz;     // 2D: Nz = 1, system length Ny; 3D: length Nz
    long field;         // elements per right-hand side
    int batch;
    const void *coef;   // 2D: [q][6] (bm,am,b,a,bp,ap); 3D: [q][12] (a..dm)
    int nop;            // operators: right-hand side b uses table b % nop
    long cstride;       // complex entries between operator tables
    const double *lamx, *lamy;   // eigenvalues 2cos(k pi/(N+1))
    int W, P, mmax;     // block shape (modes per block, chunks per mode, rows per chunk)
    int nr = 1;         // right-hand sides per thread (sharing one operator): 1, 2 or 4
    int tiled = 0;      // field in the mode-tiled layout of the row DSTs: tile width (a multiple of W), 0 = row-major
    int real = 0;       // every coefficient real: use coefr, the real parts in pairs ([q][3] / [q][6] complex slots)
    const void *coefr = nullptr;
    // Near-resonant modes (ModeSet below): the partitioned kernel does not
    // store them (bit m % 32 of fbits[o * fwords + m / 32] set); they are
    // solved by the sequential kernel from the factors in fac.
    const uint32_t *fbits = nullptr;
    int fwords = 0;
    const int *foff = nullptr, *fmode = nullptr;   // per operator: entries foff[o] .. foff[o+1]-1, their modes
    const void *fac = nullptr;                     // [row][entry](L, P, R, Up): double (real) or double2
    int nfx = 0;                                   // most entries of one operator
    long nent = 0;                                 // entries of all operators
};
hipError_t launch_tri_helm(const TriHelm &t, DstTypes ty, const void *in, void *out, hipStream_t st);
// Sequential reference-order substitution of the near-resonant modes of t
// (FP64 arithmetic; storage as ty.in / ty.out); run before launch_tri_helm
// on the same buffer.  No-op when t.nfx == 0.
hipError_t launch_tri_seq(const TriHelm &t, DstTypes ty, const void *in, void *out, hipStream_t st);

// A mode whose condition estimate exceeds this is singular to working
// precision (no correct digit is left in FP64).
constexpr double HH_SINGULAR_KAPPA = 1e14;
// Default Varah-bound threshold above which a mode is near-resonant.
constexpr double HH_KAPPA_FLAG = 4096.0;

// Classification of the modes of a plan's operators (hh_modes.hip).
struct ModeSet {
    int n = 0, nop = 0, words = 0, nfx = 0;
    long nmodes = 0;
    std::vector<int> foff, fmode;         // per operator offsets (nop + 1), flagged modes
    std::vector<uint32_t> bits;           // [op][words]
    std::vector<std::complex<double>> facc;   // [entry][row](L, P, R, Up)
    std::vector<double> facr;             // real parts (real operators only)
    std::vector<int> singular;            // per operator
    std::vector<double> kappa_worst;      // per operator: largest condition estimate of its flagged modes
};
// hcoef: the plan's device-order band table [op][q][6 | 12]
void classify_modes(int dim, int Nx, int Ny, int Nz, int nop, const std::complex<double> *hcoef,
                    const double *lamx, const double *lamy, bool real, double kappa_flag, ModeSet &ms);

// Compact fourth-order derivative line solves (band 1, diag, 1) over an
// extended-grid field; see CompactProb in hh_tri.hip for the index maps.
struct Compact {
    int n, first;
    double diag, h;
    const void *f;
    long field_in, mA, orig, sA, sB, eS;
    const void *lo, *hi;
    long field_bnd, bo, bA, bB;
    void *out;
    long field_out, oorig, oA, oB, oS;
    long nmodes;
    int batch;
};
hipError_t launch_compact(const Compact &cd, hipStream_t st);

// Choose a tridiagonal block shape for a system of length n and nmodes modes.
void tri_shape(int n, long nmodes, int prec, int *W, int *P, int *mmax);
// LDS bytes of the tridiagonal kernel's reduced system: two buffers of A, C
// (operator values of op bytes) and nr complex right-hand sides (cx bytes).
inline size_t tri_lds_bytes(size_t op, size_t cx, int P, int W, int nr) {
    return 2 * (2 * op + nr * cx) * 2 * (size_t)P * W;
}
// block width W and partitions P for mm rows per thread (256-thread blocks at most)
void tri_shape_rows(int n, long nmodes, int mm, int *W, int *P);

// Compact-stencil right-hand-side assembly (2D_Program/main_functions/rhs.c,
// 3D_Program/main_functions/rhs.c and the partial derivatives).
struct Rhs2D {
    int order, Nx, Ny, batch;
    double hx, hy;
    const void *f_e;     // (Nx+2)(Ny+2) per rhs
    const void *u_e;     // (Nx+2)(Ny+2) per rhs, Dirichlet data
    const void *fy_b;    // 2*Nx per rhs
    const void *fxx_b;   // 2(Nx+2)+2(Ny+2) per rhs (ring)
    const void *fyy_b;   // ring
    const void *k2c;     // Ny: k^2 at interior rows
    const void *k2y;     // Ny: compact (k^2)_y at interior rows
    const void *coef;    // Ny x 6 per operator
    int nop;             // right-hand side b uses operator b % nop
    void *work;          // 2 (Nx+2)(Ny+2)+Nx Ny per rhs
};
hipError_t launch_rhs2d(const Rhs2D &r, void *F, hipStream_t st);

struct Rhs3D {
    int order, Nx, Ny, Nz, batch;
    double hx, hy, hz;
    const void *f_e, *u_e;        // extended (Nx+2)(Ny+2)(Nz+2) per rhs
    const void *fxx_f, *fyy_f, *fzz_f;  // face arrays
    const void *k2e;              // Nz+2
    const void *kz;               // Nz: compact (k^2)_z
    const void *coef;             // Nz x 12 per operator
    int nop;                      // right-hand side b uses operator b % nop
    void *work;                   // 3 extended fields per rhs
};
hipError_t launch_rhs3d(const Rhs3D &r, void *F, hipStream_t st);

// res = A u - F (the reference residual.c), and the fused refinement pieces.
// coef tables: right-hand side b uses coef + (b % nop) * (6 Ny | 12 Nz)
hipError_t launch_residual2d(int Nx, int Ny, int batch, const void *coef, int nop, const void *u,
                             const void *F, void *res, int res_f32, hipStream_t st);
hipError_t launch_residual3d(int Nx, int Ny, int Nz, int batch, const void *coef, int nop, const void *u,
                             const void *F, void *res, int res_f32, hipStream_t st);
// u += s*c (c in FP32 or FP64)
hipError_t launch_axpy_update(long n, void *u, const void *c, int c_f32, double s, hipStream_t st);
// out[0] = max |x|, out[1] = max |x - y| (y may be null); out is device
// memory of norm_buffer_doubles() values, out[2..] the per-block partials
constexpr int NORM_BLOCKS = 2048;
constexpr long norm_buffer_doubles() { return 2 + 2L * NORM_BLOCKS; }
// out[0..1] = the max norms (l2: the 2-norms) of x and x - y; out[2..] scratch
hipError_t launch_norms(long n, const void *x, int x_f32, const void *y, double *out, hipStream_t st,
                        bool l2 = false);
hipError_t launch_convert(long n, const void *in, int in_f32, void *out, int out_f32, hipStream_t st);

}  // namespace hh
