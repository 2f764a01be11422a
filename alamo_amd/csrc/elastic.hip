// elastic.hip -- libalamo_elastic.so: C-ABI, operator storage and the
// multigrid driver of the nodal elastic solve on one MI355X.
//
// The driver restates the amrex::MLMG control flow that
// Solver::Nonlocal::Linear configures (src/Solver/Nonlocal/Linear.H:255-290):
// F-cycle every iteration (setMaxFmgIter(1e8)), nu1 = nu2 = 2 Fsmooth calls,
// BiCGStab bottom solve, inf-norm stopping test against
// max(tol_abs, tol_rel * max(|b|, |r0|)).  All field work runs in HIP kernels
// on the caller's stream; the host only reads one residual norm per cycle.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <type_traits>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/alamo_elastic.h"
#include "../../include/alamo_unet.h"
#include "bottom_kernels.h"
#include "cluster.h"
#include "common.h"
#include "elastic_kernels.h"
#include "inverse_kernels.h"
#include "amr_kernels.h"
#include "learned_kernels.h"
#include "psi_kernels.h"
#include "stencil.h"
#include "stencil_lds.h"

using ael::kBlock;
using ael::Lvl;

thread_local std::string alamo_detail::g_err;

namespace {

using alamo_detail::fail;
using alamo_detail::g_err;

#define HIPCHK(expr)                                                                         \
    do {                                                                                     \
        hipError_t e_ = (expr);                                                              \
        if (e_ != hipSuccess)                                                                \
            return fail(ALAMO_EL_EHIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

inline unsigned nblk(long n) { return (unsigned)((n + kBlock - 1) / kBlock); }
// grid of the XCD-remapped stencil kernels: a multiple of 8 blocks
inline unsigned nblk8(long n) { return (unsigned)((nblk(n) + ael::kXcd - 1) / ael::kXcd * ael::kXcd); }

// Trilinear prolongation of the three displacement components, fine = P crse
// (zero_first) or fine += P crse; with held, of crse + held.  One thread per
// coarse node.
template <typename T>
void launch_interp3(const T* crse, const T* held, const ael::Lvl& Lc, T* fine, const ael::Lvl& Lf, int zero_first,
                    hipStream_t s) {
    if (held)
        ael::k_interp_cell<T, true><<<nblk8(Lc.N), kBlock, 0, s>>>(crse, held, Lc, fine, Lf, zero_first);
    else
        ael::k_interp_cell<T, false><<<nblk8(Lc.N), kBlock, 0, s>>>(crse, nullptr, Lc, fine, Lf, zero_first);
}

int ncomp_of(int sym) { return sym == 0 ? 2 : (sym == 1 ? 21 : (sym == 2 ? 9 : -1)); }

template <typename T>
struct Vecs {
    T* C = nullptr;
    T* diag = nullptr;
    T* cor = nullptr;
    T* cor_hold = nullptr;
    T* res = nullptr;
    T* rescor = nullptr;
    T* tmp = nullptr;
    T* D = nullptr;  // compact grad-C matrices, 18 x nflag (MajorMinor)
    T* Ct = nullptr; // tiled levels: tile-blocked interior copy of C (ael::TileC)
    T* Cb = nullptr; // tiled levels: C of the boundary nodes, in boundary-node order
    T* pal = nullptr; // tiled levels with a palette: the table of distinct C rows (ael::kPalStride)
};

struct Level {
    Lvl L{};  // L.per: periodic axes, the handle's on every MG level
    Vecs<double> d;
    Vecs<float> f;
    unsigned char* gmask = nullptr;  // per node: axes along which C varies (k_grad_mask)
    int* dwbase = nullptr;           // per interior wave of the launch: first compact D position
    int nflag = 0;                   // interior nodes with a nonzero gmask
    bool tiled = false;              // MajorMinor level served by k_stencil_lds (else k_stencil_slab)
    double* psi = nullptr;           // SetPsi weight on this level's cells (null until set)
    ael::PalIdx* pidx = nullptr;     // palette row of every interior node, tile-blocked (null: no palette)
    ael::PalIdx* pidx_b = nullptr;   // palette row of every boundary node, in boundary_node order
    long dcap = 0, fcap = 0;         // capacities (entries) of d.D and f.D, kept across models
    long nwcap = 0;                  // capacity of dwbase
    int npal = 0;
    ael::TileGrid tg{}, tg32{};      // its tile grids (chunk lengths) of the FP64 and FP32 launches
};

template <typename T> Vecs<T>& vecs(Level& l);
template <> Vecs<double>& vecs<double>(Level& l) { return l.d; }
template <> Vecs<float>& vecs<float>(Level& l) { return l.f; }

}  // namespace

struct alamo_el_s {
    int sym = 0;
    int nc = 0;
    int per = 0;  // periodic axes (bit a), alamo_el_set_periodic
    std::vector<Level> lv;
    int* bct_dev = nullptr;
    double* bcv_dev = nullptr;
    int bct_host[78];
    double bcv_host[78];
    int uniform = 0;
    double omega = 2.0 / 3.0;
    bool model_set = false, prepared = false, have32 = false, prepared32 = false;
    double* red_partial = nullptr;
    double* red_out = nullptr;
    double* red_host = nullptr;  // pinned
    bool psi_set = false;        // SetPsi called: every operator row carries the psi weight
    double psi_small = 1e-8;
    double* red_host_vec = nullptr;  // pinned, FGMRES coefficients (two steps' copies)
    hipEvent_t kev[2] = {nullptr, nullptr};  // the copies' completion, by step parity
    double* bottom_work = nullptr;
    int* bottom_status = nullptr;
    double* gu9 = nullptr;  // Newton scratch, allocated on demand
    double* dw9 = nullptr;
    double* csoa = nullptr;
    // dense coarsest level (when it has at most kDenseMax unknowns)
    int dense_M = 0;
    double* dense_I = nullptr;     // identity, the unit vectors of the assembly
    double* dense_A = nullptr;     // operator, column-major
    // direct bottom (ALAMO_EL_BOTTOM_DIRECT): the finest MG level with at
    // most kDirectMax unknowns and the FP64 inverse of its operator
    int direct_level = -1;
    int direct_M = 0;
    double* direct_W = nullptr;     // work matrix of the inversion (M x M)
    double* direct_Ainv = nullptr;  // inverse, row-major (null when singular)
    int* direct_perm = nullptr;     // row swapped with row k at pivot step k (M)
    int* direct_map = nullptr;      // a block's row map, then the inverse's column sources (M)
    double* direct_part = nullptr;  // partial sums of the bottom product
    ael::InvState* direct_st = nullptr;
    double* direct_A = nullptr;     // the operator (M x M), kept for the check and the pivoted inversion
    double* direct_X = nullptr;     // D^-1 times a block's rows (kInvBlock x M)
    double* direct_D = nullptr;     // D^-1 of a diagonal block (kInvBlock x kInvBlock)
    double* direct_chk = nullptr;   // x, Ainv x and A Ainv x of the check (3 M)
    int direct_pivoted = 0;         // the last inverse came from the pivoted panels
    // FGMRES storage, allocated on demand
    double* kh = nullptr;          // Gram-Schmidt coefficients on the device
    double* kpart = nullptr;       // multi-dot partials
    int krylov_m = 0;
    double* kv = nullptr;  // (m+1) vectors
    double* kz = nullptr;  // m vectors
    double* kw = nullptr;
    // learned coarse-level solver (alamo_el_set_learned); the network is the caller's
    alamo_nn_t nn = nullptr;
    int nn_level = -1;
    float* nn_in = nullptr;    // network input [nc + 30][N_L]
    float* nn_out = nullptr;   // network output [3][N_L]
    double* nn_bc = nullptr;   // level-L right-hand side
    double* nn_t = nullptr;    // scaled right-hand side, capacity nn_cap nodes
    double* nn_fmean = nullptr;  // 18 face means of nn_t
    long nn_cap = 0;
    // AMR placement (alamo_amr_patch_create); a base handle has no parent
    alamo_el_t amr_parent = nullptr;
    int amr_dom[3] = {0, 0, 0};  // domain cells at this handle's level
    ael::Patch amr{};            // origin and covered parent cells (origin 0 for a base handle)
};

namespace {

constexpr int kRedBlocks = 1024;
// multi-dot partials: 4 waves per SIMD over the 256 CUs, one resident round
// (FP32 solve at 129^3: 23.9 ms against 24.1 with 2048, 24.2 with 4096,
// 24.4 with 768 and 25.6 with 512)
constexpr int kDotBlocks = 1024;
constexpr int kDenseMax = 1536;  // unknowns; LDS of the dense BiCGStab = 9 M doubles
constexpr int kDirectMax = 4096; // unknowns of the direct bottom level (an M x M FP64 inverse)
// Weights of the red-black Gauss-Seidel smoother.  Interior rows: 0.9; the
// undamped 1.0 converges fastest on a polycrystal (V-cycle factor 0.775 vs
// 0.802 at 33^3) but diverges below 1e-7 at the sharp E = 10 interface of the
// spherical inclusion, where 0.9 takes 25 V-cycles to 1e-10 (damped Jacobi
// 28).  Boundary rows: 0.5; traction rows are one-sided first-order rows, and
// at 1.0 the V-cycle on a traction-loaded polycrystal diverges.
constexpr double kRbInterior = 0.9, kRbBoundary = 0.5;

// ---------------------------------------------------------------------------
// Kernel dispatch
// ---------------------------------------------------------------------------
// Stencil kernel variant: the XCD-slab per-node kernel with the grad-C mask
// (default), the same without the mask, or the linear per-node kernel.
// ALAMO_EL_STENCIL=slab|slab-nomask|node selects one.  slab-nomask and node
// give the same result bit for bit; for MajorMinor the default slab kernel
// sums the grad-C term through the compact matrices D (one rounding order
// apart, about 1e-16 relative), for the other symmetries it agrees bit for bit.
enum { kVariantSlab = 0, kVariantSlabNoMask = 1, kVariantNode = 2, kVariantSlabOnly = 3 };

// Tile shape of the LDS-tiled kernel: 32 x 8 nodes, four waves (1-wave
// 32 x 2 and 16 x 4 tiles, 64 x 1 rows, and 64 x 4 and 128 x 2 tiles measured
// no faster).
using TileT = ael::Tile<32, 8>;
constexpr long kNtMinBytes = 64l << 20;
// Chunks of k planes sized for the resident tile workgroups: two rounds of
// 2 per CU in FP64 (200-220 VGPRs), one round of 4 per CU in FP32 (110-120
// VGPRs).  At 129^3: FP64 chunks of 8 planes (1024 workgroups; Fapply
// 104.4 us vs 108.4 with 3 planes, 107.0 with 4, 108.8 with 16), FP32 chunks
// of 8 (sweep 37.2 us vs 39.2 with chunks of 4, 41.4 with 16; the FP32
// solve 0.15 ms faster).  ALAMO_EL_TILE_TARGET (read when a level's C is
// set) replaces both targets: the tests use it to reach every chunk length
// on small boxes.
constexpr long kTileTarget64 = 2 * 2 * 256, kTileTarget32 = 4 * 256;
ael::TileGrid tile_grid_of(const Lvl& L, long target) {
    const char* e = std::getenv("ALAMO_EL_TILE_TARGET");
    const long t = e ? std::atol(e) : 0;
    return ael::tile_grid<TileT>(L, t > 0 ? t : target);
}

int stencil_variant() {
    static int v = -1;
    if (v < 0) {
        const char* e = std::getenv("ALAMO_EL_STENCIL");
        v = kVariantSlab;
        if (e && std::strcmp(e, "slab-nomask") == 0) v = kVariantSlabNoMask;
        if (e && std::strcmp(e, "node") == 0) v = kVariantNode;
        if (e && std::strcmp(e, "slab-only") == 0) v = kVariantSlabOnly;
    }
    return v;
}

// The LDS-tiled kernel on level l, when it serves the level: MajorMinor, the
// masked stencil variant, a tiled level with its compact D and C copies, and C
// the level's own.  Returns false otherwise.  x may be an FP32 field for the
// FP64 Fapply (TX = float, T = double, OP = 0).
template <typename T, int OP, typename TX>
bool launch_tiled(alamo_el_t h, const Level& l, const TX* x, const T* b, const T* diag, const T* C, T* out,
                  hipStream_t s) {
    const int var = stencil_variant();
    if (!(var == kVariantSlab || var == kVariantSlabOnly) || h->sym != 1 || !l.tiled) return false;
    Level& lv = const_cast<Level&>(l);
    const bool have_d = l.dwbase && (l.nflag == 0 || vecs<T>(lv).D);
    if (!have_d || !vecs<T>(lv).Ct || !vecs<T>(lv).Cb || C != vecs<T>(lv).C) return false;
    const Lvl L = l.L;
    const T w = (T)h->omega;
    const int u = h->uniform;
    const unsigned nbb = ael::slab_boundary_blocks(L, kBlock);
    const unsigned char* gm = l.gmask;
    // its compact D is in tile order.  C streamed with the non-temporal hint
    // on levels whose C exceeds kNtMinBytes: read once per launch, it then
    // leaves L2 to the u halos (129^3: FP64 Fapply 124 -> 108 us, FP32 sweep
    // 65 -> 62 us).  A smaller level's C stays in the 256 MB last-level cache
    // between sweeps, which the hint gives up (65^3 FP32 sweep: 11.9 us
    // without, 12.6 us with).
    const ael::TileGrid& G = sizeof(T) == 8 ? l.tg : l.tg32;
    const unsigned grid = nbb + ael::tile_blocks(G);
    // the variant with the periodic faces' wrapped rows only on a periodic level
    auto go = [&](auto per_tag) {
        constexpr bool PER = decltype(per_tag)::value;
        if (l.pidx && vecs<T>(lv).pal)
            ael::k_stencil_lds<TileT, T, OP, true, 2, TX, PER><<<grid, TileT::NT, 0, s>>>(
                x, b, diag, vecs<T>(lv).pal, vecs<T>(lv).Cb, vecs<T>(lv).C, h->bct_dev, u, L, w, gm, nbb, G, l.dwbase,
                vecs<T>(lv).D, l.nflag, out, l.pidx, l.pidx_b);
        else if ((long)sizeof(T) * 21 * L.N >= kNtMinBytes)
            ael::k_stencil_lds<TileT, T, OP, true, 1, TX, PER><<<grid, TileT::NT, 0, s>>>(
                x, b, diag, vecs<T>(lv).Ct, vecs<T>(lv).Cb, vecs<T>(lv).C, h->bct_dev, u, L, w, gm, nbb, G, l.dwbase,
                vecs<T>(lv).D, l.nflag, out, nullptr, nullptr);
        else
            ael::k_stencil_lds<TileT, T, OP, true, 0, TX, PER><<<grid, TileT::NT, 0, s>>>(
                x, b, diag, vecs<T>(lv).Ct, vecs<T>(lv).Cb, vecs<T>(lv).C, h->bct_dev, u, L, w, gm, nbb, G, l.dwbase,
                vecs<T>(lv).D, l.nflag, out, nullptr, nullptr);
    };
    if (L.per) go(std::true_type{});
    else go(std::false_type{});
    return true;
}

template <typename T, int OP>
void launch_stencil_raw(alamo_el_t h, const Level& l, const T* x, const T* b, const T* diag, const T* C, T* out,
                        hipStream_t s) {
    const Lvl L = l.L;
    const T w = (T)h->omega;
    const int u = h->uniform;
    const int var = stencil_variant();
    if (var == kVariantNode) {
        const unsigned g = nblk8(L.N);
        switch (h->sym) {
            case 0: ael::k_stencil_node<T, 0, OP><<<g, kBlock, 0, s>>>(x, b, diag, C, h->bct_dev, u, L, w, out); break;
            case 1: ael::k_stencil_node<T, 1, OP><<<g, kBlock, 0, s>>>(x, b, diag, C, h->bct_dev, u, L, w, out); break;
            default: ael::k_stencil_node<T, 2, OP><<<g, kBlock, 0, s>>>(x, b, diag, C, h->bct_dev, u, L, w, out); break;
        }
        return;
    }
    if (launch_tiled<T, OP, T>(h, l, x, b, diag, C, out, s)) return;
    const unsigned g = ael::slab_blocks(L, kBlock), nbb = ael::slab_boundary_blocks(L, kBlock);
    const bool masked = var == kVariantSlab || var == kVariantSlabOnly;
    const unsigned char* gm = masked ? l.gmask : nullptr;
    Level& lv = const_cast<Level&>(l);
    const bool have_d = l.dwbase && (l.nflag == 0 || vecs<T>(lv).D);
    // compact grad-C matrices in slab order: MajorMinor, masked variant, built for this level
    const bool dc = gm && h->sym == 1 && !l.tiled && have_d;
    const int* dwb = dc ? l.dwbase : nullptr;
    const T* Dc = dc ? vecs<T>(lv).D : nullptr;
    const int nf = dc ? l.nflag : 0;
    auto go = [&](auto per_tag) {
        constexpr bool P = decltype(per_tag)::value;
        switch (h->sym) {
            case 0:
                ael::k_stencil_slab<T, 0, OP, false, P><<<g, kBlock, 0, s>>>(x, b, diag, C, h->bct_dev, u, L, w, gm, nbb,
                                                                          dwb, Dc, nf, out);
                break;
            case 1:
                if (dc)
                    ael::k_stencil_slab<T, 1, OP, true, P><<<g, kBlock, 0, s>>>(x, b, diag, C, h->bct_dev, u, L, w, gm,
                                                                             nbb, dwb, Dc, nf, out);
                else
                    ael::k_stencil_slab<T, 1, OP, false, P><<<g, kBlock, 0, s>>>(x, b, diag, C, h->bct_dev, u, L, w, gm,
                                                                              nbb, dwb, Dc, nf, out);
                break;
            default:
                ael::k_stencil_slab<T, 2, OP, false, P><<<g, kBlock, 0, s>>>(x, b, diag, C, h->bct_dev, u, L, w, gm, nbb,
                                                                          dwb, Dc, nf, out);
                break;
        }
    };
    if (L.per) go(std::true_type{});
    else go(std::false_type{});
}

// Operator rows with the psi weight (psi_kernels.h) when SetPsi was called:
// the unweighted rows of x first, then k_psi_post finishes Fapply (OP 0), the
// residual (OP 1) or the Jacobi sweep (OP 2).  out must differ from x and b.
template <typename T, int OP>
void launch_stencil(alamo_el_t h, const Level& l, const T* x, const T* b, const T* diag, const T* C, T* out,
                    hipStream_t s) {
    if (!h->psi_set || !l.psi) {
        launch_stencil_raw<T, OP>(h, l, x, b, diag, C, out, s);
        return;
    }
    launch_stencil_raw<T, 0>(h, l, x, (const T*)nullptr, (const T*)nullptr, C, out, s);
    const Lvl L = l.L;
    const T w = (T)h->omega;
    switch (h->sym) {
        case 0: ael::k_psi_post<T, 0><<<nblk(L.N), kBlock, 0, s>>>(x, b, diag, C, h->bct_dev, l.psi, h->psi_small, L, w, OP, out); break;
        case 1: ael::k_psi_post<T, 1><<<nblk(L.N), kBlock, 0, s>>>(x, b, diag, C, h->bct_dev, l.psi, h->psi_small, L, w, OP, out); break;
        default: ael::k_psi_post<T, 2><<<nblk(L.N), kBlock, 0, s>>>(x, b, diag, C, h->bct_dev, l.psi, h->psi_small, L, w, OP, out); break;
    }
}

// The psi weight on the rows of M unit-vector columns (dense assemblies).
void psi_post_columns(alamo_el_t h, const Level& l, const double* I, double* A, int M, hipStream_t s) {
    if (!h->psi_set || !l.psi) return;
    const dim3 grid(nblk(l.L.N), M);
    switch (h->sym) {
        case 0: ael::k_psi_post<double, 0><<<grid, kBlock, 0, s>>>(I, nullptr, nullptr, l.d.C, h->bct_dev, l.psi, h->psi_small, l.L, 0.0, 0, A); break;
        case 1: ael::k_psi_post<double, 1><<<grid, kBlock, 0, s>>>(I, nullptr, nullptr, l.d.C, h->bct_dev, l.psi, h->psi_small, l.L, 0.0, 0, A); break;
        default: ael::k_psi_post<double, 2><<<grid, kBlock, 0, s>>>(I, nullptr, nullptr, l.d.C, h->bct_dev, l.psi, h->psi_small, l.L, 0.0, 0, A); break;
    }
}

template <typename T>
void launch_fapply(alamo_el_t h, const Level& l, const T* u, const T* C, T* f, hipStream_t s) {
    launch_stencil<T, 0>(h, l, u, (const T*)nullptr, (const T*)nullptr, C, f, s);
}

template <typename T>
void launch_residual(alamo_el_t h, const Level& l, const T* x, const T* b, const T* C, T* r, hipStream_t s) {
    launch_stencil<T, 1>(h, l, x, b, (const T*)nullptr, C, r, s);
}

template <typename T>
void launch_jacobi(alamo_el_t h, const Level& l, const T* x, const T* b, const T* diag, const T* C, T* xo,
                   hipStream_t s) {
    launch_stencil<T, 2>(h, l, x, b, diag, C, xo, s);
}

template <typename T>
void launch_diagonal(alamo_el_t h, const Level& l, const T* C, T* diag, hipStream_t s) {
    const Lvl L = l.L;
    switch (h->sym) {
        case 0: ael::k_diagonal<T, 0><<<nblk8(L.N), kBlock, 0, s>>>(C, h->bct_dev, L, diag); break;
        case 1: ael::k_diagonal<T, 1><<<nblk8(L.N), kBlock, 0, s>>>(C, h->bct_dev, L, diag); break;
        default: ael::k_diagonal<T, 2><<<nblk8(L.N), kBlock, 0, s>>>(C, h->bct_dev, L, diag); break;
    }
    if (h->psi_set && l.psi) ael::k_psi_diag<T><<<nblk(L.N), kBlock, 0, s>>>(diag, h->bct_dev, l.psi, h->psi_small, L);
}

template <typename T>
void launch_fill(T* p, T v, long n, hipStream_t s) {
    if (n > 0) ael::k_fill<T><<<nblk(n), kBlock, 0, s>>>(p, v, n);
}

template <typename T>
void launch_copy(T* dst, const T* src, long n, hipStream_t s) {
    if (n > 0) ael::k_cast<T, T><<<nblk(n), kBlock, 0, s>>>(src, dst, n);
}

template <typename T>
void launch_axpy(T* y, T a, const T* x, long n, hipStream_t s) {
    if (n > 0) ael::k_axpy<T><<<nblk(n), kBlock, 0, s>>>(y, a, x, n);
}

// |v|_inf (op 0) or a.b (op 1) to the host.
template <typename T>
int reduce_host(alamo_el_t h, const T* a, const T* b, long n, int op, double* out, hipStream_t s) {
    const int nb = (int)std::min<long>(kRedBlocks, std::max<long>(1, nblk(n)));
    ael::k_reduce_partial<T><<<nb, kBlock, 0, s>>>(a, b, n, op, h->red_partial);
    ael::k_reduce_final<<<1, kBlock, 0, s>>>(h->red_partial, nb, op, h->red_out);
    HIPCHK(hipMemcpyAsync(h->red_host, h->red_out, sizeof(double), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    *out = h->red_host[0];
    return 0;
}

// ---------------------------------------------------------------------------
// Multigrid cycles, templated on the working precision
// ---------------------------------------------------------------------------
template <typename T>
struct Cycle {
    alamo_el_t h;
    hipStream_t s;
    int nu1, nu2, nuf, bottom_max_iter;
    double bottom_tol;
    int bottom_solver = ALAMO_EL_BOTTOM_BICGSTAB;
    int smoother = ALAMO_EL_SMOOTH_JACOBI;
    // F-cycle: leave the finest level's correction as cor + cor_hold for the
    // caller to sum (it reads the two in one pass)
    bool leave_held = false;

    int nlev() const { return (int)h->lv.size(); }
    Level& L(int l) { return h->lv[l]; }
    Vecs<T>& V(int l) { return vecs<T>(h->lv[l]); }
    long M(int l) { return 3 * h->lv[l].L.N; }

    // Operator<Grid::Node>::Fsmooth x n: 2n Jacobi sweeps (Operator.cpp:107);
    // with the red-black smoother, n red and black half sweeps.
    void smooth(int l, T* x, const T* b, int n) {
        Vecs<T>& v = V(l);
        if (smoother == ALAMO_EL_SMOOTH_RBGS) {
            for (int q = 0; q < n; ++q)
                for (int color = 0; color < 2; ++color) {
                    launch_residual<T>(h, L(l), x, b, v.C, v.tmp, s);
                    ael::k_rb_update<T><<<nblk(L(l).L.N), kBlock, 0, s>>>(x, v.tmp, v.diag, L(l).L, color,
                                                                          T(kRbInterior), T(kRbBoundary));
                }
            return;
        }
        for (int q = 0; q < n; ++q) {
            launch_jacobi<T>(h, L(l), x, b, v.diag, v.C, v.tmp, s);
            launch_jacobi<T>(h, L(l), v.tmp, b, v.diag, v.C, x, s);
        }
    }
    // x = 0, then Fsmooth x n; the first Jacobi sweep from zero is w b / D,
    // one elementwise kernel instead of the fill and a stencil sweep (with
    // first_ready, the restriction into this level has already written it
    // to v.tmp: fz_ready).
    bool fz_ready() const { return smoother == ALAMO_EL_SMOOTH_JACOBI && nu1 > 0; }
    void smooth_from_zero(int l, T* x, const T* b, int n, bool first_ready = false) {
        Vecs<T>& v = V(l);
        if (n <= 0 || smoother == ALAMO_EL_SMOOTH_RBGS) {
            launch_fill<T>(x, T(0), M(l), s);
            smooth(l, x, b, n);
            return;
        }
        if (!first_ready) ael::k_jacobi_from_zero<T><<<nblk(M(l)), kBlock, 0, s>>>(b, v.diag, (T)h->omega, v.tmp, M(l));
        launch_jacobi<T>(h, L(l), v.tmp, b, v.diag, v.C, x, s);
        smooth(l, x, b, n - 1);
    }
    void corr_res(int l, T* r, const T* x, const T* b) { launch_residual<T>(h, L(l), x, b, V(l).C, r, s); }
    void restrict_res(int lf, const T* fine, T* crse) {
        ael::k_restrict<T><<<nblk8(L(lf + 1).L.N), kBlock, 0, s>>>(fine, L(lf).L, crse, L(lf + 1).L, 3);
    }
    // restriction into V(lf + 1).res that also writes that level's first
    // Jacobi sweep from zero into its tmp
    void restrict_res_fz(int lf, const T* fine) {
        Vecs<T>& c = V(lf + 1);
        ael::k_restrict_fz<T><<<nblk8(L(lf + 1).L.N), kBlock, 0, s>>>(fine, L(lf).L, c.res, L(lf + 1).L, 3, c.diag,
                                                                      (T)h->omega, c.tmp);
    }
    void interp(int lf, T* fine, const T* crse, int zero_first) {
        launch_interp3<T>(crse, nullptr, L(lf + 1).L, fine, L(lf).L, zero_first, s);
    }
    // Bottom level of the cycles: the direct level when its inverse is used,
    // else the coarsest level.
    bool direct() const { return bottom_solver == ALAMO_EL_BOTTOM_DIRECT && h->direct_Ainv; }
    int bot() const { return direct() ? h->direct_level : (int)h->lv.size() - 1; }
    void bottom() {
        if (direct()) {
            // exact FP64 solve with the inverse of the direct level's operator
            const int l = bot();
            const int Mb = (int)M(l);
            // the FP32 cycles keep this solve in FP64 too: the FP64 inverse
            // (row-major) times the residual widened to FP64, the sums in FP64,
            // the result rounded to the cycle's precision; one wave a row
            ael::k_dense_mv_rows<T><<<nblk(64L * Mb), kBlock, 0, s>>>(h->direct_Ainv, V(l).res, Mb, V(l).cor);
            return;
        }
        const int l = nlev() - 1;
        Level& lv = L(l);
        const int Mb = (int)M(l);
        double* x = lv.d.cor;
        const double* b = lv.d.res;
        if (!std::is_same<T, double>::value) launch_copy_cast(lv);
        if (h->dense_A) {
            ael::k_bottom_bicgstab_dense<<<1, kBlock, sizeof(double) * 9 * Mb, s>>>(
                x, b, h->dense_A, lv.d.diag, Mb, h->omega, bottom_max_iter, bottom_tol, nuf, h->bottom_status);
            if (!std::is_same<T, double>::value)
                ael::k_cast<double, float><<<nblk(Mb), kBlock, 0, s>>>(lv.d.cor, lv.f.cor, Mb);
            return;
        }
        const Lvl LL = lv.L;
        switch (h->sym) {
            case 0:
                ael::k_bottom_solve<0><<<1, ael::kBottomBlock, 0, s>>>(x, b, lv.d.C, lv.d.diag, h->bct_dev, h->uniform,
                                                                       LL, h->omega, bottom_max_iter, bottom_tol, nuf,
                                                                       h->bottom_work, h->bottom_status);
                break;
            case 1:
                ael::k_bottom_solve<1><<<1, ael::kBottomBlock, 0, s>>>(x, b, lv.d.C, lv.d.diag, h->bct_dev, h->uniform,
                                                                       LL, h->omega, bottom_max_iter, bottom_tol, nuf,
                                                                       h->bottom_work, h->bottom_status);
                break;
            default:
                ael::k_bottom_solve<2><<<1, ael::kBottomBlock, 0, s>>>(x, b, lv.d.C, lv.d.diag, h->bct_dev, h->uniform,
                                                                       LL, h->omega, bottom_max_iter, bottom_tol, nuf,
                                                                       h->bottom_work, h->bottom_status);
                break;
        }
        if (!std::is_same<T, double>::value) {
            ael::k_cast<double, float><<<nblk(M(l)), kBlock, 0, s>>>(lv.d.cor, lv.f.cor, M(l));
        }
    }
    // FP32 hierarchy: the bottom residual goes to FP64 for the bottom solve.
    void launch_copy_cast(Level& lv) {
        const long m = 3 * lv.L.N;
        ael::k_cast<float, double><<<nblk(m), kBlock, 0, s>>>(lv.f.res, lv.d.res, m);
    }

    // amrex::MLMG::mgVcycle on AMR level 0 starting at MG level top.  With
    // top_fz (and fz_ready) the caller has written the first sweep from zero
    // of level top into its tmp.
    void vcycle(int top, bool top_fz = false) {
        const int bot = this->bot();
        const bool fz = fz_ready();
        for (int l = top; l < bot; ++l) {
            Vecs<T>& v = V(l);
            smooth_from_zero(l, v.cor, v.res, nu1, fz && (l > top || top_fz));
            corr_res(l, v.rescor, v.cor, v.res);
            if (fz && l + 1 < bot) restrict_res_fz(l, v.rescor);
            else restrict_res(l, v.rescor, V(l + 1).res);
        }
        bottom();
        for (int l = bot - 1; l >= top; --l) {
            Vecs<T>& v = V(l);
            interp(l, v.cor, V(l + 1).cor, 0);
            smooth(l, v.cor, v.res, nu2);
        }
    }

    // amrex::MLMG::mgFcycle.  The residual goes down by nodal injection
    // (MLLinOp::avgDownResMG -> amrex::average_down on nodal data).
    void fcycle() {
        const int bot = this->bot();
        if (bot + 1 > ael::kInjectMaxLevels) {
            for (int l = 1; l <= bot; ++l)
                ael::k_inject<T><<<nblk(L(l).L.N), kBlock, 0, s>>>(V(l - 1).res, L(l - 1).L, V(l).res, L(l).L, 3);
        } else if (bot >= 1) {  // every level's residual from level 0 in one launch
            ael::InjectLevels<T> il{};
            il.n = bot + 1;
            il.off[1] = 0;
            for (int l = 0; l <= bot; ++l) {
                il.f[l] = V(l).res;
                il.L[l] = L(l).L;
                if (l >= 1) il.off[l + 1] = il.off[l] + L(l).L.N;
            }
            ael::k_inject_all<T><<<nblk(il.off[bot + 1]), kBlock, 0, s>>>(il, 3);
        }
        bottom();
        for (int l = bot - 1; l >= 0; --l) {
            Vecs<T>& v = V(l);
            // the level below ran a V-cycle on top of its held correction:
            // the interpolation reads their sum (the sum is not needed after)
            if (l + 1 < bot)
                launch_interp3<T>(V(l + 1).cor, V(l + 1).cor_hold, L(l + 1).L, v.cor, L(l).L, 1, s);
            else
                interp(l, v.cor, V(l + 1).cor, 1);
            corr_res(l, v.rescor, v.cor, v.res);
            std::swap(v.res, v.rescor);
            std::swap(v.cor, v.cor_hold);
            vcycle(l);
            if (l == 0 && !leave_held) launch_axpy<T>(v.cor, T(1), v.cor_hold, M(l), s);
        }
    }

    void run(int cycle, bool top_fz = false) {
        if (cycle == ALAMO_EL_CYCLE_F) fcycle();
        else vcycle(0, top_fz);
    }
};

// The cycle of outer iteration it: MLMG runs F-cycles while it < max_fmg_iter
// (every iteration when max_fmg_iter < 0), V-cycles after; a V-cycle solve
// stays on V-cycles.
int cycle_of(const alamo_el_opts& o, int it) {
    return o.cycle == ALAMO_EL_CYCLE_F && (o.max_fmg_iter < 0 || it < o.max_fmg_iter) ? ALAMO_EL_CYCLE_F
                                                                                       : ALAMO_EL_CYCLE_V;
}
// Iteration limit: fixed_iter > 0 caps max_iter.
int iter_limit(const alamo_el_opts& o) { return o.fixed_iter > 0 ? std::min(o.max_iter, o.fixed_iter) : o.max_iter; }

int alloc(void** p, size_t bytes) {
    hipError_t e = hipMalloc(p, bytes);
    if (e != hipSuccess) return fail(ALAMO_EL_EHIP, std::string("hipMalloc: ") + hipGetErrorString(e));
    return 0;
}

#define ALLOC(ptr, count)                                                         \
    do {                                                                          \
        int rc_ = alloc((void**)&(ptr), sizeof(*(ptr)) * (size_t)(count));        \
        if (rc_) return rc_;                                                      \
    } while (0)

int ensure32(alamo_el_t h) {
    if (h->have32) return 0;
    for (auto& l : h->lv) {
        const long N = l.L.N;
        ALLOC(l.f.C, (long)h->nc * N);
        ALLOC(l.f.diag, 3 * N);
        ALLOC(l.f.cor, 3 * N);
        ALLOC(l.f.cor_hold, 3 * N);
        ALLOC(l.f.res, 3 * N);
        ALLOC(l.f.rescor, 3 * N);
        ALLOC(l.f.tmp, 3 * N);
    }
    h->have32 = true;
    return 0;
}

// Gauss-Jordan inverse with partial pivoting of a column-major M x M matrix;
// false when a pivot vanishes.
// Dense coarsest-level operator (A e_q for every unit vector), for the dense
// BiCGStab bottom.
int assemble_dense(alamo_el_t h, hipStream_t s) {
    Level& lb = h->lv.back();
    const int M = (int)(3 * lb.L.N);
    if (M > kDenseMax) return 0;
    if (h->dense_M != M) {
        for (double* p : {h->dense_I, h->dense_A})
            if (p) (void)hipFree(p);
        h->dense_I = h->dense_A = nullptr;
        ALLOC(h->dense_I, (long)M * M);
        ALLOC(h->dense_A, (long)M * M);
        ael::k_eye<<<nblk((long)M * M), kBlock, 0, s>>>(h->dense_I, M);
        h->dense_M = M;
    }
    dim3 grid(nblk(lb.L.N), M);
    switch (h->sym) {
        case 0: ael::k_assemble_dense<0><<<grid, kBlock, 0, s>>>(h->dense_I, lb.d.C, h->bct_dev, h->uniform, lb.L, h->dense_A); break;
        case 1: ael::k_assemble_dense<1><<<grid, kBlock, 0, s>>>(h->dense_I, lb.d.C, h->bct_dev, h->uniform, lb.L, h->dense_A); break;
        default: ael::k_assemble_dense<2><<<grid, kBlock, 0, s>>>(h->dense_I, lb.d.C, h->bct_dev, h->uniform, lb.L, h->dense_A); break;
    }
    psi_post_columns(h, lb, h->dense_I, h->dense_A, M, s);
    HIPCHK(hipGetLastError());
    return 0;
}

// The direct bottom level: the finest MG level with at most kDirectMax
// unknowns (-1 when even the coarsest has more).
// ALAMO_EL_DIRECT_MAX lowers the limit (read at every prepare): a caller whose
// model changes every step trades a few coarse sweeps per cycle for a smaller
// inversion (at 129^3: 375 unknowns on the 5^3 level instead of 2187 on 9^3).
int direct_level_of(alamo_el_t h) {
    long lim = kDirectMax;
    if (const char* e = std::getenv("ALAMO_EL_DIRECT_MAX")) lim = std::min<long>(lim, std::max(0L, std::atol(e)));
    int l = (int)h->lv.size() - 1;
    if (3 * h->lv[l].L.N > lim) return -1;
    while (l > 0 && 3 * h->lv[l - 1].L.N <= lim) --l;
    return l;
}

void free_direct(alamo_el_t h) {
    for (void* p : {(void*)h->direct_W, (void*)h->direct_Ainv, (void*)h->direct_perm, (void*)h->direct_map,
                    (void*)h->direct_part, (void*)h->direct_st, (void*)h->direct_A, (void*)h->direct_X,
                    (void*)h->direct_D, (void*)h->direct_chk})
        if (p) (void)hipFree(p);
    h->direct_W = h->direct_Ainv = h->direct_part = h->direct_A = h->direct_X = h->direct_D = h->direct_chk = nullptr;
    h->direct_perm = h->direct_map = nullptr;
    h->direct_st = nullptr;
    h->direct_M = 0;
}

// Inverse of the M x M operator in direct_W (destroyed) by block
// Gauss-Jordan without pivoting across blocks (inverse_kernels.h); *inv is
// the buffer holding it (direct_W or direct_Ainv), or null when a block
// pivot vanished or the check of the inverse against the kept operator
// (max |A Ainv x - x| / max |x| <= kInvCheckTol) failed.
constexpr double kInvCheckTol = 1e-8;
int invert_blocked(alamo_el_t h, int M, hipStream_t s, double** inv) {
    const ael::InvState zero{0, 0.0};
    HIPCHK(hipMemcpyAsync(h->direct_st, &zero, sizeof(zero), hipMemcpyHostToDevice, s));
    double *W = h->direct_W, *W2 = h->direct_Ainv;
    const unsigned gt = (unsigned)((M + ael::kInvTile - 1) / ael::kInvTile);
    for (int k0 = 0; k0 < M; k0 += ael::kInvBlock) {
        const int nb = std::min(ael::kInvBlock, M - k0);
        ael::k_inv_diag<<<1, kBlock, 0, s>>>(W, M, k0, nb, h->direct_D, h->direct_st);
        ael::k_inv_xrow<<<gt, kBlock, 0, s>>>(W, M, k0, nb, h->direct_D, h->direct_X);
        ael::k_inv_gemm_blk<<<dim3(gt, gt), kBlock, 0, s>>>(W, W2, M, k0, nb, h->direct_X);
        std::swap(W, W2);
    }
    // the check: z = A (W x)
    double *x = h->direct_chk, *y = x + M, *z = y + M;
    const dim3 grid(nblk(M), (M + ael::kDenseMvChunk - 1) / ael::kDenseMvChunk);
    ael::k_inv_check_x<<<nblk(M), kBlock, 0, s>>>(x, M);
    ael::k_dense_mv_partial<double, double><<<grid, kBlock, 0, s>>>(W, x, M, h->direct_part);
    ael::k_dense_mv_final<double><<<nblk(M), kBlock, 0, s>>>(h->direct_part, (int)grid.y, M, y);
    ael::k_dense_mv_partial<double, double><<<grid, kBlock, 0, s>>>(h->direct_A, y, M, h->direct_part);
    ael::k_dense_mv_final<double><<<nblk(M), kBlock, 0, s>>>(h->direct_part, (int)grid.y, M, z);
    ael::k_inv_check<<<1, kBlock, 0, s>>>(z, M, h->direct_st);
    HIPCHK(hipGetLastError());
    ael::InvState st{};
    HIPCHK(hipMemcpyAsync(&st, h->direct_st, sizeof(st), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    *inv = (!st.singular && st.resid <= kInvCheckTol) ? W : nullptr;
    return 0;
}

// Inverse of the M x M operator in direct_W (destroyed) by blocked
// Gauss-Jordan with partial pivoting over all rows (inverse_kernels.h); *inv
// as above, null when the operator is singular.
int invert_pivoted(alamo_el_t h, int M, hipStream_t s, double** inv) {
    const ael::InvState zero{0, 0.0};
    HIPCHK(hipMemcpyAsync(h->direct_st, &zero, sizeof(zero), hipMemcpyHostToDevice, s));
    // blocks of kInvBlock columns: panels of pw pivot steps, each updating
    // the rest of its block, then the block's update of every other column,
    // out of place (W and the inverse's buffer alternate)
    double *W = h->direct_W, *W2 = h->direct_Ainv;
    const unsigned gt = (unsigned)((M + ael::kInvTile - 1) / ael::kInvTile);
    const int pw = ael::inv_panel_width(M);
    for (int k0 = 0; k0 < M; k0 += ael::kInvBlock) {
        const int nbk = std::min(ael::kInvBlock, M - k0), b1 = k0 + nbk;
        for (int c0 = k0; c0 < b1; c0 += pw) {
            const int nb = std::min(pw, b1 - c0);
            ael::launch_inv_panel(W, M, c0, nb, h->direct_perm, h->direct_st, s);
            if (nbk > nb)
                ael::k_inv_block<<<(unsigned)(nbk - nb), kBlock, 0, s>>>(W, M, k0, b1, c0, nb, h->direct_perm);
        }
        if (M > nbk) {
            ael::k_inv_rowmap<<<1, kBlock, 0, s>>>(h->direct_perm, M, k0, nbk, h->direct_map);
            ael::k_inv_gemm<<<dim3(gt, gt), kBlock, 0, s>>>(W, W2, M, k0, nbk, h->direct_map);
            std::swap(W, W2);
        }
    }
    HIPCHK(hipGetLastError());
    ael::InvState st{};
    std::vector<int> perm(M), src(M);
    HIPCHK(hipMemcpyAsync(&st, h->direct_st, sizeof(st), hipMemcpyDeviceToHost, s));
    HIPCHK(hipMemcpyAsync(perm.data(), h->direct_perm, sizeof(int) * M, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    if (st.singular) {
        *inv = nullptr;
        return 0;
    }
    // W holds the inverse with its columns permuted: undo the swaps from the
    // last step to the first (column c of the inverse is column src[c] of W)
    for (int c = 0; c < M; ++c) src[c] = c;
    for (int k = M - 1; k >= 0; --k) std::swap(src[k], src[perm[k]]);
    HIPCHK(hipMemcpyAsync(h->direct_map, src.data(), sizeof(int) * M, hipMemcpyHostToDevice, s));
    ael::k_inv_unpermute<<<nblk((long)M * M), kBlock, 0, s>>>(W, M, h->direct_map, W2);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(s));  // src lives on the host stack
    *inv = W2;
    return 0;
}

// Operator of the direct level (A e_q for every unit vector) and its FP64
// inverse by blocked Gauss-Jordan with partial pivoting on the device
// (inverse_kernels.h).  A singular operator leaves direct_Ainv null (the
// direct bottom is then unavailable).
int assemble_direct(alamo_el_t h, hipStream_t s) {
    const int l = direct_level_of(h);
    h->direct_level = l;
    if (l < 0) {
        free_direct(h);
        return 0;
    }
    Level& lb = h->lv[l];
    const int M = (int)(3 * lb.L.N);
    if (h->direct_M != M || !h->direct_Ainv) {
        free_direct(h);
        ALLOC(h->direct_W, (long)M * M);
        ALLOC(h->direct_Ainv, (long)M * M);
        ALLOC(h->direct_perm, M);
        ALLOC(h->direct_map, M);
        ALLOC(h->direct_part, (long)((M + ael::kDenseMvChunk - 1) / ael::kDenseMvChunk) * M);
        ALLOC(h->direct_st, 1);
        ALLOC(h->direct_A, (long)M * M);
        ALLOC(h->direct_X, (long)ael::kInvBlock * M);
        ALLOC(h->direct_D, (long)ael::kInvBlock * ael::kInvBlock);
        ALLOC(h->direct_chk, 3L * M);
        h->direct_M = M;
    }
    // the unit vectors (in the buffer the inverse will take), the operator's columns into W
    const long MM = (long)M * M;
    ael::k_eye<<<nblk(MM), kBlock, 0, s>>>(h->direct_Ainv, M);
    dim3 grid(nblk(lb.L.N), M);
    switch (h->sym) {
        case 0: ael::k_assemble_dense<0><<<grid, kBlock, 0, s>>>(h->direct_Ainv, lb.d.C, h->bct_dev, h->uniform, lb.L, h->direct_W); break;
        case 1: ael::k_assemble_dense<1><<<grid, kBlock, 0, s>>>(h->direct_Ainv, lb.d.C, h->bct_dev, h->uniform, lb.L, h->direct_W); break;
        default: ael::k_assemble_dense<2><<<grid, kBlock, 0, s>>>(h->direct_Ainv, lb.d.C, h->bct_dev, h->uniform, lb.L, h->direct_W); break;
    }
    psi_post_columns(h, lb, h->direct_Ainv, h->direct_W, M, s);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(h->direct_A, h->direct_W, sizeof(double) * MM, hipMemcpyDeviceToDevice, s));
    double* inv = nullptr;
    const char* how = std::getenv("ALAMO_EL_INVERSE");  // "pivoted": skip the block path
    int rc = how && std::strcmp(how, "pivoted") == 0 ? 0 : invert_blocked(h, M, s, &inv);
    if (rc) return rc;
    h->direct_pivoted = 0;
    if (!inv) {  // a vanishing or inaccurate block pivot: the pivoted panels, from the kept operator
        HIPCHK(hipMemcpyAsync(h->direct_W, h->direct_A, sizeof(double) * MM, hipMemcpyDeviceToDevice, s));
        rc = invert_pivoted(h, M, s, &inv);
        if (rc) return rc;
        h->direct_pivoted = 1;
    }
    if (!inv) {  // singular
        (void)hipFree(h->direct_Ainv);
        h->direct_Ainv = nullptr;
        return 0;
    }
    // the bottom reads the inverse by rows: store it row-major in the other buffer
    double* other = inv == h->direct_W ? h->direct_Ainv : h->direct_W;
    const unsigned gt = (unsigned)((M + 63) / 64);
    ael::k_transpose<<<dim3(gt, gt), kBlock, 0, s>>>(inv, M, other);
    HIPCHK(hipGetLastError());
    h->direct_Ainv = other;
    h->direct_W = inv;
    return 0;
}

// Compact grad-C matrices of one level (MajorMinor): per-wave counts of the
// flagged interior nodes, their exclusive prefix (on the host, once per
// model), then D of each flagged node at its position.
// Levels whose interior is at least kTiledMin nodes along every axis use the
// LDS-tiled MajorMinor kernel (unless ALAMO_EL_STENCIL=slab-only).
constexpr int kTiledMin = 32;

// The palette of a tiled level (ael::k_pal_index): the distinct 21-slot
// tensors of its nodes when there are at most ael::kPalMax of them and at most
// one per 16 nodes (a table near the size of C itself would only add a
// gather), found by hashing every node on the device and counting distinct
// hashes on the host; every interior node's row is then checked slot by slot
// against its own values, and any mismatch leaves the level on the C copy.
// ALAMO_EL_PALETTE=0 turns the palette off (cross-checks).
int build_palette(alamo_el_t h, Level& l, hipStream_t s) {
    const char* e = std::getenv("ALAMO_EL_PALETTE");
    if (e && std::strcmp(e, "0") == 0) return 0;
    const long N = l.L.N;
    const long limit = std::min<long>(ael::kPalMax, N / 16);
    // hashes in the level's tmp vector (3 N doubles), the runs' heads in its
    // rescor vector: the host reads the heads only (a polycrystal's rows cross
    // a few grains each), not a hash per node
    unsigned long long* hs = reinterpret_cast<unsigned long long*>(l.d.tmp);
    unsigned long long* heads = reinterpret_cast<unsigned long long*>(l.d.rescor);
    unsigned int* cnt_d = reinterpret_cast<unsigned int*>(h->red_out);
    const int cap = (int)std::min<long>(3 * N / 2, 1L << 20);
    ael::k_c_hash<<<nblk(N), kBlock, 0, s>>>(l.d.C, N, hs);
    unsigned int cnt = 0;
    HIPCHK(hipMemcpyAsync(cnt_d, &cnt, sizeof(cnt), hipMemcpyHostToDevice, s));
    ael::k_run_heads<<<nblk(N), kBlock, 0, s>>>(hs, N, cap, heads, cnt_d);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(&cnt, cnt_d, sizeof(cnt), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    int rc = 0;
    std::unordered_map<unsigned long long, long> first;  // hash -> first node
    if ((long)cnt <= cap) {
        std::vector<unsigned long long> hd(2 * (size_t)cnt);
        HIPCHK(hipMemcpyAsync(hd.data(), heads, sizeof(unsigned long long) * 2 * cnt, hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
        for (unsigned int q = 0; q < cnt && (long)first.size() <= limit; ++q) {
            auto it = first.emplace(hd[2 * q], (long)hd[2 * q + 1]).first;
            it->second = std::min<long>(it->second, (long)hd[2 * q + 1]);
        }
    } else {  // more runs than the heads buffer holds: read every hash
        std::vector<unsigned long long> hh(N);
        HIPCHK(hipMemcpyAsync(hh.data(), hs, sizeof(unsigned long long) * N, hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
        unsigned long long prev = 0;
        for (long id = 0; id < N && (long)first.size() <= limit; ++id) {
            if (id > 0 && hh[id] == prev) continue;
            prev = hh[id];
            first.emplace(prev, id);
        }
    }
    if ((long)first.size() > limit) return 0;
    std::vector<std::pair<unsigned long long, long>> rows(first.begin(), first.end());
    std::sort(rows.begin(), rows.end());
    const int npal = (int)rows.size();
    std::vector<unsigned long long> ph(npal);
    std::vector<long> rep(npal);
    for (int q = 0; q < npal; ++q) {
        ph[q] = rows[q].first;
        rep[q] = rows[q].second;
    }
    unsigned long long* ph_d = nullptr;
    long* rep_d = nullptr;
    int* bad_d = nullptr;
    auto done = [&](int r) {
        for (void* p : {(void*)ph_d, (void*)rep_d, (void*)bad_d})
            if (p) (void)hipFree(p);
        if (r) {  // an error leaves the level on its C copy, with nothing half built
            if (l.d.pal) (void)hipFree(l.d.pal);
            if (l.pidx) (void)hipFree(l.pidx);
            if (l.pidx_b) (void)hipFree(l.pidx_b);
            l.d.pal = nullptr;
            l.pidx = nullptr;
            l.pidx_b = nullptr;
        }
        return r;
    };
    if (hipMalloc((void**)&ph_d, sizeof(unsigned long long) * npal) != hipSuccess ||
        hipMalloc((void**)&rep_d, sizeof(long) * npal) != hipSuccess ||
        hipMalloc((void**)&bad_d, sizeof(int)) != hipSuccess ||
        hipMalloc((void**)&l.d.pal, sizeof(double) * npal * ael::kPalStride) != hipSuccess ||
        hipMalloc((void**)&l.pidx, sizeof(ael::PalIdx) * ael::TileIdx<TileT>::elems(l.L, l.tg)) != hipSuccess ||
        hipMalloc((void**)&l.pidx_b, sizeof(ael::PalIdx) * std::max(1L, ael::boundary_count(l.L))) != hipSuccess)
        return done(fail(ALAMO_EL_EHIP, "palette allocation failed"));
    int bad = 0;
    if (hipMemcpyAsync(ph_d, ph.data(), sizeof(unsigned long long) * npal, hipMemcpyHostToDevice, s) != hipSuccess ||
        hipMemcpyAsync(rep_d, rep.data(), sizeof(long) * npal, hipMemcpyHostToDevice, s) != hipSuccess ||
        hipMemcpyAsync(bad_d, &bad, sizeof(int), hipMemcpyHostToDevice, s) != hipSuccess)
        return done(fail(ALAMO_EL_EHIP, "palette copy failed"));
    ael::k_pal_gather<double><<<nblk((long)npal * ael::kPalStride), kBlock, 0, s>>>(l.d.C, N, rep_d, npal, l.d.pal);
    ael::k_pal_index<TileT><<<(unsigned)ael::tile_planes(l.L, l.tg), TileT::NT, 0, s>>>(
        l.d.C, l.L, l.tg, hs, ph_d, npal, l.d.pal, l.pidx, bad_d);
    if (ael::boundary_count(l.L) > 0)
        ael::k_pal_index_b<<<nblk(ael::boundary_count(l.L)), kBlock, 0, s>>>(l.d.C, l.L, hs, ph_d, npal, l.d.pal,
                                                                             l.pidx_b, bad_d);
    if (hipGetLastError() != hipSuccess ||
        hipMemcpyAsync(&bad, bad_d, sizeof(int), hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess)
        return done(fail(ALAMO_EL_EHIP, "palette build failed"));
    if (bad) {  // a hash collision: keep the C copy
        (void)hipFree(l.d.pal);
        (void)hipFree(l.pidx);
        (void)hipFree(l.pidx_b);
        l.d.pal = nullptr;
        l.pidx = nullptr;
        l.pidx_b = nullptr;
        return done(0);
    }
    l.npal = npal;
    return done(rc);
}

int build_gradc(alamo_el_t h, Level& l, hipStream_t s) {
    if (h->sym != 1) return 0;
    l.tiled = stencil_variant() == kVariantSlab && l.L.nx - 2 >= kTiledMin && l.L.ny - 2 >= kTiledMin &&
              l.L.nz - 2 >= kTiledMin;
    static_assert(TileT::NT == kBlock, "tiled and slab launches share the block size");
    l.tg = tile_grid_of(l.L, kTileTarget64);
    l.tg32 = tile_grid_of(l.L, kTileTarget32);
    const int wpb = ael::kWavesPerBlock;
    // count/fill workgroups: one per (tile, plane) on tiled levels (the compact
    // D order of both tile grids), else the slab launch's interior ones
    const unsigned nib = l.tiled ? (unsigned)ael::tile_planes(l.L, l.tg)
                                 : ael::slab_blocks(l.L, kBlock) - ael::slab_boundary_blocks(l.L, kBlock);
    const long nw = (long)nib * wpb;
    // The tile copies and dwbase depend on the grid only and stay allocated
    // from one model to the next (their contents are rebuilt here and, for the
    // FP32 copies, by prepare32); the compact D arrays grow when needed.
    for (void* p : {(void*)l.d.pal, (void*)l.f.pal, (void*)l.pidx, (void*)l.pidx_b})
        if (p) (void)hipFree(p);
    l.d.pal = nullptr;
    l.f.pal = nullptr;
    l.pidx = nullptr;
    l.pidx_b = nullptr;
    l.npal = 0;
    l.nflag = 0;
    if (!l.tiled) {
        for (void* p : {(void*)l.d.Ct, (void*)l.d.Cb, (void*)l.f.Ct, (void*)l.f.Cb})
            if (p) (void)hipFree(p);
        l.d.Ct = l.d.Cb = nullptr;
        l.f.Ct = l.f.Cb = nullptr;
    }
    if (l.tiled) {
        const long nb = ael::boundary_count(l.L);
        if (!l.d.Ct) HIPCHK(hipMalloc((void**)&l.d.Ct, sizeof(double) * ael::TileC<TileT>::elems(l.L, l.tg)));
        if (!l.d.Cb) HIPCHK(hipMalloc((void**)&l.d.Cb, sizeof(double) * h->nc * nb));
        ael::k_pack_tile_c<TileT, double><<<(unsigned)ael::tile_planes(l.L, l.tg), TileT::NT, 0, s>>>(l.d.C, l.L, l.tg,
                                                                                                   l.d.Ct);
        ael::k_pack_bnd_c<double><<<nblk(nb), kBlock, 0, s>>>(l.d.C, l.L, h->nc, l.d.Cb);
        HIPCHK(hipGetLastError());
        int rc = build_palette(h, l, s);
        if (rc) return rc;
    }
    if (nw == 0) return 0;
    if (!l.dwbase || l.nwcap < nw) {
        if (l.dwbase) (void)hipFree(l.dwbase);
        l.dwbase = nullptr;
        l.nwcap = 0;
        HIPCHK(hipMalloc((void**)&l.dwbase, sizeof(int) * nw));
        l.nwcap = nw;
    }
    if (l.tiled)
        ael::k_gradc_count_tiles<TileT><<<nib, kBlock, 0, s>>>(l.gmask, l.L, l.tg, l.dwbase);
    else
        ael::k_gradc_count<<<nib, kBlock, 0, s>>>(l.gmask, l.L, l.dwbase);
    HIPCHK(hipGetLastError());
    std::vector<int> cnt(nw);
    HIPCHK(hipMemcpyAsync(cnt.data(), l.dwbase, sizeof(int) * nw, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    long tot = 0;
    for (long w = 0; w < nw; ++w) {
        const int c = cnt[w];
        cnt[w] = (int)tot;
        tot += c;
    }
    HIPCHK(hipMemcpyAsync(l.dwbase, cnt.data(), sizeof(int) * nw, hipMemcpyHostToDevice, s));
    l.nflag = (int)tot;
    if (tot > 0) {
        if (!l.d.D || l.dcap < tot) {
            if (l.d.D) (void)hipFree(l.d.D);
            l.d.D = nullptr;
            l.dcap = 0;
            HIPCHK(hipMalloc((void**)&l.d.D, sizeof(double) * 18 * tot));
            l.dcap = tot;
        }
        if (l.tiled)
            ael::k_gradc_fill_tiles<TileT><<<nib, kBlock, 0, s>>>(l.d.C, l.gmask, l.L, l.tg, l.dwbase, l.nflag, l.d.D);
        else
            ael::k_gradc_fill<<<nib, kBlock, 0, s>>>(l.d.C, l.gmask, l.L, l.dwbase, l.nflag, l.d.D);
        HIPCHK(hipGetLastError());
    }
    HIPCHK(hipStreamSynchronize(s));  // cnt is read by the copy above
    return 0;
}

int build_gmask(alamo_el_t h, Level& l, hipStream_t s) {
    switch (h->sym) {
        case 0: ael::k_grad_mask<2><<<nblk(l.L.N), kBlock, 0, s>>>(l.d.C, l.L, l.gmask); break;
        case 1: ael::k_grad_mask<21><<<nblk(l.L.N), kBlock, 0, s>>>(l.d.C, l.L, l.gmask); break;
        default: ael::k_grad_mask<9><<<nblk(l.L.N), kBlock, 0, s>>>(l.d.C, l.L, l.gmask); break;
    }
    HIPCHK(hipGetLastError());
    return build_gradc(h, l, s);
}

int prepare(alamo_el_t h, hipStream_t s) {
    if (!h->model_set) return fail(ALAMO_EL_ESTATE, "Attempting to use operator before calling SetModel");
    // averageDownCoeffsSameAmrLevel (src/Operator/Elastic.cpp:746-910)
    for (size_t l = 1; l < h->lv.size(); ++l) {
        Level& c = h->lv[l];
        Level& f = h->lv[l - 1];
        ael::k_restrict<double><<<nblk8(c.L.N), kBlock, 0, s>>>(f.d.C, f.L, c.d.C, c.L, h->nc);
        int rc = build_gmask(h, c, s);
        if (rc) return rc;
        if (h->psi_set) {
            // (Elastic.cpp:839-907) every coarse cell covers two fine cells per axis
            if (f.L.nx - 1 != 2 * (c.L.nx - 1) || f.L.ny - 1 != 2 * (c.L.ny - 1) || f.L.nz - 1 != 2 * (c.L.nz - 1))
                return fail(ALAMO_EL_ESTATE, "psi restriction needs exact coarsening by 2");
            const long nc = (long)(c.L.nx - 1) * (c.L.ny - 1) * (c.L.nz - 1);
            ael::k_psi_restrict<<<nblk(nc), kBlock, 0, s>>>(f.psi, f.L, c.psi, c.L);
        }
    }
    // Diagonal(true) on every MG level (src/Operator/Operator.cpp:16-29)
    for (auto& l : h->lv) launch_diagonal<double>(h, l, l.d.C, l.d.diag, s);
    HIPCHK(hipGetLastError());
    int rc = assemble_dense(h, s);
    if (rc) return rc;
    rc = assemble_direct(h, s);
    if (rc) return rc;
    h->prepared = true;
    h->prepared32 = false;
    return 0;
}

int prepare32(alamo_el_t h, hipStream_t s) {
    if (h->prepared32) return 0;
    int rc = ensure32(h);
    if (rc) return rc;
    for (auto& l : h->lv) {
        const long N = l.L.N;
        ael::k_cast<double, float><<<nblk((long)h->nc * N), kBlock, 0, s>>>(l.d.C, l.f.C, (long)h->nc * N);
        ael::k_cast<double, float><<<nblk(3 * N), kBlock, 0, s>>>(l.d.diag, l.f.diag, 3 * N);
        if (l.d.Ct) {
            const long nb = ael::boundary_count(l.L);
            if (!l.f.Ct) HIPCHK(hipMalloc((void**)&l.f.Ct, sizeof(float) * ael::TileC<TileT>::elems(l.L, l.tg)));
            if (!l.f.Cb) HIPCHK(hipMalloc((void**)&l.f.Cb, sizeof(float) * h->nc * nb));
            ael::k_pack_tile_c<TileT, float><<<(unsigned)ael::tile_planes(l.L, l.tg), TileT::NT, 0, s>>>(l.d.C, l.L,
                                                                                                     l.tg, l.f.Ct);
            ael::k_pack_bnd_c<float><<<nblk(nb), kBlock, 0, s>>>(l.d.C, l.L, h->nc, l.f.Cb);
        }
        if (l.f.pal) (void)hipFree(l.f.pal);
        l.f.pal = nullptr;
        if (l.d.pal) {  // the FP32 table: the same rows rounded as the C copy rounds them
            const long np = (long)l.npal * ael::kPalStride;
            HIPCHK(hipMalloc((void**)&l.f.pal, sizeof(float) * np));
            ael::k_cast<double, float><<<nblk(np), kBlock, 0, s>>>(l.d.pal, l.f.pal, np);
        }
        if (l.d.D && l.nflag > 0) {
            if (!l.f.D || l.fcap < l.nflag) {
                if (l.f.D) (void)hipFree(l.f.D);
                l.f.D = nullptr;
                l.fcap = 0;
                HIPCHK(hipMalloc((void**)&l.f.D, sizeof(float) * 18 * (long)l.nflag));
                l.fcap = l.nflag;
            }
            ael::k_cast<double, float><<<nblk(18L * l.nflag), kBlock, 0, s>>>(l.d.D, l.f.D, 18L * l.nflag);
        }
    }
    HIPCHK(hipGetLastError());
    h->prepared32 = true;
    return 0;
}

bool bad_level(alamo_el_t h, int l) { return !h || l < 0 || l >= (int)h->lv.size(); }

// ---------------------------------------------------------------------------
// Learned coarse-level solver (no counterpart in the reference)
// ---------------------------------------------------------------------------
// Level-L right-hand side of a fine-level b: the interior rows restricted by
// the MG full weighting (src/Operator/Operator.cpp:262-363) level by level,
// the boundary rows injected from the fine node each coarse node sits on.
int coarse_rhs(alamo_el_t h, int L, const double* b, double* bc, hipStream_t s) {
    Level& l0 = h->lv[0];
    if (L == 0) {
        launch_copy<double>(bc, b, 3 * l0.L.N, s);
        HIPCHK(hipGetLastError());
        return 0;
    }
    ael::k_interior_only<<<nblk(l0.L.N), kBlock, 0, s>>>(b, l0.L, l0.d.tmp);
    const double* src = l0.d.tmp;
    for (int l = 1; l <= L; ++l) {
        double* dst = l == L ? bc : h->lv[l].d.tmp;
        ael::k_restrict<double><<<nblk8(h->lv[l].L.N), kBlock, 0, s>>>(src, h->lv[l - 1].L, dst, h->lv[l].L, 3);
        src = dst;
    }
    ael::k_inject_boundary<<<nblk(h->lv[L].L.N), kBlock, 0, s>>>(b, l0.L, 1 << L, bc, h->lv[L].L);
    HIPCHK(hipGetLastError());
    return 0;
}

// Network inputs beside the nc stiffness slots (learned_kernels.h k_features).
constexpr int kLearnedExtra = 30;

int ensure_nn_scratch(alamo_el_t h, long N) {
    if (h->nn_cap >= N) return 0;
    if (h->nn_t) (void)hipFree(h->nn_t);
    h->nn_t = nullptr;
    h->nn_cap = 0;
    HIPCHK(hipMalloc((void**)&h->nn_t, sizeof(double) * 3 * N));
    if (!h->nn_fmean) HIPCHK(hipMalloc((void**)&h->nn_fmean, sizeof(double) * 18));
    h->nn_cap = N;
    return 0;
}

// Network input of level L from its right-hand side bc (learned_kernels.h);
// *scale = s, 0 when bc is zero.
int learned_input(alamo_el_t h, int L, const double* bc, float* feat, double* scale, hipStream_t s) {
    Level& l = h->lv[L];
    const long N = l.L.N;
    int rc;
    if ((rc = ensure_nn_scratch(h, N))) return rc;
    double csum;
    if ((rc = reduce_host<double>(h, l.d.C, l.d.C, N, 2, &csum, s))) return rc;
    const double alpha = csum / (double)N;
    if (!(alpha > 0)) return fail(ALAMO_EL_EINVAL, "learned input: the mean of stiffness slot 0 is not positive");
    ael::k_prescale<<<nblk(N), kBlock, 0, s>>>(bc, l.L, 1.0 / alpha, h->nn_t);
    double smax;
    if ((rc = reduce_host<double>(h, h->nn_t, h->nn_t, 3 * N, 0, &smax, s))) return rc;
    const double sc = smax > 0 ? smax : 1.0;
    ael::k_face_means<<<18, kBlock, 0, s>>>(h->nn_t, l.L, h->nn_fmean);
    ael::k_features<<<nblk(N), kBlock, 0, s>>>(l.d.C, h->nc, h->nn_t, h->bct_dev, h->nn_fmean, l.L, 1.0 / alpha,
                                                1.0 / sc, feat);
    HIPCHK(hipGetLastError());
    *scale = smax;
    return 0;
}

// x = P^L (s * net(input(coarse_rhs(b)))), trilinear interpolation level by
// level (src/Operator/Operator.cpp:365-434); x = 0 when b has no level-L part.
int learned_apply(alamo_el_t h, const double* b, double* x, hipStream_t s) {
    const int L = h->nn_level;
    Level& lL = h->lv[L];
    const long NL = lL.L.N;
    int rc;
    if ((rc = coarse_rhs(h, L, b, h->nn_bc, s))) return rc;
    double sc;
    if ((rc = learned_input(h, L, h->nn_bc, h->nn_in, &sc, s))) return rc;
    if (sc == 0.0) {
        launch_fill<double>(x, 0.0, 3 * h->lv[0].L.N, s);
        HIPCHK(hipGetLastError());
        return 0;
    }
    if ((rc = alamo_nn_forward(h->nn, h->nn_in, h->nn_out, s))) return rc;
    ael::k_scale_output<<<nblk(3 * NL), kBlock, 0, s>>>(h->nn_out, 3 * NL, sc, L == 0 ? x : lL.d.cor);
    for (int l = L - 1; l >= 0; --l) {
        double* dst = l == 0 ? x : h->lv[l].d.cor;
        launch_interp3<double>(h->lv[l + 1].d.cor, nullptr, h->lv[l + 1].L, dst, h->lv[l].L, 1, s);
    }
    HIPCHK(hipGetLastError());
    return 0;
}

constexpr int kLearnedSmooth = 2;  // Fsmooth calls applied to a learned correction

// One judged learned correction.  On entry l0.d.res = b - A x and *rnorm is
// its inf-norm.  The network proposes e = P^L N(R^L r), smoothed by
// kLearnedSmooth Fsmooth calls on A e = r; the step length minimizes the
// 2-norm of r - g A e, and the step is taken only when it lowers |r|_inf (the
// multigrid residual is the judge).  *taken tells which.
int learned_step(alamo_el_t h, double* x, double* rnorm, int* taken, bool fp32, hipStream_t s) {
    Level& l0 = h->lv[0];
    const long M0 = 3 * l0.L.N;
    double* r = l0.d.res;
    double* e = l0.d.cor_hold;
    double* w = l0.d.rescor;
    int rc;
    *taken = 0;
    if ((rc = learned_apply(h, r, e, s))) return rc;
    if (fp32) {
        // the smoothing of the proposal in the FP32 cycles' precision (the
        // FP64 residual below judges the e it leaves)
        ael::k_cast<double, float><<<nblk(M0), kBlock, 0, s>>>(r, l0.f.res, M0);
        ael::k_cast<double, float><<<nblk(M0), kBlock, 0, s>>>(e, l0.f.cor, M0);
        for (int q = 0; q < kLearnedSmooth; ++q) {
            launch_jacobi<float>(h, l0, l0.f.cor, l0.f.res, l0.f.diag, l0.f.C, l0.f.tmp, s);
            launch_jacobi<float>(h, l0, l0.f.tmp, l0.f.res, l0.f.diag, l0.f.C, l0.f.cor, s);
        }
        ael::k_cast<float, double><<<nblk(M0), kBlock, 0, s>>>(l0.f.cor, e, M0);
    } else {
        for (int q = 0; q < kLearnedSmooth; ++q) {
            launch_jacobi<double>(h, l0, e, r, l0.d.diag, l0.d.C, l0.d.tmp, s);
            launch_jacobi<double>(h, l0, l0.d.tmp, r, l0.d.diag, l0.d.C, e, s);
        }
    }
    launch_fapply<double>(h, l0, e, l0.d.C, w, s);
    // r . w and w . w in one pass, one copy to the host
    const int nb = (int)std::min<long>(kRedBlocks, std::max<long>(1, nblk(M0)));
    ael::k_dot2_partial<<<nb, kBlock, 0, s>>>(r, w, M0, h->red_partial);
    ael::k_multidot_final<<<2, kBlock, 0, s>>>(h->red_partial, nb, h->red_out);
    HIPCHK(hipMemcpyAsync(h->red_host, h->red_out, sizeof(double) * 2, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    const double rw = h->red_host[0], ww = h->red_host[1];
    if (!(ww > 0.0) || std::isnan(rw)) return 0;
    const double g = rw / ww;
    // w = r - g A e, the residual after the step, and its max-norm
    ael::k_step_residual<<<nb, kBlock, 0, s>>>(w, g, r, M0, h->red_partial);
    ael::k_reduce_final<<<1, kBlock, 0, s>>>(h->red_partial, nb, 0, h->red_out);
    HIPCHK(hipMemcpyAsync(h->red_host, h->red_out, sizeof(double), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    const double rn = h->red_host[0];
    if (!(rn < *rnorm)) return 0;
    ael::k_take_step<<<nblk(M0), kBlock, 0, s>>>(x, g, e, r, w, M0);
    HIPCHK(hipGetLastError());
    *rnorm = rn;
    *taken = 1;
    return 0;
}

// Flexible GMRES(m) with one multigrid cycle as right preconditioner; the
// stopping test is the inf-norm of the true FP64 residual.
int fgmres_solve(alamo_el_t h, double* x, const double* b, double target, const alamo_el_opts& o, int& iters,
                 double& resnorm, int& learned_taken, hipStream_t s);

}  // namespace

// ---------------------------------------------------------------------------
// C API
// ---------------------------------------------------------------------------
extern "C" {

const char* alamo_el_last_error(void) { return g_err.c_str(); }

void alamo_el_default_opts(alamo_el_opts* o) {
    if (!o) return;
    o->max_iter = 200;
    o->nu1 = 2;
    o->nu2 = 2;
    o->cycle = ALAMO_EL_CYCLE_F;
    o->bottom_max_iter = 200;
    o->final_smooth = 8;
    o->omega = 2.0 / 3.0;
    o->bottom_tol_rel = 1e-4;
    o->precision = ALAMO_EL_PREC_FP64;
    o->learned = 0;
    o->outer = ALAMO_EL_OUTER_MLMG;
    o->restart = 30;
    o->verbose = 0;
    o->bottom_solver = ALAMO_EL_BOTTOM_BICGSTAB;
    o->smoother = ALAMO_EL_SMOOTH_JACOBI;
    o->max_fmg_iter = -1;
    o->fixed_iter = -1;
}

int alamo_el_create(const int n_cell[3], const double dx[3], int sym, int max_coarsening_level, alamo_el_t* out) {
    if (!n_cell || !dx || !out) return fail(ALAMO_EL_EINVAL, "null argument");
    if (ncomp_of(sym) < 0) return fail(ALAMO_EL_EINVAL, "unknown stiffness symmetry");
    for (int a = 0; a < 3; ++a)
        if (n_cell[a] < 2 || !(dx[a] > 0)) return fail(ALAMO_EL_EINVAL, "n_cell must be >= 2 and dx > 0");
    // The stencil kernels address each field component with a 32-bit byte offset.
    if ((long)(n_cell[0] + 1) * (n_cell[1] + 1) * (n_cell[2] + 1) >= (1L << 29))
        return fail(ALAMO_EL_EINVAL, "box too large for one handle (at most 2^29 nodes)");
    alamo_el_t h = new alamo_el_s();
    *out = nullptr;
    h->sym = sym;
    h->nc = ncomp_of(sym);
    int nc[3] = {n_cell[0], n_cell[1], n_cell[2]};
    double d[3] = {dx[0], dx[1], dx[2]};
    for (;;) {
        Level l;
        l.L.nx = nc[0] + 1;
        l.L.ny = nc[1] + 1;
        l.L.nz = nc[2] + 1;
        l.L.N = (long)l.L.nx * l.L.ny * l.L.nz;
        l.L.dx = d[0];
        l.L.dy = d[1];
        l.L.dz = d[2];
        l.L.per = 0;
        h->lv.push_back(l);
        if (max_coarsening_level >= 0 && (int)h->lv.size() > max_coarsening_level) break;
        bool ok = true;
        for (int a = 0; a < 3; ++a)
            if (nc[a] % 2 != 0 || nc[a] / 2 < 2) ok = false;
        if (!ok) break;
        for (int a = 0; a < 3; ++a) { nc[a] /= 2; d[a] *= 2.0; }
    }
    auto cleanup = [&](int rc) { alamo_el_destroy(h); return rc; };
    for (auto& l : h->lv) {
        const long N = l.L.N;
        int rc;
        if ((rc = alloc((void**)&l.d.C, sizeof(double) * h->nc * N))) return cleanup(rc);
        if ((rc = alloc((void**)&l.gmask, N))) return cleanup(rc);
        for (double** p : {&l.d.diag, &l.d.cor, &l.d.cor_hold, &l.d.res, &l.d.rescor, &l.d.tmp})
            if ((rc = alloc((void**)p, sizeof(double) * 3 * N))) return cleanup(rc);
    }
    const long Mb = 3 * h->lv.back().L.N;
    int rc;
    if ((rc = alloc((void**)&h->bct_dev, sizeof(int) * 78))) return cleanup(rc);
    if ((rc = alloc((void**)&h->bcv_dev, sizeof(double) * 78))) return cleanup(rc);
    // two slots of block partials (k_dot2_partial), kRedBlocks each
    if ((rc = alloc((void**)&h->red_partial, sizeof(double) * 2 * kRedBlocks))) return cleanup(rc);
    if ((rc = alloc((void**)&h->red_out, sizeof(double) * 2))) return cleanup(rc);
    if ((rc = alloc((void**)&h->bottom_work, sizeof(double) * 10 * Mb))) return cleanup(rc);
    if ((rc = alloc((void**)&h->bottom_status, sizeof(int) * 4))) return cleanup(rc);
    if (hipHostMalloc((void**)&h->red_host, sizeof(double) * 2) != hipSuccess)
        return cleanup(fail(ALAMO_EL_EHIP, "hipHostMalloc failed"));
    for (int q = 0; q < 78; ++q) { h->bct_host[q] = ALAMO_EL_BC_DISPLACEMENT; h->bcv_host[q] = 0.0; }
    for (int a = 0; a < 3; ++a) h->amr_dom[a] = n_cell[a];
    if (hipMemcpy(h->bct_dev, h->bct_host, sizeof(h->bct_host), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(h->bcv_dev, h->bcv_host, sizeof(h->bcv_host), hipMemcpyHostToDevice) != hipSuccess)
        return cleanup(fail(ALAMO_EL_EHIP, "hipMemcpy of the boundary table failed"));
    *out = h;
    return 0;
}

int alamo_el_destroy(alamo_el_t h) {
    if (!h) return 0;
    for (auto& l : h->lv) {
        for (double* p : {l.d.C, l.d.diag, l.d.cor, l.d.cor_hold, l.d.res, l.d.rescor, l.d.tmp, l.d.D, l.d.Ct, l.d.Cb,
                          l.d.pal})
            if (p) (void)hipFree(p);
        for (float* p : {l.f.C, l.f.diag, l.f.cor, l.f.cor_hold, l.f.res, l.f.rescor, l.f.tmp, l.f.D, l.f.Ct, l.f.Cb,
                         l.f.pal})
            if (p) (void)hipFree(p);
        if (l.pidx) (void)hipFree(l.pidx);
        if (l.pidx_b) (void)hipFree(l.pidx_b);
        if (l.gmask) (void)hipFree(l.gmask);
        if (l.dwbase) (void)hipFree(l.dwbase);
        if (l.psi) (void)hipFree(l.psi);
    }
    free_direct(h);
    for (void* p : {(void*)h->bct_dev, (void*)h->bcv_dev, (void*)h->red_partial, (void*)h->red_out,
                    (void*)h->bottom_work, (void*)h->bottom_status, (void*)h->gu9, (void*)h->dw9, (void*)h->csoa,
                    (void*)h->kv, (void*)h->kz, (void*)h->kw, (void*)h->kh, (void*)h->kpart,
                    (void*)h->dense_I, (void*)h->dense_A, (void*)h->nn_in, (void*)h->nn_out,
                    (void*)h->nn_bc, (void*)h->nn_t, (void*)h->nn_fmean})
        if (p) (void)hipFree(p);
    if (h->red_host) (void)hipHostFree(h->red_host);
    if (h->red_host_vec) (void)hipHostFree(h->red_host_vec);
    for (hipEvent_t e : h->kev)
        if (e) (void)hipEventDestroy(e);
    delete h;
    return 0;
}

int alamo_el_num_levels(alamo_el_t h, int* nlev) {
    if (!h || !nlev) return fail(ALAMO_EL_EINVAL, "null argument");
    *nlev = (int)h->lv.size();
    return 0;
}

int alamo_el_level_nodes(alamo_el_t h, int mglev, int nodes[3]) {
    if (bad_level(h, mglev) || !nodes) return fail(ALAMO_EL_EINVAL, "bad level");
    nodes[0] = h->lv[mglev].L.nx;
    nodes[1] = h->lv[mglev].L.ny;
    nodes[2] = h->lv[mglev].L.nz;
    return 0;
}

// The axes whose faces boundary region r lies on (region_of's numbering):
// faces 0-5, yz edges 6-9, zx edges 10-13, xy edges 14-17, corners 18-25.
static int region_axes(int r) {
    if (r < 6) return 1 << (r % 3);
    if (r < 10) return 6;
    if (r < 14) return 5;
    if (r < 18) return 3;
    return 7;
}

// A table entry of type periodic must lie on a periodic axis: regions on the
// faces of a periodic axis never occur (Constant::Init skips those faces,
// src/BC/Operator/Elastic/Constant.H:160-185), every other region is a
// boundary row and needs a displacement, traction, Neumann or none type.
static int periodic_table_check(const int* type, int per) {
    for (int r = 0; r < 26; ++r) {
        if (region_axes(r) & per) continue;
        for (int c = 0; c < 3; ++c)
            if (type[3 * r + c] == ALAMO_EL_BC_PERIODIC)
                return fail(ALAMO_EL_EINVAL, "a boundary region off the periodic axes has the periodic type "
                                             "(call alamo_el_set_periodic first)");
    }
    return 0;
}

int alamo_el_set_bc(alamo_el_t h, const int type_host[78], const double val_host[78]) {
    if (!h || !type_host || !val_host) return fail(ALAMO_EL_EINVAL, "null argument");
    for (int q = 0; q < 78; ++q)
        if (type_host[q] < 0 || type_host[q] > 4) return fail(ALAMO_EL_EINVAL, "bad boundary type");
    int rc;
    if ((rc = periodic_table_check(type_host, h->per))) return rc;
    std::memcpy(h->bct_host, type_host, sizeof(h->bct_host));
    std::memcpy(h->bcv_host, val_host, sizeof(h->bcv_host));
    HIPCHK(hipMemcpy(h->bct_dev, h->bct_host, sizeof(h->bct_host), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(h->bcv_dev, h->bcv_host, sizeof(h->bcv_host), hipMemcpyHostToDevice));
    h->prepared = false;
    return 0;
}

int alamo_el_set_periodic(alamo_el_t h, const int periodic[3]) {
    if (!h || !periodic) return fail(ALAMO_EL_EINVAL, "null argument");
    int per = 0;
    for (int a = 0; a < 3; ++a)
        if (periodic[a]) per |= 1 << a;
    if (per && h->amr_parent) return fail(ALAMO_EL_EINVAL, "an AMR patch cannot be periodic");
    if (per && h->psi_set) return fail(ALAMO_EL_EINVAL, "the psi weight is not supported on periodic axes");
    int rc;
    if ((rc = periodic_table_check(h->bct_host, per))) return rc;
    if (per != h->per) {
        // diagonals, coarse operators and the bottom depend on the wrap
        h->prepared = false;
        h->prepared32 = false;
    }
    h->per = per;
    for (auto& l : h->lv) l.L.per = per;
    return 0;
}

int alamo_el_set_uniform(alamo_el_t h, int uniform) {
    if (!h) return fail(ALAMO_EL_EINVAL, "null handle");
    const int u = uniform ? 1 : 0;
    if (u != h->uniform) {
        // the diagonals and the dense and direct bottom operators are built
        // with the flag: the next solve prepares again
        h->prepared = false;
        h->prepared32 = false;
    }
    h->uniform = u;
    return 0;
}

int alamo_el_set_omega(alamo_el_t h, double omega) {
    if (!h || !(omega > 0.0)) return fail(ALAMO_EL_EINVAL, "bad omega");
    h->omega = omega;
    return 0;
}

int alamo_el_set_model(alamo_el_t h, const double* C, int layout, hipStream_t s) {
    if (!h || !C) return fail(ALAMO_EL_EINVAL, "null argument");
    Level& l = h->lv[0];
    const long n = (long)h->nc * l.L.N;
    if (layout == ALAMO_EL_SOA) {
        HIPCHK(hipMemcpyAsync(l.d.C, C, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
    } else if (layout == ALAMO_EL_AOS) {
        ael::k_aos_to_soa<<<nblk(l.L.N), kBlock, 0, s>>>(C, l.d.C, l.L.N, h->nc);
        HIPCHK(hipGetLastError());
    } else {
        return fail(ALAMO_EL_EINVAL, "bad layout");
    }
    int rc = build_gmask(h, l, s);
    if (rc) return rc;
    h->model_set = true;
    h->prepared = false;
    h->prepared32 = false;
    return 0;
}

// k_bottom_solve (the coarsest-level BiCGStab for levels above kDenseMax
// unknowns) applies the unweighted operator, so with psi the bottom must be
// the dense or the direct one.
static int psi_bottom_check(alamo_el_t h, const alamo_el_opts& o) {
    if (!h->psi_set) return 0;
    if ((o.bottom_solver == ALAMO_EL_BOTTOM_DIRECT && h->direct_Ainv) || h->dense_A) return 0;
    return fail(ALAMO_EL_EINVAL, "with psi the coarsest level needs at most 1536 unknowns, or the direct bottom");
}

int alamo_el_set_psi(alamo_el_t h, const double* psi, double psi_small, hipStream_t s) {
    if (!h || !psi) return fail(ALAMO_EL_EINVAL, "null argument");
    if (h->per) return fail(ALAMO_EL_EINVAL, "the psi weight is not supported on periodic axes");
    if (!(psi_small >= 0.0 && psi_small < 1.0)) return fail(ALAMO_EL_EINVAL, "psi_small must lie in [0, 1)");
    for (auto& l : h->lv) {
        if (l.psi) continue;
        const long nc = (long)(l.L.nx - 1) * (l.L.ny - 1) * (l.L.nz - 1);
        HIPCHK(hipMalloc((void**)&l.psi, sizeof(double) * std::max(nc, 1L)));
    }
    Level& l0 = h->lv[0];
    const long nc0 = (long)(l0.L.nx - 1) * (l0.L.ny - 1) * (l0.L.nz - 1);
    HIPCHK(hipMemcpyAsync(l0.psi, psi, sizeof(double) * nc0, hipMemcpyDeviceToDevice, s));
    h->psi_set = true;
    h->psi_small = psi_small;
    h->prepared = false;
    h->prepared32 = false;
    return 0;
}

int alamo_el_palette_size(alamo_el_t h, int mglev, int* npal) {
    if (bad_level(h, mglev) || !npal) return fail(ALAMO_EL_EINVAL, "bad argument");
    *npal = h->lv[mglev].pidx ? h->lv[mglev].npal : 0;
    return 0;
}

int alamo_el_prepare(alamo_el_t h, hipStream_t s) {
    if (!h) return fail(ALAMO_EL_EINVAL, "null handle");
    return prepare(h, s);
}

int alamo_el_prepare_fp32(alamo_el_t h, hipStream_t s) {
    if (!h) return fail(ALAMO_EL_EINVAL, "null handle");
    if (!h->model_set) return fail(ALAMO_EL_ESTATE, "Attempting to use operator before calling SetModel");
    int rc = h->prepared ? 0 : prepare(h, s);
    return rc ? rc : prepare32(h, s);
}

int alamo_el_fapply(alamo_el_t h, int mglev, const double* u, double* f, hipStream_t s) {
    if (bad_level(h, mglev) || !u || !f) return fail(ALAMO_EL_EINVAL, "bad argument");
    if (!h->model_set) return fail(ALAMO_EL_ESTATE, "Attempting to use operator before calling SetModel");
    if (mglev > 0 && !h->prepared) {
        int rc = prepare(h, s);
        if (rc) return rc;
    }
    Level& l = h->lv[mglev];
    launch_fapply<double>(h, l, u, l.d.C, f, s);
    HIPCHK(hipGetLastError());
    return 0;
}

int alamo_el_diagonal(alamo_el_t h, int mglev, double* diag, hipStream_t s) {
    if (bad_level(h, mglev) || !diag) return fail(ALAMO_EL_EINVAL, "bad argument");
    if (!h->model_set) return fail(ALAMO_EL_ESTATE, "Attempting to use operator before calling SetModel");
    if (mglev > 0 && !h->prepared) {
        int rc = prepare(h, s);
        if (rc) return rc;
    }
    Level& l = h->lv[mglev];
    launch_diagonal<double>(h, l, l.d.C, diag, s);
    HIPCHK(hipGetLastError());
    return 0;
}

int alamo_el_fsmooth(alamo_el_t h, int mglev, double* x, const double* b, int nsmooth, hipStream_t s) {
    if (bad_level(h, mglev) || !x || !b || nsmooth < 0) return fail(ALAMO_EL_EINVAL, "bad argument");
    if (!h->prepared) {
        int rc = prepare(h, s);
        if (rc) return rc;
    }
    Cycle<double> c{h, s, 2, 2, 8, 200, 1e-4, ALAMO_EL_BOTTOM_BICGSTAB};
    c.smooth(mglev, x, b, nsmooth);
    HIPCHK(hipGetLastError());
    return 0;
}

int alamo_el_gs_smooth(alamo_el_t h, int mglev, double* x, const double* b, int nsmooth, hipStream_t s) {
    if (bad_level(h, mglev) || !x || !b || nsmooth < 0) return fail(ALAMO_EL_EINVAL, "bad argument");
    if (!h->prepared) {
        int rc = prepare(h, s);
        if (rc) return rc;
    }
    Cycle<double> c{h, s, 2, 2, 8, 200, 1e-4, ALAMO_EL_BOTTOM_BICGSTAB, ALAMO_EL_SMOOTH_RBGS};
    c.smooth(mglev, x, b, nsmooth);
    HIPCHK(hipGetLastError());
    return 0;
}

int alamo_el_residual(alamo_el_t h, int mglev, const double* x, const double* b, double* r, hipStream_t s) {
    if (bad_level(h, mglev) || !x || !b || !r) return fail(ALAMO_EL_EINVAL, "bad argument");
    if (!h->model_set) return fail(ALAMO_EL_ESTATE, "Attempting to use operator before calling SetModel");
    if (mglev > 0 && !h->prepared) {
        int rc = prepare(h, s);
        if (rc) return rc;
    }
    Level& l = h->lv[mglev];
    launch_residual<double>(h, l, x, b, l.d.C, r, s);
    HIPCHK(hipGetLastError());
    return 0;
}

int alamo_el_restrict(alamo_el_t h, int cmglev, double* crse, const double* fine, hipStream_t s) {
    if (bad_level(h, cmglev) || cmglev == 0 || !crse || !fine) return fail(ALAMO_EL_EINVAL, "bad argument");
    ael::k_restrict<double><<<nblk8(h->lv[cmglev].L.N), kBlock, 0, s>>>(fine, h->lv[cmglev - 1].L, crse,
                                                                        h->lv[cmglev].L, 3);
    HIPCHK(hipGetLastError());
    return 0;
}

int alamo_el_interp_add(alamo_el_t h, int fmglev, double* fine, const double* crse, hipStream_t s) {
    if (bad_level(h, fmglev) || fmglev + 1 >= (int)h->lv.size() || !crse || !fine)
        return fail(ALAMO_EL_EINVAL, "bad argument");
    launch_interp3<double>(crse, nullptr, h->lv[fmglev + 1].L, fine, h->lv[fmglev].L, 0, s);
    HIPCHK(hipGetLastError());
    return 0;
}

int alamo_el_level_op32(alamo_el_t h, int mglev, int op, const float* x, const float* b, float* out,
                        hipStream_t s) {
    if (bad_level(h, mglev) || !x || !out || op < 0 || op > 4) return fail(ALAMO_EL_EINVAL, "bad argument");
    if ((op == 1 || op == 2) && !b) return fail(ALAMO_EL_EINVAL, "bad argument");
    if (op >= 3 && mglev + 1 >= (int)h->lv.size()) return fail(ALAMO_EL_EINVAL, "no coarser level");
    if (!h->model_set) return fail(ALAMO_EL_ESTATE, "Attempting to use operator before calling SetModel");
    int rc;
    if (!h->prepared && (rc = prepare(h, s))) return rc;
    if ((rc = prepare32(h, s))) return rc;
    Level& l = h->lv[mglev];
    switch (op) {
        case 0: launch_fapply<float>(h, l, x, l.f.C, out, s); break;
        case 1: launch_residual<float>(h, l, x, b, l.f.C, out, s); break;
        case 2: launch_jacobi<float>(h, l, x, b, l.f.diag, l.f.C, out, s); break;
        case 3:
            ael::k_restrict<float><<<nblk8(h->lv[mglev + 1].L.N), kBlock, 0, s>>>(x, l.L, out, h->lv[mglev + 1].L, 3);
            break;
        default:
            launch_interp3<float>(x, nullptr, h->lv[mglev + 1].L, out, l.L, 0, s);
            break;
    }
    HIPCHK(hipGetLastError());
    return 0;
}

int alamo_el_bc_init(alamo_el_t h, double* rhs, int homogeneous, hipStream_t s) {
    if (!h || !rhs) return fail(ALAMO_EL_EINVAL, "bad argument");
    const Lvl L = h->lv[0].L;
    ael::k_bc_init<<<nblk(L.N), kBlock, 0, s>>>(rhs, h->bct_dev, h->bcv_dev, homogeneous, L);
    HIPCHK(hipGetLastError());
    return 0;
}

int alamo_el_mix_models(alamo_el_t h, const double* eta, int nmodels, const double* table_host, int ncomp,
                        double* out, hipStream_t s) {
    if (!h || !eta || !table_host || !out || nmodels < 1 || ncomp < 1)
        return fail(ALAMO_EL_EINVAL, "bad argument");
    if ((long)nmodels * ncomp > ael::kMixMax) return fail(ALAMO_EL_EINVAL, "nmodels * ncomp exceeds 256");
    const Lvl L = h->lv[0].L;
    ael::MixTable tab{};
    for (int q = 0; q < nmodels * ncomp; ++q) tab.v[q] = table_host[q];
    ael::k_mix_models<<<nblk(L.N), kBlock, 0, s>>>(eta, nmodels, tab, ncomp, L.N, out);
    HIPCHK(hipGetLastError());
    return 0;
}

namespace {
void launch_newton_dw(alamo_el_t h, const double* u, const double* C, long cs, long cn, const double* F0, int kb,
                      int np, double* gu9, double* dw9, hipStream_t s) {
    const Lvl L = h->lv[0].L;
    const unsigned g = nblk((long)np * L.nx * L.ny);
    switch (h->sym) {
        case 0: ael::k_newton_dw<0><<<g, kBlock, 0, s>>>(u, C, cs, cn, F0, L, kb, np, gu9, dw9); break;
        case 1: ael::k_newton_dw<1><<<g, kBlock, 0, s>>>(u, C, cs, cn, F0, L, kb, np, gu9, dw9); break;
        default: ael::k_newton_dw<2><<<g, kBlock, 0, s>>>(u, C, cs, cn, F0, L, kb, np, gu9, dw9); break;
    }
}
}  // namespace

int alamo_el_newton_rhs(alamo_el_t h, const double* C, int layout, const double* F0, const double* u,
                        const double* b, double* rhs, hipStream_t s) {
    if (!h || !C || !u || !b || !rhs) return fail(ALAMO_EL_EINVAL, "bad argument");
    const Lvl L = h->lv[0].L;
    if (!h->gu9) ALLOC(h->gu9, 9 * L.N);
    if (!h->dw9) ALLOC(h->dw9, 9 * L.N);
    const double* Cs = C;
    if (layout == ALAMO_EL_AOS) {
        if (!h->csoa) ALLOC(h->csoa, (long)h->nc * L.N);
        ael::k_aos_to_soa<<<nblk(L.N), kBlock, 0, s>>>(C, h->csoa, L.N, h->nc);
        Cs = h->csoa;
    } else if (layout != ALAMO_EL_SOA) {
        return fail(ALAMO_EL_EINVAL, "bad layout");
    }
    launch_newton_dw(h, u, Cs, L.N, 1, F0, 0, L.nz, h->gu9, h->dw9, s);
    ael::k_newton_rhs<<<nblk(L.N), kBlock, 0, s>>>(u, b, h->gu9, h->dw9, 0, L.nz, h->bct_dev, L, 0, L.nz, rhs);
    HIPCHK(hipGetLastError());
    return 0;
}

// --- The Newton right-hand side that allocates nothing --------------------
// Scratch of alamo_el_newton_rhs_noalloc: DW of a slab of planes.  One copy
// per GPU, held until the process ends; host code never passes the array's
// name to a kernel, only the device address hipGetSymbolAddress returns.
constexpr long kNoAllocScratchBytes = 32l << 20;
static __device__ __align__(256) double scratch[kNoAllocScratchBytes / sizeof(double)];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch_ptr[kMaxDevices] = {};

// Device address of the current device's scratch, looked up once per device.
int scratch_of_current_device(double** out) {
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) return fail(ALAMO_EL_EINVAL, "device index beyond the scratch table");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        HIPCHK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) return fail(ALAMO_EL_EHIP, "hipGetSymbolAddress returned no address for the scratch");
        g_scratch_ptr[dev] = static_cast<double*>(p);
    }
    *out = g_scratch_ptr[dev];
    return 0;
}
}  // namespace

int alamo_el_newton_rhs_noalloc(alamo_el_t h, const double* C, int layout, const double* F0, const double* u,
                                const double* b, double* rhs, hipStream_t s) {
    if (!h || !C || !u || !b || !rhs) return fail(ALAMO_EL_EINVAL, "bad argument");
    if (layout != ALAMO_EL_SOA && layout != ALAMO_EL_AOS) return fail(ALAMO_EL_EINVAL, "bad layout");
    const Lvl L = h->lv[0].L;
    const long P = (long)L.nx * L.ny;
    // planes of DW (9 doubles a node) the scratch holds; a slab of rows needs
    // its own planes and one more on each side
    const long cap = (kNoAllocScratchBytes / (long)sizeof(double)) / (9 * P);
    if (cap < 3)
        return fail(ALAMO_EL_EINVAL, "alamo_el_newton_rhs_noalloc: a plane of " + std::to_string(P) +
                                         " nodes exceeds the scratch (at most " +
                                         std::to_string(kNoAllocScratchBytes / (long)sizeof(double) / 27) +
                                         " nodes per x-y plane)");
    double* dw = nullptr;
    int rc = scratch_of_current_device(&dw);
    if (rc) return rc;
    const long cs = layout == ALAMO_EL_SOA ? L.N : 1, cn = layout == ALAMO_EL_SOA ? 1 : h->nc;
    const int rows = (int)std::min<long>(cap - 2, L.nz);
    // a periodic z wraps the divergence of the faces' rows across the box
    if ((L.per & 4) && rows < L.nz)
        return fail(ALAMO_EL_EINVAL, "alamo_el_newton_rhs_noalloc: a periodic z axis needs the whole box in the "
                                     "scratch (" + std::to_string(cap - 2) + " planes here)");
    for (int k0 = 0; k0 < L.nz; k0 += rows) {
        const int k1 = std::min(L.nz, k0 + rows);
        const int kb = std::max(0, k0 - 1), kt = std::min(L.nz - 1, k1);
        const int np = kt - kb + 1;
        launch_newton_dw(h, u, C, cs, cn, F0, kb, np, nullptr, dw, s);
        ael::k_newton_rhs<<<nblk((long)(k1 - k0) * P), kBlock, 0, s>>>(u, b, nullptr, dw, kb, np, h->bct_dev, L, k0,
                                                                        k1, rhs);
    }
    HIPCHK(hipGetLastError());
    return 0;
}

int alamo_el_stress(alamo_el_t h, const double* C, int layout, const double* F0, const double* u, double* sigma,
                    double* strain, hipStream_t s) {
    if (!h || !C || !u || !sigma) return fail(ALAMO_EL_EINVAL, "bad argument");
    const Lvl L = h->lv[0].L;
    const double* Cs = C;
    if (layout == ALAMO_EL_AOS) {
        if (!h->csoa) ALLOC(h->csoa, (long)h->nc * L.N);
        ael::k_aos_to_soa<<<nblk(L.N), kBlock, 0, s>>>(C, h->csoa, L.N, h->nc);
        Cs = h->csoa;
    } else if (layout != ALAMO_EL_SOA) {
        return fail(ALAMO_EL_EINVAL, "bad layout");
    }
    switch (h->sym) {
        case 0: ael::k_stress<0><<<nblk(L.N), kBlock, 0, s>>>(u, Cs, F0, L, sigma, strain); break;
        case 1: ael::k_stress<1><<<nblk(L.N), kBlock, 0, s>>>(u, Cs, F0, L, sigma, strain); break;
        default: ael::k_stress<2><<<nblk(L.N), kBlock, 0, s>>>(u, Cs, F0, L, sigma, strain); break;
    }
    HIPCHK(hipGetLastError());
    return 0;
}

int alamo_el_face_traction(alamo_el_t h, const double* sigma, const double* u, int axis, int side,
                           double* trac_host, double* disp_host, hipStream_t s) {
    if (!h || !sigma || !u || !trac_host || axis < 0 || axis > 2 || (side != 0 && side != 1))
        return fail(ALAMO_EL_EINVAL, "bad argument");
    const Lvl L = h->lv[0].L;
    ael::k_face_traction<<<1, ael::kFaceBlock, 0, s>>>(sigma, u, L, axis, side, h->red_partial);
    HIPCHK(hipGetLastError());
    double out[6];
    HIPCHK(hipMemcpyAsync(out, h->red_partial, sizeof(out), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    for (int r = 0; r < 3; ++r) {
        trac_host[r] = out[r];
        if (disp_host) disp_host[r] = out[3 + r];
    }
    return 0;
}

int alamo_el_norm_inf(alamo_el_t h, int mglev, const double* v, double* out_host, hipStream_t s) {
    if (bad_level(h, mglev) || !v || !out_host) return fail(ALAMO_EL_EINVAL, "bad argument");
    return reduce_host<double>(h, v, v, 3 * h->lv[mglev].L.N, 0, out_host, s);
}

int alamo_el_level_model(alamo_el_t h, int mglev, double* C, hipStream_t s) {
    if (bad_level(h, mglev) || !C) return fail(ALAMO_EL_EINVAL, "bad argument");
    if (!h->model_set) return fail(ALAMO_EL_ESTATE, "Attempting to use operator before calling SetModel");
    int rc;
    if (mglev > 0 && !h->prepared && (rc = prepare(h, s))) return rc;
    const Level& l = h->lv[mglev];
    HIPCHK(hipMemcpyAsync(C, l.d.C, sizeof(double) * h->nc * l.L.N, hipMemcpyDeviceToDevice, s));
    return 0;
}

int alamo_el_coarse_rhs(alamo_el_t h, int mglev, const double* b, double* bc, hipStream_t s) {
    if (bad_level(h, mglev) || !b || !bc || b == bc) return fail(ALAMO_EL_EINVAL, "bad argument");
    return coarse_rhs(h, mglev, b, bc, s);
}

int alamo_el_learned_input(alamo_el_t h, int mglev, const double* bc, float* features, double* scale_host,
                           hipStream_t s) {
    if (bad_level(h, mglev) || !bc || !features || !scale_host) return fail(ALAMO_EL_EINVAL, "bad argument");
    if (!h->model_set) return fail(ALAMO_EL_ESTATE, "Attempting to use operator before calling SetModel");
    int rc;
    if (mglev > 0 && !h->prepared && (rc = prepare(h, s))) return rc;
    return learned_input(h, mglev, bc, features, scale_host, s);
}

int alamo_el_set_learned(alamo_el_t h, alamo_nn_t nn, int mglev) {
    if (!h) return fail(ALAMO_EL_EINVAL, "null handle");
    for (void* p : {(void*)h->nn_in, (void*)h->nn_out, (void*)h->nn_bc})
        if (p) (void)hipFree(p);
    h->nn_in = nullptr;
    h->nn_out = nullptr;
    h->nn_bc = nullptr;
    h->nn = nullptr;
    h->nn_level = -1;
    if (!nn) return 0;
    if (bad_level(h, mglev)) return fail(ALAMO_EL_EINVAL, "bad level");
    int nodes[3], n_in, n_out;
    int rc;
    if ((rc = alamo_nn_shape(nn, nodes, &n_in, &n_out))) return rc;
    const Lvl& L = h->lv[mglev].L;
    if (nodes[0] != L.nx || nodes[1] != L.ny || nodes[2] != L.nz || n_in != h->nc + kLearnedExtra || n_out != 3)
        return fail(ALAMO_EL_EINVAL, "network grid or channels do not match the level (needs nc + 30 inputs, 3 outputs)");
    HIPCHK(hipMalloc((void**)&h->nn_in, sizeof(float) * n_in * L.N));
    HIPCHK(hipMalloc((void**)&h->nn_out, sizeof(float) * 3 * L.N));
    HIPCHK(hipMalloc((void**)&h->nn_bc, sizeof(double) * 3 * L.N));
    if ((rc = ensure_nn_scratch(h, L.N))) return rc;
    h->nn = nn;
    h->nn_level = mglev;
    return 0;
}

int alamo_el_learned_guess(alamo_el_t h, const double* b, double* x, hipStream_t s) {
    if (!h || !b || !x) return fail(ALAMO_EL_EINVAL, "null argument");
    if (!h->nn) return fail(ALAMO_EL_ESTATE, "no network set (alamo_el_set_learned)");
    if (!h->model_set) return fail(ALAMO_EL_ESTATE, "Attempting to use operator before calling SetModel");
    int rc;
    if (!h->prepared && (rc = prepare(h, s))) return rc;
    return learned_apply(h, b, x, s);
}

int alamo_el_solve(alamo_el_t h, double* x, const double* b, double tol_rel, double tol_abs,
                   const alamo_el_opts* opts_in, alamo_el_info* info, hipStream_t s) {
    if (!h || !x || !b) return fail(ALAMO_EL_EINVAL, "null argument");
    alamo_el_opts o;
    alamo_el_default_opts(&o);
    if (opts_in) o = *opts_in;
    if (o.max_iter < 0 || o.nu1 < 0 || o.nu2 < 0 || o.bottom_max_iter < 0 || !(o.omega > 0) ||
        (o.smoother != ALAMO_EL_SMOOTH_JACOBI && o.smoother != ALAMO_EL_SMOOTH_RBGS))
        return fail(ALAMO_EL_EINVAL, "bad solver options");
    const auto t0 = std::chrono::steady_clock::now();
    h->omega = o.omega;
    int rc;
    if (!h->prepared && (rc = prepare(h, s))) return rc;
    if ((rc = psi_bottom_check(h, o))) return rc;
    if (o.precision == ALAMO_EL_PREC_FP32_IR && (rc = prepare32(h, s))) return rc;
    Level& l0 = h->lv[0];
    const long M0 = 3 * l0.L.N;
    // MLMG::solve: r0, norms and the stopping target
    launch_residual<double>(h, l0, x, b, l0.d.C, l0.d.res, s);
    double resnorm0, rhsnorm0;
    if ((rc = reduce_host<double>(h, l0.d.res, l0.d.res, M0, 0, &resnorm0, s))) return rc;
    if ((rc = reduce_host<double>(h, b, b, M0, 0, &rhsnorm0, s))) return rc;
    const double target = std::max(tol_abs, std::max(tol_rel, 1.e-16) * std::max(resnorm0, rhsnorm0));
    int iters = 0, status = ALAMO_EL_NOT_CONVERGED;
    double nrm = resnorm0;  // fgmres_solve reads it as the inf-norm of r0
    if (std::isnan(resnorm0)) return fail(ALAMO_EL_ENAN, "NaN in the initial residual");
    // Learned initial guess: a judged learned correction of the caller's x
    // (learned_step); the stopping target stays the one of the caller's x.
    int learned_taken = 0;
    const bool learned_pre = (o.learned & ALAMO_EL_LEARNED_PRECOND) && h->nn;
    if ((o.learned & ALAMO_EL_LEARNED_GUESS) && h->nn && resnorm0 > target) {
        int taken;
        if ((rc = learned_step(h, x, &nrm, &taken, o.precision == ALAMO_EL_PREC_FP32_IR, s))) return rc;
        learned_taken += taken;
        if (o.verbose) std::fprintf(stderr, "alamo_el: learned guess %s, |r|_inf %.6e\n", taken ? "taken" : "rejected", nrm);
    }
    if (nrm <= target) {
        status = 0;
    } else if (o.outer == ALAMO_EL_OUTER_FGMRES) {
        status = fgmres_solve(h, x, b, target, o, iters, nrm, learned_taken, s);
        if (status < 0) return status;
    } else {
        Cycle<double> c64{h, s, o.nu1, o.nu2, o.final_smooth, o.bottom_max_iter, o.bottom_tol_rel, o.bottom_solver,
                          o.smoother};
        Cycle<float> c32{h, s, o.nu1, o.nu2, o.final_smooth, o.bottom_max_iter, o.bottom_tol_rel, o.bottom_solver,
                         o.smoother};
        for (int it = 0; it < iter_limit(o); ++it) {
            if (learned_pre && it > 0) {
                // learned coarse correction between cycles, judged like the guess
                int taken;
                if ((rc = learned_step(h, x, &nrm, &taken, o.precision == ALAMO_EL_PREC_FP32_IR, s))) return rc;
                learned_taken += taken;
                if (nrm <= target) { iters = it; status = 0; break; }
            }
            if (o.precision == ALAMO_EL_PREC_FP32_IR) {
                // FP32 cycle on the FP64 residual; FP64 correction accumulation
                ael::k_cast<double, float><<<nblk(M0), kBlock, 0, s>>>(l0.d.res, l0.f.res, M0);
                c32.run(cycle_of(o, it));
                ael::k_add_f32_to_f64<<<nblk(M0), kBlock, 0, s>>>(x, l0.f.cor, M0);
            } else {
                c64.run(cycle_of(o, it));
                launch_axpy<double>(x, 1.0, l0.d.cor, M0, s);
            }
            launch_residual<double>(h, l0, x, b, l0.d.C, l0.d.res, s);
            if ((rc = reduce_host<double>(h, l0.d.res, l0.d.res, M0, 0, &nrm, s))) return rc;
            iters = it + 1;
            if (o.verbose) std::fprintf(stderr, "alamo_el: iter %d |r|_inf = %.6e\n", iters, nrm);
            if (std::isnan(nrm)) return fail(ALAMO_EL_ENAN, "NaN in the residual");
            if (nrm <= target) { status = 0; break; }
        }
    }
    HIPCHK(hipGetLastError());
    if (status == ALAMO_EL_NOT_CONVERGED && o.fixed_iter > 0 && iters >= iter_limit(o)) status = 0;
    if (info) {
        info->status = status;
        info->learned_steps = learned_taken;
        info->iters = iters;
        info->res0 = resnorm0;
        info->rhsnorm = rhsnorm0;
        info->res = nrm;
        info->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    }
    if (status == ALAMO_EL_NOT_CONVERGED) g_err = "MLMG did not converge within max_iter";
    return status;
}

}  // extern "C"

namespace {

// Flexible GMRES(m), right preconditioned by one multigrid cycle in the
// chosen precision (the cycle is a nonlinear, changing preconditioner, so
// the preconditioned directions Z are stored).  Classical Gram-Schmidt twice:
// for m <= kMultiDotMax with the second pass of each vector delayed to the
// next step (DCGS2, elastic_kernels.h), else fused passes up to kMultiDotMax
// vectors and the DGKS test deciding a second pass beyond.  The Krylov basis
// is FP64, its vectors MP apart (M0 rounded up to whole 128-byte lines, so
// that every vector starts a line).
int fgmres_solve(alamo_el_t h, double* x, const double* b, double target, const alamo_el_opts& o, int& iters,
                 double& resnorm, int& learned_taken, hipStream_t s) {
    Level& l0 = h->lv[0];
    const long M0 = 3 * l0.L.N;
    const long MP = (M0 + 15) / 16 * 16;
    const long MPf = (M0 + 31) / 32 * 32;  // FP32 Z: whole 128-byte lines too
    const int m = std::max(2, std::min(o.restart, 100));
    if (h->krylov_m < m) {
        for (double* p : {h->kv, h->kz, h->kw, h->kh, h->kpart})
            if (p) (void)hipFree(p);
        h->kv = h->kz = h->kw = h->kh = h->kpart = nullptr;
        ALLOC(h->kv, (long)(m + 1) * MP);
        ALLOC(h->kz, (long)m * MP);
        ALLOC(h->kw, M0);
        // two CGS passes of coefficients (or DCGS2's), |w|^2 after and before,
        // then DCGS2's 2 (m + 1) dot sums
        ALLOC(h->kh, 4 * (m + 1) + 2);
        ALLOC(h->kpart, 2L * (m + 1) * kDotBlocks);
        if (h->red_host_vec) (void)hipHostFree(h->red_host_vec);
        HIPCHK(hipHostMalloc((void**)&h->red_host_vec, sizeof(double) * 2 * (2 * (m + 1) + 2)));
        h->krylov_m = m;
    }
    for (hipEvent_t& e : h->kev)
        if (!e) HIPCHK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    const bool fp32 = o.precision == ALAMO_EL_PREC_FP32_IR;
    Cycle<double> c64{h, s, o.nu1, o.nu2, o.final_smooth, o.bottom_max_iter, o.bottom_tol_rel, o.bottom_solver,
                      o.smoother};
    Cycle<float> c32{h, s, o.nu1, o.nu2, o.final_smooth, o.bottom_max_iter, o.bottom_tol_rel, o.bottom_solver,
                     o.smoother};
    c64.leave_held = c32.leave_held = true;  // Z_j takes cor + cor_hold in its copy
    // H: the Hessenberg columns as projected; R: the final columns after the
    // Givens rotations (R[i * m + j], row i, column j)
    std::vector<double> H((m + 1) * m), R((m + 1) * m), cs(m), sn(m), g(m + 1), y(m);
    auto finalize = [&](int c) {
        for (int i = 0; i <= c + 1; ++i) R[i * m + c] = H[i * m + c];
        for (int i = 0; i < c; ++i) {
            const double a = R[i * m + c], d = R[(i + 1) * m + c];
            R[i * m + c] = cs[i] * a + sn[i] * d;
            R[(i + 1) * m + c] = -sn[i] * a + cs[i] * d;
        }
        const double a = R[c * m + c], d = R[(c + 1) * m + c];
        const double rr = std::hypot(a, d);
        cs[c] = rr > 0 ? a / rr : 1.0;
        sn[c] = rr > 0 ? d / rr : 0.0;
        R[c * m + c] = rr;
        R[(c + 1) * m + c] = 0.0;
        g[c + 1] = -sn[c] * g[c];
        g[c] = cs[c] * g[c];
    };
    // |g_{c+1}| if column c were final as it stands (columns < c final)
    auto estimate = [&](int c) {
        // row c of the rotated column: rotation i mixes rows i and i+1, so
        // row i+1 after it depends only on row i after rotation i-1
        double a = H[c];
        for (int i = 0; i < c; ++i) a = -sn[i] * a + cs[i] * H[(i + 1) * m + c];
        const double d = H[(c + 1) * m + c];
        const double rr = std::hypot(a, d);
        return rr > 0 ? std::fabs(d / rr * g[c]) : std::fabs(g[c]);
    };
    // ALAMO_EL_GS=cgs2 keeps the undelayed passes (for comparison)
    const char* gs_env = std::getenv("ALAMO_EL_GS");
    const bool lag = m <= ael::kMultiDotMax && !(gs_env && std::strcmp(gs_env, "cgs2") == 0);
    // ALAMO_EL_Z=fp64 keeps the FP32 path's Z basis in FP64 (for comparison)
    const char* z_env = std::getenv("ALAMO_EL_Z");
    const bool z32 = !(z_env && std::strcmp(z_env, "fp64") == 0);
    // With the delayed passes and FP32 cycles the basis V is kept in FP32 as
    // well, MPf apart: both passes read it every step, and the sums stay FP64
    // (ALAMO_EL_V=fp64 keeps it in FP64, for comparison).  The Arnoldi
    // relation then holds to FP32 rounding relative to the residual at the
    // restart, far below what one cycle of 10 steps reduces it by; the true
    // residual is formed in FP64 at every restart and at the end.
    const char* v_env = std::getenv("ALAMO_EL_V");
    const bool v32 = lag && fp32 && z32 && !(v_env && std::strcmp(v_env, "fp64") == 0);
    int rc;
    iters = 0;
    // r = b - A x is in l0.d.res on entry
    bool first = true;
    const int limit = iter_limit(o);
    while (iters < limit) {
        if ((o.learned & ALAMO_EL_LEARNED_PRECOND) && h->nn && !first) {
            // learned coarse correction at each restart, judged like the guess
            int taken;
            if ((rc = learned_step(h, x, &resnorm, &taken, o.precision == ALAMO_EL_PREC_FP32_IR, s))) return rc;
            learned_taken += taken;
            if (resnorm <= target) return 0;
        }
        first = false;
        double beta2;
        if ((rc = reduce_host<double>(h, l0.d.res, l0.d.res, M0, 1, &beta2, s))) return rc;
        const double beta = std::sqrt(beta2);
        if (beta == 0.0) { resnorm = 0.0; return 0; }
        const double inf_at_restart = resnorm;
        std::fill(g.begin(), g.end(), 0.0);
        g[0] = beta;
        int j = 0;
        double hn_prev = 0.0;
        double* nrm_slot = h->kh + 2 * (m + 1);
        // The device part of step jj (iteration it): V_jj and its copy in the
        // cycle's precision, Z_jj = cycle(V_jj), w = A Z_jj.
        auto launch_step = [&](int jj, int it) {
            double* Vj = h->kv + (long)jj * MP;
            double* Zj = h->kz + (long)jj * MP;
            // V_j = r / beta (j = 0) or w / |w| of the last step, written with
            // its copy in the cycle's precision as the cycle's right-hand side
            const double* src = jj == 0 ? l0.d.res : h->kw;
            const double scale = jj == 0 ? 1.0 / beta : (hn_prev > 0 ? 1.0 / hn_prev : 0.0);
            // with the delayed passes the scale is taken from |w|^2 on the
            // device, so that the step can be launched ahead of the host
            const double* nrm2 = lag && jj > 0 ? nrm_slot : nullptr;
            // Z_j = cycle(V_j); an F-cycle's last correction sum is formed by
            // the copy into Z_j
            const int cyc = cycle_of(o, it);
            const bool held = cyc == ALAMO_EL_CYCLE_F;
            // a V-cycle's first sweep from zero on level 0 comes with the copy
            const bool fz = cyc == ALAMO_EL_CYCLE_V && c32.fz_ready() && c32.bot() > 0;
            if (fp32 && !z32) {  // Z_j widened to FP64 (ALAMO_EL_Z=fp64)
                ael::k_scale_copy_cast<float><<<nblk(M0), kBlock, 0, s>>>(src, scale, Vj, l0.f.res, M0, nullptr, 0.f,
                                                                             nullptr, nrm2);
                c32.run(cyc);
                if (held) ael::k_cast_sum<float, double><<<nblk(M0), kBlock, 0, s>>>(l0.f.cor, l0.f.cor_hold, Zj, M0);
                else ael::k_cast<float, double><<<nblk(M0), kBlock, 0, s>>>(l0.f.cor, Zj, M0);
                launch_fapply<double>(h, l0, Zj, l0.d.C, h->kw, s);
            } else if (fp32) {
                // Z_j is kept in FP32: the FP32 cycle's correction, whose FP64
                // widening is exact, so every product with it is the one with
                // its FP64 copy
                float* Zf = reinterpret_cast<float*>(h->kz) + (long)jj * MPf;
                // A V-cycle only reads its level-0 right-hand side, so with an
                // FP32 basis it reads V_j itself (the same FP32 values) and no
                // copy is written; an F-cycle swaps its level-0 residual
                // buffers and takes the copy.
                float* Vf = v32 ? reinterpret_cast<float*>(h->kv) + (long)jj * MPf : nullptr;
                float* rhs32 = (v32 && !held) ? nullptr : l0.f.res;
                auto copy_cast = [&](auto* V) {
                    using TY = std::remove_pointer_t<decltype(V)>;
                    if (fz)
                        ael::k_scale_copy_cast<float, TY><<<nblk(M0), kBlock, 0, s>>>(
                            src, scale, V, rhs32, M0, l0.f.diag, (float)h->omega, l0.f.tmp, nrm2);
                    else
                        ael::k_scale_copy_cast<float, TY><<<nblk(M0), kBlock, 0, s>>>(src, scale, V, rhs32, M0,
                                                                                      nullptr, 0.f, nullptr, nrm2);
                };
                if (v32) copy_cast(Vf);
                else copy_cast(Vj);
                if (held) {
                    c32.run(cyc, fz);
                    ael::k_cast_sum<float, float><<<nblk(M0), kBlock, 0, s>>>(l0.f.cor, l0.f.cor_hold, Zf, M0);
                } else {
                    // a V-cycle writes its level-0 correction straight into Z_j
                    float* cor = l0.f.cor;
                    float* res = l0.f.res;
                    l0.f.cor = Zf;
                    if (!rhs32) l0.f.res = Vf;
                    c32.run(cyc, fz);
                    l0.f.cor = cor;
                    l0.f.res = res;
                }
                // w = A Z_j: the FP64 tiled Fapply reads the FP32 field;
                // elsewhere an FP64 copy in l0.d.rescor (free during a step)
                if ((h->psi_set && l0.psi) ||
                    !launch_tiled<double, 0, float>(h, l0, Zf, nullptr, nullptr, l0.d.C, h->kw, s)) {
                    ael::k_cast<float, double><<<nblk(M0), kBlock, 0, s>>>(Zf, l0.d.rescor, M0);
                    launch_fapply<double>(h, l0, l0.d.rescor, l0.d.C, h->kw, s);
                }
            } else {
                if (fz)
                    ael::k_scale_copy_cast<double><<<nblk(M0), kBlock, 0, s>>>(src, scale, Vj, l0.d.res, M0,
                                                                              l0.d.diag, h->omega, l0.d.tmp, nrm2);
                else
                    ael::k_scale_copy_cast<double><<<nblk(M0), kBlock, 0, s>>>(src, scale, Vj, l0.d.res, M0, nullptr, 0.0,
                                                                              nullptr, nrm2);
                c64.run(cyc, fz);
                if (held) ael::k_cast_sum<double, double><<<nblk(M0), kBlock, 0, s>>>(l0.d.cor, l0.d.cor_hold, Zj, M0);
                else ael::k_cast<double, double><<<nblk(M0), kBlock, 0, s>>>(l0.d.cor, Zj, M0);
                launch_fapply<double>(h, l0, Zj, l0.d.C, h->kw, s);  // w = A Z_j
            }
        };
        // DCGS2's two passes over V for step jj, then |w|^2, and the copy of
        // the coefficients into the host buffer of jj's parity; the event of
        // that parity marks the copy done.
        auto dcgs2_passes = [&](auto* V, long ld, int jj) {
            using TV = std::remove_pointer_t<decltype(V)>;
            const int n = jj + 1;
            if (jj == 0) {
                ael::k_multidot_partial<TV><<<kDotBlocks, kBlock, 0, s>>>(V, 1, M0, ld, h->kw, h->kpart);
                ael::k_multidot_final<<<1, kBlock, 0, s>>>(h->kpart, kDotBlocks, h->kh);
                ael::k_multiaxpy_dot<false, TV><<<kDotBlocks, kBlock, 0, s>>>(h->kw, V, 1, M0, ld, h->kh, h->kpart);
            } else {
                double* sums = h->kh + 2 * (m + 1) + 2;  // the 2n sums of pass 1
                ael::k_multidot2_partial<TV><<<kDotBlocks, kBlock, 0, s>>>(V, n, M0, ld, h->kw, h->kpart);
                ael::k_multidot_final<<<2 * n, kBlock, 0, s>>>(h->kpart, kDotBlocks, sums);
                ael::k_dcgs2_update<TV><<<kDotBlocks, kBlock, 0, s>>>(V, n, M0, ld, h->kw, sums, h->kh, h->kpart);
            }
        };
        auto launch_dcgs2 = [&](int jj) -> int {
            if (v32) dcgs2_passes(reinterpret_cast<float*>(h->kv), MPf, jj);
            else dcgs2_passes(h->kv, MP, jj);
            ael::k_multidot_final<<<1, kBlock, 0, s>>>(h->kpart, kDotBlocks, nrm_slot);
            const long nv = 2 * (m + 1) + 2;
            HIPCHK(hipMemcpyAsync(h->red_host_vec + (jj & 1) * nv, h->kh, sizeof(double) * nv, hipMemcpyDeviceToHost,
                                  s));
            HIPCHK(hipEventRecord(h->kev[jj & 1], s));
            return 0;
        };
        // With the delayed passes the next step is launched before the host
        // reads this one's coefficients (it needs none of them), so the GPU
        // does not wait for the host between steps; it is not when this step
        // may end the cycle (the estimate within kAheadMargin of the target,
        // the cycle's last column or the iteration limit): a step launched
        // ahead and then not needed costs a whole step.  ALAMO_EL_AHEAD=0
        // turns it off.
        const char* ahead_env = std::getenv("ALAMO_EL_AHEAD");
        const bool ahead = lag && !(ahead_env && std::strcmp(ahead_env, "0") == 0);
        constexpr double kAheadMargin = 32.0;
        double est_prev = beta;
        int launched = -1;  // the last step launched
        for (; j < m && iters < limit; ++j) {
            if (launched < j) {
                launch_step(j, iters);
                if (lag && (rc = launch_dcgs2(j))) return rc;
                launched = j;
            }
            // classical Gram-Schmidt, twice ("twice is enough"): h = V^T w;
            // w -= V h; h' = V^T w; w -= V h'.  Up to kMultiDotMax basis
            // vectors the passes are fused: the multi-dot, then one pass that
            // projects and forms the next dots, then one that projects and
            // forms |w|^2 (V read three times); the coefficients and the norm
            // come back in one copy.  Beyond that: the DGKS criterion decides
            // the second pass, with separate dot, projection and norm passes.
            const int nb = (int)std::min<long>(kRedBlocks, std::max<long>(1, nblk(M0)));
            if (lag) {
                // DCGS2: w's first projection and V_j's second in two passes
                // over V (launch_dcgs2); V_0 = r / beta needs no second one
                const int n = j + 1;
                if (ahead && j + 1 < m && iters + 1 < limit &&
                    est_prev / beta * inf_at_restart > kAheadMargin * target) {
                    launch_step(j + 1, iters + 1);
                    if ((rc = launch_dcgs2(j + 1))) return rc;
                    launched = j + 1;
                }
                HIPCHK(hipEventSynchronize(h->kev[j & 1]));
                const double* hv = h->red_host_vec + (j & 1) * (2 * (m + 1) + 2);
                for (int i = 0; i <= m; ++i) H[i * m + j] = 0.0;
                if (j == 0) {
                    H[0] = hv[0];
                } else {
                    // V_j = (t - V P) / alpha completes column j-1:
                    // w_{j-1} = V (h + beta P) + beta alpha V_j
                    const double alpha = hv[2 * n];
                    for (int i = 0; i < j; ++i) H[i * m + (j - 1)] += hn_prev * hv[i];
                    H[j * m + (j - 1)] = hn_prev * alpha;
                    finalize(j - 1);
                    if (!(alpha > 0.0)) break;  // t in the span of V: column j-1 ends the cycle exactly
                    for (int i = 0; i <= j; ++i) H[i * m + j] = hv[n + i];
                }
                const double hn = std::sqrt(hv[2 * (m + 1)]);
                H[(j + 1) * m + j] = hn;
                hn_prev = hn;  // V_{j+1} = w / hn is written by the next step
                ++iters;
                const double est = estimate(j);
                est_prev = est;
                if (o.verbose) std::fprintf(stderr, "alamo_el: fgmres %d |r|_2 est = %.6e\n", iters, est);
                if (est / beta * inf_at_restart <= target || hn == 0.0 || j + 1 == m || iters >= limit) {
                    finalize(j);  // as it stands: its second projection is not needed
                    ++j;
                    break;
                }
                continue;
            }
            auto sq_norm = [&](double* dst) {
                ael::k_reduce_partial<double><<<nb, kBlock, 0, s>>>(h->kw, h->kw, M0, 1, h->red_partial);
                ael::k_reduce_final<<<1, kBlock, 0, s>>>(h->red_partial, nb, 1, dst);
            };
            double* nrm_after = h->kh + 2 * (m + 1);
            double hn2 = 0.0;
            for (int i = 0; i <= j; ++i) H[i * m + j] = 0.0;
            if (j + 1 <= ael::kMultiDotMax) {
                double *h0 = h->kh, *h1 = h->kh + (m + 1);
                ael::k_multidot_partial<<<kDotBlocks, kBlock, 0, s>>>(h->kv, j + 1, M0, MP, h->kw, h->kpart);
                ael::k_multidot_final<<<j + 1, kBlock, 0, s>>>(h->kpart, kDotBlocks, h0);
                ael::k_multiaxpy_dot<true><<<kDotBlocks, kBlock, 0, s>>>(h->kw, h->kv, j + 1, M0, MP, h0, h->kpart);
                ael::k_multidot_final<<<j + 1, kBlock, 0, s>>>(h->kpart, kDotBlocks, h1);
                ael::k_multiaxpy_dot<false><<<kDotBlocks, kBlock, 0, s>>>(h->kw, h->kv, j + 1, M0, MP, h1, h->kpart);
                ael::k_multidot_final<<<1, kBlock, 0, s>>>(h->kpart, kDotBlocks, nrm_after);
                HIPCHK(hipMemcpyAsync(h->red_host_vec, h->kh, sizeof(double) * (2 * (m + 1) + 2),
                                      hipMemcpyDeviceToHost, s));
                HIPCHK(hipStreamSynchronize(s));
                const double* hv = h->red_host_vec;
                for (int i = 0; i <= j; ++i) H[i * m + j] = hv[i] + hv[(m + 1) + i];
                hn2 = hv[2 * (m + 1)];
            } else {
                sq_norm(nrm_after + 1);  // |w|^2 before
                double w2_before = 0.0;
                for (int pass = 0; pass < 2; ++pass) {
                    double* hp = h->kh + (long)pass * (m + 1);
                    for (int c0 = 0; c0 <= j; c0 += ael::kMultiDotMax) {
                        const int kc = std::min(ael::kMultiDotMax, j + 1 - c0);
                        ael::k_multidot_partial<<<kDotBlocks, kBlock, 0, s>>>(h->kv + (long)c0 * MP, kc, M0, MP,
                                                                              h->kw, h->kpart + (long)c0 * kDotBlocks);
                    }
                    ael::k_multidot_final<<<j + 1, kBlock, 0, s>>>(h->kpart, kDotBlocks, hp);
                    ael::k_multiaxpy<<<nblk(M0), kBlock, 0, s>>>(h->kw, h->kv, j + 1, M0, MP, hp, -1.0);
                    sq_norm(nrm_after);
                    HIPCHK(hipMemcpyAsync(h->red_host_vec, h->kh, sizeof(double) * (2 * (m + 1) + 2),
                                          hipMemcpyDeviceToHost, s));
                    HIPCHK(hipStreamSynchronize(s));
                    const double* hv = h->red_host_vec;
                    for (int i = 0; i <= j; ++i) H[i * m + j] += hv[(long)pass * (m + 1) + i];
                    if (pass == 0) w2_before = hv[2 * (m + 1) + 1];
                    hn2 = hv[2 * (m + 1)];
                    if (!(hn2 < 0.49 * w2_before)) break;
                }
            }
            const double hn = std::sqrt(hn2);
            H[(j + 1) * m + j] = hn;
            hn_prev = hn;  // V_{j+1} = w / hn is written by the next step
            finalize(j);
            ++iters;
            if (o.verbose) std::fprintf(stderr, "alamo_el: fgmres %d |r|_2 est = %.6e\n", iters, std::fabs(g[j + 1]));
            // Test the true inf-norm residual once the 2-norm reduction since
            // the restart, applied to the inf-norm at the restart, meets the
            // target; a failed test restarts from the updated x.
            if (std::fabs(g[j + 1]) / beta * inf_at_restart <= target || hn == 0.0) { ++j; break; }
        }
        // x += Z y with H y = g
        const int k = j;
        for (int i = k - 1; i >= 0; --i) {
            double acc = g[i];
            for (int q = i + 1; q < k; ++q) acc -= R[i * m + q] * y[q];
            y[i] = acc / R[i * m + i];
        }
        // one pass over x and Z (the accumulation order of k axpys: x + y_0 z_0 + ...)
        if (k > 0) {
            HIPCHK(hipMemcpyAsync(h->kh, y.data(), sizeof(double) * k, hipMemcpyHostToDevice, s));
            if (fp32 && z32)
                ael::k_multiaxpy<<<nblk(M0), kBlock, 0, s>>>(x, reinterpret_cast<const float*>(h->kz), k, M0, MPf,
                                                             h->kh, 1.0);
            else
                ael::k_multiaxpy<<<nblk(M0), kBlock, 0, s>>>(x, h->kz, k, M0, MP, h->kh, 1.0);
        }
        launch_residual<double>(h, l0, x, b, l0.d.C, l0.d.res, s);
        if ((rc = reduce_host<double>(h, l0.d.res, l0.d.res, M0, 0, &resnorm, s))) return rc;
        if (std::isnan(resnorm)) return fail(ALAMO_EL_ENAN, "NaN in the residual");
        if (resnorm <= target) return 0;
    }
    return ALAMO_EL_NOT_CONVERGED;
}

}  // namespace

// ---------------------------------------------------------------------------
// Composite AMR solve (src/Operator/Operator.cpp:230-251, 436-448, 526-633)
// ---------------------------------------------------------------------------
namespace {

// Faces (xlo, ylo, zlo, xhi, yhi, zhi) on which boundary region r lies, in the
// region numbering of region_of (BC::Operator::Elastic::Elastic::Face).
void region_faces(int r, bool f[6]) {
    for (int q = 0; q < 6; ++q) f[q] = false;
    if (r < 6) { f[r] = true; return; }
    if (r < 10) { const int c = r - 6; f[(c & 2) ? 4 : 1] = true; f[(c & 1) ? 5 : 2] = true; return; }
    if (r < 14) { const int c = r - 10; f[(c & 2) ? 5 : 2] = true; f[(c & 1) ? 3 : 0] = true; return; }
    if (r < 18) { const int c = r - 14; f[(c & 2) ? 3 : 0] = true; f[(c & 1) ? 4 : 1] = true; return; }
    const int c = r - 18;
    f[(c & 4) ? 3 : 0] = true;
    f[(c & 2) ? 4 : 1] = true;
    f[(c & 1) ? 5 : 2] = true;
}

int amr_reduce_max(alamo_el_t h, const double* v, long n, double* out, hipStream_t s) {
    return reduce_host<double>(h, v, v, n, 0, out, s);
}

// The handles of a composite solve: h[0] is the base, every other handle a
// patch whose parent h[par[i]] comes before it; a level may hold several
// patches (they share no nodes).  order lists the handles deepest first.
struct AmrCtx {
    int n;
    const alamo_el_t* h;
    double* const* x;
    const double* const* b;
    hipStream_t s;
    std::vector<int> par, depth, order;
    std::vector<std::vector<int>> kids;

    int init() {
        par.assign(n, -1);
        depth.assign(n, 0);
        kids.assign(n, {});
        for (int i = 1; i < n; ++i) {
            for (int j = 0; j < i; ++j)
                if (h[i]->amr_parent == h[j]) par[i] = j;
            if (par[i] < 0) return fail(ALAMO_EL_EINVAL, "every patch's parent must come before it in the list");
            depth[i] = depth[par[i]] + 1;
            kids[par[i]].push_back(i);
        }
        // patches of one level share no nodes: on some axis one ends before
        // the other begins, so that neither reaches the other's ghost layer
        for (int i = 1; i < n; ++i)
            for (int j = i + 1; j < n; ++j) {
                if (depth[i] != depth[j]) continue;
                const ael::Patch &A = P(i), &B = P(j);
                bool apart = false;
                for (int a = 0; a < 3; ++a) apart |= A.chi[a] < B.clo[a] || B.chi[a] < A.clo[a];
                if (!apart) return fail(ALAMO_EL_EINVAL, "patches of one level must not share nodes");
            }
        order.resize(n);
        for (int i = 0; i < n; ++i) order[i] = i;
        std::stable_sort(order.begin(), order.end(), [&](int u, int v) { return depth[u] > depth[v]; });
        return 0;
    }
    Level& L(int i) { return h[i]->lv[0]; }
    const ael::Patch& P(int i) { return h[i]->amr; }

    void reflux(int i, const double* rf, double* rc) {  // patch i into its parent
        const ael::Patch& Pi = P(i);
        const int q = par[i];
        const long m = (long)(Pi.chi[0] - Pi.clo[0] + 1) * (Pi.chi[1] - Pi.clo[1] + 1) * (Pi.chi[2] - Pi.clo[2] + 1);
        const int* dom = h[q]->amr_dom;
        ael::k_patch_reflux<<<nblk(m), kBlock, 0, s>>>(rf, L(i).L, Pi, rc, L(q).L, P(q).ox, P(q).oy, P(q).oz, dom[0],
                                                        dom[1], dom[2], 3);
    }
    void interp(int i, const double* c, double* f, int mode) {
        const int q = par[i];
        ael::k_patch_interp<<<nblk(L(i).L.N), kBlock, 0, s>>>(c, L(q).L, P(q).ox, P(q).oy, P(q).oz, f, L(i).L, P(i),
                                                               mode);
    }
    void ghost(int i, double* v, const double* src, int mode) {
        ael::k_patch_ghost<<<nblk(L(i).L.N), kBlock, 0, s>>>(v, src, L(i).L, P(i), mode);
    }
    void sync_solution() {  // average down (deepest first), then refresh the ghost layers
        for (int i : order) {
            if (i == 0) continue;
            const ael::Patch& Pi = P(i);
            const int q = par[i];
            const long m = (long)(Pi.chi[0] - Pi.clo[0] + 1) * (Pi.chi[1] - Pi.clo[1] + 1) * (Pi.chi[2] - Pi.clo[2] + 1);
            ael::k_patch_inject<<<nblk(m), kBlock, 0, s>>>(x[i], L(i).L, Pi, x[q], L(q).L, P(q).ox, P(q).oy, P(q).oz);
        }
        for (int u = n - 1; u >= 0; --u)
            if (order[u] != 0) interp(order[u], x[par[order[u]]], x[order[u]], 2);
    }
    // Composite residual on every handle (deepest first, each patch refluxed
    // into its parent) and its masked inf-norm.
    int residuals(double* norm) {
        for (int i : order) {
            Level& lv = L(i);
            launch_residual<double>(h[i], lv, x[i], b[i], lv.d.C, lv.d.res, s);
            if (i > 0) ghost(i, lv.d.res, nullptr, 0);
            for (int c : kids[i]) reflux(c, L(c).d.res, lv.d.res);
        }
        return masked_norm(norm, true);
    }
    // max |v| over the composite grid: each handle's res (use_res) or b, ghost
    // nodes excluded, nodes strictly covered by a child excluded.
    int masked_norm(double* norm, bool use_res) {
        *norm = 0.0;
        int rc;
        for (int i = 0; i < n; ++i) {
            Level& lv = L(i);
            const long M = 3 * lv.L.N;
            const double* v = use_res ? lv.d.res : b[i];
            double* t = lv.d.rescor;
            launch_copy<double>(t, v, M, s);
            for (int c : kids[i])
                ael::k_mask_covered<<<nblk(lv.L.N), kBlock, 0, s>>>(t, lv.L, P(i).ox, P(i).oy, P(i).oz, P(c), t);
            if (i > 0) ghost(i, t, nullptr, 0);
            double m;
            if ((rc = amr_reduce_max(h[i], t, M, &m, s))) return rc;
            if (std::isnan(m)) return fail(ALAMO_EL_ENAN, "NaN in the composite residual");
            *norm = std::max(*norm, m);
        }
        return 0;
    }
};

}  // namespace

extern "C" {

int alamo_amr_patch_create(alamo_el_t parent, const int clo[3], const int chi[3], alamo_el_t* out) {
    if (!parent || !clo || !chi || !out) return fail(ALAMO_EL_EINVAL, "null argument");
    if (parent->per) return fail(ALAMO_EL_EINVAL, "AMR patches under a periodic level are not supported");
    *out = nullptr;
    ael::Patch P{};
    const Lvl& PL = parent->lv[0].L;
    const int pn[3] = {PL.nx, PL.ny, PL.nz};
    const int po[3] = {parent->amr.ox, parent->amr.oy, parent->amr.oz};
    int ncell[3];
    double dx[3] = {PL.dx / 2, PL.dy / 2, PL.dz / 2};
    for (int a = 0; a < 3; ++a) {
        if (!(0 <= clo[a] && clo[a] < chi[a] && chi[a] <= parent->amr_dom[a]))
            return fail(ALAMO_EL_EINVAL, "patch cells must satisfy 0 <= clo < chi <= domain cells");
        P.clo[a] = clo[a];
        P.chi[a] = chi[a];
        P.gl[a] = clo[a] > 0;
        P.gh[a] = chi[a] < parent->amr_dom[a];
        // proper nesting: the patch and its ghost layer lie on the parent's unknowns
        const int ulo = po[a] + parent->amr.gl[a], uhi = po[a] + pn[a] - 1 - parent->amr.gh[a];
        if (clo[a] - P.gl[a] < ulo || chi[a] + P.gh[a] > uhi)
            return fail(ALAMO_EL_EINVAL, "patch not properly nested in its parent (needs one parent node of margin)");
        ncell[a] = 2 * (chi[a] - clo[a]) + P.gl[a] + P.gh[a];
    }
    P.ox = 2 * clo[0] - P.gl[0];
    P.oy = 2 * clo[1] - P.gl[1];
    P.oz = 2 * clo[2] - P.gl[2];
    alamo_el_t h;
    int rc;
    if ((rc = alamo_el_create(ncell, dx, parent->sym, 0, &h))) return rc;
    h->amr_parent = parent;
    h->amr = P;
    for (int a = 0; a < 3; ++a) h->amr_dom[a] = 2 * parent->amr_dom[a];
    h->uniform = parent->uniform;
    h->omega = parent->omega;
    // boundary rows: the parent's table on domain faces, ghost (displacement,
    // value from the parent level) on every region touching an inner face
    int t[78];
    double v[78];
    const bool inner[6] = {P.gl[0] != 0, P.gl[1] != 0, P.gl[2] != 0, P.gh[0] != 0, P.gh[1] != 0, P.gh[2] != 0};
    for (int r = 0; r < 26; ++r) {
        bool f[6];
        region_faces(r, f);
        bool g = false;
        for (int q = 0; q < 6; ++q) g = g || (f[q] && inner[q]);
        for (int c = 0; c < 3; ++c) {
            t[3 * r + c] = g ? ALAMO_EL_BC_DISPLACEMENT : parent->bct_host[3 * r + c];
            v[3 * r + c] = g ? 0.0 : parent->bcv_host[3 * r + c];
        }
    }
    if ((rc = alamo_el_set_bc(h, t, v))) {
        alamo_el_destroy(h);
        return rc;
    }
    *out = h;
    return 0;
}

int alamo_el_tag_bbox(alamo_el_t h, const double* eta, int ncomp, double threshold, int buffer, int clo[3],
                      int chi[3], long long* ntag, hipStream_t s) {
    if (!h || !eta || ncomp < 1 || !clo || !chi || !ntag || buffer < 0) return fail(ALAMO_EL_EINVAL, "bad argument");
    const Lvl L = h->lv[0].L;
    int* bb = reinterpret_cast<int*>(h->red_partial);
    ael::k_tag_init<<<1, kBlock, 0, s>>>(bb);
    ael::k_tag_eta<<<nblk(L.N), kBlock, 0, s>>>(eta, ncomp, L, threshold, bb);
    HIPCHK(hipGetLastError());
    int out[7];
    HIPCHK(hipMemcpyAsync(out, bb, sizeof(out), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    *ntag = out[6];
    if (out[6] == 0) return 0;
    const int nc[3] = {L.nx - 1, L.ny - 1, L.nz - 1};
    for (int a = 0; a < 3; ++a) {
        clo[a] = std::max(out[a] - buffer, 0);
        chi[a] = std::min(out[3 + a] + 1 + buffer, nc[a]);
    }
    return 0;
}

int alamo_el_tag_cells(alamo_el_t h, const double* eta, int ncomp, double threshold, unsigned char* tags,
                       hipStream_t s) {
    if (!h || !eta || ncomp < 1 || !tags) return fail(ALAMO_EL_EINVAL, "bad argument");
    const Lvl L = h->lv[0].L;
    ael::k_tag_cells<<<nblk(L.N), kBlock, 0, s>>>(eta, ncomp, L, threshold, tags);
    HIPCHK(hipGetLastError());
    return 0;
}

int alamo_amr_cluster(const unsigned char* tags_host, const int ncell[3], const int origin[3], int buffer,
                      int blocking_factor, double grid_eff, int max_boxes, int* boxes, int* nboxes) {
    if (!tags_host || !ncell || !origin || !boxes || !nboxes || buffer < 0 || blocking_factor < 0 ||
        !(grid_eff > 0.0 && grid_eff <= 1.0) || max_boxes < 1)
        return fail(ALAMO_EL_EINVAL, "bad argument");
    for (int a = 0; a < 3; ++a)
        if (ncell[a] < 1) return fail(ALAMO_EL_EINVAL, "empty tag array");
    *nboxes = 0;
    const std::vector<ael::cluster::Box> bs =
        ael::cluster::make_boxes(tags_host, ncell, origin, buffer, blocking_factor, grid_eff);
    if ((int)bs.size() > max_boxes) return fail(ALAMO_EL_EINVAL, "more boxes than max_boxes");
    for (size_t q = 0; q < bs.size(); ++q)
        for (int a = 0; a < 3; ++a) {
            boxes[6 * q + a] = bs[q].lo[a];
            boxes[6 * q + 3 + a] = bs[q].hi[a] + 1;
        }
    *nboxes = (int)bs.size();
    return 0;
}

int alamo_amr_reflux(alamo_el_t patch, const double* fine_res, double* parent_res, hipStream_t s) {
    if (!patch || !fine_res || !parent_res) return fail(ALAMO_EL_EINVAL, "null argument");
    if (!patch->amr_parent) return fail(ALAMO_EL_EINVAL, "reflux needs a patch (alamo_amr_patch_create)");
    Level& lv = patch->lv[0];
    const alamo_el_t par = patch->amr_parent;
    const ael::Patch& P = patch->amr;
    // the patch residual with its ghost rows zeroed, in the patch's own scratch
    launch_copy<double>(lv.d.rescor, fine_res, 3 * lv.L.N, s);
    ael::k_patch_ghost<<<nblk(lv.L.N), kBlock, 0, s>>>(lv.d.rescor, nullptr, lv.L, P, 0);
    const long m = (long)(P.chi[0] - P.clo[0] + 1) * (P.chi[1] - P.clo[1] + 1) * (P.chi[2] - P.clo[2] + 1);
    ael::k_patch_reflux<<<nblk(m), kBlock, 0, s>>>(lv.d.rescor, lv.L, P, parent_res, par->lv[0].L, par->amr.ox,
                                                    par->amr.oy, par->amr.oz, par->amr_dom[0], par->amr_dom[1],
                                                    par->amr_dom[2], 3);
    HIPCHK(hipGetLastError());
    return 0;
}

int alamo_amr_average_down_coeffs(alamo_el_t patch, const double* patch_C, double* parent_C, hipStream_t s) {
    if (!patch || !patch_C || !parent_C) return fail(ALAMO_EL_EINVAL, "null argument");
    if (!patch->amr_parent) return fail(ALAMO_EL_EINVAL, "averaging down needs a patch (alamo_amr_patch_create)");
    const alamo_el_t par = patch->amr_parent;
    const ael::Patch& P = patch->amr;
    const long m = (long)(P.chi[0] - P.clo[0] + 1) * (P.chi[1] - P.clo[1] + 1) * (P.chi[2] - P.clo[2] + 1);
    // the patch's C planes, ghost layer included, restricted onto the parent's
    // nodes clo..chi with the reflux weights
    ael::k_patch_reflux<<<nblk(m), kBlock, 0, s>>>(patch_C, patch->lv[0].L, P, parent_C, par->lv[0].L, par->amr.ox,
                                                    par->amr.oy, par->amr.oz, par->amr_dom[0], par->amr_dom[1],
                                                    par->amr_dom[2], patch->nc);
    HIPCHK(hipGetLastError());
    return 0;
}

int alamo_el_box_copy(double* dst, const int dst_n[3], const int dst_origin[3], const double* src, const int src_n[3],
                      const int src_origin[3], const int lo[3], const int hi[3], int ncomp, hipStream_t s) {
    if (!dst || !src || !dst_n || !dst_origin || !src_n || !src_origin || !lo || !hi || ncomp < 1)
        return fail(ALAMO_EL_EINVAL, "null argument");
    int b[3];
    for (int a = 0; a < 3; ++a) {
        if (dst_n[a] < 1 || src_n[a] < 1 || hi[a] < lo[a]) return fail(ALAMO_EL_EINVAL, "bad extents");
        if (lo[a] < dst_origin[a] || hi[a] > dst_origin[a] + dst_n[a] - 1 || lo[a] < src_origin[a] ||
            hi[a] > src_origin[a] + src_n[a] - 1)
            return fail(ALAMO_EL_EINVAL, "the box must lie inside both arrays");
        b[a] = hi[a] - lo[a] + 1;
    }
    const long n = (long)b[0] * b[1] * b[2];
    ael::k_box_copy<<<nblk(n), kBlock, 0, s>>>(dst, dst_n[0], dst_n[1], dst_n[2], dst_origin[0], dst_origin[1],
                                                dst_origin[2], src, src_n[0], src_n[1], src_n[2], src_origin[0],
                                                src_origin[1], src_origin[2], lo[0], lo[1], lo[2], b[0], b[1], b[2],
                                                ncomp);
    HIPCHK(hipGetLastError());
    return 0;
}

int alamo_amr_patch_info(alamo_el_t h, int origin[3], int ghost_lo[3], int ghost_hi[3]) {
    if (!h) return fail(ALAMO_EL_EINVAL, "null handle");
    const int o[3] = {h->amr.ox, h->amr.oy, h->amr.oz};
    for (int a = 0; a < 3; ++a) {
        if (origin) origin[a] = o[a];
        if (ghost_lo) ghost_lo[a] = h->amr.gl[a];
        if (ghost_hi) ghost_hi[a] = h->amr.gh[a];
    }
    return 0;
}

int alamo_amr_fillpatch(alamo_el_t dst, const double* parent_u, alamo_el_t src, const double* src_u, double* dst_u,
                        hipStream_t s) {
    if (!dst || !parent_u || !dst_u) return fail(ALAMO_EL_EINVAL, "null argument");
    if (!dst->amr_parent) return fail(ALAMO_EL_EINVAL, "the destination must be a patch (alamo_amr_patch_create)");
    if (src) {
        if (!src_u) return fail(ALAMO_EL_EINVAL, "an old patch needs its field");
        if (!src->amr_parent) return fail(ALAMO_EL_EINVAL, "the old patch must be a patch (alamo_amr_patch_create)");
        for (int a = 0; a < 3; ++a)
            if (src->amr_dom[a] != dst->amr_dom[a])
                return fail(ALAMO_EL_EINVAL, "the old and new patches must belong to the same AMR level");
    }
    const alamo_el_t par = dst->amr_parent;
    const Lvl& Lf = dst->lv[0].L;
    const Lvl& Ls = src ? src->lv[0].L : Lf;
    ael::k_patch_fill<<<nblk(Lf.N), kBlock, 0, s>>>(parent_u, par->lv[0].L, par->amr.ox, par->amr.oy, par->amr.oz,
                                                     src ? src_u : nullptr, Ls, src ? src->amr : dst->amr, dst_u, Lf,
                                                     dst->amr);
    HIPCHK(hipGetLastError());
    return 0;
}

int alamo_amr_solve(int nlev, const alamo_el_t* hs, double* const* x, const double* const* b, double tol_rel,
                    double tol_abs, const alamo_el_opts* opts_in, alamo_el_info* info, hipStream_t s) {
    if (nlev < 1 || !hs || !x || !b) return fail(ALAMO_EL_EINVAL, "bad argument");
    for (int l = 0; l < nlev; ++l) {
        if (!hs[l] || !x[l] || !b[l]) return fail(ALAMO_EL_EINVAL, "null level");
        if ((l == 0) != (hs[l]->amr_parent == nullptr))
            return fail(ALAMO_EL_EINVAL, "the list must hold a base handle followed by its nested patches");
        if (!hs[l]->model_set) return fail(ALAMO_EL_ESTATE, "Attempting to use operator before calling SetModel");
    }
    if (nlev == 1) return alamo_el_solve(hs[0], x[0], b[0], tol_rel, tol_abs, opts_in, info, s);
    alamo_el_opts o;
    alamo_el_default_opts(&o);
    if (opts_in) o = *opts_in;
    if (o.max_iter < 0 || o.nu1 < 0 || o.nu2 < 0 || !(o.omega > 0) ||
        (o.smoother != ALAMO_EL_SMOOTH_JACOBI && o.smoother != ALAMO_EL_SMOOTH_RBGS))
        return fail(ALAMO_EL_EINVAL, "bad solver options");
    const auto t0 = std::chrono::steady_clock::now();
    int rc;
    AmrCtx A{nlev, hs, x, b, s};
    if ((rc = A.init())) return rc;
    for (int l = 0; l < nlev; ++l) {
        hs[l]->omega = o.omega;
        if (!hs[l]->prepared && (rc = prepare(hs[l], s))) return rc;
        if (l == 0 && (rc = psi_bottom_check(hs[0], o))) return rc;
    }
    const bool fp32 = o.precision == ALAMO_EL_PREC_FP32_IR;
    if (fp32 && (rc = prepare32(hs[0], s))) return rc;
    A.sync_solution();
    double r0, bn;
    if ((rc = A.residuals(&r0))) return rc;
    if ((rc = A.masked_norm(&bn, false))) return rc;
    const double target = std::max(tol_abs, std::max(tol_rel, 1.e-16) * std::max(r0, bn));
    Cycle<double> c64{hs[0], s, o.nu1, o.nu2, o.final_smooth, o.bottom_max_iter, o.bottom_tol_rel, o.bottom_solver,
                      o.smoother};
    Cycle<float> c32{hs[0], s, o.nu1, o.nu2, o.final_smooth, o.bottom_max_iter, o.bottom_tol_rel, o.bottom_solver,
                     o.smoother};
    double nrm = r0;
    int iters = 0, status = nrm <= target ? 0 : ALAMO_EL_NOT_CONVERGED;
    const long M0 = 3 * A.L(0).L.N;
    // nu Fsmooth calls of patch i on its correction (ghost rows held)
    auto smooth_patch = [&](int i, int nu) {
        Level& lv = A.L(i);
        if (o.smoother == ALAMO_EL_SMOOTH_RBGS) {
            Cycle<double> ci{hs[i], s, o.nu1, o.nu2, o.final_smooth, o.bottom_max_iter, o.bottom_tol_rel,
                             o.bottom_solver, o.smoother};
            ci.smooth(0, lv.d.cor, lv.d.res, nu);
            return;
        }
        for (int q = 0; q < 2 * nu; ++q) {
            double* src = (q % 2 == 0) ? lv.d.cor : lv.d.tmp;
            double* dst = (q % 2 == 0) ? lv.d.tmp : lv.d.cor;
            launch_jacobi<double>(hs[i], lv, src, lv.d.res, lv.d.diag, lv.d.C, dst, s);
        }
    };
    for (int it = 0; it < iter_limit(o) && status != 0; ++it) {
        // down: pre-smooth each patch on its composite residual (deepest
        // first), reflux the remaining residual into its parent
        for (int i : A.order) {
            if (i == 0) continue;
            Level& lv = A.L(i);
            const long M = 3 * lv.L.N;
            launch_fill<double>(lv.d.cor, 0.0, M, s);
            A.ghost(i, lv.d.res, lv.d.cor, 1);
            smooth_patch(i, o.nu1);
            launch_residual<double>(hs[i], lv, lv.d.cor, lv.d.res, lv.d.C, lv.d.rescor, s);
            A.ghost(i, lv.d.rescor, nullptr, 0);
            A.reflux(i, lv.d.rescor, A.L(A.par[i]).d.res);
        }
        // base level: one multigrid cycle on its composite residual
        if (fp32) {
            ael::k_cast<double, float><<<nblk(M0), kBlock, 0, s>>>(A.L(0).d.res, A.L(0).f.res, M0);
            c32.run(cycle_of(o, it));
            ael::k_cast<float, double><<<nblk(M0), kBlock, 0, s>>>(A.L(0).f.cor, A.L(0).d.cor, M0);
        } else {
            c64.run(cycle_of(o, it));
        }
        // up: add the interpolated parent correction, post-smooth (parents first)
        for (int u = nlev - 1; u >= 0; --u) {
            const int i = A.order[u];
            if (i == 0) continue;
            Level& lv = A.L(i);
            A.interp(i, A.L(A.par[i]).d.cor, lv.d.cor, 1);
            A.ghost(i, lv.d.res, lv.d.cor, 1);
            smooth_patch(i, o.nu2);
        }
        for (int l = 0; l < nlev; ++l) launch_axpy<double>(x[l], 1.0, A.L(l).d.cor, 3 * A.L(l).L.N, s);
        A.sync_solution();
        if ((rc = A.residuals(&nrm))) return rc;
        iters = it + 1;
        if (o.verbose) std::fprintf(stderr, "alamo_amr: iter %d |r|_inf = %.6e\n", iters, nrm);
        if (nrm <= target) status = 0;
    }
    HIPCHK(hipGetLastError());
    if (status == ALAMO_EL_NOT_CONVERGED && o.fixed_iter > 0 && iters >= iter_limit(o)) status = 0;
    if (info) {
        info->status = status;
        info->iters = iters;
        info->res0 = r0;
        info->rhsnorm = bn;
        info->res = nrm;
        info->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        info->learned_steps = 0;
    }
    if (status == ALAMO_EL_NOT_CONVERGED) g_err = "composite MLMG did not converge within max_iter";
    return status;
}

}  // extern "C"
