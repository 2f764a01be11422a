// inverse_kernels.h -- FP64 inverse of the direct bottom level's operator
// (M x M, column-major, M <= 4096) by blocked Gauss-Jordan elimination with
// partial pivoting, in place.
//
// In-place Gauss-Jordan with row swaps processes pivot column k as: swap rows
// k and p (p >= k, the largest |W[i][k]|), W[k][k] = 1, row k /= pivot, and
// for every other row i: f = W[i][k], W[i][k] = 0, row i -= f row k.  Column
// k then holds the k-th column of the transform accumulated so far, and after
// all M columns W holds the inverse with its columns permuted; swapping
// columns k and p back, from the last step to the first, gives the inverse.
//
// The steps of a block of NB columns touch the other columns only through
// the block's row swaps and one rank-NB update: after the swaps, a column y
// becomes y with rows k0 .. k0+NB-1 zeroed plus G times those rows of y, G
// being the block's columns once the block is factored.  So the M pivot steps
// run in one workgroup with a panel of 16 columns (8 when M > 3072) in
// registers (k_inv_panel); each panel then updates the other columns of its
// block (k_inv_block, rank 16), and each block of kInvBlock columns all the
// columns outside it (k_inv_rowmap, k_inv_gemm: rank kInvBlock, an LDS-tiled
// product, out of place).  The pivot steps are the sequential part: one
// barrier each.
#pragma once
#include <type_traits>
#include <utility>

#include "common.h"

namespace ael {

constexpr int kInvPanelMax = 16;   // widest panel factored in registers
constexpr int kInvBlock = 64;      // columns per trailing update
constexpr int kInvThreads = 512;   // threads of the panel kernel
constexpr int kInvMax = 4096;      // largest M
constexpr int kInvTile = 64;       // output tile of the trailing products

struct InvState {
    int singular;  // set when a pivot column holds no nonzero below the diagonal
    double resid;  // max |A (Ainv x) - x| / max |x| of the check (k_inv_check)
};

// Panel width for an M x M inverse: 16 columns while a thread's rows of the
// panel fit in registers (5 rows x 16 at M <= 2560), 8 beyond.
inline int inv_panel_width(int M) { return M <= 5 * kInvThreads ? 16 : 8; }

// f(std::integral_constant<int, 0>{}), ..., f(std::integral_constant<int, N - 1>{}).
template <typename F, int... I>
__device__ __forceinline__ void static_for_impl(F&& f, std::integer_sequence<int, I...>) {
    (f(std::integral_constant<int, I>{}), ...);
}
template <int N, typename F>
__device__ __forceinline__ void static_for(F&& f) {
    static_for_impl(f, std::make_integer_sequence<int, N>{});
}

// Pivot key of row r: the upper half of |a|'s bits (exponent and the top 8
// mantissa bits after the low 12 bits are cleared) above 4095 - r, so that
// the unsigned maximum is a row whose |a| is within 2^-8 of the largest,
// the lowest such row on ties; 0 value bits mean |a| below 2^-1022.
__device__ __forceinline__ unsigned inv_pivot_key(double a, int r) {
    const unsigned hi = (unsigned)((unsigned long long)__double_as_longlong(fabs(a)) >> 32);
    return (hi & ~0xFFFu) | (unsigned)(4095 - r);
}
__device__ __forceinline__ int inv_key_row(unsigned key) { return 4095 - (int)(key & 0xFFFu); }

// Maximum over the 64 lanes of a wave (every lane active), uniform result:
// cross-lane moves within 16-lane rows (quad xor 1, quad xor 2, half-row
// mirror, row mirror), then the four rows' lane 0 through readlane.
__device__ __forceinline__ unsigned wave_max_u32(unsigned x) {
    x = max(x, (unsigned)__builtin_amdgcn_update_dpp(0, (int)x, 0xB1, 0xF, 0xF, false));
    x = max(x, (unsigned)__builtin_amdgcn_update_dpp(0, (int)x, 0x4E, 0xF, 0xF, false));
    x = max(x, (unsigned)__builtin_amdgcn_update_dpp(0, (int)x, 0x141, 0xF, 0xF, false));
    x = max(x, (unsigned)__builtin_amdgcn_update_dpp(0, (int)x, 0x140, 0xF, 0xF, false));
    const unsigned a = (unsigned)__builtin_amdgcn_readlane((int)x, 0), b = (unsigned)__builtin_amdgcn_readlane((int)x, 16),
                   c = (unsigned)__builtin_amdgcn_readlane((int)x, 32), d = (unsigned)__builtin_amdgcn_readlane((int)x, 48);
    return max(max(a, b), max(c, d));
}

// The nb <= NB pivot steps of columns c0 .. c0 + nb - 1 on those columns of
// W only.  Thread t holds rows t + q kInvThreads (q < ROWS) in registers;
// perm[k] receives the row swapped with row k.  One barrier a step: before
// it, each wave publishes its best pivot key and that row, and the owner of
// row k publishes row k (double-buffered by step parity); after it, every
// thread takes the best of the waves' keys and reads the pivot row and row k.
template <int ROWS, int NB>
__global__ void __launch_bounds__(kInvThreads) k_inv_panel(double* __restrict__ W, int M, int c0, int nb,
                                                           int* __restrict__ perm, InvState* __restrict__ st) {
    constexpr int NW = kInvThreads / 64;
    __shared__ unsigned keys[2][NW];
    __shared__ double cand[2][NW][NB];
    __shared__ double krow[2][NB];
    const int t = (int)threadIdx.x, lane = t & 63, wave = t >> 6;
    double v[ROWS][NB];
#pragma unroll
    for (int q = 0; q < ROWS; ++q) {
        const int r = t + q * kInvThreads;
#pragma unroll
        for (int c = 0; c < NB; ++c) v[q][c] = (r < M && c < nb) ? W[(long)(c0 + c) * M + r] : 0.0;
    }
    // the steps as a compile-time sequence: column j of the registers is then
    // a fixed register, never an indexed access
    static_for<NB>([&](auto jc) __attribute__((always_inline)) {
        constexpr int j = decltype(jc)::value;
        if (j >= nb) return;  // uniform
        const int k = c0 + j, par = j & 1;
        unsigned best = 0u;
#pragma unroll
        for (int q = 0; q < ROWS; ++q) {
            const int r = t + q * kInvThreads;
            if (r >= k && r < M) best = max(best, inv_pivot_key(v[q][j], r));
        }
        const unsigned wbest = wave_max_u32(best);
        // rows leave the registers through one branch per row slot, taken by
        // one lane (an indexed read would move v to scratch memory, a select
        // per element would cost every wave)
        if ((wbest >> 12) != 0u) {  // a real row of this wave: its owner publishes it
            const int rw = inv_key_row(wbest);
#pragma unroll
            for (int q = 0; q < ROWS; ++q)
                if (t + q * kInvThreads == rw) {
#pragma unroll
                    for (int c = 0; c < NB; ++c) cand[par][wave][c] = v[q][c];
                }
        }
#pragma unroll
        for (int q = 0; q < ROWS; ++q)
            if (t + q * kInvThreads == k) {
#pragma unroll
                for (int c = 0; c < NB; ++c) krow[par][c] = v[q][c];
            }
        if (lane == 0) keys[par][wave] = wbest;
        __syncthreads();
        unsigned g = keys[par][0];
        int gw = 0;
#pragma unroll
        for (int w = 1; w < NW; ++w) {
            const unsigned kw = keys[par][w];
            if (kw > g) {
                g = kw;
                gw = w;
            }
        }
        const bool zero = (g >> 12) == 0u;  // no usable pivot: singular, row k stays
        const int p = zero ? k : inv_key_row(g);
        double pr[NB];
#pragma unroll
        for (int c = 0; c < NB; ++c) pr[c] = zero ? krow[par][c] : cand[par][gw][c];
        if (t == 0) {
            perm[k] = p;
            if (zero || !(pr[j] != 0.0)) st->singular = 1;
        }
        // the new row k: the pivot row times 1/d (one division a step)
        const double rd = 1.0 / pr[j];
#pragma unroll
        for (int c = 0; c < NB; ++c) pr[c] = c == j ? rd : pr[c] * rd;
#pragma unroll
        for (int q = 0; q < ROWS; ++q) {
            if (p != k && t + q * kInvThreads == p) {
#pragma unroll
                for (int c = 0; c < NB; ++c) v[q][c] = krow[par][c];  // old row k lands on row p
            }
            const double f = v[q][j];
#pragma unroll
            for (int c = 0; c < NB; ++c) v[q][c] = c == j ? -f * rd : v[q][c] - f * pr[c];
            if (t + q * kInvThreads == k) {
#pragma unroll
                for (int c = 0; c < NB; ++c) v[q][c] = pr[c];  // the new row k
            }
        }
    });
#pragma unroll
    for (int q = 0; q < ROWS; ++q) {
        const int r = t + q * kInvThreads;
        if (r < M) {
#pragma unroll
            for (int c = 0; c < NB; ++c)
                if (c < nb) W[(long)(c0 + c) * M + r] = v[q][c];
        }
    }
}

// k_inv_panel for an M x M inverse: the row count a thread holds and the
// panel width (inv_panel_width) follow from M <= kInvMax (at most 5 rows of
// 16 columns a thread, else up to 8 rows of 8).
inline void launch_inv_panel(double* W, int M, int c0, int nb, int* perm, InvState* st, hipStream_t s) {
    if (M <= kInvThreads) k_inv_panel<1, 16><<<1, kInvThreads, 0, s>>>(W, M, c0, nb, perm, st);
    else if (M <= 2 * kInvThreads) k_inv_panel<2, 16><<<1, kInvThreads, 0, s>>>(W, M, c0, nb, perm, st);
    else if (M <= 3 * kInvThreads) k_inv_panel<3, 16><<<1, kInvThreads, 0, s>>>(W, M, c0, nb, perm, st);
    else if (M <= 4 * kInvThreads) k_inv_panel<4, 16><<<1, kInvThreads, 0, s>>>(W, M, c0, nb, perm, st);
    else if (M <= 5 * kInvThreads) k_inv_panel<5, 16><<<1, kInvThreads, 0, s>>>(W, M, c0, nb, perm, st);
    else if (M <= 6 * kInvThreads) k_inv_panel<6, 8><<<1, kInvThreads, 0, s>>>(W, M, c0, nb, perm, st);
    else k_inv_panel<8, 8><<<1, kInvThreads, 0, s>>>(W, M, c0, nb, perm, st);
}

// Row map of the swaps of steps k0 .. k0 + n - 1 (n <= kInvBlock) in LDS:
// afterwards row r holds the value row map[r] held before (identity where
// untouched).  The caller's threads fill the identity and stage the swaps;
// thread 0 applies them in order, from LDS.
__device__ __forceinline__ void swap_map(const int* __restrict__ perm, int M, int k0, int n, int* map, int* ps) {
    for (int r = (int)threadIdx.x; r < M; r += blockDim.x) map[r] = r;
    if ((int)threadIdx.x < n) ps[threadIdx.x] = perm[k0 + threadIdx.x];
    __syncthreads();
    if (threadIdx.x == 0)
        for (int j = 0; j < n; ++j) {
            const int k = k0 + j, p = ps[j];
            const int a = map[k];
            map[k] = map[p];
            map[p] = a;
        }
    __syncthreads();
}

// The panel's swaps and rank-nb update on the other columns of its block
// [b0, b1): one workgroup per column.  Only rows c0 .. c0 + nb - 1 and the
// rows swapped with them change under the swaps; thread j moves rows c0 + j
// and perm[c0 + j] (a row reached twice takes the same value twice).  The
// update then takes kInvBatch rows a thread at a time, every load of a batch
// issued before its arithmetic.
constexpr int kInvBatch = 4;
__global__ void __launch_bounds__(kBlock) k_inv_block(double* __restrict__ W, int M, int b0, int b1, int c0, int nb,
                                                      const int* __restrict__ perm) {
    __shared__ int map[kInvMax], ps[kInvPanelMax];
    __shared__ double ys[kInvPanelMax];
    int col = b0 + (int)blockIdx.x;
    if (col >= c0) col += nb;  // skip the panel itself
    if (col >= b1) return;     // uniform over the workgroup
    // the column and the panel's columns never overlap
    double* __restrict__ y = W + (long)col * M;
    const double* __restrict__ G = W + (long)c0 * M;
    swap_map(perm, M, c0, nb, map, ps);
    const int j = (int)threadIdx.x;
    double a = 0.0, b = 0.0;
    int ra = 0, rb = 0;
    if (j < nb) {
        ra = c0 + j;
        rb = ps[j];
        a = y[map[ra]];
        b = y[map[rb]];
    }
    __syncthreads();
    if (j < nb) {
        y[ra] = a;
        y[rb] = b;
    }
    if (j < kInvPanelMax) ys[j] = j < nb ? a : 0.0;
    __syncthreads();
    double yq[kInvPanelMax];
#pragma unroll
    for (int q = 0; q < kInvPanelMax; ++q) yq[q] = ys[q];
    for (int r0 = (int)threadIdx.x; r0 < M; r0 += kInvBatch * kBlock) {
        double yv[kInvBatch], g[kInvBatch][kInvPanelMax];
#pragma unroll
        for (int e = 0; e < kInvBatch; ++e) {
            const int r = r0 + e * kBlock;
            const bool in = r < M;
            yv[e] = (in && !(r >= c0 && r < c0 + nb)) ? y[r] : 0.0;
#pragma unroll
            for (int q = 0; q < kInvPanelMax; ++q) g[e][q] = (in && q < nb) ? G[(long)q * M + r] : 0.0;
        }
#pragma unroll
        for (int e = 0; e < kInvBatch; ++e) {
            const int r = r0 + e * kBlock;
            if (r >= M) break;
            double acc = yv[e];
#pragma unroll
            for (int q = 0; q < kInvPanelMax; ++q) acc += g[e][q] * yq[q];
            y[r] = acc;
        }
    }
}

// Row map of a block's kInvBlock swaps, to global memory.
__global__ void __launch_bounds__(kBlock) k_inv_rowmap(const int* __restrict__ perm, int M, int k0, int nbk,
                                                       int* __restrict__ map) {
    __shared__ int sm[kInvMax], ps[kInvBlock];
    swap_map(perm, M, k0, nbk, sm, ps);
    for (int r = (int)threadIdx.x; r < M; r += blockDim.x) map[r] = sm[r];
}

// Out of place, Wo from Wi: the block's columns [k0, k0 + kInvBlock) are
// copied; every other column y becomes, with ys its rows after the block's
// swaps (ys[r] = y[map[r]]), ys with the block rows zeroed plus G ys[block
// rows], G = Wi[:, block].  A 64 x 64 tile of outputs per workgroup (rows
// r0 .., columns c0 ..), G's and the columns' block rows staged in LDS, 4 x 4
// outputs per thread.
__global__ void __launch_bounds__(kBlock) k_inv_gemm(const double* __restrict__ Wi, double* __restrict__ Wo, int M,
                                                     int k0, int nbk, const int* __restrict__ map) {
    __shared__ double gs[kInvBlock][kInvTile + 1];  // gs[j][r]
    __shared__ double ps[kInvBlock][kInvTile + 1];  // ps[j][c]
    const int r0 = (int)blockIdx.x * kInvTile, cb = (int)blockIdx.y * kInvTile;
    const int t = (int)threadIdx.x;
    if (cb >= k0 && cb < k0 + kInvBlock) {  // the block's own columns (k0 is a multiple of kInvTile)
        for (int e = t; e < kInvTile * kInvTile; e += kBlock) {
            const int c = cb + e / kInvTile, r = r0 + e % kInvTile;
            if (c < M && r < M) Wo[(long)c * M + r] = Wi[(long)c * M + r];
        }
        return;
    }
    for (int e = t; e < kInvBlock * kInvTile; e += kBlock) {
        const int j = e / kInvTile, rr = e % kInvTile;
        const int r = r0 + rr, c = cb + rr;
        gs[j][rr] = (r < M && j < nbk) ? Wi[(long)(k0 + j) * M + r] : 0.0;
        ps[j][rr] = (c < M && j < nbk) ? Wi[(long)c * M + map[k0 + j]] : 0.0;
    }
    __syncthreads();
    const int tr = (t % 16) * 4, tc = (t / 16) * 4;
    double acc[4][4];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) acc[a][b] = 0.0;
    for (int j = 0; j < kInvBlock; ++j) {
        double g[4], pv[4];
#pragma unroll
        for (int a = 0; a < 4; ++a) g[a] = gs[j][tr + a];
#pragma unroll
        for (int b = 0; b < 4; ++b) pv[b] = ps[j][tc + b];
#pragma unroll
        for (int a = 0; a < 4; ++a)
#pragma unroll
            for (int b = 0; b < 4; ++b) acc[a][b] += g[a] * pv[b];
    }
#pragma unroll
    for (int b = 0; b < 4; ++b) {
        const int c = cb + tc + b;
        if (c >= M) continue;
        const double* yi = Wi + (long)c * M;
        double* yo = Wo + (long)c * M;
#pragma unroll
        for (int a = 0; a < 4; ++a) {
            const int r = r0 + tr + a;
            if (r >= M) continue;
            const bool blk = r >= k0 && r < k0 + nbk;
            yo[r] = (blk ? 0.0 : yi[map[r]]) + acc[a][b];
        }
    }
}

// --- Block Gauss-Jordan without pivoting across blocks ------------------------
// The fast path.  For each block b of kInvBlock rows and columns, with D the
// block's diagonal block of the current W:
//   W[b, b] <- D^-1,  W[b, j] <- D^-1 W[b, j],  W[i, b] <- -W[i, b] D^-1,
//   W[i, j] <- W[i, j] - W[i, b] D^-1 W[b, j]          (i, j outside b);
// after the last block W = A^-1 (no permutation).  D^-1 is formed with
// partial pivoting inside the block (k_inv_diag); the rows are never swapped
// across blocks, which is stable for the elastic operators here (the
// inverse's residual matches a partially pivoted one), and k_inv_check
// measures the residual of every inverse so that a failing case can fall
// back to the pivoted panels above.

// D^-1 of the nb x nb diagonal block at (k0, k0) of Wi (nb <= kInvBlock) by
// Gauss-Jordan with partial pivoting, rows never moved: step j pivots on the
// unused row p_j with the largest |a[p][j]| (pivot keys as above), scales
// that row and eliminates column j from all others.  Afterwards D^-1[j][k] =
// a[p_j][q_k] with q the inverse of p.  Wave w holds columns 16 w .. 16 w + 15
// of all 64 rows in registers (lane r: row r); one barrier a step, after the
// wave that holds column j has published it and the pivot row.  Dinv
// receives D^-1 column-major with leading dimension kInvBlock (zeros outside
// nb x nb); a vanishing pivot sets st->singular.
__global__ void __launch_bounds__(kBlock) k_inv_diag(const double* __restrict__ Wi, int M, int k0, int nb,
                                                     double* __restrict__ Dinv, InvState* __restrict__ st) {
    static_assert(kBlock == 4 * kInvBlock && kInvBlock == 64, "four waves of 16 columns");
    __shared__ double colj[2][kInvBlock];
    __shared__ int pis[2];
    __shared__ int prow[kInvBlock], pinv[kInvBlock];
    const int t = (int)threadIdx.x, r = t & 63, w = t >> 6;
    double v[16];
#pragma unroll
    for (int c = 0; c < 16; ++c) {
        const int col = 16 * w + c;
        v[c] = (r < nb && col < nb) ? Wi[(long)(k0 + col) * M + k0 + r] : (r == col ? 1.0 : 0.0);
    }
    bool used = r >= nb;  // rows beyond the block never pivot
    static_for<kInvBlock>([&](auto jc_) __attribute__((always_inline)) {
        constexpr int j = decltype(jc_)::value, jw = j / 16, jc = j % 16, par = j & 1;
        if (j >= nb) return;  // uniform
        if (w == jw) {        // the wave of column j: the pivot among the unused rows, and the column
            // an unused row's key is at least 4095 - 63: p is always an unused row
            const unsigned m = wave_max_u32(used ? 0u : inv_pivot_key(v[jc], r));
            colj[par][r] = v[jc];
            if (r == 0) {
                pis[par] = inv_key_row(m);
                if ((m >> 12) == 0u) st->singular = 1;
            }
        }
        __syncthreads();
        const int p = pis[par];
        const double rd = 1.0 / colj[par][p], f = colj[par][r];
        double sv[16];  // the pivot row scaled, this wave's columns
#pragma unroll
        for (int c = 0; c < 16; ++c) {
            const unsigned long long bits = (unsigned long long)__double_as_longlong(v[c]);
            const int lo = __builtin_amdgcn_readlane((int)(unsigned)bits, p);
            const int hi = __builtin_amdgcn_readlane((int)(unsigned)(bits >> 32), p);
            const double x = __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo));
            sv[c] = (c == jc && w == jw) ? rd : x * rd;
        }
        if (r == p) {
            used = true;
#pragma unroll
            for (int c = 0; c < 16; ++c) v[c] = sv[c];
        } else {
#pragma unroll
            for (int c = 0; c < 16; ++c) v[c] = (c == jc && w == jw) ? -f * rd : v[c] - f * sv[c];
        }
        if (t == 0) {
            prow[j] = p;
            pinv[p] = j;
        }
    });
    __syncthreads();
#pragma unroll
    for (int c = 0; c < 16; ++c) {
        const int col = 16 * w + c;
        if (r < nb && col < nb) Dinv[(long)prow[col] * kInvBlock + pinv[r]] = v[c];
        else Dinv[(long)col * kInvBlock + r] = 0.0;
    }
}

// X = D^-1 Wi[block rows, :] (kInvBlock x M, column-major, leading dimension
// kInvBlock), with X[:, block] = D^-1: a 64-column tile per workgroup.
__global__ void __launch_bounds__(kBlock) k_inv_xrow(const double* __restrict__ Wi, int M, int k0, int nb,
                                                     const double* __restrict__ Dinv, double* __restrict__ X) {
    __shared__ double ds[kInvBlock][kInvBlock + 1];  // ds[j][r] = D^-1[r][j]
    __shared__ double ws[kInvBlock][kInvTile + 1];   // ws[j][c] = Wi[k0 + j][cb + c]
    const int cb = (int)blockIdx.x * kInvTile, t = (int)threadIdx.x;
    if (cb == k0) {  // the block's nb columns
        for (int e = t; e < nb * kInvBlock; e += kBlock) X[(long)k0 * kInvBlock + e] = Dinv[e];
        return;
    }
    // every load of the staging issued before the LDS stores
    constexpr int kSt = kInvBlock * kInvTile / kBlock;
    double dv[kSt], wv[kSt];
#pragma unroll
    for (int q = 0; q < kSt; ++q) {
        const int e = t + q * kBlock, c = e / kInvBlock, j = e - c * kInvBlock;
        dv[q] = Dinv[e];
        wv[q] = (j < nb && cb + c < M) ? Wi[(long)(cb + c) * M + k0 + j] : 0.0;
    }
#pragma unroll
    for (int q = 0; q < kSt; ++q) {
        const int e = t + q * kBlock, c = e / kInvBlock, j = e - c * kInvBlock;
        ds[c][j] = dv[q];  // column c of D^-1, row j
        ws[j][c] = wv[q];
    }
    __syncthreads();
    const int tr = (t % 16) * 4, tc = (t / 16) * 4;
    double acc[4][4] = {};
    for (int j = 0; j < kInvBlock; ++j) {
        double d[4], w[4];
#pragma unroll
        for (int a = 0; a < 4; ++a) d[a] = ds[j][tr + a];
#pragma unroll
        for (int b = 0; b < 4; ++b) w[b] = ws[j][tc + b];
#pragma unroll
        for (int a = 0; a < 4; ++a)
#pragma unroll
            for (int b = 0; b < 4; ++b) acc[a][b] += d[a] * w[b];
    }
#pragma unroll
    for (int b = 0; b < 4; ++b) {
        const int c = cb + tc + b;
        if (c >= M) continue;
#pragma unroll
        for (int a = 0; a < 4; ++a) X[(long)c * kInvBlock + tr + a] = acc[a][b];
    }
}

// Out of place, Wo from Wi: the block rows take X; every other row i takes
// (Wi[i, j], or 0 in the block columns) - Wi[i, block] X[:, j].  A 64 x 64
// tile of outputs per workgroup, 4 x 4 per thread, operands staged in LDS.
__global__ void __launch_bounds__(kBlock) k_inv_gemm_blk(const double* __restrict__ Wi, double* __restrict__ Wo,
                                                         int M, int k0, int nb, const double* __restrict__ X) {
    __shared__ double gs[kInvBlock][kInvTile + 1];  // gs[j][r] = Wi[r0 + r][k0 + j]
    __shared__ double xs[kInvBlock][kInvTile + 1];  // xs[j][c] = X[j][cb + c]
    const int r0 = (int)blockIdx.x * kInvTile, cb = (int)blockIdx.y * kInvTile;
    const int t = (int)threadIdx.x;
    if (r0 == k0) {  // the block rows (k0 is a multiple of kInvTile)
        for (int e = t; e < kInvTile * kInvTile; e += kBlock) {
            const int c = cb + e / kInvTile, rr = e % kInvTile;
            if (c < M && rr < nb) Wo[(long)c * M + k0 + rr] = X[(long)c * kInvBlock + rr];
        }
        return;
    }
    // every load of the staging issued before the LDS stores; X is read
    // along its columns (j fastest), the block columns of Wi along rows
    constexpr int kSt = kInvBlock * kInvTile / kBlock;
    double gv[kSt], xv[kSt];
#pragma unroll
    for (int q = 0; q < kSt; ++q) {
        const int e = t + q * kBlock, hi = e / kInvTile, lo = e % kInvTile;
        const int r = r0 + lo, c = cb + hi;
        gv[q] = (r < M && hi < nb) ? Wi[(long)(k0 + hi) * M + r] : 0.0;
        xv[q] = (c < M && lo < nb) ? X[(long)c * kInvBlock + lo] : 0.0;
    }
#pragma unroll
    for (int q = 0; q < kSt; ++q) {
        const int e = t + q * kBlock, hi = e / kInvTile, lo = e % kInvTile;
        gs[hi][lo] = gv[q];
        xs[lo][hi] = xv[q];
    }
    __syncthreads();
    const int tr = (t % 16) * 4, tc = (t / 16) * 4;
    double acc[4][4] = {};
    for (int j = 0; j < kInvBlock; ++j) {
        double g[4], x[4];
#pragma unroll
        for (int a = 0; a < 4; ++a) g[a] = gs[j][tr + a];
#pragma unroll
        for (int b = 0; b < 4; ++b) x[b] = xs[j][tc + b];
#pragma unroll
        for (int a = 0; a < 4; ++a)
#pragma unroll
            for (int b = 0; b < 4; ++b) acc[a][b] += g[a] * x[b];
    }
    const bool blkcol = cb == k0;
#pragma unroll
    for (int b = 0; b < 4; ++b) {
        const int c = cb + tc + b;
        if (c >= M) continue;
        const double* yi = Wi + (long)c * M;
        double* yo = Wo + (long)c * M;
#pragma unroll
        for (int a = 0; a < 4; ++a) {
            const int r = r0 + tr + a;
            if (r >= M) continue;
            yo[r] = (blkcol ? 0.0 : yi[r]) - acc[a][b];
        }
    }
}

// The check vector x_r (|x_r| <= 1, deterministic) and the residual of the
// inverse: st->resid = max |z - x| / max |x| with z = A (Ainv x), one
// workgroup over the M entries (y and z come from the dense products).
__device__ __forceinline__ double inv_check_x(int r) {
    const unsigned h = (unsigned)r * 2654435761u + 12345u;
    return (double)(int)(h >> 8) / 8388608.0 - 1.0;
}
__global__ void __launch_bounds__(kBlock) k_inv_check_x(double* __restrict__ x, int M) {
    const int r = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (r < M) x[r] = inv_check_x(r);
}
__global__ void __launch_bounds__(kBlock) k_inv_check(const double* __restrict__ z, int M, InvState* __restrict__ st) {
    __shared__ double red[2][kBlock / 64];
    double e = 0.0, n = 0.0;
    for (int r = (int)threadIdx.x; r < M; r += kBlock) {
        const double x = inv_check_x(r), d = fabs(z[r] - x);
        e = (d > e || d != d) ? d : e;  // NaN propagates
        n = fmax(n, fabs(x));
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        const double oe = __shfl_xor(e, off), on = __shfl_xor(n, off);
        e = (oe > e || oe != oe) ? oe : e;
        n = fmax(n, on);
    }
    if ((threadIdx.x & 63) == 0) {
        red[0][threadIdx.x >> 6] = e;
        red[1][threadIdx.x >> 6] = n;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        e = red[0][0];
        n = red[1][0];
        for (int w = 1; w < kBlock / 64; ++w) {
            e = (red[0][w] > e || red[0][w] != red[0][w]) ? red[0][w] : e;
            n = fmax(n, red[1][w]);
        }
        st->resid = n > 0.0 ? e / n : e;
    }
}

// Undo the column permutation: inv[:, c] = W[:, src[c]].
__global__ void __launch_bounds__(kBlock) k_inv_unpermute(const double* __restrict__ W, int M,
                                                          const int* __restrict__ src, double* __restrict__ inv) {
    const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (long)M * M) return;
    const int c = (int)(e / M), r = (int)(e - (long)c * M);
    inv[e] = W[(long)src[c] * M + r];
}

// Identity (column-major M x M): the unit vectors of the dense assemblies.
__global__ void __launch_bounds__(kBlock) k_eye(double* __restrict__ I, int M) {
    const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (long)M * M) return;
    const int c = (int)(e / M), r = (int)(e - (long)c * M);
    I[e] = r == c ? 1.0 : 0.0;
}

}  // namespace ael
