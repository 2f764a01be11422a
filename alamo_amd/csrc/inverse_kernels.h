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
// run in one workgroup with a panel of kInvPanel columns in registers
// (k_inv_panel); each panel then updates the other columns of its block
// (k_inv_block, rank kInvPanel), and each block of kInvBlock columns all the
// columns outside it (k_inv_rowmap, k_inv_gemm: rank kInvBlock, an LDS-tiled
// product, out of place).  The pivot steps are the sequential part: two
// barriers each.
#pragma once
#include "common.h"

namespace ael {

constexpr int kInvPanel = 8;      // columns factored together in registers
constexpr int kInvBlock = 64;     // columns per trailing update
constexpr int kInvThreads = 1024; // threads of the panel kernel
constexpr int kInvRows = 4;       // rows per thread of the panel kernel: M <= 4096
constexpr int kInvMax = kInvThreads * kInvRows;

struct InvState {
    int singular;  // set when a pivot column holds no nonzero below the diagonal
};

// The nb <= kInvPanel pivot steps of columns c0 .. c0 + nb - 1 on those
// columns of W only.  Thread t holds rows t + q kInvThreads in registers;
// perm[k] receives the row swapped with row k.
__global__ void __launch_bounds__(kInvThreads) k_inv_panel(double* __restrict__ W, int M, int c0, int nb,
                                                           int* __restrict__ perm, InvState* __restrict__ st) {
    __shared__ double wv[2][kInvThreads / 64];
    __shared__ int wi[2][kInvThreads / 64];
    __shared__ double prow[2][kInvPanel], krow[2][kInvPanel];
    const int t = (int)threadIdx.x, lane = t & 63, wave = t >> 6;
    double v[kInvRows][kInvPanel];
#pragma unroll
    for (int q = 0; q < kInvRows; ++q) {
        const int r = t + q * kInvThreads;
#pragma unroll
        for (int c = 0; c < kInvPanel; ++c) v[q][c] = (r < M && c < nb) ? W[(long)(c0 + c) * M + r] : 0.0;
    }
#pragma unroll
    for (int j = 0; j < kInvPanel; ++j) {
        if (j >= nb) break;  // uniform
        const int k = c0 + j, par = j & 1;
        // pivot: the largest |W[r][k]| over rows r >= k, the first on ties
        double best = -1.0;
        int idx = M;
#pragma unroll
        for (int q = 0; q < kInvRows; ++q) {
            const int r = t + q * kInvThreads;
            const double a = fabs(v[q][j]);
            if (r >= k && r < M && (a > best || (a == best && r < idx))) {
                best = a;
                idx = r;
            }
        }
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
            const double ob = __shfl_xor(best, off);
            const int oi = __shfl_xor(idx, off);
            if (ob > best || (ob == best && oi < idx)) {
                best = ob;
                idx = oi;
            }
        }
        if (lane == 0) {
            wv[par][wave] = best;
            wi[par][wave] = idx;
        }
        __syncthreads();
        best = wv[par][0];
        idx = wi[par][0];
#pragma unroll
        for (int w = 1; w < kInvThreads / 64; ++w) {
            const double ob = wv[par][w];
            const int oi = wi[par][w];
            if (ob > best || (ob == best && oi < idx)) {
                best = ob;
                idx = oi;
            }
        }
        const int p = idx;
        // rows k and p trade places; every thread needs the new row k (row p)
        const int ok = k % kInvThreads, qk = k / kInvThreads, op = p % kInvThreads, qp = p / kInvThreads;
        if (t == op) {
#pragma unroll
            for (int q = 0; q < kInvRows; ++q)
                if (q == qp) {
#pragma unroll
                    for (int c = 0; c < kInvPanel; ++c) prow[par][c] = v[q][c];
                }
        }
        if (t == ok) {
#pragma unroll
            for (int q = 0; q < kInvRows; ++q)
                if (q == qk) {
#pragma unroll
                    for (int c = 0; c < kInvPanel; ++c) krow[par][c] = v[q][c];
                }
        }
        __syncthreads();
        if (t == 0) {
            perm[k] = p;
            if (!(best > 0.0)) st->singular = 1;
        }
        // the new row k: the pivot row times 1/d (one division a step)
        const double rd = 1.0 / prow[par][j];
        double pr[kInvPanel];
#pragma unroll
        for (int c = 0; c < kInvPanel; ++c) pr[c] = c == j ? rd : prow[par][c] * rd;
#pragma unroll
        for (int q = 0; q < kInvRows; ++q) {
            const int r = t + q * kInvThreads;
            if (r == p && p != k) {
#pragma unroll
                for (int c = 0; c < kInvPanel; ++c) v[q][c] = krow[par][c];  // old row k lands on row p
            }
            const double f = v[q][j];
#pragma unroll
            for (int c = 0; c < kInvPanel; ++c)
                v[q][c] = r == k ? pr[c] : (c == j ? -f * rd : v[q][c] - f * pr[c]);
        }
    }
#pragma unroll
    for (int q = 0; q < kInvRows; ++q) {
        const int r = t + q * kInvThreads;
        if (r < M) {
#pragma unroll
            for (int c = 0; c < kInvPanel; ++c)
                if (c < nb) W[(long)(c0 + c) * M + r] = v[q][c];
        }
    }
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

// The panel's swaps and rank-kInvPanel update on the other columns of its
// block [b0, b1): one workgroup per column.  Only rows c0 .. c0 + kInvPanel - 1
// and the rows swapped with them change under the swaps; thread j moves rows
// c0 + j and perm[c0 + j] (a row reached twice takes the same value twice).
__global__ void __launch_bounds__(kBlock) k_inv_block(double* __restrict__ W, int M, int b0, int b1, int c0, int nb,
                                                      const int* __restrict__ perm) {
    __shared__ int map[kInvMax], ps[kInvPanel];
    __shared__ double ys[kInvPanel];
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
    if (j < kInvPanel) ys[j] = j < nb ? a : 0.0;
    __syncthreads();
    for (int r = (int)threadIdx.x; r < M; r += blockDim.x) {
        double acc = (r >= c0 && r < c0 + nb) ? 0.0 : y[r];
#pragma unroll
        for (int q = 0; q < kInvPanel; ++q)
            if (q < nb) acc += G[(long)q * M + r] * ys[q];
        y[r] = acc;
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
constexpr int kInvTile = 64;
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
