// psi_kernels.h -- the psi (diffuse-domain) weight of Operator::Elastic
// (SetPsi, src/Operator/Elastic.cpp:118-137).
//
// psi is a cell field (nx-1) x (ny-1) x (nz-1) per MG level, x fastest; cells
// outside the domain read 1 (the operator fills its psi with 1 at define and
// SetPsi copies only the domain's cells, Elastic.cpp:49-52,124-134).  At node
// (i, j, k):
//   psi_avg = (1 - small) * mean of the 8 cells around the node + small
//   (Numeric::Interpolate::CellToNodeAverage, src/Numeric/Stencil.H:1390-1406);
// Fapply (Elastic.cpp:190-259):
//   interior   f = psi_avg (C : grad grad u + grad C . grad u)
//                  + (1 - small) (C : grad u) . grad psi
//              (grad psi: Numeric::CellGradientOnNode, src/Numeric/Stencil.H:699-715);
//   boundary   the BC row with sigma = psi_avg C : grad u, so traction rows
//              scale by psi_avg, displacement and Neumann rows do not
//              (Constant::set, src/BC/Operator/Elastic/Constant.H:408-425);
// Diagonal (Elastic.cpp:340-363): the same scaling without the grad psi term.
//
// The operator kernels compute the unweighted rows; k_psi_post applies the
// weight to them, so every level, precision and kernel variant share one
// formula.
#pragma once
#include "elastic_kernels.h"

namespace ael {

__device__ __forceinline__ double psi_cell(const double* __restrict__ psi, const Lvl& L, int ci, int cj, int ck) {
    const int cx = L.nx - 1, cy = L.ny - 1, cz = L.nz - 1;
    if (ci < 0 || cj < 0 || ck < 0 || ci >= cx || cj >= cy || ck >= cz) return 1.0;
    return psi[((long)ck * cy + cj) * cx + ci];
}

// The 8 cells around node (i, j, k): p[c] with c = dx + 2 dy + 4 dz, cell
// (i - 1 + dx, j - 1 + dy, k - 1 + dz).
__device__ __forceinline__ void psi_cells8(const double* __restrict__ psi, const Lvl& L, int i, int j, int k,
                                           double p[8]) {
#pragma unroll
    for (int c = 0; c < 8; ++c) p[c] = psi_cell(psi, L, i - 1 + (c & 1), j - 1 + ((c >> 1) & 1), k - 1 + (c >> 2));
}

__device__ __forceinline__ double psi_avg_of(const double p[8], double small) {
    return (1.0 - small) * ((p[1] + p[0] + p[3] + p[2] + p[5] + p[4] + p[7] + p[6]) * 0.125) + small;
}

// Weighted rows of the operator at every node from the unweighted rows in
// out (f): mode 0 out = A_psi x, mode 1 out = b - A_psi x, mode 2 out = x +
// w (b - A_psi x) / diag (one damped Jacobi sweep with the weighted diagonal).
// blockIdx.y selects a column when the rows of several inputs are weighted at
// once (dense assembly): x and out advance by 3 N per column.
template <typename T, int SYM>
__global__ void __launch_bounds__(kBlock) k_psi_post(const T* __restrict__ x, const T* __restrict__ b,
                                                     const T* __restrict__ diag, const T* __restrict__ C,
                                                     const int* __restrict__ bct, const double* __restrict__ psi,
                                                     double small, Lvl L, T w, int mode, T* __restrict__ out) {
    constexpr int NC = NComp<SYM>::value;
    const long id = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= L.N) return;
    x += 3 * L.N * (long)blockIdx.y;
    out += 3 * L.N * (long)blockIdx.y;
    int i, j, k;
    node_ijk(id, L, i, j, k);
    double pc[8];
    psi_cells8(psi, L, i, j, k, pc);
    const double pa = psi_avg_of(pc, small);
    T f[3];
#pragma unroll
    for (int p = 0; p < 3; ++p) f[p] = out[p * L.N + id];
    T nrm[3];
    const int r = region_of<T>(i, j, k, L, nrm);
    if (r >= 0) {
#pragma unroll
        for (int p = 0; p < 3; ++p)
            if (bct[3 * r + p] == 1) f[p] = (T)(pa * (double)f[p]);
    } else {
        const long sy = L.nx, sz = (long)L.nx * L.ny;
        T gu[3][3], s[3][3], c[NC];
#pragma unroll
        for (int p = 0; p < 3; ++p) {
            const T* u = x + p * L.N;
            gu[p][0] = d1(u, id, 1, i, L.nx, (T)L.dx);
            gu[p][1] = d1(u, id, sy, j, L.ny, (T)L.dy);
            gu[p][2] = d1(u, id, sz, k, L.nz, (T)L.dz);
        }
#pragma unroll
        for (int q = 0; q < NC; ++q) c[q] = C[q * L.N + id];
        contract2<T, SYM>(c, gu, s);
        const double gp[3] = {
            0.25 * (pc[1] - pc[0] + pc[3] - pc[2] + pc[5] - pc[4] + pc[7] - pc[6]) / L.dx,
            0.25 * (pc[3] - pc[1] + pc[2] - pc[0] + pc[7] - pc[5] + pc[6] - pc[4]) / L.dy,
            0.25 * (pc[7] - pc[3] + pc[5] - pc[1] + pc[6] - pc[2] + pc[4] - pc[0]) / L.dz};
#pragma unroll
        for (int p = 0; p < 3; ++p) {
            double v = pa * (double)f[p];
#pragma unroll
            for (int a = 0; a < 3; ++a) v += (double)s[p][a] * ((1.0 - small) * gp[a]);
            f[p] = (T)v;
        }
    }
#pragma unroll
    for (int p = 0; p < 3; ++p) {
        const long q = p * L.N + id;
        out[q] = mode == 0 ? f[p] : mode == 1 ? b[q] - f[p] : x[q] + w * (b[q] - f[p]) / diag[q];
    }
}

// Weighted diagonal from the unweighted one (Elastic.cpp:340-363).
template <typename T>
__global__ void __launch_bounds__(kBlock) k_psi_diag(T* __restrict__ diag, const int* __restrict__ bct,
                                                     const double* __restrict__ psi, double small, Lvl L) {
    const long id = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= L.N) return;
    int i, j, k;
    node_ijk(id, L, i, j, k);
    double pc[8];
    psi_cells8(psi, L, i, j, k, pc);
    const double pa = psi_avg_of(pc, small);
    T nrm[3];
    const int r = region_of<T>(i, j, k, L, nrm);
#pragma unroll
    for (int p = 0; p < 3; ++p)
        if (r < 0 || bct[3 * r + p] == 1) diag[p * L.N + id] = (T)(pa * (double)diag[p * L.N + id]);
}

// psi on the next MG level (averageDownCoeffsSameAmrLevel,
// src/Operator/Elastic.cpp:839-907): coarse cell (I, J, K) from the fine
// cells around (2I, 2J, 2K).  Cells on the low domain faces (I, J or K = 0)
// take the corner value, the 1-2-1 edge average or the 1-2-1 x 1-2-1 face
// average in the face; every other cell, those on the high faces included
// (the reference compares the cell index with the node domain's upper end,
// which a cell never reaches), the 27-cell full weighting 1/64, 1/32, 1/16,
// 1/8 by distance.
__global__ void __launch_bounds__(kBlock) k_psi_restrict(const double* __restrict__ pf, Lvl Lf,
                                                         double* __restrict__ pc, Lvl Lc) {
    const int cx = Lc.nx - 1, cy = Lc.ny - 1, cz = Lc.nz - 1;
    const int fx = Lf.nx - 1, fy = Lf.ny - 1;
    const long id = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= (long)cx * cy * cz) return;
    const long t = id / cx;
    const int I = (int)(id - t * cx), K = (int)(t / cy), J = (int)(t - (long)K * cy);
    const int i = 2 * I, j = 2 * J, k = 2 * K;
    auto F = [&](int a, int b, int c) { return pf[((long)c * fy + b) * fx + a]; };
    double v;
    if (I == 0 && J == 0 && K == 0) v = F(i, j, k);
    else if (J == 0 && K == 0) v = F(i - 1, j, k) * 0.25 + F(i, j, k) * 0.5 + F(i + 1, j, k) * 0.25;
    else if (K == 0 && I == 0) v = F(i, j - 1, k) * 0.25 + F(i, j, k) * 0.5 + F(i, j + 1, k) * 0.25;
    else if (I == 0 && J == 0) v = F(i, j, k - 1) * 0.25 + F(i, j, k) * 0.5 + F(i, j, k + 1) * 0.25;
    else if (I == 0)
        v = (F(i, j - 1, k - 1) + F(i, j, k - 1) * 2.0 + F(i, j + 1, k - 1) + F(i, j - 1, k) * 2.0 + F(i, j, k) * 4.0 +
             F(i, j + 1, k) * 2.0 + F(i, j - 1, k + 1) + F(i, j, k + 1) * 2.0 + F(i, j + 1, k + 1)) / 16.0;
    else if (J == 0)
        v = (F(i - 1, j, k - 1) + F(i - 1, j, k) * 2.0 + F(i - 1, j, k + 1) + F(i, j, k - 1) * 2.0 + F(i, j, k) * 4.0 +
             F(i, j, k + 1) * 2.0 + F(i + 1, j, k - 1) + F(i + 1, j, k) * 2.0 + F(i + 1, j, k + 1)) / 16.0;
    else if (K == 0)
        v = (F(i - 1, j - 1, k) + F(i, j - 1, k) * 2.0 + F(i + 1, j - 1, k) + F(i - 1, j, k) * 2.0 + F(i, j, k) * 4.0 +
             F(i + 1, j, k) * 2.0 + F(i - 1, j + 1, k) + F(i, j + 1, k) * 2.0 + F(i + 1, j + 1, k)) / 16.0;
    else
        v = (F(i - 1, j - 1, k - 1) + F(i - 1, j - 1, k + 1) + F(i - 1, j + 1, k - 1) + F(i - 1, j + 1, k + 1) +
             F(i + 1, j - 1, k - 1) + F(i + 1, j - 1, k + 1) + F(i + 1, j + 1, k - 1) + F(i + 1, j + 1, k + 1)) / 64.0 +
            (F(i, j - 1, k - 1) + F(i, j - 1, k + 1) + F(i, j + 1, k - 1) + F(i, j + 1, k + 1) + F(i - 1, j, k - 1) +
             F(i + 1, j, k - 1) + F(i - 1, j, k + 1) + F(i + 1, j, k + 1) + F(i - 1, j - 1, k) + F(i - 1, j + 1, k) +
             F(i + 1, j - 1, k) + F(i + 1, j + 1, k)) / 32.0 +
            (F(i - 1, j, k) + F(i, j - 1, k) + F(i, j, k - 1) + F(i + 1, j, k) + F(i, j + 1, k) + F(i, j, k + 1)) / 16.0 +
            F(i, j, k) / 8.0;
    pc[id] = v;
}

}  // namespace ael
