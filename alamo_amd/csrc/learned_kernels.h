// learned_kernels.h -- device code that connects the elastic multigrid
// hierarchy to the learned coarse-level solver (alamo_unet.h): the coarse
// right-hand side, the network input features and its output.
//
// On MG level L the network sees, per node, the stiffness slots divided by
// alpha (the mean of slot 0 over the level), the interior right-hand side
// divided by alpha * s, the boundary values divided by s, per component
// whether the boundary row is a displacement row, the node coordinates
// scaled to [0, 1] (the solutions of boundary loads are close to affine
// fields, which translation-invariant convolutions cannot place without
// them) and, as constant channels, the mean scaled boundary value of each
// face and component (the load, visible at every node instead of only on
// the boundary shell); it predicts u_L / s.  The
// scale s is the largest magnitude of the scaled right-hand side, so the
// network sees inputs of order one whatever the load, and the output is
// multiplied back by s.  Interior rows of A scale with C, displacement rows
// do not: dividing only the interior right-hand side by alpha makes the
// scaled problem the one of the normalized stiffness C / alpha.
#pragma once
#include "elastic_kernels.h"

namespace ael {

__device__ __forceinline__ bool on_boundary(const Lvl& L, int i, int j, int k) {
    return i == 0 || j == 0 || k == 0 || i == L.nx - 1 || j == L.ny - 1 || k == L.nz - 1;
}

// out = b on interior nodes, 0 on boundary nodes (3 components).
__global__ void __launch_bounds__(kBlock) k_interior_only(const double* __restrict__ b, Lvl L,
                                                          double* __restrict__ out) {
    const long id = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= L.N) return;
    int i, j, k;
    node_ijk(id, L, i, j, k);
    const bool bd = on_boundary(L, i, j, k);
#pragma unroll
    for (int p = 0; p < 3; ++p) out[p * L.N + id] = bd ? 0.0 : b[p * L.N + id];
}

// Boundary nodes of the coarse level take the value of the fine node they
// sit on (fine index = factor * coarse index); interior nodes are kept.
__global__ void __launch_bounds__(kBlock) k_inject_boundary(const double* __restrict__ bf, Lvl Lf, int factor,
                                                            double* __restrict__ bc, Lvl Lc) {
    const long id = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= Lc.N) return;
    int I, J, K;
    node_ijk(id, Lc, I, J, K);
    if (!on_boundary(Lc, I, J, K)) return;
    const long f = ((long)factor * K * Lf.ny + (long)factor * J) * Lf.nx + (long)factor * I;
#pragma unroll
    for (int p = 0; p < 3; ++p) bc[p * Lc.N + id] = bf[p * Lf.N + f];
}

// t = b / alpha on interior nodes, b on boundary nodes.
__global__ void __launch_bounds__(kBlock) k_prescale(const double* __restrict__ b, Lvl L, double inv_alpha,
                                                     double* __restrict__ t) {
    const long id = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= L.N) return;
    int i, j, k;
    node_ijk(id, L, i, j, k);
    const double f = on_boundary(L, i, j, k) ? 1.0 : inv_alpha;
#pragma unroll
    for (int p = 0; p < 3; ++p) t[p * L.N + id] = f * b[p * L.N + id];
}

// Mean of t over face f (0..5: xlo, ylo, zlo, xhi, yhi, zhi) for component c,
// one workgroup per (f, c): fmean[3 f + c].
__global__ void __launch_bounds__(kBlock) k_face_means(const double* __restrict__ t, Lvl L,
                                                       double* __restrict__ fmean) {
    __shared__ double sh[kBlock];
    const int f = blockIdx.x / 3, c = blockIdx.x % 3;
    const int ax = f % 3;
    const bool hi = f >= 3;
    const int n1 = ax == 0 ? L.ny : L.nx, n2 = ax == 2 ? L.ny : L.nz;
    const long cnt = (long)n1 * n2;
    double acc = 0.0;
    for (long q = threadIdx.x; q < cnt; q += blockDim.x) {
        const int a = (int)(q % n1), b = (int)(q / n1);
        int i, j, k;
        if (ax == 0) { i = hi ? L.nx - 1 : 0; j = a; k = b; }
        else if (ax == 1) { j = hi ? L.ny - 1 : 0; i = a; k = b; }
        else { k = hi ? L.nz - 1 : 0; i = a; j = b; }
        acc += t[c * L.N + ((long)k * L.ny + j) * L.nx + i];
    }
    sh[threadIdx.x] = acc;
    __syncthreads();
    for (int w = kBlock / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) fmean[blockIdx.x] = sh[0] / (double)cnt;
}

// Network input [nc + 30][N] (float32): C / alpha, interior t / s, boundary
// t / s, displacement-row indicators, coordinates i/(nx-1), j/(ny-1),
// k/(nz-1), and the 18 face means / s.
__global__ void __launch_bounds__(kBlock) k_features(const double* __restrict__ C, int nc, const double* __restrict__ t,
                                                     const int* __restrict__ bct, const double* __restrict__ fmean,
                                                     Lvl L, double inv_alpha, double inv_s, float* __restrict__ feat) {
    const long id = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= L.N) return;
    int i, j, k;
    node_ijk(id, L, i, j, k);
    const long N = L.N;
    for (int q = 0; q < nc; ++q) feat[q * N + id] = (float)(C[q * N + id] * inv_alpha);
    double nrm[3];
    const int reg = region_of<double>(i, j, k, L, nrm);
#pragma unroll
    for (int p = 0; p < 3; ++p) {
        const float v = (float)(t[p * N + id] * inv_s);
        feat[(nc + p) * N + id] = reg < 0 ? v : 0.f;
        feat[(nc + 3 + p) * N + id] = reg < 0 ? 0.f : v;
        feat[(nc + 6 + p) * N + id] = (reg >= 0 && bct[3 * reg + p] == 0) ? 1.f : 0.f;
    }
    feat[(nc + 9) * N + id] = (float)i / (float)(L.nx - 1);
    feat[(nc + 10) * N + id] = (float)j / (float)(L.ny - 1);
    feat[(nc + 11) * N + id] = (float)k / (float)(L.nz - 1);
    for (int q = 0; q < 18; ++q) feat[(nc + 12 + q) * N + id] = (float)(fmean[q] * inv_s);
}

// u = s * y (float32 network output -> FP64 level field).
__global__ void __launch_bounds__(kBlock) k_scale_output(const float* __restrict__ y, long n, double s,
                                                         double* __restrict__ u) {
    const long id = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (id < n) u[id] = s * (double)y[id];
}

}  // namespace ael
