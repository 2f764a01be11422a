// amr_kernels.h -- device code of the composite (AMR) solve: nested single-box
// patches, refinement ratio 2, nodal.
//
// A patch of AMR level l covers the parent-level cells [clo, chi] (global
// indices).  Its handle holds the fine nodes 2 clo .. 2 chi plus one ghost
// node layer on each face that lies inside the domain; `origin` is the
// global level-l node index of handle node 0.  Ghost nodes carry values
// interpolated from the parent level (the "ghostnodes" coarse-fine
// treatment), the nodes 2 clo .. 2 chi are the patch's unknowns.
#pragma once
#include "elastic_kernels.h"

namespace ael {

struct Patch {
    int ox, oy, oz;        // origin: global fine node index of handle node 0
    int clo[3], chi[3];    // covered parent cells [clo, chi] (global parent indices)
    int gl[3], gh[3];      // 1 where the lo / hi face of the handle is a ghost layer
};

__device__ __forceinline__ bool ghost_node(const Lvl& L, const Patch& P, int i, int j, int k) {
    return (P.gl[0] && i == 0) || (P.gh[0] && i == L.nx - 1) || (P.gl[1] && j == 0) || (P.gh[1] && j == L.ny - 1) ||
           (P.gl[2] && k == 0) || (P.gh[2] && k == L.nz - 1);
}

// Integrator::Mechanics::TagCellsForRefinement (src/Integrator/Mechanics.H:298-313):
// node (i, j, k) tags cell (i, j, k) when |grad eta_n| |dx| > thr for some
// component n, grad central inside the box and one-sided on its faces
// (Numeric::Gradient with GetStencil on the box, as d1), |dx| the 2-norm of
// the cell size.  Only the box's cells are kept (the top nodes' tags fall
// outside it).  The tagged cells' bounding box and count accumulate with
// atomics into bb[0..2] (min i, j, k), bb[3..5] (max) and bb[6] (count),
// initialised by k_tag_init.
__global__ void __launch_bounds__(kBlock) k_tag_init(int* __restrict__ bb) {
    if (threadIdx.x < 7) bb[threadIdx.x] = threadIdx.x < 3 ? 0x7fffffff : (threadIdx.x < 6 ? -1 : 0);
}
// Whether node (i, j, k) tags cell (i, j, k).
__device__ __forceinline__ bool eta_tagged(const double* __restrict__ eta, int ncomp, const Lvl& L, double thr,
                                           long id, int i, int j, int k) {
#pragma clang fp contract(off)
    const double dxn = sqrt(L.dx * L.dx + L.dy * L.dy + L.dz * L.dz);
    const long sy = L.nx, sz = (long)L.nx * L.ny;
    bool tag = false;
    for (int n = 0; n < ncomp; ++n) {
        const double* f = eta + (long)n * L.N;
        const double g0 = d1(f, id, 1, i, L.nx, L.dx), g1 = d1(f, id, sy, j, L.ny, L.dy),
                     g2 = d1(f, id, sz, k, L.nz, L.dz);
        tag |= sqrt(g0 * g0 + g1 * g1 + g2 * g2) * dxn > thr;
    }
    return tag;
}
__global__ void __launch_bounds__(kBlock) k_tag_eta(const double* __restrict__ eta, int ncomp, Lvl L, double thr,
                                                    int* __restrict__ bb) {
    const long id = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= L.N) return;
    int i, j, k;
    node_ijk(id, L, i, j, k);
    if (i >= L.nx - 1 || j >= L.ny - 1 || k >= L.nz - 1) return;
    if (!eta_tagged(eta, ncomp, L, thr, id, i, j, k)) return;
    atomicMin(bb + 0, i);
    atomicMin(bb + 1, j);
    atomicMin(bb + 2, k);
    atomicMax(bb + 3, i);
    atomicMax(bb + 4, j);
    atomicMax(bb + 5, k);
    atomicAdd(bb + 6, 1);
}

// The same tags, one byte per cell of the box ((nx-1)(ny-1)(nz-1), x fastest).
__global__ void __launch_bounds__(kBlock) k_tag_cells(const double* __restrict__ eta, int ncomp, Lvl L, double thr,
                                                      unsigned char* __restrict__ tags) {
    const long id = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= L.N) return;
    int i, j, k;
    node_ijk(id, L, i, j, k);
    if (i >= L.nx - 1 || j >= L.ny - 1 || k >= L.nz - 1) return;
    tags[((long)k * (L.ny - 1) + j) * (L.nx - 1) + i] = eta_tagged(eta, ncomp, L, thr, id, i, j, k) ? 1 : 0;
}

// v = 0 on ghost nodes (mode 0) or v = src on ghost nodes (mode 1).
__global__ void __launch_bounds__(kBlock) k_patch_ghost(double* __restrict__ v, const double* __restrict__ src, Lvl L,
                                                        Patch P, int mode) {
    const long id = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= L.N) return;
    int i, j, k;
    node_ijk(id, L, i, j, k);
    if (!ghost_node(L, P, i, j, k)) return;
#pragma unroll
    for (int p = 0; p < 3; ++p) v[p * L.N + id] = mode ? src[p * L.N + id] : 0.0;
}

// Trilinear value of component p of the parent field c (level Lc, whole
// parent handle, parent origin po) at the global fine node (gx, gy, gz)
// (src/Operator/Operator.cpp:365-434 weights; node_bilinear_interp).
__device__ __forceinline__ double parent_trilinear(const double* __restrict__ c, const Lvl& Lc, int pox, int poy,
                                                   int poz, int gx, int gy, int gz, int p) {
    const int xs[2] = {(gx >> 1) - pox, ((gx + 1) >> 1) - pox};
    const int ys[2] = {(gy >> 1) - poy, ((gy + 1) >> 1) - poy};
    const int zs[2] = {(gz >> 1) - poz, ((gz + 1) >> 1) - poz};
    const int nxs = (gx & 1) ? 2 : 1, nys = (gy & 1) ? 2 : 1, nzs = (gz & 1) ? 2 : 1;
    const double w = 1.0 / (double)(nxs * nys * nzs);
    double acc = 0.0;
    for (int a = 0; a < nzs; ++a)
        for (int b = 0; b < nys; ++b)
            for (int d = 0; d < nxs; ++d) acc += c[p * Lc.N + ((long)zs[a] * Lc.ny + ys[b]) * Lc.nx + xs[d]];
    return acc * w;
}

// Trilinear interpolation of the parent field c onto the patch handle field f:
// mode 0: f = P c everywhere, mode 1: f += P c everywhere, mode 2: f = P c on
// ghost nodes only.
__global__ void __launch_bounds__(kBlock) k_patch_interp(const double* __restrict__ c, Lvl Lc, int pox, int poy,
                                                         int poz, double* __restrict__ f, Lvl Lf, Patch P, int mode) {
    const long id = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= Lf.N) return;
    int i, j, k;
    node_ijk(id, Lf, i, j, k);
    if (mode == 2 && !ghost_node(Lf, P, i, j, k)) return;
#pragma unroll
    for (int p = 0; p < 3; ++p) {
        const double acc = parent_trilinear(c, Lc, pox, poy, poz, P.ox + i, P.oy + j, P.oz + k, p);
        double* q = f + p * Lf.N + id;
        *q = mode == 1 ? *q + acc : acc;
    }
}

// FillPatchTwoLevels with node_bilinear_interp (src/Integrator/Integrator.cpp:
// 394-446) onto a new patch P of handle shape Lf: a node takes the old patch's
// value where the old patch's unknowns (global nodes 2 Ps.clo .. 2 Ps.chi of
// the same AMR level) hold it, else the trilinear interpolation of the parent
// field c.  Without an old patch (so == nullptr) every node is interpolated:
// InterpFromCoarseLevel (FillCoarsePatch, :452-478).
__global__ void __launch_bounds__(kBlock) k_patch_fill(const double* __restrict__ c, Lvl Lc, int pox, int poy,
                                                       int poz, const double* __restrict__ so, Lvl Ls, Patch Ps,
                                                       double* __restrict__ f, Lvl Lf, Patch P) {
    const long id = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= Lf.N) return;
    int i, j, k;
    node_ijk(id, Lf, i, j, k);
    const int gx = P.ox + i, gy = P.oy + j, gz = P.oz + k;
    const bool old = so && 2 * Ps.clo[0] <= gx && gx <= 2 * Ps.chi[0] && 2 * Ps.clo[1] <= gy &&
                     gy <= 2 * Ps.chi[1] && 2 * Ps.clo[2] <= gz && gz <= 2 * Ps.chi[2];
    const long is = old ? ((long)(gz - Ps.oz) * Ls.ny + (gy - Ps.oy)) * Ls.nx + (gx - Ps.ox) : 0;
#pragma unroll
    for (int p = 0; p < 3; ++p)
        f[p * Lf.N + id] = old ? so[p * Ls.N + is] : parent_trilinear(c, Lc, pox, poy, poz, gx, gy, gz, p);
}

// averageDownSolutionRHS (src/Operator/Operator.cpp:436-448, nodal
// average_down = injection): parent nodes clo..chi take the patch values.
__global__ void __launch_bounds__(kBlock) k_patch_inject(const double* __restrict__ f, Lvl Lf, Patch P,
                                                         double* __restrict__ c, Lvl Lc, int pox, int poy, int poz) {
    const int nx = P.chi[0] - P.clo[0] + 1, ny = P.chi[1] - P.clo[1] + 1, nz = P.chi[2] - P.clo[2] + 1;
    const long id = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= (long)nx * ny * nz) return;
    const int a = (int)(id % nx), b = (int)((id / nx) % ny), d = (int)(id / ((long)nx * ny));
    const int I = P.clo[0] + a, J = P.clo[1] + b, K = P.clo[2] + d;  // global parent node
    const long fi = ((long)(2 * K - P.oz) * Lf.ny + (2 * J - P.oy)) * Lf.nx + (2 * I - P.ox);
    const long ci = ((long)(K - poz) * Lc.ny + (J - poy)) * Lc.nx + (I - pox);
#pragma unroll
    for (int p = 0; p < 3; ++p) c[p * Lc.N + ci] = f[p * Lf.N + fi];
}

// Operator::reflux (src/Operator/Operator.cpp:526-633): the parent residual
// at the nodes clo..chi (covered and coarse-fine nodes) is replaced by the
// full-weighting restriction of the patch residual.  At a domain face the
// stencil falls back to the face, edge and corner weights of the MG
// restriction (Operator.cpp:280-353), since no fine node exists outside the
// domain; inside the domain the fine ghost layer supplies the outer nodes.
// The same restriction of the ncomp planes of a patch's C onto its parent's
// is Elastic::averageDownCoeffsDifferentAmrLevels (Elastic.cpp:628-745,
// whose corner, edge, face and interior weights are these).
__global__ void __launch_bounds__(kBlock) k_patch_reflux(const double* __restrict__ rf, Lvl Lf, Patch P,
                                                         double* __restrict__ rc, Lvl Lc, int pox, int poy, int poz,
                                                         int dnx, int dny, int dnz, int ncomp) {
    const int nx = P.chi[0] - P.clo[0] + 1, ny = P.chi[1] - P.clo[1] + 1, nz = P.chi[2] - P.clo[2] + 1;
    const long id = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= (long)nx * ny * nz) return;
    const int a = (int)(id % nx), b = (int)((id / nx) % ny), d = (int)(id / ((long)nx * ny));
    const int I = P.clo[0] + a, J = P.clo[1] + b, K = P.clo[2] + d;  // global parent node
    // domain faces of the parent level: dnx = parent-level domain cells
    const bool bi = (I == 0 || I == dnx), bj = (J == 0 || J == dny), bk = (K == 0 || K == dnz);
    const long sy = Lf.nx, sz = (long)Lf.nx * Lf.ny;
    const long f0 = ((long)(2 * K - P.oz) * Lf.ny + (2 * J - P.oy)) * Lf.nx + (2 * I - P.ox);
    const long ci = ((long)(K - poz) * Lc.ny + (J - poy)) * Lc.nx + (I - pox);
    for (int p = 0; p < ncomp; ++p) {
        const double* F = rf + p * Lf.N + f0;
        double v;
        if (bi && bj && bk) v = F[0];
        else if (bj && bk) v = 0.25 * F[-1] + 0.5 * F[0] + 0.25 * F[1];
        else if (bk && bi) v = 0.25 * F[-sy] + 0.5 * F[0] + 0.25 * F[sy];
        else if (bi && bj) v = 0.25 * F[-sz] + 0.5 * F[0] + 0.25 * F[sz];
        else if (bi)
            v = (F[-sy - sz] + 2.0 * F[-sz] + F[sy - sz] + 2.0 * F[-sy] + 4.0 * F[0] + 2.0 * F[sy] + F[-sy + sz] +
                 2.0 * F[sz] + F[sy + sz]) / 16.0;
        else if (bj)
            v = (F[-1 - sz] + 2.0 * F[-1] + F[-1 + sz] + 2.0 * F[-sz] + 4.0 * F[0] + 2.0 * F[sz] + F[1 - sz] +
                 2.0 * F[1] + F[1 + sz]) / 16.0;
        else if (bk)
            v = (F[-1 - sy] + 2.0 * F[-sy] + F[1 - sy] + 2.0 * F[-1] + 4.0 * F[0] + 2.0 * F[1] + F[-1 + sy] +
                 2.0 * F[sy] + F[1 + sy]) / 16.0;
        else
            v = (F[-1 - sy - sz] + F[-1 - sy + sz] + F[-1 + sy - sz] + F[-1 + sy + sz] + F[1 - sy - sz] +
                 F[1 - sy + sz] + F[1 + sy - sz] + F[1 + sy + sz]) / 64.0 +
                (F[-sy - sz] + F[-sy + sz] + F[sy - sz] + F[sy + sz] + F[-1 - sz] + F[1 - sz] + F[-1 + sz] +
                 F[1 + sz] + F[-1 - sy] + F[-1 + sy] + F[1 - sy] + F[1 + sy]) / 32.0 +
                (F[-1] + F[-sy] + F[-sz] + F[1] + F[sy] + F[sz]) / 16.0 + F[0] / 8.0;
        rc[p * Lc.N + ci] = v;
    }
}

// Copy the nodes [lo, hi] (global node indices, inclusive) of ncomp planes
// from src (ns nodes per axis, node 0 at global so) into dst (nd nodes, node 0
// at global dO): the gather of one box of a multi-box field into a level's
// single-box layout, or the scatter back.
__global__ void __launch_bounds__(kBlock) k_box_copy(double* __restrict__ dst, int ndx, int ndy, int ndz, int dox,
                                                     int doy, int doz, const double* __restrict__ src, int nsx, int nsy,
                                                     int nsz, int sox, int soy, int soz, int lox, int loy, int loz,
                                                     int bx, int by, int bz, int ncomp) {
    const long id = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= (long)bx * by * bz) return;
    const int a = (int)(id % bx), b = (int)((id / bx) % by), c = (int)(id / ((long)bx * by));
    const int gx = lox + a, gy = loy + b, gz = loz + c;
    const long di = ((long)(gz - doz) * ndy + (gy - doy)) * ndx + (gx - dox);
    const long si = ((long)(gz - soz) * nsy + (gy - soy)) * nsx + (gx - sox);
    const long Nd = (long)ndx * ndy * ndz, Ns = (long)nsx * nsy * nsz;
    for (int p = 0; p < ncomp; ++p) dst[p * Nd + di] = src[p * Ns + si];
}

// out = v with the parent nodes strictly inside the patch (covered by the
// finer level, excluded from the composite norm) set to 0.
__global__ void __launch_bounds__(kBlock) k_mask_covered(const double* v, Lvl L, int pox, int poy, int poz, Patch P,
                                                         double* out) {  // out may be v
    const long id = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= L.N) return;
    int i, j, k;
    node_ijk(id, L, i, j, k);
    const int I = pox + i, J = poy + j, K = poz + k;
    const bool covered = I > P.clo[0] && I < P.chi[0] && J > P.clo[1] && J < P.chi[1] && K > P.clo[2] && K < P.chi[2];
#pragma unroll
    for (int p = 0; p < 3; ++p) out[p * L.N + id] = covered ? 0.0 : v[p * L.N + id];
}

}  // namespace ael
