// coarse_cycle.h -- the coarse tail of a multigrid V-cycle in one workgroup.
//
// On the small MG levels every smoothing sweep, residual and transfer is a
// launch of a few workgroups whose time is the launch itself (4-6 us each,
// some 50 of them per F-cycle on the 5^3 and 3^3 levels of a 129^3 problem).
// k_vcycle_tail runs amrex::MLMG::mgVcycle from MG level f down to the dense
// bottom and back up in one workgroup of kTailBlock threads with every
// operand in LDS, the steps separated by workgroup barriers.  Each step is
// the per-node body of the launched kernel it replaces (apply_node,
// finish_op, restrict_node, interp_node; the bottom is the same FP64
// product with the dense inverse as k_bottom_direct).  The host uses the
// tail only when its levels fit in the LDS of one workgroup.
#pragma once
#include "stencil.h"

namespace ael {

constexpr int kTailBlock = 512;
constexpr int kTailMaxLevels = 8;
// Levels with at most this many nodes join the tail.  One CU runs a sweep of
// the 5^3 level faster than a launch over several, but not one of the 9^3
// level (729 nodes, measured 115 us per tail call against about 90 us for
// the launches it replaced).
constexpr long kTailNodes = 256;

// One level of the tail: its global buffers and, in elements of T (bytes for
// gm), the offsets of its LDS copies.
template <typename T>
struct TailLevel {
    Lvl L;
    T* cor;                      // global; written back for the first level only
    const T* res;                // global; read for the first level only
    const T* diag;
    const T* C;
    const unsigned char* gmask;  // optional
    int oC, odiag, ores, ocor, otmp, orescor, ogm;
};

template <typename T>
struct TailArgs {
    TailLevel<T> lv[kTailMaxLevels];  // lv[0]: first level of the tail ... lv[n-1]: the bottom
    int n, nc;
    int nu1, nu2;
    T omega;
    int uniform;
    const int* bct;
    const double* Ainv;  // inverse of the bottom operator, column-major, 3 N x 3 N
};

// out = OP(x) over every node of one level (LDS operands), the workgroup
// striding over nodes.
template <typename T, int SYM, int OP>
__device__ __forceinline__ void tail_sweep(const Lvl& L, const T* C, const T* diag, const unsigned char* gm,
                                           const T* x, const T* b, T* out, T omega, int uniform,
                                           const int* __restrict__ bct) {
    for (long id = threadIdx.x; id < L.N; id += blockDim.x) {
        int i, j, k;
        node_ijk(id, L, i, j, k);
        T o[3];
        apply_node<T, SYM>(x, C, bct, uniform, L, i, j, k, id, o, gm ? (unsigned)gm[id] : 7u);
#pragma unroll
        for (int p = 0; p < 3; ++p) {
            const long q = p * L.N + id;
            const T xv = (OP == 2) ? x[q] : T(0);
            const T bv = (OP >= 1) ? b[q] : T(0);
            const T dv = (OP == 2) ? diag[q] : T(1);
            out[q] = finish_op<T, OP>(o[p], xv, bv, dv, omega);
        }
    }
}

template <typename T>
__device__ __forceinline__ void tail_copy(T* dst, const T* src, long n) {
    for (long q = threadIdx.x; q < n; q += blockDim.x) dst[q] = src[q];
}

// mgVcycle over the tail levels with every operand in LDS (dynamic shared
// memory laid out by the host): the coefficients, diagonals and gradient
// masks and the first level's residual are copied in, the first level's
// correction is copied out; the coarser levels' vectors never leave LDS
// (every mgVcycle call zeroes their corrections and restricts their
// residuals afresh).
template <typename T, int SYM>
__global__ void __launch_bounds__(kTailBlock) k_vcycle_tail(TailArgs<T> a) {
    extern __shared__ __align__(16) unsigned char tail_sm[];
    T* sm = reinterpret_cast<T*>(tail_sm);
    const int n = a.n;
    const int nt = (int)blockDim.x;
    for (int l = 0; l < n; ++l) {
        const TailLevel<T>& c = a.lv[l];
        if (l < n - 1) {
            tail_copy<T>(sm + c.oC, c.C, (long)a.nc * c.L.N);
            tail_copy<T>(sm + c.odiag, c.diag, 3 * c.L.N);
            if (c.gmask)
                for (long q = threadIdx.x; q < c.L.N; q += nt) tail_sm[c.ogm + q] = c.gmask[q];
        }
    }
    tail_copy<T>(sm + a.lv[0].ores, a.lv[0].res, 3 * a.lv[0].L.N);
    __syncthreads();
    // down: zero the correction, pre-smooth, residual, restrict
    for (int l = 0; l < n - 1; ++l) {
        const TailLevel<T>& c = a.lv[l];
        const T* C = sm + c.oC;
        const T* dg = sm + c.odiag;
        const unsigned char* gm = c.gmask ? tail_sm + c.ogm : nullptr;
        T *cor = sm + c.ocor, *tmp = sm + c.otmp, *res = sm + c.ores, *rc = sm + c.orescor;
        // zero guess; its first sweep is w b / D (k_jacobi_from_zero)
        for (long q = threadIdx.x; q < 3 * c.L.N; q += nt)
            if (a.nu1 > 0) tmp[q] = a.omega * res[q] / dg[q];
            else cor[q] = T(0);
        __syncthreads();
        for (int s = 0; s < a.nu1; ++s) {
            if (s > 0) {
                tail_sweep<T, SYM, 2>(c.L, C, dg, gm, cor, res, tmp, a.omega, a.uniform, a.bct);
                __syncthreads();
            }
            tail_sweep<T, SYM, 2>(c.L, C, dg, gm, tmp, res, cor, a.omega, a.uniform, a.bct);
            __syncthreads();
        }
        tail_sweep<T, SYM, 1>(c.L, C, dg, gm, cor, res, rc, a.omega, a.uniform, a.bct);
        __syncthreads();
        const TailLevel<T>& cc = a.lv[l + 1];
        for (long q = threadIdx.x; q < cc.L.N; q += nt) restrict_node<T>(rc, c.L, sm + cc.ores, cc.L, 3, q);
        __syncthreads();
    }
    // bottom: cor = Ainv res in FP64 (k_bottom_direct's sum order)
    {
        const TailLevel<T>& c = a.lv[n - 1];
        const int M = (int)(3 * c.L.N);
        const T* res = sm + c.ores;
        T* cor = sm + c.ocor;
        for (int row = threadIdx.x; row < M; row += nt) {
            double acc = 0.0;
            for (int col = 0; col < M; ++col) acc += a.Ainv[(long)col * M + row] * (double)res[col];
            cor[row] = (T)acc;
        }
        __syncthreads();
    }
    // up: add the interpolated correction, post-smooth
    for (int l = n - 2; l >= 0; --l) {
        const TailLevel<T>& c = a.lv[l];
        const TailLevel<T>& cc = a.lv[l + 1];
        const T* C = sm + c.oC;
        const T* dg = sm + c.odiag;
        const unsigned char* gm = c.gmask ? tail_sm + c.ogm : nullptr;
        T *cor = sm + c.ocor, *tmp = sm + c.otmp, *res = sm + c.ores;
        for (long q = threadIdx.x; q < c.L.N; q += nt) interp_node<T>(sm + cc.ocor, cc.L, cor, c.L, 3, 0, q);
        __syncthreads();
        for (int s = 0; s < a.nu2; ++s) {
            tail_sweep<T, SYM, 2>(c.L, C, dg, gm, cor, res, tmp, a.omega, a.uniform, a.bct);
            __syncthreads();
            tail_sweep<T, SYM, 2>(c.L, C, dg, gm, tmp, res, cor, a.omega, a.uniform, a.bct);
            __syncthreads();
        }
    }
    tail_copy<T>(a.lv[0].cor, sm + a.lv[0].ocor, 3 * a.lv[0].L.N);
}

}  // namespace ael
