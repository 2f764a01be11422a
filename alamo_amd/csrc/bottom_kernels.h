// bottom_kernels.h -- single-workgroup bottom solver for the coarsest MG level.
//
// The coarsest level of the hierarchy holds 3^3 nodes for power-of-two
// boxes, so the whole bottom solve (BiCGStab with the operator diagonal as
// left preconditioner, then the post-smoothing of a failed solve) runs in one
// workgroup: no host round trips and no launch per BiCGStab step.
// Restates amrex::MLMG::actualBottomSolve + MLCGSolver::solve_bicgstab with
// Operator<Grid::Node>::normalize (src/Operator/Operator.cpp:156-165) and the
// bicgstab choice of src/Solver/Nonlocal/Linear.H:258.
#pragma once
#include "elastic_kernels.h"

namespace ael {

constexpr int kBottomBlock = 1024;

// Block-wide max (op 0) or sum (op 1); every thread returns the same value.
__device__ __forceinline__ double block_reduce(double v, int op, double* red) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        const double o = __shfl_xor(v, off);
        v = op == 0 ? fmax(v, o) : v + o;
    }
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = blockDim.x >> 6;
    __syncthreads();
    if (lane == 0) red[wid] = v;
    __syncthreads();
    double r = 0.0;
    for (int q = 0; q < nw; ++q) r = op == 0 ? fmax(r, red[q]) : r + red[q];
    return r;
}

template <int SYM>
__device__ void blk_apply(const double* __restrict__ in, double* __restrict__ out, const double* __restrict__ C,
                          const int* __restrict__ bct, int uniform, const Lvl& L) {
    for (long id = threadIdx.x; id < L.N; id += blockDim.x) {
        int i, j, k;
        node_ijk(id, L, i, j, k);
        double o[3];
        apply_node<double, SYM>(in, C, bct, uniform, L, i, j, k, id, o);
        out[id] = o[0];
        out[L.N + id] = o[1];
        out[2 * L.N + id] = o[2];
    }
    __syncthreads();
}

// w: 10 vectors of 3N doubles of scratch.  status[0] receives the BiCGStab
// return code (0 converged, 1..4 breakdown, 8 not converged).
template <int SYM>
__global__ void __launch_bounds__(kBottomBlock)
    k_bottom_solve(double* __restrict__ x, const double* __restrict__ rhs, const double* __restrict__ C,
                   const double* __restrict__ diag, const int* __restrict__ bct, int uniform, Lvl L, double omega,
                   int maxiter, double eps_rel, int nuf, double* __restrict__ w, int* __restrict__ status) {
    __shared__ double red[kBottomBlock / 64];
    const long M = 3 * L.N;
    double *ph = w, *sh = w + M, *p = w + 2 * M, *r = w + 3 * M, *s = w + 4 * M, *rh = w + 5 * M, *v = w + 6 * M,
           *t = w + 7 * M, *tmp = w + 8 * M;
    const int tid = threadIdx.x, nt = blockDim.x;
    // actualBottomSolve: x = 0; MLCGSolver: r = b - A x, normalized.
    for (long q = tid; q < M; q += nt) x[q] = 0.0;
    __syncthreads();
    blk_apply<SYM>(x, v, C, bct, uniform, L);
    double loc = 0.0;
    for (long q = tid; q < M; q += nt) {
        const double rq = (rhs[q] + (-1.0) * v[q]) / diag[q];
        r[q] = rq;
        rh[q] = rq;
        loc = fmax(loc, fabs(rq));
    }
    const double rnorm0 = block_reduce(loc, 0, red);
    double rnorm = rnorm0, rho_1 = 0.0, alpha = 0.0, omg = 0.0;
    int ret = 0;
    if (rnorm0 != 0.0) {
        for (int iter = 1; iter <= maxiter; ++iter) {
            loc = 0.0;
            for (long q = tid; q < M; q += nt) loc += rh[q] * r[q];
            const double rho = block_reduce(loc, 1, red);
            if (rho == 0.0) { ret = 1; break; }
            if (iter == 1) {
                for (long q = tid; q < M; q += nt) p[q] = r[q];
            } else {
                const double beta = (rho / rho_1) * (alpha / omg);
                for (long q = tid; q < M; q += nt) {
                    const double pq = p[q] + (-omg) * v[q];
                    p[q] = r[q] + beta * pq;
                }
            }
            for (long q = tid; q < M; q += nt) ph[q] = p[q];
            __syncthreads();
            blk_apply<SYM>(ph, v, C, bct, uniform, L);
            loc = 0.0;
            for (long q = tid; q < M; q += nt) {
                v[q] /= diag[q];
                loc += rh[q] * v[q];
            }
            const double rhTv = block_reduce(loc, 1, red);
            if (rhTv != 0.0) alpha = rho / rhTv;
            else { ret = 2; break; }
            loc = 0.0;
            for (long q = tid; q < M; q += nt) {
                x[q] += alpha * ph[q];
                const double sq = 1.0 * r[q] + (-alpha) * v[q];
                s[q] = sq;
                sh[q] = sq;
                loc = fmax(loc, fabs(sq));
            }
            rnorm = block_reduce(loc, 0, red);
            if (rnorm < eps_rel * rnorm0) break;
            blk_apply<SYM>(sh, t, C, bct, uniform, L);
            double l1 = 0.0, l2 = 0.0;
            for (long q = tid; q < M; q += nt) {
                t[q] /= diag[q];
                l1 += t[q] * t[q];
                l2 += t[q] * s[q];
            }
            const double tt = block_reduce(l1, 1, red);
            const double ts = block_reduce(l2, 1, red);
            if (tt != 0.0) omg = ts / tt;
            else { ret = 3; break; }
            loc = 0.0;
            for (long q = tid; q < M; q += nt) {
                x[q] += omg * sh[q];
                const double rq = 1.0 * s[q] + (-omg) * t[q];
                r[q] = rq;
                loc = fmax(loc, fabs(rq));
            }
            rnorm = block_reduce(loc, 0, red);
            if (rnorm < eps_rel * rnorm0) break;
            if (omg == 0.0) { ret = 4; break; }
            rho_1 = rho;
        }
        if (ret == 0 && rnorm > eps_rel * rnorm0) ret = 8;
        // sorig is the zero initial guess: keep x only when it reduced the residual
        if (!((ret == 0 || ret == 8) && rnorm < rnorm0))
            for (long q = tid; q < M; q += nt) x[q] = 0.0;
        // actualBottomSolve: a failed solve (ret other than 0, 9) is discarded
        if (ret != 0 && ret != 9)
            for (long q = tid; q < M; q += nt) x[q] = 0.0;
        __syncthreads();
        // then nuf Fsmooth calls after a failure (nub = 0 after success)
        const int nsm = ret == 0 ? 0 : nuf;
        for (int sm = 0; sm < 2 * nsm; ++sm) {
            const double* xin = (sm & 1) ? tmp : x;
            double* xout = (sm & 1) ? x : tmp;
            blk_apply<SYM>(xin, v, C, bct, uniform, L);
            for (long q = tid; q < M; q += nt) {
                const double xv = xin[q], dv = diag[q];
                const double Rx = v[q] - xv * dv;
                xout[q] = (1.0 - omega) * xv + omega * (rhs[q] - Rx) / dv;
            }
            __syncthreads();
        }
    }
    if (tid == 0) status[0] = ret;
}

}  // namespace ael
