// stencil.h -- the nodal elastic operator kernels used by every MG level.
//
// One thread per node.  Boundary nodes get workgroups of their own; interior
// nodes are XCD-slab ordered (slab_interior_node) so that the planes a node
// reads stay in one XCD's L2, and the optional grad-C mask skips the
// neighbour loads of axes along which C does not change.
//
// OP 0: out = A x            (Elastic::Fapply, src/Operator/Elastic.cpp:139-292)
// OP 1: out = b - A x        (correctionResidual, src/Operator/Operator.cpp:647-657)
// OP 2: out = (1-w) x + w (b - (A x - D x)) / D   (one Fsmooth sweep,
//                              src/Operator/Operator.cpp:109-146)
#pragma once
#include "elastic_kernels.h"

namespace ael {

template <typename T, int OP>
__device__ __forceinline__ T finish_op(T Ax, T xv, T bv, T dv, T omega) {
    if constexpr (OP == 0) return Ax;
    else if constexpr (OP == 1) return bv + T(-1) * Ax;
    else return (T(1) - omega) * xv + omega * (bv - (Ax - xv * dv)) / dv;
}

// One thread per node over the whole box (XCD-remapped), any OP.
template <typename T, int SYM, int OP>
__global__ void __launch_bounds__(kBlock)
    k_stencil_node(const T* __restrict__ x, const T* __restrict__ b, const T* __restrict__ diag,
                   const T* __restrict__ C, const int* __restrict__ bct, int uniform, Lvl L, T omega,
                   T* __restrict__ out) {
    const long id = xcd_node();
    if (id >= L.N) return;
    int i, j, k;
    node_ijk(id, L, i, j, k);
    T o[3];
    apply_node<T, SYM>(x, C, bct, uniform, L, i, j, k, id, o);
#pragma unroll
    for (int p = 0; p < 3; ++p) {
        const long qq = p * L.N + id;
        const T xv = (OP == 2) ? x[qq] : T(0);
        const T bv = (OP >= 1) ? b[qq] : T(0);
        const T dv = (OP == 2) ? diag[qq] : T(1);
        out[qq] = finish_op<T, OP>(o[p], xv, bv, dv, omega);
    }
}

// Launch layout of the stencil kernels.  The first nbb workgroups (a
// multiple of kXcd) take the nodes on the faces of the box, the rest the
// interior box [1, nx-2] x [1, ny-2] x [1, nz-2].  Boundary rows need
// one-sided gradients and the BC table; a wave that mixed them with interior
// nodes would run both code paths, and with rows of nx nodes straddling
// 64-lane waves about half of all waves would.
//
// Interior nodes are XCD-slab ordered.  Workgroups are dispatched to the kXcd
// XCDs round robin (workgroup b runs on XCD b % kXcd), each XCD with its own
// L2.  XCD x is given the interior rows j in [x*nj/kXcd, (x+1)*nj/kXcd) of
// every plane and walks that slab in (k, j, i) order, so a node's k-1 and k+1
// neighbours were read a few workgroups earlier, or will be read a few later,
// on the same L2: u and C are fetched from HBM about once (plus the rows each
// slab shares with its neighbours) instead of once per plane that reads them.
__host__ __device__ inline long interior_count(const Lvl& L) {
    return L.nx > 2 && L.ny > 2 && L.nz > 2 ? (long)(L.nx - 2) * (L.ny - 2) * (L.nz - 2) : 0;
}
__host__ __device__ inline long boundary_count(const Lvl& L) { return L.N - interior_count(L); }

// The t-th boundary node, t < boundary_count(L): the planes k = 0 and
// k = nz-1 whole, then for each inner plane its ring of face nodes.
__device__ __forceinline__ void boundary_node(const Lvl& L, long t, int& i, int& j, int& k) {
    const long plane = (long)L.nx * L.ny;
    if (t < 2 * plane) {
        const bool top = t >= plane;
        const long r = top ? t - plane : t;
        k = top ? L.nz - 1 : 0;
        j = (int)(r / L.nx);
        i = (int)(r - (long)j * L.nx);
        return;
    }
    t -= 2 * plane;
    const long ring = 2L * L.nx + 2L * (L.ny - 2);
    const long kk = t / ring;
    long r = t - kk * ring;
    k = 1 + (int)kk;
    if (r < L.nx) {
        j = 0;
        i = (int)r;
    } else if (r < 2L * L.nx) {
        j = L.ny - 1;
        i = (int)(r - L.nx);
    } else {
        r -= 2L * L.nx;
        j = 1 + (int)(r >> 1);
        i = (r & 1) ? L.nx - 1 : 0;
    }
}

// Interior node of interior workgroup ib (blockIdx.x - nbb), XCD-slab order.
__device__ __forceinline__ bool slab_interior_node(const Lvl& L, long ib, int& i, int& j, int& k, long& id) {
    const int ni = L.nx - 2, nj = L.ny - 2, nk = L.nz - 2;
    if (ni <= 0 || nj <= 0 || nk <= 0) return false;
    const int x = (int)(ib % kXcd);
    const long lb = ib / kXcd;
    const int j0 = (int)((long)x * nj / kXcd), j1 = (int)((long)(x + 1) * nj / kXcd);
    const long per_plane = (long)(j1 - j0) * ni;
    const long t = lb * blockDim.x + threadIdx.x;
    if (t >= per_plane * nk) return false;
    const long kk = t / per_plane;
    const long r = t - kk * per_plane;
    const int jj = (int)(r / ni);
    i = 1 + (int)(r - (long)jj * ni);
    j = 1 + j0 + jj;
    k = 1 + (int)kk;
    id = ((long)k * L.ny + j) * L.nx + i;
    return true;
}

// Boundary workgroups (rounded up to a multiple of kXcd).
inline unsigned slab_boundary_blocks(const Lvl& L, int block) {
    const long nb = (boundary_count(L) + block - 1) / block;
    return (unsigned)((nb + kXcd - 1) / kXcd * kXcd);
}
// All workgroups: the boundary ones, then kXcd times those of the tallest slab.
inline unsigned slab_blocks(const Lvl& L, int block) {
    const long nj = L.ny > 2 ? L.ny - 2 : 0;
    const long rows = (nj + kXcd - 1) / kXcd;
    const long per_xcd = (rows * (L.nx > 2 ? L.nx - 2 : 0) * (L.nz > 2 ? L.nz - 2 : 0) + block - 1) / block;
    return slab_boundary_blocks(L, block) + (unsigned)(per_xcd * kXcd);
}

// Compact grad-C matrices (MajorMinor).  D (gradc_matrix_mm) is stored only
// for the interior nodes with a nonzero gmask, in the order the interior
// waves of the slab launch meet them: in interior wave w, the flagged lanes
// take positions dwbase[w] + (number of flagged lanes below them), and entry
// q of D sits at Dc[q * nflag + p].  A wave reads its D entries with 18
// coalesced loads instead of the up to 2 x 15 neighbour loads per axis of the
// central differences of C.
constexpr int kWavesPerBlock = kBlock / 64;

__device__ __forceinline__ unsigned long long lanes_below() {
    const unsigned lane = threadIdx.x & 63u;
    return lane ? (~0ull >> (64u - lane)) : 0ull;
}

// Flagged lanes per interior wave (grid: the interior workgroups only).
__global__ void __launch_bounds__(kBlock) k_gradc_count(const unsigned char* __restrict__ gmask, Lvl L,
                                                        int* __restrict__ wcount) {
    int i, j, k;
    long id = 0;
    const bool v = slab_interior_node(L, (long)blockIdx.x, i, j, k, id);
    const unsigned long long bal = __ballot(v && gmask[id] != 0);
    if ((threadIdx.x & 63u) == 0)
        wcount[(long)blockIdx.x * kWavesPerBlock + threadIdx.x / 64] = (int)__popcll(bal);
}

// D of every flagged interior node at its compact position (FP64).
__global__ void __launch_bounds__(kBlock) k_gradc_fill(const double* __restrict__ C,
                                                       const unsigned char* __restrict__ gmask, Lvl L,
                                                       const int* __restrict__ dwbase, int nflag,
                                                       double* __restrict__ Dc) {
    int i, j, k;
    long id = 0;
    const bool v = slab_interior_node(L, (long)blockIdx.x, i, j, k, id);
    const bool f = v && gmask[id] != 0;
    const unsigned long long bal = __ballot(f);
    if (!f) return;
    const int p = dwbase[(long)blockIdx.x * kWavesPerBlock + threadIdx.x / 64] + (int)__popcll(bal & lanes_below());
    double D[18];
    gradc_matrix_mm(C, L, (unsigned)id, D);
#pragma unroll
    for (int q = 0; q < 18; ++q) Dc[(size_t)q * nflag + p] = D[q];
}

// One thread per node in the layout above, any OP.  gmask (optional, one
// byte per node) holds bit a set where C(+a) != C(-a) in some slot; the
// grad-C contribution of a clear axis is exactly zero and its 2 x NC
// neighbour loads are skipped.  DC (MajorMinor): the grad-C term of the
// flagged nodes is D e1 with D read from the compact array Dc.
// PER: L has periodic axes (their face nodes take the wrapped interior row);
// the variant without is the one every non-periodic level runs.
template <typename T, int SYM, int OP, bool DC = false, bool PER = false>
__global__ void __launch_bounds__(kBlock)
    k_stencil_slab(const T* __restrict__ x, const T* __restrict__ b, const T* __restrict__ diag,
                   const T* __restrict__ C, const int* __restrict__ bct, int uniform, Lvl L, T omega,
                   const unsigned char* __restrict__ gmask, unsigned nbb, const int* __restrict__ dwbase,
                   const T* __restrict__ Dc, int nflag, T* __restrict__ out) {
    int i, j, k;
    long id;
    T o[3];
    if (blockIdx.x < nbb) {
        const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
        if (t >= boundary_count(L)) return;
        boundary_node(L, t, i, j, k);
        id = ((long)k * L.ny + j) * L.nx + i;
        T nrm[3];
        const int reg = region_of<T, PER>(i, j, k, L, nrm);
        if constexpr (PER) {
            if (reg >= 0) apply_boundary<T, SYM, T, true>(x, C, bct, L, i, j, k, id, reg, nrm, o);
            else apply_interior_wrap<T, SYM>(x, C, uniform, L, i, j, k, id, o);  // a periodic face
        } else {
            apply_boundary<T, SYM, T, false>(x, C, bct, L, i, j, k, id, reg, nrm, o);
        }
    } else {
        const long ib = (long)(blockIdx.x - nbb);
        if (!slab_interior_node(L, ib, i, j, k, id)) return;
        if constexpr (SYM == 1 && DC) {
            // Every load of the node leaves first (the 19-point u neighbourhood,
            // C, the mask, the wave's first D position and the OP operands),
            // ahead of the arithmetic: on the latency-bound small levels the
            // kernel then waits on memory twice (these, then D), not once per
            // group of loads the compiler would otherwise interleave.
            const unsigned o32 = (unsigned)id;
            // wave-uniform; readfirstlane (convergent) keeps the load ahead of
            // the flagged-lanes branch, where the compiler would otherwise sink it
            const int dwb = __builtin_amdgcn_readfirstlane(dwbase[ib * kWavesPerBlock + threadIdx.x / 64]);
            const unsigned gm = gmask[id];  // the compact-D launch always passes the mask
            T uv[3][19], cv[21], xv[3], bv[3], dv[3];
            mm_load<T>(x, C, L, o32, uv, cv);
#pragma unroll
            for (int p = 0; p < 3; ++p) {
                const size_t off = (size_t)p * L.N;
                xv[p] = (OP == 2) ? uv[p][0] : T(0);
                bv[p] = (OP >= 1) ? ld32(b + off, o32) : T(0);
                dv[p] = (OP == 2) ? ld32(diag + off, o32) : T(1);
            }
            __builtin_amdgcn_sched_barrier(0);
            T e1[6];
            mm_compute<T>(uv, cv, L, o, e1);
            const unsigned long long bal = __ballot(gm != 0u);
            if (!uniform && gm != 0u) {
                const unsigned p = (unsigned)dwb + (unsigned)__popcll(bal & lanes_below());
#pragma unroll
                for (int r = 0; r < 3; ++r)
#pragma unroll
                    for (int Q = 0; Q < 6; ++Q) o[r] += ld32(Dc + (size_t)(6 * r + Q) * nflag, p) * e1[Q];
            }
#pragma unroll
            for (int p = 0; p < 3; ++p)
                st32(out + (size_t)p * L.N, o32, finish_op<T, OP>(o[p], xv[p], bv[p], dv[p], omega));
            return;
        } else {
            apply_interior<T, SYM>(x, C, uniform, L, id, o, gmask ? (unsigned)gmask[id] : 7u);
        }
    }
    const unsigned n32 = (unsigned)id;
#pragma unroll
    for (int p = 0; p < 3; ++p) {
        const size_t off = (size_t)p * L.N;
        const T xv = (OP == 2) ? ld32(x + off, n32) : T(0);
        const T bv = (OP >= 1) ? ld32(b + off, n32) : T(0);
        const T dv = (OP == 2) ? ld32(diag + off, n32) : T(1);
        st32(out + off, n32, finish_op<T, OP>(o[p], xv, bv, dv, omega));
    }
}

// gmask of every interior node: bit a set iff some slot of C differs between
// the two neighbours along axis a (bitwise compare of the FP64 values, so the
// mask also holds for any rounding of C to FP32).  Boundary nodes get 7.
template <int NC>
__global__ void __launch_bounds__(kBlock) k_grad_mask(const double* __restrict__ C, Lvl L,
                                                      unsigned char* __restrict__ gmask) {
    const long id = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= L.N) return;
    int i, j, k;
    node_ijk(id, L, i, j, k);
    if (i == 0 || j == 0 || k == 0 || i == L.nx - 1 || j == L.ny - 1 || k == L.nz - 1) {
        gmask[id] = 7;
        return;
    }
    const long st[3] = {1, (long)L.nx, (long)L.nx * L.ny};
    unsigned m = 0;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        bool diff = false;
        for (int s = 0; s < NC; ++s)
            diff |= __double_as_longlong(C[s * L.N + id + st[a]]) != __double_as_longlong(C[s * L.N + id - st[a]]);
        m |= diff ? (1u << a) : 0u;
    }
    gmask[id] = (unsigned char)m;
}

}  // namespace ael
