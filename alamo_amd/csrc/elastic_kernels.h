// elastic_kernels.h -- HIP/CDNA4 device code of the nodal elastic operator.
//
// Kernels are templated on the arithmetic type T (double or float) and the
// stiffness symmetry SYM (0 isotropic, 1 major-minor, 2 diagonal).  Field
// layout: one box per level, node index i fastest, vector fields component
// slowest (v[c*N + node]), stiffness structure-of-arrays (C[s*N + node]).
// Reference paths are relative to the reference root.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

namespace ael {

constexpr int kBlock = 256;

// Workgroups are dispatched round-robin over the 8 XCDs of the MI355X (block
// b runs on XCD b % 8), each XCD with its own L2.  Stencil kernels remap the
// block index so that every XCD walks one contiguous run of nodes: the y and
// z neighbours a workgroup reads were then loaded into the same L2 by the
// workgroups just before it.  Grids are launched with a multiple of 8 blocks
// so that the remap is a bijection; blocks past the field exit.
constexpr int kXcd = 8;
__device__ __forceinline__ long xcd_node() {
    const long per = (long)gridDim.x / kXcd;
    const long lb = (long)(blockIdx.x % kXcd) * per + blockIdx.x / kXcd;
    return lb * blockDim.x + threadIdx.x;
}

template <int SYM> struct NComp;
template <> struct NComp<0> { static constexpr int value = 2; };
template <> struct NComp<1> { static constexpr int value = 21; };
template <> struct NComp<2> { static constexpr int value = 9; };

struct Lvl {
    int nx, ny, nz;
    int per;  // periodic axes: bit a set when axis a is periodic (node n-1 is node 0)
    long N;
    double dx, dy, dz;
};

// Neighbour offsets of node (i, j, k) along each axis: om[a] to the node at
// -1, op[a] to the node at +1.  Across the faces of a periodic axis they wrap
// (the nodes 0 and n-1 are the same point, so the node below 0 is n-2 and the
// node above n-1 is 1: the ghost nodes AMReX fills with the periodic images).
__device__ __forceinline__ void nb_offsets(const Lvl& L, int i, int j, int k, long om[3], long op[3]) {
    const int pos[3] = {i, j, k}, n[3] = {L.nx, L.ny, L.nz};
    const long st[3] = {1, (long)L.nx, (long)L.nx * L.ny};
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const bool per = (L.per >> a) & 1;
        const long w = (long)(n[a] - 2) * st[a];
        om[a] = (per && pos[a] == 0) ? w : -st[a];
        op[a] = (per && pos[a] == n[a] - 1) ? -w : st[a];
    }
}
__device__ __forceinline__ bool on_face(const Lvl& L, int i, int j, int k) {
    return i == 0 || j == 0 || k == 0 || i == L.nx - 1 || j == L.ny - 1 || k == L.nz - 1;
}

__host__ __device__ constexpr int vpair(int a, int b) {
    return a == b ? (a == 0 ? 0 : (a == 1 ? 3 : 5)) : ((a + b) == 1 ? 1 : ((a + b) == 2 ? 2 : 4));
}
// 21 MajorMinor slots = upper triangle (row by row) of the 6x6 matrix over the
// symmetric pairs (00,01,02,11,12,22): src/Set/Matrix4_MajorMinor.H:327-375.
__host__ __device__ constexpr int tri6(int P, int Q) {
    return P <= Q ? P * 6 - P * (P - 1) / 2 + (Q - P) : Q * 6 - Q * (Q - 1) / 2 + (P - Q);
}

// sigma = C : A  (src/Set/Matrix4_Isotropic.H:137-157,
// src/Set/Matrix4_MajorMinor.H:676-691, src/Set/Matrix4_Diagonal.H:119-132).
template <typename T, int SYM>
__device__ __forceinline__ void contract2(const T* c, const T A[3][3], T s[3][3]) {
    if constexpr (SYM == 0) {
        const T lam = c[0], mu = c[1];
        s[0][0] = (lam + T(2) * mu) * A[0][0] + lam * A[1][1] + lam * A[2][2];
        s[1][1] = lam * A[0][0] + (lam + T(2) * mu) * A[1][1] + lam * A[2][2];
        s[2][2] = lam * A[0][0] + lam * A[1][1] + (lam + T(2) * mu) * A[2][2];
        s[1][2] = s[2][1] = mu * (A[1][2] + A[2][1]);
        s[2][0] = s[0][2] = mu * (A[2][0] + A[0][2]);
        s[0][1] = s[1][0] = mu * (A[0][1] + A[1][0]);
    } else if constexpr (SYM == 1) {
        const T e[6] = {A[0][0], A[0][1] + A[1][0], A[0][2] + A[2][0], A[1][1], A[1][2] + A[2][1], A[2][2]};
        constexpr int ra[6] = {0, 0, 0, 1, 1, 2}, rb[6] = {0, 1, 2, 1, 2, 2};
#pragma unroll
        for (int P = 0; P < 6; ++P) {
            T acc = T(0);
#pragma unroll
            for (int Q = 0; Q < 6; ++Q) acc += c[tri6(P, Q)] * e[Q];
            s[ra[P]][rb[P]] = acc;
            s[rb[P]][ra[P]] = acc;
        }
    } else {
#pragma unroll
        for (int a = 0; a < 3; ++a)
#pragma unroll
            for (int b = 0; b < 3; ++b) s[a][b] = c[3 * a + b] * A[a][b];
    }
}

// Column col of C : A only (the grad-C term needs one column per direction).
template <typename T, int SYM>
__device__ __forceinline__ void contract2_col(const T* c, const T A[3][3], int col, T out[3]) {
    if constexpr (SYM == 0) {
        const T lam = c[0], mu = c[1];
        const T tr = A[0][0] + A[1][1] + A[2][2];
#pragma unroll
        for (int r = 0; r < 3; ++r) out[r] = mu * (A[r][col] + A[col][r]) + (r == col ? lam * tr : T(0));
    } else if constexpr (SYM == 1) {
        const T e[6] = {A[0][0], A[0][1] + A[1][0], A[0][2] + A[2][0], A[1][1], A[1][2] + A[2][1], A[2][2]};
#pragma unroll
        for (int r = 0; r < 3; ++r) {
            T acc = T(0);
            const int P = vpair(r, col);
#pragma unroll
            for (int Q = 0; Q < 6; ++Q) acc += c[tri6(P, Q)] * e[Q];
            out[r] = acc;
        }
    } else {
#pragma unroll
        for (int r = 0; r < 3; ++r) out[r] = c[3 * r + col] * A[r][col];
    }
}

// f_i = C_ijkl B(k,l,j), B(k,l,j) = u_{k,lj}  (src/Set/Matrix4_Isotropic.H:163-173,
// src/Set/Matrix4_MajorMinor.H:709-767, src/Set/Matrix4_Diagonal.H:134-143).
template <typename T, int SYM>
__device__ __forceinline__ void contract3(const T* c, const T B[3][3][3], T f[3]) {
    if constexpr (SYM == 0) {
        const T lam = c[0], mu = c[1];
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            T r = T(0);
#pragma unroll
            for (int j = 0; j < 3; ++j) r += mu * (B[i][j][j] + B[j][i][j]) + lam * B[j][j][i];
            f[i] = r;
        }
    } else if constexpr (SYM == 1) {
        T e[3][6];
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            e[j][0] = B[0][0][j];
            e[j][1] = B[0][1][j] + B[1][0][j];
            e[j][2] = B[0][2][j] + B[2][0][j];
            e[j][3] = B[1][1][j];
            e[j][4] = B[1][2][j] + B[2][1][j];
            e[j][5] = B[2][2][j];
        }
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            T r = T(0);
#pragma unroll
            for (int j = 0; j < 3; ++j) {
                const int P = vpair(i, j);
#pragma unroll
                for (int Q = 0; Q < 6; ++Q) r += c[tri6(P, Q)] * e[j][Q];
            }
            f[i] = r;
        }
    } else {
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            T r = T(0);
#pragma unroll
            for (int j = 0; j < 3; ++j) r += c[3 * i + j] * B[i][j][j];
            f[i] = r;
        }
    }
}

// Boundary region in BC::Operator::Elastic::Elastic::Face order, classified
// as Constant::operator() does (src/BC/Operator/Elastic/Constant.H:373-400);
// normals use the truncated constants of src/BC/Operator/Elastic/Elastic.H:79-80.
// PER false: the caller knows that no axis of L is periodic (the stencil
// kernels' variants without the wrap).
template <typename T, bool PER = true>
__device__ __forceinline__ int region_of(int i, int j, int k, const Lvl& L, T nrm[3]) {
    // the faces of a periodic axis are no boundary (Elastic<SYM>::Fapply
    // classifies on growPeriodicDomain(1), src/Operator/Elastic.cpp:145)
    const bool px = PER && (L.per & 1), py = PER && ((L.per >> 1) & 1), pz = PER && ((L.per >> 2) & 1);
    const bool xl = !px && i == 0, xh = !px && i == L.nx - 1, yl = !py && j == 0, yh = !py && j == L.ny - 1,
               zl = !pz && k == 0, zh = !pz && k == L.nz - 1;
    const bool bx = xl || xh, by = yl || yh, bz = zl || zh;
    if (!(bx || by || bz)) return -1;
    const T S3 = T(0.57735026919), S2 = T(0.70710678118);
    const T sx = xl ? T(-1) : T(1), sy = yl ? T(-1) : T(1), sz = zl ? T(-1) : T(1);
    if (bx && by && bz) {
        nrm[0] = sx * S3; nrm[1] = sy * S3; nrm[2] = sz * S3;
        return 18 + (xh ? 4 : 0) + (yh ? 2 : 0) + (zh ? 1 : 0);
    }
    if (by && bz) { nrm[0] = T(0); nrm[1] = sy * S2; nrm[2] = sz * S2; return 6 + (yh ? 2 : 0) + (zh ? 1 : 0); }
    if (bz && bx) { nrm[0] = sx * S2; nrm[1] = T(0); nrm[2] = sz * S2; return 10 + (zh ? 2 : 0) + (xh ? 1 : 0); }
    if (bx && by) { nrm[0] = sx * S2; nrm[1] = sy * S2; nrm[2] = T(0); return 14 + (xh ? 2 : 0) + (yh ? 1 : 0); }
    nrm[0] = nrm[1] = nrm[2] = T(0);
    if (bx) { nrm[0] = sx; return xl ? 0 : 3; }
    if (by) { nrm[1] = sy; return yl ? 1 : 4; }
    nrm[2] = sz;
    return zl ? 2 : 5;
}

// Constant::set (src/BC/Operator/Elastic/Constant.H:408-425).
template <typename T>
__device__ __forceinline__ void bc_set(const int* __restrict__ t3, const T u[3], const T gu[3][3], const T sg[3][3],
                                       const T n[3], T f[3]) {
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const int ty = t3[c];
        T v = T(0);
        if (ty == 0) v = u[c];
        else if (ty == 1) v = sg[c][0] * n[0] + sg[c][1] * n[1] + sg[c][2] * n[2];
        else if (ty == 3) v = gu[c][0] * n[0] + gu[c][1] * n[1] + gu[c][2] * n[2];
        f[c] = v;
    }
}

__device__ __forceinline__ void node_ijk(long id, const Lvl& L, int& i, int& j, int& k) {
    const long t = id / L.nx;
    i = (int)(id - t * L.nx);
    k = (int)(t / L.ny);
    j = (int)(t - (long)k * L.ny);
}

// One-sided first derivative at the faces, central inside (src/Numeric/Stencil.H:70-191).
// Branch-free: both loads are always issued (and always in the box), so a
// wave with lanes on different faces waits on memory once, not once per case;
// the one-sided factor 1 is exact, so each case rounds as (a - b) / h or
// (a - b) * 0.5 / h.
// f may be narrower than T (an FP32 field in an FP64 row): widened exactly.
// On a periodic axis the derivative is central everywhere, the neighbours
// across the faces wrapped (nb_offsets).
template <typename T, typename TU>
__device__ __forceinline__ T d1(const TU* __restrict__ f, long id, long stride, int pos, int n, T h,
                                bool per = false) {
    const bool lo = pos == 0, hi = pos == n - 1;
    if (per) {
        const long w = (long)(n - 2) * stride;
        const T a = (T)f[hi ? id - w : id + stride], b = (T)f[lo ? id + w : id - stride];
        return (a - b) * T(0.5) / h;
    }
    const T a = (T)f[hi ? id : id + stride], b = (T)f[lo ? id : id - stride];
    return (a - b) * (lo || hi ? T(1) : T(0.5)) / h;
}

// Pair index P <-> (pair_a(P), pair_b(P)) of (00,01,02,11,12,22), and the
// MajorMinor slot s <-> (slot_p(s), slot_q(s)), P <= Q, of the 6x6 pair matrix.
__host__ __device__ constexpr int pair_a(int P) { return P < 3 ? 0 : (P < 5 ? 1 : 2); }
__host__ __device__ constexpr int pair_b(int P) { return P < 3 ? P : (P < 5 ? P - 2 : 2); }
// The r with vpair(r, a) == P, or -1 when a is not an index of pair P.
__host__ __device__ constexpr int pair_other(int P, int a) {
    return pair_a(P) == a ? pair_b(P) : (pair_b(P) == a ? pair_a(P) : -1);
}
__host__ __device__ constexpr int slot_p(int s) { return s < 6 ? 0 : s < 11 ? 1 : s < 15 ? 2 : s < 18 ? 3 : s < 20 ? 4 : 5; }
__host__ __device__ constexpr int slot_q(int s) {
    return s < 6 ? s : s < 11 ? s - 5 : s < 15 ? s - 9 : s < 18 ? s - 12 : s < 20 ? s - 14 : 5;
}
constexpr bool slot_table_ok() {
    for (int s = 0; s < 21; ++s)
        if (tri6(slot_p(s), slot_q(s)) != s || slot_p(s) > slot_q(s)) return false;
    for (int P = 0; P < 6; ++P)
        if (vpair(pair_a(P), pair_b(P)) != P) return false;
    return true;
}
static_assert(slot_table_ok(), "slot/pair tables disagree with tri6/vpair");

// Slot s of dC/dx_a enters column a of (dC/dx_a) : grad u.
__host__ __device__ constexpr bool slot_in_col(int s, int a) {
    return pair_other(slot_p(s), a) >= 0 || pair_other(slot_q(s), a) >= 0;
}

// Load base[idx] with a 32-bit element index: with base uniform, the address
// is a scalar base plus a 32-bit byte offset (no 64-bit vector address math).
// Every array indexed this way holds fewer than 2^32 bytes (checked at create).
template <typename T>
__device__ __forceinline__ T ld32(const T* __restrict__ base, unsigned idx) {
    return *reinterpret_cast<const T*>(reinterpret_cast<const char*>(base) + (size_t)(idx * (unsigned)sizeof(T)));
}
template <typename T>
__device__ __forceinline__ void st32(T* __restrict__ base, unsigned idx, T v) {
    *reinterpret_cast<T*>(reinterpret_cast<char*>(base) + (size_t)(idx * (unsigned)sizeof(T))) = v;
}
// The same with the non-temporal hint (global_load/store ... nt): for data
// streamed once per launch, so that it does not displace reused lines (the u
// halos) from L2.
template <typename T>
__device__ __forceinline__ T ld32nt(const T* __restrict__ base, unsigned idx) {
    return __builtin_nontemporal_load(
        reinterpret_cast<const T*>(reinterpret_cast<const char*>(base) + (size_t)(idx * (unsigned)sizeof(T))));
}
template <typename T>
__device__ __forceinline__ void st32nt(T* __restrict__ base, unsigned idx, T v) {
    __builtin_nontemporal_store(v, reinterpret_cast<T*>(reinterpret_cast<char*>(base) + (size_t)(idx * (unsigned)sizeof(T))));
}

// + sum_a ((dC/dx_a) : grad u).col(a)  (src/Operator/Elastic.cpp:245-254) for
// MajorMinor C at interior node o, for the axes whose bit is set in gmask;
// e1 = symmetric gradient pairs of u at the node.  Accumulates into out.
template <typename T>
__device__ __forceinline__ void gradc_mm(const T* __restrict__ C, const Lvl& L, unsigned o, unsigned gmask,
                                         const T e1[6], T out[3]) {
    const unsigned N = (unsigned)L.N, sy = (unsigned)L.nx, sz = (unsigned)(L.nx * L.ny);
    const unsigned st[3] = {1u, sy, sz};
    const T h2[3] = {T(0.5) * (T(1) / T(L.dx)), T(0.5) * (T(1) / T(L.dy)), T(0.5) * (T(1) / T(L.dz))};
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        if (!((gmask >> a) & 1u)) continue;
#pragma unroll
        for (int s = 0; s < 21; ++s) {
            if (!slot_in_col(s, a)) continue;
            const int P = slot_p(s), Q = slot_q(s);
            const T* Cs = C + (size_t)s * N;
            const T d = (ld32(Cs, o + st[a]) - ld32(Cs, o - st[a])) * h2[a];
            if (pair_other(P, a) >= 0) out[pair_other(P, a)] += d * e1[Q];
            if (Q != P && pair_other(Q, a) >= 0) out[pair_other(Q, a)] += d * e1[P];
        }
    }
}

// Interior row of Fapply for MajorMinor C, the same sums as contract3 and
// contract2_col but streamed over the 21 slots: each slot is loaded once,
// used for every term it enters, and dropped, so only the 18 second
// derivative combinations e[j][Q], the 6 symmetric gradients e1[Q] and the
// result stay live.  Grid spacings enter through reciprocals computed once.
// The operands of mm_core at interior node o: the 19-point neighbourhood of
// each component, uv[p][n] with n = centre, x+, x-, y+, y-, z+, z-, then the
// xy, xz and yz corners (++, --, +-, -+), and the 21 slots of C.
template <typename T>
__device__ __forceinline__ void mm_load(const T* __restrict__ u, const T* __restrict__ C, const Lvl& L, unsigned o,
                                        T uv[3][19], T cv[21]) {
    const unsigned N = (unsigned)L.N, sy = (unsigned)L.nx, sz = (unsigned)(L.nx * L.ny);
    const unsigned off[19] = {o,           o + 1,       o - 1,       o + sy,      o - sy,      o + sz,      o - sz,
                              o + 1 + sy,  o - 1 - sy,  o + 1 - sy,  o - 1 + sy,  o + 1 + sz,  o - 1 - sz,  o + 1 - sz,
                              o - 1 + sz,  o + sy + sz, o - sy - sz, o + sy - sz, o - sy + sz};
#pragma unroll
    for (int p = 0; p < 3; ++p)
#pragma unroll
        for (int n = 0; n < 19; ++n) uv[p][n] = ld32(u + (size_t)p * N, off[n]);
#pragma unroll
    for (int s = 0; s < 21; ++s) cv[s] = ld32(C + (size_t)s * N, o);
}
template <typename T>
__device__ __forceinline__ void mm_compute(const T uv[3][19], const T cv[21], const Lvl& L, T out[3], T e1[6]) {
    const T ihx = T(1) / T(L.dx), ihy = T(1) / T(L.dy), ihz = T(1) / T(L.dz);
    const T hx2 = T(0.5) * ihx, hy2 = T(0.5) * ihy, hz2 = T(0.5) * ihz;
    const T ixx = ihx * ihx, iyy = ihy * ihy, izz = ihz * ihz;
    const T ixy = T(0.25) * ihx * ihy, ixz = T(0.25) * ihx * ihz, iyz = T(0.25) * ihy * ihz;
    T gu[3][3], B[3][3][3];
#pragma unroll
    for (int p = 0; p < 3; ++p) {
        const T* f = uv[p];
        const T c0 = f[0], xp = f[1], xm = f[2], yp = f[3], ym = f[4], zp = f[5], zm = f[6];
        gu[p][0] = (xp - xm) * hx2;
        gu[p][1] = (yp - ym) * hy2;
        gu[p][2] = (zp - zm) * hz2;
        B[p][0][0] = (xp - T(2) * c0 + xm) * ixx;
        B[p][1][1] = (yp - T(2) * c0 + ym) * iyy;
        B[p][2][2] = (zp - T(2) * c0 + zm) * izz;
        B[p][0][1] = (f[7] + f[8] - f[9] - f[10]) * ixy;
        B[p][0][2] = (f[11] + f[12] - f[13] - f[14]) * ixz;
        B[p][1][2] = (f[15] + f[16] - f[17] - f[18]) * iyz;
        B[p][1][0] = B[p][0][1];
        B[p][2][0] = B[p][0][2];
        B[p][2][1] = B[p][1][2];
    }
    T e[3][6];
#pragma unroll
    for (int Q = 0; Q < 6; ++Q) {
        const int qa = pair_a(Q), qb = pair_b(Q);
#pragma unroll
        for (int j = 0; j < 3; ++j) e[j][Q] = qa == qb ? B[qa][qa][j] : B[qa][qb][j] + B[qb][qa][j];
        e1[Q] = qa == qb ? gu[qa][qa] : gu[qa][qb] + gu[qb][qa];
    }
    out[0] = out[1] = out[2] = T(0);
    // C : grad grad u  (src/Set/Matrix4_MajorMinor.H:709-767)
#pragma unroll
    for (int s = 0; s < 21; ++s) {
        const int P = slot_p(s), Q = slot_q(s);
        const T c = cv[s];
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            if (pair_other(P, j) >= 0) out[pair_other(P, j)] += c * e[j][Q];
            if (Q != P && pair_other(Q, j) >= 0) out[pair_other(Q, j)] += c * e[j][P];
        }
    }
}
template <typename T>
__device__ __forceinline__ void mm_core(const T* __restrict__ u, const T* __restrict__ C, const Lvl& L, unsigned o,
                                        T out[3], T e1[6]) {
    T uv[3][19], cv[21];
    mm_load<T>(u, C, L, o, uv, cv);
    mm_compute<T>(uv, cv, L, out, e1);
}

template <typename T>
__device__ __forceinline__ void apply_interior_mm(const T* __restrict__ u, const T* __restrict__ C, const Lvl& L,
                                                  unsigned o, T out[3], unsigned gmask, bool uniform) {
    T e1[6];
    mm_core<T>(u, C, L, o, out, e1);
    if (uniform) return;
    gradc_mm<T>(C, L, o, gmask, e1, out);
}

// The grad-C term of gradc_mm as a matrix on the symmetric gradient pairs:
// sum_a ((dC/dx_a) : grad u).col(a) = D e1 with D[r][Q] (r = output row,
// Q = pair) the sum of the central differences of the slots that meet row r
// and pair Q.  D is zero where C(+a) == C(-a) for every axis a, and it
// depends on C only, so it is computed once per model (FP64).
__device__ __forceinline__ void gradc_matrix_mm(const double* __restrict__ C, const Lvl& L, unsigned o,
                                                double D[18]) {
    const unsigned N = (unsigned)L.N, sy = (unsigned)L.nx, sz = (unsigned)(L.nx * L.ny);
    const unsigned st[3] = {1u, sy, sz};
    const double h2[3] = {0.5 * (1.0 / L.dx), 0.5 * (1.0 / L.dy), 0.5 * (1.0 / L.dz)};
#pragma unroll
    for (int q = 0; q < 18; ++q) D[q] = 0.0;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
#pragma unroll
        for (int s = 0; s < 21; ++s) {
            if (!slot_in_col(s, a)) continue;
            const int P = slot_p(s), Q = slot_q(s);
            const double* Cs = C + (size_t)s * N;
            const double d = (ld32(Cs, o + st[a]) - ld32(Cs, o - st[a])) * h2[a];
            if (pair_other(P, a) >= 0) D[6 * pair_other(P, a) + Q] += d;
            if (Q != P && pair_other(Q, a) >= 0) D[6 * pair_other(Q, a) + P] += d;
        }
    }
}

// sigma = C : A for MajorMinor C, streamed over the slots (contract2).
template <typename T>
__device__ __forceinline__ void contract2_mm_stream(const T* __restrict__ C, long N, long id, const T A[3][3],
                                                    T sg[3][3]) {
    T e1[6], sp[6];
#pragma unroll
    for (int Q = 0; Q < 6; ++Q) {
        const int qa = pair_a(Q), qb = pair_b(Q);
        e1[Q] = qa == qb ? A[qa][qa] : A[qa][qb] + A[qb][qa];
        sp[Q] = T(0);
    }
#pragma unroll
    for (int s = 0; s < 21; ++s) {
        const int P = slot_p(s), Q = slot_q(s);
        const T c = C[s * N + id];
        sp[P] += c * e1[Q];
        if (Q != P) sp[Q] += c * e1[P];
    }
#pragma unroll
    for (int P = 0; P < 6; ++P) sg[pair_a(P)][pair_b(P)] = sg[pair_b(P)][pair_a(P)] = sp[P];
}

// Boundary row of Fapply at node (i, j, k) of region reg with normal nrm:
// one-sided gradients, sigma = C : grad u, and the Constant BC row.  u may be
// narrower than T (widened exactly).  PER false: no axis of L is periodic.
template <typename T, int SYM, typename TU = T, bool PER = true>
__device__ __forceinline__ void apply_boundary(const TU* __restrict__ u, const T* __restrict__ C,
                                               const int* __restrict__ bct, const Lvl& L, int i, int j, int k,
                                               long id, int reg, const T nrm[3], T out[3], long cN = -1,
                                               long cid = -1) {
    // C slot s of the node at C[s * cN + cid] (a separate layout of C), or at
    // C[s * N + id] by default
    const long N = L.N, sy = L.nx, sz = (long)L.nx * L.ny;
    if (cN < 0) {
        cN = N;
        cid = id;
    }
    const T hx = T(L.dx), hy = T(L.dy), hz = T(L.dz);
    T uu[3], gu[3][3], sg[3][3];
#pragma unroll
    for (int p = 0; p < 3; ++p) {
        const TU* f = u + p * N;
        uu[p] = (T)f[id];
        gu[p][0] = d1(f, id, 1, i, L.nx, hx, PER && (L.per & 1));
        gu[p][1] = d1(f, id, sy, j, L.ny, hy, PER && ((L.per >> 1) & 1));
        gu[p][2] = d1(f, id, sz, k, L.nz, hz, PER && ((L.per >> 2) & 1));
    }
    if constexpr (SYM == 1) {
        contract2_mm_stream<T>(C, cN, cid, gu, sg);
    } else {
        constexpr int NC = NComp<SYM>::value;
        T c[NC];
#pragma unroll
        for (int s = 0; s < NC; ++s) c[s] = C[s * cN + cid];
        contract2<T, SYM>(c, gu, sg);
    }
    bc_set<T>(bct + 3 * reg, uu, gu, sg, nrm, out);
}

// Interior row of Fapply at node id (no face of the box through it).
template <typename T, int SYM>
__device__ __forceinline__ void apply_interior(const T* __restrict__ u, const T* __restrict__ C, int uniform,
                                               const Lvl& L, long id, T out[3], unsigned gmask) {
    if constexpr (SYM == 1) {
        apply_interior_mm<T>(u, C, L, (unsigned)id, out, gmask, uniform != 0);
    } else {
        constexpr int NC = NComp<SYM>::value;
        const long N = L.N, sy = L.nx, sz = (long)L.nx * L.ny;
        const T hx = T(L.dx), hy = T(L.dy), hz = T(L.dz);
        T c[NC];
#pragma unroll
        for (int s = 0; s < NC; ++s) c[s] = C[s * N + id];
        T gu[3][3], B[3][3][3];
#pragma unroll
        for (int p = 0; p < 3; ++p) {
            const T* f = u + p * N;
            const T c0 = f[id];
            const T xp = f[id + 1], xm = f[id - 1], yp = f[id + sy], ym = f[id - sy], zp = f[id + sz], zm = f[id - sz];
            gu[p][0] = (xp - xm) * T(0.5) / hx;
            gu[p][1] = (yp - ym) * T(0.5) / hy;
            gu[p][2] = (zp - zm) * T(0.5) / hz;
            B[p][0][0] = (xp - T(2) * c0 + xm) / hx / hx;
            B[p][1][1] = (yp - T(2) * c0 + ym) / hy / hy;
            B[p][2][2] = (zp - T(2) * c0 + zm) / hz / hz;
            B[p][0][1] = T(0.25) * (f[id + 1 + sy] + f[id - 1 - sy] - f[id + 1 - sy] - f[id - 1 + sy]) / (hx * hy);
            B[p][0][2] = T(0.25) * (f[id + 1 + sz] + f[id - 1 - sz] - f[id + 1 - sz] - f[id - 1 + sz]) / (hx * hz);
            B[p][1][2] = T(0.25) * (f[id + sy + sz] + f[id - sy - sz] - f[id + sy - sz] - f[id - sy + sz]) / (hy * hz);
            B[p][1][0] = B[p][0][1];
            B[p][2][0] = B[p][0][2];
            B[p][2][1] = B[p][1][2];
        }
        contract3<T, SYM>(c, B, out);
        if (!uniform) {
            // f += sum_a ((dC/dx_a) : grad u).col(a)   (src/Operator/Elastic.cpp:245-254)
            const long st[3] = {1, sy, sz};
            const T h[3] = {hx, hy, hz};
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                // Bit a of gmask clear: C(+a) and C(-a) are equal slot by slot, so
                // the central difference and its contribution are exactly zero.
                if (!((gmask >> a) & 1u)) continue;
                T cg[NC], col[3];
#pragma unroll
                for (int s = 0; s < NC; ++s) cg[s] = (C[s * N + id + st[a]] - C[s * N + id - st[a]]) * T(0.5) / h[a];
                contract2_col<T, SYM>(cg, gu, a, col);
                out[0] += col[0];
                out[1] += col[1];
                out[2] += col[2];
            }
        }
    }
}

// Interior row at a node on the face of a periodic axis (and on no other
// face): the row of apply_interior with the neighbours across the periodic
// faces wrapped (nb_offsets), from the level's C (SoA) and every axis's grad-C
// term.  Only such nodes take it; the others keep their unwrapped kernels.
template <typename T, int SYM, typename TU = T>
__device__ __forceinline__ void apply_interior_wrap(const TU* __restrict__ u, const T* __restrict__ C, int uniform,
                                                 const Lvl& L, int i, int j, int k, long id, T out[3]) {
    constexpr int NC = NComp<SYM>::value;
    const long N = L.N;
    const T hx = T(L.dx), hy = T(L.dy), hz = T(L.dz);
    long om[3], op[3];
    nb_offsets(L, i, j, k, om, op);
    T c[NC];
    for (int s = 0; s < NC; ++s) c[s] = C[s * N + id];
    T gu[3][3], B[3][3][3];
    for (int p = 0; p < 3; ++p) {
        const TU* f = u + p * N;
        const T c0 = (T)f[id];
        const T xp = (T)f[id + op[0]], xm = (T)f[id + om[0]], yp = (T)f[id + op[1]], ym = (T)f[id + om[1]];
        const T zp = (T)f[id + op[2]], zm = (T)f[id + om[2]];
        gu[p][0] = (xp - xm) * T(0.5) / hx;
        gu[p][1] = (yp - ym) * T(0.5) / hy;
        gu[p][2] = (zp - zm) * T(0.5) / hz;
        B[p][0][0] = (xp - T(2) * c0 + xm) / hx / hx;
        B[p][1][1] = (yp - T(2) * c0 + ym) / hy / hy;
        B[p][2][2] = (zp - T(2) * c0 + zm) / hz / hz;
        B[p][0][1] = T(0.25) * ((T)f[id + op[0] + op[1]] + (T)f[id + om[0] + om[1]] - (T)f[id + op[0] + om[1]] -
                                (T)f[id + om[0] + op[1]]) / (hx * hy);
        B[p][0][2] = T(0.25) * ((T)f[id + op[0] + op[2]] + (T)f[id + om[0] + om[2]] - (T)f[id + op[0] + om[2]] -
                                (T)f[id + om[0] + op[2]]) / (hx * hz);
        B[p][1][2] = T(0.25) * ((T)f[id + op[1] + op[2]] + (T)f[id + om[1] + om[2]] - (T)f[id + op[1] + om[2]] -
                                (T)f[id + om[1] + op[2]]) / (hy * hz);
        B[p][1][0] = B[p][0][1];
        B[p][2][0] = B[p][0][2];
        B[p][2][1] = B[p][1][2];
    }
    contract3<T, SYM>(c, B, out);
    if (!uniform) {
        // f += sum_a ((dC/dx_a) : grad u).col(a)   (src/Operator/Elastic.cpp:245-254)
        const T h[3] = {hx, hy, hz};
        for (int a = 0; a < 3; ++a) {
            T cg[NC], col[3];
            for (int s = 0; s < NC; ++s) cg[s] = (C[s * N + id + op[a]] - C[s * N + id + om[a]]) * T(0.5) / h[a];
            contract2_col<T, SYM>(cg, gu, a, col);
            out[0] += col[0];
            out[1] += col[1];
            out[2] += col[2];
        }
    }
}

// Elastic<SYM>::Fapply at one node (src/Operator/Elastic.cpp:165-290); psi
// unset (psi_avg = 1).
template <typename T, int SYM>
__device__ __forceinline__ void apply_node(const T* __restrict__ u, const T* __restrict__ C,
                                           const int* __restrict__ bct, int uniform, const Lvl& L, int i, int j,
                                           int k, long id, T out[3], unsigned gmask = 7u) {
    T nrm[3];
    const int reg = region_of<T>(i, j, k, L, nrm);
    if (reg >= 0) apply_boundary<T, SYM>(u, C, bct, L, i, j, k, id, reg, nrm, out);
    else if (L.per && on_face(L, i, j, k)) apply_interior_wrap<T, SYM>(u, C, uniform, L, i, j, k, id, out);
    else apply_interior<T, SYM>(u, C, uniform, L, id, out, gmask);
}

// Integrator::Mechanics::UpdateModel (src/Integrator/Mechanics.H:249-253):
// model(node) = sum_n eta_n(node) * model_n, component by component, from zero
// and in model order, each product rounded before it is added (the
// reference's `model += eta * models[n]`), so no fused multiply-add.
constexpr int kMixMax = 256;  // nmodels * ncomp, passed by value
struct MixTable {
    double v[kMixMax];
};
__global__ void __launch_bounds__(kBlock) k_mix_models(const double* __restrict__ eta, int nmodels, MixTable tab,
                                                       int ncomp, long N, double* __restrict__ out) {
#pragma clang fp contract(off)
    const long id = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= N) return;
    for (int c = 0; c < ncomp; ++c) {
        double acc = 0.0;
        for (int n = 0; n < nmodels; ++n) acc += eta[n * N + id] * tab.v[n * ncomp + c];
        out[c * N + id] = acc;
    }
}

// Elastic<SYM>::Diagonal (src/Operator/Elastic.cpp:321-374) with
// Gradient_Diagonal (src/Numeric/Stencil.H:900-963).
template <typename T, int SYM>
__global__ void __launch_bounds__(kBlock) k_diagonal(const T* __restrict__ C, const int* __restrict__ bct, Lvl L,
                                                     T* __restrict__ diag) {
    constexpr int NC = NComp<SYM>::value;
    const long id = xcd_node();
    if (id >= L.N) return;
    int i, j, k;
    node_ijk(id, L, i, j, k);
    T c[NC];
#pragma unroll
    for (int s = 0; s < NC; ++s) c[s] = C[s * L.N + id];
    T nrm[3];
    const int reg = region_of<T>(i, j, k, L, nrm);
    const int pos[3] = {i, j, k}, nn[3] = {L.nx, L.ny, L.nz};
    const T h[3] = {T(L.dx), T(L.dy), T(L.dz)};
#pragma unroll
    for (int p = 0; p < 3; ++p) {
        T val;
        if (reg >= 0) {
            T gu[3][3], sg[3][3], uu[3] = {T(0), T(0), T(0)}, f[3];
#pragma unroll
            for (int a = 0; a < 3; ++a)
#pragma unroll
                for (int q = 0; q < 3; ++q) gu[q][a] = T(0);
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                // one-sided at the faces of a non-periodic axis (Gradient_Diagonal)
                const bool per = (L.per >> a) & 1;
                const bool mn = !per && pos[a] == 0, mx = !per && pos[a] == nn[a] - 1;
                gu[p][a] = ((!mx ? T(0) : T(1)) - (!mn ? T(0) : T(1))) / ((mn || mx ? T(1) : T(2)) * h[a]);
            }
            contract2<T, SYM>(c, gu, sg);
            uu[p] = T(1);
            bc_set<T>(bct + 3 * reg, uu, gu, sg, nrm, f);
            val = f[p];
        } else {
            T B[3][3][3], f[3];
#pragma unroll
            for (int a = 0; a < 3; ++a)
#pragma unroll
                for (int b2 = 0; b2 < 3; ++b2)
#pragma unroll
                    for (int d = 0; d < 3; ++d) B[a][b2][d] = T(0);
#pragma unroll
            for (int l = 0; l < 3; ++l) B[p][l][l] = T(-2) / h[l] / h[l];
            contract3<T, SYM>(c, B, f);
            val = f[p];
        }
        diag[p * L.N + id] = val;
    }
}

// Nodal full-weighting restriction of ncomp components from level Lf to Lc
// (src/Operator/Operator.cpp:280-353); the same weights average C between MG
// levels (src/Operator/Elastic.cpp:786-834).
// NC > 0 fixes the component count at compile time: the loop unrolls and
// every component's loads are issued before the first one's arithmetic (the
// coarse levels' launches are set by that latency, not by their bytes).
template <typename T, bool WRAP, int NC = 0>
__device__ __forceinline__ void restrict_node_w(const T* __restrict__ fine, const Lvl& Lf, T* __restrict__ crse,
                                                const Lvl& Lc, int ncomp, long cid) {
    int I, J, K;
    node_ijk(cid, Lc, I, J, K);
    const int hx = Lc.nx - 1, hy = Lc.ny - 1, hz = Lc.nz - 1;
    // faces of a periodic axis take the full weights, their fine neighbours
    // across the face wrapped (the coarse domain is grown there,
    // src/Operator/Operator.cpp:268)
    const bool px = WRAP && (Lc.per & 1), py = WRAP && ((Lc.per >> 1) & 1), pz = WRAP && ((Lc.per >> 2) & 1);
    const bool bi = !px && (I == 0 || I == hx), bj = !py && (J == 0 || J == hy), bk = !pz && (K == 0 || K == hz);
    const long sy = Lf.nx, sz = (long)Lf.nx * Lf.ny;
    // the fine node's -1 and +1 neighbours, wrapped on a periodic axis
    long om[3] = {-1, -sy, -sz}, op[3] = {1, sy, sz};
    if constexpr (WRAP) nb_offsets(Lf, 2 * I, 2 * J, 2 * K, om, op);
    const long x1m = om[0], x1p = op[0], sym = om[1], syp = op[1], szm = om[2], szp = op[2];
    const long f0 = (2L * K) * sz + (2L * J) * sy + 2L * I;
    const int nc = NC > 0 ? NC : ncomp;
    T vout[NC > 0 ? NC : 1];
#pragma unroll
    for (int n = 0; n < nc; ++n) {
        const T* F = fine + n * Lf.N + f0;
        T v;
        if (bi && bj && bk) v = F[0];
        else if (bj && bk) v = T(0.25) * F[x1m] + T(0.5) * F[0] + T(0.25) * F[x1p];
        else if (bk && bi) v = T(0.25) * F[sym] + T(0.5) * F[0] + T(0.25) * F[syp];
        else if (bi && bj) v = T(0.25) * F[szm] + T(0.5) * F[0] + T(0.25) * F[szp];
        else if (bi)
            v = (+F[sym + szm] + T(2) * F[szm] + F[syp + szm] + T(2) * F[sym] + T(4) * F[0] + T(2) * F[syp] +
                 F[sym + szp] + T(2) * F[szp] + F[syp + szp]) / T(16);
        else if (bj)
            v = (+F[x1m + szm] + T(2) * F[x1m] + F[x1m + szp] + T(2) * F[szm] + T(4) * F[0] + T(2) * F[szp] +
                 F[x1p + szm] + T(2) * F[x1p] + F[x1p + szp]) / T(16);
        else if (bk)
            v = (+F[x1m + sym] + T(2) * F[sym] + F[x1p + sym] + T(2) * F[x1m] + T(4) * F[0] + T(2) * F[x1p] +
                 F[x1m + syp] + T(2) * F[syp] + F[x1p + syp]) / T(16);
        else
            v = (F[x1m + sym + szm] + F[x1m + sym + szp] + F[x1m + syp + szm] + F[x1m + syp + szp] +
                 F[x1p + sym + szm] + F[x1p + sym + szp] + F[x1p + syp + szm] + F[x1p + syp + szp]) / T(64) +
                (F[sym + szm] + F[sym + szp] + F[syp + szm] + F[syp + szp] + F[x1m + szm] + F[x1p + szm] +
                 F[x1m + szp] + F[x1p + szp] + F[x1m + sym] + F[x1m + syp] + F[x1p + sym] + F[x1p + syp]) / T(32) +
                (F[x1m] + F[sym] + F[szm] + F[x1p] + F[syp] + F[szp]) / T(16) + F[0] / T(8);
        if constexpr (NC > 0) vout[n] = v;
        else crse[n * Lc.N + cid] = v;
    }
    if constexpr (NC > 0) {
#pragma unroll
        for (int n = 0; n < NC; ++n) crse[n * Lc.N + cid] = vout[n];
    }
}
template <typename T, int NC = 0>
__device__ __forceinline__ void restrict_node(const T* __restrict__ fine, const Lvl& Lf, T* __restrict__ crse,
                                              const Lvl& Lc, int ncomp, long cid) {
    if (Lc.per) restrict_node_w<T, true, NC>(fine, Lf, crse, Lc, ncomp, cid);
    else restrict_node_w<T, false, NC>(fine, Lf, crse, Lc, ncomp, cid);
}
template <typename T>
__global__ void __launch_bounds__(kBlock) k_restrict(const T* __restrict__ fine, Lvl Lf, T* __restrict__ crse,
                                                     Lvl Lc, int ncomp) {
    const long cid = xcd_node();
    if (cid >= Lc.N) return;
    if (ncomp == 3) restrict_node<T, 3>(fine, Lf, crse, Lc, 3, cid);
    else restrict_node<T>(fine, Lf, crse, Lc, ncomp, cid);
}

// Restriction followed by the first Jacobi sweep from zero on the coarse
// level, tmp = w crse / D (k_jacobi_from_zero's arithmetic on the restricted
// values), in one launch.
template <typename T>
__global__ void __launch_bounds__(kBlock) k_restrict_fz(const T* __restrict__ fine, Lvl Lf, T* __restrict__ crse,
                                                        Lvl Lc, int ncomp, const T* __restrict__ diag, T omega,
                                                        T* __restrict__ tmp) {
    const long cid = xcd_node();
    if (cid >= Lc.N) return;
    if (ncomp == 3) {
        restrict_node<T, 3>(fine, Lf, crse, Lc, 3, cid);
#pragma unroll
        for (int n = 0; n < 3; ++n) {
            const long q = n * Lc.N + cid;
            tmp[q] = omega * crse[q] / diag[q];
        }
        return;
    }
    restrict_node<T>(fine, Lf, crse, Lc, ncomp, cid);
    for (int n = 0; n < ncomp; ++n) {
        const long q = n * Lc.N + cid;
        tmp[q] = omega * crse[q] / diag[q];
    }
}

// Nodal injection crse(I,J,K) = fine(2I,2J,2K): amrex::average_down on nodal
// data, which MLLinOp::avgDownResMG applies to the residual in the F-cycle.
template <typename T>
__global__ void __launch_bounds__(kBlock) k_inject(const T* __restrict__ fine, Lvl Lf, T* __restrict__ crse, Lvl Lc,
                                                   int ncomp) {
    const long cid = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (cid >= Lc.N) return;
    int I, J, K;
    node_ijk(cid, Lc, I, J, K);
    const long f0 = (2L * K) * ((long)Lf.nx * Lf.ny) + (2L * J) * Lf.nx + 2L * I;
    for (int n = 0; n < ncomp; ++n) crse[n * Lc.N + cid] = fine[n * Lf.N + f0];
}

// Nodal injection from level 0 to every coarser level 1 .. n-1 at once:
// node (I, J, K) of level l takes node (2^l I, 2^l J, 2^l K) of level 0,
// which is what l successive injections give (each copies a node of the
// level above it).  Launch over the sum of the coarse levels' nodes.
constexpr int kInjectMaxLevels = 16;
template <typename T>
struct InjectLevels {
    int n;  // levels, the finest included
    T* f[kInjectMaxLevels];
    Lvl L[kInjectMaxLevels];
    long off[kInjectMaxLevels + 1];  // off[l]: first flat index of level l (off[1] = 0)
};
template <typename T>
__global__ void __launch_bounds__(kBlock) k_inject_all(InjectLevels<T> lv, int ncomp) {
    const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= lv.off[lv.n]) return;
    int l = 1;
    while (e >= lv.off[l + 1]) ++l;
    const Lvl& Lc = lv.L[l];
    const Lvl& L0 = lv.L[0];
    const long cid = e - lv.off[l];
    int I, J, K;
    node_ijk(cid, Lc, I, J, K);
    const long f0 = ((long)K << l) * ((long)L0.nx * L0.ny) + ((long)J << l) * L0.nx + ((long)I << l);
    for (int n = 0; n < ncomp; ++n) lv.f[l][n * Lc.N + cid] = lv.f[0][n * L0.N + f0];
}

// Trilinear prolongation of the three components, fine += P crse, or
// fine = P crse with zero_first (src/Operator/Operator.cpp:399-428), by
// coarse node: thread (I, J, K) loads the 8 coarse values of its cell once
// and writes the fine nodes (2I + a, 2J + b, 2K + c), a, b, c in {0, 1}, that
// exist.  A fine node at even (i, j, k) takes its coarse node, one with one
// odd index the mean of 2 coarse nodes, two odd the mean of 4, three odd the
// mean of 8.  No branch on the fine node's parity, and one load round trip
// for 8 fine nodes (one thread per fine node ran both parities' branches in
// every wave, each with its own loads).  With held the coarse value is
// crse + held: the interpolation of a V-cycle's correction on top of the
// held one, without a launch forming their sum.
template <typename T, bool SUM>
__global__ void __launch_bounds__(kBlock) k_interp_cell(const T* __restrict__ crse, const T* __restrict__ held, Lvl Lc,
                                                        T* __restrict__ fine, Lvl Lf, int zero_first) {
    const long cid = xcd_node();
    if (cid >= Lc.N) return;
    int I, J, K;
    node_ijk(cid, Lc, I, J, K);
    const bool hx = I + 1 < Lc.nx, hy = J + 1 < Lc.ny, hz = K + 1 < Lc.nz;
    // offsets of the cell's corners; 0 where the corner does not exist (its
    // value is then read but never used)
    const long ox = hx ? 1 : 0, oy = hy ? (long)Lc.nx : 0, oz = hz ? (long)Lc.nx * Lc.ny : 0;
    const long fsy = Lf.nx, fsz = (long)Lf.nx * Lf.ny;
    const long f0 = (2L * K) * fsz + (2L * J) * fsy + 2L * I;
    T v[3][8], old[3][8];
#pragma unroll
    for (int n = 0; n < 3; ++n) {
        const long q = n * Lc.N + cid;
        auto at = [&](long o) { return SUM ? crse[q + o] + T(1) * held[q + o] : crse[q + o]; };
        const T c000 = at(0), c100 = at(ox), c010 = at(oy), c001 = at(oz);
        const T c011 = at(oy + oz), c101 = at(ox + oz), c110 = at(ox + oy), c111 = at(ox + oy + oz);
        v[n][0] = c000;
        v[n][1] = T(0.5) * (c000 + c100);
        v[n][2] = T(0.5) * (c000 + c010);
        v[n][3] = T(0.25) * (c000 + c100 + c010 + c110);
        v[n][4] = T(0.5) * (c000 + c001);
        v[n][5] = T(0.25) * (c000 + c001 + c100 + c101);
        v[n][6] = T(0.25) * (c000 + c010 + c001 + c011);
        v[n][7] = T(0.125) * (c000 + c100 + c010 + c001 + c011 + c101 + c110 + c111);
        if (!zero_first) {
#pragma unroll
            for (int e = 0; e < 8; ++e) {
                const bool a = e & 1, b = e & 2, c = e & 4;
                const bool ex = (!a || hx) && (!b || hy) && (!c || hz);
                old[n][e] = ex ? fine[n * Lf.N + f0 + (a ? 1 : 0) + (b ? fsy : 0) + (c ? fsz : 0)] : T(0);
            }
        }
    }
#pragma unroll
    for (int n = 0; n < 3; ++n)
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            const bool a = e & 1, b = e & 2, c = e & 4;
            if ((a && !hx) || (b && !hy) || (c && !hz)) continue;
            T* out = fine + n * Lf.N + f0 + (a ? 1 : 0) + (b ? fsy : 0) + (c ? fsz : 0);
            *out = zero_first ? v[n][e] : old[n][e] + v[n][e];
        }
}

template <typename T>
__global__ void __launch_bounds__(kBlock) k_fill(T* __restrict__ p, T v, long n) {
    const long id = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (id < n) p[id] = v;
}

// The first damped Jacobi sweep from a zero guess, elementwise:
// (1-w) 0 + w (b - (A 0 - 0 D)) / D = w b / D  (A 0 is zero on every row,
// boundary rows included), rounded as the full sweep rounds it.
template <typename T>
__global__ void __launch_bounds__(kBlock) k_jacobi_from_zero(const T* __restrict__ b, const T* __restrict__ diag,
                                                             T omega, T* __restrict__ out, long n) {
    const long id = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (id < n) out[id] = omega * b[id] / diag[id];
}

// Red-black Gauss-Seidel half sweep (no counterpart in the reference, whose
// smoother is damped Jacobi, src/Operator/Operator.cpp:88-154): the nodes with
// (i + j + k) % 2 == color take x += w r / D, r = b - A x of the current x,
// with w = w_in on interior rows and w_bd on boundary rows; the other color
// keeps its values.  Colors split the face neighbours of the 19-point
// stencil; edge neighbours share a color and see each other's old values.
template <typename T>
__global__ void __launch_bounds__(kBlock) k_rb_update(T* __restrict__ x, const T* __restrict__ r,
                                                      const T* __restrict__ diag, Lvl L, int color, T w_in,
                                                      T w_bd) {
    const long id = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= L.N) return;
    int i, j, k;
    node_ijk(id, L, i, j, k);
    if (((i + j + k) & 1) != color) return;
    const bool bnd = i == 0 || j == 0 || k == 0 || i == L.nx - 1 || j == L.ny - 1 || k == L.nz - 1;
    const T w = bnd ? w_bd : w_in;
#pragma unroll
    for (int p = 0; p < 3; ++p) {
        const long q = (long)p * L.N + id;
        x[q] += w * r[q] / diag[q];
    }
}

template <typename T>
__global__ void __launch_bounds__(kBlock) k_axpy(T* __restrict__ y, T a, const T* __restrict__ x, long n) {
    const long id = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (id < n) y[id] += a * x[id];
}

// The judged step of a learned correction e (w = A e): partials of r . w
// (slot 0) and w . w (slot 1), one read of r and w.
__global__ void __launch_bounds__(kBlock) k_dot2_partial(const double* __restrict__ r, const double* __restrict__ w,
                                                         long n, double* __restrict__ partial) {
    __shared__ double sh[2][kBlock / 64];
    double rw = 0.0, ww = 0.0;
    for (long q = (long)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (long)gridDim.x * blockDim.x) {
        const double wq = w[q];
        rw += r[q] * wq;
        ww += wq * wq;
    }
    const int lane = (int)(threadIdx.x & 63u), wave = (int)(threadIdx.x >> 6);
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        rw += __shfl_xor(rw, off, 64);
        ww += __shfl_xor(ww, off, 64);
    }
    if (lane == 0) {
        sh[0][wave] = rw;
        sh[1][wave] = ww;
    }
    __syncthreads();
    if (threadIdx.x < 2) {
        double t = 0.0;
#pragma unroll
        for (int v = 0; v < kBlock / 64; ++v) t += sh[threadIdx.x][v];
        partial[(long)threadIdx.x * gridDim.x + blockIdx.x] = t;
    }
}

// w <- r - g w (the residual after the step x += g e) and partials of its
// max-norm (NaN propagated, as k_reduce_partial's op 0).
__global__ void __launch_bounds__(kBlock) k_step_residual(double* __restrict__ w, double g,
                                                          const double* __restrict__ r, long n,
                                                          double* __restrict__ partial) {
    __shared__ double sh[kBlock];
    double m = 0.0;
    for (long q = (long)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (long)gridDim.x * blockDim.x) {
        const double v = (-g) * w[q] + 1.0 * r[q];
        w[q] = v;
        m = (isnan(v) || isnan(m)) ? NAN : fmax(m, fabs(v));
    }
    sh[threadIdx.x] = m;
    __syncthreads();
    for (int st = kBlock / 2; st > 0; st >>= 1) {
        if ((int)threadIdx.x < st) {
            const double a = sh[threadIdx.x], b = sh[threadIdx.x + st];
            sh[threadIdx.x] = (isnan(a) || isnan(b)) ? NAN : fmax(a, b);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = sh[0];
}

// x += g e and r = w, one pass (the step taken).
__global__ void __launch_bounds__(kBlock) k_take_step(double* __restrict__ x, double g, const double* __restrict__ e,
                                                      double* __restrict__ r, const double* __restrict__ w, long n) {
    const long id = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (id < n) {
        x[id] += g * e[id];
        r[id] = w[id];
    }
}

// y = a * x
template <typename T>
__global__ void __launch_bounds__(kBlock) k_scale_copy(const T* __restrict__ x, T a, T* __restrict__ y, long n) {
    const long id = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (id < n) y[id] = a * x[id];
}

template <typename Tin, typename Tout>
__global__ void __launch_bounds__(kBlock) k_cast(const Tin* __restrict__ a, Tout* __restrict__ b, long n) {
    const long id = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (id < n) b[id] = (Tout)a[id];
}

// y = a x (FP64) and yc = y in the cycle's precision, one pass (k_scale_copy
// then k_cast); yc may be x itself.  With fz, also the first Jacobi sweep
// from zero on yc, fz = w yc / D (k_jacobi_from_zero's arithmetic).
// yc may be null (no copy in the cycle's precision).  With nrm2, a is
// 1 / sqrt(*nrm2) (0 when *nrm2 is 0), read on the device:
// the scale of the next Krylov vector, w / |w|, can then be launched before
// the host has read |w|^2.
template <typename Tc, typename TY = double>
__global__ void __launch_bounds__(kBlock) k_scale_copy_cast(const double* x, double a, TY* __restrict__ y, Tc* yc,
                                                            long n, const Tc* __restrict__ diag = nullptr,
                                                            Tc omega = Tc(0), Tc* __restrict__ fz = nullptr,
                                                            const double* __restrict__ nrm2 = nullptr) {
    const long id = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (nrm2) {
        const double hn = sqrt(*nrm2);
        a = hn > 0.0 ? 1.0 / hn : 0.0;
    }
    if (id < n) {
        const double v = a * x[id];
        y[id] = (TY)v;
        const Tc c = (Tc)v;
        if (yc) yc[id] = c;
        if (fz) fz[id] = omega * c / diag[id];
    }
}

// out = (Tout)(a + 1 b): k_axpy(a, 1, b) then k_cast, one pass.
template <typename Tin, typename Tout>
__global__ void __launch_bounds__(kBlock) k_cast_sum(const Tin* __restrict__ a, const Tin* __restrict__ b,
                                                     Tout* __restrict__ out, long n) {
    const long id = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (id < n) out[id] = (Tout)(a[id] + Tin(1) * b[id]);
}

// y += (double) x  (FP64 correction accumulation of an FP32 correction)
__global__ void __launch_bounds__(kBlock) k_add_f32_to_f64(double* __restrict__ y, const float* __restrict__ x,
                                                           long n) {
    const long id = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (id < n) y[id] += (double)x[id];
}

// AoS (BaseFab<Matrix4> order) -> SoA copy of a stiffness field.
__global__ void __launch_bounds__(kBlock) k_aos_to_soa(const double* __restrict__ a, double* __restrict__ s, long N,
                                                       int nc) {
    const long id = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= N) return;
    for (int c = 0; c < nc; ++c) s[c * N + id] = a[id * nc + c];
}

// Two-pass reductions: per-block partials written with vector stores, then
// one block folds them.  op 0: max |v|, op 1: sum a*b, op 2: sum a.
template <typename T>
__global__ void __launch_bounds__(kBlock) k_reduce_partial(const T* __restrict__ a, const T* __restrict__ b, long n,
                                                           int op, double* __restrict__ partial) {
    __shared__ double sh[kBlock];
    double acc = 0.0;
    for (long id = (long)blockIdx.x * blockDim.x + threadIdx.x; id < n; id += (long)gridDim.x * blockDim.x) {
        // max |v| propagates NaN (fmax alone would drop it and hide a blow-up)
        if (op == 0) acc = (isnan((double)a[id]) || isnan(acc)) ? NAN : fmax(acc, fabs((double)a[id]));
        else if (op == 1) acc += (double)a[id] * (double)b[id];
        else acc += (double)a[id];
    }
    sh[threadIdx.x] = acc;
    __syncthreads();
    for (int s = kBlock / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s)
            sh[threadIdx.x] = op == 0 ? ((isnan(sh[threadIdx.x]) || isnan(sh[threadIdx.x + s]))
                                             ? NAN
                                             : fmax(sh[threadIdx.x], sh[threadIdx.x + s]))
                                      : sh[threadIdx.x] + sh[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = sh[0];
}

__global__ void __launch_bounds__(kBlock) k_reduce_final(const double* __restrict__ partial, int nb, int op,
                                                         double* __restrict__ out) {
    __shared__ double sh[kBlock];
    double acc = 0.0;
    for (int q = threadIdx.x; q < nb; q += blockDim.x)
        acc = op == 0 ? ((isnan(partial[q]) || isnan(acc)) ? NAN : fmax(acc, partial[q])) : acc + partial[q];
    sh[threadIdx.x] = acc;
    __syncthreads();
    for (int s = kBlock / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s)
            sh[threadIdx.x] = op == 0 ? ((isnan(sh[threadIdx.x]) || isnan(sh[threadIdx.x + s]))
                                             ? NAN
                                             : fmax(sh[threadIdx.x], sh[threadIdx.x + s]))
                                      : sh[threadIdx.x] + sh[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = sh[0];
}

// BC::Operator::Elastic::Constant::Init (src/BC/Operator/Elastic/Constant.H:197-286).
__global__ void __launch_bounds__(kBlock) k_bc_init(double* __restrict__ rhs, const int* __restrict__ bct,
                                                    const double* __restrict__ bcv, int homogeneous, Lvl L) {
    const long id = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= L.N) return;
    int i, j, k;
    node_ijk(id, L, i, j, k);
    double nrm[3];
    const int reg = region_of<double>(i, j, k, L, nrm);
    if (reg < 0) return;
    for (int d = 0; d < 3; ++d) {
        const int ty = bct[3 * reg + d];
        rhs[d * L.N + id] = (homogeneous && ty == 0) ? 0.0 : bcv[3 * reg + d];
    }
}

// Newton set-up, pass 1 (src/Solver/Nonlocal/Newton.H:168-184): per node
// grad u (one-sided at faces) and DW = C:(grad u - F0).  The pass covers the
// planes [kb, kb + np): node q of that slab (q = id - kb nx ny) writes
// gu9 (when not null) and dw9 at (3p + a) S + q, S = np nx ny.  C slot s of
// node id is read at C[s cs + id cn] (SoA: cs = N, cn = 1; AoS: cs = 1,
// cn = ncomp).
__device__ __forceinline__ void newton_gradu(const double* __restrict__ u, const Lvl& L, long id, int i, int j,
                                             int k, double gu[3][3]) {
    const long sy = L.nx, sz = (long)L.nx * L.ny;
    for (int p = 0; p < 3; ++p) {
        const double* f = u + p * L.N;
        gu[p][0] = d1(f, id, 1, i, L.nx, L.dx, L.per & 1);
        gu[p][1] = d1(f, id, sy, j, L.ny, L.dy, (L.per >> 1) & 1);
        gu[p][2] = d1(f, id, sz, k, L.nz, L.dz, (L.per >> 2) & 1);
    }
}

template <int SYM>
__global__ void __launch_bounds__(kBlock) k_newton_dw(const double* __restrict__ u, const double* __restrict__ C,
                                                      long cs, long cn, const double* __restrict__ F0, Lvl L, int kb,
                                                      int np, double* __restrict__ gu9, double* __restrict__ dw9) {
    constexpr int NC = NComp<SYM>::value;
    const long S = (long)np * L.nx * L.ny;
    const long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= S) return;
    const long id = (long)kb * L.nx * L.ny + q;
    int i, j, k;
    node_ijk(id, L, i, j, k);
    double gu[3][3], A[3][3], s[3][3], c[NC];
    newton_gradu(u, L, id, i, j, k, gu);
    for (int p = 0; p < 3; ++p)
        for (int a = 0; a < 3; ++a) A[p][a] = gu[p][a] - (F0 ? F0[(3 * p + a) * L.N + id] : 0.0);
    for (int t = 0; t < NC; ++t) c[t] = C[t * cs + id * cn];
    contract2<double, SYM>(c, F0 ? A : gu, s);
    for (int p = 0; p < 3; ++p)
        for (int a = 0; a < 3; ++a) {
            if (gu9) gu9[(3 * p + a) * S + q] = gu[p][a];
            dw9[(3 * p + a) * S + q] = s[p][a];
        }
}

// Base::Mechanics post-processing (src/Integrator/Base/Mechanics.H:210-245),
// small-strain models: stress = DW(grad u) = C:(grad u - F0) and strain =
// (grad u + grad u^T) / 2 at every node, grad u central inside and one-sided
// on the domain faces (Numeric::GetStencil on the domain-clipped box).
// sigma and strain: 9 planes, (3p+q).
template <int SYM>
__global__ void __launch_bounds__(kBlock) k_stress(const double* __restrict__ u, const double* __restrict__ C,
                                                   const double* __restrict__ F0, Lvl L, double* __restrict__ sigma,
                                                   double* __restrict__ strain) {
    constexpr int NC = NComp<SYM>::value;
    const long id = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= L.N) return;
    int i, j, k;
    node_ijk(id, L, i, j, k);
    const long sy = L.nx, sz = (long)L.nx * L.ny;
    double gu[3][3], A[3][3], s[3][3], c[NC];
    for (int p = 0; p < 3; ++p) {
        const double* f = u + p * L.N;
        gu[p][0] = d1(f, id, 1, i, L.nx, L.dx, L.per & 1);
        gu[p][1] = d1(f, id, sy, j, L.ny, L.dy, (L.per >> 1) & 1);
        gu[p][2] = d1(f, id, sz, k, L.nz, L.dz, (L.per >> 2) & 1);
    }
    for (int p = 0; p < 3; ++p)
        for (int a = 0; a < 3; ++a) A[p][a] = gu[p][a] - (F0 ? F0[(3 * p + a) * L.N + id] : 0.0);
    for (int q = 0; q < NC; ++q) c[q] = C[q * L.N + id];
    contract2<double, SYM>(c, F0 ? A : gu, s);
    for (int p = 0; p < 3; ++p)
        for (int a = 0; a < 3; ++a) {
            sigma[(3 * p + a) * L.N + id] = s[p][a];
            if (strain) strain[(3 * p + a) * L.N + id] = 0.5 * (gu[p][a] + gu[a][p]);
        }
}

// Base::Mechanics::Integrate, 3D (src/Integrator/Base/Mechanics.H:391-448):
// the traction on a domain face, sum over the face's cells of
// 0.25 (sigma at the 4 corner nodes) . da with da = e_axis h_b h_c, and the
// displacement at the face node whose two in-face indices are n - 2 (the
// reference's disp_hi, last written in its loop order).  The reference
// integrates the x-hi face; any face is allowed here.  One workgroup: out[0..2]
// traction, out[3..5] displacement.
constexpr int kFaceBlock = 1024;
__global__ void __launch_bounds__(kFaceBlock) k_face_traction(const double* __restrict__ sigma,
                                                              const double* __restrict__ u, Lvl L, int axis, int side,
                                                              double* __restrict__ out) {
    __shared__ double red[3][kFaceBlock];
    const int n[3] = {L.nx, L.ny, L.nz};
    const double hh[3] = {L.dx, L.dy, L.dz};
    const int b = axis == 0 ? 1 : 0, c = axis == 2 ? 1 : 2;  // the in-face axes, b < c
    const int ia = side ? n[axis] - 1 : 0;
    const long st[3] = {1, (long)L.nx, (long)L.nx * L.ny};
    const double da = hh[b] * hh[c];
    const long nb = n[b] - 1, ncell = nb * (n[c] - 1);
    double acc[3] = {0.0, 0.0, 0.0};
    for (long q = threadIdx.x; q < ncell; q += kFaceBlock) {
        const long jb = q % nb, jc = q / nb;
        const long id = ia * st[axis] + jb * st[b] + jc * st[c];
        const long corner[4] = {id, id + st[b], id + st[c], id + st[b] + st[c]};
#pragma unroll
        for (int r = 0; r < 3; ++r) {
            const double* sr = sigma + (3 * r + axis) * L.N;
            acc[r] += 0.25 * (sr[corner[0]] + sr[corner[1]] + sr[corner[2]] + sr[corner[3]]) * da;
        }
    }
    for (int r = 0; r < 3; ++r) red[r][threadIdx.x] = acc[r];
    __syncthreads();
    for (int w = kFaceBlock / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w)
            for (int r = 0; r < 3; ++r) red[r][threadIdx.x] += red[r][threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const long jb = n[b] >= 2 ? n[b] - 2 : 0, jc = n[c] >= 2 ? n[c] - 2 : 0;
        const long id = ia * st[axis] + jb * st[b] + jc * st[c];
        for (int r = 0; r < 3; ++r) {
            out[r] = red[r][0];
            out[3 + r] = u[r * L.N + id];
        }
    }
}

// Newton set-up, pass 2 (src/Solver/Nonlocal/Newton.H:216-231) on the rows
// of planes [k0, k1), reading gu9 and dw9 of the slab [kb, kb + np) that pass
// 1 wrote (it must hold planes k0 - 1 .. k1 inside the domain).  With gu9 null
// the boundary rows take grad u from u again, by the same expressions.
__global__ void __launch_bounds__(kBlock) k_newton_rhs(const double* __restrict__ u, const double* __restrict__ b,
                                                       const double* __restrict__ gu9, const double* __restrict__ dw9,
                                                       int kb, int np, const int* __restrict__ bct, Lvl L, int k0,
                                                       int k1, double* __restrict__ rhs) {
    const long P = (long)L.nx * L.ny, S = (long)np * P;
    const long id = (long)k0 * P + (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= (long)k1 * P) return;
    const long q = id - (long)kb * P;  // position in the slab of gu9 and dw9
    int i, j, k;
    node_ijk(id, L, i, j, k);
    double nrm[3];
    const int reg = region_of<double>(i, j, k, L, nrm);
    if (reg >= 0) {
        double uu[3], gu[3][3], s[3][3], f[3];
        if (!gu9) newton_gradu(u, L, id, i, j, k, gu);
        for (int p = 0; p < 3; ++p) {
            uu[p] = u[p * L.N + id];
            for (int a = 0; a < 3; ++a) {
                if (gu9) gu[p][a] = gu9[(3 * p + a) * S + q];
                s[p][a] = dw9[(3 * p + a) * S + q];
            }
        }
        bc_set<double>(bct + 3 * reg, uu, gu, s, nrm, f);
        for (int d = 0; d < 3; ++d) rhs[d * L.N + id] = b[d * L.N + id] - f[d];
    } else {
        // Numeric::Divergence, central (src/Numeric/Stencil.H:588-657); on
        // the faces of a periodic axis the neighbours wrap (a wrap along z
        // needs the slab to hold the whole box)
        long om[3], op[3];
        nb_offsets(L, i, j, k, om, op);
        const double h[3] = {L.dx, L.dy, L.dz};
        double dv[3] = {0.0, 0.0, 0.0};
        for (int a = 0; a < 3; ++a)
            for (int r = 0; r < 3; ++r)
                dv[r] += (dw9[(3 * r + a) * S + q + op[a]] - dw9[(3 * r + a) * S + q + om[a]]) / 2. / h[a];
        for (int d = 0; d < 3; ++d) rhs[d * L.N + id] = b[d * L.N + id] - dv[d];
    }
}

// ---------------------------------------------------------------------------
// Dense coarsest level
// ---------------------------------------------------------------------------
// Column q of the coarsest-level operator: A e_q, for all q at once.  The
// identity I (M x M, column-major) supplies the unit vectors; Acol receives
// A column-major (Acol[q*M + row]).
template <int SYM>
__global__ void __launch_bounds__(kBlock) k_assemble_dense(const double* __restrict__ I, const double* __restrict__ C,
                                                           const int* __restrict__ bct, int uniform, Lvl L,
                                                           double* __restrict__ Acol) {
    const long id = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long M = 3 * L.N;
    if (id >= L.N) return;
    const long q = blockIdx.y;
    int i, j, k;
    node_ijk(id, L, i, j, k);
    double o[3];
    apply_node<double, SYM>(I + q * M, C, bct, uniform, L, i, j, k, id, o);
    for (int c = 0; c < 3; ++c) Acol[q * M + c * L.N + id] = o[c];
}

// Direct bottom solve: x = Ainv rhs with the inverse (column-major, M x M)
// of the operator of the direct bottom level, FP64 for the FP64 cycles and
// rounded to FP32 for the FP32 cycles (TA = T); the sums are in FP64.  Column
// chunks of kDenseMvChunk run in parallel workgroups (partial sums), one more
// launch folds them.
constexpr int kDenseMvChunk = 32;
template <typename T, typename TA>
__global__ void __launch_bounds__(kBlock) k_dense_mv_partial(const TA* __restrict__ A, const T* __restrict__ x,
                                                             int M, double* __restrict__ part) {
    const int r = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (r >= M) return;
    const int c0 = (int)blockIdx.y * kDenseMvChunk, c1 = min(M, c0 + kDenseMvChunk);
    double acc = 0.0;
#pragma unroll 8
    for (int c = c0; c < c1; ++c) acc += (double)A[(long)c * M + r] * (double)x[c];
    part[(long)blockIdx.y * M + r] = acc;
}
template <typename T>
__global__ void __launch_bounds__(kBlock) k_dense_mv_final(const double* __restrict__ part, int nch, int M,
                                                           T* __restrict__ y) {
    const int r = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (r >= M) return;
    double acc = 0.0;
#pragma unroll 8
    for (int q = 0; q < nch; ++q) acc += part[(long)q * M + r];
    y[r] = (T)acc;
}

// Direct bottom solve with the inverse stored row-major (the transpose of
// the column-major inverse): one wave per row, a dot product of contiguous
// FP64 values with x widened to FP64, summed in FP64 and rounded to T.
template <typename T>
__global__ void __launch_bounds__(kBlock) k_dense_mv_rows(const double* __restrict__ At, const T* __restrict__ x,
                                                          int M, T* __restrict__ y) {
    const int r = (int)((blockIdx.x * blockDim.x + threadIdx.x) >> 6), lane = (int)(threadIdx.x & 63);
    if (r >= M) return;  // whole waves
    const double* __restrict__ row = At + (long)r * M;
    double acc0 = 0.0, acc1 = 0.0, acc2 = 0.0, acc3 = 0.0;
    int c = lane;
    for (; c + 192 < M; c += 256) {
        acc0 += row[c] * (double)x[c];
        acc1 += row[c + 64] * (double)x[c + 64];
        acc2 += row[c + 128] * (double)x[c + 128];
        acc3 += row[c + 192] * (double)x[c + 192];
    }
    for (; c < M; c += 64) acc0 += row[c] * (double)x[c];
    double acc = (acc0 + acc1) + (acc2 + acc3);
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off);
    if (lane == 0) y[r] = (T)acc;
}

// Out-of-place transpose of an M x M matrix (column-major in, row-major out:
// the same bytes read as the other order), 64 x 64 tiles through LDS.
__global__ void __launch_bounds__(kBlock) k_transpose(const double* __restrict__ A, int M, double* __restrict__ At) {
    __shared__ double tile[64][65];
    const int r0 = (int)blockIdx.x * 64, c0 = (int)blockIdx.y * 64;
    for (int e = (int)threadIdx.x; e < 64 * 64; e += kBlock) {
        const int c = e >> 6, r = e & 63;
        tile[c][r] = (r0 + r < M && c0 + c < M) ? A[(long)(c0 + c) * M + r0 + r] : 0.0;
    }
    __syncthreads();
    for (int e = (int)threadIdx.x; e < 64 * 64; e += kBlock) {
        const int r = e >> 6, c = e & 63;
        if (r0 + r < M && c0 + c < M) At[(long)(r0 + r) * M + c0 + c] = tile[c][r];
    }
}

// Block-wide max (op 0) or sum (op 1) for kBlock threads; all threads get it.
__device__ __forceinline__ double blk_reduce256(double v, int op, double* red) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        const double o = __shfl_xor(v, off);
        v = op == 0 ? fmax(v, o) : v + o;
    }
    __syncthreads();
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    double r = 0.0;
    for (int q = 0; q < kBlock / 64; ++q) r = op == 0 ? fmax(r, red[q]) : r + red[q];
    return r;
}

// The reference bottom solve on the dense coarsest operator: MLCGSolver
// BiCGStab with the diagonal as left preconditioner, then the smoothing of a
// failed solve (see k_bottom_solve for the stencil form of the same
// algorithm).  Vectors live in LDS (9 M doubles).  One workgroup of kBlock.
__global__ void __launch_bounds__(kBlock)
    k_bottom_bicgstab_dense(double* __restrict__ xg, const double* __restrict__ rhs, const double* __restrict__ Acol,
                            const double* __restrict__ diag, int M, double omega, int maxiter, double eps_rel, int nuf,
                            int* __restrict__ status) {
    extern __shared__ double sm[];
    __shared__ double red[kBlock / 64];
    double *x = sm, *r = sm + M, *rh = sm + 2 * M, *p = sm + 3 * M, *v = sm + 4 * M, *s = sm + 5 * M,
           *t = sm + 6 * M, *b = sm + 7 * M, *d = sm + 8 * M;
    const int tid = threadIdx.x, nt = blockDim.x;
    auto matvec = [&](const double* in, double* out) {
        __syncthreads();
        for (int row = tid; row < M; row += nt) {
            double acc = 0.0;
            for (int c = 0; c < M; ++c) acc += Acol[(long)c * M + row] * in[c];
            out[row] = acc;
        }
        __syncthreads();
    };
    for (int q = tid; q < M; q += nt) { x[q] = 0.0; b[q] = rhs[q]; d[q] = diag[q]; }
    __syncthreads();
    matvec(x, v);
    double loc = 0.0;
    for (int q = tid; q < M; q += nt) {
        const double rq = (b[q] + (-1.0) * v[q]) / d[q];
        r[q] = rq;
        rh[q] = rq;
        loc = fmax(loc, fabs(rq));
    }
    const double rnorm0 = blk_reduce256(loc, 0, red);
    double rnorm = rnorm0, rho_1 = 0.0, alpha = 0.0, omg = 0.0;
    int ret = 0;
    if (rnorm0 != 0.0) {
        for (int iter = 1; iter <= maxiter; ++iter) {
            loc = 0.0;
            for (int q = tid; q < M; q += nt) loc += rh[q] * r[q];
            const double rho = blk_reduce256(loc, 1, red);
            if (rho == 0.0) { ret = 1; break; }
            if (iter == 1) {
                for (int q = tid; q < M; q += nt) p[q] = r[q];
            } else {
                const double beta = (rho / rho_1) * (alpha / omg);
                for (int q = tid; q < M; q += nt) p[q] = r[q] + beta * (p[q] + (-omg) * v[q]);
            }
            matvec(p, v);
            loc = 0.0;
            for (int q = tid; q < M; q += nt) {
                v[q] /= d[q];
                loc += rh[q] * v[q];
            }
            const double rhTv = blk_reduce256(loc, 1, red);
            if (rhTv != 0.0) alpha = rho / rhTv;
            else { ret = 2; break; }
            loc = 0.0;
            for (int q = tid; q < M; q += nt) {
                x[q] += alpha * p[q];
                const double sq = 1.0 * r[q] + (-alpha) * v[q];
                s[q] = sq;
                loc = fmax(loc, fabs(sq));
            }
            rnorm = blk_reduce256(loc, 0, red);
            if (rnorm < eps_rel * rnorm0) break;
            matvec(s, t);
            double l1 = 0.0, l2 = 0.0;
            for (int q = tid; q < M; q += nt) {
                t[q] /= d[q];
                l1 += t[q] * t[q];
                l2 += t[q] * s[q];
            }
            const double tt = blk_reduce256(l1, 1, red);
            const double ts = blk_reduce256(l2, 1, red);
            if (tt != 0.0) omg = ts / tt;
            else { ret = 3; break; }
            loc = 0.0;
            for (int q = tid; q < M; q += nt) {
                x[q] += omg * s[q];
                const double rq = 1.0 * s[q] + (-omg) * t[q];
                r[q] = rq;
                loc = fmax(loc, fabs(rq));
            }
            rnorm = blk_reduce256(loc, 0, red);
            if (rnorm < eps_rel * rnorm0) break;
            if (omg == 0.0) { ret = 4; break; }
            rho_1 = rho;
        }
        if (ret == 0 && rnorm > eps_rel * rnorm0) ret = 8;
        if (!((ret == 0 || ret == 8) && rnorm < rnorm0))
            for (int q = tid; q < M; q += nt) x[q] = 0.0;
        if (ret != 0 && ret != 9)
            for (int q = tid; q < M; q += nt) x[q] = 0.0;
        const int nsm = ret == 0 ? 0 : nuf;
        for (int sw = 0; sw < 2 * nsm; ++sw) {
            matvec(x, v);
            for (int q = tid; q < M; q += nt) {
                const double Rx = v[q] - x[q] * d[q];
                t[q] = (1.0 - omega) * x[q] + omega * (b[q] - Rx) / d[q];
            }
            __syncthreads();
            for (int q = tid; q < M; q += nt) x[q] = t[q];
        }
    }
    __syncthreads();
    for (int q = tid; q < M; q += nt) xg[q] = x[q];
    if (tid == 0) status[0] = ret;
}

// h[i] partials of V_i . w for i < k <= kMultiDotMax (V: k vectors of length
// M, stride M).  Each thread reads w[q] once and keeps the k running sums in
// registers, so one pass streams V and w once; the block then reduces each
// sum through the waves (shuffles) and LDS.
constexpr int kMultiDotMax = 12;
template <typename TV = double>
__global__ void __launch_bounds__(kBlock) k_multidot_partial(const TV* __restrict__ V, int k, long M, long ld,
                                                             const double* __restrict__ w,
                                                             double* __restrict__ partial) {
    __shared__ double sh[kBlock / 64][kMultiDotMax];
    double acc[kMultiDotMax];
#pragma unroll
    for (int i = 0; i < kMultiDotMax; ++i) acc[i] = 0.0;
    // four grid strides per step with independent loads (the accumulation
    // order is the plain loop's), then the remainder
    const long st = (long)gridDim.x * blockDim.x;
    long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
    for (; q + 3 * st < M; q += 4 * st) {
        double wq[4], vq[kMultiDotMax][4];
#pragma unroll
        for (int u = 0; u < 4; ++u) wq[u] = w[q + u * st];
#pragma unroll
        for (int i = 0; i < kMultiDotMax; ++i)
#pragma unroll
            for (int u = 0; u < 4; ++u) vq[i][u] = i < k ? (double)V[(long)i * ld + q + u * st] : 0.0;
#pragma unroll
        for (int i = 0; i < kMultiDotMax; ++i)
#pragma unroll
            for (int u = 0; u < 4; ++u)
                if (i < k) acc[i] += vq[i][u] * wq[u];
    }
    for (; q < M; q += st) {
        const double wq = w[q];
#pragma unroll
        for (int i = 0; i < kMultiDotMax; ++i)
            if (i < k) acc[i] += (double)V[(long)i * ld + q] * wq;
    }
    const int lane = (int)(threadIdx.x & 63u), wave = (int)(threadIdx.x >> 6);
#pragma unroll
    for (int i = 0; i < kMultiDotMax; ++i) {
        if (i < k) {
            double v = acc[i];
#pragma unroll
            for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
            if (lane == 0) sh[wave][i] = v;
        }
    }
    __syncthreads();
    if ((int)threadIdx.x < k) {
        double t = 0.0;
#pragma unroll
        for (int wv = 0; wv < kBlock / 64; ++wv) t += sh[wv][threadIdx.x];
        partial[(long)threadIdx.x * gridDim.x + blockIdx.x] = t;
    }
}

// One pass of w <- w - sum_{i<k} h[i] V_i (k <= kMultiDotMax, h on the
// device) that also forms, from the updated w, the partials of V_i . w
// (slots 0 .. k-1, when DOTS) and of w . w (slot k when DOTS, else slot 0),
// each slot holding gridDim.x block partials: the projection and the dots
// or the norm of classical Gram-Schmidt's next step read V once instead of
// twice.  Grid-stride, four strides of independent loads per step.
template <bool DOTS, typename TV = double>
__global__ void __launch_bounds__(kBlock) k_multiaxpy_dot(double* __restrict__ w, const TV* __restrict__ V, int k,
                                                          long M, long ld, const double* __restrict__ h,
                                                          double* __restrict__ partial) {
    __shared__ double sh[kBlock / 64][kMultiDotMax + 1];
    double hc[kMultiDotMax];
#pragma unroll
    for (int i = 0; i < kMultiDotMax; ++i) hc[i] = i < k ? h[i] : 0.0;
    double acc[kMultiDotMax], nrm = 0.0;
#pragma unroll
    for (int i = 0; i < kMultiDotMax; ++i) acc[i] = 0.0;
    const long st = (long)gridDim.x * blockDim.x;
    long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
    for (; q + 3 * st < M; q += 4 * st) {
        double wq[4], vq[kMultiDotMax][4];
#pragma unroll
        for (int u = 0; u < 4; ++u) wq[u] = w[q + u * st];
#pragma unroll
        for (int i = 0; i < kMultiDotMax; ++i)
#pragma unroll
            for (int u = 0; u < 4; ++u) vq[i][u] = i < k ? (double)V[(long)i * ld + q + u * st] : 0.0;
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            double a = wq[u];
#pragma unroll
            for (int i = 0; i < kMultiDotMax; ++i)
                if (i < k) a += -hc[i] * vq[i][u];
            w[q + u * st] = a;
            if constexpr (DOTS) {
#pragma unroll
                for (int i = 0; i < kMultiDotMax; ++i)
                    if (i < k) acc[i] += vq[i][u] * a;
            }
            nrm += a * a;
        }
    }
    for (; q < M; q += st) {
        double a = w[q];
        double vq[kMultiDotMax];
#pragma unroll
        for (int i = 0; i < kMultiDotMax; ++i) vq[i] = i < k ? (double)V[(long)i * ld + q] : 0.0;
#pragma unroll
        for (int i = 0; i < kMultiDotMax; ++i)
            if (i < k) a += -hc[i] * vq[i];
        w[q] = a;
        if constexpr (DOTS) {
#pragma unroll
            for (int i = 0; i < kMultiDotMax; ++i)
                if (i < k) acc[i] += vq[i] * a;
        }
        nrm += a * a;
    }
    const int lane = (int)(threadIdx.x & 63u), wave = (int)(threadIdx.x >> 6);
    const int nd = DOTS ? k : 0;
#pragma unroll
    for (int i = 0; i <= kMultiDotMax; ++i) {
        if (i <= nd) {
            double v = i < nd ? acc[i < kMultiDotMax ? i : 0] : nrm;
#pragma unroll
            for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
            if (lane == 0) sh[wave][i] = v;
        }
    }
    __syncthreads();
    if ((int)threadIdx.x <= nd) {
        double t = 0.0;
#pragma unroll
        for (int wv = 0; wv < kBlock / 64; ++wv) t += sh[wv][threadIdx.x];
        partial[(long)threadIdx.x * gridDim.x + blockIdx.x] = t;
    }
}

// Classical Gram-Schmidt twice with the second pass delayed by one step
// (DCGS2: Bielich, Langou, Thomas, Swirydowicz, Yamazaki and Boman, "Low-synch
// Gram-Schmidt with delayed reorthogonalization for Krylov solvers", Parallel
// Computing 112, 2022).  Of the n vectors in V, V_0 .. V_{n-2} are
// orthonormal and t = V_{n-1} is the last step's vector after one projection,
// normalized.  Its second projection shares the two passes over V with the
// first projection of the new w:
//   pass 1 (k_multidot2_partial, then k_multidot_final over its 2n slots):
//     Q_i = V_i . w and R_i = V_i . t, i < n;
//   pass 2 (k_dcgs2_update): P_i = R_i (i < n-1), alpha = sqrt(R_{n-1} - |P|^2),
//     c_i = Q_i (i < n-1), c_{n-1} = (Q_{n-1} - P . Q) / alpha, then
//     V_{n-1} <- (t - sum P_i V_i) / alpha and
//     w <- w - sum_{i<n-1} c_i V_i - c_{n-1} V_{n-1}, with |w|^2.
// V is read twice a step where CGS2 reads it three times.  n <= kMultiDotMax.
//
// V may hold FP32 vectors (TV = float): they are widened exactly, and every
// sum is formed in FP64.
//
// Pass 1: slot i < n of the partials holds Q_i, slot n + i holds R_i.
template <typename TV = double>
__global__ void __launch_bounds__(kBlock) k_multidot2_partial(const TV* __restrict__ V, int n, long M, long ld,
                                                              const double* __restrict__ w,
                                                              double* __restrict__ partial) {
    __shared__ double sh[kBlock / 64][2 * kMultiDotMax];
    const TV* __restrict__ t = V + (long)(n - 1) * ld;
    const int k = n - 1;  // vectors before t
    double aw[kMultiDotMax - 1], at[kMultiDotMax - 1], awl = 0.0, atl = 0.0;
#pragma unroll
    for (int i = 0; i < kMultiDotMax - 1; ++i) aw[i] = at[i] = 0.0;
    // two grid strides per step with independent loads (four held 256 VGPRs
    // with an FP64 basis, and ran at half speed with an FP32 one)
    constexpr int kU = 2;
    const long st = (long)gridDim.x * blockDim.x;
    long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
    for (; q + (kU - 1) * st < M; q += kU * st) {
        double wq[kU], tq[kU], vq[kMultiDotMax - 1][kU];
#pragma unroll
        for (int u = 0; u < kU; ++u) {
            wq[u] = w[q + u * st];
            tq[u] = (double)t[q + u * st];
        }
#pragma unroll
        for (int i = 0; i < kMultiDotMax - 1; ++i)
#pragma unroll
            for (int u = 0; u < kU; ++u) vq[i][u] = i < k ? (double)V[(long)i * ld + q + u * st] : 0.0;
#pragma unroll
        for (int i = 0; i < kMultiDotMax - 1; ++i)
#pragma unroll
            for (int u = 0; u < kU; ++u)
                if (i < k) {
                    aw[i] += vq[i][u] * wq[u];
                    at[i] += vq[i][u] * tq[u];
                }
#pragma unroll
        for (int u = 0; u < kU; ++u) {
            awl += tq[u] * wq[u];
            atl += tq[u] * tq[u];
        }
    }
    for (; q < M; q += st) {
        const double wq = w[q], tq = (double)t[q];
#pragma unroll
        for (int i = 0; i < kMultiDotMax - 1; ++i)
            if (i < k) {
                const double v = (double)V[(long)i * ld + q];
                aw[i] += v * wq;
                at[i] += v * tq;
            }
        awl += tq * wq;
        atl += tq * tq;
    }
    const int lane = (int)(threadIdx.x & 63u), wave = (int)(threadIdx.x >> 6);
    auto wave_sum = [](double v) {
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
        return v;
    };
#pragma unroll
    for (int i = 0; i < kMultiDotMax - 1; ++i) {
        if (i < k) {
            const double a = wave_sum(aw[i]), b = wave_sum(at[i]);
            if (lane == 0) {
                sh[wave][i] = a;
                sh[wave][n + i] = b;
            }
        }
    }
    {
        const double a = wave_sum(awl), b = wave_sum(atl);
        if (lane == 0) {
            sh[wave][k] = a;
            sh[wave][n + k] = b;
        }
    }
    __syncthreads();
    if ((int)threadIdx.x < 2 * n) {
        double s = 0.0;
#pragma unroll
        for (int wv = 0; wv < kBlock / 64; ++wv) s += sh[wv][threadIdx.x];
        partial[(long)threadIdx.x * gridDim.x + blockIdx.x] = s;
    }
}

// Pass 2: V_{n-1} and w updated in place, |w|^2 partials in slot 0.  d holds
// the 2n sums of pass 1 (Q, then R); every thread forms the coefficients from
// them, and workgroup 0 stores them for the host: coef[0 .. n-2] = P,
// coef[n-1] = 1/alpha, coef[n .. 2n-1] = c, coef[2n] = alpha (0 when t lies in
// the span of the other vectors).  With FP32 vectors the rewritten V_{n-1}
// is rounded to FP32 and w is projected on the rounded vector, the one the
// later passes read.
template <typename TV = double>
__global__ void __launch_bounds__(kBlock) k_dcgs2_update(TV* __restrict__ V, int n, long M, long ld,
                                                         double* __restrict__ w, const double* __restrict__ d,
                                                         double* __restrict__ coef, double* __restrict__ partial) {
    __shared__ double sh[kBlock / 64];
    TV* __restrict__ t = V + (long)(n - 1) * ld;
    const int k = n - 1;
    double p[kMultiDotMax - 1], c[kMultiDotMax - 1];
    double pp = 0.0, pq = 0.0;
#pragma unroll
    for (int i = 0; i < kMultiDotMax - 1; ++i) {
        p[i] = i < k ? d[n + i] : 0.0;
        c[i] = i < k ? d[i] : 0.0;
        pp += p[i] * p[i];
        pq += p[i] * c[i];
    }
    const double a2 = d[n + k] - pp;
    const double alpha = a2 > 0.0 ? sqrt(a2) : 0.0;
    const double inv = alpha > 0.0 ? 1.0 / alpha : 0.0;
    const double cl = (d[k] - pq) * inv;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        for (int i = 0; i < k; ++i) {
            coef[i] = d[n + i];
            coef[n + i] = d[i];
        }
        coef[k] = inv;
        coef[n + k] = cl;
        coef[2 * n] = alpha;
    }
    double nrm = 0.0;
    // two grid strides per step with independent loads (four held 256 VGPRs)
    constexpr int kU = 2;
    const long st = (long)gridDim.x * blockDim.x;
    long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
    for (; q + (kU - 1) * st < M; q += kU * st) {
        double wq[kU], tq[kU], vq[kMultiDotMax - 1][kU];
#pragma unroll
        for (int u = 0; u < kU; ++u) {
            wq[u] = w[q + u * st];
            tq[u] = (double)t[q + u * st];
        }
#pragma unroll
        for (int i = 0; i < kMultiDotMax - 1; ++i)
#pragma unroll
            for (int u = 0; u < kU; ++u) vq[i][u] = i < k ? (double)V[(long)i * ld + q + u * st] : 0.0;
#pragma unroll
        for (int u = 0; u < kU; ++u) {
            double v = tq[u], a = wq[u];
#pragma unroll
            for (int i = 0; i < kMultiDotMax - 1; ++i)
                if (i < k) {
                    v += -p[i] * vq[i][u];
                    a += -c[i] * vq[i][u];
                }
            const TV vs = (TV)(v * inv);
            a += -cl * (double)vs;
            t[q + u * st] = vs;
            w[q + u * st] = a;
            nrm += a * a;
        }
    }
    for (; q < M; q += st) {
        double v = (double)t[q], a = w[q];
#pragma unroll
        for (int i = 0; i < kMultiDotMax - 1; ++i)
            if (i < k) {
                const double vi = (double)V[(long)i * ld + q];
                v += -p[i] * vi;
                a += -c[i] * vi;
            }
        const TV vs = (TV)(v * inv);
        a += -cl * (double)vs;
        t[q] = vs;
        w[q] = a;
        nrm += a * a;
    }
    const int lane = (int)(threadIdx.x & 63u), wave = (int)(threadIdx.x >> 6);
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) nrm += __shfl_xor(nrm, off, 64);
    if (lane == 0) sh[wave] = nrm;
    __syncthreads();
    if (threadIdx.x == 0) {
        double s = 0.0;
#pragma unroll
        for (int wv = 0; wv < kBlock / 64; ++wv) s += sh[wv];
        partial[blockIdx.x] = s;
    }
}

// h[i] = sum of partials; one block per i.
__global__ void __launch_bounds__(kBlock) k_multidot_final(const double* __restrict__ partial, int nb,
                                                           double* __restrict__ h) {
    __shared__ double sh[kBlock];
    double acc = 0.0;
    for (int q = threadIdx.x; q < nb; q += blockDim.x) acc += partial[(long)blockIdx.x * nb + q];
    sh[threadIdx.x] = acc;
    __syncthreads();
    for (int st = kBlock / 2; st > 0; st >>= 1) {
        if ((int)threadIdx.x < st) sh[threadIdx.x] += sh[threadIdx.x + st];
        __syncthreads();
    }
    if (threadIdx.x == 0) h[blockIdx.x] = sh[0];
}

// w += sign sum_i h[i] V_i  (h on the device); V may be FP32 vectors
// (widened exactly, so the sums are those of their FP64 copies).
template <typename TV>
__global__ void __launch_bounds__(kBlock) k_multiaxpy(double* __restrict__ w, const TV* __restrict__ V, int k, long M,
                                                      long ld, const double* __restrict__ h, double sign) {
    const long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= M) return;
    double acc = w[q];
    for (int i = 0; i < k; ++i) acc += sign * h[i] * (double)V[(long)i * ld + q];
    w[q] = acc;
}

}  // namespace ael
