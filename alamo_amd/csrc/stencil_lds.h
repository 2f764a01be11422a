// stencil_lds.h -- LDS-staged MajorMinor stencil kernel for the large MG
// levels.
//
// A workgroup owns a TX x TY tile of (i, j) interior columns and marches it
// over a chunk of kc planes in k, one node per thread per plane.  u lives in
// an LDS ring of 3 planes (k-1, k, k+1) with a one-node halo; while plane k
// is computed, each thread holds its share of plane k+2 (PRE values) in
// registers, so the 19-point neighbourhood costs LDS reads and the global
// traffic of u is the halo tile once per plane.  C, the compact grad-C
// matrices D and the OP operands stream from HBM as coalesced rows of TX
// values.  All global loads of a plane are issued together, ahead of the LDS
// work (a scheduling barrier keeps the compiler from sinking each load to
// its use, which would cost one memory round trip per load); the chunk's
// gradient masks and compact D bases are fetched once, in the prologue,
// so that no load of a plane waits on another.  Boundary nodes take nbb
// workgroups of their own, after the tiles in FP64 and before them in FP32
// (see the kernel).  Jacobi sweeps compute an interior node's diagonal from
// its C instead of reading diag.
//
// The kernel reads C from two copies of its own: the interior nodes from a
// tile-blocked copy (TileC: the 21 slots of a tile's plane in one contiguous
// block, whole 128-byte lines per wave and slot), and the boundary nodes from
// a compact array in the order of the boundary workgroups, so that the two
// never share a cache line.
//
// Tiles are XCD-slab ordered: XCD x (workgroups b with b % kXcd == x) owns
// the j-tiles [x*nty/kXcd, (x+1)*nty/kXcd) of every chunk, so the halo rows
// two tiles share are read through one L2.
#pragma once
#include "stencil.h"

namespace ael {

// Tile shape: TX x TY nodes (i x j), one thread per node, whole waves.
template <int TX_, int TY_>
struct Tile {
    static constexpr int TX = TX_, TY = TY_, NT = TX * TY;
    static constexpr int HX = TX + 2, HY = TY + 2, HP = HX * HY;  // halo tile of one component
    static constexpr int PRE = (3 * HP + NT - 1) / NT;            // u values per thread per plane
    static constexpr int RING = PRE * NT;                         // LDS values per ring slot (>= 3 HP)
    static constexpr int WPB = NT / 64;                           // waves per workgroup
    static_assert(NT % 64 == 0 && NT <= 1024, "whole waves");
};

struct TileGrid {
    int ntx, nty, nkc, kc;  // tiles along i and j, k chunks, planes per chunk
};
// Longest chunk: the chunk's gradient masks (3 bits each) travel packed 4 bits
// a plane in one 64-bit register.
constexpr int kMaxChunk = 16;

// Interior copy of C in tile blocks: the 21 slots of the TX x TY nodes of
// one (tile, plane) are contiguous, slot-major, so that the workgroup's 21
// loads of a plane read one block (21 NT values) and each wave's load of a
// slot is 64 consecutive values.  Slot s of the node (tx, ty) of tile
// (it, jt) in plane k is at block(G, k, it, jt) + s NT + ty TX + tx; nodes of
// partial tiles outside the interior hold zeros.
template <class TS>
struct TileC {
    static constexpr int NC = 21;
    __host__ __device__ static long block(const TileGrid& G, int k, int it, int jt) {
        return (((long)(k - 1) * G.nty + jt) * G.ntx + it) * NC * TS::NT;
    }
    static long elems(const Lvl& L, const TileGrid& G) { return (long)(L.nz - 2) * G.nty * G.ntx * NC * TS::NT; }
};
// The tile-blocked copy from the FP64 C (rounded to T).  Grid: one workgroup
// of NT threads per (tile, plane), in block order.
template <class TS, typename T>
__global__ void __launch_bounds__(TS::NT) k_pack_tile_c(const double* __restrict__ C, Lvl L, TileGrid G,
                                                        T* __restrict__ Ct) {
    const long tp = blockIdx.x;  // ((k - 1) * nty + jt) * ntx + it
    const int it = (int)(tp % G.ntx), jt = (int)(tp / G.ntx % G.nty), k = 1 + (int)(tp / ((long)G.ntx * G.nty));
    const int i = 1 + it * TS::TX + (int)threadIdx.x % TS::TX, j = 1 + jt * TS::TY + (int)threadIdx.x / TS::TX;
    const bool v = i <= L.nx - 2 && j <= L.ny - 2;
    const long id = ((long)k * L.ny + j) * L.nx + i;
    T* blk = Ct + TileC<TS>::block(G, k, it, jt) + threadIdx.x;
    for (int s = 0; s < TileC<TS>::NC; ++s) blk[(long)s * TS::NT] = v ? (T)C[s * L.N + id] : T(0);
}
// Palette of a level's C: when the nodes' 21-slot tensors take few distinct
// values (a polycrystal: one per grain; its averaged coarse levels: the grains
// and their boundary mixtures), the tiled kernel reads a two-byte index per
// node, tile-blocked like TileC, and the tensor from a table of
// kPalStride-value rows that stays in the caches.  The table holds the nodes'
// own values, so every product is the same.
using PalIdx = unsigned short;
constexpr int kPalMax = 65536;
constexpr int kPalStride = 24;  // row length: 21 slots, padded to 16-byte multiples in FP64 and FP32
template <class TS>
struct TileIdx {
    __host__ __device__ static long block(const TileGrid& G, int k, int it, int jt) {
        return (((long)(k - 1) * G.nty + jt) * G.ntx + it) * TS::NT;
    }
    static long elems(const Lvl& L, const TileGrid& G) { return (long)(L.nz - 2) * G.nty * G.ntx * TS::NT; }
};
// 64-bit hash of each node's 21 slots (bit patterns).
__global__ void __launch_bounds__(kBlock) k_c_hash(const double* __restrict__ C, long N,
                                                   unsigned long long* __restrict__ hs) {
    const long id = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= N) return;
    unsigned long long h = 0xcbf29ce484222325ull;
    for (int q = 0; q < 21; ++q) {
        h ^= (unsigned long long)__double_as_longlong(C[q * N + id]);
        h *= 0x100000001b3ull;
        h ^= h >> 31;
    }
    hs[id] = h;
}
// Heads of the runs of equal hashes along the node order (a node whose hash
// differs from the previous node's): their (hash, node) pairs, compacted in any
// order; *count receives the number of heads, pairs beyond cap are dropped.
__global__ void __launch_bounds__(kBlock) k_run_heads(const unsigned long long* __restrict__ hs, long N, int cap,
                                                      unsigned long long* __restrict__ heads,
                                                      unsigned int* __restrict__ count) {
    const long id = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= N) return;
    const unsigned long long h = hs[id];
    if (id > 0 && hs[id - 1] == h) return;
    const unsigned int q = atomicAdd(count, 1u);
    if ((int)q < cap) {
        heads[2 * (long)q] = h;
        heads[2 * (long)q + 1] = (unsigned long long)id;
    }
}
// Row q of the table from node rep[q] (rows padded with zeros).
template <typename T>
__global__ void __launch_bounds__(kBlock) k_pal_gather(const double* __restrict__ C, long N, const long* __restrict__ rep,
                                                       int npal, T* __restrict__ pal) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= npal * kPalStride) return;
    const int q = t / kPalStride, s = t % kPalStride;
    pal[t] = s < 21 ? (T)C[s * N + rep[q]] : T(0);
}
// Tile-blocked index of every interior node: the row whose hash matches
// (ph sorted, binary search), checked slot by slot against the node's own
// values; a mismatch (hash collision) sets bad[0].  Grid as k_pack_tile_c.
template <class TS>
__global__ void __launch_bounds__(TS::NT) k_pal_index(const double* __restrict__ C, Lvl L, TileGrid G,
                                                      const unsigned long long* __restrict__ hs,
                                                      const unsigned long long* __restrict__ ph, int npal,
                                                      const double* __restrict__ pal, PalIdx* __restrict__ pidx,
                                                      int* __restrict__ bad) {
    const long tp = blockIdx.x;
    const int it = (int)(tp % G.ntx), jt = (int)(tp / G.ntx % G.nty), k = 1 + (int)(tp / ((long)G.ntx * G.nty));
    const int i = 1 + it * TS::TX + (int)threadIdx.x % TS::TX, j = 1 + jt * TS::TY + (int)threadIdx.x / TS::TX;
    PalIdx q8 = 0;
    if (i <= L.nx - 2 && j <= L.ny - 2) {
        const long id = ((long)k * L.ny + j) * L.nx + i;
        const unsigned long long h = hs[id];
        int lo = 0, hi = npal - 1;
        while (lo < hi) {
            const int mid = (lo + hi) >> 1;
            if (ph[mid] < h) lo = mid + 1;
            else hi = mid;
        }
        bool ok = ph[lo] == h;
        for (int q = 0; q < 21 && ok; ++q)
            ok = __double_as_longlong(C[q * L.N + id]) == __double_as_longlong(pal[lo * kPalStride + q]);
        if (!ok) bad[0] = 1;
        q8 = (PalIdx)lo;
    }
    pidx[TileIdx<TS>::block(G, k, it, jt) + threadIdx.x] = q8;
}
// The palette row of every boundary node, in boundary_node order (the
// boundary workgroups' order), checked like k_pal_index.
__global__ void __launch_bounds__(kBlock) k_pal_index_b(const double* __restrict__ C, Lvl L,
                                                        const unsigned long long* __restrict__ hs,
                                                        const unsigned long long* __restrict__ ph, int npal,
                                                        const double* __restrict__ pal, PalIdx* __restrict__ pidx_b,
                                                        int* __restrict__ bad) {
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= boundary_count(L)) return;
    int i, j, k;
    boundary_node(L, t, i, j, k);
    const long id = ((long)k * L.ny + j) * L.nx + i;
    const unsigned long long h = hs[id];
    int lo = 0, hi = npal - 1;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (ph[mid] < h) lo = mid + 1;
        else hi = mid;
    }
    bool ok = ph[lo] == h;
    for (int q = 0; q < 21 && ok; ++q)
        ok = __double_as_longlong(C[q * L.N + id]) == __double_as_longlong(pal[lo * kPalStride + q]);
    if (!ok) bad[0] = 1;
    pidx_b[t] = (PalIdx)lo;
}
// The 21 slots of palette row pr (16-byte loads).
template <typename T>
__device__ __forceinline__ void pal_row(const T* __restrict__ pr, T cv[21]) {
    if constexpr (sizeof(T) == 8) {
        const double2* p2 = reinterpret_cast<const double2*>(pr);
#pragma unroll
        for (int v = 0; v < 10; ++v) {
            const double2 t = p2[v];
            cv[2 * v] = t.x;
            cv[2 * v + 1] = t.y;
        }
        cv[20] = pr[20];
    } else {
        const float4* p4 = reinterpret_cast<const float4*>(pr);
#pragma unroll
        for (int v = 0; v < 5; ++v) {
            const float4 t = p4[v];
            cv[4 * v] = t.x;
            cv[4 * v + 1] = t.y;
            cv[4 * v + 2] = t.z;
            cv[4 * v + 3] = t.w;
        }
        cv[20] = pr[20];
    }
}

template <typename T>
__global__ void __launch_bounds__(kBlock) k_pack_bnd_c(const double* __restrict__ C, Lvl L, int nc,
                                                       T* __restrict__ Cb) {
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long nb = boundary_count(L);
    if (t >= nb) return;
    int i, j, k;
    boundary_node(L, t, i, j, k);
    const long id = ((long)k * L.ny + j) * L.nx + i;
    for (int s = 0; s < nc; ++s) Cb[s * nb + t] = (T)C[s * L.N + id];
}

// Tile grid of the interior of L: chunks of kc planes, with kc chosen so
// that the launch has about target workgroups (1 to 16 planes per chunk).
template <class TS>
inline TileGrid tile_grid(const Lvl& L, long target) {
    TileGrid G{0, 0, 0, 1};
    const int ni = L.nx - 2, nj = L.ny - 2, nk = L.nz - 2;
    if (ni <= 0 || nj <= 0 || nk <= 0) return G;
    G.ntx = (ni + TS::TX - 1) / TS::TX;
    G.nty = (nj + TS::TY - 1) / TS::TY;
    const long per_plane = (long)G.ntx * G.nty;
    long kc = ((long)nk * per_plane + target / 2) / target;  // nearest
    kc = kc < 1 ? 1 : (kc > kMaxChunk ? kMaxChunk : kc);
    G.kc = (int)kc;
    G.nkc = (nk + G.kc - 1) / G.kc;
    return G;
}
// Interior workgroups: kXcd times those of the XCD with the most j-tiles.
inline unsigned tile_blocks(const TileGrid& G) {
    if (G.ntx == 0) return 0;
    const long per = (long)((G.nty + kXcd - 1) / kXcd) * G.ntx * G.nkc;
    return (unsigned)(per * kXcd);
}

// Tile (it, jt) and chunk kc of interior workgroup ib; false for the
// padding workgroups of XCDs with fewer j-tiles (uniform over the workgroup).
__device__ __forceinline__ bool tile_of(const TileGrid& G, long ib, int& it, int& jt, int& kc) {
    const int x = (int)(ib % kXcd);
    const long lb = ib / kXcd;
    const int t0 = (int)((long)x * G.nty / kXcd), t1 = (int)((long)(x + 1) * G.nty / kXcd);
    const long per = (long)(t1 - t0) * G.ntx;
    if (per == 0 || lb >= per * G.nkc) return false;
    kc = (int)(lb / per);
    const long r = lb - (long)kc * per;
    jt = t0 + (int)(r / G.ntx);
    it = (int)(r - (long)(jt - t0) * G.ntx);
    return true;
}

// Compact D order of a tiled level: (plane k, j-tile, i-tile, wave), the
// same for every chunk length, so that launches that group planes into
// chunks differently (FP64 and FP32) share one D.  Index of the wave-plane:
__host__ __device__ inline long tile_wave_plane(const TileGrid& G, int k, int it, int jt, int wave, int wpb) {
    return (((long)(k - 1) * G.nty + jt) * G.ntx + it) * wpb + wave;
}
// (tile, plane) pairs and wave-planes of a tiled level.
inline long tile_planes(const Lvl& L, const TileGrid& G) { return (long)(L.nz - 2) * G.nty * G.ntx; }
inline long tile_wave_planes(const Lvl& L, const TileGrid& G, int wpb) { return tile_planes(L, G) * wpb; }

// Flagged lanes per wave-plane.  Grid: one workgroup per (tile, plane), in
// wave-plane order.
template <class TS>
__global__ void __launch_bounds__(TS::NT) k_gradc_count_tiles(const unsigned char* __restrict__ gmask, Lvl L,
                                                              TileGrid G, int* __restrict__ wcount) {
    const long tp = blockIdx.x;  // ((k - 1) * nty + jt) * ntx + it
    const int it = (int)(tp % G.ntx), jt = (int)(tp / G.ntx % G.nty), k = 1 + (int)(tp / ((long)G.ntx * G.nty));
    const int i = 1 + it * TS::TX + (int)threadIdx.x % TS::TX, j = 1 + jt * TS::TY + (int)threadIdx.x / TS::TX;
    const bool v = i <= L.nx - 2 && j <= L.ny - 2;
    const long id = ((long)k * L.ny + j) * L.nx + i;
    const unsigned long long bal = __ballot(v && gmask[id] != 0);
    if ((threadIdx.x & 63u) == 0) wcount[tp * TS::WPB + threadIdx.x / 64] = (int)__popcll(bal);
}

template <class TS>
__global__ void __launch_bounds__(TS::NT) k_gradc_fill_tiles(const double* __restrict__ C,
                                                             const unsigned char* __restrict__ gmask, Lvl L,
                                                             TileGrid G, const int* __restrict__ dwbase, int nflag,
                                                             double* __restrict__ Dc) {
    const long tp = blockIdx.x;
    const int it = (int)(tp % G.ntx), jt = (int)(tp / G.ntx % G.nty), k = 1 + (int)(tp / ((long)G.ntx * G.nty));
    const int i = 1 + it * TS::TX + (int)threadIdx.x % TS::TX, j = 1 + jt * TS::TY + (int)threadIdx.x / TS::TX;
    const bool v = i <= L.nx - 2 && j <= L.ny - 2;
    const long id = ((long)k * L.ny + j) * L.nx + i;
    const bool f = v && gmask[id] != 0;
    const unsigned long long bal = __ballot(f);
    if (f) {
        const int p = dwbase[tp * TS::WPB + threadIdx.x / 64] + (int)__popcll(bal & lanes_below());
        double D[18];
        gradc_matrix_mm(C, L, (unsigned)id, D);
#pragma unroll
        for (int q = 0; q < 18; ++q) Dc[(size_t)q * nflag + p] = D[q];
    }
}

// One plane of the u halo tile: this thread's PRE values (in range of the
// box; the others are never read by a node of the tile).
template <class TS, typename T, typename TX>
__device__ __forceinline__ void tile_fetch(const TX* __restrict__ u, const Lvl& L, int it, int jt, int k,
                                           T pre[TS::PRE]) {
    const unsigned N = (unsigned)L.N;
    const unsigned plane = (unsigned)k * (unsigned)(L.nx * L.ny);
#pragma unroll
    for (int m = 0; m < TS::PRE; ++m) {
        // Every lane loads (no register default, so no wait on the previous
        // plane's loads): halo positions past the grid or past 3 HP read a
        // clamped in-grid node, and only nodes outside the interior, whose
        // results are never stored, read those positions.
        const int q = min((int)threadIdx.x + m * TS::NT, 3 * TS::HP - 1);
        const int c = q / TS::HP, pos = q - c * TS::HP;
        const int hy = pos / TS::HX, hx = pos - hy * TS::HX;
        const int gi = min(it * TS::TX + hx, L.nx - 1), gj = min(jt * TS::TY + hy, L.ny - 1);
        pre[m] = (T)ld32(u, (unsigned)c * N + plane + (unsigned)(gj * L.nx + gi));
    }
}
template <class TS, typename T>
__device__ __forceinline__ void tile_store(T* __restrict__ s, const T pre[TS::PRE]) {
#pragma unroll
    for (int m = 0; m < TS::PRE; ++m) {
        s[threadIdx.x + m * TS::NT] = pre[m];  // ring slots hold RING >= 3 HP values
    }
}

// CM: how the interior nodes read C -- 0 the tile-blocked copy Ct, 1 the same
// with the non-temporal hint, 2 the palette (Ct is then the table, pidx the
// tile-blocked row indices).  TX: the type of x, narrower than T only for
// Fapply (an FP32 field in FP64 rows, widened exactly on load).  PER: L has
// periodic axes (their face nodes, in the boundary workgroups, take the
// wrapped interior row from the level's C, Cf).
template <class TS, typename T, int OP, bool DC, int CM, typename TX = T, bool PER = false>
__global__ void __launch_bounds__(TS::NT)
    k_stencil_lds(const TX* __restrict__ x, const T* __restrict__ b, const T* __restrict__ diag,
                  const T* __restrict__ Ct, const T* __restrict__ Cb, const T* __restrict__ Cf,
                  const int* __restrict__ bct, int uniform,
                  Lvl L, T omega, const unsigned char* __restrict__ gmask, unsigned nbb, TileGrid G,
                  const int* __restrict__ dwbase, const T* __restrict__ Dc, int nflag, T* __restrict__ out,
                  const PalIdx* __restrict__ pidx, const PalIdx* __restrict__ pidx_b) {
    static_assert(OP == 0 || std::is_same<T, TX>::value, "a narrower x is for Fapply only");
    __shared__ T su[3][TS::RING];  // [ring slot][component][hy][hx], padded to RING
    // The nbb boundary workgroups (a multiple of kXcd, so that tile ib keeps
    // ib % kXcd == XCD) come after the tile ones in FP64 and before them in
    // FP32.  FP64 runs 2 workgroups per CU and 2 rounds of tiles: the latency-
    // bound boundary workgroups fill the slots the second round frees (Fapply
    // at 129^3: 101.5 us, vs 106.3 first and 110.4 spread among the tiles).
    // FP32 runs 4 per CU: placed last they form a trailing round of their own
    // (65^3 sweep 12.9 us vs 11.9 first; 129^3 residual 63.5 vs 61.8 us).
    constexpr bool kBoundaryLast = sizeof(T) == 8;
    const unsigned ntb = gridDim.x - nbb;  // tile workgroups
    const bool bnd = kBoundaryLast ? blockIdx.x >= ntb : blockIdx.x < nbb;
    if (bnd) {
        const long t = (long)(kBoundaryLast ? blockIdx.x - ntb : blockIdx.x) * blockDim.x + threadIdx.x;
        if (t >= boundary_count(L)) return;
        int i, j, k;
        boundary_node(L, t, i, j, k);
        const long id = ((long)k * L.ny + j) * L.nx + i;
        T nrm[3], o[3];
        const int reg = region_of<T, PER>(i, j, k, L, nrm);
        if constexpr (PER) {
            if (reg >= 0) apply_boundary<T, 1, TX, true>(x, Cb, bct, L, i, j, k, id, reg, nrm, o, boundary_count(L), t);
            else apply_interior_wrap<T, 1, TX>(x, Cf, uniform, L, i, j, k, id, o);  // a periodic face: the level's C
        } else if constexpr (CM == 2) {
            // the node's tensor from the palette (Ct is the table): the values
            // of the compact copy, so the same bits, from a row in the caches
            apply_boundary<T, 1, TX, false>(x, Ct + (size_t)pidx_b[t] * kPalStride, bct, L, i, j, k, id, reg, nrm, o,
                                            1, 0);
        } else {
            apply_boundary<T, 1, TX, false>(x, Cb, bct, L, i, j, k, id, reg, nrm, o, boundary_count(L), t);
        }
        const unsigned n32 = (unsigned)id;
#pragma unroll
        for (int p = 0; p < 3; ++p) {
            const size_t off = (size_t)p * L.N;
            const T xv = (OP == 2) ? (T)ld32(x + off, n32) : T(0);
            const T bv = (OP >= 1) ? ld32(b + off, n32) : T(0);
            const T dv = (OP == 2) ? ld32(diag + off, n32) : T(1);
            st32(out + off, n32, finish_op<T, OP>(o[p], xv, bv, dv, omega));
        }
        return;
    }
    const long ib = (long)(kBoundaryLast ? blockIdx.x : blockIdx.x - nbb);
    int it, jt, kc;
    if (!tile_of(G, ib, it, jt, kc)) return;  // whole workgroup, before any barrier
    constexpr int kHX = TS::HX, kHP = TS::HP;
    const int tx = (int)threadIdx.x % TS::TX, ty = (int)threadIdx.x / TS::TX;
    const int i = 1 + it * TS::TX + tx, j = 1 + jt * TS::TY + ty;
    const bool valid = i <= L.nx - 2 && j <= L.ny - 2;
    const int k0 = 1 + kc * G.kc, k1 = min(k0 + G.kc, L.nz - 1);  // planes [k0, k1)
    const unsigned N = (unsigned)L.N, sxy = (unsigned)(L.nx * L.ny);
    const unsigned col = (unsigned)(j * L.nx + i);
    // Even chunks march down from k1 - 1, odd chunks up from k0: two adjacent
    // chunks then read the planes around their common boundary at the same
    // time (both at the start of their march, or both at its end), and the
    // second read of those planes hits L2.  Step kk centres on plane
    // kfirst + d kk.
    const int d = (kc & 1) ? 1 : -1;
    const int kfirst = d > 0 ? k0 : k1 - 1, nstep = k1 - k0;

    // prologue: planes kfirst - d, kfirst, kfirst + d (behind, centre, ahead)
    // into ring slots 0, 1, 2; the chunk's gradient masks (packed, step kk in
    // bits 4 kk .. 4 kk + 3) and, in lane kk of each wave, the first compact D
    // position of the plane of step kk
    {
        T pre[TS::PRE];
#pragma unroll
        for (int s = 0; s < 3; ++s) {
            tile_fetch<TS, T, TX>(x, L, it, jt, kfirst + d * (s - 1), pre);
            tile_store<TS, T>(su[s], pre);
        }
    }
    // palette: the row index of plane kfirst now, each next plane's one step ahead
    unsigned pi_next = 0;
    if constexpr (CM == 2) pi_next = pidx[TileIdx<TS>::block(G, kfirst, it, jt) + threadIdx.x];
    unsigned long long gpk = 0;
    int dwb_lane = 0;
    if constexpr (DC) {
        if (valid) {
#pragma unroll
            for (int kk = 0; kk < kMaxChunk; ++kk)
                if (kk < nstep)
                    gpk |= (unsigned long long)(gmask[(unsigned)(kfirst + d * kk) * sxy + col] & 0xfu) << (4 * kk);
        }
        const int lane = (int)(threadIdx.x & 63u);
        if (lane < nstep)
            dwb_lane = dwbase[tile_wave_plane(G, kfirst + d * lane, it, jt, (int)(threadIdx.x / 64), TS::WPB)];
    }
    // Every prologue load has landed (dwb_lane included): without this the
    // compiler cannot prove dwb_lane complete inside the loop and waits for
    // all loads in flight before its readlane on every plane.
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();

    const T ihx = T(1) / T(L.dx), ihy = T(1) / T(L.dy), ihz = T(1) / T(L.dz);
    const T hx2 = T(0.5) * ihx, hy2 = T(0.5) * ihy, hz2 = T(0.5) * ihz;
    const T ixx = ihx * ihx, iyy = ihy * ihy, izz = ihz * ihz;
    const T ixy = T(0.25) * ihx * ihy, ixz = T(0.25) * ihx * ihz, iyz = T(0.25) * ihy * ihz;
    const int c0 = (ty + 1) * kHX + (tx + 1);  // centre position in the halo tile

    for (int kk = 0; kk < nstep; ++kk) {
        const int k = kfirst + d * kk;
        const int sb = kk % 3, sc = (kk + 1) % 3, sa = (kk + 2) % 3;  // slots behind, centre, ahead
        const int sm = d > 0 ? sb : sa, sp = d > 0 ? sa : sb;         // slots of planes k-1 and k+1
        const bool more = kk + 1 < nstep;
        const unsigned o = (unsigned)k * sxy + col;
        // every global load of the plane leaves here, before the LDS work
        T cv[21], dv[18], pre[TS::PRE], bv[3], gv[3];
        unsigned pi = 0;
        if constexpr (CM == 2) {
            pi = pi_next;
            if (more) pi_next = pidx[TileIdx<TS>::block(G, k + d, it, jt) + threadIdx.x];
        }
        if (valid) {
            if constexpr (CM == 2) {
                pal_row<T>(Ct + (size_t)pi * kPalStride, cv);
            } else {
                const T* cblk = Ct + TileC<TS>::block(G, k, it, jt);  // this (tile, plane) of the C copy
#pragma unroll
                for (int s = 0; s < 21; ++s)
                    cv[s] = CM == 1 ? ld32nt(cblk + s * TS::NT, threadIdx.x) : ld32(cblk + s * TS::NT, threadIdx.x);
            }
#pragma unroll
            for (int p = 0; p < 3; ++p) bv[p] = (OP >= 1) ? ld32(b + (size_t)p * N, o) : T(0);
        }
        bool flagged = false;
        if constexpr (DC) {
            const unsigned gm = (unsigned)(gpk & 0xfu);
            gpk >>= 4;
            flagged = !uniform && gm != 0u;
            const unsigned long long bal = __ballot(flagged);
            const int base = __builtin_amdgcn_readlane(dwb_lane, kk);
            if (flagged) {
                const unsigned pidx = (unsigned)base + (unsigned)__popcll(bal & lanes_below());
#pragma unroll
                for (int q = 0; q < 18; ++q) dv[q] = ld32(Dc + (size_t)q * nflag, pidx);
            }
        }
        if (more) tile_fetch<TS, T, TX>(x, L, it, jt, k + 2 * d, pre);
        __builtin_amdgcn_sched_barrier(0);

        if (valid) {
            // symmetric gradient pairs e1 and second derivative combinations e (LDS)
            T gu[3][3], B[3][3][3];
#pragma unroll
            for (int p = 0; p < 3; ++p) {
                const T* um = su[sm] + p * kHP;
                const T* uc = su[sc] + p * kHP;
                const T* up = su[sp] + p * kHP;
                const T cc = uc[c0];
                const T xp = uc[c0 + 1], xm = uc[c0 - 1], yp = uc[c0 + kHX], ym = uc[c0 - kHX];
                const T zp = up[c0], zm = um[c0];
                gu[p][0] = (xp - xm) * hx2;
                gu[p][1] = (yp - ym) * hy2;
                gu[p][2] = (zp - zm) * hz2;
                B[p][0][0] = (xp - T(2) * cc + xm) * ixx;
                B[p][1][1] = (yp - T(2) * cc + ym) * iyy;
                B[p][2][2] = (zp - T(2) * cc + zm) * izz;
                B[p][0][1] = (uc[c0 + 1 + kHX] + uc[c0 - 1 - kHX] - uc[c0 + 1 - kHX] - uc[c0 - 1 + kHX]) * ixy;
                B[p][0][2] = (up[c0 + 1] + um[c0 - 1] - um[c0 + 1] - up[c0 - 1]) * ixz;
                B[p][1][2] = (up[c0 + kHX] + um[c0 - kHX] - um[c0 + kHX] - up[c0 - kHX]) * iyz;
                B[p][1][0] = B[p][0][1];
                B[p][2][0] = B[p][0][2];
                B[p][2][1] = B[p][1][2];
            }
            T e[3][6], e1[6];
#pragma unroll
            for (int Q = 0; Q < 6; ++Q) {
                const int qa = pair_a(Q), qb = pair_b(Q);
#pragma unroll
                for (int jj = 0; jj < 3; ++jj) e[jj][Q] = qa == qb ? B[qa][qa][jj] : B[qa][qb][jj] + B[qb][qa][jj];
                e1[Q] = qa == qb ? gu[qa][qa] : gu[qa][qb] + gu[qb][qa];
            }
            // C : grad grad u  (src/Set/Matrix4_MajorMinor.H:709-767)
            T outv[3] = {T(0), T(0), T(0)};
            if constexpr (OP == 2) {
                // Diagonal of an interior row, from the C already in registers
                // instead of a load: (C : Gradient_Diagonal)_p = -2 sum_l C_plpl / h_l^2
                // (src/Operator/Elastic.cpp:321-374, src/Numeric/Stencil.H:900-963).
                const T m[3] = {T(-2) * ixx, T(-2) * iyy, T(-2) * izz};
#pragma unroll
                for (int p = 0; p < 3; ++p) {
                    T d = T(0);
#pragma unroll
                    for (int l = 0; l < 3; ++l) d += cv[tri6(vpair(p, l), vpair(p, l))] * m[l];
                    gv[p] = d;
                }
            } else {
#pragma unroll
                for (int p = 0; p < 3; ++p) gv[p] = T(1);
            }
#pragma unroll
            for (int s = 0; s < 21; ++s) {
                const int P = slot_p(s), Q = slot_q(s);
#pragma unroll
                for (int jj = 0; jj < 3; ++jj) {
                    if (pair_other(P, jj) >= 0) outv[pair_other(P, jj)] += cv[s] * e[jj][Q];
                    if (Q != P && pair_other(Q, jj) >= 0) outv[pair_other(Q, jj)] += cv[s] * e[jj][P];
                }
            }
            if (DC && flagged) {
#pragma unroll
                for (int r = 0; r < 3; ++r)
#pragma unroll
                    for (int Q = 0; Q < 6; ++Q) outv[r] += dv[6 * r + Q] * e1[Q];
            }
#pragma unroll
            for (int p = 0; p < 3; ++p) {
                const T xv = (OP == 2) ? su[sc][p * kHP + c0] : T(0);
                st32(out + (size_t)p * N, o, finish_op<T, OP>(outv[p], xv, bv[p], gv[p], omega));
            }
        }
        if (more) {
            __syncthreads();  // every read of slot sb (the plane behind) is done
            tile_store<TS, T>(su[sb], pre);
            __syncthreads();
        }
    }
}

}  // namespace ael
