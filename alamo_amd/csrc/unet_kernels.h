// unet_kernels.h -- device kernels of the learned coarse-level solver (a 3D
// U-Net on a nodal grid), bfloat16 storage, float32 accumulation.
//
// Activations are channels-last: X[node * C + c], C a multiple of 16, node
// index i fastest.  A 3x3x3 convolution is an implicit GEMM
//     Y[v][n] = sum_{t, c} X[v + d_t][c] * W[n][t * Cin + c]       (K = 27 Cin)
// on the gfx950 matrix cores: v_mfma_f32_16x16x32_bf16 takes a 16 x 32 slice
// of A (16 nodes x 32 of K) and a 32 x 16 slice of B (32 of K x 16 output
// channels), each lane holding 8 consecutive K of one row/column, and
// accumulates a 16 x 16 float32 tile (lane l: rows 4(l/16)..+3, column l%16).
// A K-slice of 8 never straddles two taps because Cin is a multiple of 8.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace aunet {

typedef uint16_t bf16_t;
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

struct G3 {
    int nx, ny, nz, N;
};

// float -> bfloat16, round to nearest even (finite inputs).
__device__ __forceinline__ bf16_t f2bf(float f) {
    uint32_t u = __float_as_uint(f);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (bf16_t)(u >> 16);
}
__device__ __forceinline__ float bf2f(bf16_t b) { return __uint_as_float(((uint32_t)b) << 16); }

__device__ __forceinline__ void unpack8(const uint4& q, float f[8]) {
    const uint32_t w[4] = {q.x, q.y, q.z, q.w};
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        f[2 * i] = __uint_as_float(w[i] << 16);
        f[2 * i + 1] = __uint_as_float(w[i] & 0xFFFF0000u);
    }
}
__device__ __forceinline__ uint4 pack8(const float f[8]) {
    uint32_t w[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) w[i] = (uint32_t)f2bf(f[2 * i]) | ((uint32_t)f2bf(f[2 * i + 1]) << 16);
    return make_uint4(w[0], w[1], w[2], w[3]);
}

__device__ __forceinline__ bool tap_node(const G3& g, int x, int y, int z, int t, int& nb) {
    const int xx = x + t % 3 - 1, yy = y + (t / 3) % 3 - 1, zz = z + t / 9 - 1;
    nb = (zz * g.ny + yy) * g.nx + xx;
    return xx >= 0 && xx < g.nx && yy >= 0 && yy < g.ny && zz >= 0 && zz < g.nz;
}

// The K rounds of one wave's 16 nodes x 16*NB output channels, starting at
// k = kfirst and stepping by kstep (a multiple of 32 * KU): KU steps of 32
// along K per round, every A and B load of the round issued before its MFMAs
// (one wait per round instead of one per step; the small levels have few
// waves to hide it), the MFMAs in step order.
constexpr int kConvKU = 4;
template <int NB>
__device__ __forceinline__ void conv_rounds(const bf16_t* __restrict__ X, int Cin, const bf16_t* __restrict__ Wt,
                                            const G3& g, bool vin, int x, int y, int z, int row, int grp, int nbase,
                                            int kfirst, int kstep, f32x4 acc[NB]) {
    constexpr int KU = kConvKU;
    const int K = 27 * Cin;
    for (int k0 = kfirst; k0 < K; k0 += kstep) {
        uint4 a[KU], b[KU][NB];
#pragma unroll
        for (int u = 0; u < KU; ++u) {
            const int k = k0 + 32 * u + grp * 8;
            a[u] = make_uint4(0, 0, 0, 0);
            if (vin && k < K) {
                const int t = k / Cin, c = k - t * Cin;
                int nbr;
                if (tap_node(g, x, y, z, t, nbr)) a[u] = *reinterpret_cast<const uint4*>(X + (size_t)nbr * Cin + c);
            }
#pragma unroll
            for (int nb = 0; nb < NB; ++nb) {
                const int n = nbase + nb * 16 + row;
                b[u][nb] = make_uint4(0, 0, 0, 0);
                if (k < K) b[u][nb] = *reinterpret_cast<const uint4*>(Wt + (size_t)n * K + k);
            }
        }
#pragma unroll
        for (int u = 0; u < KU; ++u) {
            if (k0 + 32 * u >= K) break;
            const bf16x8 A = __builtin_bit_cast(bf16x8, a[u]);
#pragma unroll
            for (int nb = 0; nb < NB; ++nb)
                acc[nb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A, __builtin_bit_cast(bf16x8, b[u][nb]), acc[nb], 0,
                                                                 0, 0);
        }
    }
}

// The epilogue of a 16-node x 16*NB-channel tile: bias, ReLU, and the bf16
// channels-last store, or the float32 component-major one for the output.
template <int NB>
__device__ __forceinline__ void conv_store(const f32x4 acc[NB], const float* __restrict__ bias, int Cout,
                                           const G3& g, int relu, int v0, int row, int grp, int nbase,
                                           bf16_t* __restrict__ Y, float* __restrict__ Yf, int yf_ch) {
#pragma unroll
    for (int nb = 0; nb < NB; ++nb) {
        const int n = nbase + nb * 16 + row;
        // the output layer's bias holds yf_ch entries, its Cout is padded to 16
        if (Yf && n >= yf_ch) continue;
        const float bv = bias ? bias[n] : 0.f;
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const int v = v0 + grp * 4 + e;
            if (v >= g.N) continue;
            float val = acc[nb][e] + bv;
            if (relu) val = fmaxf(val, 0.f);
            if (Yf) {
                Yf[(size_t)n * g.N + v] = val;
            } else {
                Y[(size_t)v * Cout + n] = f2bf(val);
            }
        }
    }
}

__device__ __forceinline__ void conv_node(const G3& g, int vox, bool vin, int& x, int& y, int& z) {
    x = y = z = 0;
    if (vin) {
        x = vox % g.nx;
        const int t = vox / g.nx;
        y = t % g.ny;
        z = t / g.ny;
    }
}

// Y = act(conv3(X, W) + bias).  Workgroup: 4 waves x 16 nodes; blockIdx.y
// selects 16*NB output channels.  Wt[n][27 * Cin] (bf16).  bias may be null.
// Yf != null: write float32 component-major Yf[n * N + v] for n < yf_ch
// instead of the bf16 Y (the network output).
template <int NB>
__global__ void __launch_bounds__(256) k_conv_fwd(const bf16_t* __restrict__ X, int Cin, const bf16_t* __restrict__ Wt,
                                                  const float* __restrict__ bias, int Cout, G3 g, int relu,
                                                  bf16_t* __restrict__ Y, float* __restrict__ Yf, int yf_ch) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int v0 = (blockIdx.x * 4 + wave) * 16;
    if (v0 >= g.N) return;
    const int nbase = blockIdx.y * 16 * NB;
    const int row = lane & 15, grp = lane >> 4;
    const int vox = v0 + row;
    const bool vin = vox < g.N;
    int x, y, z;
    conv_node(g, vox, vin, x, y, z);
    f32x4 acc[NB];
#pragma unroll
    for (int nb = 0; nb < NB; ++nb) acc[nb] = f32x4{0.f, 0.f, 0.f, 0.f};
    conv_rounds<NB>(X, Cin, Wt, g, vin, x, y, z, row, grp, nbase, 0, 32 * kConvKU, acc);
    conv_store<NB>(acc, bias, Cout, g, relu, v0, row, grp, nbase, Y, Yf, yf_ch);
}

// k_conv_fwd with K split over the KS waves of a workgroup, for the levels
// with few nodes (where k_conv_fwd runs a few dozen waves, each through all
// of K = 27 Cin one round after another).  Every wave of the workgroup takes
// the same 16 nodes and 16*NB channels and every KS-th round of K; wave 0
// adds the others' float32 tiles in wave order through LDS and stores.
template <int NB, int KS>
__global__ void __launch_bounds__(64 * KS) k_conv_fwd_ks(const bf16_t* __restrict__ X, int Cin,
                                                         const bf16_t* __restrict__ Wt, const float* __restrict__ bias,
                                                         int Cout, G3 g, int relu, bf16_t* __restrict__ Y,
                                                         float* __restrict__ Yf, int yf_ch) {
    __shared__ f32x4 part[KS - 1][NB][64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int v0 = blockIdx.x * 16;
    const int nbase = blockIdx.y * 16 * NB;
    const int row = lane & 15, grp = lane >> 4;
    const int vox = v0 + row;
    const bool vin = vox < g.N;
    int x, y, z;
    conv_node(g, vox, vin, x, y, z);
    f32x4 acc[NB];
#pragma unroll
    for (int nb = 0; nb < NB; ++nb) acc[nb] = f32x4{0.f, 0.f, 0.f, 0.f};
    conv_rounds<NB>(X, Cin, Wt, g, vin, x, y, z, row, grp, nbase, wave * 32 * kConvKU, 32 * kConvKU * KS, acc);
    if (wave > 0) {
#pragma unroll
        for (int nb = 0; nb < NB; ++nb) part[wave - 1][nb][lane] = acc[nb];
    }
    __syncthreads();
    if (wave > 0) return;
#pragma unroll
    for (int w = 0; w < KS - 1; ++w)
#pragma unroll
        for (int nb = 0; nb < NB; ++nb) acc[nb] += part[w][nb][lane];
    conv_store<NB>(acc, bias, Cout, g, relu, v0, row, grp, nbase, Y, Yf, yf_ch);
}

// Partial weight gradient over one chunk of nodes:
//   part[chunk][co][t][ci] = sum_{v in chunk} dZ[v][co] * X[v + d_t][ci].
// One wave per (co block of 16, ci block of 16, chunk), all 27 taps; the
// reduction over nodes is the MFMA K dimension (32 nodes per step).
__global__ void __launch_bounds__(64) k_conv_wgrad(const bf16_t* __restrict__ X, int Cin, const bf16_t* __restrict__ dZ,
                                                   int Cout, G3 g, int chunk, float* __restrict__ part) {
    const int lane = threadIdx.x;
    const int ncib = Cin / 16;
    const int cib = blockIdx.x % ncib, cob = blockIdx.x / ncib;
    const int ch = blockIdx.y;
    const int vbeg = ch * chunk, vend = min(g.N, vbeg + chunk);
    const int row = lane & 15, grp = lane >> 4;
    const int co = cob * 16 + row, ci = cib * 16 + row;
    f32x4 acc[27];
#pragma unroll
    for (int t = 0; t < 27; ++t) acc[t] = f32x4{0.f, 0.f, 0.f, 0.f};
    for (int vb = vbeg; vb < vend; vb += 32) {
        int xs[8], ys[8], zs[8];
        bool ok[8];
        uint16_t av[8];
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            const int v = vb + grp * 8 + e;
            ok[e] = v < vend;
            const int vv = ok[e] ? v : 0;
            xs[e] = vv % g.nx;
            const int t = vv / g.nx;
            ys[e] = t % g.ny;
            zs[e] = t / g.ny;
            av[e] = ok[e] ? dZ[(size_t)vv * Cout + co] : (uint16_t)0;
        }
        const uint4 aq = make_uint4(av[0] | ((uint32_t)av[1] << 16), av[2] | ((uint32_t)av[3] << 16),
                                    av[4] | ((uint32_t)av[5] << 16), av[6] | ((uint32_t)av[7] << 16));
        const bf16x8 A = __builtin_bit_cast(bf16x8, aq);
#pragma unroll
        for (int t = 0; t < 27; ++t) {
            uint16_t bv[8];
#pragma unroll
            for (int e = 0; e < 8; ++e) {
                int nbr;
                const bool in = ok[e] && tap_node(g, xs[e], ys[e], zs[e], t, nbr);
                bv[e] = in ? X[(size_t)nbr * Cin + ci] : (uint16_t)0;
            }
            const uint4 bq = make_uint4(bv[0] | ((uint32_t)bv[1] << 16), bv[2] | ((uint32_t)bv[3] << 16),
                                        bv[4] | ((uint32_t)bv[5] << 16), bv[6] | ((uint32_t)bv[7] << 16));
            acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A, __builtin_bit_cast(bf16x8, bq), acc[t], 0, 0, 0);
        }
    }
    // acc[t][e] = D[co_local = 4 grp + e][ci_local = row]
#pragma unroll
    for (int t = 0; t < 27; ++t)
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const int o = cob * 16 + grp * 4 + e;
            part[(((size_t)ch * Cout + o) * 27 + t) * Cin + ci] = acc[t][e];
        }
}

// out[i] = sum_c part[c * M + i]
__global__ void __launch_bounds__(256) k_sum_chunks(const float* __restrict__ part, int nchunk, long M,
                                                    float* __restrict__ out) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= M) return;
    float s = 0.f;
    for (int c = 0; c < nchunk; ++c) s += part[(size_t)c * M + i];
    out[i] = s;
}

// db[n] = sum_v dZ[v][n]; one workgroup per channel.
__global__ void __launch_bounds__(256) k_bias_grad(const bf16_t* __restrict__ dZ, int C, int N, float* __restrict__ db) {
    __shared__ float sh[256];
    const int n = blockIdx.x;
    float s = 0.f;
    for (int v = threadIdx.x; v < N; v += 256) s += bf2f(dZ[(size_t)v * C + n]);
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) db[n] = sh[0];
}

// Canonical float32 W[Cout][27][Cin] -> bf16 Wt[Coutp][27 * Cinp] (zero
// padded) and the flipped transpose Wb[Cinp][27 * Coutp],
// Wb[ci][t][co] = W[co][26 - t][ci], used by the input gradient.
__global__ void __launch_bounds__(256) k_pack_weights(const float* __restrict__ W, int Cin, int Cout, int Cinp,
                                                      int Coutp, bf16_t* __restrict__ Wt, bf16_t* __restrict__ Wb) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long M = (long)Coutp * 27 * Cinp;
    if (i >= M) return;
    const int ci = (int)(i % Cinp);
    const int t = (int)((i / Cinp) % 27);
    const int co = (int)(i / ((long)27 * Cinp));
    const float w = (ci < Cin && co < Cout) ? W[((size_t)co * 27 + t) * Cin + ci] : 0.f;
    const bf16_t b = f2bf(w);
    Wt[i] = b;
    Wb[((size_t)ci * 27 + (26 - t)) * Coutp + co] = b;
}

// Padded float32 dW[Coutp][27][Cinp] -> canonical dWc[Cout][27][Cin].
__global__ void __launch_bounds__(256) k_unpack_grad(const float* __restrict__ dWp, int Cin, int Cout, int Cinp,
                                                     float* __restrict__ dWc) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long M = (long)Cout * 27 * Cin;
    if (i >= M) return;
    const int ci = (int)(i % Cin);
    const int t = (int)((i / Cin) % 27);
    const int co = (int)(i / ((long)27 * Cin));
    dWc[i] = dWp[((size_t)co * 27 + t) * Cinp + ci];
}

// Full weighting with scale: Y[I][c] = scale * sum_d w(d) X[2I + d][c],
// w(d) = prod_a (1 - |d_a| / 2), zero outside the fine grid.  8 channels per thread.
__global__ void __launch_bounds__(256) k_down(const bf16_t* __restrict__ X, int C, G3 gf, G3 gc, float scale,
                                              bf16_t* __restrict__ Y) {
    const int c8 = C / 8;
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long)gc.N * c8) return;
    const int q = (int)(i % c8);
    const int I = (int)(i / c8);
    const int X0 = I % gc.nx, Y0 = (I / gc.nx) % gc.ny, Z0 = I / (gc.nx * gc.ny);
    float acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (int t = 0; t < 27; ++t) {
        const int dx = t % 3 - 1, dy = (t / 3) % 3 - 1, dz = t / 9 - 1;
        const int xx = 2 * X0 + dx, yy = 2 * Y0 + dy, zz = 2 * Z0 + dz;
        if (xx < 0 || xx >= gf.nx || yy < 0 || yy >= gf.ny || zz < 0 || zz >= gf.nz) continue;
        const float w = (dx ? 0.5f : 1.f) * (dy ? 0.5f : 1.f) * (dz ? 0.5f : 1.f);
        float f[8];
        unpack8(*reinterpret_cast<const uint4*>(X + ((size_t)(zz * gf.ny + yy) * gf.nx + xx) * C + 8 * q), f);
#pragma unroll
        for (int e = 0; e < 8; ++e) acc[e] += w * f[e];
    }
#pragma unroll
    for (int e = 0; e < 8; ++e) acc[e] *= scale;
    *reinterpret_cast<uint4*>(Y + (size_t)I * C + 8 * q) = pack8(acc);
}

// Trilinear nodal interpolation with scale and optional addend:
// Y[i][c] = scale * (P X)[i][c] + (add ? add[i][c] : 0).
__global__ void __launch_bounds__(256) k_up(const bf16_t* __restrict__ X, int C, G3 gc, G3 gf, float scale,
                                            const bf16_t* __restrict__ add, bf16_t* __restrict__ Y) {
    const int c8 = C / 8;
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long)gf.N * c8) return;
    const int q = (int)(i % c8);
    const int v = (int)(i / c8);
    const int x = v % gf.nx, y = (v / gf.nx) % gf.ny, z = v / (gf.nx * gf.ny);
    const int xs[2] = {x >> 1, (x + 1) >> 1}, ys[2] = {y >> 1, (y + 1) >> 1}, zs[2] = {z >> 1, (z + 1) >> 1};
    const int nx_ = (x & 1) ? 2 : 1, ny_ = (y & 1) ? 2 : 1, nz_ = (z & 1) ? 2 : 1;
    const float w = 1.f / (float)(nx_ * ny_ * nz_);
    float acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (int a = 0; a < nz_; ++a)
        for (int b = 0; b < ny_; ++b)
            for (int c = 0; c < nx_; ++c) {
                float f[8];
                unpack8(*reinterpret_cast<const uint4*>(X + ((size_t)(zs[a] * gc.ny + ys[b]) * gc.nx + xs[c]) * C + 8 * q), f);
#pragma unroll
                for (int e = 0; e < 8; ++e) acc[e] += w * f[e];
            }
    float ad[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    if (add) unpack8(*reinterpret_cast<const uint4*>(add + (size_t)v * C + 8 * q), ad);
#pragma unroll
    for (int e = 0; e < 8; ++e) acc[e] = scale * acc[e] + ad[e];
    *reinterpret_cast<uint4*>(Y + (size_t)v * C + 8 * q) = pack8(acc);
}

// Y[v] = [A[v] (Ca channels), B[v] (Cb channels)]; split is the inverse.
__global__ void __launch_bounds__(256) k_concat(const bf16_t* __restrict__ A, int Ca, const bf16_t* __restrict__ B,
                                                int Cb, int N, bf16_t* __restrict__ Y) {
    const int c8 = (Ca + Cb) / 8;
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long)N * c8) return;
    const int q = (int)(i % c8), v = (int)(i / c8);
    const int c = 8 * q;
    const uint4 val = c < Ca ? *reinterpret_cast<const uint4*>(A + (size_t)v * Ca + c)
                             : *reinterpret_cast<const uint4*>(B + (size_t)v * Cb + (c - Ca));
    *reinterpret_cast<uint4*>(Y + (size_t)v * (Ca + Cb) + c) = val;
}
__global__ void __launch_bounds__(256) k_split(const bf16_t* __restrict__ Y, int Ca, int Cb, int N,
                                               bf16_t* __restrict__ A, bf16_t* __restrict__ B) {
    const int c8 = (Ca + Cb) / 8;
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long)N * c8) return;
    const int q = (int)(i % c8), v = (int)(i / c8);
    const int c = 8 * q;
    const uint4 val = *reinterpret_cast<const uint4*>(Y + (size_t)v * (Ca + Cb) + c);
    if (c < Ca) *reinterpret_cast<uint4*>(A + (size_t)v * Ca + c) = val;
    else *reinterpret_cast<uint4*>(B + (size_t)v * Cb + (c - Ca)) = val;
}

// g = (y > 0) ? g : 0 for the bf16 output y of a ReLU layer, 8 values per thread.
__global__ void __launch_bounds__(256) k_relu_mask(bf16_t* __restrict__ g, const bf16_t* __restrict__ y, long n8) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n8) return;
    uint4 gq = reinterpret_cast<uint4*>(g)[i];
    const uint4 yq = reinterpret_cast<const uint4*>(y)[i];
    uint32_t gw[4] = {gq.x, gq.y, gq.z, gq.w};
    const uint32_t yw[4] = {yq.x, yq.y, yq.z, yq.w};
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const uint32_t lo = yw[k] & 0xFFFFu, hi = yw[k] >> 16;
        const bool plo = (lo & 0x7FFFu) != 0 && !(lo & 0x8000u), phi = (hi & 0x7FFFu) != 0 && !(hi & 0x8000u);
        gw[k] = (plo ? (gw[k] & 0xFFFFu) : 0u) | (phi ? (gw[k] & 0xFFFF0000u) : 0u);
    }
    reinterpret_cast<uint4*>(g)[i] = make_uint4(gw[0], gw[1], gw[2], gw[3]);
}

// Component-major float32 in[c * N + v] (c < n) -> bf16 X[v * Cp + c] (zero padded).
__global__ void __launch_bounds__(256) k_pack_soa(const float* __restrict__ in, int n, int Cp, int N,
                                                  bf16_t* __restrict__ X) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long)N * Cp) return;
    const int c = (int)(i % Cp), v = (int)(i / Cp);
    X[i] = c < n ? f2bf(in[(size_t)c * N + v]) : (bf16_t)0;
}
// bf16 X[v * Cp + c] -> float32 out[c * N + v] for c < n.
__global__ void __launch_bounds__(256) k_unpack_soa(const bf16_t* __restrict__ X, int n, int Cp, int N,
                                                    float* __restrict__ out) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long)N * n) return;
    const int v = (int)(i % N), c = (int)(i / N);
    out[i] = bf2f(X[(size_t)v * Cp + c]);
}

// Mean squared error over n components: loss partials and d = 2 (y - t) / (n N).
__global__ void __launch_bounds__(256) k_mse(const float* __restrict__ y, const float* __restrict__ t, long M,
                                             float inv, float* __restrict__ d, float* __restrict__ part) {
    __shared__ float sh[256];
    float s = 0.f;
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < M; i += (long)gridDim.x * blockDim.x) {
        const float e = y[i] - t[i];
        s += e * e;
        if (d) d[i] = 2.f * e * inv;
    }
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

// Adam on float32 parameters (bias-corrected, no weight decay).
__global__ void __launch_bounds__(256) k_adam(float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m,
                                              float* __restrict__ v, long n, float lr, float b1, float b2, float eps,
                                              float c1, float c2) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

}  // namespace aunet
