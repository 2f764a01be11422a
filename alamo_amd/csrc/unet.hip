// unet.hip -- host side of the learned coarse-level solver: the U-Net of
// alamo_unet.h, its forward pass (activations kept for training), its
// backward pass, the MSE loss and Adam, all as launches of unet_kernels.h on
// the caller's stream.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "../../include/alamo_elastic.h"
#include "../../include/alamo_unet.h"
#include "common.h"
#include "unet_kernels.h"

using aunet::bf16_t;
using aunet::G3;
using alamo_detail::fail;

#define NNCHK(expr)                                                                          \
    do {                                                                                     \
        hipError_t e_ = (expr);                                                              \
        if (e_ != hipSuccess)                                                                \
            return fail(ALAMO_EL_EHIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

namespace {

constexpr int kLevels = 4;
constexpr int kCh[kLevels] = {16, 32, 64, 128};
constexpr int kLossBlocks = 256;
constexpr int kWgradWaves = 2048;  // target waves of one weight-gradient launch

inline unsigned nb256(long n) { return (unsigned)((n + 255) / 256); }
inline int pad16(int c) { return (c + 15) / 16 * 16; }

struct ConvL {
    int cin, cout, cinp, coutp, level;
    bool relu;
    long woff, boff;  // offsets of W and b in the canonical parameter vector
    bf16_t* wt = nullptr;  // [coutp][27 cinp]
    bf16_t* wb = nullptr;  // [cinp][27 coutp], taps flipped
};

}  // namespace

struct alamo_nn_s {
    int n_in = 0, n_out = 0, cin0p = 0, coutp_out = 16;
    G3 g[kLevels];
    std::vector<ConvL> cv;
    long nparams = 0;
    float* p = nullptr;     // parameters, float32 canonical layout
    float* grad = nullptr;  // gradient scratch of train_step
    float* m = nullptr;     // Adam moments
    float* v = nullptr;
    int adam_t = 0;
    bool packed = false;    // bf16 weights up to date with p
    bool have_fwd = false;
    float* dwp = nullptr;   // padded weight gradient of one layer
    float* part = nullptr;  // per-chunk partial weight gradients
    long part_cap = 0;
    float* loss_part = nullptr;
    float* dtmp = nullptr;  // output gradient of train_step [n_out][N0]
    float* otmp = nullptr;  // output of train_step [n_out][N0]
    // activations (bf16, channels last)
    bf16_t* X0 = nullptr;
    bf16_t *A[kLevels] = {}, *E[kLevels] = {}, *D[kLevels] = {};
    bf16_t *U[kLevels - 1] = {}, *K[kLevels - 1] = {}, *B[kLevels - 1] = {}, *G[kLevels - 1] = {};
    // gradients of the activations
    bf16_t *gX0 = nullptr, *gY = nullptr;
    bf16_t *gA[kLevels] = {}, *gE[kLevels] = {}, *gD[kLevels] = {};
    bf16_t *gU[kLevels - 1] = {}, *gK[kLevels - 1] = {}, *gB[kLevels - 1] = {}, *gG[kLevels - 1] = {},
           *gEs[kLevels - 1] = {};
    std::vector<void*> allocs;
};

namespace {

int nalloc(alamo_nn_t nn, void** p, size_t bytes) {
    hipError_t e = hipMalloc(p, bytes < 256 ? 256 : bytes);
    if (e != hipSuccess) return fail(ALAMO_EL_EHIP, std::string("hipMalloc: ") + hipGetErrorString(e));
    nn->allocs.push_back(*p);
    return 0;
}

template <typename T>
int nalloc_t(alamo_nn_t nn, T** p, size_t count) {
    return nalloc(nn, (void**)p, sizeof(T) * count);
}

void launch_conv(const bf16_t* X, int Cin, const bf16_t* W, const float* bias, int Cout, const G3& g, int relu,
                 bf16_t* Y, float* Yf, int yf_ch, hipStream_t s) {
    const dim3 grid((unsigned)((g.N + 63) / 64), 1);
    // Levels whose one-wave-per-16-nodes grid leaves most of the 256 CUs
    // idle split K over 4 waves a workgroup, 16 channels a workgroup
    // (ALAMO_NN_KSPLIT=0 never, 2 on every level; read at every launch).
    const char* ks_env = std::getenv("ALAMO_NN_KSPLIT");
    const int ksplit = ks_env ? std::atoi(ks_env) : 1;
    const long plain_wg = (long)grid.x * (Cout % 64 == 0 ? Cout / 64 : Cout % 32 == 0 ? Cout / 32 : Cout / 16);
    if (ksplit == 2 || (ksplit == 1 && plain_wg < 256)) {
        aunet::k_conv_fwd_ks<1, 4><<<dim3((unsigned)((g.N + 15) / 16), Cout / 16), 256, 0, s>>>(X, Cin, W, bias, Cout, g,
                                                                                                relu, Y, Yf, yf_ch);
        return;
    }
    if (Cout % 64 == 0) {
        aunet::k_conv_fwd<4><<<dim3(grid.x, Cout / 64), 256, 0, s>>>(X, Cin, W, bias, Cout, g, relu, Y, Yf, yf_ch);
    } else if (Cout % 32 == 0) {
        aunet::k_conv_fwd<2><<<dim3(grid.x, Cout / 32), 256, 0, s>>>(X, Cin, W, bias, Cout, g, relu, Y, Yf, yf_ch);
    } else {
        aunet::k_conv_fwd<1><<<dim3(grid.x, Cout / 16), 256, 0, s>>>(X, Cin, W, bias, Cout, g, relu, Y, Yf, yf_ch);
    }
}

// Chunking of the weight-gradient reduction over N nodes for `tiles` waves per chunk.
void wgrad_chunks(int N, int tiles, int& chunk, int& nchunk) {
    const int maxc = (N + 31) / 32;
    nchunk = std::max(1, std::min(maxc, kWgradWaves / std::max(1, tiles)));
    chunk = ((N + nchunk - 1) / nchunk + 31) / 32 * 32;
    nchunk = (N + chunk - 1) / chunk;
}

int pack_weights(alamo_nn_t nn, hipStream_t s) {
    if (nn->packed) return 0;
    for (auto& L : nn->cv) {
        const long M = (long)L.coutp * 27 * L.cinp;
        aunet::k_pack_weights<<<nb256(M), 256, 0, s>>>(nn->p + L.woff, L.cin, L.cout, L.cinp, L.coutp, L.wt, L.wb);
    }
    NNCHK(hipGetLastError());
    nn->packed = true;
    return 0;
}

void conv_fwd(alamo_nn_t nn, int li, const bf16_t* X, bf16_t* Y, float* Yf, hipStream_t s) {
    const ConvL& L = nn->cv[li];
    launch_conv(X, L.cinp, L.wt, nn->p + L.boff, L.coutp, nn->g[L.level], L.relu ? 1 : 0, Y, Yf, nn->n_out, s);
}

// Backward of conv li: dY (bf16, masked in place for ReLU layers), X its
// input, Y its output.  Writes the weight and bias gradients into grad and,
// when dX is given, the input gradient.
void conv_back(alamo_nn_t nn, int li, const bf16_t* X, const bf16_t* Y, bf16_t* dY, bf16_t* dX, float* grad,
               hipStream_t s) {
    const ConvL& L = nn->cv[li];
    const G3& g = nn->g[L.level];
    if (L.relu) aunet::k_relu_mask<<<nb256((long)g.N * L.coutp / 8), 256, 0, s>>>(dY, Y, (long)g.N * L.coutp / 8);
    const int tiles = (L.cinp / 16) * (L.coutp / 16);
    int chunk, nchunk;
    wgrad_chunks(g.N, tiles, chunk, nchunk);
    aunet::k_conv_wgrad<<<dim3(tiles, nchunk), 64, 0, s>>>(X, L.cinp, dY, L.coutp, g, chunk, nn->part);
    const long M = (long)L.coutp * 27 * L.cinp;
    aunet::k_sum_chunks<<<nb256(M), 256, 0, s>>>(nn->part, nchunk, M, nn->dwp);
    const long Mc = (long)L.cout * 27 * L.cin;
    aunet::k_unpack_grad<<<nb256(Mc), 256, 0, s>>>(nn->dwp, L.cin, L.cout, L.cinp, grad + L.woff);
    aunet::k_bias_grad<<<L.cout, 256, 0, s>>>(dY, L.coutp, g.N, grad + L.boff);
    if (dX) launch_conv(dY, L.coutp, L.wb, nullptr, L.cinp, g, 0, dX, nullptr, 0, s);
}

int forward(alamo_nn_t nn, const float* in, float* out, hipStream_t s) {
    int rc = pack_weights(nn, s);
    if (rc) return rc;
    const G3* g = nn->g;
    aunet::k_pack_soa<<<nb256((long)g[0].N * nn->cin0p), 256, 0, s>>>(in, nn->n_in, nn->cin0p, g[0].N, nn->X0);
    const bf16_t* h = nn->X0;
    for (int l = 0; l < kLevels; ++l) {
        conv_fwd(nn, 2 * l, h, nn->A[l], nullptr, s);
        conv_fwd(nn, 2 * l + 1, nn->A[l], nn->E[l], nullptr, s);
        if (l + 1 < kLevels) {
            const long n = (long)g[l + 1].N * kCh[l] / 8;
            aunet::k_down<<<nb256(n), 256, 0, s>>>(nn->E[l], kCh[l], g[l], g[l + 1], 0.125f, nn->D[l + 1]);
            h = nn->D[l + 1];
        }
    }
    const bf16_t* up_in = nn->E[kLevels - 1];
    for (int l = kLevels - 2; l >= 0; --l) {
        const int cu = kCh[l + 1], ce = kCh[l];
        aunet::k_up<<<nb256((long)g[l].N * cu / 8), 256, 0, s>>>(up_in, cu, g[l + 1], g[l], 1.f, nullptr, nn->U[l]);
        aunet::k_concat<<<nb256((long)g[l].N * (cu + ce) / 8), 256, 0, s>>>(nn->U[l], cu, nn->E[l], ce, g[l].N,
                                                                            nn->K[l]);
        const int li = 2 * kLevels + 2 * (kLevels - 2 - l);
        conv_fwd(nn, li, nn->K[l], nn->B[l], nullptr, s);
        conv_fwd(nn, li + 1, nn->B[l], nn->G[l], nullptr, s);
        up_in = nn->G[l];
    }
    conv_fwd(nn, (int)nn->cv.size() - 1, nn->G[0], nullptr, out, s);
    NNCHK(hipGetLastError());
    nn->have_fwd = true;
    return 0;
}

int backward(alamo_nn_t nn, const float* dout, float* grad, float* din, hipStream_t s) {
    if (!nn->have_fwd) return fail(ALAMO_EL_ESTATE, "alamo_nn_backward needs a preceding alamo_nn_forward");
    const G3* g = nn->g;
    const int last = (int)nn->cv.size() - 1;
    aunet::k_pack_soa<<<nb256((long)g[0].N * nn->coutp_out), 256, 0, s>>>(dout, nn->n_out, nn->coutp_out, g[0].N,
                                                                          nn->gY);
    conv_back(nn, last, nn->G[0], nullptr, nn->gY, nn->gG[0], grad, s);
    for (int l = 0; l <= kLevels - 2; ++l) {
        const int li = 2 * kLevels + 2 * (kLevels - 2 - l);
        const int cu = kCh[l + 1], ce = kCh[l];
        conv_back(nn, li + 1, nn->B[l], nn->G[l], nn->gG[l], nn->gB[l], grad, s);
        conv_back(nn, li, nn->K[l], nn->B[l], nn->gB[l], nn->gK[l], grad, s);
        aunet::k_split<<<nb256((long)g[l].N * (cu + ce) / 8), 256, 0, s>>>(nn->gK[l], cu, ce, g[l].N, nn->gU[l],
                                                                           nn->gEs[l]);
        // adjoint of trilinear up: 8 x full weighting
        bf16_t* dst = (l + 1 < kLevels - 1) ? nn->gG[l + 1] : nn->gE[kLevels - 1];
        aunet::k_down<<<nb256((long)g[l + 1].N * cu / 8), 256, 0, s>>>(nn->gU[l], cu, g[l], g[l + 1], 1.f, dst);
    }
    for (int l = kLevels - 1; l >= 0; --l) {
        if (l < kLevels - 1) {
            // adjoint of down (full weighting / 8): up / 8, plus the skip-path gradient
            aunet::k_up<<<nb256((long)g[l].N * kCh[l] / 8), 256, 0, s>>>(nn->gD[l + 1], kCh[l], g[l + 1], g[l],
                                                                         0.125f, nn->gEs[l], nn->gE[l]);
        }
        conv_back(nn, 2 * l + 1, nn->A[l], nn->E[l], nn->gE[l], nn->gA[l], grad, s);
        const bf16_t* Xl = l == 0 ? nn->X0 : nn->D[l];
        bf16_t* dXl = l == 0 ? (din ? nn->gX0 : nullptr) : nn->gD[l];
        conv_back(nn, 2 * l, Xl, nn->A[l], nn->gA[l], dXl, grad, s);
    }
    if (din)
        aunet::k_unpack_soa<<<nb256((long)g[0].N * nn->n_in), 256, 0, s>>>(nn->gX0, nn->n_in, nn->cin0p, g[0].N, din);
    NNCHK(hipGetLastError());
    return 0;
}

int mse(alamo_nn_t nn, const float* out, const float* target, float* dout, double* loss_host, hipStream_t s) {
    const long M = (long)nn->n_out * nn->g[0].N;
    aunet::k_mse<<<kLossBlocks, 256, 0, s>>>(out, target, M, 1.f / (float)M, dout, nn->loss_part);
    NNCHK(hipGetLastError());
    if (loss_host) {
        float h[kLossBlocks];
        NNCHK(hipMemcpyAsync(h, nn->loss_part, sizeof(h), hipMemcpyDeviceToHost, s));
        NNCHK(hipStreamSynchronize(s));
        double t = 0;
        for (float v : h) t += v;
        *loss_host = t / (double)M;
    }
    return 0;
}

int adam(alamo_nn_t nn, const float* grad, float lr, hipStream_t s) {
    nn->adam_t += 1;
    const double b1 = 0.9, b2 = 0.999;
    const float c1 = (float)(1.0 - std::pow(b1, nn->adam_t)), c2 = (float)(1.0 - std::pow(b2, nn->adam_t));
    aunet::k_adam<<<nb256(nn->nparams), 256, 0, s>>>(nn->p, grad, nn->m, nn->v, nn->nparams, lr, (float)b1,
                                                     (float)b2, 1e-8f, c1, c2);
    NNCHK(hipGetLastError());
    nn->packed = false;
    return 0;
}

}  // namespace

extern "C" {

int alamo_nn_create(const int nodes[3], int n_in, int n_out, alamo_nn_t* out) {
    if (!nodes || !out) return fail(ALAMO_EL_EINVAL, "null argument");
    *out = nullptr;
    if (n_in < 1 || n_in > 256 || n_out < 1 || n_out > 16) return fail(ALAMO_EL_EINVAL, "n_in in 1..256, n_out in 1..16");
    for (int a = 0; a < 3; ++a)
        if (nodes[a] < 9 || (nodes[a] - 1) % 8 != 0)
            return fail(ALAMO_EL_EINVAL, "every nodes[a] - 1 must be a positive multiple of 8");
    if ((long)nodes[0] * nodes[1] * nodes[2] > (1L << 27)) return fail(ALAMO_EL_EINVAL, "network grid too large");
    alamo_nn_t nn = new alamo_nn_s();
    nn->n_in = n_in;
    nn->n_out = n_out;
    nn->cin0p = pad16(n_in);
    int n[3] = {nodes[0], nodes[1], nodes[2]};
    for (int l = 0; l < kLevels; ++l) {
        nn->g[l] = G3{n[0], n[1], n[2], n[0] * n[1] * n[2]};
        for (int a = 0; a < 3; ++a) n[a] = (n[a] - 1) / 2 + 1;
    }
    // convolution table (oracle/unet_oracle.py conv_specs)
    auto add = [&](int cin, int cinp, int cout, int coutp, bool relu, int level) {
        ConvL L{cin, cout, cinp, coutp, level, relu, 0, 0};
        L.woff = nn->nparams;
        nn->nparams += (long)cout * 27 * cin;
        L.boff = nn->nparams;
        nn->nparams += cout;
        nn->cv.push_back(L);
    };
    for (int l = 0; l < kLevels; ++l) {
        const int ci = l == 0 ? n_in : kCh[l - 1], cip = l == 0 ? nn->cin0p : kCh[l - 1];
        add(ci, cip, kCh[l], kCh[l], true, l);
        add(kCh[l], kCh[l], kCh[l], kCh[l], true, l);
    }
    for (int l = kLevels - 2; l >= 0; --l) {
        add(kCh[l + 1] + kCh[l], kCh[l + 1] + kCh[l], kCh[l], kCh[l], true, l);
        add(kCh[l], kCh[l], kCh[l], kCh[l], true, l);
    }
    add(kCh[0], kCh[0], n_out, nn->coutp_out, false, 0);

    auto cleanup = [&](int rc) { alamo_nn_destroy(nn); return rc; };
    int rc;
    const long np = nn->nparams;
    if ((rc = nalloc_t(nn, &nn->p, np)) || (rc = nalloc_t(nn, &nn->grad, np)) || (rc = nalloc_t(nn, &nn->m, np)) ||
        (rc = nalloc_t(nn, &nn->v, np)))
        return cleanup(rc);
    NNCHK(hipMemset(nn->p, 0, sizeof(float) * np));
    NNCHK(hipMemset(nn->m, 0, sizeof(float) * np));
    NNCHK(hipMemset(nn->v, 0, sizeof(float) * np));
    long dwmax = 0, partmax = 0;
    for (auto& L : nn->cv) {
        const long M = (long)L.coutp * 27 * L.cinp;
        if ((rc = nalloc_t(nn, &L.wt, M)) || (rc = nalloc_t(nn, &L.wb, M))) return cleanup(rc);
        dwmax = std::max(dwmax, M);
        int chunk, nchunk;
        wgrad_chunks(nn->g[L.level].N, (L.cinp / 16) * (L.coutp / 16), chunk, nchunk);
        partmax = std::max(partmax, M * nchunk);
    }
    if ((rc = nalloc_t(nn, &nn->dwp, dwmax)) || (rc = nalloc_t(nn, &nn->part, partmax)) ||
        (rc = nalloc_t(nn, &nn->loss_part, kLossBlocks)))
        return cleanup(rc);
    nn->part_cap = partmax;
    const G3* g = nn->g;
    const long N0 = g[0].N;
    if ((rc = nalloc_t(nn, &nn->dtmp, (long)n_out * N0)) || (rc = nalloc_t(nn, &nn->otmp, (long)n_out * N0)) ||
        (rc = nalloc_t(nn, &nn->X0, N0 * nn->cin0p)) || (rc = nalloc_t(nn, &nn->gX0, N0 * nn->cin0p)) ||
        (rc = nalloc_t(nn, &nn->gY, N0 * nn->coutp_out)))
        return cleanup(rc);
    for (int l = 0; l < kLevels; ++l) {
        const long N = g[l].N;
        if ((rc = nalloc_t(nn, &nn->A[l], N * kCh[l])) || (rc = nalloc_t(nn, &nn->E[l], N * kCh[l])) ||
            (rc = nalloc_t(nn, &nn->gA[l], N * kCh[l])) || (rc = nalloc_t(nn, &nn->gE[l], N * kCh[l])))
            return cleanup(rc);
        if (l > 0 && ((rc = nalloc_t(nn, &nn->D[l], N * kCh[l - 1])) || (rc = nalloc_t(nn, &nn->gD[l], N * kCh[l - 1]))))
            return cleanup(rc);
        if (l < kLevels - 1) {
            const int cu = kCh[l + 1], ce = kCh[l];
            if ((rc = nalloc_t(nn, &nn->U[l], N * cu)) || (rc = nalloc_t(nn, &nn->K[l], N * (cu + ce))) ||
                (rc = nalloc_t(nn, &nn->B[l], N * ce)) || (rc = nalloc_t(nn, &nn->G[l], N * ce)) ||
                (rc = nalloc_t(nn, &nn->gU[l], N * cu)) || (rc = nalloc_t(nn, &nn->gK[l], N * (cu + ce))) ||
                (rc = nalloc_t(nn, &nn->gB[l], N * ce)) || (rc = nalloc_t(nn, &nn->gG[l], N * ce)) ||
                (rc = nalloc_t(nn, &nn->gEs[l], N * ce)))
                return cleanup(rc);
        }
    }
    *out = nn;
    return 0;
}

int alamo_nn_destroy(alamo_nn_t nn) {
    if (!nn) return 0;
    for (void* p : nn->allocs) (void)hipFree(p);
    delete nn;
    return 0;
}

int alamo_nn_num_params(alamo_nn_t nn, long* n) {
    if (!nn || !n) return fail(ALAMO_EL_EINVAL, "null argument");
    *n = nn->nparams;
    return 0;
}

int alamo_nn_shape(alamo_nn_t nn, int nodes[3], int* n_in, int* n_out) {
    if (!nn) return fail(ALAMO_EL_EINVAL, "null argument");
    if (nodes) {
        nodes[0] = nn->g[0].nx;
        nodes[1] = nn->g[0].ny;
        nodes[2] = nn->g[0].nz;
    }
    if (n_in) *n_in = nn->n_in;
    if (n_out) *n_out = nn->n_out;
    return 0;
}

int alamo_nn_init_params(alamo_nn_t nn, unsigned long long seed, hipStream_t s) {
    if (!nn) return fail(ALAMO_EL_EINVAL, "null argument");
    std::vector<float> h(nn->nparams, 0.f);
    std::mt19937_64 rng(seed);
    std::normal_distribution<double> nd(0.0, 1.0);
    for (auto& L : nn->cv) {
        const double sd = std::sqrt(2.0 / (27.0 * L.cin));
        for (long i = 0; i < (long)L.cout * 27 * L.cin; ++i) h[L.woff + i] = (float)(sd * nd(rng));
    }
    NNCHK(hipMemcpyAsync(nn->p, h.data(), sizeof(float) * nn->nparams, hipMemcpyHostToDevice, s));
    NNCHK(hipStreamSynchronize(s));
    nn->packed = false;
    nn->adam_t = 0;
    NNCHK(hipMemsetAsync(nn->m, 0, sizeof(float) * nn->nparams, s));
    NNCHK(hipMemsetAsync(nn->v, 0, sizeof(float) * nn->nparams, s));
    return 0;
}

int alamo_nn_set_params(alamo_nn_t nn, const float* params, hipStream_t s) {
    if (!nn || !params) return fail(ALAMO_EL_EINVAL, "null argument");
    NNCHK(hipMemcpyAsync(nn->p, params, sizeof(float) * nn->nparams, hipMemcpyDeviceToDevice, s));
    nn->packed = false;
    return 0;
}

int alamo_nn_get_params(alamo_nn_t nn, float* params, hipStream_t s) {
    if (!nn || !params) return fail(ALAMO_EL_EINVAL, "null argument");
    NNCHK(hipMemcpyAsync(params, nn->p, sizeof(float) * nn->nparams, hipMemcpyDeviceToDevice, s));
    return 0;
}

int alamo_nn_forward(alamo_nn_t nn, const float* in, float* out, hipStream_t s) {
    if (!nn || !in || !out) return fail(ALAMO_EL_EINVAL, "null argument");
    return forward(nn, in, out, s);
}

int alamo_nn_backward(alamo_nn_t nn, const float* dout, float* dparams, float* din, hipStream_t s) {
    if (!nn || !dout || !dparams) return fail(ALAMO_EL_EINVAL, "null argument");
    return backward(nn, dout, dparams, din, s);
}

int alamo_nn_tape(alamo_nn_t nn, int conv, float* x, float* y, hipStream_t s) {
    if (!nn || !x) return fail(ALAMO_EL_EINVAL, "null argument");
    const int ncv = (int)nn->cv.size();
    if (conv < 0 || conv >= ncv) return fail(ALAMO_EL_EINVAL, "alamo_nn_tape: no such convolution");
    if (y && conv == ncv - 1) return fail(ALAMO_EL_EINVAL, "alamo_nn_tape: the output layer keeps no bf16 output");
    if (!nn->have_fwd) return fail(ALAMO_EL_ESTATE, "alamo_nn_tape needs a preceding alamo_nn_forward");
    // the buffers forward() reads and writes for each layer
    const bf16_t *in = nullptr, *out = nullptr;
    if (conv < 2 * kLevels) {
        const int l = conv / 2;
        in = (conv & 1) ? nn->A[l] : (l == 0 ? nn->X0 : nn->D[l]);
        out = (conv & 1) ? nn->E[l] : nn->A[l];
    } else if (conv < ncv - 1) {
        const int q = conv - 2 * kLevels, l = kLevels - 2 - q / 2;
        in = (q & 1) ? nn->B[l] : nn->K[l];
        out = (q & 1) ? nn->G[l] : nn->B[l];
    } else {
        in = nn->G[0];
    }
    const ConvL& c = nn->cv[conv];
    const int N = nn->g[c.level].N;
    aunet::k_unpack_soa<<<nb256((long)N * c.cin), 256, 0, s>>>(in, c.cin, c.cinp, N, x);
    if (y) aunet::k_unpack_soa<<<nb256((long)N * c.cout), 256, 0, s>>>(out, c.cout, c.coutp, N, y);
    NNCHK(hipGetLastError());
    return 0;
}

int alamo_nn_mse(alamo_nn_t nn, const float* out, const float* target, float* dout, double* loss_host, hipStream_t s) {
    if (!nn || !out || !target) return fail(ALAMO_EL_EINVAL, "null argument");
    return mse(nn, out, target, dout, loss_host, s);
}

int alamo_nn_adam(alamo_nn_t nn, const float* dparams, float lr, hipStream_t s) {
    if (!nn || !dparams || !(lr > 0.f)) return fail(ALAMO_EL_EINVAL, "bad argument");
    return adam(nn, dparams, lr, s);
}

int alamo_nn_train_step(alamo_nn_t nn, const float* in, const float* target, float lr, double* loss_host,
                        hipStream_t s) {
    if (!nn || !in || !target || !(lr > 0.f)) return fail(ALAMO_EL_EINVAL, "bad argument");
    int rc;
    if ((rc = forward(nn, in, nn->otmp, s))) return rc;
    if ((rc = mse(nn, nn->otmp, target, nn->dtmp, loss_host, s))) return rc;
    if ((rc = backward(nn, nn->dtmp, nn->grad, nullptr, s))) return rc;
    return adam(nn, nn->grad, lr, s);
}

}  // extern "C"
