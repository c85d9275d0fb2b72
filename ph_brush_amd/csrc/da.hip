// da.hip -- learned surrogate for the move's Delta-E and the delayed-acceptance
// chain.  No counterpart in the reference (it has no learned part); the chain
// keeps the proposal rule and draw order of the trial loop of
// phase_gc_mu_control (MC/src/mc.cu:379-503), and the exact Delta-E decides
// every accepted move.
//
// Descriptor of a move (site i, ion j): d = h_i - h_j, 48 shell counts per
// centre (12 shells of r_desc/12 x 4 species), each centre excluding the other
// one, so a move and its reverse see the same d.  Per-atom histograms are kept
// in HBM (uint16 x 48) and updated incrementally from the accepted atoms'
// r_desc rows.  f is the MLP 48 -> 128 -> 128 -> 1 (GELU; BF16 weights and
// activations on MFMA, FP32 accumulation); the surrogate is dE* = B + s f(d)
// with B the physics terms below (s = +1 forward).  Second stage (Christen &
// Fox 2005), with r1 the clamped stage-1 exponent of the reverse move:
//   e2 = (-beta (dE - s mu) + logH) + min(0, r1) - min(0, e1).
//
// The surrogate adds to f the physics it can get exactly and cheaply: the
// reciprocal-space potential at the two atoms from the maintained potential
// grid, the mesh self terms, a pair term and the O(1) corrections (k_da_base).
// That part does not flip sign under the reverse move, so stage 2 uses the
// general rule with the reverse move's stage-1 probability, computed from the
// same quantities the exact Delta-E yields.
//
// Chain round (one workgroup of 16 waves): wave 0 draws up to 64 proposals
// assuming each is rejected at stage 1; all waves build their descriptors and
// physics terms and run the MLP on MFMA; lane k takes candidate k's stage-1
// decision, a ballot finds the first stage-1 acceptance; only that candidate
// gets the exact Delta-E (all 16 waves), its stage-2 decision, and on
// acceptance the state, pool and histogram updates.  The engine's stream
// position then rolls back to just after that candidate's own draws, as in the
// plain chain.  A launch ends at its accepted move; the refresh that follows
// folds it into the potential grid.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <limits>
#include <memory>
#include <numeric>
#include <random>
#include <vector>

#include "mc_host.hpp"
#include "mc_kernels.hpp"

namespace phb {

constexpr int NN_IN = 48, NN_H = 128;
constexpr int OFF_W1 = 0, OFF_B1 = NN_IN * NN_H, OFF_W2 = OFF_B1 + NN_H, OFF_B2 = OFF_W2 + NN_H * NN_H;
constexpr int OFF_W3 = OFF_B2 + NN_H, OFF_B3 = OFF_W3 + NN_H;
constexpr int NN_TRAIN = OFF_B3 + 1;                      // 22,913
constexpr int OFF_SCALE = NN_TRAIN, OFF_MEAN = OFF_SCALE + NN_IN, OFF_STD = OFF_MEAN + 1;
constexpr int NN_PARAM = OFF_STD + 1;                     // 22,963
constexpr int NTP = NN_TRAIN + 1;                         // partial-gradient record: gradient + loss
constexpr int DA_M = 64;                                  // candidates per round (max)
constexpr int DA_THREADS = 512;                           // chain workgroup: 8 waves, 256 VGPRs per lane
constexpr int DA_WAVES = DA_THREADS / 64;
constexpr int DA_FEAT = (DA_M * 48 + DA_THREADS - 1) / DA_THREADS;   // descriptor features per thread
constexpr int XP = NN_IN + 8;                             // LDS row pitch (bf16) of input tiles
constexpr int HP = NN_H + 8;                              // LDS row pitch (bf16) of hidden tiles
constexpr int DA_SHIFT = 24;                              // r_desc row entry: atom | shell << 24

static_assert(PHB_NN_TRAIN == NN_TRAIN && PHB_NN_PARAMS == NN_PARAM, "parameter layout");

typedef short s4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ unsigned short f2bf(float f) {      // round to nearest even
    unsigned u = __float_as_uint(f);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (unsigned short)(u >> 16);
}
__device__ __forceinline__ float bf2f(unsigned short h) { return __uint_as_float((unsigned)h << 16); }
// erf by the rational approximation of Abramowitz & Stegun 7.1.26 (Hastings):
// |error| <= 1.5e-7 plus the hardware reciprocal's and exp2's last-bit errors,
// far below the BF16 rounding of the activations; branch-free, unlike erff.
__device__ __forceinline__ float erf_as(float x) {
    const float ax = fabsf(x);
    const float t = __builtin_amdgcn_rcpf(1.0f + 0.3275911f * ax);
    float p = 1.061405429f;
    p = p * t - 1.453152027f;
    p = p * t + 1.421413741f;
    p = p * t - 0.284496736f;
    p = p * t + 0.254829592f;
    const float r = 1.0f - p * t * __expf(-ax * ax);
    return copysignf(r, x);
}
__device__ __forceinline__ float gelu_f(float z) { return 0.5f * z * (1.0f + erf_as(z * 0.70710678118654752f)); }
__device__ __forceinline__ float gelu_grad_f(float z) {
    return 0.5f * (1.0f + erf_as(z * 0.70710678118654752f)) + z * __expf(-0.5f * z * z) * 0.39894228040143268f;
}
__device__ __forceinline__ s4 ld4(const unsigned short* p) { return *reinterpret_cast<const s4*>(p); }
__device__ __forceinline__ f4 mfma16(s4 a, s4 b, f4 c) { return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0); }
// a 16x16x16 tile: A[i][k] in lane i + 16 (k/4), element k%4; B[k][j] in lane j + 16 (k/4);
// D[i][j] in lane j + 16 (i/4), element i%4.
typedef __bf16 bf8 __attribute__((ext_vector_type(8)));
__device__ __forceinline__ bf8 ld8(const unsigned short* p) { return *reinterpret_cast<const bf8*>(p); }
// 16x16x32: A[i][k] in lane i + 16 (k/8), element k%8; B[k][j] in lane j + 16 (k/8); D as mfma16.
__device__ __forceinline__ f4 mfma32(bf8 a, bf8 b, f4 c) { return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0); }

struct NetDev {
    const unsigned short *w1t;   // [128 out][48 in]   bf16 (B operand of the forward pass)
    const unsigned short *w2t;   // [128 out][128 in]  bf16
    const unsigned short *w2;    // [128 in][128 out]  bf16 (B operand of dh1 = dz2 W2^T)
    const unsigned short *w3;    // [128]              bf16
    const float* theta;          // FP32 master copy: biases, in_scale, out_mean, out_std
};

// ---------------------------------------------------------------- 4-wave tile passes (64 rows)
// Layer 1 for rows r0..r0+15 of X: z[ct] holds D-layout tiles (rows r0 + 4g + e, column 16 ct + rr).
// Three 16x16x16 steps over the 48 inputs (a 16x16x32 step feeding a 16x16x16 step's
// accumulator gave wrong values in two of the four accumulator registers, so the two
// shapes are not chained); the delayed-acceptance chain issues the same steps.
__device__ __forceinline__ void layer1(const unsigned short* X, int r0, const NetDev& n, f4 (&z)[8], int lane) {
    const int rr = lane & 15, g = lane >> 4;
    s4 a[3];
#pragma unroll
    for (int kb = 0; kb < 3; ++kb) a[kb] = ld4(X + (r0 + rr) * XP + kb * 16 + 4 * g);
#pragma unroll
    for (int ct = 0; ct < 8; ++ct) {
        f4 acc = {0.f, 0.f, 0.f, 0.f};
        const unsigned short* wr = n.w1t + (ct * 16 + rr) * NN_IN + 4 * g;
#pragma unroll
        for (int kb = 0; kb < 3; ++kb) acc = mfma16(a[kb], ld4(wr + kb * 16), acc);
        const float b = n.theta[OFF_B1 + ct * 16 + rr];
        z[ct] = acc + b;
    }
}

// Layer 2: four 16x16x32 steps over the 128 hidden units.
__device__ __forceinline__ void layer2(const unsigned short* H1, int r0, const NetDev& n, f4 (&z)[8], int lane) {
    const int rr = lane & 15, g = lane >> 4;
    bf8 a[4];
#pragma unroll
    for (int kb = 0; kb < 4; ++kb) a[kb] = ld8(H1 + (r0 + rr) * HP + 32 * kb + 8 * g);
#pragma unroll
    for (int ct = 0; ct < 8; ++ct) {
        f4 acc = {0.f, 0.f, 0.f, 0.f};
        const unsigned short* wr = n.w2t + (ct * 16 + rr) * NN_H + 8 * g;
#pragma unroll
        for (int kb = 0; kb < 4; ++kb) acc = mfma32(a[kb], ld8(wr + 32 * kb), acc);
        const float b = n.theta[OFF_B2 + ct * 16 + rr];
        z[ct] = acc + b;
    }
}

__device__ __forceinline__ float xor_sum16(float v) {   // sum over the 16 lanes sharing lane >> 4
    v += __shfl_xor(v, 1, 64);
    v += __shfl_xor(v, 2, 64);
    v += __shfl_xor(v, 4, 64);
    v += __shfl_xor(v, 8, 64);
    return v;
}
// Sum over the 16 lanes of a DPP row (lane >> 4) by four row-local DPP adds;
// the association differs between the row's quads, so the sum may differ in
// the last bit between lanes: for consumers that read one lane.
template <int CTRL>
__device__ __forceinline__ float dpp_f(float v) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, 0xF, 0xF, false));
}
__device__ __forceinline__ float row_sum16(float v) {
    v += dpp_f<0xB1>(v);    // quad_perm [1,0,3,2]
    v += dpp_f<0x4E>(v);    // quad_perm [2,3,0,1]
    v += dpp_f<0x124>(v);   // row_ror:4
    v += dpp_f<0x128>(v);   // row_ror:8
    return v;
}
__device__ __forceinline__ float xor_sum_groups(float v) {  // sum over the 4 lanes sharing lane & 15
    v += __shfl_xor(v, 16, 64);
    v += __shfl_xor(v, 32, 64);
    return v;
}

// Loads 64 descriptor rows (dataset indices from idx, or rows first.. when idx is null) into X,
// scaled by in_scale and rounded to BF16: four threads per row, each with three 8-byte loads
// of 12 features issued together (a row is 96 bytes, 8-byte aligned).
__device__ __forceinline__ void load_rows(const short* __restrict__ dd, const int* __restrict__ idx, long long first,
                                          int nrows, const float* __restrict__ theta, unsigned short* X, int tid,
                                          int nthr) {
    typedef short sh4 __attribute__((ext_vector_type(4)));
    for (int u = tid; u < 64 * 4; u += nthr) {
        const int r = u >> 2, c0 = 12 * (u & 3);
        sh4 v[3] = {{0, 0, 0, 0}, {0, 0, 0, 0}, {0, 0, 0, 0}};
        if (r < nrows) {
            const long long s = idx ? (long long)idx[r] : first + r;
            const sh4* src = reinterpret_cast<const sh4*>(dd + s * NN_IN + c0);
            v[0] = src[0]; v[1] = src[1]; v[2] = src[2];
        }
#pragma unroll
        for (int k = 0; k < 12; ++k) X[r * XP + c0 + k] = f2bf((float)v[k >> 2][k & 3] * theta[OFF_SCALE + c0 + k]);
    }
}

// Forward only: y = out_mean + out_std * net(d), 64 rows per workgroup of 256.
__global__ __launch_bounds__(256) void k_mlp_forward(NetDev n, const short* __restrict__ dd, long long nrows,
                                                     float* __restrict__ y) {
    __shared__ __align__(16) unsigned short X[64 * XP];
    __shared__ __align__(16) unsigned short H1[64 * HP];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, rr = lane & 15, g = lane >> 4;
    const long long first = (long long)blockIdx.x * 64;
    const int nr = (int)std::min<long long>(64, nrows - first);
    load_rows(dd, nullptr, first, nr, n.theta, X, tid, 256);
    __syncthreads();
    const int r0 = wave * 16;
    f4 z[8];
    layer1(X, r0, n, z, lane);
#pragma unroll
    for (int ct = 0; ct < 8; ++ct)
#pragma unroll
        for (int e = 0; e < 4; ++e) H1[(r0 + 4 * g + e) * HP + ct * 16 + rr] = f2bf(gelu_f(z[ct][e]));
    __syncthreads();
    layer2(H1, r0, n, z, lane);
    float yp[4] = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int ct = 0; ct < 8; ++ct) {
        const float w = bf2f(n.w3[ct * 16 + rr]);
#pragma unroll
        for (int e = 0; e < 4; ++e) yp[e] += bf2f(f2bf(gelu_f(z[ct][e]))) * w;
    }
    const float b3 = n.theta[OFF_B3], om = n.theta[OFF_MEAN], os = n.theta[OFF_STD];
#pragma unroll
    for (int e = 0; e < 4; ++e) {
        const float v = row_sum16(yp[e]);
        const int r = r0 + 4 * g + e;
        if (rr == 0 && r < nr) y[first + r] = om + os * (v + b3);
    }
}

// Training step of the loss sum_rows (y - t_hat)^2 / B, in three launches:
//   k_mlp_act   forward and backward activations, 64 rows per workgroup (4 waves, wave w owns
//               rows 16w..16w+15), stored transposed in BF16 (row index fastest) for the weight
//               products, and the workgroup's sums for the vector parameters (db1, db2, dw3, db3)
//               and the loss;
//   k_mlp_wgrad dW2 = H1^T DZ2 and dW1 = X^T DZ1 over the whole batch, one workgroup per 16x16
//               output tile (16x16x32 BF16 MFMA along the rows, the four waves' sums added in a
//               fixed order), plus one workgroup adding the vector sums in workgroup order;
//   k_mlp_adam  Adam on the finished gradient.
// Every sum has a fixed order: the gradient does not depend on scheduling.
constexpr int VP = 3 * NN_H + 2;   // vector-parameter record per workgroup: db1, db2, dw3, db3, loss


// LDS copies of the weights in the training step (pitches in BF16 elements; the padding spreads
// the 16 rows an MFMA operand reads over the LDS banks)
constexpr int W1P = NN_IN + 8, W2P = NN_H + 8;
constexpr size_t ACT_WLDS = (size_t)(NN_H * W1P + 2 * NN_H * W2P) * sizeof(unsigned short);

struct TrainBufs {
    unsigned short *xt, *h1t, *dz1t, *dz2t;   // [48][ld], [128][ld] x 3: column-major activations
    float* vpart;                             // [nparts][VP]
    int ld;                                   // rows of the batch
};

__global__ __launch_bounds__(256) void k_mlp_act(NetDev n, const short* __restrict__ dd, const float* __restrict__ dt,
                                                 const int* __restrict__ idx, long long first, int B, TrainBufs tb) {
    extern __shared__ __align__(16) unsigned char s_wraw[];
    unsigned short* sW1 = reinterpret_cast<unsigned short*>(s_wraw);   // [128 out][W1P]
    unsigned short* sW2t = sW1 + NN_H * W1P;                            // [128 out][W2P]
    unsigned short* sW2 = sW2t + NN_H * W2P;                            // [128 in][W2P]
    __shared__ __align__(16) unsigned short X[64 * XP];
    __shared__ __align__(16) unsigned short H1[64 * HP];
    __shared__ __align__(16) unsigned short DZ1[64 * HP];
    __shared__ __align__(16) unsigned short DZ2[64 * HP];
    __shared__ float s_t[64];
    __shared__ float s_red[4][3 * NN_H];      // per wave: db1, db2, dw3
    __shared__ float s_red2[4][2];            // per wave: db3, loss
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, rr = lane & 15, g = lane >> 4;
    const int b = blockIdx.x;
    const int* bidx = idx ? idx + (long long)b * 64 : nullptr;
    const long long bfirst = first + (long long)b * 64;
    // weights into LDS (8-byte chunks), every load of the copy issued with the rows' loads
    typedef unsigned short us4 __attribute__((ext_vector_type(4)));
    for (int c = tid; c < NN_H * (NN_IN / 4); c += 256) {
        const int row = c / (NN_IN / 4), k = c - row * (NN_IN / 4);
        *reinterpret_cast<us4*>(sW1 + row * W1P + 4 * k) = *reinterpret_cast<const us4*>(n.w1t + row * NN_IN + 4 * k);
    }
    for (int c = tid; c < NN_H * (NN_H / 4); c += 256) {
        const int row = c / (NN_H / 4), k = c - row * (NN_H / 4);
        *reinterpret_cast<us4*>(sW2t + row * W2P + 4 * k) = *reinterpret_cast<const us4*>(n.w2t + row * NN_H + 4 * k);
        *reinterpret_cast<us4*>(sW2 + row * W2P + 4 * k) = *reinterpret_cast<const us4*>(n.w2 + row * NN_H + 4 * k);
    }
    load_rows(dd, bidx, bfirst, 64, n.theta, X, tid, 256);
    const float om = n.theta[OFF_MEAN], os = n.theta[OFF_STD];
    if (tid < 64) {
        const long long s = bidx ? (long long)bidx[tid] : bfirst + tid;
        s_t[tid] = (dt[s] - om) / os;
    }
    __syncthreads();
    const int r0 = wave * 16;
    f4 z1[8], z2[8];
    {   // layer 1: 16x16x16 over the 48 inputs
        s4 a[3];
#pragma unroll
        for (int kb = 0; kb < 3; ++kb) a[kb] = ld4(X + (r0 + rr) * XP + kb * 16 + 4 * g);
#pragma unroll
        for (int ct = 0; ct < 8; ++ct) {
            f4 acc = {0.f, 0.f, 0.f, 0.f};
            const unsigned short* wr = sW1 + (ct * 16 + rr) * W1P + 4 * g;
#pragma unroll
            for (int kb = 0; kb < 3; ++kb) acc = mfma16(a[kb], ld4(wr + kb * 16), acc);
            z1[ct] = acc + n.theta[OFF_B1 + ct * 16 + rr];
        }
    }
#pragma unroll
    for (int ct = 0; ct < 8; ++ct)
#pragma unroll
        for (int e = 0; e < 4; ++e) H1[(r0 + 4 * g + e) * HP + ct * 16 + rr] = f2bf(gelu_f(z1[ct][e]));
    __syncthreads();
    {   // layer 2: 16x16x32 over the 128 hidden units
        bf8 a[4];
#pragma unroll
        for (int kb = 0; kb < 4; ++kb) a[kb] = ld8(H1 + (r0 + rr) * HP + kb * 32 + 8 * g);
#pragma unroll
        for (int ct = 0; ct < 8; ++ct) {
            f4 acc = {0.f, 0.f, 0.f, 0.f};
            const unsigned short* wr = sW2t + (ct * 16 + rr) * W2P + 8 * g;
#pragma unroll
            for (int kb = 0; kb < 4; ++kb) acc = mfma32(a[kb], ld8(wr + kb * 32), acc);
            z2[ct] = acc + n.theta[OFF_B2 + ct * 16 + rr];
        }
    }
    float yp[4] = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int ct = 0; ct < 8; ++ct) {
        const float w = bf2f(n.w3[ct * 16 + rr]);
#pragma unroll
        for (int e = 0; e < 4; ++e) yp[e] += bf2f(f2bf(gelu_f(z2[ct][e]))) * w;
    }
    const float b3 = n.theta[OFF_B3], invB = 1.0f / (float)B;
    float gy[4], lsum = 0.f, gsum = 0.f;
#pragma unroll
    for (int e = 0; e < 4; ++e) {
        const float y = xor_sum16(yp[e]) + b3;          // every lane of the group holds the sum
        const float err = y - s_t[r0 + 4 * g + e];
        gy[e] = 2.0f * err * invB;
        lsum += err * err;
        gsum += gy[e];
    }
    // backward through layer 3 and GELU 2; column reductions over this wave's 16 rows
#pragma unroll
    for (int ct = 0; ct < 8; ++ct) {
        const int col = ct * 16 + rr;
        const float w = bf2f(n.w3[col]);
        float dw3 = 0.f, db2 = 0.f;
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const float h2 = bf2f(f2bf(gelu_f(z2[ct][e])));
            dw3 += gy[e] * h2;
            const float dz = gy[e] * w * gelu_grad_f(z2[ct][e]);
            db2 += dz;
            DZ2[(r0 + 4 * g + e) * HP + col] = f2bf(dz);
        }
        dw3 = xor_sum_groups(dw3);
        db2 = xor_sum_groups(db2);
        if (g == 0) { s_red[wave][NN_H + col] = db2; s_red[wave][2 * NN_H + col] = dw3; }
    }
    if (lane == 0) { s_red2[wave][0] = 0.f; s_red2[wave][1] = 0.f; }
    __syncthreads();
    // dh1 = dz2 W2^T (this wave's rows; 16x16x32), then GELU 1
    {
        bf8 a[4];
#pragma unroll
        for (int kb = 0; kb < 4; ++kb) a[kb] = ld8(DZ2 + (r0 + rr) * HP + kb * 32 + 8 * g);
#pragma unroll
        for (int ct = 0; ct < 8; ++ct) {
            f4 acc = {0.f, 0.f, 0.f, 0.f};
            const unsigned short* wr = sW2 + (ct * 16 + rr) * W2P + 8 * g;
#pragma unroll
            for (int kb = 0; kb < 4; ++kb) acc = mfma32(a[kb], ld8(wr + kb * 32), acc);
            const int col = ct * 16 + rr;
            float db1 = 0.f;
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const float dz = acc[e] * gelu_grad_f(z1[ct][e]);
                db1 += dz;
                DZ1[(r0 + 4 * g + e) * HP + col] = f2bf(dz);
            }
            db1 = xor_sum_groups(db1);
            if (g == 0) s_red[wave][col] = db1;
        }
    }
    // gy and loss of this wave (each group of 16 lanes holds 4 rows; lanes rr == 0 contribute)
    {
        float l = rr == 0 ? lsum : 0.f, gs = rr == 0 ? gsum : 0.f;
        l = xor_sum_groups(l);
        gs = xor_sum_groups(gs);
        if (lane == 0) { s_red2[wave][0] = gs; s_red2[wave][1] = l; }
    }
    __syncthreads();
    // the 64 rows' activations, transposed (row index fastest): 4 rows per 8-byte store
    const int r0g = b * 64;
    for (int e = tid; e < 16 * NN_IN; e += 256) {
        const int c = e >> 4, r = 4 * (e & 15);
        const us4 v = {X[r * XP + c], X[(r + 1) * XP + c], X[(r + 2) * XP + c], X[(r + 3) * XP + c]};
        *reinterpret_cast<us4*>(tb.xt + (size_t)c * tb.ld + r0g + r) = v;
    }
    for (int e = tid; e < 16 * NN_H; e += 256) {
        const int c = e >> 4, r = 4 * (e & 15);
        const size_t o = (size_t)c * tb.ld + r0g + r;
        const us4 h = {H1[r * HP + c], H1[(r + 1) * HP + c], H1[(r + 2) * HP + c], H1[(r + 3) * HP + c]};
        const us4 d1 = {DZ1[r * HP + c], DZ1[(r + 1) * HP + c], DZ1[(r + 2) * HP + c], DZ1[(r + 3) * HP + c]};
        const us4 d2 = {DZ2[r * HP + c], DZ2[(r + 1) * HP + c], DZ2[(r + 2) * HP + c], DZ2[(r + 3) * HP + c]};
        *reinterpret_cast<us4*>(tb.h1t + o) = h;
        *reinterpret_cast<us4*>(tb.dz1t + o) = d1;
        *reinterpret_cast<us4*>(tb.dz2t + o) = d2;
    }
    float* vp = tb.vpart + (size_t)b * VP;
    for (int c = tid; c < NN_H; c += 256) {
        vp[c] = ((s_red[0][c] + s_red[1][c]) + s_red[2][c]) + s_red[3][c];
        vp[NN_H + c] = ((s_red[0][NN_H + c] + s_red[1][NN_H + c]) + s_red[2][NN_H + c]) + s_red[3][NN_H + c];
        vp[2 * NN_H + c] =
            ((s_red[0][2 * NN_H + c] + s_red[1][2 * NN_H + c]) + s_red[2][2 * NN_H + c]) + s_red[3][2 * NN_H + c];
    }
    if (tid == 0) {
        vp[3 * NN_H] = ((s_red2[0][0] + s_red2[1][0]) + s_red2[2][0]) + s_red2[3][0];
        vp[3 * NN_H + 1] = ((s_red2[0][1] + s_red2[1][1]) + s_red2[2][1]) + s_red2[3][1];
    }
}

constexpr int WG_TILES = 64 + 24;   // dW2: 8 x 8 tiles, dW1: 3 x 8 tiles of 16 x 16

constexpr int WG_THREADS = 1024, WG_WAVES = WG_THREADS / 64, WG_UNROLL = 4;

// g[W2 / W1 entries] and, by the last workgroup, g[vector parameters] and g[NN_TRAIN] (loss sum).
// Wave w of a tile's workgroup takes the 32-row chunks w, w + 16, ...; each wave issues the loads
// of WG_UNROLL chunks before their MFMAs.
__global__ __launch_bounds__(WG_THREADS) void k_mlp_wgrad(TrainBufs tb, int nparts, float* __restrict__ g) {
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, r = lane & 15, q = lane >> 4;
    const int t = blockIdx.x;
    if (t == WG_TILES) {
        for (int k = tid; k < VP; k += WG_THREADS) {
            // the partials in workgroup order, loaded 16 at a time ahead of their (ordered) additions
            float s = 0.f;
            int p = 0;
            for (; p + 16 <= nparts; p += 16) {
                float v[16];
#pragma unroll
                for (int u = 0; u < 16; ++u) v[u] = tb.vpart[(size_t)(p + u) * VP + k];
#pragma unroll
                for (int u = 0; u < 16; ++u) s += v[u];
            }
            for (; p < nparts; ++p) s += tb.vpart[(size_t)p * VP + k];
            const int o = k < NN_H ? OFF_B1 + k : k < 2 * NN_H ? OFF_B2 + (k - NN_H)
                        : k < 3 * NN_H ? OFF_W3 + (k - 2 * NN_H) : k == 3 * NN_H ? OFF_B3 : NN_TRAIN;
            g[o] = s;
        }
        return;
    }
    const bool w2 = t < 64;
    const int tt = w2 ? t : t - 64, ti = tt >> 3, tj = tt & 7;
    const unsigned short* At = (w2 ? tb.h1t : tb.xt) + (size_t)(ti * 16 + r) * tb.ld + 8 * q;
    const unsigned short* Bt = (w2 ? tb.dz2t : tb.dz1t) + (size_t)(tj * 16 + r) * tb.ld + 8 * q;
    f4 acc = {0.f, 0.f, 0.f, 0.f};
    const int nchunk = tb.ld / 32;
    int c = wave;
    for (; c + (WG_UNROLL - 1) * WG_WAVES < nchunk; c += WG_UNROLL * WG_WAVES) {
        bf8 a[WG_UNROLL], b[WG_UNROLL];
#pragma unroll
        for (int u = 0; u < WG_UNROLL; ++u) {
            a[u] = ld8(At + 32 * (c + u * WG_WAVES));
            b[u] = ld8(Bt + 32 * (c + u * WG_WAVES));
        }
#pragma unroll
        for (int u = 0; u < WG_UNROLL; ++u) acc = mfma32(a[u], b[u], acc);
    }
    for (; c < nchunk; c += WG_WAVES) acc = mfma32(ld8(At + 32 * c), ld8(Bt + 32 * c), acc);
    __shared__ f4 red[WG_WAVES][64];
    red[wave][lane] = acc;
    __syncthreads();
    if (wave == 0) {
        f4 s = red[0][lane];
#pragma unroll
        for (int w = 1; w < WG_WAVES; ++w) s += red[w][lane];
        const int off = w2 ? OFF_W2 : OFF_W1;
#pragma unroll
        for (int e = 0; e < 4; ++e) g[off + (ti * 16 + 4 * q + e) * NN_H + tj * 16 + r] = s[e];
    }
}

__device__ __forceinline__ void cast_param(int k, float v, unsigned short* w1t, unsigned short* w2t,
                                           unsigned short* w2, unsigned short* w3) {
    if (k < OFF_B1) {
        const int in = k / NN_H, o = k - in * NN_H;
        w1t[o * NN_IN + in] = f2bf(v);
    } else if (k >= OFF_W2 && k < OFF_B2) {
        const int r = k - OFF_W2, in = r / NN_H, o = r - in * NN_H;
        w2t[o * NN_H + in] = f2bf(v);
        w2[in * NN_H + o] = f2bf(v);
    } else if (k >= OFF_W3 && k < OFF_B3) {
        w3[k - OFF_W3] = f2bf(v);
    }
}

__global__ void k_mlp_cast(const float* __restrict__ theta, unsigned short* w1t, unsigned short* w2t,
                           unsigned short* w2, unsigned short* w3) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < NN_TRAIN) cast_param(k, theta[k], w1t, w2t, w2, w3);
}

// Adam (Kingma & Ba 2015, Algorithm 1) on the gradient g, then the BF16 copies.
__global__ void k_mlp_adam(float* __restrict__ theta, float* __restrict__ m, float* __restrict__ v,
                           const float* __restrict__ g, float lr, float b1, float b2, float eps, float bc1, float bc2,
                           unsigned short* w1t, unsigned short* w2t, unsigned short* w2, unsigned short* w3,
                           float* __restrict__ loss_out, float invB) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k == NN_TRAIN && loss_out) *loss_out = g[NN_TRAIN] * invB;
    if (k >= NN_TRAIN) return;
    const float gk = g[k];
    const float mk = b1 * m[k] + (1.0f - b1) * gk;
    const float vk = b2 * v[k] + (1.0f - b2) * gk * gk;
    m[k] = mk;
    v[k] = vk;
    const float th = theta[k] - lr * (mk / bc1) / (sqrtf(vk / bc2) + eps);
    theta[k] = th;
    cast_param(k, th, w1t, w2t, w2, w3);
}

// ---------------------------------------------------------------- descriptors
// What the surrogate's physics terms read about a pool atom, in one 64-byte
// record (four 16-byte loads): position and self term (fixed), and the
// potential term, charge, type and species, rewritten after every refresh
// (a delayed-acceptance launch holds at most one accepted move and ends with
// it, so the dynamic part is current for the whole launch).
struct __align__(16) AtomRec {
    double x, y, z, selfw;
    double phiw, q;
    int type, sp;
    int pad_[2];
};

struct DaDev {
    const int* dhead;             // r_desc rows (both directions)
    const unsigned* dlist;        // atom | shell << 24
    unsigned short* H;            // per-atom histograms [N][48]
    const signed char* species;   // per 0-based type
    const double *x, *y;
    double Lx, Ly, Lz, rd2;
    double th[12];                // shell thresholds (s r_desc / 12)^2
    double rc2;                   // real-space cutoff^2 of the pair term
    int M;                        // candidates per round
    int prof;                     // accumulate per-phase times into McCtl::da_prof
    double* rec_fy;               // per-trial record of the reverse surrogate dE*(y->x)
    AtomRec* rec;                 // per atom (pool atoms: dynamic part current)
    int upd_one;                  // the site's and the ion's species sets are disjoint: one update pass
    long long* psi;               // per atom: real-space potential of the listed neighbours (fixed point)
    int use_psi;                  // the physics part includes dq_i psi_i + dq_j psi_j
    // local-pair form: rows of the pool atoms within r_local of each pool atom
    // (ascending), their coefficients C_ab in fixed point, and per atom
    // psil_a = sum_b q_b C_ab in fixed point (rewritten in place by accepted moves)
    int local;
    double rl2;                   // r_local^2
    const int* lhead;
    const int* latom;
    const long long* lcoef;
    long long* psil;
    const double* q0;             // charges at the freeze
    double hx, hy, hz;            // mesh spacings
    double cw;                    // half-width of the lateral column that L also covers
    const double* gs;             // mesh kernel smoothed by two grid-point stencils, folded as gf
    double* rqn;                  // per pending entry: the atom's charge after the change (written at launch end)
    int* fold_from;               // first pending entry a fold still has to add to psil
    unsigned* fapplied;           // per fold workgroup: the pending list (pend_seq) it has added
};

// Phase clock of the chain (100 MHz real-time counter), thread 0 only.
struct PhaseClock {
    long long t = 0;
    __device__ __forceinline__ void start(bool on) { if (on) t = (long long)__builtin_amdgcn_s_memrealtime(); }
    __device__ __forceinline__ void lap(bool on, long long* acc, int k) {
        if (!on) return;
        const long long n = (long long)__builtin_amdgcn_s_memrealtime();
        acc[k] += n - t;
        t = n;
    }
};

__device__ __forceinline__ int shell_of(const DaDev& A, double r2) {
    int s = 0;
#pragma unroll
    for (int k = 1; k < 12; ++k) s += r2 >= A.th[k] ? 1 : 0;
    return s;
}

// r_ij and the two exclusion slots of a move: exA (j's entry in h_i) and exB (i's entry in h_j).
__device__ __forceinline__ void pair_terms(const DaDev& A, const McDev& D, int i, int j, double& r, int& exA, int& exB) {
    const double dx = min_image_exact(A.x[i] - A.x[j], A.Lx), dy = min_image_exact(A.y[i] - A.y[j], A.Ly),
                 dz = min_image_exact(D.z[i] - D.z[j], A.Lz);
    const double r2 = r2_exact(dx, dy, dz);
    r = sqrt(r2);
    exA = -1; exB = -1;
    if (r2 <= A.rd2) {
        const int sh = shell_of(A, r2);
        const int sj = A.species[D.type[j]], si = A.species[D.type[i]];
        if (sj >= 0) exA = sh * 4 + sj;
        if (si >= 0) exB = sh * 4 + si;
    }
}

__device__ __forceinline__ int desc_value(const DaDev& A, int i, int j, int c, int exA, int exB) {
    return (int)A.H[(size_t)i * NN_IN + c] - (int)A.H[(size_t)j * NN_IN + c] - (c == exA ? 1 : 0) + (c == exB ? 1 : 0);
}

__global__ void k_da_row_entries(int N, const double* __restrict__ x, const double* __restrict__ y,
                                 const double* __restrict__ z, DaDev A, const int* __restrict__ nbr,
                                 unsigned* __restrict__ out) {
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= N) return;
    for (int e = A.dhead[a]; e < A.dhead[a + 1]; ++e) {
        const int b = nbr[e];
        const double dx = min_image_exact(x[a] - x[b], A.Lx), dy = min_image_exact(y[a] - y[b], A.Ly),
                     dz = min_image_exact(z[a] - z[b], A.Lz);
        out[e] = (unsigned)b | ((unsigned)shell_of(A, r2_exact(dx, dy, dz)) << DA_SHIFT);
    }
}

// One wave per atom: counts by shell and species of its row.
__global__ __launch_bounds__(256) void k_da_hist(int N, DaDev A, const int* __restrict__ type) {
    __shared__ unsigned cnt[4][NN_IN];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int a = blockIdx.x * 4 + w;
    if (lane < NN_IN) cnt[w][lane] = 0;
    __syncthreads();
    if (a < N) {
        for (int e = A.dhead[a] + lane; e < A.dhead[a + 1]; e += 64) {
            const unsigned v = A.dlist[e];
            const int sp = A.species[type[v & 0xFFFFFFu]];
            if (sp >= 0) atomicAdd(&cnt[w][(v >> DA_SHIFT) * 4 + sp], 1u);
        }
    }
    __syncthreads();
    if (a < N && lane < NN_IN) A.H[(size_t)a * NN_IN + lane] = (unsigned short)cnt[w][lane];
}

// Descriptors and pair distances of given moves (one thread per feature).
__global__ void k_da_desc(DaDev A, McDev D, int n, const int* __restrict__ mi, const int* __restrict__ mj,
                          short* __restrict__ d, double* __restrict__ rij) {
    const long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (e >= (long long)n * NN_IN) return;
    const int m = (int)(e / NN_IN), c = (int)(e - (long long)m * NN_IN);
    double r;
    int exA, exB;
    pair_terms(A, D, mi[m], mj[m], r, exA, exB);
    d[e] = (short)desc_value(A, mi[m], mj[m], c, exA, exB);
    if (c == 0 && rij) rij[m] = r;
}

// ---------------------------------------------------------------- physics part of the surrogate
// dE*(x -> y) = lin + selfq + pr + corr + s f(d) for the move of site i and ion j:
//   lin   = sum_a dq_a W_a . phi        reciprocal-space potential at the two atoms (phi exact:
//                                       no accepted change is pending while the surrogate runs)
//   selfq = 0.5 sum_a dq_a^2 W_a^T G W_a
//   pr    = (qi' qj' - qi qj) elF erfc(alpha r)/r [r <= rc] + dqi dqj elF erf(alpha r)/r
//           (real-space pair, and the mesh pair term in its continuum form)
//   corr  = the exact self, net-charge and slab change
// and f the MLP of the shell-count descriptor, which learns the rest (the
// real-space energy of the environment).
struct MoveTerms {
    double dqi, dqj, selfq, perfc, perf, corr, psi;
};

// Real-space potential of the listed neighbours at each atom, psi_a = sum over
// a's neighbour row of q_k elF erfc(alpha r)/r, in fixed point (2^-32 units):
// each pair contributes llrint(q_k llrint(kreal 2^32)), so psi is an exact
// integer sum, a function of the charges alone whatever order the changes of an
// accepted move are added in.
constexpr double PSI_SCALE = 4294967296.0, PSI_INV = 1.0 / 4294967296.0;
__device__ __forceinline__ long long psi_k(double kreal) { return __double2ll_rn(kreal * PSI_SCALE); }
__device__ __forceinline__ long long psi_c(double q, long long K) { return __double2ll_rn(q * (double)K); }

__device__ __forceinline__ long long psi_of(const McDev& D, int a) {
    long long s = 0;
    for (int e = D.fhead[a]; e < D.fhead[a + 1]; ++e) s += psi_c(D.q[D.fnbr[e]], psi_k(D.kreal[e]));
    return s;
}

// The surrogate's real-space term of a move: dq_i psi_i + dq_j psi_j.
__device__ __forceinline__ double psi_term(double dqi, long long pi, double dqj, long long pj) {
    return (dqi * (double)pi + dqj * (double)pj) * PSI_INV;
}

__global__ void k_psi_all(McDev D, long long* __restrict__ psi) {
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a < D.N) psi[a] = psi_of(D, a);
}

__device__ __forceinline__ MoveTerms move_terms(const McDev& D, const DaDev& A, int i, int j, int fwd, double r,
                                                double Q, double Mz, double Qz2) {
    MoveTerms m;
    const double qi = D.q[i], qj = D.q[j];
    const double qin = fwd ? 1.0 : 0.0, qjn = fwd ? -1.0 : 0.0;
    m.dqi = qin - qi;
    m.dqj = qjn - qj;
    m.selfq = 0.5 * (m.dqi * m.dqi * D.selfw[i] + m.dqj * m.dqj * D.selfw[j]);
    // the real-space change is dq_i psi_i + dq_j psi_j + dq_i dq_j k_ij (k_ij: the listed
    // pair's coefficient); the pair terms are part of the surrogate only: single
    // precision is enough
    // (without psi the pair term is the whole listed pair's change, (q_i' q_j' - q_i q_j) k_ij)
    const float ec = erfcf((float)(D.alpha * r)), kr = (float)(D.elF / r);
    const double pq = A.use_psi ? m.dqi * m.dqj : qin * qjn - qi * qj;
    m.perfc = r * r <= A.rc2 ? (double)((float)pq * kr * ec) : 0.0;
    m.perf = (double)((float)(m.dqi * m.dqj) * kr * (1.0f - ec));
    m.corr = corr_move(D, qi, qin, qj, qjn, D.z[i], D.z[j], Q, Mz, Qz2);
    m.psi = A.use_psi ? psi_term(m.dqi, psi_of(D, i), m.dqj, psi_of(D, j)) : 0.0;
    return m;
}

// pair_terms and move_terms from the atoms' records (same inputs, same results).
__device__ __forceinline__ void pair_terms_rec(const DaDev& A, const AtomRec& ri, const AtomRec& rj, double& r,
                                               int& exA, int& exB) {
    const double dx = min_image_exact(ri.x - rj.x, A.Lx), dy = min_image_exact(ri.y - rj.y, A.Ly),
                 dz = min_image_exact(ri.z - rj.z, A.Lz);
    const double r2 = r2_exact(dx, dy, dz);
    r = sqrt(r2);
    exA = -1; exB = -1;
    if (r2 <= A.rd2) {
        const int sh = shell_of(A, r2);
        if (rj.sp >= 0) exA = sh * 4 + rj.sp;
        if (ri.sp >= 0) exB = sh * 4 + ri.sp;
    }
}

__device__ __forceinline__ MoveTerms move_terms_rec(const McDev& D, const DaDev& A, const AtomRec& ri,
                                                    const AtomRec& rj, int fwd, double r, double Q, double Mz,
                                                    double Qz2, long long pi, long long pj) {
    MoveTerms m;
    const double qi = ri.q, qj = rj.q;
    const double qin = fwd ? 1.0 : 0.0, qjn = fwd ? -1.0 : 0.0;
    m.dqi = qin - qi;
    m.dqj = qjn - qj;
    m.selfq = 0.5 * (m.dqi * m.dqi * ri.selfw + m.dqj * m.dqj * rj.selfw);
    const float ec = erfcf((float)(D.alpha * r)), kr = (float)(D.elF / r);
    const double pq = A.use_psi ? m.dqi * m.dqj : qin * qjn - qi * qj;
    m.perfc = r * r <= A.rc2 ? (double)((float)pq * kr * ec) : 0.0;
    m.perf = (double)((float)(m.dqi * m.dqj) * kr * (1.0f - ec));
    m.corr = corr_move(D, qi, qin, qj, qjn, ri.z, rj.z, Q, Mz, Qz2);
    m.psi = A.use_psi ? psi_term(m.dqi, pi, m.dqj, pj) : 0.0;
    return m;
}

__global__ void k_rec_static(int N, const double* __restrict__ x, const double* __restrict__ y,
                             const double* __restrict__ z, const double* __restrict__ selfw, AtomRec* __restrict__ rec) {
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= N) return;
    rec[a].x = x[a]; rec[a].y = y[a]; rec[a].z = z[a]; rec[a].selfw = selfw[a];
}

// One (z, y) row r < P^2 of W_a . phi.
template <int P>
__device__ __forceinline__ double stencil_row(const McDev& D, int a, int r) {
    const int wz = r / P, wy = r - wz * P;
    const int* sb = D.sbase + 3 * a;
    const double* w = D.sw + (size_t)a * 3 * P;
    const int gz = wrap1(sb[2] + wz, D.nz), gy = wrap1(sb[1] + wy, D.ny);
    const double* row = D.phi + ((size_t)gz * D.ny + gy) * D.nx;
    double sx = 0.0;
#pragma unroll
    for (int u = 0; u < P; ++u) sx += w[u] * row[wrap1(sb[0] + u, D.nx)];
    return w[2 * P + wz] * w[P + wy] * sx;
}

template <int P>
__device__ double stencil_dot(const McDev& D, int a) {
    double s = 0.0;
    for (int r = 0; r < P * P; ++r) s += stencil_row<P>(D, a, r);
    return s;
}

// W_a . phi of pool atom list[e] (the site and ion pools, a fixed set) into its
// record with its charge, type and species: 8 lanes per atom (lanes u = 8 e + 0..7
// of the caller's numbering), lane u < P sums the x column u of the stencil; the
// list is in stencil-origin order, so neighbouring groups read neighbouring grid
// rows.
template <int P>
__device__ __forceinline__ void phi_dot_atom(const McDev& D, const signed char* __restrict__ species,
                                             const int* __restrict__ list, int e, int u, AtomRec* __restrict__ rec) {
    const int a = list[e];
    const int* sb = D.sbase + 3 * a;
    const double* w = D.sw + (size_t)a * 3 * P;
    double t = 0.0;
    if (u < P) {
        const int gx = wrap1(sb[0] + u, D.nx);
        double v[P * P];
#pragma unroll
        for (int wz = 0; wz < P; ++wz) {
            const size_t zrow = (size_t)wrap1(sb[2] + wz, D.nz) * D.ny;
#pragma unroll
            for (int wy = 0; wy < P; ++wy) v[wz * P + wy] = D.phi[(zrow + wrap1(sb[1] + wy, D.ny)) * D.nx + gx];
        }
#pragma unroll
        for (int wz = 0; wz < P; ++wz) {
            double sy = 0.0;
#pragma unroll
            for (int wy = 0; wy < P; ++wy) sy += w[P + wy] * v[wz * P + wy];
            t += w[2 * P + wz] * sy;
        }
        t *= w[u];
    }
    t += __shfl_xor(t, 4, 8);
    t += __shfl_xor(t, 2, 8);
    t += __shfl_xor(t, 1, 8);
    if (u == 0) {
        const int ty = D.type[a];
        rec[a].phiw = t;
        rec[a].q = D.q[a];
        rec[a].type = ty;
        rec[a].sp = species[ty];
    }
}

// Workgroups are dealt to the 8 XCDs in turn (workgroup b on XCD b mod 8), each
// XCD with its own L2: XCD x takes the x-th eighth of the list (neighbouring grid
// rows), so that each XCD's L2 holds one slab of the grid instead of all of it.
// The grid is a multiple of 8 workgroups; logical workgroups past the list leave.
template <int P>
__global__ __launch_bounds__(256) void k_phi_dot(McDev D, const signed char* __restrict__ species,
                                                 const int* __restrict__ list, int n, AtomRec* __restrict__ rec) {
    const int per = gridDim.x >> 3, lb = (blockIdx.x & 7) * per + (blockIdx.x >> 3);
    const int gt = lb * blockDim.x + threadIdx.x, e = gt >> 3;
    if (e >= n) return;
    phi_dot_atom<P>(D, species, list, e, gt & 7, rec);
}

// Physics part B = lin + selfq + pr + corr of n moves (one thread per move).
template <int P>
__global__ void k_da_base(McDev D, DaDev A, int n, const int* __restrict__ mi, const int* __restrict__ mj,
                          const int* __restrict__ mf, double* __restrict__ base) {
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= n) return;
    const McCtl* C = D.ctl;
    const int i = mi[m], j = mj[m];
    double r;
    int exA, exB;
    pair_terms(A, D, i, j, r, exA, exB);
    const MoveTerms t = move_terms(D, A, i, j, mf[m], r, C->Q, C->Mz, C->Qz2);
    const double lin = t.dqi * stencil_dot<P>(D, i) + t.dqj * stencil_dot<P>(D, j);
    base[m] = lin + t.selfq + t.perfc + t.perf + t.corr + t.psi;
}

// Training target t = s (dE - B).
__global__ void k_da_target(int n, const int* __restrict__ mf, const double* __restrict__ dE,
                            const double* __restrict__ base, float* __restrict__ t) {
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= n) return;
    const double v = dE[m] - base[m];
    t[m] = (float)(mf[m] ? v : -v);
}

// Surrogate dE* = B + s f(d).
__global__ void k_da_surrogate(int n, const int* __restrict__ mf, const float* __restrict__ y,
                               const double* __restrict__ base, double* __restrict__ out) {
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= n) return;
    const double f = (double)y[m];
    out[m] = base[m] + (mf[m] ? f : -f);
}

// Every stride-th dataset row (normalisation statistics).
__global__ void k_gather_rows(const short* __restrict__ dd, const float* __restrict__ dt, long long stride, long long ns,
                              short* __restrict__ od, float* __restrict__ ot) {
    const long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (e >= ns * NN_IN) return;
    const long long r = e / NN_IN, c = e - r * NN_IN;
    od[e] = dd[r * stride * NN_IN + c];
    if (c == 0) ot[r] = dt[r * stride];
}

__global__ void k_da_set_target(McCtl* C, long long target) {
    if (threadIdx.x == 0) C->trials_target = target;
}

// ---------------------------------------------------------------- local-pair form
// The potential at a pool atom a (mesh and real space, its own mesh self term
// included) is written as
//   pot_a = q_a s_a + sum_{b in L(a)} q_b C_ab + sum_{b pool, not in L(a)} q_b G_ab + rest_a
// with L(a) the pool atoms within r_local of a or in its lateral column (see
// local_member), C_ab = W_a^T G W_b + k_ab (k_ab the
// real-space coefficient when a and b are listed neighbours), G_ab = W_a^T G W_b,
// rest_a the potential of the fixed charges.  The surrogate freezes everything
// at a state x0 and tracks the change since then in fixed point (2^-32 units,
// each pair rounded on its own, so that the sums are exact integers that do not
// depend on the order or time at which changes are added):
//   T_a(x) = sum_{b in L(a)} c(q_b(x), C_ab) + sum_{b not in L(a)} [c(q_b(x), Gt_ab) - c(q_b(x0), Gt_ab)]
//   P_a(x) = far_a + q_a(x) s_a + T_a(x),   far_a = pot_a(x0) - q_a(x0) s_a - T_a(x0)
// with c(q, K) = rint(q K) and Gt_ab the mesh kernel at the pair's displacement
// (trilinear in the folded table, |offset| per axis: the same bits for (a, b)
// and (b, a)) in place of W_a^T G W_b beyond r_local.  P_a is a function of the
// state for x0 fixed before the chain starts, equals pot_a at x0, and differs
// from it later only by the difference of Gt and W^T G W over the pool atoms
// that changed beyond r_local.  The move's physics part is
//   dq_i P_i + dq_j P_j + 0.5 (dq_i^2 s_i + dq_j^2 s_j) + dq_i dq_j K_ij + corr,
// K_ij = C_ij or Gt_ij (the same coefficient T carries), so the reverse move at y
// is its negative up to the fixed-point rounding.  The chain adds the L rows of an
// accepted move to T at once; the far part of its changes is added to every pool
// atom by k_loc_fold after the launch, and until then the chain adds it for the
// atoms it evaluates.  No contraction in these sums: the chain, the unchained
// evaluation and the reverse value form the same ones.

// W_a^T G W_b of two stencils (one thread; the separable correlations times the folded mesh kernel).
template <int P>
__device__ double mesh_pair(const McDev& D, int a, int b) {
    constexpr int CW = 2 * P - 1;
    const int* sba = D.sbase + 3 * a;
    const int* sbb = D.sbase + 3 * b;
    const double* wa = D.sw + (size_t)a * 3 * P;
    const double* wb = D.sw + (size_t)b * 3 * P;
    double cx[CW], cy[CW], cz[CW];
#pragma unroll
    for (int d = 0; d < CW; ++d) {
        cx[d] = corr<P>(wa, wb, d);
        cy[d] = corr<P>(wa + P, wb + P, d);
        cz[d] = corr<P>(wa + 2 * P, wb + 2 * P, d);
    }
    const int x0 = wrap1(sba[0] - sbb[0], D.nx) - (P - 1);
    const int y0 = wrap1(sba[1] - sbb[1], D.ny) - (P - 1);
    const int z0 = wrap1(sba[2] - sbb[2], D.nz) - (P - 1);
    int gx[CW];
#pragma unroll
    for (int d = 0; d < CW; ++d) gx[d] = fold(wrap1(x0 + d, D.nx), D.nx);
    double s = 0.0;
    for (int dz = 0; dz < CW; ++dz) {
        const int gz = fold(wrap1(z0 + dz, D.nz), D.nz);
        double sz = 0.0;
        for (int dy = 0; dy < CW; ++dy) {
            const int gy = fold(wrap1(y0 + dy, D.ny), D.ny);
            const double* row = D.gf + ((size_t)gz * D.fny + gy) * D.fnx;
            double sx = 0.0;
#pragma unroll
            for (int d = 0; d < CW; ++d) sx += cx[d] * row[gx[d]];
            sz += cy[dy] * sx;
        }
        s += cz[dz] * sz;
    }
    return s;
}

// Minimum-image displacement and its square (the arithmetic of the neighbour lists).
__device__ __forceinline__ double pair_disp(const DaDev& A, double xa, double ya, double za, double xb, double yb,
                                            double zb, double& dx, double& dy, double& dz) {
    dx = min_image_exact(xa - xb, A.Lx);
    dy = min_image_exact(ya - yb, A.Ly);
    dz = min_image_exact(za - zb, A.Lz);
    return r2_exact(dx, dy, dz);
}

__device__ __forceinline__ double pair_r2(const DaDev& A, const McDev& D, int a, int b) {
    double dx, dy, dz;
    return pair_disp(A, A.x[a], A.y[a], D.z[a], A.x[b], A.y[b], D.z[b], dx, dy, dz);
}

__device__ __forceinline__ int fold_index(int k, int n) { return fold(k < n ? k : k - n, n); }

// Gt of a displacement: trilinear at |offset| per axis in the mesh kernel smoothed
// by the stencils of two atoms on grid points (the kernel itself oscillates at the
// grid scale: the spline moduli are divided out of it).
__device__ __forceinline__ long long far_coef(const McDev& D, const DaDev& A, double dx, double dy, double dz) {
#pragma clang fp contract(off)
    const double ax = fabs(dx) / A.hx, ay = fabs(dy) / A.hy, az = fabs(dz) / A.hz;
    const int ix = (int)ax, iy = (int)ay, iz = (int)az;
    const double tx = ax - (double)ix, ty = ay - (double)iy, tz = az - (double)iz;
    const int x0 = fold_index(ix, D.nx), x1 = fold_index(ix + 1, D.nx);
    const int y0 = fold_index(iy, D.ny), y1 = fold_index(iy + 1, D.ny);
    const int z0 = fold_index(iz, D.nz), z1 = fold_index(iz + 1, D.nz);
    const double* g = A.gs;
    const size_t fx = (size_t)D.fnx, fxy = (size_t)D.fnx * D.fny;
    auto at = [&](int z, int y, int x) { return g[(size_t)z * fxy + (size_t)y * fx + x]; };
    const double c00 = at(z0, y0, x0) + tx * (at(z0, y0, x1) - at(z0, y0, x0));
    const double c01 = at(z0, y1, x0) + tx * (at(z0, y1, x1) - at(z0, y1, x0));
    const double c10 = at(z1, y0, x0) + tx * (at(z1, y0, x1) - at(z1, y0, x0));
    const double c11 = at(z1, y1, x0) + tx * (at(z1, y1, x1) - at(z1, y1, x0));
    const double c0 = c00 + ty * (c01 - c00), c1 = c10 + ty * (c11 - c10);
    return psi_k(c0 + tz * (c1 - c0));
}

// Pool pairs that L carries exactly: within r_local, or in the same lateral
// column (|dx|, |dy| <= cw, any dz).  The column holds what the interpolated
// kernel misses: the spline moduli divided out of the mesh kernel amplify its
// near-Nyquist lateral modes, and those with kz = 0 couple atoms at nearly the
// same (x, y) at any height.  A function of the displacement, the same for (a, b)
// and (b, a).
__device__ __forceinline__ bool local_member(const DaDev& A, double r2, double dx, double dy) {
    return r2 <= A.rl2 || (fabs(dx) <= A.cw && fabs(dy) <= A.cw);
}

// The change of T_a from atom p's change qo -> qn when L(a) does not hold p (0
// when it does, and for p == a).
__device__ __forceinline__ long long far_delta(const McDev& D, const DaDev& A, int a, double xa, double ya,
                                               double za, int p, double xp, double yp, double zp, double qo,
                                               double qn) {
    if (p == a) return 0;
    double dx, dy, dz;
    const double r2 = pair_disp(A, xa, ya, za, xp, yp, zp, dx, dy, dz);
    if (local_member(A, r2, dx, dy)) return 0;
    const long long K = far_coef(D, A, dx, dy, dz);
    return psi_c(qn, K) - psi_c(qo, K);
}

__device__ __forceinline__ double local_pot(double far, double q, double s, long long T) {
#pragma clang fp contract(off)
    return far + q * s + (double)T * PSI_INV;
}

__device__ __forceinline__ double local_selfq(double dqi, double si, double dqj, double sj) {
#pragma clang fp contract(off)
    return 0.5 * (dqi * dqi * si + dqj * dqj * sj);
}

__device__ __forceinline__ double local_pair(double dqi, double dqj, long long K) {
#pragma clang fp contract(off)
    return (dqi * dqj) * ((double)K * PSI_INV);
}

__device__ __forceinline__ double local_phys(double dqi, double Pi, double dqj, double Pj, double selfq, double pair,
                                             double corr) {
#pragma clang fp contract(off)
    return dqi * Pi + dqj * Pj + selfq + pair + corr;
}

// C_ij from i's ascending L row by one wave: 64 samples at a stride, then the
// segment that must hold j (two dependent loads instead of a binary search's
// chain).  All lanes of the wave call it with the same i and j.
__device__ __forceinline__ long long wave_find_coef(const DaDev& A, int i, int j, int lane) {
    const int h = A.lhead[i], n = A.lhead[i + 1] - h;
    const int stride = (n + 63) >> 6;
    const int k = lane * stride;
    const bool le = k < n && A.latom[h + k] <= j;
    const unsigned long long m = __ballot(le);
    const int seg = m ? 63 - __clzll((long long)m) : 0;
    const int e = seg * stride + lane;
    const bool hit = lane < stride && e < n && A.latom[h + e] == j;
    const unsigned long long mh = __ballot(hit);
    const int src = mh ? __ffsll((long long)mh) - 1 : 0;
    long long K = 0;
    if (hit) K = A.lcoef[h + e];
    return __shfl(K, src, 64);
}

// K_ij of a move's pair: C_ij from i's L row (ascending: binary search) when L(i)
// holds j, else Gt_ij.
__device__ __forceinline__ long long local_pair_coef(const McDev& D, const DaDev& A, int i, int j, double xi,
                                                     double yi, double zi, double xj, double yj, double zj) {
    double dx, dy, dz;
    const double r2 = pair_disp(A, xi, yi, zi, xj, yj, zj, dx, dy, dz);
    if (local_member(A, r2, dx, dy)) {
        int l = A.lhead[i], h = A.lhead[i + 1];
        while (l < h) {
            const int m = (l + h) >> 1;
            if (A.latom[m] < j) l = m + 1;
            else h = m;
        }
        if (l < A.lhead[i + 1] && A.latom[l] == j) return A.lcoef[l];
    }
    return far_coef(D, A, dx, dy, dz);
}

// T of the atoms in L(i) and L(j) after a move (charges qio -> qin, qjo -> qjn):
// integer atomics without return, both rows in one pass, each thread's row
// entries loaded before any atomic is issued.  The caller fences before anything
// reads T again; the chain reads T with agent-scope loads, which see the atomics
// (performed in L2) past the L1.
__device__ __forceinline__ void local_update(const DaDev& A, int i, double qio, double qin, int j, double qjo,
                                             double qjn, int tid, int bd) {
    constexpr int U = 4;
    const int hi = A.lhead[i], ni = A.lhead[i + 1] - hi, hj = A.lhead[j], nt = ni + (A.lhead[j + 1] - hj);
    for (int base = 0; base < nt; base += U * bd) {
        long long K[U];
        int b[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int e = base + tid + u * bd;
            b[u] = -1;
            K[u] = 0;
            if (e < nt) {
                const int x = e < ni ? hi + e : hj + (e - ni);
                K[u] = A.lcoef[x];
                b[u] = A.latom[x];
            }
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int e = base + tid + u * bd;
            if (b[u] >= 0) {
                const bool first = e < ni;
                const long long d = first ? psi_c(qin, K[u]) - psi_c(qio, K[u]) : psi_c(qjn, K[u]) - psi_c(qjo, K[u]);
                if (d != 0) atomicAdd(reinterpret_cast<unsigned long long*>(A.psil + b[u]), (unsigned long long)d);
            }
        }
    }
}

__device__ __forceinline__ long long load_T(const DaDev& A, int a) {
    return __hip_atomic_load(A.psil + a, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// L rows, one wave per pool atom a (list, n atoms): the pool atoms b (ps, np of
// them in ascending order) that local_member admits, counted (out = cnt[a]) or
// written in ascending order from lhead[a] (fill).
template <bool FILL>
__global__ __launch_bounds__(256) void k_loc_rows(McDev D, DaDev A, const int* __restrict__ list, int n,
                                                  const int* __restrict__ ps, int np, int* __restrict__ cnt,
                                                  int* __restrict__ latom, int* __restrict__ lown) {
    const int w = (blockIdx.x * blockDim.x + threadIdx.x) >> 6, lane = threadIdx.x & 63;
    if (w >= n) return;
    const int a = list[w];
    const double xa = A.x[a], ya = A.y[a], za = D.z[a];
    int o = FILL ? A.lhead[a] : 0;
    for (int k0 = 0; k0 < np; k0 += 64) {
        const int k = k0 + lane;
        bool in = false;
        int b = -1;
        if (k < np) {
            b = ps[k];
            double dx, dy, dz;
            const double r2 = pair_disp(A, xa, ya, za, A.x[b], A.y[b], D.z[b], dx, dy, dz);
            in = b != a && local_member(A, r2, dx, dy);
        }
        const unsigned long long m = __ballot(in);
        if (FILL && in) {
            const int pos = o + __popcll(m & ((1ull << lane) - 1ull));
            latom[pos] = b;
            lown[pos] = a;
        }
        o += __popcll(m);
    }
    if (!FILL && lane == 0) cnt[a] = o;
}

// C_ab per L entry, formed from the lower-numbered atom's side (the same bits
// for (a, b) and (b, a)): mesh pair plus the real-space coefficient of the
// listed pair (a binary search of the ascending full row).
template <int P>
__global__ void k_loc_coef(McDev D, long long n, const int* __restrict__ latom, const int* __restrict__ lown,
                           long long* __restrict__ lcoef) {
    const long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (e >= n) return;
    const int a = lown[e], b = latom[e];
    const int lo = a < b ? a : b, hi = a < b ? b : a;
    double c = mesh_pair<P>(D, lo, hi);
    int l = D.fhead[lo], h = D.fhead[lo + 1];
    while (l < h) {
        const int m = (l + h) >> 1;
        if (D.fnbr[m] < hi) l = m + 1;
        else h = m;
    }
    if (l < D.fhead[lo + 1] && D.fnbr[l] == hi) c += D.kreal[l];
    lcoef[e] = psi_k(c);
}

// T of every atom from the current charges: the L rows, plus the far part of the
// pool atoms whose charge differs from the freeze's (list chg, nchg entries).
// Exact integers: equal to what the chain and the folds keep.
__global__ void k_psil_all(McDev D, DaDev A, const int* __restrict__ list, int n, const int* __restrict__ chg,
                           int nchg) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const int a = list[k];
    long long s = 0;
    for (int e = A.lhead[a]; e < A.lhead[a + 1]; ++e) s += psi_c(D.q[A.latom[e]], A.lcoef[e]);
    const double xa = A.x[a], ya = A.y[a], za = D.z[a];
    for (int c = 0; c < nchg; ++c) {
        const int p = chg[c];
        s += far_delta(D, A, a, xa, ya, za, p, A.x[p], A.y[p], D.z[p], A.q0[p], D.q[p]);
    }
    A.psil[a] = s;
}

// far_a = W_a . phi + psi_real_a - q_a s_a - T_a at the freeze (the record's
// potential slot holds W_a . phi on entry; T_a holds only its L rows then).
__global__ void k_loc_far(McDev D, DaDev A, const int* __restrict__ list, int n) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const int a = list[k];
    double pr = 0.0;
    for (int e = D.fhead[a]; e < D.fhead[a + 1]; ++e) pr += D.q[D.fnbr[e]] * D.kreal[e];
    AtomRec& r = A.rec[a];
    r.phiw = ((r.phiw + pr) - D.q[a] * D.selfw[a]) - (double)A.psil[a] * PSI_INV;
}

// Adds the far part of a chain launch's accepted changes (pending entries
// fold_from .. pending_n, new charges in rqn) to T of every pool atom, four lanes
// per atom (entries dealt in turn, summed by shuffles).  Gate as the refresh's: a
// workgroup adds a list only if its own fapplied differs from pend_seq, then
// records pend_seq (a launch that accepted nothing leaves pend_seq).
constexpr int FOLD_LANES = 4, FOLD_THREADS = 256;

// Fold block blk, run by FOLD_THREADS threads (tid its thread); threads of a
// workgroup that are not active (no such block) only meet the barrier.
__device__ __forceinline__ void loc_fold_block(const McDev& D, const DaDev& A, const int* __restrict__ list, int n,
                                               int blk, int tid, bool active) {
    const McCtl* C = D.ctl;
    const unsigned seq = C->pend_seq;
    const bool run = active && A.fapplied[blk] != seq;
    const int gt = blk * FOLD_THREADS + tid, k = gt / FOLD_LANES, g = gt - k * FOLD_LANES;
    long long s = 0;
    int a = -1;
    if (run && k < n) {
        const int f0 = *A.fold_from, pn = C->pending_n;
        a = list[k];
        const double xa = A.x[a], ya = A.y[a], za = D.z[a];
        for (int e = f0 + g; e < pn; e += FOLD_LANES) {
            const int p = C->pend_atom[e];
            const double qn = A.rqn[e];
            s += far_delta(D, A, a, xa, ya, za, p, A.x[p], A.y[p], D.z[p], qn - C->pend_dq[e], qn);
        }
    }
    s += __shfl_xor(s, 1, FOLD_LANES);
    s += __shfl_xor(s, 2, FOLD_LANES);
    if (a >= 0 && g == 0 && s != 0) A.psil[a] += s;
    __syncthreads();
    if (run && tid == 0) A.fapplied[blk] = seq;
}

__global__ __launch_bounds__(FOLD_THREADS) void k_loc_fold(McDev D, DaDev A, const int* __restrict__ list, int n) {
    loc_fold_block(D, A, list, n, blockIdx.x, threadIdx.x, true);
}

// The fold and the direct potential refresh of the same accepted changes in one
// launch: workgroups below ntiles refresh listed tile tiles[w] (as k_refresh_direct),
// the others run two fold blocks each, one per half.  The fold reads charges,
// positions and the pending list; the refresh reads the pending list and writes
// the potential grid, so the two parts share nothing they write.
template <int P, int TY, int TZ>
__global__ __launch_bounds__(RD_THREADS) void k_loc_fold_refresh(McDev D, DaDev A, double* __restrict__ phi,
                                                                 float* __restrict__ phi32,
                                                                 const int* __restrict__ tiles, int ntiles,
                                                                 const int* __restrict__ list, int n, int nfold) {
    static_assert(RD_THREADS == 2 * FOLD_THREADS, "two fold blocks per refresh workgroup");
    if ((int)blockIdx.x < ntiles) {
        refresh_direct_tile_fixed<P, TY, TZ>(D, phi, phi32, tiles[blockIdx.x]);
        return;
    }
    const int blk = 2 * ((int)blockIdx.x - ntiles) + (int)(threadIdx.x / FOLD_THREADS);
    loc_fold_block(D, A, list, n, blk, threadIdx.x % FOLD_THREADS, blk < nfold);
}

__global__ void k_loc_fold_init(McDev D, DaDev A, int nblk) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < nblk) A.fapplied[k] = D.ctl->pend_seq;
}

// 1 in *flag when a pool atom's record holds another charge than the atom (the
// identities changed outside the delayed-acceptance chain).
__global__ void k_loc_dirty(McDev D, DaDev A, const int* __restrict__ list, int n, int* __restrict__ flag) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n && A.rec[list[k]].q != D.q[list[k]]) *flag = 1;
}

// The dynamic part of the pool atoms' records from the current identities.
__global__ void k_rec_dyn(McDev D, DaDev A, const int* __restrict__ list, int n) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const int a = list[k], ty = D.type[a];
    A.rec[a].q = D.q[a];
    A.rec[a].type = ty;
    A.rec[a].sp = A.species[ty];
}

// Physics part of n moves in the local-pair form (T current, far frozen).
template <int P>
__global__ void k_da_base_local(McDev D, DaDev A, int n, const int* __restrict__ mi, const int* __restrict__ mj,
                                const int* __restrict__ mf, double* __restrict__ base) {
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= n) return;
    const McCtl* C = D.ctl;
    const int i = mi[m], j = mj[m], fwd = mf[m];
    const double qi = D.q[i], qj = D.q[j];
    const double qin = fwd ? 1.0 : 0.0, qjn = fwd ? -1.0 : 0.0;
    const double dqi = qin - qi, dqj = qjn - qj;
    const long long K = local_pair_coef(D, A, i, j, A.x[i], A.y[i], D.z[i], A.x[j], A.y[j], D.z[j]);
    const double Pi = local_pot(A.rec[i].phiw, qi, D.selfw[i], A.psil[i]);
    const double Pj = local_pot(A.rec[j].phiw, qj, D.selfw[j], A.psil[j]);
    const double corr = corr_move(D, qi, qin, qj, qjn, D.z[i], D.z[j], C->Q, C->Mz, C->Qz2);
    base[m] = local_phys(dqi, Pi, dqj, Pj, local_selfq(dqi, D.selfw[i], dqj, D.selfw[j]), local_pair(dqi, dqj, K),
                         corr);
}

// ---------------------------------------------------------------- delayed-acceptance chain
// The accepted move's pool changes, O(1) per pool: the erased entry takes its
// pool's last entry (the delayed-acceptance chain's own pool order, which the
// oracle's restatement follows; the plain chain keeps the reference's
// erase-and-shift order), and the atom joins its partner pool at the end.
// Thread s < 2 handles the site pools (s = 0) or the ion pools (s = 1): each
// pair of partner pools shares one region of slots, so one thread orders its
// erase before its push.
template <typename PoolT>
__device__ __forceinline__ void da_pool_swap(const McDev& D, PoolT* pbuf, const DecisionLds& t, int tid) {
    if (tid < 2) {
        const int p = t.er_pool[tid], k = t.pos[tid], n = t.er_n[tid];
        const PoolT last = pbuf[pool_slot(D, p, n - 1)];
        pbuf[pool_slot(D, p, k)] = last;
        pbuf[pool_slot(D, t.pb_pool[tid], t.pb_n[tid])] = (PoolT)t.pb_val[tid];
    }
}

// The shell histograms after an accepted move, when the species a site's change
// touches and those an ion's change touches are disjoint (no histogram counter
// is written by both rows): both rows' read-modify-writes in one load phase and
// one store phase.  hrow: the rows' bounds in dlist (i's, then j's).
__device__ void da_hist_update(const DaDev& A, const int* hrow, int so_i, int sn_i, int so_j, int sn_j, int tid,
                               int bd) {
    const int hi = hrow[0], ni = so_i != sn_i ? hrow[1] - hrow[0] : 0;
    const int hj = hrow[2], nj = so_j != sn_j ? hrow[3] - hrow[2] : 0;
    const int NH = ni + nj;
    constexpr int MH = 4;
    for (int hb = 0; hb < NH; hb += MH * bd) {
        unsigned short* hp[MH];
        int ho[MH], hn[MH], so[MH], sn[MH];
#pragma unroll
        for (int m = 0; m < MH; ++m) {
            const int e = hb + tid + m * bd;
            hp[m] = nullptr;
            ho[m] = hn[m] = 0;
            so[m] = sn[m] = -1;
            if (e < NH) {
                const bool first = e < ni;
                const unsigned en = A.dlist[first ? hi + e : hj + (e - ni)];
                so[m] = first ? so_i : so_j;
                sn[m] = first ? sn_i : sn_j;
                hp[m] = A.H + (size_t)(en & 0xFFFFFFu) * NN_IN + (en >> DA_SHIFT) * 4;
                if (so[m] >= 0) ho[m] = hp[m][so[m]];
                if (sn[m] >= 0) hn[m] = hp[m][sn[m]];
            }
        }
#pragma unroll
        for (int m = 0; m < MH; ++m)
            if (hp[m]) {
                if (so[m] >= 0) hp[m][so[m]] = (unsigned short)(ho[m] - 1);
                if (sn[m] >= 0) hp[m][sn[m]] = (unsigned short)(hn[m] + 1);
            }
    }
}

// local_update and da_hist_update of an accepted move in one pass: the L-row heads
// (lrow) and the histogram rows' bounds (hrow) are known, so the L entries and the
// neighbour list are loaded together, then the counts; the T additions and the
// count stores follow (about three dependent memory trips instead of six).
__device__ __forceinline__ void local_hist_update(const DaDev& A, const int* lrow, const int* hrow, double qio,
                                                  double qin, double qjo, double qjn, int so_i, int sn_i, int so_j,
                                                  int sn_j, int tid, int bd) {
    constexpr int U = 6, MH = 4;
    const int hi = lrow[0], ni = lrow[1] - lrow[0], hj = lrow[2], nt = ni + (lrow[3] - lrow[2]);
    const int di = hrow[0], dni = so_i != sn_i ? hrow[1] - hrow[0] : 0;
    const int dj = hrow[2], NH = dni + (so_j != sn_j ? hrow[3] - hrow[2] : 0);
    for (int base = 0, hb = 0; base < nt || hb < NH; base += U * bd, hb += MH * bd) {
        long long K[U];
        int b[U];
        unsigned en[MH];
        bool hv[MH];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int e = base + tid + u * bd;
            b[u] = -1;
            K[u] = 0;
            if (e < nt) {
                const int x = e < ni ? hi + e : hj + (e - ni);
                K[u] = A.lcoef[x];
                b[u] = A.latom[x];
            }
        }
#pragma unroll
        for (int m = 0; m < MH; ++m) {
            const int e = hb + tid + m * bd;
            hv[m] = e < NH;
            en[m] = hv[m] ? A.dlist[e < dni ? di + e : dj + (e - dni)] : 0u;
        }
        unsigned short* hp[MH];
        int ho[MH], hn[MH], so[MH], sn[MH];
#pragma unroll
        for (int m = 0; m < MH; ++m) {
            const int e = hb + tid + m * bd;
            hp[m] = nullptr;
            ho[m] = hn[m] = 0;
            so[m] = sn[m] = -1;
            if (hv[m]) {
                const bool first = e < dni;
                so[m] = first ? so_i : so_j;
                sn[m] = first ? sn_i : sn_j;
                hp[m] = A.H + (size_t)(en[m] & 0xFFFFFFu) * NN_IN + (en[m] >> DA_SHIFT) * 4;
                if (so[m] >= 0) ho[m] = hp[m][so[m]];
                if (sn[m] >= 0) hn[m] = hp[m][sn[m]];
            }
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int e = base + tid + u * bd;
            if (b[u] >= 0) {
                const bool first = e < ni;
                const long long d = first ? psi_c(qin, K[u]) - psi_c(qio, K[u]) : psi_c(qjn, K[u]) - psi_c(qjo, K[u]);
                if (d != 0) atomicAdd(reinterpret_cast<unsigned long long*>(A.psil + b[u]), (unsigned long long)d);
            }
        }
#pragma unroll
        for (int m = 0; m < MH; ++m)
            if (hp[m]) {
                if (so[m] >= 0) hp[m][so[m]] = (unsigned short)(ho[m] - 1);
                if (sn[m] >= 0) hp[m][sn[m]] = (unsigned short)(hn[m] + 1);
            }
    }
}

// psi of the neighbours of i and j after an accepted move (integer atomics: the
// order of the additions does not change the sums).
template <int P>
__device__ __forceinline__ void psi_update(const McDev& D, const DaDev& A, const TrialT<P>& t, int tid, int bd) {
    const int ni = t.nrow[0], n = ni + t.nrow[1];
    for (int it = tid; it < n; it += bd) {
        const bool first = it < ni;
        const int e = first ? t.hrow[0] + it : t.hrow[1] + (it - ni);
        const long long K = psi_k(D.kreal[e]);
        const long long d = first ? psi_c(t.qi_new, K) - psi_c(t.qi_old, K) : psi_c(t.qj_new, K) - psi_c(t.qj_old, K);
        if (d != 0) atomicAdd(reinterpret_cast<unsigned long long*>(A.psi + D.fnbr[e]), (unsigned long long)d);
    }
}

struct CandLds {
    int i[DA_M], j[DA_M], fwd[DA_M], exA[DA_M], exB[DA_M];
    int pos[DA_M];                // stream position after the trial's draws
    int ki[DA_M], kj[DA_M];       // pool positions the proposal drew for the site and the ion
    double u1[DA_M], u2[DA_M], rij[DA_M], e1[DA_M], Fx[DA_M];
    double lin[DA_M][2];          // W_a . phi of the site and the ion
    MoveTerms mt[DA_M];
    long long pi[DA_M], pj[DA_M]; // psi (local-pair form: psil) of the site and the ion
    float sf[DA_M];               // s f(d)
    // local-pair form (pi, pj: T_i, T_j)
    double far[DA_M][2];          // far_i, far_j
    double cq[DA_M][2], cs[DA_M][2];             // charges, self terms
    double cx[DA_M][2], cy[DA_M][2], cz[DA_M][2];  // positions
    long long kq[DA_M];           // K_ij in fixed point
    int kin[DA_M];                // L(i) holds j: kq is found by a row search after the physics pass
};

// The chain of one replica.  W . phi form (LOCAL false): each launch ends after
// one accepted move, so that the refresh between launches keeps phi exact for
// the surrogate's potential term.  Local-pair form: the surrogate reads only the
// frozen field, the record and psil, which the chain updates itself, so a launch
// runs until its pending budget is full; the exact Delta-E takes the pending
// changes as the plain chain does.
template <int P, typename PoolT, bool LDS_POOLS, bool LOCAL>
__global__ __launch_bounds__(DA_THREADS) void k_da_chain(const McDev* __restrict__ Dp, const DaDev* __restrict__ Ap,
                                                         const NetDev* __restrict__ Np) {
    // the views live in device memory: read field by field instead of holding every field in registers
    const McDev& D = *Dp;
    const DaDev& A = *Ap;
    const NetDev& net = *Np;
    // network weights staged in LDS for the whole launch.  W1 rows are padded to
    // XP elements; the 8-byte chunks of W2's 256-byte rows are XOR-swizzled by
    // row (chunk c of row j at c ^ (2j mod 32)), so that the 16 rows one MFMA
    // B operand reads fall into distinct banks.
    __shared__ __align__(16) unsigned short s_w1t[NN_H * XP];
    __shared__ __align__(16) unsigned short s_w2t[NN_H * NN_H];
    __shared__ float s_b1[NN_H], s_b2[NN_H], s_w3[NN_H], s_scale[NN_IN];
    __shared__ RngLds s_rng;
    __shared__ TrialT<P> s_tr[1];
    __shared__ DecisionLds s_dc;
    __shared__ ChainState s_st;
    __shared__ CandLds s_c;
    __shared__ double s_wsum[DA_WAVES], s_wwin[DA_WAVES];
    __shared__ __align__(16) unsigned short s_X[DA_M * XP];
    __shared__ __align__(16) unsigned short s_H1[DA_M * HP];
    __shared__ float s_y[DA_M][8];
    __shared__ int s_first, s_start;
    __shared__ int s_defer;                // local-pair form: an accepted move's T and histogram updates are due
    __shared__ int s_hrow[4];              // dlist bounds of the staged trial's i and j
    __shared__ int s_lrow[4];              // local-pair form: L-row bounds of the staged trial's i and j
    __shared__ double s_kij;               // the staged pair's real-space coefficient (0: not listed)
    __shared__ unsigned long long s_msk;   // stage-1 acceptances of the round
    __shared__ long long s_done0;          // trials done at the round's start
    __shared__ long long s_s1, s_rounds;
    __shared__ long long s_prof[8];
    __shared__ float s_out[3];             // output bias, mean and scale of the network
    extern __shared__ __align__(16) unsigned char s_dyn[];
    PhaseClock clk;
    const bool prof = A.prof != 0 && threadIdx.x == 0;
    clk.start(prof);
    McCtl* C = D.ctl;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int M = A.M;
    // W . phi form: pending is empty for the whole launch and trial_partial never
    // reads this view; local-pair form: the changes accepted since the last refresh
    __shared__ PendT<LOCAL ? P : 1> s_pdl;
    // local-pair form: position and charges of each pending entry (the far part of the
    // entries this launch adds is not yet in T)
    constexpr int RK = LOCAL ? MC_KMAX : 1;
    __shared__ double s_rx[RK], s_ry[RK], s_rz[RK], s_rqo[RK], s_rqn[RK];
    const PendT<P>* pdp;
    if constexpr (LOCAL) pdp = &s_pdl;
    else pdp = reinterpret_cast<const PendT<P>*>(s_wsum);
    const PendT<P>& s_pd = *pdp;
    const int pend_max = C->pending_max;
    int pn0 = 0;   // local-pair form: changes pending (not in phi) at the launch's start
    PoolT* pbuf = LDS_POOLS ? reinterpret_cast<PoolT*>(s_dyn) : reinterpret_cast<PoolT*>(D.pbuf);
    long long target, rec_cap;
    {   // engine state and network into LDS: every load (with the control words the exit
        // checks read) issued before the first decision and the first store
        static_assert(NN_H * NN_IN / 4 == 3 * DA_THREADS && NN_H * NN_H / 4 == 8 * DA_THREADS && NN_H <= DA_THREADS,
                      "prologue layout");
        const uint2* w1 = reinterpret_cast<const uint2*>(net.w1t);
        const uint2* w2 = reinterpret_cast<const uint2*>(net.w2t);
        uint2 v1[3], v2[8];
        unsigned mt[2];
        float b1 = 0.f, b2 = 0.f, w3 = 0.f, sc = 0.f;
#pragma unroll
        for (int m = 0; m < 3; ++m) v1[m] = w1[tid + m * DA_THREADS];
#pragma unroll
        for (int m = 0; m < 8; ++m) v2[m] = w2[tid + m * DA_THREADS];
#pragma unroll
        for (int m = 0; m < 2; ++m) mt[m] = tid + m * DA_THREADS < 624 ? C->mt[tid + m * DA_THREADS] : 0u;
        if (tid < NN_H) {
            b1 = net.theta[OFF_B1 + tid];
            b2 = net.theta[OFF_B2 + tid];
            w3 = bf2f(net.w3[tid]);
        }
        if (tid < NN_IN) sc = net.theta[OFF_SCALE + tid];
        target = C->trials_target;
        rec_cap = C->rec_cap;
        const long long done_in = C->trials_done;
        const int status_in = C->status, pend_in = C->pending_n;
        const unsigned seq_in = C->pend_seq, applied_in = D.applied[0];
        // a launch queued beyond the target (batches are sized ahead) leaves at once
        if (done_in >= target || status_in != 0) return;
        if constexpr (LOCAL) {
            pn0 = applied_in == seq_in ? 0 : pend_in;
        } else if (applied_in != seq_in && pend_in > 0) {
            // phi lags the state: the surrogate's potential term would not be a function of the state
            if (tid == 0) C->status = 3;
            return;
        }
        if (LDS_POOLS)
            for (int k = tid; k < D.nsites + D.nions; k += blockDim.x) pbuf[k] = (PoolT)D.pbuf[k];
#pragma unroll
        for (int m = 0; m < 3; ++m) {   // 12 chunks per row
            const int k = tid + m * DA_THREADS, j = k / (NN_IN / 4), c = k - j * (NN_IN / 4);
            reinterpret_cast<uint2*>(s_w1t)[j * (XP / 4) + c] = v1[m];
        }
#pragma unroll
        for (int m = 0; m < 8; ++m) {   // 32 chunks per row
            const int k = tid + m * DA_THREADS, j = k >> 5, c = k & 31;
            reinterpret_cast<uint2*>(s_w2t)[j * 32 + (c ^ ((2 * j) & 31))] = v2[m];
        }
#pragma unroll
        for (int m = 0; m < 2; ++m) {
            const int k = tid + m * DA_THREADS;
            if (k < 624) {
                s_rng.cur[k] = mt[m];
                s_rng.out[0][k] = mt_temper(mt[m]);
            }
        }
        if (tid < NN_H) {
            s_b1[tid] = b1;
            s_b2[tid] = b2;
            s_w3[tid] = w3;
        }
        if (tid < NN_IN) s_scale[tid] = sc;
    }
    if constexpr (LOCAL)
        for (int k = wave; k < pn0; k += DA_WAVES) stage_pending<P>(D, s_pdl, k, C->pend_atom[k], C->pend_dq[k], lane, 64);
    if (tid == 0) {
        s_st.done = C->trials_done;
        s_st.acc_n = C->accepted; s_st.acc_f = C->acc_fwd; s_st.acc_b = C->acc_bwd;
        for (int k = 0; k < 4; ++k) s_st.sz[k] = C->sizes[k];
        s_st.status = C->status; s_st.nprot = C->nprot;
        s_st.blk = 0; s_st.pos = C->mti; s_st.pend_n = pn0;
        s_st.E = C->E; s_st.Q = C->Q; s_st.Q2 = C->Q2; s_st.Mz = C->Mz; s_st.Qz2 = C->Qz2;
        s_s1 = C->da_stage1; s_rounds = C->da_rounds;
        s_defer = 0;
        for (int k = 0; k < 8; ++k) s_prof[k] = C->da_prof[k];
    }
    // values used once per round live in LDS, not in registers across the round loop
    if (tid == 0) {
        s_out[0] = net.theta[OFF_B3];
        s_out[1] = net.theta[OFF_MEAN];
        s_out[2] = net.theta[OFF_STD];
    }
    __syncthreads();
    clk.lap(prof, s_prof, 0);
    for (;;) {
        // ---- 1. proposals (wave 0), each assumed rejected at stage 1; the other waves
        //         meanwhile apply the last accepted move's T-row and histogram updates
        if (LOCAL && wave != 0 && s_defer) {
            const TrialT<P>& t = s_tr[0];
            local_hist_update(A, s_lrow, s_hrow, t.qi_old, t.qi_new, t.qj_old, t.qj_new, A.species[t.ti_old],
                              A.species[t.ti_new], A.species[t.tj_old], A.species[t.tj_new], tid - 64,
                              DA_THREADS - 64);
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");   // complete before the barrier
        }
        if (wave == 0) {
            int blk = s_st.blk, pos = s_st.pos, status = s_st.status;
            const long long done = s_st.done;
            const int sz0 = s_st.sz[0], sz1 = s_st.sz[1], sz2 = s_st.sz[2], sz3 = s_st.sz[3];
            int ntr = 0, stop = 0;
            if (done >= target || status != 0 || (LOCAL ? s_st.pend_n + 2 > pend_max : s_st.pend_n > 0)) stop = 1;
            else {
                const long long wf = (long long)sz0 * sz2, wb = (long long)sz1 * sz3;
                if (wf == 0 && wb == 0) { status = 1; stop = 1; }
                else {
                    const long long room = target - done;
                    ntr = (int)(room < M ? room : M);
                    const double sum = (double)(wf + wb);
                    const bool both = wf > 0 && wb > 0;
                    // direction (when both are possible), site, ion, u1, u2: fixed per trial
                    const int per = (both ? 2 : 0) + 6;
                    rng_ensure(s_rng, blk, pos + ntr * per, lane);
                    bool ok = true;
                    int fwd = 0, ii = 0, jj = 0, ki = 0, kj = 0;
                    double u = 0.0, uu = 0.0;
                    if (lane < ntr) {
                        int q = pos + lane * per;
                        if (both) { fwd = canonical_at(s_rng, q) < (double)wf / sum ? 1 : 0; q += 2; }
                        else fwd = wf > 0 ? 1 : 0;
                        unsigned a1, a2;
                        ok = index_at(s_rng, q, (unsigned)(fwd ? sz0 : sz1), a1);
                        ok = index_at(s_rng, q + 1, (unsigned)(fwd ? sz2 : sz3), a2) && ok;
                        ki = (int)a1;
                        kj = (int)a2;
                        ii = (int)pbuf[pool_slot(D, fwd ? 0 : 1, ki)];
                        jj = (int)pbuf[pool_slot(D, fwd ? 2 : 3, kj)];
                        u = canonical_at(s_rng, q + 2);
                        uu = canonical_at(s_rng, q + 4);
                    }
                    if (__all(ok)) {
                        if (lane < ntr) {
                            s_c.i[lane] = ii; s_c.j[lane] = jj; s_c.fwd[lane] = fwd;
                            s_c.ki[lane] = ki; s_c.kj[lane] = kj;
                            s_c.pos[lane] = pos + (lane + 1) * per; s_c.u1[lane] = u; s_c.u2[lane] = uu;
                        }
                        pos += ntr * per;
                    } else {
                        for (int k = 0; k < ntr; ++k) {
                            int f2;
                            if (both) f2 = draw_canonical(s_rng, blk, pos, lane) < (double)wf / sum ? 1 : 0;
                            else f2 = wf > 0 ? 1 : 0;
                            const int ki2 = (int)draw_index(s_rng, blk, pos, lane, (unsigned)(f2 ? sz0 : sz1));
                            const int kj2 = (int)draw_index(s_rng, blk, pos, lane, (unsigned)(f2 ? sz2 : sz3));
                            const int i2 = (int)pbuf[pool_slot(D, f2 ? 0 : 1, ki2)];
                            const int j2 = (int)pbuf[pool_slot(D, f2 ? 2 : 3, kj2)];
                            const double v1 = draw_canonical(s_rng, blk, pos, lane);
                            const double v2 = draw_canonical(s_rng, blk, pos, lane);
                            if (lane == 0) {
                                s_c.i[k] = i2; s_c.j[k] = j2; s_c.fwd[k] = f2; s_c.pos[k] = pos;
                                s_c.u1[k] = v1; s_c.u2[k] = v2;
                                s_c.ki[k] = ki2; s_c.kj[k] = kj2;
                            }
                        }
                    }
                }
            }
            if (lane == 0) {
                s_dc.ntr = ntr; s_dc.stop = stop;
                s_st.blk = blk; s_st.pos = pos; s_st.status = status;
            }
        }
        __syncthreads();
        clk.lap(prof, s_prof, 1);
        if (LOCAL && tid == 0) s_defer = 0;   // every reader passed the barrier
        if (s_dc.stop) break;
        const int ntr = s_dc.ntr;
        // ---- 2. per-candidate physics terms (W . phi from the per-atom table) and the raw
        //         shell-count differences
        // (the histogram rows' loads are issued first; the records' follow)
        int raw[DA_FEAT];
        // the thread's feature indices, formed each round from an opaque copy of tid so
        // that they are not hoisted out of the round loop and held in registers
        int tid_r = tid;
        asm volatile("" : "+v"(tid_r));
#pragma unroll
        for (int q = 0; q < DA_FEAT; ++q) {
            const int e = tid_r + q * DA_THREADS, k = e / NN_IN, c = e - k * NN_IN;
            raw[q] = 0;
            if (e < DA_M * NN_IN && k < ntr)
                raw[q] = (int)A.H[(size_t)s_c.i[k] * NN_IN + c] - (int)A.H[(size_t)s_c.j[k] * NN_IN + c];
        }
        if (tid < ntr) {
            double r;
            int exA, exB;
            const int ci = s_c.i[tid], cj = s_c.j[tid];
            const AtomRec ri = A.rec[ci], rj = A.rec[cj];
            if constexpr (LOCAL) {
                const long long pi = load_T(A, ci), pj = load_T(A, cj);
                pair_terms_rec(A, ri, rj, r, exA, exB);
                s_c.rij[tid] = r; s_c.exA[tid] = exA; s_c.exB[tid] = exB;
                s_c.pi[tid] = pi; s_c.pj[tid] = pj;   // the unfolded changes are added below
                double dx, dy, dz;
                const double r2 = pair_disp(A, ri.x, ri.y, ri.z, rj.x, rj.y, rj.z, dx, dy, dz);
                s_c.kin[tid] = local_member(A, r2, dx, dy) ? 1 : 0;   // K_ij after the barrier
                s_c.far[tid][0] = ri.phiw; s_c.far[tid][1] = rj.phiw;
                s_c.cq[tid][0] = ri.q; s_c.cq[tid][1] = rj.q;
                s_c.cs[tid][0] = ri.selfw; s_c.cs[tid][1] = rj.selfw;
                s_c.cx[tid][0] = ri.x; s_c.cx[tid][1] = rj.x;
                s_c.cy[tid][0] = ri.y; s_c.cy[tid][1] = rj.y;
                s_c.cz[tid][0] = ri.z; s_c.cz[tid][1] = rj.z;
            } else {
                const long long pi = A.use_psi ? A.psi[ci] : 0, pj = A.use_psi ? A.psi[cj] : 0;
                pair_terms_rec(A, ri, rj, r, exA, exB);
                s_c.rij[tid] = r; s_c.exA[tid] = exA; s_c.exB[tid] = exB;
                s_c.pi[tid] = pi; s_c.pj[tid] = pj;
                s_c.mt[tid] = move_terms_rec(D, A, ri, rj, s_c.fwd[tid], r, s_st.Q, s_st.Mz, s_st.Qz2, pi, pj);
                s_c.lin[tid][0] = ri.phiw;
                s_c.lin[tid][1] = rj.phiw;
            }
        }
        __syncthreads();
        // formed again (not kept from the loop above across the records' part)
        int tid_s = tid;
        asm volatile("" : "+v"(tid_s));
#pragma unroll
        for (int q = 0; q < DA_FEAT; ++q) {
            const int e = tid_s + q * DA_THREADS, k = e / NN_IN, c = e - k * NN_IN;
            if (e < DA_M * NN_IN) {
                float v = 0.f;
                if (k < ntr) {
                    const int d = raw[q] - (c == s_c.exA[k] ? 1 : 0) + (c == s_c.exB[k] ? 1 : 0);
                    v = (float)d * s_scale[c];
                }
                s_X[k * XP + c] = f2bf(v);
            }
        }
        for (int k = tid; k < DA_M * 8; k += DA_THREADS) (&s_y[0][0])[k] = 0.f;
        if constexpr (LOCAL) {
            // K_ij: Gt_ij (one thread per candidate), or C_ij when L holds the pair (one
            // wave per such candidate)
            if (tid >= DA_THREADS - DA_M) {
                const int c = tid - (DA_THREADS - DA_M);
                if (c < ntr && !s_c.kin[c]) {
                    double dx, dy, dz;
                    pair_disp(A, s_c.cx[c][0], s_c.cy[c][0], s_c.cz[c][0], s_c.cx[c][1], s_c.cy[c][1], s_c.cz[c][1],
                              dx, dy, dz);
                    s_c.kq[c] = far_coef(D, A, dx, dy, dz);
                }
            }
            for (int c = wave; c < ntr; c += DA_WAVES)
                if (s_c.kin[c]) {
                    const long long K = wave_find_coef(A, s_c.i[c], s_c.j[c], lane);
                    if (lane == 0) s_c.kq[c] = K;
                }
            // T of the site and the ion: the far part of this launch's accepted changes
            const int nrec = s_st.pend_n - pn0;
            for (int idx = tid; idx < ntr * 2 * nrec; idx += DA_THREADS) {
                const int c = idx / (2 * nrec), rem = idx - c * 2 * nrec, side = rem / nrec;
                const int e = pn0 + (rem - side * nrec);
                const int a = side ? s_c.j[c] : s_c.i[c];
                const long long d = far_delta(D, A, a, s_c.cx[c][side], s_c.cy[c][side], s_c.cz[c][side],
                                              s_pdl.atom[e], s_rx[e], s_ry[e], s_rz[e], s_rqo[e], s_rqn[e]);
                if (d != 0)
                    atomicAdd(reinterpret_cast<unsigned long long*>(side ? &s_c.pj[c] : &s_c.pi[c]),
                              (unsigned long long)d);
            }
        }
        __syncthreads();
        clk.lap(prof, s_prof, 2);
        // ---- 3. MLP on MFMA (the steps of layer1/layer2): wave w owns the 16 candidates of row
        //      block w & 3 and the column tiles (w >> 2) + 2u, u = 0..3, all four in flight at once
        {
            const int rr = lane & 15, g = lane >> 4, r0 = (wave & 3) * 16, c0 = wave >> 2;
            {
                s4 a0[3];
#pragma unroll
                for (int kb = 0; kb < 3; ++kb) a0[kb] = ld4(s_X + (r0 + rr) * XP + kb * 16 + 4 * g);
                f4 acc[4];
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    const unsigned short* wr = s_w1t + ((c0 + 2 * u) * 16 + rr) * XP + 4 * g;
                    acc[u] = f4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
                    for (int kb = 0; kb < 3; ++kb) acc[u] = mfma16(a0[kb], ld4(wr + kb * 16), acc[u]);
                }
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    const int ct = c0 + 2 * u;
                    const float b = s_b1[ct * 16 + rr];
#pragma unroll
                    for (int e = 0; e < 4; ++e)
                        s_H1[(r0 + 4 * g + e) * HP + ct * 16 + rr] = f2bf(gelu_f(acc[u][e] + b));
                }
            }
            __syncthreads();
            {
                bf8 a[4];
#pragma unroll
                for (int kb = 0; kb < 4; ++kb) a[kb] = ld8(s_H1 + (r0 + rr) * HP + 32 * kb + 8 * g);
                f4 acc[4];
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    const int row = (c0 + 2 * u) * 16 + rr, swz = (2 * row) & 31;
                    const unsigned short* wr = s_w2t + row * NN_H;
                    acc[u] = f4{0.f, 0.f, 0.f, 0.f};
                    // k 32kb + 8g .. +7 are the swizzled chunks 8kb + 2g and 8kb + 2g + 1: swz is
                    // even, so they stay adjacent (one 16-byte load)
#pragma unroll
                    for (int kb = 0; kb < 4; ++kb)
                        acc[u] = mfma32(a[kb], ld8(wr + (((8 * kb + 2 * g) ^ swz) << 2)), acc[u]);
                }
                // the row's four tiles summed in registers, one reduction into this wave's
                // slot s_y[row][wave] (the other slots stay zero)
                float yp[4] = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    const int row = (c0 + 2 * u) * 16 + rr;
                    const float b = s_b2[row], w = s_w3[row];
#pragma unroll
                    for (int e = 0; e < 4; ++e) yp[e] += bf2f(f2bf(gelu_f(acc[u][e] + b))) * w;
                }
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    const float v = row_sum16(yp[e]);
                    if (rr == 0) s_y[r0 + 4 * g + e][wave] = v;
                }
            }
        }
        __syncthreads();
        clk.lap(prof, s_prof, 3);
        // ---- 4. stage-1 decisions of the round (wave 0)
        if (wave == 0) {
            const long long a = s_st.sz[0], b = s_st.sz[1], c = s_st.sz[2], d = s_st.sz[3];
            const long long Sb = a * c + b * d;
            const int k = lane;
            bool acc1 = false;
            if (k < ntr) {
                float y = s_out[0];
#pragma unroll
                for (int ct = 0; ct < 8; ++ct) y += s_y[k][ct];
                const float f = s_out[1] + s_out[2] * y;
                const int fwd = s_c.fwd[k];
                const float sf = fwd ? f : -f;
                double Fx;
                if constexpr (LOCAL) {
                    const double qi = s_c.cq[k][0], qj = s_c.cq[k][1];
                    const double qin = fwd ? 1.0 : 0.0, qjn = fwd ? -1.0 : 0.0;
                    const double dqi = qin - qi, dqj = qjn - qj;
                    const double Pi = local_pot(s_c.far[k][0], qi, s_c.cs[k][0], s_c.pi[k]);
                    const double Pj = local_pot(s_c.far[k][1], qj, s_c.cs[k][1], s_c.pj[k]);
                    const double corr = corr_move(D, qi, qin, qj, qjn, s_c.cz[k][0], s_c.cz[k][1], s_st.Q, s_st.Mz,
                                                  s_st.Qz2);
                    Fx = local_phys(dqi, Pi, dqj, Pj, local_selfq(dqi, s_c.cs[k][0], dqj, s_c.cs[k][1]),
                                    local_pair(dqi, dqj, s_c.kq[k]), corr) +
                         (double)sf;
                } else {
                    const MoveTerms& m = s_c.mt[k];
                    Fx = m.dqi * s_c.lin[k][0] + m.dqj * s_c.lin[k][1] + m.selfq + m.perfc + m.perf + m.corr + m.psi +
                         (double)sf;
                }
                const long long Sa = fwd ? (a - 1) * (c - 1) + (b + 1) * (d + 1) : (a + 1) * (c + 1) + (b - 1) * (d - 1);
                const double e1 = mh_exponent(D.beta, Fx, fwd ? 1 : -1, D.mu, Sb, Sa);
                acc1 = e1 >= 0.0 || exp(e1) > s_c.u1[k];
                s_c.sf[k] = sf;
                s_c.e1[k] = e1;
                s_c.Fx[k] = Fx;
            }
            const unsigned long long msk = __ballot(acc1);
            if (lane == 0) {
                s_msk = msk;
                s_start = 0;
                s_done0 = s_st.done;
                ++s_rounds;
            }
        }
        __syncthreads();
        clk.lap(prof, s_prof, 4);
        // ---- 5./6. the round's stage-1 acceptances in order, until stage 2 accepts one
        //         (a stage-2 rejection leaves the state, so the later trials stand)
        for (;;) {
            if (wave == 0) {
                const int start = s_start;
                const unsigned long long m = start < 64 ? (s_msk & (~0ull << start)) : 0ull;
                const int first = m ? __ffsll((long long)m) - 1 : ntr;
                const long long done0 = s_done0;
                const int k = lane;
                if (k >= start && k < first && k < ntr && done0 + k < rec_cap) {
                    const double nan = __longlong_as_double(0x7FF8000000000000LL);
                    D.rec_fwd[done0 + k] = (signed char)s_c.fwd[k]; D.rec_i[done0 + k] = s_c.i[k];
                    D.rec_j[done0 + k] = s_c.j[k]; D.rec_acc[done0 + k] = 0; D.rec_dE[done0 + k] = nan;
                    D.rec_u[done0 + k] = s_c.Fx[k]; A.rec_fy[done0 + k] = nan;
                }
                if (lane == 0) {
                    s_first = first;
                    if (first >= ntr) s_st.done = done0 + ntr;
                    else {
                        ++s_s1;
                        const int i1 = s_c.i[first], j1 = s_c.j[first];
                        stage_trial<P>(D, s_tr[0], i1, j1, s_c.fwd[first]);
                        s_hrow[0] = A.dhead[i1]; s_hrow[1] = A.dhead[i1 + 1];
                        s_hrow[2] = A.dhead[j1]; s_hrow[3] = A.dhead[j1 + 1];
                        if constexpr (LOCAL) {
                            s_lrow[0] = A.lhead[i1]; s_lrow[1] = A.lhead[i1 + 1];
                            s_lrow[2] = A.lhead[j1]; s_lrow[3] = A.lhead[j1 + 1];
                        }
                        s_kij = 0.0;
                    }
                }
            }
            __syncthreads();
            const int first = s_first;
            if (first >= ntr) break;
            // ---- 5. exact Delta-E of the stage-1 acceptance (all waves); its mesh-kernel pair part apart
            stage_trial_data<P>(D, s_tr[0], tid, DA_THREADS);
            __syncthreads();
            {
                double win = 0.0;
                const double part = wave_sum(
                    trial_partial<P>(D, s_tr[0], s_pd, LOCAL ? s_st.pend_n : 0, tid, DA_THREADS, &win, &s_kij));
                win = wave_sum(win);
                if (lane == 0) { s_wsum[wave] = part; s_wwin[wave] = win; }
            }
            __syncthreads();
            clk.lap(prof, s_prof, 5);
            if (wave == 0 && lane == 0) {
                const TrialT<P>& t = s_tr[0];
                const long long done0 = s_done0;
                const double Q = s_st.Q, Q2 = s_st.Q2, Mz = s_st.Mz, Qz2 = s_st.Qz2;
                const long long a = s_st.sz[0], b = s_st.sz[1], c = s_st.sz[2], d = s_st.sz[3];
                const long long Sb = a * c + b * d;
                double dsum = 0.0, wsum = 0.0;
                for (int w = 0; w < DA_WAVES; ++w) { dsum += s_wsum[w]; wsum += s_wwin[w]; }
                const double dE = dsum + correction_delta<P>(D, t, Q, Q2, Mz, Qz2);
                const int fwd = t.fwd, sg = fwd ? 1 : -1;
                const long long Sa = fwd ? (a - 1) * (c - 1) + (b + 1) * (d + 1) : (a + 1) * (c + 1) + (b - 1) * (d - 1);
                const double e1 = s_c.e1[first];
                // the surrogate of the reverse move from the proposed state: the potential
                // there is phi + G dQ, so W_a . phi_y = W_a . phi + sum_b dq_b W_a^T G W_b
                const MoveTerms& m = s_c.mt[first];
                const double lin = m.dqi * s_c.lin[first][0] + m.dqj * s_c.lin[first][1];
                double Fy;
                if constexpr (LOCAL) {
                    // the chain's own stage-1 value of the reverse move at y: psil of i and j at y
                    // (the pair's change when j is in L(i)), the records' charges at y, the
                    // moments as the update below forms them
                    const long long K = s_c.kq[first];
                    const long long piy = s_c.pi[first] + (psi_c(t.qj_new, K) - psi_c(t.qj_old, K));
                    const long long pjy = s_c.pj[first] + (psi_c(t.qi_new, K) - psi_c(t.qi_old, K));
                    const double dqi = t.qi_new - t.qi_old, dqj = t.qj_new - t.qj_old;
                    const double Qy = Q + (dqi + dqj);
                    const double Mzy = Mz + (dqi * t.zi + dqj * t.zj);
                    const double Qz2y = Qz2 + (dqi * t.zi * t.zi + dqj * t.zj * t.zj);
                    const double rdqi = t.qi_old - t.qi_new, rdqj = t.qj_old - t.qj_new;
                    const double Pi = local_pot(s_c.far[first][0], t.qi_new, t.self[0], piy);
                    const double Pj = local_pot(s_c.far[first][1], t.qj_new, t.self[1], pjy);
                    const double corr = corr_move(D, t.qi_new, t.qi_old, t.qj_new, t.qj_old, t.zi, t.zj, Qy, Mzy, Qz2y);
                    Fy = local_phys(rdqi, Pi, rdqj, Pj, local_selfq(rdqi, t.self[0], rdqj, t.self[1]),
                                    local_pair(rdqi, rdqj, K), corr) +
                         (double)(-s_c.sf[first]);
                } else if (A.use_psi) {
                    // its real-space term from psi at y, formed in fixed point as the table will hold it;
                    // the pair term dq_i dq_j k_ij is even in the move
                    const long long K = psi_k(s_kij);
                    const long long piy = s_c.pi[first] + (psi_c(t.qj_new, K) - psi_c(t.qj_old, K));
                    const long long pjy = s_c.pj[first] + (psi_c(t.qi_new, K) - psi_c(t.qi_old, K));
                    Fy = -lin - m.selfq - 2.0 * wsum + m.perfc + m.perf - m.corr - (double)s_c.sf[first] +
                         psi_term(-m.dqi, piy, -m.dqj, pjy);
                } else {   // the pair term (q_i' q_j' - q_i q_j) k_ij is odd in the move
                    Fy = -lin - m.selfq - 2.0 * wsum - m.perfc + m.perf - m.corr - (double)s_c.sf[first];
                }
                double e2;
                {
#pragma clang fp contract(off)
                    double logH = 0.0;
                    if (Sb > 0 && Sa > 0) logH = log((double)Sb) - log((double)Sa);
                    double r1 = (double)((-D.beta) * (Fy + (double)sg * D.mu)) - logH;
                    r1 = fmin(fmax(r1, -40.0), 40.0);
                    e2 = (((double)((-D.beta) * (dE - (double)sg * D.mu)) + logH) + fmin(0.0, r1)) - fmin(0.0, e1);
                }
                const bool acc2 = e2 >= 0.0 || exp(e2) > s_c.u2[first];
                const long long tk = done0 + first;
                if (tk < rec_cap) {
                    D.rec_fwd[tk] = (signed char)fwd; D.rec_i[tk] = t.i; D.rec_j[tk] = t.j;
                    D.rec_acc[tk] = acc2 ? 1 : 2; D.rec_dE[tk] = dE; D.rec_u[tk] = s_c.Fx[first]; A.rec_fy[tk] = Fy;
                }
                s_st.done = tk + 1;
                s_start = first + 1;
                s_dc.acc_k = -1;
                if (acc2) {
                    // the trials after this one are proposed again from the new state:
                    // the stream resumes right after this trial's draws
                    const int pos = s_c.pos[first];
                    if (pos < s_st.blk * 624 - 624) s_st.status = 2;
                    else s_st.pos = pos;
                    const double dqi = t.qi_new - t.qi_old, dqj = t.qj_new - t.qj_old;
                    s_st.E += dE;
                    s_st.Q2 = Q2 + ((t.qi_new * t.qi_new - t.qi_old * t.qi_old) + (t.qj_new * t.qj_new - t.qj_old * t.qj_old));
                    s_st.Q = Q + (dqi + dqj);
                    s_st.Mz = Mz + (dqi * t.zi + dqj * t.zj);
                    s_st.Qz2 = Qz2 + (dqi * t.zi * t.zi + dqj * t.zj * t.zj);
                    ++s_st.acc_n;
                    if (fwd) { ++s_st.acc_f; ++s_st.nprot; } else { ++s_st.acc_b; --s_st.nprot; }
                    D.q[t.i] = t.qi_new; D.type[t.i] = t.ti_new;
                    D.q[t.j] = t.qj_new; D.type[t.j] = t.tj_new;
                    const int ps = fwd ? 0 : 1, pdd = fwd ? 1 : 0, is = fwd ? 2 : 3, id = fwd ? 3 : 2;
                    s_dc.er_pool[0] = ps; s_dc.er_val[0] = t.i; s_dc.er_n[0] = s_st.sz[ps];
                    s_dc.er_pool[1] = is; s_dc.er_val[1] = t.j; s_dc.er_n[1] = s_st.sz[is];
                    s_dc.pb_pool[0] = pdd; s_dc.pb_val[0] = t.i; s_dc.pb_n[0] = s_st.sz[pdd];
                    s_dc.pb_pool[1] = id; s_dc.pb_val[1] = t.j; s_dc.pb_n[1] = s_st.sz[id];
                    s_dc.pos[0] = s_c.ki[first]; s_dc.pos[1] = s_c.kj[first];
                    if (fwd) { --s_st.sz[0]; ++s_st.sz[1]; --s_st.sz[2]; ++s_st.sz[3]; }
                    else { ++s_st.sz[0]; --s_st.sz[1]; ++s_st.sz[2]; --s_st.sz[3]; }
                    s_dc.acc_k = 0;
                    if (!LOCAL) s_st.pend_n = 2;   // ends the launch: the refresh folds the move into phi
                }
            }
            __syncthreads();
            const bool accepted = s_dc.acc_k == 0;
            // ---- 6. accepted: pools and the histograms of the atoms around i and j
            if (accepted) {
                da_pool_swap<PoolT>(D, pbuf, s_dc, tid);
                if constexpr (LOCAL) {
                    // the move joins the pending list; the records and psil of the atoms
                    // within r_local of i, then of j (a row has no repeated atom)
                    const TrialT<P>& t = s_tr[0];
                    const int pn = s_st.pend_n;
                    if (wave == 1) stage_pending<P>(D, s_pdl, pn, t.i, t.qi_new - t.qi_old, lane, 64);
                    if (wave == 2) stage_pending<P>(D, s_pdl, pn + 1, t.j, t.qj_new - t.qj_old, lane, 64);
                    if (tid == 64 || tid == 65) {
                        const int side = tid - 64, a = side ? t.j : t.i, e = pn + side;
                        s_rx[e] = A.x[a]; s_ry[e] = A.y[a]; s_rz[e] = side ? t.zj : t.zi;
                        s_rqo[e] = side ? t.qj_old : t.qi_old;
                        s_rqn[e] = side ? t.qj_new : t.qi_new;
                    }
                    if (tid == 0) {
                        A.rec[t.i].q = t.qi_new; A.rec[t.i].type = t.ti_new; A.rec[t.i].sp = A.species[t.ti_new];
                    } else if (tid == 1) {
                        A.rec[t.j].q = t.qj_new; A.rec[t.j].type = t.tj_new; A.rec[t.j].sp = A.species[t.tj_new];
                    }
                    if (!A.upd_one) {
                        local_update(A, t.i, t.qi_old, t.qi_new, t.j, t.qj_old, t.qj_new, tid, DA_THREADS);
                        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");   // the atomics complete before the barrier
                    }
                } else if (A.use_psi) {
                    psi_update<P>(D, A, s_tr[0], tid, DA_THREADS);
                }
            }
            if (accepted && A.upd_one && LOCAL) {
                // T rows and histograms: waves 1-7 during the next round's proposals
                if (tid == 0) s_defer = 1;
            } else if (accepted && A.upd_one) {
                const TrialT<P>& t = s_tr[0];
                da_hist_update(A, s_hrow, A.species[t.ti_old], A.species[t.ti_new], A.species[t.tj_old],
                               A.species[t.tj_new], tid, DA_THREADS);
            } else if (accepted) {
                const TrialT<P>& t = s_tr[0];
                for (int side = 0; side < 2; ++side) {
                    const int atom = side ? t.j : t.i;
                    const int so = A.species[side ? t.tj_old : t.ti_old], sn = A.species[side ? t.tj_new : t.ti_new];
                    if (so != sn)
                        for (int e = A.dhead[atom] + tid; e < A.dhead[atom + 1]; e += DA_THREADS) {
                            const unsigned v = A.dlist[e];
                            unsigned short* h = A.H + (size_t)(v & 0xFFFFFFu) * NN_IN + (v >> DA_SHIFT) * 4;
                            if (so >= 0) h[so] = (unsigned short)(h[so] - 1);
                            if (sn >= 0) h[sn] = (unsigned short)(h[sn] + 1);
                        }
                    __syncthreads();
                }
            }
            if (accepted) __syncthreads();   // the pool swap (LDS pools are copied out below)
            if (LOCAL && accepted && tid == 0) s_st.pend_n += 2;
            clk.lap(prof, s_prof, 6);
            if (accepted) break;
        }
    }
    // write back
    if (LDS_POOLS)
        for (int k = tid; k < D.nsites + D.nions; k += blockDim.x) D.pbuf[k] = (int)pbuf[k];
    if (wave == 0) {
        const int blk = s_st.blk, pos = s_st.pos;
        const bool in_cur = pos >= blk * 624;
        const unsigned* st = in_cur ? s_rng.cur : s_rng.prev;
        for (int k = lane; k < 624; k += 64) C->mt[k] = st[k];
        if (lane == 0) {
            C->mti = in_cur ? pos - blk * 624 : pos - (blk - 1) * 624;
            for (int k = 0; k < 4; ++k) C->sizes[k] = s_st.sz[k];
            C->trials_done = s_st.done;
            C->accepted = s_st.acc_n; C->acc_fwd = s_st.acc_f; C->acc_bwd = s_st.acc_b;
            C->status = s_st.status; C->nprot = s_st.nprot;
            C->E = s_st.E; C->Q = s_st.Q; C->Q2 = s_st.Q2; C->Mz = s_st.Mz; C->Qz2 = s_st.Qz2;
            C->da_stage1 = s_s1; C->da_rounds = s_rounds;
            if (!LOCAL && s_st.pend_n > 0) {
                const TrialT<P>& t = s_tr[0];
                C->pend_atom[0] = t.i; C->pend_dq[0] = t.qi_new - t.qi_old;
                C->pend_atom[1] = t.j; C->pend_dq[1] = t.qj_new - t.qj_old;
                C->pending_n = 2;
                C->pend_seq = C->pend_seq + 1u;   // a new list: no workgroup has folded it
            }
        }
    }
    if constexpr (LOCAL) {
        const int pend_n = s_st.pend_n;
        if (tid == 0) {
            if (pend_n > pn0) C->pend_seq = C->pend_seq + 1u;   // a new list: no workgroup has folded it
            C->pending_n = pend_n;
            *A.fold_from = pn0;
        }
        for (int k = tid; k < pend_n; k += DA_THREADS) { C->pend_atom[k] = s_pdl.atom[k]; C->pend_dq[k] = s_pdl.dq[k]; }
        for (int k = pn0 + tid; k < pend_n; k += DA_THREADS) A.rqn[k] = s_rqn[k];
    }
    if (prof) {
        clk.lap(prof, s_prof, 7);
        for (int k = 0; k < 8; ++k) C->da_prof[k] = s_prof[k];
    }
}

}  // namespace phb

using namespace phb;

// ----------------------------------------------------------------------- host
struct phb_surrogate_s {
    DevBuf<float> theta, m, v, vpart, grad, loss_hist;
    DevBuf<unsigned short> act;   // transposed training activations of one batch: X^T, H1^T, DZ1^T, DZ2^T
    DevBuf<unsigned short> w1t, w2t, w2, w3;
    DevBuf<short> dd;          // dataset descriptors [cap][48]
    DevBuf<float> dt;          // dataset targets [cap]
    DevBuf<int> idx;
    long long cap = 0, n = 0, steps = 0;
    int rows_cap = 0;
    NetDev net() const { return NetDev{w1t.p, w2t.p, w2.p, w3.p, theta.p}; }
    void cast() {
        hipLaunchKernelGGL(k_mlp_cast, dim3(ceil_div(NN_TRAIN, 256)), dim3(256), 0, stream(), theta.p, w1t.p, w2t.p,
                           w2.p, w3.p);
        PHB_HIP(hipGetLastError());
    }
    // buffers of a training batch of `rows` rows (a multiple of 64)
    TrainBufs bufs(int rows) {
        if (rows > rows_cap) {
            act.alloc((size_t)(NN_IN + 3 * NN_H) * rows);
            vpart.alloc((size_t)(rows / 64) * VP);
            rows_cap = rows;
        }
        unsigned short* a = act.p;
        return TrainBufs{a, a + (size_t)NN_IN * rows, a + (size_t)(NN_IN + NN_H) * rows,
                         a + (size_t)(NN_IN + 2 * NN_H) * rows, vpart.p, rows};
    }
    // gradient of one batch into grad (W1, W2 from the tile products, the rest and the loss sum from
    // the workgroups' vector sums)
    void batch_grad(const short* dd_, const float* dt_, const int* idx_, int rows) {
        const TrainBufs tb = bufs(rows);
        const int nb = rows / 64;
        static bool lds_set = false;
        if (!lds_set) {
            PHB_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(k_mlp_act),
                                        hipFuncAttributeMaxDynamicSharedMemorySize, (int)ACT_WLDS));
            lds_set = true;
        }
        hipLaunchKernelGGL(k_mlp_act, dim3(nb), dim3(256), ACT_WLDS, stream(), net(), dd_, dt_, idx_, 0LL, rows, tb);
        hipLaunchKernelGGL(k_mlp_wgrad, dim3(WG_TILES + 1), dim3(WG_THREADS), 0, stream(), tb, nb, grad.p);
        PHB_HIP(hipGetLastError());
    }
};

namespace {

struct DaState {
    DevBuf<McDev> dmc;         // device copies of the chain's views (kernel reads them by pointer)
    DevBuf<DaDev> dda;
    DevBuf<NetDev> dnet;
    phb_da_desc d{};
    std::vector<signed char> species;
    DevBuf<signed char> sp;
    DevBuf<int> dhead;
    DevBuf<unsigned> dlist;
    DevBuf<unsigned short> H;
    DevBuf<double> x, y;       // copies of the positions (fixed)
    DevBuf<double> rec_fy;
    DevBuf<AtomRec> rec;       // per atom: what the chain's physics terms read
    DevBuf<long long> psi;     // per atom: real-space potential of the listed neighbours (fixed point)
    DevBuf<int> cand;          // the pool atoms, in stencil-origin order
    int ncand = 0;
    // local-pair form: L rows, coefficients, psil; far frozen in the records' potential slot
    DevBuf<int> lhead, latom;
    DevBuf<long long> lcoef, psil;
    DevBuf<double> q0, rqn, gs;
    DevBuf<int> fold_from, flag;
    DevBuf<unsigned> fapplied;
    long long nlocal = 0;
    bool frozen = false;
    DaDev dev{};
    long long nentries = 0;
    long long launches = 0, trials_seen = 0;   // batch sizing of phb_da_run
};

void da_free(void* p) { delete static_cast<DaState*>(p); }

DaState* da_of(phb_mc_s* M) {
    require(M->da != nullptr, "delayed acceptance: call phb_da_attach first", PHB_ERR_STATE);
    return static_cast<DaState*>(M->da);
}

void rebuild_hist(phb_mc_s* M, DaState* S) {
    const int N = M->atoms->n;
    hipLaunchKernelGGL(k_da_hist, dim3(ceil_div(N, 4)), dim3(256), 0, stream(), N, S->dev, M->atoms->type.p);
    PHB_HIP(hipGetLastError());
}

// W_a . phi of the pool atoms from the current potential grid.
void phi_dot(phb_mc_s* M, DaState* S) {
    if (S->ncand == 0) return;
    auto go = [&](auto kern) {
        hipLaunchKernelGGL(kern, dim3(8 * ceil_div(ceil_div(8 * S->ncand, 256), 8)), dim3(256), 0, stream(), M->dev,
                           S->sp.p, S->cand.p, S->ncand, S->rec.p);
    };
    switch (M->dev.P) {
        case 3: go(k_phi_dot<3>); break;
        case 4: go(k_phi_dot<4>); break;
        case 5: go(k_phi_dot<5>); break;
        case 6: go(k_phi_dot<6>); break;
        case 7: go(k_phi_dot<7>); break;
        default: go(k_phi_dot<8>); break;
    }
    PHB_HIP(hipGetLastError());
}

// psil of every atom from the current charges (exact integers: equal to what the chain keeps).
void local_psil(phb_mc_s* M, DaState* S, const int* chg, int nchg) {
    if (S->ncand > 0)
        hipLaunchKernelGGL(k_psil_all, dim3(ceil_div(S->ncand, 256)), dim3(256), 0, stream(), M->dev, S->dev,
                           S->cand.p, S->ncand, chg, nchg);
    PHB_HIP(hipGetLastError());
}

// The records' charge, type and species from the current identities (the far field stays).
void local_rec_dyn(phb_mc_s* M, DaState* S) {
    if (S->ncand > 0)
        hipLaunchKernelGGL(k_rec_dyn, dim3(ceil_div(S->ncand, 256)), dim3(256), 0, stream(), M->dev, S->dev,
                           S->cand.p, S->ncand);
    PHB_HIP(hipGetLastError());
}

int fold_blocks(const DaState* S) { return std::max(ceil_div(S->ncand * FOLD_LANES, FOLD_THREADS), 1); }

// Every fold workgroup takes the current pending list as added (T is current).
void local_fold_init(phb_mc_s* M, DaState* S) {
    const int nb = fold_blocks(S);
    hipLaunchKernelGGL(k_loc_fold_init, dim3(ceil_div(nb, 256)), dim3(256), 0, stream(), M->dev, S->dev, nb);
    PHB_HIP(hipGetLastError());
}

// After a chain launch: the far part of its accepted changes into T of every pool atom.
void local_fold(phb_mc_s* M, DaState* S) {
    if (S->ncand > 0)
        hipLaunchKernelGGL(k_loc_fold, dim3(fold_blocks(S)), dim3(FOLD_THREADS), 0, stream(), M->dev, S->dev,
                           S->cand.p, S->ncand);
    PHB_HIP(hipGetLastError());
}

template <int P, int TY, int TZ>
void launch_fold_refresh(phb_mc_s* M, DaState* S) {
    static bool lds_set = false;   // per instantiation
    if (!lds_set) {
        PHB_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(k_loc_fold_refresh<P, TY, TZ>),
                                    hipFuncAttributeMaxDynamicSharedMemorySize, (int)refresh_direct_lds(P)));
        lds_set = true;
    }
    const int nf = fold_blocks(S);
    hipLaunchKernelGGL((k_loc_fold_refresh<P, TY, TZ>), dim3(M->rd_ntiles + (nf + 1) / 2), dim3(RD_THREADS),
                       refresh_direct_lds(P), stream(), M->dev, S->dev, M->phi.p,
                       M->d.precision == PHB_FP32 ? M->phi32.p : nullptr, M->rd_tiles.p, M->rd_ntiles, S->cand.p,
                       S->ncand, nf);
}

template <int P>
void fold_refresh_p(phb_mc_s* M, DaState* S) {
    const int ty = M->dev.rd_ty, tz = M->dev.rd_tz;
    if (ty == 8 && tz == 4) launch_fold_refresh<P, 8, 4>(M, S);
    else if (ty == 8) launch_fold_refresh<P, 8, 2>(M, S);
    else if (tz == 4) launch_fold_refresh<P, 4, 4>(M, S);
    else launch_fold_refresh<P, 4, 2>(M, S);
    PHB_HIP(hipGetLastError());
    ++M->refreshes;
}

// After a local-pair chain launch: the fold, then the refresh; one launch when
// the refresh is direct and the order 4 or 5 (the others: two launches).
void local_fold_refresh(phb_mc_s* M, DaState* S) {
    if (S->ncand > 0 && use_direct_refresh(M) && (M->dev.P == 4 || M->dev.P == 5)) {
        if (M->dev.P == 4) fold_refresh_p<4>(M, S);
        else fold_refresh_p<5>(M, S);
        return;
    }
    local_fold(M, S);
    refresh(M);
}

// far_a at the current state: phi with every accepted change, W . phi into the
// records, the charges kept as the freeze's, T (its L rows), then
// far = W . phi + psi_real - q s - T.
void local_freeze(phb_mc_s* M, DaState* S) {
    const int N = M->atoms->n;
    refresh(M);
    phi_dot(M, S);
    if (N > 0) PHB_HIP(hipMemcpyAsync(S->q0.p, M->atoms->q.p, sizeof(double) * N, hipMemcpyDeviceToDevice, stream()));
    local_psil(M, S, nullptr, 0);
    if (S->ncand > 0)
        hipLaunchKernelGGL(k_loc_far, dim3(ceil_div(S->ncand, 256)), dim3(256), 0, stream(), M->dev, S->dev,
                           S->cand.p, S->ncand);
    PHB_HIP(hipGetLastError());
    local_fold_init(M, S);
    S->frozen = true;
}

// Run start of the local-pair form: freeze once; later, when the identities
// changed outside the chain (plain runs, uploads), the records' dynamic part and
// T from the charges (the far part over the pool atoms changed since the freeze).
void local_start(phb_mc_s* M, DaState* S) {
    if (!S->frozen) {
        local_freeze(M, S);
        return;
    }
    refresh(M);
    int dirty = 0;
    if (S->ncand > 0) {
        PHB_HIP(hipMemsetAsync(S->flag.p, 0, sizeof(int), stream()));
        hipLaunchKernelGGL(k_loc_dirty, dim3(ceil_div(S->ncand, 256)), dim3(256), 0, stream(), M->dev, S->dev,
                           S->cand.p, S->ncand, S->flag.p);
        PHB_HIP(hipGetLastError());
        S->flag.download(&dirty, 1);
        sync();
    }
    if (dirty) {
        const int N = M->atoms->n;
        std::vector<double> q(N), q0(N);
        std::vector<int> pool(S->ncand), chg;
        M->atoms->q.download(q.data(), N);
        S->q0.download(q0.data(), N);
        S->cand.download(pool.data(), S->ncand);
        sync();
        for (int a : pool)
            if (q[a] != q0[a]) chg.push_back(a);
        DevBuf<int> dchg(std::max<size_t>(chg.size(), 1));
        if (!chg.empty()) dchg.upload(chg.data(), chg.size());
        local_rec_dyn(M, S);
        local_psil(M, S, dchg.p, (int)chg.size());
        sync();   // dchg is freed on return
    }
    local_fold_init(M, S);
}

template <int P>
void build_local_p(phb_mc_s* M, DaState* S, const int* lown) {
    hipLaunchKernelGGL(k_loc_coef<P>, dim3((unsigned)((S->nlocal + 255) / 256)), dim3(256), 0, stream(), M->dev,
                       S->nlocal, S->latom.p, lown, S->lcoef.p);
}

// Cardinal B-spline M_n at x (recursion over the order).
double cardinal_bspline(int n, double x) {
    if (n == 1) return x >= 0.0 && x < 1.0 ? 1.0 : 0.0;
    return (x * cardinal_bspline(n - 1, x) + (n - x) * cardinal_bspline(n - 1, x - 1.0)) / (n - 1);
}

// The mesh kernel smoothed by the stencils of two atoms on grid points,
// sum_{u,v} w(u) w(v) g(r + u - v): along each axis a convolution with the
// discrete autocorrelation of the stencil samples M_P(1..P-1), whose DFT is the
// squared spline modulus that g has divided out, so the table is the continuum
// reciprocal-space pair potential sampled at the grid offsets.  Folded to |offset|
// per axis like gf (the result is even).
void build_smoothed_kernel(phb_mc_s* M, DaState* S) {
    const int nx = M->dev.nx, ny = M->dev.ny, nz = M->dev.nz, P = M->dev.P;
    const size_t G = (size_t)nx * ny * nz;
    std::vector<double> g(G), t(G);
    M->gk.download(g.data(), G);
    sync();
    std::vector<double> sm(P + 1, 0.0), a(2 * P - 1, 0.0);
    for (int k = 1; k < P; ++k) sm[k] = cardinal_bspline(P, (double)k);
    for (int d = -(P - 1); d <= P - 1; ++d)
        for (int k = 1; k < P; ++k) {
            const int k2 = k + d;
            if (k2 >= 1 && k2 < P) a[d + P - 1] += sm[k] * sm[k2];
        }
    auto pass = [&](std::vector<double>& in, std::vector<double>& out, int axis) {
        const int n = axis == 0 ? nx : axis == 1 ? ny : nz;
        const size_t stride = axis == 0 ? 1 : axis == 1 ? (size_t)nx : (size_t)nx * ny;
        for (size_t k = 0; k < G; ++k) {
            const int c = (int)((k / stride) % n);
            double v = 0.0;
            for (int d = -(P - 1); d <= P - 1; ++d) {
                int cc = c + d;
                cc += cc < 0 ? n : 0;
                cc -= cc >= n ? n : 0;
                v += a[d + P - 1] * in[k + ((long long)cc - c) * (long long)stride];
            }
            out[k] = v;
        }
    };
    pass(g, t, 0);
    pass(t, g, 1);
    pass(g, t, 2);
    const int fx = nx / 2 + 1, fy = ny / 2 + 1, fz = nz / 2 + 1;
    std::vector<double> f((size_t)fx * fy * fz);
    for (int z = 0; z < fz; ++z)
        for (int y = 0; y < fy; ++y)
            for (int x = 0; x < fx; ++x) f[((size_t)z * fy + y) * fx + x] = t[((size_t)z * ny + y) * nx + x];
    S->gs.alloc(f.size());
    S->gs.upload(f.data(), f.size());
    S->dev.gs = S->gs.p;
    sync();
}

// L rows: for each pool atom the pool atoms within r_local of it or in its lateral
// column, ascending, and their C_ab.
void build_local(phb_mc_s* M, DaState* S) {
    build_smoothed_kernel(M, S);
    const int N = M->atoms->n, np = S->ncand;
    DaDev& A = S->dev;
    A.hx = M->box.Lx / M->dev.nx; A.hy = M->box.Ly / M->dev.ny; A.hz = M->box.Lz / M->dev.nz;
    A.cw = 5.0 * std::max(A.hx, A.hy);
    std::vector<int> ps(np);
    S->cand.download(ps.data(), np);
    sync();
    std::sort(ps.begin(), ps.end());
    DevBuf<int> dps(std::max(np, 1));
    if (np > 0) dps.upload(ps.data(), np);
    DevBuf<int> cnt(std::max(N, 1));
    PHB_HIP(hipMemsetAsync(cnt.p, 0, sizeof(int) * std::max(N, 1), stream()));
    const unsigned rows_grid = (unsigned)ceil_div(np, 4);   // four waves per workgroup
    if (np > 0)
        hipLaunchKernelGGL(k_loc_rows<false>, dim3(rows_grid), dim3(256), 0, stream(), M->dev, A, S->cand.p, np, dps.p,
                           np, cnt.p, nullptr, nullptr);
    PHB_HIP(hipGetLastError());
    std::vector<int> c(N), head(N + 1, 0);
    cnt.download(c.data(), N);
    sync();
    long long tot = 0;
    for (int a = 0; a < N; ++a) {
        head[a] = (int)tot;
        tot += c[a];
    }
    require(tot < (1LL << 31), "phb_da_attach: too many local pairs");
    head[N] = (int)tot;
    S->nlocal = tot;
    S->lhead.alloc(N + 1);
    S->lhead.upload(head.data(), N + 1);
    S->latom.alloc(std::max<long long>(tot, 1));
    S->lcoef.alloc(std::max<long long>(tot, 1));
    S->psil.alloc(std::max(N, 1));
    S->q0.alloc(std::max(N, 1));
    S->rqn.alloc(MC_KMAX);
    S->fold_from.alloc(1);
    S->flag.alloc(1);
    S->fapplied.alloc(fold_blocks(S));
    A.lhead = S->lhead.p; A.latom = S->latom.p; A.lcoef = S->lcoef.p; A.psil = S->psil.p;
    A.q0 = S->q0.p; A.rqn = S->rqn.p; A.fold_from = S->fold_from.p; A.fapplied = S->fapplied.p;
    PHB_HIP(hipMemsetAsync(S->fold_from.p, 0, sizeof(int), stream()));
    DevBuf<int> lown(std::max<long long>(tot, 1));
    if (np > 0)
        hipLaunchKernelGGL(k_loc_rows<true>, dim3(rows_grid), dim3(256), 0, stream(), M->dev, A, S->cand.p, np, dps.p,
                           np, nullptr, S->latom.p, lown.p);
    if (tot > 0) {
        switch (M->dev.P) {
            case 3: build_local_p<3>(M, S, lown.p); break;
            case 4: build_local_p<4>(M, S, lown.p); break;
            case 5: build_local_p<5>(M, S, lown.p); break;
            case 6: build_local_p<6>(M, S, lown.p); break;
            case 7: build_local_p<7>(M, S, lown.p); break;
            default: build_local_p<8>(M, S, lown.p); break;
        }
    }
    PHB_HIP(hipGetLastError());
    sync();   // dps, cnt and lown are freed on return
}

// Uploads the views the chain kernel reads by pointer (before a run; the stream orders it).
void upload_views(phb_mc_s* M, DaState* S, phb_surrogate_s* sur) {
    S->dev.prof = M->profiling && M->phase_clock ? 1 : 0;
    const NetDev net = sur->net();
    S->dmc.upload(&M->dev, 1);
    S->dda.upload(&S->dev, 1);
    S->dnet.upload(&net, 1);
    sync();   // the host copies are stack/member values
}

template <int P>
void launch_da_p(phb_mc_s* M, DaState* S) {
    (void)M;
    // Pools stay in HBM: a launch covers few accepted moves, so copying them
    // into LDS and back each launch costs more than an in-LDS shift saves (the
    // shift is one merged pass at known positions).
    if (S->dev.local)
        hipLaunchKernelGGL((k_da_chain<P, int, false, true>), dim3(1), dim3(DA_THREADS), 0, stream(), S->dmc.p,
                           S->dda.p, S->dnet.p);
    else
        hipLaunchKernelGGL((k_da_chain<P, int, false, false>), dim3(1), dim3(DA_THREADS), 0, stream(), S->dmc.p,
                           S->dda.p, S->dnet.p);
}

void launch_da(phb_mc_s* M, DaState* S) {
    switch (M->dev.P) {
        case 3: launch_da_p<3>(M, S); break;
        case 4: launch_da_p<4>(M, S); break;
        case 5: launch_da_p<5>(M, S); break;
        case 6: launch_da_p<6>(M, S); break;
        case 7: launch_da_p<7>(M, S); break;
        default: launch_da_p<8>(M, S); break;
    }
}

void check_moves(phb_mc_s* M, int n, const int* i, const int* j, const int* f) {
    for (int k = 0; k < n; ++k)
        require(i[k] >= 0 && i[k] < M->atoms->n && j[k] >= 0 && j[k] < M->atoms->n && i[k] != j[k] &&
                    (f == nullptr || f[k] == 0 || f[k] == 1),
                "delayed acceptance: move index out of range");
}

// physics part B of n moves (device buffers); phi must hold every accepted change
void base_terms(phb_mc_s* M, DaState* S, int n, const int* di, const int* dj, const int* df, double* out) {
    if (S->dev.local) {
        local_start(M, S);   // freezes on first use
        auto gl = [&](auto kern) {
            hipLaunchKernelGGL(kern, dim3(ceil_div(n, 128)), dim3(128), 0, stream(), M->dev, S->dev, n, di, dj, df,
                               out);
        };
        switch (M->dev.P) {
            case 3: gl(k_da_base_local<3>); break;
            case 4: gl(k_da_base_local<4>); break;
            case 5: gl(k_da_base_local<5>); break;
            case 6: gl(k_da_base_local<6>); break;
            case 7: gl(k_da_base_local<7>); break;
            default: gl(k_da_base_local<8>); break;
        }
        PHB_HIP(hipGetLastError());
        return;
    }
    refresh(M);   // folds pending accepted changes into phi (no-op when none are pending)
    auto go = [&](auto kern) {
        hipLaunchKernelGGL(kern, dim3(ceil_div(n, 128)), dim3(128), 0, stream(), M->dev, S->dev, n, di, dj, df, out);
    };
    switch (M->dev.P) {
        case 3: go(k_da_base<3>); break;
        case 4: go(k_da_base<4>); break;
        case 5: go(k_da_base<5>); break;
        case 6: go(k_da_base<6>); break;
        case 7: go(k_da_base<7>); break;
        default: go(k_da_base<8>); break;
    }
    PHB_HIP(hipGetLastError());
}

// descriptors + r_ij of n moves (device buffers)
void descriptors(phb_mc_s* M, DaState* S, int n, const int* di, const int* dj, short* dd, double* drij) {
    rebuild_hist(M, S);
    const long long tot = (long long)n * NN_IN;
    hipLaunchKernelGGL(k_da_desc, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, stream(), S->dev, M->dev, n, di,
                       dj, dd, drij);
    PHB_HIP(hipGetLastError());
}

}  // namespace

extern "C" int phb_surrogate_create(long long capacity, unsigned seed, phb_surrogate* out) {
    return guard([&] {
        require(out != nullptr && capacity >= 0, "phb_surrogate_create: bad arguments");
        auto S = std::make_unique<phb_surrogate_s>();
        S->theta.alloc(NN_PARAM); S->m.alloc(NN_TRAIN); S->v.alloc(NN_TRAIN);
        S->grad.alloc(NTP);
        S->w1t.alloc(NN_IN * NN_H); S->w2t.alloc(NN_H * NN_H); S->w2.alloc(NN_H * NN_H); S->w3.alloc(NN_H);
        S->cap = capacity;
        S->dd.alloc(std::max<long long>(capacity, 1) * NN_IN);
        S->dt.alloc(std::max<long long>(capacity, 1));
        // He-uniform initial weights, zero biases, unit scales (std::mt19937_64(seed))
        std::vector<float> th(NN_PARAM, 0.0f);
        std::mt19937_64 g(seed);
        std::uniform_real_distribution<double> u(-1.0, 1.0);
        for (int k = 0; k < NN_IN * NN_H; ++k) th[OFF_W1 + k] = (float)(u(g) * std::sqrt(6.0 / NN_IN));
        for (int k = 0; k < NN_H * NN_H; ++k) th[OFF_W2 + k] = (float)(u(g) * std::sqrt(6.0 / NN_H));
        for (int k = 0; k < NN_H; ++k) th[OFF_W3 + k] = (float)(u(g) * std::sqrt(6.0 / NN_H) * 0.1);
        for (int k = 0; k < NN_IN; ++k) th[OFF_SCALE + k] = 1.0f;
        th[OFF_MEAN] = 0.0f; th[OFF_STD] = 1.0f;
        S->theta.upload(th.data(), NN_PARAM);
        PHB_HIP(hipMemsetAsync(S->m.p, 0, NN_TRAIN * sizeof(float), stream()));
        PHB_HIP(hipMemsetAsync(S->v.p, 0, NN_TRAIN * sizeof(float), stream()));
        S->cast();
        sync();
        *out = S.release();
        return PHB_OK;
    });
}

extern "C" int phb_surrogate_destroy(phb_surrogate s) {
    return guard([&] { delete s; return PHB_OK; });
}

extern "C" int phb_surrogate_set_params(phb_surrogate s, const float* theta) {
    return guard([&] {
        require(s && theta, "phb_surrogate_set_params: null");
        require(theta[OFF_STD] > 0.0f, "phb_surrogate_set_params: out_std must be > 0");
        s->theta.upload(theta, NN_PARAM);
        PHB_HIP(hipMemsetAsync(s->m.p, 0, NN_TRAIN * sizeof(float), stream()));
        PHB_HIP(hipMemsetAsync(s->v.p, 0, NN_TRAIN * sizeof(float), stream()));
        s->steps = 0;
        s->cast();
        sync();
        return PHB_OK;
    });
}

extern "C" int phb_surrogate_get_params(phb_surrogate s, float* theta) {
    return guard([&] {
        require(s && theta, "phb_surrogate_get_params: null");
        s->theta.download(theta, NN_PARAM);
        sync();
        return PHB_OK;
    });
}

extern "C" int phb_surrogate_forward(phb_surrogate s, int n, const short* d, float* y) {
    return guard([&] {
        require(s && n >= 0 && (n == 0 || (d && y)), "phb_surrogate_forward: bad arguments");
        if (n == 0) return PHB_OK;
        DevBuf<short> dd((size_t)n * NN_IN);
        DevBuf<float> dy(n);
        dd.upload(d, (size_t)n * NN_IN);
        hipLaunchKernelGGL(k_mlp_forward, dim3(ceil_div(n, 64)), dim3(256), 0, stream(), s->net(), dd.p, (long long)n,
                           dy.p);
        PHB_HIP(hipGetLastError());
        dy.download(y, n);
        sync();
        return PHB_OK;
    });
}

extern "C" int phb_surrogate_grad(phb_surrogate s, int n, const short* d, const float* t, float* grad, double* loss) {
    return guard([&] {
        require(s && n > 0 && n % 64 == 0 && d && t && grad, "phb_surrogate_grad: n must be a positive multiple of 64");
        DevBuf<short> dd((size_t)n * NN_IN);
        DevBuf<float> dt(n);
        dd.upload(d, (size_t)n * NN_IN);
        dt.upload(t, n);
        s->batch_grad(dd.p, dt.p, nullptr, n);
        std::vector<float> g(NTP);
        s->grad.download(g.data(), NTP);
        sync();
        std::copy(g.begin(), g.begin() + NN_TRAIN, grad);
        if (loss) *loss = (double)g[NN_TRAIN] / n;
        return PHB_OK;
    });
}

extern "C" int phb_surrogate_data_add(phb_surrogate s, int n, const short* d, const float* t) {
    return guard([&] {
        require(s && n >= 0 && (n == 0 || (d && t)), "phb_surrogate_data_add: bad arguments");
        require(s->n + n <= s->cap, "phb_surrogate_data_add: dataset capacity exceeded");
        if (n == 0) return PHB_OK;
        PHB_HIP(hipMemcpyAsync(s->dd.p + s->n * NN_IN, d, (size_t)n * NN_IN * sizeof(short), hipMemcpyHostToDevice,
                               stream()));
        PHB_HIP(hipMemcpyAsync(s->dt.p + s->n, t, (size_t)n * sizeof(float), hipMemcpyHostToDevice, stream()));
        sync();
        s->n += n;
        return PHB_OK;
    });
}

extern "C" int phb_surrogate_data_get(phb_surrogate s, long long first, int n, short* d, float* t) {
    return guard([&] {
        require(s && first >= 0 && n >= 0 && first + n <= s->n, "phb_surrogate_data_get: range outside the dataset");
        if (n == 0) return PHB_OK;
        if (d) PHB_HIP(hipMemcpyAsync(d, s->dd.p + first * NN_IN, (size_t)n * NN_IN * sizeof(short),
                                      hipMemcpyDeviceToHost, stream()));
        if (t) PHB_HIP(hipMemcpyAsync(t, s->dt.p + first, (size_t)n * sizeof(float), hipMemcpyDeviceToHost, stream()));
        sync();
        return PHB_OK;
    });
}

extern "C" int phb_surrogate_data_size(phb_surrogate s, long long* n) {
    return guard([&] {
        require(s && n, "phb_surrogate_data_size: null");
        *n = s->n;
        return PHB_OK;
    });
}

extern "C" int phb_surrogate_data_clear(phb_surrogate s) {
    return guard([&] {
        require(s != nullptr, "phb_surrogate_data_clear: null");
        s->n = 0;
        return PHB_OK;
    });
}

extern "C" int phb_surrogate_train(phb_surrogate s, const phb_train_desc* td, phb_train_stats* st) {
    return guard([&] {
        require(s && td, "phb_surrogate_train: null");
        const long long ntr = td->n_train > 0 ? std::min(td->n_train, s->n) : s->n;
        const int B = td->batch > 0 ? td->batch : 4096;
        require(B % 64 == 0, "phb_surrogate_train: batch must be a multiple of 64");
        require(ntr >= B, "phb_surrogate_train: fewer training samples than one batch");
        require(td->epochs >= 0 && td->lr > 0.0, "phb_surrogate_train: bad epochs or learning rate");
        const auto t0 = std::chrono::steady_clock::now();
        if (td->normalize) {
            // input scales and target normalisation from (up to) 65,536 evenly spaced training samples
            const long long ns = std::min<long long>(ntr, 65536);
            std::vector<short> hd((size_t)ns * NN_IN);
            std::vector<float> ht(ns);
            const long long stride = ntr / ns;
            {
                DevBuf<short> gd((size_t)ns * NN_IN);
                DevBuf<float> gt(ns);
                hipLaunchKernelGGL(k_gather_rows, dim3((unsigned)((ns * NN_IN + 255) / 256)), dim3(256), 0, stream(),
                                   s->dd.p, s->dt.p, stride, ns, gd.p, gt.p);
                PHB_HIP(hipGetLastError());
                gd.download(hd.data(), (size_t)ns * NN_IN);
                gt.download(ht.data(), ns);
                sync();
            }
            std::vector<float> th(NN_PARAM);
            s->theta.download(th.data(), NN_PARAM);
            sync();
            for (int c = 0; c < NN_IN; ++c) {
                double s1 = 0, s2 = 0;
                for (long long k = 0; k < ns; ++k) { const double v = hd[k * NN_IN + c]; s1 += v; s2 += v * v; }
                const double var = s2 / ns - (s1 / ns) * (s1 / ns);
                th[OFF_SCALE + c] = (float)(1.0 / std::sqrt(std::max(var, 1.0)));
            }
            double s1 = 0, s2 = 0;
            for (long long k = 0; k < ns; ++k) { s1 += ht[k]; s2 += (double)ht[k] * ht[k]; }
            const double mean = s1 / ns, var = s2 / ns - mean * mean;
            th[OFF_MEAN] = (float)mean;
            th[OFF_STD] = (float)std::sqrt(std::max(var, 1e-12));
            s->theta.upload(th.data(), NN_PARAM);
        }
        const long long spe = ntr / B;
        const long long total = spe * td->epochs;
        s->idx.alloc(std::max<long long>(spe * B, 1));
        s->loss_hist.alloc(std::max<long long>(total, 1));
        s->bufs(B);
        std::vector<int> perm(ntr);
        std::mt19937_64 g(td->seed);
        const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
        long long k = 0;
        for (int ep = 0; ep < td->epochs; ++ep) {
            std::iota(perm.begin(), perm.end(), 0);
            std::shuffle(perm.begin(), perm.end(), g);
            s->idx.upload(perm.data(), spe * B);
            for (long long b = 0; b < spe; ++b, ++k) {
                ++s->steps;
                const double bc1 = 1.0 - std::pow((double)b1, (double)s->steps);
                const double bc2 = 1.0 - std::pow((double)b2, (double)s->steps);
                s->batch_grad(s->dd.p, s->dt.p, s->idx.p + b * B, B);
                hipLaunchKernelGGL(k_mlp_adam, dim3(ceil_div(NTP, 256)), dim3(256), 0, stream(), s->theta.p, s->m.p,
                                   s->v.p, s->grad.p, (float)td->lr, b1, b2, eps, (float)bc1, (float)bc2, s->w1t.p,
                                   s->w2t.p, s->w2.p, s->w3.p, s->loss_hist.p + k, 1.0f / B);
            }
            PHB_HIP(hipGetLastError());
            sync();   // the index buffer is reused by the next epoch
        }
        sync();
        if (st) {
            st->steps = total;
            st->first_loss = st->last_loss = std::numeric_limits<double>::quiet_NaN();
            if (total > 0) {
                float a = 0, b = 0;
                s->loss_hist.download(&a, 1);
                PHB_HIP(hipMemcpyAsync(&b, s->loss_hist.p + total - 1, sizeof(float), hipMemcpyDeviceToHost, stream()));
                sync();
                st->first_loss = a; st->last_loss = b;
            }
            st->ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        }
        return PHB_OK;
    });
}

extern "C" int phb_surrogate_loss_history(phb_surrogate s, long long n, float* loss) {
    return guard([&] {
        require(s && loss && n >= 0 && (size_t)n <= s->loss_hist.n, "phb_surrogate_loss_history: bad arguments");
        if (n) s->loss_hist.download(loss, n);
        sync();
        return PHB_OK;
    });
}

extern "C" int phb_surrogate_eval(phb_surrogate s, long long first, long long n, double* rmse, double* mae) {
    return guard([&] {
        require(s && first >= 0 && n > 0 && first + n <= s->n, "phb_surrogate_eval: range outside the dataset");
        DevBuf<float> dy(n);
        hipLaunchKernelGGL(k_mlp_forward, dim3((unsigned)((n + 63) / 64)), dim3(256), 0, stream(), s->net(),
                           s->dd.p + first * NN_IN, n, dy.p);
        PHB_HIP(hipGetLastError());
        std::vector<float> y(n), t(n);
        dy.download(y.data(), n);
        PHB_HIP(hipMemcpyAsync(t.data(), s->dt.p + first, n * sizeof(float), hipMemcpyDeviceToHost, stream()));
        sync();
        double se = 0, ae = 0;
        for (long long k = 0; k < n; ++k) { const double e = (double)y[k] - t[k]; se += e * e; ae += std::fabs(e); }
        if (rmse) *rmse = std::sqrt(se / n);
        if (mae) *mae = ae / n;
        return PHB_OK;
    });
}

extern "C" int phb_da_attach(phb_mc M, const phb_da_desc* d) {
    return guard([&] {
        require(M && d && d->species_of_type, "phb_da_attach: null argument");
        require(d->ntypes == M->params->d.ntypes, "phb_da_attach: ntypes must match the parameter set");
        require(M->d.precision == PHB_FP64, "phb_da_attach: the delayed-acceptance chain runs the FP64 form",
                PHB_ERR_UNSUPPORTED);
        require(M->params->d.coulombtype == PHB_COUL_PME,
                "phb_da_attach: the surrogate's physics part needs coulombtype=pme", PHB_ERR_UNSUPPORTED);
        require(d->r_desc > 0.0, "phb_da_attach: r_desc must be > 0");
        const double rl = d->r_local > 0.0 ? d->r_local : d->r_desc;
        require(!d->local_pairs || (rl >= M->params->d.rc && rl <= d->r_desc),
                "phb_da_attach: r_local must lie in [rc, r_desc]");
        const int cand = d->candidates > 0 ? d->candidates : DA_M;
        require(cand >= 1 && cand <= DA_M, "phb_da_attach: candidates must be in [1, 64]");
        const int N = M->atoms->n;
        require(N < (1 << DA_SHIFT), "phb_da_attach: more than 2^24 atoms");
        for (int t = 0; t < d->ntypes; ++t)
            require(d->species_of_type[t] >= -1 && d->species_of_type[t] <= 3,
                    "phb_da_attach: species must be -1 or 0..3");
        auto S = std::make_unique<DaState>();
        S->dmc.alloc(1); S->dda.alloc(1); S->dnet.alloc(1);
        S->d = *d;
        S->d.candidates = cand;
        S->species.assign(d->species_of_type, d->species_of_type + d->ntypes);
        S->d.species_of_type = nullptr;
        S->sp.alloc(d->ntypes);
        S->sp.upload(S->species.data(), d->ntypes);
        // r_desc rows: half list without bond exclusion, expanded, entries packed with their shell
        phb_csr half = nullptr;
        if (phb_nlist_build_half(M->atoms, &M->box, nullptr, d->r_desc, &half) != PHB_OK)
            throw Error(PHB_ERR_INTERNAL, std::string("phb_da_attach: ") + phb_last_error());
        std::unique_ptr<phb_csr_s> half_owner(half);
        DevBuf<int> nbr;
        S->nentries = build_full_rows(N, half, S->dhead, nbr);
        half_owner.reset();
        S->dlist.alloc(std::max<long long>(S->nentries, 1));
        S->H.alloc((size_t)std::max(N, 1) * NN_IN);
        DaDev& A = S->dev;
        A.dhead = S->dhead.p; A.dlist = S->dlist.p; A.H = S->H.p; A.species = S->sp.p;
        {   // species a site's change touches vs those an ion's change touches
            auto spc = [&](int t1) { return t1 >= 1 && t1 <= (int)S->species.size() ? (int)S->species[t1 - 1] : -1; };
            const int a1 = spc(M->d.nh_n_type), a2 = spc(M->d.nh_p_type), b1 = spc(M->d.w_type), b2 = spc(M->d.cl_type);
            auto meet = [](int x, int y) { return x >= 0 && x == y; };
            A.upd_one = (meet(a1, b1) || meet(a1, b2) || meet(a2, b1) || meet(a2, b2)) ? 0 : 1;
        }
        A.x = M->atoms->x.p; A.y = M->atoms->y.p;
        A.Lx = M->box.Lx; A.Ly = M->box.Ly; A.Lz = M->box.Lz; A.rd2 = d->r_desc * d->r_desc;
        for (int k = 0; k < 12; ++k) {
            const double rk = (double)k * d->r_desc / 12.0;
            A.th[k] = rk * rk;
        }
        A.M = cand;
        A.rc2 = M->params->d.rc * M->params->d.rc;
        S->rec_fy.alloc(std::max<long long>(M->d.record_cap, 1));
        A.rec_fy = S->rec_fy.p;
        S->rec.alloc(std::max(N, 1));
        A.rec = S->rec.p;
        A.use_psi = d->real_space ? 1 : 0;
        if (A.use_psi) {
            S->psi.alloc(std::max(N, 1));
            A.psi = S->psi.p;
        }
        if (N > 0)
            hipLaunchKernelGGL(k_rec_static, dim3(ceil_div(N, 256)), dim3(256), 0, stream(), N, M->atoms->x.p,
                               M->atoms->y.p, M->atoms->z.p, M->selfw.p, S->rec.p);
        {   // pool atoms (sites and ions: a fixed set) sorted by stencil origin
            const int np = M->nsites + M->nions;
            std::vector<int> pool(np), sb(3 * (size_t)N);
            M->pbuf.download(pool.data(), np);
            M->sbase.download(sb.data(), 3 * (size_t)N);
            sync();
            std::sort(pool.begin(), pool.end(), [&](int x, int y) {
                const int* a = &sb[3 * (size_t)x];
                const int* b = &sb[3 * (size_t)y];
                if (a[2] != b[2]) return a[2] < b[2];
                if (a[1] != b[1]) return a[1] < b[1];
                return a[0] != b[0] ? a[0] < b[0] : x < y;
            });
            S->ncand = np;
            S->cand.alloc(std::max(np, 1));
            S->cand.upload(pool.data(), np);
        }
        if (S->nentries > 0)
            hipLaunchKernelGGL(k_da_row_entries, dim3(ceil_div(N, 256)), dim3(256), 0, stream(), N, M->atoms->x.p,
                               M->atoms->y.p, M->atoms->z.p, A, nbr.p, S->dlist.p);
        PHB_HIP(hipGetLastError());
        rebuild_hist(M, S.get());
        sync();
        A.local = d->local_pairs ? 1 : 0;
        A.rl2 = rl * rl;
        if (A.local) build_local(M, S.get());
        if (M->da_free) M->da_free(M->da);
        M->da = S.release();
        M->da_free = da_free;
        return PHB_OK;
    });
}

extern "C" int phb_da_freeze(phb_mc M) {
    return guard([&] {
        require(M != nullptr, "phb_da_freeze: null argument");
        DaState* S = da_of(M);
        require(S->dev.local != 0, "phb_da_freeze: the chain is not in the local-pair form", PHB_ERR_STATE);
        local_freeze(M, S);
        sync();
        return PHB_OK;
    });
}

extern "C" int phb_da_descriptors(phb_mc M, int n, const int* i, const int* j, short* d, double* rij) {
    return guard([&] {
        require(M && n >= 0 && (n == 0 || (i && j && d)), "phb_da_descriptors: bad arguments");
        DaState* S = da_of(M);
        if (n == 0) return PHB_OK;
        check_moves(M, n, i, j, nullptr);
        DevBuf<int> di(n), dj(n);
        DevBuf<short> dd((size_t)n * NN_IN);
        DevBuf<double> dr(n);
        di.upload(i, n); dj.upload(j, n);
        descriptors(M, S, n, di.p, dj.p, dd.p, dr.p);
        dd.download(d, (size_t)n * NN_IN);
        if (rij) dr.download(rij, n);
        sync();
        return PHB_OK;
    });
}

extern "C" int phb_da_collect(phb_mc M, phb_surrogate s, int n, const int* i, const int* j, const int* fwd) {
    return guard([&] {
        require(M && s && n >= 0 && (n == 0 || (i && j && fwd)), "phb_da_collect: bad arguments");
        DaState* S = da_of(M);
        require(s->n + n <= s->cap, "phb_da_collect: dataset capacity exceeded");
        if (n == 0) return PHB_OK;
        check_moves(M, n, i, j, fwd);
        DevBuf<int> di(n), dj(n), df(n);
        DevBuf<double> dr(n), dE(n), base(n);
        di.upload(i, n); dj.upload(j, n); df.upload(fwd, n);
        descriptors(M, S, n, di.p, dj.p, s->dd.p + s->n * NN_IN, dr.p);
        base_terms(M, S, n, di.p, dj.p, df.p, base.p);
        launch_probe(M, n, di.p, dj.p, df.p, dE.p);
        hipLaunchKernelGGL(k_da_target, dim3(ceil_div(n, 256)), dim3(256), 0, stream(), n, df.p, dE.p, base.p,
                           s->dt.p + s->n);
        PHB_HIP(hipGetLastError());
        sync();
        s->n += n;
        return PHB_OK;
    });
}

extern "C" int phb_da_predict(phb_mc M, phb_surrogate s, int n, const int* i, const int* j, const int* fwd,
                              double* dE_star) {
    return guard([&] {
        require(M && s && n >= 0 && (n == 0 || (i && j && fwd && dE_star)), "phb_da_predict: bad arguments");
        DaState* S = da_of(M);
        if (n == 0) return PHB_OK;
        check_moves(M, n, i, j, fwd);
        DevBuf<int> di(n), dj(n), df(n);
        DevBuf<short> dd((size_t)n * NN_IN);
        DevBuf<double> dr(n), out(n), base(n);
        DevBuf<float> dy(n);
        di.upload(i, n); dj.upload(j, n); df.upload(fwd, n);
        descriptors(M, S, n, di.p, dj.p, dd.p, dr.p);
        base_terms(M, S, n, di.p, dj.p, df.p, base.p);
        hipLaunchKernelGGL(k_mlp_forward, dim3(ceil_div(n, 64)), dim3(256), 0, stream(), s->net(), dd.p, (long long)n,
                           dy.p);
        hipLaunchKernelGGL(k_da_surrogate, dim3(ceil_div(n, 256)), dim3(256), 0, stream(), n, df.p, dy.p, base.p,
                           out.p);
        PHB_HIP(hipGetLastError());
        out.download(dE_star, n);
        sync();
        return PHB_OK;
    });
}

extern "C" int phb_da_run(phb_mc M, phb_surrogate s, long long ntrials, phb_da_stats* out) {
    return guard([&] {
        require(M && s && ntrials >= 0, "phb_da_run: bad arguments");
        DaState* S = da_of(M);
        const bool local = S->dev.local != 0;
        rebuild_hist(M, S);
        upload_views(M, S, s);
        if (local) local_start(M, S);   // far frozen once; records and psil to the current identities
        else {
            refresh(M);   // the surrogate needs phi to hold every accepted change
            phi_dot(M, S);
        }
        if (!local && S->dev.use_psi) {   // psi from the current charges (identities may have changed since the last run)
            hipLaunchKernelGGL(k_psi_all, dim3(ceil_div(M->atoms->n, 256)), dim3(256), 0, stream(), M->dev,
                               S->psi.p);
            PHB_HIP(hipGetLastError());
        }
        McCtl c = read_ctl(M);
        const long long target = c.trials_done + ntrials;
        long long done = c.trials_done;
        int status = c.status;
        while (done < target && status == 0) {
            long long chunk_end = target;
            if (M->next_recompute > 0) chunk_end = std::min(chunk_end, M->next_recompute);
            hipLaunchKernelGGL(k_da_set_target, dim3(1), dim3(64), 0, stream(), M->ctl.p, chunk_end);
            while (done < chunk_end && status == 0) {
                // a launch ends at its accepted move (local-pair form: when its pending budget is full)
                const int batch = launch_batch(c, chunk_end - done, local ? M->d.pending_max : 2);
                const long long before = done;
                for (int b = 0; b < batch; ++b) {
                    const size_t e0 = M->ev_used;
                    if (M->profiling) M->event();
                    launch_da(M, S);
                    ++M->launches;
                    ++S->launches;
                    if (M->profiling) M->event();
                    if (local) local_fold_refresh(M, S);
                    else {
                        refresh(M);
                        phi_dot(M, S);
                    }
                    if (M->profiling) { M->event(); M->marks.push_back({e0, 0}); }
                }
                PHB_HIP(hipGetLastError());
                c = read_ctl(M);
                S->trials_seen += c.trials_done - before;
                require(c.trials_done > done || c.status != 0 || c.trials_done >= chunk_end,
                        "phb_da_run: the chain made no progress", PHB_ERR_INTERNAL);
                require(c.status != 2, "phb_da_run: engine stream rolled back past its kept block", PHB_ERR_INTERNAL);
                require(c.status != 3, "phb_da_run: accepted changes pending at a chain launch", PHB_ERR_INTERNAL);
                done = c.trials_done;
                status = c.status;
            }
            if (done >= chunk_end && M->next_recompute > 0 && chunk_end == M->next_recompute) {
                full_recompute(M);
                if (!local) phi_dot(M, S);
                M->next_recompute += M->d.recompute_interval;
            }
        }
        sync();
        if (out) {
            c = read_ctl(M);
            out->attempted = c.trials_done;
            out->stage1_accepted = c.da_stage1;
            out->accepted = c.accepted;
            out->rounds = c.da_rounds;
            out->launches = M->launches;
            out->nprot = c.nprot;
            out->nsites = M->nsites;
            out->status = c.status;
            out->E = c.E;
            out->chain_ms = M->chain_ms;
            out->refresh_ms = M->refresh_ms;
            for (int k = 0; k < 8; ++k) out->phase_us[k] = (double)c.da_prof[k] * 0.01;   // 100 MHz ticks
        }
        return PHB_OK;
    });
}

extern "C" int phb_da_record(phb_mc M, long long n, signed char* fwd, int* i, int* j, signed char* stage,
                             double* dE_star, double* dE_star_rev, double* dE) {
    // the delayed-acceptance chain stores the stage in the acceptance record and dE* in the uniform record
    const int rc = phb_mc_record(M, n, fwd, i, j, stage, dE, dE_star);
    if (rc != PHB_OK || !dE_star_rev || n == 0) return rc;
    return guard([&] {
        DaState* S = da_of(M);
        S->rec_fy.download(dE_star_rev, n);
        sync();
        return PHB_OK;
    });
}
