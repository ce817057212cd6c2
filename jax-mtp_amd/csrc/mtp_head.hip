// Hand-written kernels of the distilled head (network layout in
// mtp_head.hpp).  The reference has no learned head; these kernels implement
// the MLP's inference (energy and dE/db), its parameter gradient for
// force-and-energy matching, and the basis-Jacobian products used in
// training.
//
// Mapping: one wavefront per atom at a time, lane j = hidden unit j (64
// hidden units = 64 lanes).  Weight matrices sit in LDS in both orientations,
// so every lane reads consecutive addresses; activations move between lanes
// with readlane broadcasts (uniform lane index, no LDS traffic).  The work
// per atom is ~0.2 MFLOP in training, ~30 kFLOP in inference: small next to
// the MTP descriptor, so the kernels favour simplicity over peak FLOP rate.
#include <hip/hip_runtime.h>

#include "mtp_device.hpp"
#include "mtp_head.hpp"

namespace mtp {
namespace {

constexpr int H = kHeadHidden, IN = kHeadIn;
// offsets in theta
constexpr int oW1 = 0, oB1 = IN * H, oW2 = oB1 + H, oB2 = oW2 + H * H, oW3 = oB2 + H, oB3 = oW3 + H;
// offsets in the prepared transposes
constexpr int oW1T = 0, oW2T = H * IN;
constexpr int kLdsDoubles = 2 * IN * H + 2 * H * H + 3 * H;
static_assert(oB3 + 1 == kHeadTheta, "theta layout");

__device__ __forceinline__ double bcast(double v, int lane)
{
    const int lo = __builtin_amdgcn_readlane(__double2loint(v), lane);
    const int hi = __builtin_amdgcn_readlane(__double2hiint(v), lane);
    return __hiloint2double(hi, lo);
}

// silu(a) = a * sigmoid(a) and its first two derivatives
struct Silu {
    double h, s, q;
};
__device__ __forceinline__ Silu silu(double a)
{
    const double g = 1.0 / (1.0 + exp(-a));
    Silu r;
    r.h = a * g;
    r.s = g * (1.0 + a * (1.0 - g));
    r.q = g * (1.0 - g) * (2.0 + a * (1.0 - 2.0 * g));
    return r;
}

struct HeadLds {
    double *W1, *W2, *W1T, *W2T, *b1, *b2, *w3;
};

__device__ HeadLds load_weights(const HeadNet& Hn, double* w)
{
    HeadLds L;
    L.W1 = w; L.W2 = L.W1 + IN * H; L.W1T = L.W2 + H * H; L.W2T = L.W1T + H * IN;
    L.b1 = L.W2T + H * H; L.b2 = L.b1 + H; L.w3 = L.b2 + H;
    for (int q = threadIdx.x; q < IN * H; q += blockDim.x) {
        L.W1[q] = Hn.theta[oW1 + q];
        L.W1T[q] = Hn.wt[oW1T + q];
    }
    for (int q = threadIdx.x; q < H * H; q += blockDim.x) {
        L.W2[q] = Hn.theta[oW2 + q];
        L.W2T[q] = Hn.wt[oW2T + q];
    }
    for (int q = threadIdx.x; q < H; q += blockDim.x) {
        L.b1[q] = Hn.theta[oB1 + q];
        L.b2[q] = Hn.theta[oB2 + q];
        L.w3[q] = Hn.theta[oW3 + q];
    }
    __syncthreads();
    return L;
}

// this lane's input z (normalised basis value, or the species one-hot)
__device__ __forceinline__ double input_of(const HeadNet& Hn, const double* basis, int a, int ts, int lane)
{
    if (lane < Hn.S) return (basis[(size_t)a * Hn.S + lane] - Hn.mean[lane]) * Hn.inv_std[lane];
    return lane == Hn.S + ts ? 1.0 : 0.0;
}

__global__ void prepare_kernel(HeadNet Hn, double* wt)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < IN * H) {
        const int l = t / IN, i = t - l * IN;
        wt[oW1T + t] = Hn.theta[oW1 + i * H + l];
    }
    if (t < H * H) {
        const int j = t / H, l = t - j * H;
        wt[oW2T + t] = Hn.theta[oW2 + l * H + j];
    }
}

__global__ void zero_count_kernel(int* c) { *c = 0; }

// Inference on FP64 matrix cores: one wavefront per 16 atoms.  Each layer
// is computed transposed (hidden units on the 16 rows of an MFMA tile,
// atoms on its 16 columns), out = W^T z with v_mfma_f64_16x16x4f64: the
// A operand (lane l: row l % 16, k l / 16) is read from the weights in LDS,
// the B operand (lane l: k l / 16, column l % 16) is the previous layer's
// output.  The output of a 16x16 tile holds, on lane l, rows 4 q + l / 16
// (q = 0..3) of column l % 16, which is exactly the B operand of k-step
// 4 t + q of the next product (t = the tile), so activations never leave
// the registers.  Backward to the inputs uses the same weights read along
// the other index (row stride kWS, odd, so both orientations read
// conflict-free).
typedef double dbl4 __attribute__((ext_vector_type(4)));
constexpr int kWS = H + 1;
constexpr int kEvalLdsDoubles = IN * kWS + H * kWS + 3 * H;
constexpr int kEvalBlock = 256;
static_assert(IN == kWave && H == kWave, "one hidden unit / input per 16x16 tile row");

__device__ __forceinline__ dbl4 mfma4(double a, double b, dbl4 c)
{
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

__global__ void __launch_bounds__(kEvalBlock) __attribute__((amdgpu_waves_per_eu(2))) head_eval_kernel(const HeadNet Hn, int n, const int* itypes,
                                                               const double* basis, double* E, double* xi,
                                                               int* flags, int* rows, int* count, int* exact_mask)
{
    extern __shared__ double w[];
    double* W1s = w;               // W1[i][h] at i * kWS + h
    double* W2s = W1s + IN * kWS;  // W2[l][j] at l * kWS + j
    double* b1 = W2s + H * kWS;
    double* b2 = b1 + H;
    double* w3 = b2 + H;
    for (int q = threadIdx.x; q < IN * H; q += blockDim.x) W1s[(q / H) * kWS + q % H] = Hn.theta[oW1 + q];
    for (int q = threadIdx.x; q < H * H; q += blockDim.x) W2s[(q / H) * kWS + q % H] = Hn.theta[oW2 + q];
    for (int q = threadIdx.x; q < H; q += blockDim.x) {
        b1[q] = Hn.theta[oB1 + q];
        b2[q] = Hn.theta[oB2 + q];
        w3[q] = Hn.theta[oW3 + q];
    }
    __syncthreads();
    const double b3 = Hn.theta[oB3];
    const int lane = threadIdx.x & (kWave - 1);
    const int col = lane & 15, rq = lane >> 4;
    const int S = Hn.S;
    const int k1 = (S + Hn.Sp + 3) / 4;  // k-steps of the first layer (inputs in use)
    const int tz = (S + 15) / 16;        // tiles of dE/dz that hold basis inputs
    const int ngroups = (n + 15) / 16;
    const int nwb = blockDim.x / kWave;
    for (int g = blockIdx.x * nwb + (int)(threadIdx.x / kWave); g < ngroups; g += gridDim.x * nwb) {
        const int a = g * 16 + col;  // this lane's atom (tile column)
        const bool va = a < n;
        const int t = va ? itypes[a] : 0;
        const int ts = wrap_species(t, Hn.Sp);
        // inputs z[a][4 kb + rq] (normalised basis values, species one-hot)
        double z[16];
        bool out = false;
        double lin_p = 0.0;
#pragma unroll
        for (int kb = 0; kb < 16; ++kb) {
            const int i = 4 * kb + rq;
            double v = 0.0;
            if (va && i < S) {
                const double bv = basis[(size_t)a * S + i];
                v = (bv - Hn.mean[i]) * Hn.inv_std[i];
                if (Hn.lo) out = out || v < Hn.lo[ts * S + i] || v > Hn.hi[ts * S + i];
                if (Hn.lin) lin_p += Hn.lin[i] * bv;
            } else if (va && i == S + ts) {
                v = 1.0;
            }
            z[kb] = v;
        }
        // layer 1: a1^T = W1^T z^T
        dbl4 acc[4];
#pragma unroll
        for (int tt = 0; tt < 4; ++tt) acc[tt] = dbl4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int tp = 0; tp < 4; tp += 2) {
#pragma unroll
            for (int kb = 0; kb < 16; ++kb) {
                if (kb < k1) {
                    acc[tp] = mfma4(W1s[(4 * kb + rq) * kWS + 16 * tp + col], z[kb], acc[tp]);
                    acc[tp + 1] = mfma4(W1s[(4 * kb + rq) * kWS + 16 * tp + 16 + col], z[kb], acc[tp + 1]);
                }
            }
            __builtin_amdgcn_sched_barrier(0);
        }
        double h1[16], s1[16];
#pragma unroll
        for (int tt = 0; tt < 4; ++tt)
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const Silu sv = silu(acc[tt][q] + b1[16 * tt + 4 * q + rq]);
                h1[4 * tt + q] = sv.h;
                s1[4 * tt + q] = sv.s;
            }
        // layer 2: a2^T = W2^T h1^T
#pragma unroll
        for (int tt = 0; tt < 4; ++tt) acc[tt] = dbl4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int tp = 0; tp < 4; tp += 2) {
#pragma unroll
            for (int kb = 0; kb < 16; ++kb) {
                acc[tp] = mfma4(W2s[(4 * kb + rq) * kWS + 16 * tp + col], h1[kb], acc[tp]);
                acc[tp + 1] = mfma4(W2s[(4 * kb + rq) * kWS + 16 * tp + 16 + col], h1[kb], acc[tp + 1]);
            }
            __builtin_amdgcn_sched_barrier(0);
        }
        double da2[16];
        double net = 0.0;
#pragma unroll
        for (int tt = 0; tt < 4; ++tt)
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int h = 16 * tt + 4 * q + rq;
                const Silu sv = silu(acc[tt][q] + b2[h]);
                net += w3[h] * sv.h;
                da2[4 * tt + q] = w3[h] * sv.s;
            }
        net += __shfl_xor(net, 16);
        net += __shfl_xor(net, 32);
        // backward: dh1^T = W2 da2^T, da1 = dh1 silu'(a1)
#pragma unroll
        for (int tt = 0; tt < 4; ++tt) acc[tt] = dbl4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int tp = 0; tp < 4; tp += 2) {
#pragma unroll
            for (int kb = 0; kb < 16; ++kb) {
                acc[tp] = mfma4(W2s[(16 * tp + col) * kWS + 4 * kb + rq], da2[kb], acc[tp]);
                acc[tp + 1] = mfma4(W2s[(16 * tp + 16 + col) * kWS + 4 * kb + rq], da2[kb], acc[tp + 1]);
            }
            __builtin_amdgcn_sched_barrier(0);
        }
        double da1[16];
#pragma unroll
        for (int tt = 0; tt < 4; ++tt)
#pragma unroll
            for (int q = 0; q < 4; ++q) da1[4 * tt + q] = acc[tt][q] * s1[4 * tt + q];
        // dz^T = W1 da1^T over the tiles that hold basis inputs
#pragma unroll
        for (int tt = 0; tt < 4; ++tt) acc[tt] = dbl4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int tp = 0; tp < 4; tp += 2) {
            if (tp < tz) {
#pragma unroll
                for (int kb = 0; kb < 16; ++kb) {
                    acc[tp] = mfma4(W1s[(16 * tp + col) * kWS + 4 * kb + rq], da1[kb], acc[tp]);
                    acc[tp + 1] = mfma4(W1s[(16 * tp + 16 + col) * kWS + 4 * kb + rq], da1[kb], acc[tp + 1]);
                }
            }
            __builtin_amdgcn_sched_barrier(0);
        }
        if (va) {
#pragma unroll
            for (int tt = 0; tt < 4; ++tt)
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const int i = 16 * tt + 4 * q + rq;
                    if (tt < tz && i < S)
                        xi[(size_t)a * S + i] = Hn.e_scale * acc[tt][q] * Hn.inv_std[i] + (Hn.lin ? Hn.lin[i] : 0.0);
                }
        }
        int o = out ? 1 : 0;
        o |= __shfl_xor(o, 16);
        o |= __shfl_xor(o, 32);
        lin_p += __shfl_xor(lin_p, 16);
        lin_p += __shfl_xor(lin_p, 32);
        if (va && rq == 0) {
            const double lin_e = Hn.lin ? lin_p + Hn.lin_sc[ts] : 0.0;
            E[a] = Hn.e_shift[ts] + Hn.e_scale * (net + b3) + lin_e;
            if (flags) flags[a] = o;
            if (count && o && t >= 0) rows[atomicAdd(count, 1)] = a;
            if (exact_mask) exact_mask[a] = (o && t >= 0) ? 1 : 0;
        }
    }
}

constexpr int kGradBlock = 256;  // one wave per SIMD: the per-lane gradient accumulators need ~300 VGPRs

int grad_grid(int n)
{
    const int per_block = (kGradBlock / kWave) * 16;  // >= 16 atoms per wave
    int g = (n + per_block - 1) / per_block;
    if (g > 128) g = 128;
    if (g < 1) g = 1;
    return g;
}

// Per wave: the parameter gradient of sum_a [alpha_a net_a + vz_a . dnet_a/dz]
// over its atoms (forward-over-reverse through the two SiLU layers), written
// as one partial vector per wave.
__global__ void __launch_bounds__(kGradBlock) head_grad_kernel(const HeadNet Hn, int n, const int* itypes,
                                                               const double* basis, const double* alpha,
                                                               const double* v, double* part)
{
    extern __shared__ double w[];
    const HeadLds L = load_weights(Hn, w);
    const int lane = threadIdx.x & (kWave - 1);
    const int wave = (blockIdx.x * blockDim.x + threadIdx.x) / kWave;
    const int nw = (gridDim.x * blockDim.x) / kWave;
    double gW1[IN], gW2[H];
#pragma unroll
    for (int q = 0; q < IN; ++q) gW1[q] = 0.0;
#pragma unroll
    for (int q = 0; q < H; ++q) gW2[q] = 0.0;
    double gb1 = 0.0, gb2 = 0.0, gw3 = 0.0, gb3 = 0.0;
    const double w3l = L.w3[lane];
    for (int a = wave; a < n; a += nw) {
        const int ts = wrap_species(itypes[a], Hn.Sp);
        const double z = input_of(Hn, basis, a, ts, lane);
        const double vz = lane < Hn.S ? v[(size_t)a * Hn.S + lane] * Hn.inv_std[lane] : 0.0;
        const double al = alpha[a];
        // primal and tangent forward
        double a1 = L.b1[lane], ad1 = 0.0;
#pragma unroll
        for (int q = 0; q < IN; ++q) {
            const double wq = L.W1[q * H + lane];
            a1 += wq * bcast(z, q);
            ad1 += wq * bcast(vz, q);
        }
        const Silu s1 = silu(a1);
        const double hd1 = s1.s * ad1;
        double a2 = L.b2[lane], ad2 = 0.0;
#pragma unroll 16
        for (int l = 0; l < H; ++l) {
            const double wl = L.W2[l * H + lane];
            a2 += wl * bcast(s1.h, l);
            ad2 += wl * bcast(hd1, l);
        }
        const Silu s2 = silu(a2);
        const double hd2 = s2.s * ad2;
        // reverse
        gw3 += al * s2.h + hd2;
        gb3 += al;
        const double ac2 = s2.s * w3l;                       // adjoint of the tangent a2
        const double ab2 = w3l * (al * s2.s + s2.q * ad2);   // adjoint of a2
        gb2 += ab2;
#pragma unroll
        for (int l = 0; l < H; ++l) gW2[l] += bcast(s1.h, l) * ab2 + bcast(hd1, l) * ac2;
        double hb1 = 0.0, hcb1 = 0.0;
#pragma unroll 16
        for (int j = 0; j < H; ++j) {
            const double wj = L.W2T[j * H + lane];
            hb1 += wj * bcast(ab2, j);
            hcb1 += wj * bcast(ac2, j);
        }
        const double ac1 = s1.s * hcb1;
        const double ab1 = s1.s * hb1 + s1.q * ad1 * hcb1;
        gb1 += ab1;
#pragma unroll
        for (int q = 0; q < IN; ++q) gW1[q] += bcast(z, q) * ab1 + bcast(vz, q) * ac1;
    }
    double* p = part + (size_t)wave * kHeadTheta;
#pragma unroll
    for (int q = 0; q < IN; ++q) p[oW1 + q * H + lane] = gW1[q];
    p[oB1 + lane] = gb1;
#pragma unroll
    for (int l = 0; l < H; ++l) p[oW2 + l * H + lane] = gW2[l];
    p[oB2 + lane] = gb2;
    p[oW3 + lane] = gw3;
    if (lane == 0) p[oB3] = gb3;
}

__global__ void grad_reduce_kernel(const double* part, int nwaves, double scale, double* grad)
{
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= kHeadTheta) return;
    double s = 0.0;
    for (int w = 0; w < nwaves; ++w) s += part[(size_t)w * kHeadTheta + q];
    grad[q] = scale * s;
}

__global__ void jac_apply_kernel(int n, int K3, int S, const double* J, const double* xi, double* g)
{
    const long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (t >= (long long)n * K3) return;
    const int a = (int)(t / K3), q = (int)(t - (long long)a * K3);
    const double* Ja = J + (size_t)a * S * K3 + q;
    const double* xa = xi + (size_t)a * S;
    double acc = 0.0;
    for (int s = 0; s < S; ++s) acc += xa[s] * Ja[(size_t)s * K3];
    g[t] = acc;
}

__global__ void __launch_bounds__(256) jac_tapply_kernel(int n, int K, int S, int n_f, const double* J, const int* js,
                                                         const double* gF, double* v)
{
    extern __shared__ double d[];  // [K * 3]
    const int a = blockIdx.x;
    const int K3 = K * 3;
    for (int kc = threadIdx.x; kc < K3; kc += blockDim.x) {
        const int k = kc / 3, c = kc - 3 * k;
        const int j = js[(size_t)a * K + k];
        double x = 0.0;
        if (a < n_f) x = gF[(size_t)a * 3 + c] - ((j >= 0 && j < n_f) ? gF[(size_t)j * 3 + c] : 0.0);
        d[kc] = x;
    }
    __syncthreads();
    const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave, nwv = blockDim.x / kWave;
    for (int s = wave; s < S; s += nwv) {
        const double* Js = J + ((size_t)a * S + s) * K3;
        double acc = 0.0;
        for (int kc = lane; kc < K3; kc += kWave) acc += Js[kc] * d[kc];
        acc = wave_sum(acc);
        if (lane == 0) v[(size_t)a * S + s] = acc;
    }
}

}  // namespace

hipError_t head_prepare(const HeadNet& H_, double* wt, hipStream_t s)
{
    hipLaunchKernelGGL(prepare_kernel, dim3((H * H + 255) / 256), dim3(256), 0, s, H_, wt);
    return hipGetLastError();
}

hipError_t head_zero_count(int* count, hipStream_t s)
{
    hipLaunchKernelGGL(zero_count_kernel, dim3(1), dim3(1), 0, s, count);
    return hipGetLastError();
}

hipError_t head_eval(const HeadNet& Hn, int n, const int* itypes, const double* basis, double* E, double* xi,
                     int* flags, int* rows, int* count, hipStream_t s, int* exact_mask)
{
    if (n <= 0) return hipSuccess;
    static bool attr = false;
    if (!attr) {
        hipError_t e = hipFuncSetAttribute((const void*)head_eval_kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                                           kEvalLdsDoubles * (int)sizeof(double));
        if (e != hipSuccess) return e;
        attr = true;
    }
    const int ngroups = (n + 15) / 16;
    const int per_block = kEvalBlock / kWave;
    int grid = (ngroups + per_block - 1) / per_block;
    if (grid > 1024) grid = 1024;
    hipLaunchKernelGGL(head_eval_kernel, dim3(grid), dim3(kEvalBlock), kEvalLdsDoubles * sizeof(double), s, Hn, n,
                       itypes, basis, E, xi, flags, rows, count, exact_mask);
    return hipGetLastError();
}

size_t head_grad_partials(int n) { return (size_t)grad_grid(n) * (kGradBlock / kWave) * kHeadTheta; }

hipError_t head_param_grad(const HeadNet& Hn, int n, const int* itypes, const double* basis, const double* alpha,
                           const double* v, double* part, double* grad, hipStream_t s)
{
    static bool attr = false;
    if (!attr) {
        hipError_t e = hipFuncSetAttribute((const void*)head_grad_kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                                           kLdsDoubles * (int)sizeof(double));
        if (e != hipSuccess) return e;
        attr = true;
    }
    const int grid = grad_grid(n);
    hipLaunchKernelGGL(head_grad_kernel, dim3(grid), dim3(kGradBlock), kLdsDoubles * sizeof(double), s, Hn, n,
                       itypes, basis, alpha, v, part);
    hipLaunchKernelGGL(grad_reduce_kernel, dim3((kHeadTheta + 255) / 256), dim3(256), 0, s, part,
                       grid * (kGradBlock / kWave), Hn.e_scale, grad);
    return hipGetLastError();
}

hipError_t jacobian_apply(int n, int K, int S, const double* J, const double* xi, double* g, hipStream_t s)
{
    const long long tot = (long long)n * K * 3;
    if (tot <= 0) return hipSuccess;
    hipLaunchKernelGGL(jac_apply_kernel, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, s, n, K * 3, S, J, xi, g);
    return hipGetLastError();
}

hipError_t jacobian_tapply(int n, int K, int S, int n_f, const double* J, const int* js, const double* gF, double* v,
                           hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(jac_tapply_kernel, dim3(n), dim3(256), (size_t)K * 3 * sizeof(double), s, n, K, S, n_f, J, js,
                       gF, v);
    return hipGetLastError();
}

}  // namespace mtp
