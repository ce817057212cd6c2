// Training kernels: packed-batch totals and the hand-written parameter
// gradient (the second backward).
//
// For upstream gradients gE[C] (energies), gF[N][3] (forces) and gS[C][6]
// (stress, reference Voigt order) the parameter gradient of
//     sum_c gE_c E_c + sum_a gF_a . F_a + sum_c gS_c . sigma_c
// equals that of
//     J = sum_i [ alpha_i E_i + sum_k g_ik . Delta_ik ],
//     alpha_i  = gE_c(i),
//     Delta_ik = gF_i - [j_ik < N] gF_{j_ik} + Lambda_c r_ik,
//     Lambda_c = (M + M^T) / (2 V_c),  M built from gS_c in Voigt order,
// because F_a = sum_k g_ak - sum_{(i,k): j_ik = a} g_ik (jax_backend.py:263-303)
// and sigma = sym(sum r (x) g) / V (jax_backend.py:355-360).  The second
// term is the derivative of E_i along the slot perturbation Delta, so
// train_kernel propagates moments m and their tangents mt = dm/dt along
// Delta through the contraction layers, then runs one reverse pass over
// (m, mt) to obtain d J / d {moment, species, radial} coefficients.
// This replaces JAX's reverse-over-reverse differentiation of the training
// loss (motep_jax_train_function.py:191).
#include <hip/hip_runtime.h>

#include <cstdint>

#include "mtp_device.hpp"
#include "mtp_kernels.hpp"
#include "mtp_train.hpp"

namespace mtp {

constexpr int kTB = 256;

__host__ __device__ inline TrainLds train_lds_layout(int amc, int K, int R, int N, int nu_max, int n_parts)
{
    TrainLds L;
    int mlen = amc;
    if (mlen < 3 * n_parts * K * R) mlen = 3 * n_parts * K * R;
    int dlen = amc;
    if (dlen < K * 3 * (nu_max + 1)) dlen = K * 3 * (nu_max + 1);
    L.m = 0;
    L.mt = L.m + mlen;
    L.dm = L.mt + amc;
    L.dmt = L.dm + dlen;
    L.f = L.dmt + amc;
    L.fp = L.f + R * K;
    L.vx = L.fp + R * K;
    L.vy = L.vx + K;
    L.vz = L.vy + K;
    L.r = L.vz + K;
    L.dx = L.r + K;
    L.dy = L.dx + K;
    L.dz = L.dy + K;
    L.ud = L.dz + K;
    L.a = L.ud + K;
    L.b = L.a + N * K;
    L.red = L.b + N * K;
    L.ints = L.red + kTB;
    L.total = L.ints + (2 * K + 8) / 2 + 1;
    return L;
}

size_t train_lds_bytes(int amc, int K, int R, int N, int nu_max, int n_parts)
{
    return sizeof(double) * (size_t)train_lds_layout(amc, K, R, N, nu_max, n_parts).total;
}

__global__ void __launch_bounds__(kTB) train_kernel(const DevPlan P, TrainArgs A)
{
    extern __shared__ double lds[];
    const int row = blockIdx.x;
    const int tid = threadIdx.x;
    const int K = A.nslots, R = P.R, N = P.N;
    const TrainLds L = train_lds_layout(P.amc, K, R, N, P.nu_max, P.n_parts);
    double* m = lds + L.m;
    double* mt = lds + L.mt;
    double* dm = lds + L.dm;
    double* dmt = lds + L.dmt;
    double* f = lds + L.f;
    double* fp = lds + L.fp;
    double* vx = lds + L.vx;
    double* vy = lds + L.vy;
    double* vz = lds + L.vz;
    double* rr = lds + L.r;
    double* Dx = lds + L.dx;
    double* Dy = lds + L.dy;
    double* Dz = lds + L.dz;
    double* ud = lds + L.ud;
    double* ca = lds + L.a;
    double* cb = lds + L.b;
    int* ints = reinterpret_cast<int*>(lds + L.ints);
    int* jtc = ints;              // [K] neighbor type of compact slot
    int* wave_cnt = ints + K;     // [4]
    int* nvp = wave_cnt + 4;

    const int ti = A.itypes[row];
    const int conf = A.conf_of[row];
    const double alpha = A.gE[conf];
    const double* rij_row = A.rij + (size_t)row * K * 3;
    const int* jt_row = A.jtypes + (size_t)row * K;
    const int* js_row = A.js + (size_t)row * K;
    // Lambda = (M + M^T) / (2V) from gS in Voigt order xx yy zz yz xz xy
    const double* gs = A.gS + (size_t)conf * 6;
    const double iv = 0.5 / A.volumes[conf];
    const double Lxx = 2.0 * gs[0] * iv, Lyy = 2.0 * gs[1] * iv, Lzz = 2.0 * gs[2] * iv;
    const double Lyz = gs[3] * iv, Lxz = gs[4] * iv, Lxy = gs[5] * iv;
    const double gFx = A.gF[3 * (size_t)row], gFy = A.gF[3 * (size_t)row + 1], gFz = A.gF[3 * (size_t)row + 2];

    // ---- A: radial basis, coefficient factors, perturbation Delta per slot
    {
        const int k = tid;
        double v0 = 0, v1 = 0, v2 = 0, r = 1.0;
        bool valid = false;
        if (k < K) {
            v0 = rij_row[3 * k]; v1 = rij_row[3 * k + 1]; v2 = rij_row[3 * k + 2];
            r = sqrt(v0 * v0 + v1 * v1 + v2 * v2);
            valid = (jt_row[k] >= 0) && (r < P.max_dist);
        }
        const uint64_t mask = __ballot(valid);
        const int lane = tid & (kWave - 1), wave = tid / kWave;
        if (lane == 0) wave_cnt[wave] = __popcll(mask);
        __syncthreads();
        int off = 0;
        for (int w = 0; w < wave; ++w) off += wave_cnt[w];
        const int pos = off + __popcll(mask & ((1ull << lane) - 1ull));
        if (valid) {
            double fv[kMaxR], dfv[kMaxR];
            radial_slot_coef(P, r, ti, jt_row[k], fv, dfv, ca + pos, cb + pos, K);
#pragma unroll
            for (int mu = 0; mu < kMaxR; ++mu)
                if (mu < R) { f[mu * K + pos] = fv[mu]; fp[mu * K + pos] = dfv[mu]; }
            vx[pos] = v0; vy[pos] = v1; vz[pos] = v2; rr[pos] = r;
            jtc[pos] = jt_row[k];
            const int j = js_row[k];
            double dx = gFx, dy = gFy, dz = gFz;
            if (j >= 0 && j < A.nrows) {
                dx -= A.gF[3 * (size_t)j]; dy -= A.gF[3 * (size_t)j + 1]; dz -= A.gF[3 * (size_t)j + 2];
            }
            dx += Lxx * v0 + Lxy * v1 + Lxz * v2;
            dy += Lxy * v0 + Lyy * v1 + Lyz * v2;
            dz += Lxz * v0 + Lyz * v1 + Lzz * v2;
            Dx[pos] = dx; Dy[pos] = dy; Dz[pos] = dz;
            ud[pos] = (v0 * dx + v1 * dy + v2 * dz) / r;
        }
        if (tid == 0) {
            int tot = 0;
            for (int w = 0; w < kTB / kWave; ++w) tot += wave_cnt[w];
            *nvp = tot;
        }
        __syncthreads();
    }
    const int nv = *nvp;
    for (int x = P.B + tid; x < P.amc; x += kTB) { m[x] = 0.0; mt[x] = 0.0; }

    // ---- B: basic moments and their tangents along Delta
    const int np1 = P.nu_max + 1;
    double* pw = dm;
    for (int t = tid; t < nv * 3; t += kTB) {
        const int k = t / 3, q = t - 3 * k;
        const double u = (q == 0 ? vx[k] : (q == 1 ? vy[k] : vz[k])) / rr[k];
        double* row_pw = pw + (size_t)t * np1;
        double v = 1.0;
        row_pw[0] = 1.0;
        for (int p = 1; p < np1; ++p) { v *= u; row_pw[p] = v; }
    }
    __syncthreads();
    for (int t = tid; t < P.n_mono; t += kTB) {
        const I4 mo = P.mono[t];
        const int nu = mo.w & 0xff, alpha_i = (mo.w >> 8) & 0xffff;
        const int mumax = P.mu_max_of_nu[nu];
        double acc[kMaxR], acct[kMaxR];
#pragma unroll
        for (int mu = 0; mu < kMaxR; ++mu) { acc[mu] = 0.0; acct[mu] = 0.0; }
        for (int k = 0; k < nv; ++k) {
            const double* pk = pw + (size_t)k * 3 * np1;
            const double px = pk[mo.x], py = pk[np1 + mo.y], pz = pk[2 * np1 + mo.z];
            const double Pm = px * py * pz;
            // grad_u P . Delta
            double gd = 0.0;
            if (mo.x > 0) gd += mo.x * pk[mo.x - 1] * py * pz * Dx[k];
            if (mo.y > 0) gd += mo.y * px * pk[np1 + mo.y - 1] * pz * Dy[k];
            if (mo.z > 0) gd += mo.z * px * py * pk[2 * np1 + mo.z - 1] * Dz[k];
            const double ir = 1.0 / rr[k];
            const double tP = ud[k] * Pm;                      // (u . Delta) P
            const double tG = (gd - (double)nu * tP) * ir;     // grad_r P . Delta
#pragma unroll
            for (int mu = 0; mu < kMaxR; ++mu)
                if (mu <= mumax) {
                    acc[mu] += f[mu * K + k] * Pm;
                    acct[mu] += fp[mu * K + k] * tP + f[mu * K + k] * tG;
                }
        }
#pragma unroll
        for (int mu = 0; mu < kMaxR; ++mu)
            if (mu <= mumax) {
                m[P.blk_off[mu][nu] + alpha_i] = acc[mu];
                mt[P.blk_off[mu][nu] + alpha_i] = acct[mu];
            }
    }
    __syncthreads();

    // ---- C: contraction layers for m and mt
    for (int d = 1; d <= P.n_depth; ++d) {
        const int t0 = P.fwd_layer[d], t1 = P.fwd_layer[d + 1];
        for (int t = t0 + tid; t < t1; t += kTB) {
            double s = 0.0, st = 0.0;
            const int e1 = P.fwd_beg[t + 1];
            for (int e = P.fwd_beg[t]; e < e1; ++e) {
                const uint32_t pr = P.fwd_pair[e];
                const int i1 = pr & 0xffffu, i2 = pr >> 16;
                const double w = P.fwd_mult[e];
                s += w * m[i1] * m[i2];
                st += w * (mt[i1] * m[i2] + m[i1] * mt[i2]);
            }
            m[P.fwd_tgt[t]] = s;
            mt[P.fwd_tgt[t]] = st;
        }
        __syncthreads();
    }

    // ---- D: dJ/dxi_s partial for this atom
    for (int s = tid; s < P.S; s += kTB) {
        const int x = P.mapping[s];
        A.gxi_part[(size_t)row * P.S + s] = alpha * m[x] + mt[x];
    }

    // ---- E: reverse over (m, mt): dmt = dJ/dmt, dm = dJ/dm
    for (int li = 0; li <= P.n_depth; ++li) {
        const int t0 = P.rev_layer[li], t1 = P.rev_layer[li + 1];
        for (int t = t0 + tid; t < t1; t += kTB) {
            const int x = P.rev_x[t];
            const double xi = P.xi_of[x];
            double s = alpha * xi, st = xi;
            const int e1 = P.rev_beg[t + 1];
            for (int e = P.rev_beg[t]; e < e1; ++e) {
                const uint32_t pr = P.rev_pair[e];
                const int o = pr & 0xffffu, tg = pr >> 16;
                const double w = P.rev_mult[e];
                s += w * (m[o] * dm[tg] + mt[o] * dmt[tg]);
                st += w * m[o] * dmt[tg];
            }
            dm[x] = s;
            dmt[x] = st;
        }
        __syncthreads();
    }

    // ---- F: per slot, per mu: A = sum dm P, Ap = sum dmt P, Q = sum dmt grad_r P . Delta
    double* accs = m;  // [part][slot][3][R]
    {
        const int kk = tid & (kWave - 1), part = tid / kWave;
        for (int k = kk; k < nv && part < P.n_parts; k += kWave) {
            const double ir = 1.0 / rr[k];
            const double ux = vx[k] * ir, uy = vy[k] * ir, uz = vz[k] * ir;
            double px[kMaxNu + 1], py[kMaxNu + 1], pz[kMaxNu + 1];
            px[0] = py[0] = pz[0] = 1.0;
#pragma unroll
            for (int p = 1; p <= kMaxNu; ++p) { px[p] = px[p - 1] * ux; py[p] = py[p - 1] * uy; pz[p] = pz[p - 1] * uz; }
            double Am[kMaxR], Apm[kMaxR], Qm[kMaxR];
#pragma unroll
            for (int mu = 0; mu < kMaxR; ++mu) { Am[mu] = 0.0; Apm[mu] = 0.0; Qm[mu] = 0.0; }
            const double ddx = Dx[k], ddy = Dy[k], ddz = Dz[k], udk = ud[k];
            const int b1 = P.part_beg[part + 1];
            for (int bi = P.part_beg[part]; bi < b1; ++bi) {
                const I2 bl = P.part_blk[bi];
                const int mu = bl.x & 0xff, nu = bl.y;
                const double* d = dm + (bl.x >> 8);
                const double* dt = dmt + (bl.x >> 8);
                double a0, a1, q;
                if (nu == 0) {
                    a0 = d[0]; a1 = dt[0]; q = 0.0;
                } else {
                    double Hx = 0, Hy = 0, Hz = 0, Tx = 0, Ty = 0, Tz = 0;
                    grad_block_dispatch(nu, d, px, py, pz, Hx, Hy, Hz);
                    grad_block_dispatch(nu, dt, px, py, pz, Tx, Ty, Tz);
                    const double uH = ux * Hx + uy * Hy + uz * Hz;
                    const double uT = ux * Tx + uy * Ty + uz * Tz;
                    a0 = uH / (double)nu;
                    a1 = uT / (double)nu;
                    q = (Tx * ddx + Ty * ddy + Tz * ddz - uT * udk) * ir;
                }
#pragma unroll
                for (int mm = 0; mm < kMaxR; ++mm)
                    if (mm == mu) { Am[mm] += a0; Apm[mm] += a1; Qm[mm] += q; }
            }
            double* o = accs + ((size_t)part * K + k) * 3 * R;
#pragma unroll
            for (int mu = 0; mu < kMaxR; ++mu)
                if (mu < R) { o[mu] = Am[mu]; o[R + mu] = Apm[mu]; o[2 * R + mu] = Qm[mu]; }
        }
    }
    __syncthreads();
    // dJ/dc[ti][tk][mu][n] for this atom, plus alpha for the species gradient
    {
        const int ncols = P.Sp * R * N;
        double* out = A.gc_part + (size_t)row * (1 + ncols);
        if (tid == 0) out[0] = alpha;
        for (int c = tid; c < ncols; c += kTB) {
            const int n = c % N, mu = (c / N) % R, tk = c / (N * R);
            double s = 0.0;
            for (int k = 0; k < nv; ++k) {
                if (wrap_species(jtc[k], P.Sp) != tk) continue;
                double aq = 0.0, ap = 0.0;
                for (int part = 0; part < P.n_parts; ++part) {
                    const double* o = accs + ((size_t)part * K + k) * 3 * R;
                    aq += o[mu] + o[2 * R + mu];
                    ap += o[R + mu];
                }
                s += aq * ca[n * K + k] + ap * ud[k] * cb[n * K + k];
            }
            out[1 + c] = s;
        }
    }
}

// Column sums over row chunks; with rowtype the sums are split by type.
// partial[chunk][t][col], chunk = 1024 rows, fixed summation order.
__global__ void __launch_bounds__(256) colsum_stage1(const double* X, int rows, int cols, const int* rowtype, int ntypes,
                                                     int Sp, double* partial)
{
    __shared__ double buf[4][64];
    const int col = blockIdx.x * 64 + (threadIdx.x & 63);
    const int lane = threadIdx.x >> 6;
    const int chunk = blockIdx.y;
    const int r0 = chunk * 1024, r1 = min(rows, r0 + 1024);
    for (int t = 0; t < ntypes; ++t) {
        double s = 0.0;
        if (col < cols)
            for (int r = r0 + lane; r < r1; r += 4) {
                if (rowtype && wrap_species(rowtype[r], Sp) != t) continue;
                s += X[(size_t)r * cols + col];
            }
        buf[lane][threadIdx.x & 63] = s;
        __syncthreads();
        if (lane == 0 && col < cols) {
            const int c = threadIdx.x & 63;
            partial[((size_t)chunk * ntypes + t) * cols + col] = ((buf[0][c] + buf[1][c]) + buf[2][c]) + buf[3][c];
        }
        __syncthreads();
    }
}

__global__ void colsum_stage2(const double* partial, int nchunk, int width, double* out)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= width) return;
    double s = 0.0;
    for (int q = 0; q < nchunk; ++q) s += partial[(size_t)q * width + c];
    out[c] = s;
}

// grad = [dxi (S) | dspecies (Sp) | dradial (Sp*Sp*R*N)] from the reduced parts.
__global__ void assemble_grad(const double* xi_sum, const double* typed_sum, int S, int Sp, int ncols, double* grad)
{
    const int total = S + Sp + Sp * ncols;
    for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < total; q += gridDim.x * blockDim.x) {
        double v;
        if (q < S) v = xi_sum[q];
        else if (q < S + Sp) v = typed_sum[(size_t)(q - S) * (1 + ncols)];
        else {
            const int c = q - S - Sp, t = c / ncols, rest = c % ncols;
            v = typed_sum[(size_t)t * (1 + ncols) + 1 + rest];
        }
        grad[q] = v;
    }
}

// Per-configuration energy and stress (reference Voigt order, / V_c).
__global__ void __launch_bounds__(256) seg_totals_kernel(const double* Ei, const double* Wi, const int* conf_off,
                                                         const double* volumes, double* E, double* S6)
{
    __shared__ double buf[256];
    __shared__ double acc[10];
    const int c = blockIdx.x, tid = threadIdx.x;
    const int r0 = conf_off[c], r1 = conf_off[c + 1];
    for (int q = 0; q < 10; ++q) {
        double s = 0.0;
        for (int i = r0 + tid; i < r1; i += 256) s += q == 0 ? Ei[i] : Wi[(size_t)i * 9 + q - 1];
        buf[tid] = s;
        __syncthreads();
        for (int w = 128; w > 0; w >>= 1) {
            if (tid < w) buf[tid] += buf[tid + w];
            __syncthreads();
        }
        if (tid == 0) acc[q] = buf[0];
        __syncthreads();
    }
    if (tid == 0) {
        const double* W = acc + 1;
        E[c] = acc[0];
        const double iv = 0.5 / volumes[c];
        const int idx[6] = {0, 4, 8, 5, 2, 1};
        for (int q = 0; q < 6; ++q) {
            const int a = idx[q] / 3, b = idx[q] % 3;
            S6[(size_t)c * 6 + q] = (W[3 * a + b] + W[3 * b + a]) * iv;
        }
    }
}

// xi_of[x] = sum of the moment coefficients mapped to moment x, summed in
// increasing scalar index (the host's order), for the mapped moments only
// (the others stay zero).
__global__ void xi_of_kernel(const double* moment, int n_mapped, const int* map_x, const int* map_beg,
                             const int* map_s, double* xi)
{
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n_mapped) return;
    double v = 0.0;
    for (int t = map_beg[q]; t < map_beg[q + 1]; ++t) v += moment[map_s[t]];
    xi[map_x[q]] = v;
}

hipError_t launch_xi_of(const DevPlan& P, double* xi, hipStream_t stream)
{
    if (P.n_mapped <= 0) return hipSuccess;
    hipLaunchKernelGGL(xi_of_kernel, dim3((P.n_mapped + 255) / 256), dim3(256), 0, stream, P.moment_coeffs,
                       P.n_mapped, P.map_x, P.map_beg, P.map_s, xi);
    return hipGetLastError();
}

// Raises (never lowers) the kernel's dynamic-LDS limit, shared by every open
// potential.
hipError_t set_train_kernel_lds(size_t bytes)
{
    static size_t cur = 0;
    if (bytes <= cur) return hipSuccess;
    hipError_t e = hipFuncSetAttribute((const void*)train_kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
    if (e == hipSuccess) cur = bytes;
    return e;
}

hipError_t launch_train_grad(const DevPlan& P, const TrainArgs& A, size_t lds_bytes, double* scratch_partial,
                             double* xi_sum, double* typed_sum, double* grad, hipStream_t stream)
{
    const int ncols = P.Sp * P.R * P.N;
    hipLaunchKernelGGL(train_kernel, dim3(A.nrows), dim3(kTB), lds_bytes, stream, P, A);
    const int nchunk = (A.nrows + 1023) / 1024;
    // xi columns (no types)
    hipLaunchKernelGGL(colsum_stage1, dim3((P.S + 63) / 64, nchunk), dim3(256), 0, stream, A.gxi_part, A.nrows, P.S,
                       (const int*)nullptr, 1, P.Sp, scratch_partial);
    hipLaunchKernelGGL(colsum_stage2, dim3((P.S + 255) / 256), dim3(256), 0, stream, scratch_partial, nchunk, P.S,
                       xi_sum);
    // alpha + radial columns, split by the row's species
    const int w = 1 + ncols;
    hipLaunchKernelGGL(colsum_stage1, dim3((w + 63) / 64, nchunk), dim3(256), 0, stream, A.gc_part, A.nrows, w,
                       A.itypes, P.Sp, P.Sp, scratch_partial);
    hipLaunchKernelGGL(colsum_stage2, dim3((P.Sp * w + 255) / 256), dim3(256), 0, stream, scratch_partial, nchunk,
                       P.Sp * w, typed_sum);
    hipLaunchKernelGGL(assemble_grad, dim3(8), dim3(256), 0, stream, xi_sum, typed_sum, P.S, P.Sp, ncols, grad);
    return hipGetLastError();
}

hipError_t launch_seg_totals(const double* Ei, const double* Wi, const int* conf_off, const double* volumes, int nconf,
                             double* E, double* S6, hipStream_t stream)
{
    if (nconf > 0)
        hipLaunchKernelGGL(seg_totals_kernel, dim3(nconf), dim3(256), 0, stream, Ei, Wi, conf_off, volumes, E, S6);
    return hipGetLastError();
}

}  // namespace mtp
